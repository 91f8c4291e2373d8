// SENSE coil expansion / combination and the conjugate-gradient update of
// CG-SENSE for a group of reconstructions that share the image size, with
// every reduction in a fixed order.
//
// The encoding operator of a reconstruction is E = F * C (NUFFT of the coil
// images of C x, src/deepmr/recon/alg/linop.py:108-148) and CG runs on
// E^H E x = E^H y (src/deepmr/optim/cg.py:150-212). The NUFFT stages batch
// any mix of reconstructions; what is left per reconstruction is
//   expand   img[c] = S[c] * p                      (SenseOp,
//                                                    src/deepmr/linops/coil.py:11-98)
//   combine  ap = sum_c conj(S[c]) * img[c]           (SenseAdjointOp, :101-205)
//   CG step  alpha = rs / <p, ap>, x += alpha p, r -= alpha ap,
//            rs' = <r, r>, p = r + (rs' / rs) p        (cg.py:203-212)
//
// A reconstruction's result must not depend on which others share a launch,
// so every kernel indexes one reconstruction per blockIdx.y and its block
// partition of the image depends only on the image size: each block reduces
// kSenseEpb pixels in a fixed tree, and every block that needs a total sums
// the per-block partials in block order with the same fixed butterfly.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tn_internal.h"

namespace tn {
namespace {

constexpr int kSenseEpl = 4;                     // pixels per lane
constexpr int kSenseEpb = kBlock * kSenseEpl;    // pixels per block

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // conj(a) * b
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Fixed-order sum over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Fixed-order block sum of (a, b) per thread: waves by butterfly, then the
// four wave sums in wave order. Valid in thread 0.
__device__ __forceinline__ float2 block_sum2(float a, float b) {
    __shared__ float red[2][kBlock / kWave];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
        red[0][w] = a;
        red[1][w] = b;
    }
    __syncthreads();
    float2 s = make_float2(0.0f, 0.0f);
    if (threadIdx.x == 0) {
        s = make_float2(red[0][0], red[1][0]);
#pragma unroll
        for (int k = 1; k < kBlock / kWave; ++k) {
            s.x += red[0][k];
            s.y += red[1][k];
        }
    }
    return s;
}

// Sum of the nblk per-block partials (stride `stride` floats, `nc` of them
// per partial) in a fixed order, computed by every wave of the block; the
// result is the same in every block that calls it.
__device__ __forceinline__ float2 total_of(const float *part, int nblk, int stride, int nc) {
    const int lane = threadIdx.x % kWave;
    float a = 0.0f, b = 0.0f;
    for (int k = lane; k < nblk; k += kWave) {
        a += part[(int64_t)k * stride];
        if (nc > 1) b += part[(int64_t)k * stride + 1];
    }
    return make_float2(wave_sum(a), wave_sum(b));
}

// img[c] = S[c] * p for every coil of a reconstruction in CG mode.
__global__ __launch_bounds__(kBlock) void sense_expand_kernel(const tn_sense_job *__restrict__ jobs,
                                                              float2 *__restrict__ img, int64_t npix) {
    const tn_sense_job &J = jobs[blockIdx.y];
    if (J.mode != 1) return;
    const float2 *S = (const float2 *)J.smaps;
    const float2 *p = (const float2 *)J.p;
    float2 *out = img + J.img_off;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npix; i += (int64_t)gridDim.x * kBlock) {
        const float2 v = p[i];
        for (int c = 0; c < J.ncoils; ++c) out[c * npix + i] = cmul(S[c * npix + i], v);
    }
}

// ap = sum_c conj(S[c]) * img[c], coils in order.
__global__ __launch_bounds__(kBlock) void sense_combine_kernel(const tn_sense_job *__restrict__ jobs,
                                                               const float2 *__restrict__ img, int64_t npix) {
    const tn_sense_job &J = jobs[blockIdx.y];
    if (J.mode > 1) return;
    const float2 *S = (const float2 *)J.smaps;
    const float2 *in = img + J.img_off;
    float2 *ap = (float2 *)J.ap;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npix; i += (int64_t)gridDim.x * kBlock) {
        float2 s = cmulc(S[i], in[i]);
        for (int c = 1; c < J.ncoils; ++c) {
            const float2 t = cmulc(S[c * npix + i], in[c * npix + i]);
            s.x += t.x;
            s.y += t.y;
        }
        ap[i] = s;
    }
}

// Partial (block) sums of <p, ap> = sum conj(p) * ap (CG mode); also moves
// the rs written by the previous step's direction kernel into place.
__global__ __launch_bounds__(kBlock) void cg_dot_kernel(const tn_sense_job *__restrict__ jobs, int64_t npix,
                                                        float *__restrict__ work, int nblk) {
    const tn_sense_job &J = jobs[blockIdx.y];
    if (J.mode != 1) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) J.scal[0] = J.scal[1];
    const float2 *p = (const float2 *)J.p;
    const float2 *ap = (const float2 *)J.ap;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
    float a = 0.0f, b = 0.0f;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < npix) {
            const float2 t = cmulc(p[i], ap[i]);
            a += t.x;
            b += t.y;
        }
    }
    const float2 s = block_sum2(a, b);
    if (threadIdx.x == 0) {
        float *w = work + ((int64_t)blockIdx.y * nblk + blockIdx.x) * 4;
        w[0] = s.x;
        w[1] = s.y;
    }
}

// CG mode: alpha = rs / <p, ap>, x += alpha p, r -= alpha ap. Start mode:
// x = 0, r = ap. Both leave the block partials of <r, r>.
__global__ __launch_bounds__(kBlock) void cg_update_kernel(const tn_sense_job *__restrict__ jobs, int64_t npix,
                                                           float *__restrict__ work, int nblk) {
    const tn_sense_job &J = jobs[blockIdx.y];
    if (J.mode > 1) return;
    float *wrow = work + (int64_t)blockIdx.y * nblk * 4;
    float2 *x = (float2 *)J.x;
    float2 *r = (float2 *)J.r;
    const float2 *p = (const float2 *)J.p;
    const float2 *ap = (const float2 *)J.ap;
    float2 alpha = make_float2(0.0f, 0.0f);
    if (J.mode == 1) {
        const float2 d = total_of(wrow, nblk, 4, 2);
        const float rs = J.scal[0];
        const float den = d.x * d.x + d.y * d.y;
        alpha = make_float2(rs * d.x / den, -rs * d.y / den);
    }
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < npix) {
            float2 rn;
            if (J.mode == 1) {
                const float2 ax = cmul(alpha, p[i]);
                const float2 aa = cmul(alpha, ap[i]);
                const float2 xo = x[i], ro = r[i];
                x[i] = make_float2(xo.x + ax.x, xo.y + ax.y);
                rn = make_float2(ro.x - aa.x, ro.y - aa.y);
            } else {
                x[i] = make_float2(0.0f, 0.0f);
                rn = ap[i];
            }
            r[i] = rn;
            a += rn.x * rn.x + rn.y * rn.y;
        }
    }
    const float2 s = block_sum2(a, 0.0f);
    if (threadIdx.x == 0) wrow[(int64_t)blockIdx.x * 4 + 2] = s.x;
}

// rs' = <r, r>; CG mode: p = r + (rs' / rs) p; start mode: p = r. The new rs
// is parked in scal[1] (scal[0] is still read by the other blocks here).
__global__ __launch_bounds__(kBlock) void cg_direction_kernel(const tn_sense_job *__restrict__ jobs, int64_t npix,
                                                              const float *__restrict__ work, int nblk) {
    const tn_sense_job &J = jobs[blockIdx.y];
    if (J.mode > 1) return;
    const float *wrow = work + (int64_t)blockIdx.y * nblk * 4;
    const float rs_new = total_of(wrow + 2, nblk, 4, 1).x;
    const float beta = J.mode == 1 ? rs_new / J.scal[0] : 0.0f;
    const float2 *r = (const float2 *)J.r;
    float2 *p = (float2 *)J.p;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < npix) {
            const float2 rv = r[i];
            if (J.mode == 1) {
                const float2 pv = p[i];
                p[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
            } else {
                p[i] = rv;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) J.scal[1] = rs_new;
}

int sense_blocks(int64_t npix) { return (int)((npix + kSenseEpb - 1) / kSenseEpb); }

int check_group(const tn_sense_job *jobs, int64_t njobs, int64_t npix) {
    if (njobs < 0 || njobs > 65535) return fail(TN_E_ARG, "njobs must be in [0, 65535] (got %lld)", (long long)njobs);
    if (njobs > 0 && !jobs) return fail(TN_E_ARG, "jobs is NULL");
    if (npix < 1 || npix > ((int64_t)1 << 30)) return fail(TN_E_ARG, "npix out of range (%lld)", (long long)npix);
    return TN_OK;
}

// Coil combination of SenseAdjointOp (src/deepmr/linops/coil.py:101-205):
// x[b, k, p] = sum_s sum_c conj(S[s, b, c, p]) * y[s, b, c, k, p], coils
// summed first, then sets (deepmr's .sum(coil).sum(0) order). One thread per
// output element; the maps are read once per contrast.
__global__ __launch_bounds__(kBlock) void coil_combine_kernel(const float2 *__restrict__ smaps,
                                                             const float2 *__restrict__ y,
                                                             float2 *__restrict__ x, int64_t nsets,
                                                             int64_t nb, int64_t ncoils, int64_t nk,
                                                             int64_t npix) {
    const int64_t total = nb * nk * npix;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
        const int64_t p = i % npix;
        const int64_t k = (i / npix) % nk;
        const int64_t b = i / (npix * nk);
        float2 acc = make_float2(0.0f, 0.0f);
        for (int64_t st = 0; st < nsets; ++st) {
            float2 cs = make_float2(0.0f, 0.0f);
            const float2 *sm = smaps + ((st * nb + b) * ncoils) * npix + p;
            const float2 *yy = y + (((st * nb + b) * ncoils) * nk + k) * npix + p;
            for (int64_t c = 0; c < ncoils; ++c) {
                const float2 v = cmulc(sm[c * npix], yy[c * nk * npix]);
                cs.x += v.x;
                cs.y += v.y;
            }
            acc.x += cs.x;
            acc.y += cs.y;
        }
        x[i] = acc;
    }
}

}  // namespace
}  // namespace tn

using namespace tn;

extern "C" {

int64_t tn_cg_work_floats(int64_t njobs, int64_t npix) {
    if (njobs < 0 || npix < 1) return 0;
    return njobs * sense_blocks(npix) * 4;
}

int tn_sense_expand(const tn_sense_job *jobs, int64_t njobs, int64_t npix, void *img, tn_stream stream) {
    clear_error();
    int rc = check_group(jobs, njobs, npix);
    if (rc || njobs == 0) return rc;
    if (!img) return fail(TN_E_ARG, "img is NULL");
    const unsigned bx = (unsigned)std::min<int64_t>((npix + kBlock - 1) / kBlock, 1024);
    hipLaunchKernelGGL(sense_expand_kernel, dim3(bx, (unsigned)njobs), dim3(kBlock), 0, (hipStream_t)stream, jobs,
                       (float2 *)img, npix);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_sense_combine(const tn_sense_job *jobs, int64_t njobs, int64_t npix, const void *img, tn_stream stream) {
    clear_error();
    int rc = check_group(jobs, njobs, npix);
    if (rc || njobs == 0) return rc;
    if (!img) return fail(TN_E_ARG, "img is NULL");
    const unsigned bx = (unsigned)std::min<int64_t>((npix + kBlock - 1) / kBlock, 1024);
    hipLaunchKernelGGL(sense_combine_kernel, dim3(bx, (unsigned)njobs), dim3(kBlock), 0, (hipStream_t)stream, jobs,
                       (const float2 *)img, npix);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_cg_step(const tn_sense_job *jobs, int64_t njobs, int64_t npix, float *work, tn_stream stream) {
    clear_error();
    int rc = check_group(jobs, njobs, npix);
    if (rc || njobs == 0) return rc;
    if (!work) return fail(TN_E_ARG, "work is NULL");
    const int nblk = sense_blocks(npix);
    const dim3 grid((unsigned)nblk, (unsigned)njobs);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(cg_dot_kernel, grid, dim3(kBlock), 0, s, jobs, npix, work, nblk);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(cg_update_kernel, grid, dim3(kBlock), 0, s, jobs, npix, work, nblk);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(cg_direction_kernel, grid, dim3(kBlock), 0, s, jobs, npix, (const float *)work, nblk);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_coil_combine(const void *smaps, const void *y, void *x, int64_t nsets, int64_t nb, int64_t ncoils,
                    int64_t nk, int64_t npix, tn_stream stream) {
    clear_error();
    if (nsets < 1 || nb < 1 || ncoils < 1 || nk < 1 || npix < 0)
        return fail(TN_E_ARG, "coil combination sizes must be >= 1");
    const int64_t total = nb * nk * npix;
    if (total == 0) return TN_OK;
    if (!smaps || !y || !x) return fail(TN_E_ARG, "NULL data pointer");
    const unsigned blocks = (unsigned)std::min<int64_t>((total + kBlock - 1) / kBlock, 65535);
    hipLaunchKernelGGL(coil_combine_kernel, dim3(blocks), dim3(kBlock), 0, (hipStream_t)stream,
                       (const float2 *)smaps, (const float2 *)y, (float2 *)x, nsets, nb, ncoils, nk, npix);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

}  // extern "C"
