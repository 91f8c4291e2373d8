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

// ---- one differentiable CG iteration (CGStep.forward, cg.py:203-212) ------
// B independent systems of n complex elements, (B, n) contiguous; rho, rho1
// real per system; ad = (alpha, d) per system, kept for the backward. The
// block partition (kSenseEpb elements per block) and the order of every sum
// depend only on n.
constexpr int kCgStride = 8;  // floats per block partial record

__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {  // a / b
    const float den = b.x * b.x + b.y * b.y;
    return make_float2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// d = <p, q> = sum conj(p) q, block partials
__global__ __launch_bounds__(kBlock) void cgf_dot_kernel(const float2 *__restrict__ p, const float2 *__restrict__ q,
                                                         int64_t n, float *__restrict__ work, int nblk) {
    const int64_t off = (int64_t)blockIdx.y * n;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
    float a = 0.0f, b = 0.0f;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < n) {
            const float2 t = cmulc(p[off + i], q[off + i]);
            a += t.x;
            b += t.y;
        }
    }
    const float2 s = block_sum2(a, b);
    if (threadIdx.x == 0) {
        float *w = work + ((int64_t)blockIdx.y * nblk + blockIdx.x) * kCgStride;
        w[0] = s.x;
        w[1] = s.y;
    }
}

// alpha = rho / d, x1 = x + alpha p, r1 = r - alpha q; block partials of |r1|^2
__global__ __launch_bounds__(kBlock) void cgf_update_kernel(const float2 *__restrict__ x, const float2 *__restrict__ r,
                                                            const float2 *__restrict__ p, const float2 *__restrict__ q,
                                                            const float *__restrict__ rho, float2 *__restrict__ x1,
                                                            float2 *__restrict__ r1, float2 *__restrict__ ad,
                                                            int64_t n, float *__restrict__ work, int nblk) {
    const int64_t bsys = blockIdx.y;
    float *wrow = work + bsys * nblk * kCgStride;
    const float2 d = total_of(wrow, nblk, kCgStride, 2);
    const float2 alpha = cdiv(make_float2(rho[bsys], 0.0f), d);
    const int64_t off = bsys * n;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < n) {
            const float2 ax = cmul(alpha, p[off + i]);
            const float2 aq = cmul(alpha, q[off + i]);
            const float2 xo = x[off + i], ro = r[off + i];
            x1[off + i] = make_float2(xo.x + ax.x, xo.y + ax.y);
            const float2 rn = make_float2(ro.x - aq.x, ro.y - aq.y);
            r1[off + i] = rn;
            a += rn.x * rn.x + rn.y * rn.y;
        }
    }
    const float2 s = block_sum2(a, 0.0f);
    if (threadIdx.x == 0) wrow[(int64_t)blockIdx.x * kCgStride + 2] = s.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ad[2 * bsys] = alpha;
        ad[2 * bsys + 1] = d;
    }
}

// rho1 = |r1|^2, p1 = r1 + (rho1 / rho) p
__global__ __launch_bounds__(kBlock) void cgf_direction_kernel(const float2 *__restrict__ r1,
                                                               const float2 *__restrict__ p,
                                                               const float *__restrict__ rho,
                                                               float2 *__restrict__ p1, float *__restrict__ rho1,
                                                               int64_t n, const float *__restrict__ work, int nblk) {
    const int64_t bsys = blockIdx.y;
    const float rs = total_of(work + bsys * nblk * kCgStride + 2, nblk, kCgStride, 1).x;
    const float beta = rs / rho[bsys];
    const int64_t off = bsys * n;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < n) {
            const float2 rv = r1[off + i], pv = p[off + i];
            p1[off + i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rho1[bsys] = rs;
}

// Backward, pass 1: block partials of
//   S1 = Re sum gp1 conj(p), S2 = sum (gr1 + gp1) conj(q), S3 = sum r1 conj(q),
//   S4 = sum gx1 conj(p)   (missing output gradients are zeros)
__global__ __launch_bounds__(kBlock) void cgb_sums_kernel(const float2 *__restrict__ p, const float2 *__restrict__ q,
                                                          const float2 *__restrict__ r1,
                                                          const float2 *__restrict__ gx1,
                                                          const float2 *__restrict__ gr1,
                                                          const float2 *__restrict__ gp1, int64_t n,
                                                          float *__restrict__ work, int nblk) {
    const int64_t off = (int64_t)blockIdx.y * n;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
    float s1 = 0.0f, s2r = 0.0f, s2i = 0.0f, s3r = 0.0f, s3i = 0.0f, s4r = 0.0f, s4i = 0.0f;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < n) {
            const float2 pv = p[off + i], qv = q[off + i], rv = r1[off + i];
            const float2 gp = gp1 ? gp1[off + i] : make_float2(0.0f, 0.0f);
            const float2 gr = gr1 ? gr1[off + i] : make_float2(0.0f, 0.0f);
            const float2 gx = gx1 ? gx1[off + i] : make_float2(0.0f, 0.0f);
            s1 += gp.x * pv.x + gp.y * pv.y;
            const float2 g = make_float2(gr.x + gp.x, gr.y + gp.y);
            const float2 t2 = cmulc(qv, g);  // g conj(q)
            s2r += t2.x;
            s2i += t2.y;
            const float2 t3 = cmulc(qv, rv);
            s3r += t3.x;
            s3i += t3.y;
            const float2 t4 = cmulc(pv, gx);
            s4r += t4.x;
            s4i += t4.y;
        }
    }
    float *w = work + ((int64_t)blockIdx.y * nblk + blockIdx.x) * kCgStride;
    float2 t = block_sum2(s1, s2r);
    if (threadIdx.x == 0) {
        w[0] = t.x;
        w[1] = t.y;
    }
    __syncthreads();  // block_sum2's LDS is reused
    t = block_sum2(s2i, s3r);
    if (threadIdx.x == 0) {
        w[2] = t.x;
        w[3] = t.y;
    }
    __syncthreads();
    t = block_sum2(s3i, s4r);
    if (threadIdx.x == 0) {
        w[4] = t.x;
        w[5] = t.y;
    }
    __syncthreads();
    t = block_sum2(s4i, 0.0f);
    if (threadIdx.x == 0) w[6] = t.x;
}

// Backward, pass 2: the scalar gradients from the totals, then
//   gr = gr1 + gp1 + 2 g_rho1 r1,  gq = -gr conj(alpha) + g_d p,
//   gp = beta gp1 + conj(alpha) gx1 + conj(g_d) q,  grho (real)
// with g_rho1 = grho1 + S1 / rho, g_alpha = S4 - (S2 + 2 g_rho1 S3),
// g_d = g_alpha conj(-rho / d^2), grho = -S1 rho1 / rho^2 + Re(g_alpha conj(1 / d))
// (torch's conjugate-Wirtinger convention for complex gradients).
__global__ __launch_bounds__(kBlock) void cgb_apply_kernel(const float2 *__restrict__ p, const float2 *__restrict__ q,
                                                           const float2 *__restrict__ r1,
                                                           const float2 *__restrict__ gx1,
                                                           const float2 *__restrict__ gr1,
                                                           const float2 *__restrict__ gp1,
                                                           const float *__restrict__ rho,
                                                           const float *__restrict__ rho1,
                                                           const float2 *__restrict__ ad,
                                                           const float *__restrict__ grho1,
                                                           float2 *__restrict__ gr, float2 *__restrict__ gp,
                                                           float2 *__restrict__ gq, float *__restrict__ grho, int64_t n,
                                                           const float *__restrict__ work, int nblk) {
    const int64_t bsys = blockIdx.y;
    const float *wrow = work + bsys * nblk * kCgStride;
    const float2 t01 = total_of(wrow, nblk, kCgStride, 2);
    const float2 t23 = total_of(wrow + 2, nblk, kCgStride, 2);
    const float2 t45 = total_of(wrow + 4, nblk, kCgStride, 2);
    const float t6 = total_of(wrow + 6, nblk, kCgStride, 1).x;
    const float S1 = t01.x;
    const float2 S2 = make_float2(t01.y, t23.x), S3 = make_float2(t23.y, t45.x), S4 = make_float2(t45.y, t6);
    const float rh = rho[bsys], rh1 = rho1[bsys];
    const float2 alpha = ad[2 * bsys], d = ad[2 * bsys + 1];
    const float beta = rh1 / rh;
    const float g_rho1 = (grho1 ? grho1[bsys] : 0.0f) + S1 / rh;
    const float2 g_alpha = make_float2(S4.x - (S2.x + 2.0f * g_rho1 * S3.x), S4.y - (S2.y + 2.0f * g_rho1 * S3.y));
    const float2 d2 = cmul(d, d);
    const float2 da = cdiv(make_float2(-rh, 0.0f), d2);               // d alpha / d d
    const float2 g_d = make_float2(g_alpha.x * da.x + g_alpha.y * da.y, g_alpha.y * da.x - g_alpha.x * da.y);  // g_alpha conj(da)
    const float2 inv_d = cdiv(make_float2(1.0f, 0.0f), d);
    const float g_rho = -S1 * rh1 / (rh * rh) + (g_alpha.x * inv_d.x + g_alpha.y * inv_d.y);
    const float2 ca = make_float2(alpha.x, -alpha.y);
    const int64_t off = bsys * n;
    const int64_t i0 = (int64_t)blockIdx.x * kSenseEpb + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kSenseEpl; ++k) {
        const int64_t i = i0 + k * kBlock;
        if (i < n) {
            const float2 pv = p[off + i], qv = q[off + i], rv = r1[off + i];
            const float2 gpo = gp1 ? gp1[off + i] : make_float2(0.0f, 0.0f);
            const float2 gro = gr1 ? gr1[off + i] : make_float2(0.0f, 0.0f);
            const float2 gxo = gx1 ? gx1[off + i] : make_float2(0.0f, 0.0f);
            const float2 g = make_float2(gro.x + gpo.x + 2.0f * g_rho1 * rv.x, gro.y + gpo.y + 2.0f * g_rho1 * rv.y);
            gr[off + i] = g;
            const float2 a1 = cmul(g, ca), a2 = cmul(g_d, pv);
            gq[off + i] = make_float2(a2.x - a1.x, a2.y - a1.y);
            const float2 b2 = cmul(gxo, ca), b3 = cmulc(g_d, qv);  // conj(g_d) q
            gp[off + i] = make_float2(beta * gpo.x + b2.x + b3.x, beta * gpo.y + b2.y + b3.y);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) grho[bsys] = g_rho;
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

int64_t tn_cg_grad_work_floats(int64_t nsys, int64_t n) {
    if (nsys < 1 || n < 1) return 0;
    return nsys * sense_blocks(n) * kCgStride;
}

int tn_cg_iterate(const void *x, const void *r, const void *p, const void *q, const float *rho, void *x1, void *r1,
                  void *p1, float *rho1, void *ad, float *work, int64_t nsys, int64_t n, tn_stream stream) {
    clear_error();
    if (nsys < 1 || nsys > 65535 || n < 1 || n > ((int64_t)1 << 30))
        return fail(TN_E_ARG, "CG sizes out of range (nsys %lld, n %lld)", (long long)nsys, (long long)n);
    if (!x || !r || !p || !q || !rho || !x1 || !r1 || !p1 || !rho1 || !ad || !work)
        return fail(TN_E_ARG, "NULL data pointer");
    const int nblk = sense_blocks(n);
    const dim3 grid((unsigned)nblk, (unsigned)nsys);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(cgf_dot_kernel, grid, dim3(kBlock), 0, s, (const float2 *)p, (const float2 *)q, n, work, nblk);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(cgf_update_kernel, grid, dim3(kBlock), 0, s, (const float2 *)x, (const float2 *)r,
                       (const float2 *)p, (const float2 *)q, rho, (float2 *)x1, (float2 *)r1, (float2 *)ad, n, work,
                       nblk);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(cgf_direction_kernel, grid, dim3(kBlock), 0, s, (const float2 *)r1, (const float2 *)p, rho,
                       (float2 *)p1, rho1, n, (const float *)work, nblk);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int tn_cg_iterate_backward(const void *p, const void *q, const void *r1, const float *rho, const float *rho1,
                           const void *ad, const void *gx1, const void *gr1, const void *gp1, const float *grho1,
                           void *gr, void *gp, void *gq, float *grho, float *work, int64_t nsys, int64_t n,
                           tn_stream stream) {
    clear_error();
    if (nsys < 1 || nsys > 65535 || n < 1 || n > ((int64_t)1 << 30))
        return fail(TN_E_ARG, "CG sizes out of range (nsys %lld, n %lld)", (long long)nsys, (long long)n);
    if (!p || !q || !r1 || !rho || !rho1 || !ad || !gr || !gp || !gq || !grho || !work)
        return fail(TN_E_ARG, "NULL data pointer");
    const int nblk = sense_blocks(n);
    const dim3 grid((unsigned)nblk, (unsigned)nsys);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(cgb_sums_kernel, grid, dim3(kBlock), 0, s, (const float2 *)p, (const float2 *)q,
                       (const float2 *)r1, (const float2 *)gx1, (const float2 *)gr1, (const float2 *)gp1, n, work,
                       nblk);
    TN_HIP(hipGetLastError());
    hipLaunchKernelGGL(cgb_apply_kernel, grid, dim3(kBlock), 0, s, (const float2 *)p, (const float2 *)q,
                       (const float2 *)r1, (const float2 *)gx1, (const float2 *)gr1, (const float2 *)gp1, rho, rho1,
                       (const float2 *)ad, grho1, (float2 *)gr, (float2 *)gp, (float2 *)gq, grho, n,
                       (const float *)work, nblk);
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
