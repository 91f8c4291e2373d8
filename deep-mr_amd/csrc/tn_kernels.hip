// Interpolation, gridding and the fused pre/post kernels for gfx950.
//
// Interpolation (grid -> samples): one lane per sample in trajectory order,
// taps evaluated in registers once per sample and reused for every coil of
// the work item; coordinate reads and sample writes are coalesced, the grid
// gathers are served from L2 because consecutive samples of a trajectory are
// neighbours in k-space.
//
// Gridding (samples -> grid): one workgroup owns one tile of one frame for a
// group of coils. Samples whose footprint touches the tile (plan-time bins,
// sorted by sample index) are spread into an LDS accumulator; each wave owns
// whole coils, so a tile cell is only ever updated by one wave and the
// updates arrive in the fixed bin order. The finished tile is written once,
// with coalesced rows, zeros included. No global atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tn_internal.h"
#include "tn_kb.h"

namespace tn {
namespace {

template <int NDIM, int WC>
struct Taps {
    int idx[3][WC];
    float w[3][WC];
    int n[3];
};

template <int NDIM, int WC>
__device__ __forceinline__ void compute_taps(const Job &J, float cx, float cy, float cz,
                                             Taps<NDIM, WC> &t) {
    const float c[3] = {cx, cy, cz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d < NDIM) {
            int base;
            axis_taps<WC>(c[d], J.W[d], J.G[d], J.shift[d], J.beta[d], base, t.w[d]);
            t.n[d] = J.W[d];
#pragma unroll
            for (int k = 0; k < WC; ++k) t.idx[d][k] = wrap_add(base, k, J.G[d]);
        } else {
            t.n[d] = 1;
#pragma unroll
            for (int k = 0; k < WC; ++k) {
                t.idx[d][k] = 0;
                t.w[d][k] = k == 0 ? 1.0f : 0.0f;
            }
        }
    }
}

template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void interp_kernel(const Job *__restrict__ jobs,
                                                        const Item *__restrict__ items,
                                                        const float2 *__restrict__ grid,
                                                        float2 *__restrict__ samp,
                                                        const float *__restrict__ weight) {
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int64_t p = (int64_t)it.unit * kInterpBlockPts + threadIdx.x;
    if (p >= J.npts) return;

    const float *c = J.coord + p * NDIM;
    const float cx = c[0];
    const float cy = NDIM > 1 ? c[1] : 0.0f;
    const float cz = NDIM > 2 ? c[2] : 0.0f;
    Taps<NDIM, WC> t;
    compute_taps<NDIM, WC>(J, cx, cy, cz, t);

    float sc = J.inv_scale;
    if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];

    const int Gx = J.G[0], Gy = J.G[1];
    for (int b = it.c0; b < it.c1; ++b) {
        const float2 *g = grid + J.grid_off + (int64_t)b * J.grid_bstride;
        float re = 0.0f, im = 0.0f;
#pragma unroll
        for (int kz = 0; kz < WC; ++kz) {
            if (kz < t.n[2]) {
#pragma unroll
                for (int ky = 0; ky < WC; ++ky) {
                    if (ky < t.n[1]) {
                        const float vzy = t.w[2][kz] * t.w[1][ky];
                        const float2 *row = g + ((int64_t)t.idx[2][kz] * Gy + t.idx[1][ky]) * Gx;
#pragma unroll
                        for (int kx = 0; kx < WC; ++kx) {
                            if (kx < t.n[0]) {
                                const float v = vzy * t.w[0][kx];
                                const float2 s = row[t.idx[0][kx]];
                                re += v * s.x;
                                im += v * s.y;
                            }
                        }
                    }
                }
            }
        }
        samp[J.samp_off + (int64_t)b * J.samp_bstride + p] = make_float2(re * sc, im * sc);
    }
}

// Staged interpolation: one workgroup loads the (tile + W-1 halo) region of
// the grid for a few coils into LDS with coalesced row reads, then every lane
// interpolates one sample whose footprint origin lies in the tile, gathering
// its taps from LDS. Samples come from the plan's primary bins (each sample
// exactly once); hot tiles are split into several work items by entry range.
template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void interp_staged_kernel(const Job *__restrict__ jobs,
                                                               const Item *__restrict__ items,
                                                               const float2 *__restrict__ grid,
                                                               float2 *__restrict__ samp,
                                                               const float *__restrict__ weight) {
    extern __shared__ float2 region[];
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int tile = it.unit;
    const int Tx = J.IT[0], Ty = J.IT[1], Tz = J.IT[2];
    const int tx = tile % J.INT_[0];
    const int ty = (tile / J.INT_[0]) % J.INT_[1];
    const int tz = tile / (J.INT_[0] * J.INT_[1]);
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
    const int Rx = Tx + J.W[0] - 1, Ry = Ty + J.W[1] - 1, Rz = Tz + J.W[2] - 1;
    const int Gx = J.G[0], Gy = J.G[1], Gz = J.G[2];
    const int cells = Rx * Ry * Rz;
    const int ncoil = it.c1 - it.c0;

    for (int i = threadIdx.x; i < ncoil * cells; i += kBlock) {
        const int c = i / cells;
        const int cell = i - c * cells;
        const int lx = cell % Rx;
        const int ly = (cell / Rx) % Ry;
        const int lz = cell / (Rx * Ry);
        int gx = ox + lx, gy = oy + ly, gz = oz + lz;
        while (gx >= Gx) gx -= Gx;
        while (gy >= Gy) gy -= Gy;
        while (gz >= Gz) gz -= Gz;
        region[i] = grid[J.grid_off + (int64_t)(it.c0 + c) * J.grid_bstride +
                         ((int64_t)gz * Gy + gy) * Gx + gx];
    }
    __syncthreads();

    for (int e = it.e0 + (int)threadIdx.x; e < it.e1; e += kBlock) {
        const float4 en = J.pentries[e];
        const int p = __float_as_int(en.w);
        float w[3][WC];
        int base[3];
        const float c3[3] = {en.x, en.y, en.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d < NDIM) {
                axis_taps<WC>(c3[d], J.W[d], J.G[d], J.shift[d], J.beta[d], base[d], w[d]);
            } else {
                base[d] = 0;
#pragma unroll
                for (int k = 0; k < WC; ++k) w[d][k] = k == 0 ? 1.0f : 0.0f;
            }
        }
        // local origin of the footprint inside the region (clamped: a sample
        // always bins into the tile holding its origin, the clamp only keeps
        // LDS addressing in range)
        const int lx0 = min(max(base[0] - ox, 0), Tx - 1);
        const int ly0 = min(max(base[1] - oy, 0), Ty - 1);
        const int lz0 = min(max(base[2] - oz, 0), Tz - 1);
        const int nx = J.W[0], ny = NDIM > 1 ? J.W[1] : 1, nz = NDIM > 2 ? J.W[2] : 1;
        float sc = J.inv_scale;
        if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];
        for (int b = 0; b < ncoil; ++b) {
            const float2 *r = region + (size_t)b * cells;
            float re = 0.0f, im = 0.0f;
#pragma unroll
            for (int kz = 0; kz < WC; ++kz) {
                if (kz < nz) {
#pragma unroll
                    for (int ky = 0; ky < WC; ++ky) {
                        if (ky < ny) {
                            const float vzy = w[2][kz] * w[1][ky];
                            const float2 *row = r + ((lz0 + kz) * Ry + (ly0 + ky)) * Rx + lx0;
#pragma unroll
                            for (int kx = 0; kx < WC; ++kx) {
                                if (kx < nx) {
                                    const float v = vzy * w[0][kx];
                                    const float2 s = row[kx];
                                    re += v * s.x;
                                    im += v * s.y;
                                }
                            }
                        }
                    }
                }
            }
            samp[J.samp_off + (int64_t)(it.c0 + b) * J.samp_bstride + p] = make_float2(re * sc, im * sc);
        }
    }
}

template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void grid_kernel(const Job *__restrict__ jobs,
                                                      const Item *__restrict__ items,
                                                      const float2 *__restrict__ samp,
                                                      float2 *__restrict__ grid,
                                                      const float *__restrict__ weight,
                                                      float *__restrict__ scratch,
                                                      const int64_t *__restrict__ slot_base) {
    extern __shared__ float lds[];
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int tile = it.unit;
    const int Tx = J.T[0], Ty = J.T[1], Tz = J.T[2];
    const int tx = tile % J.NT[0];
    const int ty = (tile / J.NT[0]) % J.NT[1];
    const int tz = tile / (J.NT[0] * J.NT[1]);
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
    const int ex = min(Tx, J.G[0] - ox), ey = min(Ty, J.G[1] - oy), ez = min(Tz, J.G[2] - oz);
    const int cells = Tx * Ty * Tz;
    const int ncoil = it.c1 - it.c0;

    for (int i = threadIdx.x; i < 2 * ncoil * cells; i += kBlock) lds[i] = 0.0f;
    __syncthreads();

    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int e0 = it.e0;
    const int e1 = it.e1;
    if (wave < ncoil) {
        for (int e = e0 + lane; e < e1; e += kWave) {
            const float4 en = J.entries[e];
            const int p = __float_as_int(en.w);
            Taps<NDIM, WC> t;
            compute_taps<NDIM, WC>(J, en.x, en.y, en.z, t);
            // local tile coordinates (-1 outside this tile)
            int lx[WC], ly[WC], lz[WC];
#pragma unroll
            for (int k = 0; k < WC; ++k) {
                const int ax = t.idx[0][k] - ox, ay = t.idx[1][k] - oy, az = t.idx[2][k] - oz;
                lx[k] = (k < t.n[0] && ax >= 0 && ax < ex) ? ax : -1;
                ly[k] = (k < t.n[1] && ay >= 0 && ay < ey) ? ay : -1;
                lz[k] = (k < t.n[2] && az >= 0 && az < ez) ? az : -1;
            }
            float sc = J.inv_scale;
            if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];
            for (int b = it.c0 + wave; b < it.c1; b += kBlock / kWave) {
                float2 s = samp[J.samp_off + (int64_t)b * J.samp_bstride + p];
                s.x *= sc;
                s.y *= sc;
                float *acc = lds + (size_t)2 * (b - it.c0) * cells;
#pragma unroll
                for (int kz = 0; kz < WC; ++kz) {
                    if (lz[kz] >= 0) {
#pragma unroll
                        for (int ky = 0; ky < WC; ++ky) {
                            if (ly[ky] >= 0) {
                                const float vzy = t.w[2][kz] * t.w[1][ky];
                                const int rowc = (lz[kz] * Ty + ly[ky]) * Tx;
#pragma unroll
                                for (int kx = 0; kx < WC; ++kx) {
                                    if (lx[kx] >= 0) {
                                        const float v = vzy * t.w[0][kx];
                                        float *a = acc + 2 * (rowc + lx[kx]);
                                        __hip_atomic_fetch_add(a, v * s.x, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                                        __hip_atomic_fetch_add(a + 1, v * s.y, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    if (it.slot >= 0) {
        // split tile: store this partial; the fold kernel sums the parts in
        // part order and writes the tile
        float *dst = scratch + slot_base[it.slot] + (int64_t)it.part * 2 * ncoil * cells;
        for (int i = threadIdx.x; i < 2 * ncoil * cells; i += kBlock) dst[i] = lds[i];
        return;
    }
    const int Gx = J.G[0], Gy = J.G[1];
    for (int i = threadIdx.x; i < ncoil * cells; i += kBlock) {
        const int b = i / cells;
        const int cell = i - b * cells;
        const int lx = cell % Tx;
        const int ly = (cell / Tx) % Ty;
        const int lz = cell / (Tx * Ty);
        if (lx < ex && ly < ey && lz < ez) {
            const float *a = lds + 2 * ((size_t)b * cells + cell);
            grid[J.grid_off + (int64_t)(it.c0 + b) * J.grid_bstride +
                 ((int64_t)(oz + lz) * Gy + (oy + ly)) * Gx + (ox + lx)] = make_float2(a[0], a[1]);
        }
    }
}

// Fold of split gridding tiles: out = part 0 + part 1 + ... in that order.
__global__ __launch_bounds__(kBlock) void fold_kernel(const Job *__restrict__ jobs,
                                                      const Item *__restrict__ items,
                                                      const float *__restrict__ scratch,
                                                      const int64_t *__restrict__ slot_base,
                                                      float2 *__restrict__ grid) {
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int tile = it.unit;
    const int Tx = J.T[0], Ty = J.T[1], Tz = J.T[2];
    const int tx = tile % J.NT[0];
    const int ty = (tile / J.NT[0]) % J.NT[1];
    const int tz = tile / (J.NT[0] * J.NT[1]);
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
    const int ex = min(Tx, J.G[0] - ox), ey = min(Ty, J.G[1] - oy), ez = min(Tz, J.G[2] - oz);
    const int cells = Tx * Ty * Tz;
    const int ncoil = it.c1 - it.c0;
    const int nparts = it.e1;
    const float *src = scratch + slot_base[it.e0];
    const int64_t pstride = 2 * (int64_t)ncoil * cells;
    const int Gx = J.G[0], Gy = J.G[1];
    for (int i = threadIdx.x; i < ncoil * cells; i += kBlock) {
        const int b = i / cells;
        const int cell = i - b * cells;
        const int lx = cell % Tx;
        const int ly = (cell / Tx) % Ty;
        const int lz = cell / (Tx * Ty);
        if (lx < ex && ly < ey && lz < ez) {
            float re = src[2 * i], im = src[2 * i + 1];
            for (int q = 1; q < nparts; ++q) {
                re += src[q * pstride + 2 * i];
                im += src[q * pstride + 2 * i + 1];
            }
            grid[J.grid_off + (int64_t)(it.c0 + b) * J.grid_bstride +
                 ((int64_t)(oz + lz) * Gy + (oy + ly)) * Gx + (ox + lx)] = make_float2(re, im);
        }
    }
}

// ---- fused pre/post processing -------------------------------------------
struct PrePost {
    int N[3];   // image (x, y, z)
    int G[3];   // grid  (x, y, z)
    const float *apod[3];  // (x, y, z), NULL = ones
};

PrePost make_prepost(const tn_plan *plan) {
    PrePost pp{};
    for (int d = 0; d < 3; ++d) {
        const int zyx = 2 - d;
        pp.N[d] = plan->p.shape[zyx];
        pp.G[d] = plan->p.grid[zyx];
        pp.apod[d] = plan->d_apod[zyx];
    }
    return pp;
}

__device__ __forceinline__ float2 apodize(float2 v, const PrePost &pp, int mx, int my, int mz) {
    // same axis order as deepmr's _apodize (x first, then y, then z)
    if (pp.apod[0]) { const float a = pp.apod[0][mx]; v.x *= a; v.y *= a; }
    if (pp.apod[1]) { const float a = pp.apod[1][my]; v.x *= a; v.y *= a; }
    if (pp.apod[2]) { const float a = pp.apod[2][mz]; v.x *= a; v.y *= a; }
    return v;
}

// grid index i <-> image index m = (i + N/2) mod G (valid when m < N)
__global__ __launch_bounds__(kBlock) void apod_pad_kernel(PrePost pp, const float2 *__restrict__ img,
                                                          float2 *__restrict__ grid, int64_t nimg) {
    const int64_t gvol = (int64_t)pp.G[0] * pp.G[1] * pp.G[2];
    const int64_t nvol = (int64_t)pp.N[0] * pp.N[1] * pp.N[2];
    const int64_t total = nimg * gvol;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kBlock) {
        const int64_t b = i / gvol;
        const int64_t r = i - b * gvol;
        const int x = (int)(r % pp.G[0]);
        const int y = (int)((r / pp.G[0]) % pp.G[1]);
        const int z = (int)(r / ((int64_t)pp.G[0] * pp.G[1]));
        int mx = x + pp.N[0] / 2; if (mx >= pp.G[0]) mx -= pp.G[0];
        int my = y + pp.N[1] / 2; if (my >= pp.G[1]) my -= pp.G[1];
        int mz = z + pp.N[2] / 2; if (mz >= pp.G[2]) mz -= pp.G[2];
        float2 v = make_float2(0.0f, 0.0f);
        if (mx < pp.N[0] && my < pp.N[1] && mz < pp.N[2]) {
            v = img[b * nvol + ((int64_t)mz * pp.N[1] + my) * pp.N[0] + mx];
            v = apodize(v, pp, mx, my, mz);
        }
        grid[i] = v;
    }
}

__global__ __launch_bounds__(kBlock) void crop_apod_kernel(PrePost pp, const float2 *__restrict__ grid,
                                                           float2 *__restrict__ img, int64_t nimg) {
    const int64_t gvol = (int64_t)pp.G[0] * pp.G[1] * pp.G[2];
    const int64_t nvol = (int64_t)pp.N[0] * pp.N[1] * pp.N[2];
    const int64_t total = nimg * nvol;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kBlock) {
        const int64_t b = i / nvol;
        const int64_t r = i - b * nvol;
        const int mx = (int)(r % pp.N[0]);
        const int my = (int)((r / pp.N[0]) % pp.N[1]);
        const int mz = (int)(r / ((int64_t)pp.N[0] * pp.N[1]));
        int x = mx - pp.N[0] / 2; if (x < 0) x += pp.G[0];
        int y = my - pp.N[1] / 2; if (y < 0) y += pp.G[1];
        int z = mz - pp.N[2] / 2; if (z < 0) z += pp.G[2];
        float2 v = grid[b * gvol + ((int64_t)z * pp.G[1] + y) * pp.G[0] + x];
        img[i] = apodize(v, pp, mx, my, mz);
    }
}

unsigned stride_blocks(int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 256 * 64));
}

// ---- dispatch ----------------------------------------------------------------
using InterpFn = void (*)(const Job *, const Item *, const float2 *, float2 *, const float *);
using GridFn = void (*)(const Job *, const Item *, const float2 *, float2 *, const float *, float *,
                        const int64_t *);

template <int NDIM>
InterpFn pick_interp(int wc) {
    switch (wc) {
        case 1: return interp_kernel<NDIM, 1>;
        case 2: return interp_kernel<NDIM, 2>;
        case 3: return interp_kernel<NDIM, 3>;
        case 4: return interp_kernel<NDIM, 4>;
        case 5: return interp_kernel<NDIM, 5>;
        case 6: return interp_kernel<NDIM, 6>;
        case 7: return interp_kernel<NDIM, 7>;
        case 8: return interp_kernel<NDIM, 8>;
        default: return nullptr;
    }
}

template <int NDIM>
InterpFn pick_staged(int wc) {
    switch (wc) {
        case 1: return interp_staged_kernel<NDIM, 1>;
        case 2: return interp_staged_kernel<NDIM, 2>;
        case 3: return interp_staged_kernel<NDIM, 3>;
        case 4: return interp_staged_kernel<NDIM, 4>;
        case 5: return interp_staged_kernel<NDIM, 5>;
        case 6: return interp_staged_kernel<NDIM, 6>;
        case 7: return interp_staged_kernel<NDIM, 7>;
        case 8: return interp_staged_kernel<NDIM, 8>;
        default: return nullptr;
    }
}

template <int NDIM>
GridFn pick_grid(int wc) {
    switch (wc) {
        case 1: return grid_kernel<NDIM, 1>;
        case 2: return grid_kernel<NDIM, 2>;
        case 3: return grid_kernel<NDIM, 3>;
        case 4: return grid_kernel<NDIM, 4>;
        case 5: return grid_kernel<NDIM, 5>;
        case 6: return grid_kernel<NDIM, 6>;
        case 7: return grid_kernel<NDIM, 7>;
        case 8: return grid_kernel<NDIM, 8>;
        default: return nullptr;
    }
}

}  // namespace

// kind = ndim * 16 + max width
int interp_kind(const std::vector<Job> &jobs) {
    int ndim = 0, wc = 1;
    for (const Job &j : jobs) {
        ndim = std::max(ndim, j.ndim);
        for (int d = 0; d < 3; ++d) wc = std::max(wc, j.W[d]);
    }
    return ndim * 16 + wc;
}

int launch_interp(const LaunchTables &t, const float2 *grid, float2 *samp, const float *weight,
                  hipStream_t stream) {
    if (t.n_interp == 0) return TN_OK;
    const int ndim = t.kind_interp / 16, wc = t.kind_interp % 16;
    const bool staged = t.interp_mode == kInterpStaged;
    InterpFn fn = staged ? (ndim == 2 ? pick_staged<2>(wc) : ndim == 3 ? pick_staged<3>(wc) : nullptr)
                         : (ndim == 2 ? pick_interp<2>(wc) : ndim == 3 ? pick_interp<3>(wc) : nullptr);
    if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d)", ndim, wc);
    hipLaunchKernelGGL(fn, dim3((unsigned)t.n_interp), dim3(kBlock), staged ? t.interp_lds : 0, stream,
                       t.d_jobs, t.d_interp, grid, samp, weight);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int launch_grid(const LaunchTables &t, const float2 *samp, float2 *grid, const float *weight,
                hipStream_t stream) {
    if (t.n_grid == 0) return TN_OK;
    const int ndim = t.kind_grid / 16, wc = t.kind_grid % 16;
    GridFn fn = ndim == 2 ? pick_grid<2>(wc) : ndim == 3 ? pick_grid<3>(wc) : nullptr;
    if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d)", ndim, wc);
    hipLaunchKernelGGL(fn, dim3((unsigned)t.n_grid), dim3(kBlock), t.grid_lds, stream, t.d_jobs,
                       t.d_grid, samp, grid, weight, t.d_scratch, t.d_slot_base);
    TN_HIP(hipGetLastError());
    if (t.n_fold > 0) {
        hipLaunchKernelGGL(fold_kernel, dim3((unsigned)t.n_fold), dim3(kBlock), 0, stream, t.d_jobs,
                           t.d_fold, t.d_scratch, t.d_slot_base, grid);
        TN_HIP(hipGetLastError());
    }
    return TN_OK;
}

int launch_apod_pad(const tn_plan *plan, const float2 *image, float2 *grid, int64_t nimg,
                    hipStream_t stream) {
    const PrePost pp = make_prepost(plan);
    const int64_t total = nimg * (int64_t)pp.G[0] * pp.G[1] * pp.G[2];
    if (total == 0) return TN_OK;
    hipLaunchKernelGGL(apod_pad_kernel, dim3(stride_blocks(total)), dim3(kBlock), 0, stream, pp,
                       image, grid, nimg);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int launch_crop_apod(const tn_plan *plan, const float2 *grid, float2 *image, int64_t nimg,
                     hipStream_t stream) {
    const PrePost pp = make_prepost(plan);
    const int64_t total = nimg * (int64_t)pp.N[0] * pp.N[1] * pp.N[2];
    if (total == 0) return TN_OK;
    hipLaunchKernelGGL(crop_apod_kernel, dim3(stride_blocks(total)), dim3(kBlock), 0, stream, pp,
                       grid, image, nimg);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

}  // namespace tn
