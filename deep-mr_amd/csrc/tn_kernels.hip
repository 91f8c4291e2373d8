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
#include <mutex>

#include "tn_internal.h"
#include "tn_kb.h"

namespace tn {
namespace {

// Incremental decomposition of a linear index into (b, z, y, x) over the
// dims (.., Z, Y, X) for a loop that advances by a fixed stride: the start
// and the stride are divided once, then each step propagates carries. This
// replaces 3-4 integer divisions by runtime values per element (each some
// 30 VALU instructions on CDNA) in the tile, region and image loops.
struct Idx4 {
    int x, y, z;
    int64_t b;
    int sx, sy, sz;
    int64_t sb;
    int X, Y, Z;
    __device__ __forceinline__ Idx4(int64_t start, int64_t stride, int X_, int Y_, int Z_)
        : X(X_), Y(Y_), Z(Z_) {
        split(start, x, y, z, b);
        split(stride, sx, sy, sz, sb);
    }
    __device__ __forceinline__ void split(int64_t i, int &x_, int &y_, int &z_, int64_t &b_) const {
        int64_t t = i / X;
        x_ = (int)(i - t * X);
        int64_t u = t / Y;
        y_ = (int)(t - u * Y);
        b_ = u / Z;
        z_ = (int)(u - b_ * Z);
    }
    __device__ __forceinline__ void next() {  // sx < X, sy < Y, sz < Z: one carry each
        x += sx;
        int c = x >= X;
        x -= c ? X : 0;
        y += sy + c;
        c = y >= Y;
        y -= c ? Y : 0;
        z += sz + c;
        c = z >= Z;
        z -= c ? Z : 0;
        b += sb + c;
    }
};

// Work item of this workgroup, items dealt to the XCDs in contiguous ranges:
// the hardware deals workgroups out round-robin over the XCDs, so without
// the remap neighbouring items (which share grid cache lines, halos and
// samples) would land in eight different L2s. Launches are rounded up to a
// multiple of kXcds; the padding workgroups get -1.
__device__ __forceinline__ int64_t xcd_item(int64_t nitems) {
    const int64_t per = (nitems + kXcds - 1) / kXcds;
    const int64_t idx = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    return idx < nitems ? idx : -1;
}

// Clang vector types for the nontemporal load/store builtins.
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

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
                                                        const Item *__restrict__ items, int64_t nitems,
                                                        const float2 *__restrict__ grid,
                                                        float2 *__restrict__ samp,
                                                        const float *__restrict__ weight) {
    const int64_t item = xcd_item(nitems);
    if (item < 0) return;
    const Item it = items[item];
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

// Low-rank (subspace) interpolation: the grid holds `ncoeff` coefficient
// images per coil, (coil, coeff, *G), and the samples of frame f are
// sum_c coef[f, c] * interp(grid[coil, c]) with coef = basis_adjoint
// (nframes, ncoeff) (src/deepmr/fft/_interp/interp.py:268-357). The plan
// holds every frame's samples as one trajectory (frame = p / frame_npts),
// so the per-frame sample counts do not fragment the launch. The taps are
// evaluated once per sample and reused for every coil and coefficient.
template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void interp_lowrank_kernel(const Job *__restrict__ jobs,
                                                                const Item *__restrict__ items, int64_t nitems,
                                                                const float2 *__restrict__ grid,
                                                                float2 *__restrict__ samp,
                                                                const float *__restrict__ weight,
                                                                LowRank lr) {
    const int64_t item = xcd_item(nitems);
    if (item < 0) return;
    const Item it = items[item];
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
    const float2 *cf = lr.coef + (p / lr.frame_npts) * lr.ncoeff;

    const int Gx = J.G[0], Gy = J.G[1];
    for (int b = it.c0; b < it.c1; ++b) {
        float re = 0.0f, im = 0.0f;
        for (int k = 0; k < lr.ncoeff; ++k) {
            const float2 *g = grid + J.grid_off + ((int64_t)b * lr.ncoeff + k) * J.grid_bstride;
            float tr = 0.0f, ti = 0.0f;
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
                                    tr += v * s.x;
                                    ti += v * s.y;
                                }
                            }
                        }
                    }
                }
            }
            const float2 w = cf[k];
            re += w.x * tr - w.y * ti;
            im += w.x * ti + w.y * tr;
        }
        samp[J.samp_off + (int64_t)b * J.samp_bstride + p] = make_float2(re * sc, im * sc);
    }
}

// Tap sum of one sample over an LDS region (origin r, row pitch Rx, plane
// pitch Rx * Ry); FULL: every axis has all WC taps (no per-tap tests).
template <int NDIM, int WC, bool FULL>
__device__ __forceinline__ float2 staged_taps(const float2 *r, const float (&w)[3][WC], int nx, int ny, int nz,
                                              int Rx, int Ry) {
    float re = 0.0f, im = 0.0f;
    auto plane = [&](int kz, float wz) {
#pragma unroll
        for (int ky = 0; ky < WC; ++ky) {
            if (FULL || ky < ny) {
                const float vzy = wz * w[1][ky];
                const float2 *row = r + (kz * Ry + ky) * Rx;
#pragma unroll
                for (int kx = 0; kx < WC; ++kx) {
                    if (FULL || kx < nx) {
                        const float v = vzy * w[0][kx];
                        const float2 sv = row[kx];
                        re += v * sv.x;
                        im += v * sv.y;
                    }
                }
            }
        }
    };
    if constexpr (NDIM > 2) {
        // one z plane of the footprint at a time: its W^2 LDS reads are in
        // flight together, not all W^3 (which would hold 2 W^3 registers)
#pragma unroll 1
        for (int kz = 0; kz < WC; ++kz)
            if (FULL || kz < nz) plane(kz, w[2][kz]);
    } else {
        plane(0, w[2][0]);
    }
    return make_float2(re, im);
}

// The same for two coils of the staged region (planes r0 and r1): each tap's
// weight is formed once for both, and each coil's sum is the same sequence
// of operations as staged_taps gives it.
template <int NDIM, int WC, bool FULL>
__device__ __forceinline__ void staged_taps2(const float2 *r0, const float2 *r1, const float (&w)[3][WC], int nx,
                                             int ny, int nz, int Rx, int Ry, float2 &t0, float2 &t1) {
    float re0 = 0.0f, im0 = 0.0f, re1 = 0.0f, im1 = 0.0f;
    auto plane = [&](int kz, float wz) {
#pragma unroll
        for (int ky = 0; ky < WC; ++ky) {
            if (FULL || ky < ny) {
                const float vzy = wz * w[1][ky];
                const int off = (kz * Ry + ky) * Rx;
#pragma unroll
                for (int kx = 0; kx < WC; ++kx) {
                    if (FULL || kx < nx) {
                        const float v = vzy * w[0][kx];
                        const float2 a = r0[off + kx], b = r1[off + kx];
                        re0 += v * a.x;
                        im0 += v * a.y;
                        re1 += v * b.x;
                        im1 += v * b.y;
                    }
                }
            }
        }
    };
    if constexpr (NDIM > 2) {
#pragma unroll 1
        for (int kz = 0; kz < WC; ++kz)
            if (FULL || kz < nz) plane(kz, w[2][kz]);
    } else {
        plane(0, w[2][0]);
    }
    t0 = make_float2(re0, im0);
    t1 = make_float2(re1, im1);
}

// Staged interpolation: a workgroup takes one segment (<= kBlock * staged_ept(NDIM)
// samples) of one interpolation tile. Each lane evaluates the taps of its
// samples once and keeps them in registers; then, coil chunk by coil chunk,
// the workgroup loads the (tile + W-1 halo) grid region of the chunk into
// LDS with coalesced row reads and every lane gathers its taps from LDS.
// Samples come from the plan's primary bins (each sample exactly once; in
// 3D ordered by footprint origin inside the tile).
template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void interp_staged_kernel(const Job *__restrict__ jobs,
                                                               const Item *__restrict__ items, int64_t nitems,
                                                               const float2 *__restrict__ grid,
                                                               float2 *__restrict__ samp,
                                                               const float *__restrict__ weight,
                                                               int chunk) {
    extern __shared__ float2 region[];
    const int64_t item = xcd_item(nitems);
    if (item < 0) return;
    const Item it = items[item];
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
    const int nx = J.W[0], ny = NDIM > 1 ? J.W[1] : 1, nz = NDIM > 2 ? J.W[2] : 1;

    constexpr int kStagedEpt = staged_ept(NDIM);
    // taps of this lane's samples, kept across coil chunks
    float w[kStagedEpt][3][WC];
    int loc[kStagedEpt];  // region offset of the footprint origin, -1 = none
    int pnt[kStagedEpt];
    float scl[kStagedEpt];
#pragma unroll
    for (int q = 0; q < kStagedEpt; ++q) {
        const int e = it.e0 + q * kBlock + (int)threadIdx.x;
        loc[q] = -1;
        pnt[q] = 0;
        scl[q] = 0.0f;
        if (e < it.e1) {
            const float4 en = J.pentries[e];
            const float c3[3] = {en.x, en.y, en.z};
            int base[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (d < NDIM) {
                    axis_taps<WC>(c3[d], J.W[d], J.G[d], J.shift[d], J.beta[d], base[d], w[q][d]);
                } else {
#pragma unroll
                    for (int k = 0; k < WC; ++k) w[q][d][k] = k == 0 ? 1.0f : 0.0f;
                }
            }
            // a sample always bins into the tile holding its origin; the
            // clamp only keeps LDS addressing in range
            const int lx0 = min(max(base[0] - ox, 0), Tx - 1);
            const int ly0 = min(max(base[1] - oy, 0), Ty - 1);
            const int lz0 = min(max(base[2] - oz, 0), Tz - 1);
            loc[q] = (lz0 * Ry + ly0) * Rx + lx0;
            pnt[q] = __float_as_int(en.w);
            float sc = J.inv_scale;
            if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + pnt[q]];
            scl[q] = sc;
        }
    }

    // wrapped grid offset of the first cell of each (z, y) row of the region,
    // once per item, in the LDS words after the chunk's region
    int *rowtab = reinterpret_cast<int *>(region + (size_t)chunk * cells);
    const int RZY = Ry * Rz;
    for (int i = threadIdx.x; i < RZY; i += kBlock) {
        const int z = i / Ry, y = i - z * Ry;
        rowtab[i] = (((oz + z) % Gz) * Gy + (oy + y) % Gy) * Gx;
    }
    // lane (xl, rr) loads column xl of region rows rr, rr + step, ... (rows
    // of Rx cells; a row per lane group, no per-cell index arithmetic)
    const bool rowwise = Rx <= kBlock;
    const int step = rowwise ? kBlock / Rx : 1;
    const int rr = rowwise ? (int)threadIdx.x / Rx : 0;
    const int xl = (int)threadIdx.x - rr * Rx;
    const int xoff = (ox + xl) % Gx;
    // every axis at the full compiled width: branch-free tap sums
    const bool full = nx == WC && (NDIM < 2 || ny == WC) && (NDIM < 3 || nz == WC);
    for (int cc = it.c0; cc < it.c1; cc += chunk) {
        const int ncoil = min(chunk, it.c1 - cc);
        __syncthreads();  // previous chunk consumed; the row table written
        const int rows = ncoil * RZY;
        if (rowwise) {
            if (rr < step) {
                // kLoadBatch independent loads per lane in flight before their LDS writes
                int r = rr, yz = rr;
                int64_t coff = J.grid_off + (int64_t)cc * J.grid_bstride;
                while (yz >= RZY) {
                    yz -= RZY;
                    coff += J.grid_bstride;
                }
                while (r < rows) {
                    // the batch's rows first, then their table reads, loads
                    // and LDS writes, each kind issued together
                    int dst[kLoadBatch], row[kLoadBatch];
                    int64_t co[kLoadBatch];
#pragma unroll
                    for (int u = 0; u < kLoadBatch; ++u) {
                        dst[u] = r < rows ? r * Rx + xl : -1;
                        row[u] = yz;
                        co[u] = coff;
                        r += step;
                        yz += step;
                        while (yz >= RZY) {
                            yz -= RZY;
                            coff += J.grid_bstride;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kLoadBatch; ++u) row[u] = rowtab[row[u]];
                    float2 v[kLoadBatch];
#pragma unroll
                    for (int u = 0; u < kLoadBatch; ++u)
                        if (dst[u] >= 0) v[u] = grid[co[u] + row[u] + xoff];
#pragma unroll
                    for (int u = 0; u < kLoadBatch; ++u)
                        if (dst[u] >= 0) region[dst[u]] = v[u];
                }
            }
        } else {  // rows longer than the workgroup: cell by cell
            for (int i = threadIdx.x; i < rows * Rx; i += kBlock) {
                const int r = i / Rx, x = i - r * Rx;
                const int b = r / RZY, yz = r - b * RZY;
                region[i] = grid[J.grid_off + (int64_t)(cc + b) * J.grid_bstride + rowtab[yz] + (ox + x) % Gx];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kStagedEpt; ++q) {
            if (loc[q] < 0) continue;
            int b = 0;
            for (; b + 2 <= ncoil; b += 2) {  // coil pairs share each tap's weight
                const float2 *r = region + (size_t)b * cells + loc[q];
                float2 t0, t1;
                if (full)
                    staged_taps2<NDIM, WC, true>(r, r + cells, w[q], nx, ny, nz, Rx, Ry, t0, t1);
                else
                    staged_taps2<NDIM, WC, false>(r, r + cells, w[q], nx, ny, nz, Rx, Ry, t0, t1);
                samp[J.samp_off + (int64_t)(cc + b) * J.samp_bstride + pnt[q]] =
                    make_float2(t0.x * scl[q], t0.y * scl[q]);
                samp[J.samp_off + (int64_t)(cc + b + 1) * J.samp_bstride + pnt[q]] =
                    make_float2(t1.x * scl[q], t1.y * scl[q]);
            }
            if (b < ncoil) {
                const float2 *r = region + (size_t)b * cells + loc[q];
                const float2 t = full ? staged_taps<NDIM, WC, true>(r, w[q], nx, ny, nz, Rx, Ry)
                                      : staged_taps<NDIM, WC, false>(r, w[q], nx, ny, nz, Rx, Ry);
                samp[J.samp_off + (int64_t)(cc + b) * J.samp_bstride + pnt[q]] =
                    make_float2(t.x * scl[q], t.y * scl[q]);
            }
        }
    }
}

// Writes a finished LDS tile accumulator ([coil][re, im][rows][P]) of a
// gridding item: straight into the grid (every in-grid cell, zeros
// included), or, for a split tile, as partial `part` of fold slot `slot`
// ([coil][cell](re, im)), summed later by fold_kernel in part order.
template <bool INTERLEAVED = false>
__device__ __forceinline__ void store_tile(const Job &J, const Item &it, const float *acc, int plane,
                                           int P, int ox, int oy, int oz, int ex, int ey, int ez,
                                           float2 *__restrict__ grid, float *__restrict__ scratch,
                                           const int64_t *__restrict__ slot_base) {
    // planar: acc[coil][re, im][cell], `plane` floats per part; interleaved:
    // acc[coil][cell](re, im), `plane` cells per coil
    auto at = [&](int b, int cell_off) -> float2 {
        if (INTERLEAVED) return ((const float2 *)acc)[b * plane + cell_off];
        const float *a = acc + 2 * b * plane + cell_off;
        return make_float2(a[0], a[plane]);
    };
    const int Tx = J.T[0], Ty = J.T[1], Tz = J.T[2];
    const int Gx = J.G[0], Gy = J.G[1];
    const int ncoil = it.c1 - it.c0;
    const int cells = Tx * Ty * Tz;
    const int tid = threadIdx.x;
    Idx4 q(tid, kBlock, Tx, Ty, Tz);
    if (it.slot >= 0) {
        float *dst = scratch + slot_base[it.slot] + (int64_t)it.part * 2 * ncoil * cells;
        for (int i = tid; i < ncoil * cells; i += kBlock, q.next()) {
            const float2 v = at((int)q.b, (q.z * Ty + q.y) * P + q.x);
            dst[2 * i] = v.x;
            dst[2 * i + 1] = v.y;
        }
        return;
    }
    for (int i = tid; i < ncoil * cells; i += kBlock, q.next()) {
        if (q.x < ex && q.y < ey && q.z < ez) {
            grid[J.grid_off + (int64_t)(it.c0 + q.b) * J.grid_bstride +
                 ((int64_t)(oz + q.z) * Gy + (oy + q.y)) * Gx + (ox + q.x)] =
                at((int)q.b, (q.z * Ty + q.y) * P + q.x);
        }
    }
}

// Scaled sample value of accumulator coil `vc` (0-based within the job) for
// sample p. Low rank: accumulator coils are (coil, coeff) pairs and the
// sample of frame f enters coefficient k times basis[k, f].
__device__ __forceinline__ float2 grid_sample(const Job &J, const float2 *__restrict__ samp, int vc,
                                              int p, float sc, const LowRank &lr) {
    float2 v;
    if (lr.ncoeff > 0) {
        const int b = vc / lr.ncoeff, k = vc - b * lr.ncoeff;
        const float2 s = samp[J.samp_off + (int64_t)b * J.samp_bstride + p];
        const float2 w = lr.coef[(int64_t)k * lr.nframes + p / lr.frame_npts];
        v = make_float2(w.x * s.x - w.y * s.y, w.x * s.y + w.y * s.x);
    } else {
        v = samp[J.samp_off + (int64_t)vc * J.samp_bstride + p];
    }
    return make_float2(v.x * sc, v.y * sc);
}

// Gridding. One workgroup owns one tile for a set of coils and walks the
// tile's entries in batches of kGridBatch:
//   A. one lane per entry evaluates the taps (per axis: wrapped origin + W
//      weights) and all lanes gather the scaled sample values of every coil;
//      both are staged in LDS once for the four waves;
//   B. every wave spreads the whole batch for the coils it owns, one entry
//      per lane: the plan orders each tile's entries round-robin over
//      footprint origins, so the 64 lanes of an LDS atomic add mostly hit
//      distinct cells even on dense trajectory centres. Accumulators are
//      planar with a padded row pitch.
// Each accumulator cell is updated by a single wave (waves own whole coils,
// or row bands when there are fewer than four coils), in the fixed entry
// order, so the result is deterministic. The tile is written once (or
// stored as a partial of a split tile, folded by fold_kernel).
template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void grid_kernel(const Job *__restrict__ jobs,
                                                      const Item *__restrict__ items, int64_t nitems,
                                                      const float2 *__restrict__ samp,
                                                      float2 *__restrict__ grid,
                                                      const float *__restrict__ weight,
                                                      float *__restrict__ scratch,
                                                      const int64_t *__restrict__ slot_base,
                                                      LowRank lr) {
    extern __shared__ float smem[];
    constexpr int B = kGridBatch;
    const int64_t item = xcd_item(nitems);
    if (item < 0) return;
    const Item it = items[item];
    const Job &J = jobs[it.job];
    const int tile = it.unit;
    const int Tx = J.T[0], Ty = J.T[1], Tz = J.T[2];
    const int tx = tile % J.NT[0];
    const int ty = (tile / J.NT[0]) % J.NT[1];
    const int tz = tile / (J.NT[0] * J.NT[1]);
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
    const int ex = min(Tx, J.G[0] - ox), ey = min(Ty, J.G[1] - oy), ez = min(Tz, J.G[2] - oz);
    const int P = Tx + kGridPad;
    const int rowsP = Tz * Ty * P;
    const int plane = rowsP + kWave;  // + one trash cell per lane
    const int ncoil = it.c1 - it.c0;
    const int Gx = J.G[0], Gy = J.G[1], Gz = J.G[2];

    float *acc = smem;                                     // [ncoil][re, im][Tz*Ty*P + trash]
    int *sbase = (int *)(acc + 2 * ncoil * plane);         // [3][B]
    int *sidx = sbase + 3 * B;                             // [B]
    float *sw = (float *)(sidx + B);                       // [3][WC][B]
    float2 *sval = (float2 *)(sw + 3 * WC * B);            // [ncoil][B]

    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = tid % kWave;
    const int nx = J.W[0];
    const int ny = NDIM > 1 ? J.W[1] : 1;
    const int nz = NDIM > 2 ? J.W[2] : 1;

    // wave ownership: whole coils (>= 4 coils) or (coil, row band)
    const int rows = Tz * Ty;
    int coil0, cstep, row_lo = 0, row_hi = rows;
    if (ncoil >= kBlock / kWave) {
        coil0 = wave;
        cstep = kBlock / kWave;
    } else {
        const int nb = (kBlock / kWave) / ncoil;
        const int band = wave / ncoil;
        coil0 = band < nb ? wave % ncoil : ncoil;
        cstep = ncoil;  // one coil per wave
        const int rpb = (rows + nb - 1) / nb;
        row_lo = min(rows, band * rpb);
        row_hi = min(rows, row_lo + rpb);
    }

    for (int i = tid; i < 2 * ncoil * plane; i += kBlock) acc[i] = 0.0f;

    for (int b0 = it.e0; b0 < it.e1; b0 += B) {
        const int nb_ = min(B, it.e1 - b0);
        __syncthreads();  // previous batch consumed (and accumulators zeroed)
        // A1. taps of the batch, one entry per lane
        if (tid < nb_) {
            const float4 en = J.entries[b0 + tid];
            const float c3[3] = {en.x, en.y, en.z};
            sidx[tid] = __float_as_int(en.w);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                float w[WC];
                int base = 0;
                if (d < NDIM) {
                    axis_taps<WC>(c3[d], J.W[d], J.G[d], J.shift[d], J.beta[d], base, w);
                } else {
#pragma unroll
                    for (int k = 0; k < WC; ++k) w[k] = k == 0 ? 1.0f : 0.0f;
                }
                sbase[d * B + tid] = base;
#pragma unroll
                for (int k = 0; k < WC; ++k) sw[(d * WC + k) * B + tid] = w[k];
            }
        }
        __syncthreads();
        // A2. scaled sample values of every coil of the item. Low-rank:
        // accumulator "coils" are (coil, coeff) pairs and the sample of
        // frame f enters coefficient k times basis[k, f].
        for (int i = tid; i < nb_ * ncoil; i += kBlock) {
            const int c = i / nb_;
            const int j = i - c * nb_;
            const int p = sidx[j];
            float sc = J.inv_scale;
            if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];
            sval[c * B + j] = grid_sample(J, samp, it.c0 + c, p, sc, lr);
        }
        __syncthreads();
        // B. spread, one entry per lane. The tap loop is branch-free: taps
        // outside the tile (or outside this wave's row band) get weight 0
        // and a per-lane trash address, so one wait for the sample value
        // is followed by back-to-back LDS atomics.
        if (coil0 < ncoil) {
            for (int r = 0; r < nb_; r += kWave) {
                const int j = min(r + lane, nb_ - 1);
                const bool live = r + lane < nb_;
                const int bx = sbase[j], by = sbase[B + j], bz = sbase[2 * B + j];
                int lx[WC], lrow[WC * WC];
                float wx[WC], wzy[WC * WC];
#pragma unroll
                for (int k = 0; k < WC; ++k) {
                    const int ax = wrap_add(bx, k, Gx) - ox;
                    const bool okx = live && k < nx && ax >= 0 && ax < ex;
                    lx[k] = okx ? ax : -1;
                    wx[k] = okx ? sw[k * B + j] : 0.0f;
                }
#pragma unroll
                for (int kz = 0; kz < WC; ++kz) {
                    const int az = wrap_add(bz, kz, Gz) - oz;
                    const bool okz = kz < nz && az >= 0 && az < ez;
                    const float wz = sw[(2 * WC + kz) * B + j];
#pragma unroll
                    for (int ky = 0; ky < WC; ++ky) {
                        const int ay = wrap_add(by, ky, Gy) - oy;
                        const int row = az * Ty + ay;
                        const bool ok = okz && ky < ny && ay >= 0 && ay < ey && row >= row_lo && row < row_hi;
                        lrow[kz * WC + ky] = ok ? row * P : -1;
                        wzy[kz * WC + ky] = ok ? wz * sw[(WC + ky) * B + j] : 0.0f;
                    }
                }
                const int trash = rowsP + lane;
                for (int c = coil0; c < ncoil; c += cstep) {
                    const float2 sv = sval[c * B + j];
                    float *are = acc + 2 * c * plane;
#pragma unroll
                    for (int q = 0; q < WC * WC; ++q) {
                        if (NDIM < 3 && q >= WC) break;
#pragma unroll
                        for (int kx = 0; kx < WC; ++kx) {
                            const bool ok = lrow[q] >= 0 && lx[kx] >= 0;
                            const int a = ok ? lrow[q] + lx[kx] : trash;
                            const float v = wzy[q] * wx[kx];
                            __hip_atomic_fetch_add(are + a, v * sv.x, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(are + plane + a, v * sv.y, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    store_tile(J, it, acc, plane, P, ox, oy, oz, ex, ey, ez, grid, scratch, slot_base);
}

// Gridding by register runs. The plan sorts each tile's entries by
// footprint origin, so the W^2 tap sums of one origin and footprint z plane
// -- the complete contribution of all samples sharing that origin -- can be
// formed in registers and added into the LDS tile once.
//
// Lanes: a wave owns Cw coils of the item (Cw a power of two) and its 64
// lanes are 64/Cw groups of Cw lanes, one lane per coil. The item's entries
// are taken in batches of kRunsBatch:
//   A. the threads stage the batch: each entry record is loaded, its taps
//      are evaluated once and stored in SoA LDS rows (origin key, per axis
//      the tile-local offset of tap 0 and the tap at which the footprint
//      wraps around the grid, the weights), and the scaled sample values of
//      every coil of the item are gathered (independent loads, so their
//      latencies overlap);
//   B. group g walks batch entries g*R .. g*R+R-1 (R = kRunsBatch * Cw / 64;
//      slots are laid out so that consecutive groups read consecutive LDS
//      words and a group's lanes read the same word) and each lane sums the
//      taps of its coil per origin, one z plane of the footprint per pass.
//      An origin that starts and ends inside the group's run is unique to
//      the group and is added at once. The run's first and last origins may
//      continue in the neighbouring groups: the first-origin sums move to the
//      previous group when it ends on the same origin, and a segmented scan
//      over the groups (keyed by origin, which is sorted) gathers each
//      last-origin sum in the last group of its segment, which adds it. Both
//      steps are skipped (wave ballot) when no origin spans two groups.
// Within one add no two lanes address the same cell of the same coil (lanes
// with nothing to add use per-lane trash cells), so the tile is updated with
// plain LDS read-add-write instead of atomics, and dense trajectory centres
// -- thousands of samples on a few origins -- cost one add per run. Each
// accumulator cell is updated by one wave in program order (waves own coils,
// or row bands when there are fewer coils than waves): the result is
// deterministic.
template <int NDIM, int WC>
__global__ __launch_bounds__(kBlock) void grid_runs_kernel(const Job *__restrict__ jobs,
                                                           const Item *__restrict__ items, int64_t nitems,
                                                           const float2 *__restrict__ samp,
                                                           float2 *__restrict__ grid,
                                                           const float *__restrict__ weight,
                                                           float *__restrict__ scratch,
                                                           const int64_t *__restrict__ slot_base,
                                                           LowRank lr) {
    constexpr int TAPS = WC * WC;  // taps of one z plane: sums kept in registers per pass
    constexpr int B = kRunsBatch;
    constexpr int NREC = runs_record_words(NDIM, WC);  // key, (d0, kw) per axis, weights
    static_assert(B == kBlock, "one staged entry per thread");
    extern __shared__ float smem[];
    const int64_t item = xcd_item(nitems);
    if (item < 0) return;
    const Item it = items[item];
    const Job &J = jobs[it.job];
    const int tile = it.unit;
    const int Tx = J.T[0], Ty = J.T[1], Tz = J.T[2];
    const int tx = tile % J.NT[0];
    const int ty = (tile / J.NT[0]) % J.NT[1];
    const int tz = tile / (J.NT[0] * J.NT[1]);
    const int o3[3] = {tx * Tx, ty * Ty, tz * Tz};
    const int e3[3] = {min(Tx, J.G[0] - o3[0]), min(Ty, J.G[1] - o3[1]), min(Tz, J.G[2] - o3[2])};
    const int P = Tx + kGridPad;
    const int rowsP = Tz * Ty * P;
    const int plane = rowsP + kWave;  // + one trash cell per lane
    const int ncoil = it.c1 - it.c0;
    const int SV = ncoil + 1;         // sample-value row pitch (float2), padded against bank conflicts
    const int Gx = J.G[0], Gy = J.G[1];
    float2 *acc = (float2 *)smem;                                        // [ncoil][plane](re, im)
    int *rec = (int *)(smem + ((2 * ncoil * plane + 3) & ~3));           // [NREC][B]
    float2 *sval = (float2 *)(rec + NREC * B);                           // [B][SV]

    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = tid % kWave;
    const int rows = Tz * Ty;
    // lane -> (group, coil); waves own coils (cw per wave) or row bands
    int lcw = 0;  // log2(cw)
    while ((kBlock / kWave) << lcw < ncoil && lcw < 4) ++lcw;
    const int cw = 1 << lcw;
    const int ng = kWave >> lcw;             // groups per wave
    const int grp = lane >> lcw;
    const int R = B / ng;                    // entries per group per batch
    int coil, row_lo = 0, row_hi = rows;
    if (ncoil >= kBlock / kWave) {
        coil = wave * cw + (lane & (cw - 1));  // >= ncoil: idle lane
    } else {
        const int nb = (kBlock / kWave) / ncoil;
        const int band = wave / ncoil;
        coil = band < nb ? wave % ncoil : ncoil;
        const int rpb = (rows + nb - 1) / nb;
        row_lo = min(rows, band * rpb);
        row_hi = min(rows, row_lo + rpb);
    }
    const bool active = coil < ncoil;
    const int cs = active ? coil : 0;  // a valid coil index for idle lanes' reads
    for (int i = tid; i < ncoil * plane; i += kBlock) acc[i] = make_float2(0.0f, 0.0f);
    const int trash = rowsP + lane;

    // adds the tap sums (sr, si) of an origin (tile-local tap-0 offsets d0,
    // wrap taps kw), z plane kz, into this lane's coil tile; lanes with
    // do_add false add zeros to their trash cell. The lanes of one call add
    // distinct origins or distinct coils, so no two lanes touch the same cell
    // and a plain read-add-write is race-free; the compiler barrier keeps
    // each tap's read behind the previous tap's write (a wave's LDS
    // operations complete in order), as another lane's next tap may land on
    // a cell this tap just wrote.
    auto add_origin = [&](bool do_add, const int (&d0)[NDIM], const int (&kw)[NDIM], int kz,
                          const float (&sr)[TAPS], const float (&si)[TAPS]) {
        int lx[WC], lrow[WC];
        do_add = do_add && active;
#pragma unroll
        for (int k = 0; k < WC; ++k) {
            const int a = d0[0] + k - (k >= kw[0] ? Gx : 0);
            lx[k] = (do_add && k < J.W[0] && a >= 0 && a < e3[0]) ? a : -1;
        }
        int az = 0;
        bool okz = true;
        if (NDIM > 2) {
            constexpr int ZD = NDIM > 2 ? 2 : 0;
            az = d0[ZD] + kz - (kz >= kw[ZD] ? J.G[2] : 0);
            okz = kz < J.W[2] && az >= 0 && az < e3[2];
        }
#pragma unroll
        for (int ky = 0; ky < WC; ++ky) {
            const int ay = d0[1] + ky - (ky >= kw[1] ? Gy : 0);
            const int row = az * Ty + ay;
            const bool ok = okz && ky < J.W[1] && ay >= 0 && ay < e3[1] && row >= row_lo && row < row_hi;
            lrow[ky] = ok ? row * P : -1;
        }
        float2 *are = acc + cs * plane;
#pragma unroll
        for (int ky = 0; ky < WC; ++ky) {
#pragma unroll
            for (int kx = 0; kx < WC; ++kx) {
                const bool ok = lrow[ky] >= 0 && lx[kx] >= 0;
                const int a = ok ? lrow[ky] + lx[kx] : trash;
                float2 v = are[a];
                v.x += ok ? sr[ky * WC + kx] : 0.0f;
                v.y += ok ? si[ky * WC + kx] : 0.0f;
                are[a] = v;
                __asm__ __volatile__("" ::: "memory");
            }
        }
    };

    const int nzp = NDIM > 2 ? J.W[2] : 1;  // z planes of the taps, one pass each
    for (int b0 = it.e0; b0 < it.e1; b0 += B) {
        __syncthreads();  // previous batch consumed (and accumulators zeroed)
        {   // A. slot t holds the entry that group t % ng walks at step t / ng
            const int slot = tid;
            const int e = b0 + (tid % ng) * R + tid / ng;
            if (e < it.e1) {
                const float4 en = J.entries[e];
                const float c3[3] = {en.x, en.y, en.z};
                int key = 0;
#pragma unroll
                for (int d = NDIM - 1; d >= 0; --d) {
                    int base;
                    float w[WC];
                    axis_taps<WC>(c3[d], J.W[d], J.G[d], J.shift[d], J.beta[d], base, w);
                    key = key * J.G[d] + base;
                    rec[(1 + 2 * d) * B + slot] = base - o3[d];
                    rec[(2 + 2 * d) * B + slot] = J.G[d] - base;
#pragma unroll
                    for (int k = 0; k < WC; ++k) rec[(1 + 2 * NDIM + d * WC + k) * B + slot] = __float_as_int(w[k]);
                }
                rec[slot] = key;
                const int p = __float_as_int(en.w);
                float sc = J.inv_scale;
                if (weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];
                for (int c = 0; c < ncoil; ++c) sval[slot * SV + c] = grid_sample(J, samp, it.c0 + c, p, sc, lr);
            }
        }
        __syncthreads();
        if (ncoil >= kBlock / kWave ? wave * cw >= ncoil : coil >= ncoil) continue;  // wave without work
        const int nvalid = min(R, max(0, it.e1 - (b0 + grp * R)));  // entries in this group's run
        for (int kz = 0; kz < nzp; ++kz) {
            // B. walk the run (z plane kz of the taps): H = sums of the first
            //    origin, C = current origin
            float hr[TAPS], hi[TAPS], cr[TAPS], ci[TAPS];
#pragma unroll
            for (int t = 0; t < TAPS; ++t) hr[t] = hi[t] = cr[t] = ci[t] = 0.0f;
            int hd0[NDIM], hkw[NDIM], qd0[NDIM], qkw[NDIM];
#pragma unroll
            for (int d = 0; d < NDIM; ++d) hd0[d] = hkw[d] = qd0[d] = qkw[d] = 0;
            int ka = -1, kq = -1;     // first / current origin key (-1: none yet)
            bool first_open = true;   // the current origin is still the first one
            for (int i = 0; i < nvalid; ++i) {
                const int s = i * ng + grp;
                const int key = rec[s];
                if (kq < 0) {
                    ka = kq = key;
#pragma unroll
                    for (int d = 0; d < NDIM; ++d) {
                        hd0[d] = qd0[d] = rec[(1 + 2 * d) * B + s];
                        hkw[d] = qkw[d] = rec[(2 + 2 * d) * B + s];
                    }
                } else if (key != kq) {
                    if (first_open) {  // the first origin ends inside the run: keep it in H
#pragma unroll
                        for (int t = 0; t < TAPS; ++t) {
                            hr[t] = cr[t];
                            hi[t] = ci[t];
                        }
                        first_open = false;
                    } else {  // an interior origin: unique to this group
                        add_origin(true, qd0, qkw, kz, cr, ci);
                    }
#pragma unroll
                    for (int t = 0; t < TAPS; ++t) cr[t] = ci[t] = 0.0f;
                    kq = key;
#pragma unroll
                    for (int d = 0; d < NDIM; ++d) {
                        qd0[d] = rec[(1 + 2 * d) * B + s];
                        qkw[d] = rec[(2 + 2 * d) * B + s];
                    }
                }
                const float2 sv = sval[s * SV + cs];
                float wx[WC], wy[WC];
#pragma unroll
                for (int t = 0; t < WC; ++t) {
                    wx[t] = __int_as_float(rec[(1 + 2 * NDIM + t) * B + s]);
                    wy[t] = __int_as_float(rec[(1 + 2 * NDIM + WC + t) * B + s]);
                }
                const float wz = NDIM > 2 ? __int_as_float(rec[(1 + 2 * NDIM + 2 * WC + kz) * B + s]) : 1.0f;
#pragma unroll
                for (int ky = 0; ky < WC; ++ky) {
                    const float wzy = wz * wy[ky];
                    const float fr = wzy * sv.x, fi = wzy * sv.y;
#pragma unroll
                    for (int kx = 0; kx < WC; ++kx) {
                        cr[ky * WC + kx] += wx[kx] * fr;
                        ci[ky * WC + kx] += wx[kx] * fi;
                    }
                }
            }
            // kb: key of the last origin (-1: empty run). H holds the first
            // origin's sums only if the run has more than one origin.
            const int kb = kq;
            const bool has_h = kq >= 0 && !first_open;
            const int kb_prev = __shfl_up(kb, cw, kWave);
            const bool h_to_prev = has_h && grp > 0 && kb_prev == ka;
            // 1. the first-origin sums of group g+1 join group g when g ends on that origin
            if (__ballot(h_to_prev) != 0) {
                const int ka_next = __shfl_down(ka, cw, kWave);
                const bool hn_next = __shfl_down(has_h ? 1 : 0, cw, kWave) != 0;
                const bool take_next = grp < ng - 1 && kb >= 0 && hn_next && ka_next == kb;
#pragma unroll
                for (int t = 0; t < TAPS; ++t) {
                    const float nr = __shfl_down(hr[t], cw, kWave);
                    const float ni = __shfl_down(hi[t], cw, kWave);
                    if (take_next) {
                        cr[t] += nr;
                        ci[t] += ni;
                    }
                }
            }
            if (__ballot(has_h && !h_to_prev) != 0) add_origin(has_h && !h_to_prev, hd0, hkw, kz, hr, hi);
            // 2. segmented inclusive scan of the last-origin sums over the groups
            if (__ballot(kb >= 0 && grp > 0 && kb_prev == kb) != 0) {
                for (int d = 1; d < ng; d <<= 1) {
                    const int kp = __shfl_up(kb, d * cw, kWave);
                    const bool join = grp >= d && kp == kb && kb >= 0;
#pragma unroll
                    for (int t = 0; t < TAPS; ++t) {
                        const float pr = __shfl_up(cr[t], d * cw, kWave);
                        const float pi = __shfl_up(ci[t], d * cw, kWave);
                        if (join) {
                            cr[t] += pr;
                            ci[t] += pi;
                        }
                    }
                }
            }
            const int kb_next = __shfl_down(kb, cw, kWave);
            const bool seg_end = kb >= 0 && (grp == ng - 1 || kb_next != kb);
            add_origin(seg_end, qd0, qkw, kz, cr, ci);
        }
    }
    __syncthreads();
    store_tile<true>(J, it, (const float *)acc, plane, P, o3[0], o3[1], o3[2], e3[0], e3[1], e3[2], grid,
                     scratch, slot_base);
}

// Gridding by the plan's sparse gridding matrix: per frame, every (cell,
// sample, weight) triple of the kernel footprints, sorted by output cell
// (CSR), samples ascending inside a cell. Two kinds of work item, both
// summing weight * scaled sample for up to CM coils in registers and writing
// each cell of every coil once (zeros included) with streaming stores:
//   part 0: a run of (it.e1 & 0xffff) <= kBlock consecutive cells from
//           it.unit, one lane per cell, walking its row with the next entry
//           prefetched; rows longer than J.mlong are skipped;
//   part 1: entries [it.e1 * kMatrixPart, + kMatrixPart) of one such long
//           row (dense trajectory centres hold hundreds to 10^5 entries per
//           cell), strided over kLongGroups lanes, each summing its
//           entries in order, the lane sums then added in lane order; a row
//           of several parts (it.slot >= 0) leaves partial record
//           it.slot + it.e1 for matrix_fold_kernel, else the cell is
//           written.
// No atomics, and a fixed summation order for every cell. Long rows are
// queued first; items are dealt to the XCDs in contiguous ranges (the
// hardware deals workgroups out round-robin), so the neighbouring rows of
// one frame, which read mostly the same samples, share one XCD's L2.
// Low rank (LR): accumulator coils are (coil, coeff) pairs; the sample of
// frame f enters coefficient k times basis[k, f].
template <int CM, bool LR>
__global__ __launch_bounds__(kBlock) void grid_matrix_kernel(const Job *__restrict__ jobs,
                                                             const Item *__restrict__ items, int64_t nitems,
                                                             const float2 *__restrict__ samp,
                                                             float2 *__restrict__ grid,
                                                             const float *__restrict__ weight, LowRank lr,
                                                             float2 *__restrict__ parts) {
    __shared__ float red[kLongGroups][2 * CM];
    const int64_t idx = xcd_item(nitems);
    if (idx < 0) return;
    const Item it = items[idx];
    const Job &J = jobs[it.job];
    const int ncoil = it.c1 - it.c0;
    const bool dcf = weight != nullptr && J.w_off >= 0;
    // per-coil sample rows (uniform); coils past the item's last repeat it
    // (loads stay in bounds and unconditional, their sums are not stored)
    const char *cb[CM];
    const float2 *qb[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) {
        const int vc = it.c0 + min(c, ncoil - 1);
        const int b = LR ? vc / lr.ncoeff : vc;
        cb[c] = (const char *)(samp + J.samp_off + (int64_t)b * J.samp_bstride);
        qb[c] = LR ? lr.coef + (int64_t)(vc - b * lr.ncoeff) * lr.nframes : nullptr;
    }
    float ar[CM], ai[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) ar[c] = ai[c] = 0.0f;
    auto add_entry = [&](i32x2 m) {
        const uint32_t boff = (uint32_t)m.x * (uint32_t)sizeof(float2);  // npts < 2^28
        float2 v[CM];
#pragma unroll
        for (int c = 0; c < CM; ++c) v[c] = *(const float2 *)(cb[c] + boff);
        if (!LR) {
            // the sample scaled first, then weighted: the same operations as
            // the coil-minor copy and its kernel, so both give equal sums
            float sc = J.inv_scale;
            if (dcf) sc *= weight[J.w_off + m.x];
            const float w = __int_as_float(m.y);
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                ar[c] += w * (v[c].x * sc);
                ai[c] += w * (v[c].y * sc);
            }
        } else {
            float w = __int_as_float(m.y) * J.inv_scale;
            if (dcf) w *= weight[J.w_off + m.x];
            const int f = m.x / lr.frame_npts;
#pragma unroll
            for (int c = 0; c < CM; ++c) {
                const float2 q = qb[c][f];
                ar[c] += w * (q.x * v[c].x - q.y * v[c].y);
                ai[c] += w * (q.x * v[c].y + q.y * v[c].x);
            }
        }
    };
    const i32x2 *mnz = (const i32x2 *)J.mnz;
    if (it.part == 0) {  // cells it.unit .. it.unit + (it.e1 & 0xffff) - 1 (<= kBlock)
        const int64_t cell = (int64_t)it.unit + threadIdx.x;
        if ((int)threadIdx.x >= (it.e1 & 0xffff)) return;
        const int r0 = J.mrow[cell], r1 = J.mrow[cell + 1];
        if (r1 - r0 > J.mlong) return;  // written by its long-row item
        if (r0 < r1) {
            i32x2 m = __builtin_nontemporal_load(mnz + r0);
            for (int e = r0; e < r1; ++e) {
                const i32x2 mn = __builtin_nontemporal_load(mnz + min(e + 1, r1 - 1));
                add_entry(m);
                m = mn;
            }
        }
        float2 *g = grid + J.grid_off + cell;
#pragma unroll
        for (int c = 0; c < CM; ++c) {
            if (c < ncoil) {
                const f32x2 v = {ar[c], ai[c]};
                __builtin_nontemporal_store(v, (f32x2 *)(g + (int64_t)(it.c0 + c) * J.grid_bstride));
            }
        }
        return;
    }
    // a long row: lane k < kLongGroups sums the entries r0 + k, r0 + k +
    // kLongGroups, ... in order, then the lane sums are added in lane order
    // (the partition and order of the coil-minor kernel's long rows)
    const int64_t cell = it.unit;
    const int r0 = J.mrow[cell] + it.e1 * kMatrixPart, r1 = min(J.mrow[cell + 1], r0 + kMatrixPart);
    if ((int)threadIdx.x < kLongGroups) {
        int e = r0 + (int)threadIdx.x;
        for (; e + 3 * kLongGroups < r1; e += 4 * kLongGroups) {  // four entries in flight
            i32x2 m[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) m[u] = __builtin_nontemporal_load(mnz + e + u * kLongGroups);
#pragma unroll
            for (int u = 0; u < 4; ++u) add_entry(m[u]);
        }
        for (; e < r1; e += kLongGroups) add_entry(__builtin_nontemporal_load(mnz + e));
#pragma unroll
        for (int c = 0; c < CM; ++c) {
            red[threadIdx.x][2 * c] = ar[c];
            red[threadIdx.x][2 * c + 1] = ai[c];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < ncoil) {
        const int c = threadIdx.x;
        float re = red[0][2 * c], im = red[0][2 * c + 1];
#pragma unroll 1
        for (int k = 1; k < kLongGroups; ++k) {
            re += red[k][2 * c];
            im += red[k][2 * c + 1];
        }
        if (it.slot < 0)
            grid[J.grid_off + (int64_t)(it.c0 + c) * J.grid_bstride + cell] = make_float2(re, im);
        else
            parts[((int64_t)it.slot + it.e1) * kMatrixCoils + c] = make_float2(re, im);
    }
}

// Coil-minor copy of the samples for the matrix gridding: chunk q of a job
// (coils c0 .. c1 - 1 of the item) becomes st[q][p][CW] = samp[c0 + j][p]
// * inv_scale * weight[p], zeros for j >= c1 - c0. One item = kBlock
// samples of one chunk, transposed through LDS: coalesced reads along the
// samples of each coil, contiguous writes.
template <int LCW>
__global__ __launch_bounds__(kBlock) void samp_transpose_kernel(const Job *__restrict__ jobs,
                                                                const Item *__restrict__ items,
                                                                const float2 *__restrict__ samp,
                                                                const float *__restrict__ weight,
                                                                float2 *__restrict__ st) {
    constexpr int CW = 1 << LCW;
    __shared__ float2 tile[kBlock * (CW + 1)];
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)it.unit * kBlock;
    const int64_t p = p0 + t;
    const int nc = it.c1 - it.c0;
    float sc = J.inv_scale;
    if (p < J.npts && weight != nullptr && J.w_off >= 0) sc *= weight[J.w_off + p];
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        float2 v = make_float2(0.0f, 0.0f);
        if (j < nc && p < J.npts) {
            const float2 a = samp[J.samp_off + (int64_t)(it.c0 + j) * J.samp_bstride + p];
            v = make_float2(a.x * sc, a.y * sc);
        }
        tile[t * (CW + 1) + j] = v;
    }
    __syncthreads();
    float2 *dst = st + J.st_off + ((int64_t)it.e0 * J.npts + p0) * CW;
    const int np = (int)min<int64_t>(kBlock, J.npts - p0);
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const int flat = k * kBlock + t;
        const int pp = flat >> LCW, j = flat & (CW - 1);
        if (pp < np) dst[flat] = tile[pp * (CW + 1) + j];
    }
}

// Matrix gridding over the coil-minor sample copy: CW lanes per cell, each
// lane summing CPL adjacent coils, so that the cell's lanes read one
// contiguous RC = CW * CPL coil record per entry (one cache line for 16
// coils instead of one line per coil; with CPL = 2 each lane reads 16 bytes).
// Items as for grid_matrix_kernel (it.e0 = chunk index). Part 0, a run of
// (it.e1 & 0xffff) cells from it.unit whose entries (it.e1 >> 16 of them,
// at most STAGE, contiguous in the CSR order from it.slot) and row
// offsets are first staged in LDS with independent coalesced loads, so that
// the only dependent global loads left are the sample records; kBlock / CW
// cells per pass, each row U entries at a time. Part 1 strides one long
// row's entries over kLongGroups cell groups and adds the group sums in
// group order. One coil per lane fits 64 VGPRs, so eight waves per SIMD are
// asked for (kooshball 3D, 8 coils: 3.48 -> 3.35 ms); two coils per lane
// would spill at that bound.
template <int LCW, int CPL, int U, int STAGE>
__global__ __launch_bounds__(kBlock, CPL == 1 ? 8 : 1) void grid_matrix_t_kernel(const Job *__restrict__ jobs,
                                                               const Item *__restrict__ items, int64_t nitems,
                                                               const float2 *__restrict__ st,
                                                               float2 *__restrict__ grid,
                                                               float2 *__restrict__ parts) {
    static_assert(CPL == 1 || CPL == 2, "one or two coils per lane");
    constexpr int CW = 1 << LCW;      // lanes per cell
    constexpr int RC = CW * CPL;      // coils per record
    constexpr int CPP = kBlock / CW;  // cells per pass
    // cells per item (the host builds items of at most this many cells for
    // this record width); the item's sums are staged per coil plane in LDS
    // and stored as contiguous runs (row pitch CPI + 1 against bank conflicts)
    constexpr int CPI = CPP * kMatrixPasses < kBlock ? CPP * kMatrixPasses : kBlock;
    __shared__ float2 sout[RC * (CPI + 1)];
    __shared__ i32x2 sent[STAGE];
    __shared__ int srow[kBlock + 1];
    const int64_t idx = xcd_item(nitems);
    if (idx < 0) return;
    const Item it = items[idx];
    const Job &J = jobs[it.job];
    const int j = threadIdx.x & (CW - 1), g = threadIdx.x >> LCW;
    // this lane's coils j * CPL .. j * CPL + CPL - 1 of the record
    const float2 *sq = st + J.st_off + (int64_t)it.e0 * J.npts * RC + j * CPL;
    const i32x2 *mnz = (const i32x2 *)J.mnz;
    struct Acc {
        float re[CPL], im[CPL];
    };
    // entries e0, e0 + stride, ... below r1 of one row (from `ent`), U at a time
    auto row_sum = [&](const i32x2 *ent, bool global, int e0, int r1, int stride, Acc &a) {
        for (int e = e0; e < r1; e += U * stride) {
            i32x2 m[U];
            float2 v[U][CPL];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int ei = min(e + u * stride, r1 - 1);
                m[u] = global ? __builtin_nontemporal_load(ent + ei) : ent[ei];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (CPL == 2) {
                    const float4 r = *(const float4 *)(sq + (int64_t)m[u].x * RC);
                    v[u][0] = make_float2(r.x, r.y);
                    v[u][1] = make_float2(r.z, r.w);
                } else {
                    v[u][0] = sq[(int64_t)m[u].x * RC];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float w = e + u * stride < r1 ? __int_as_float(m[u].y) : 0.0f;
#pragma unroll
                for (int k = 0; k < CPL; ++k) {
                    a.re[k] += w * v[u][k].x;
                    a.im[k] += w * v[u][k].y;
                }
            }
        }
    };
    if (it.part == 0) {
        const int64_t cell0 = it.unit;
        const int ncells = min(it.e1 & 0xffff, CPI);
        const int rb = it.slot;
        const int nent = min(it.e1 >> 16, STAGE);  // <= STAGE by construction (matrix_stage)
        for (int i = threadIdx.x; i <= ncells; i += kBlock) srow[i] = min(max(J.mrow[cell0 + i] - rb, 0), nent);
        for (int i0 = threadIdx.x; i0 < nent; i0 += kLoadBatch * kBlock) {  // loads first, then the LDS writes
            i32x2 m[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (i0 + u * kBlock < nent) m[u] = __builtin_nontemporal_load(mnz + rb + i0 + u * kBlock);
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (i0 + u * kBlock < nent) sent[i0 + u * kBlock] = m[u];
        }
        __syncthreads();
        for (int c = g; c < ncells; c += CPP) {
            Acc a;
#pragma unroll
            for (int k = 0; k < CPL; ++k) a.re[k] = a.im[k] = 0.0f;
            row_sum(sent, false, srow[c], srow[c + 1], 1, a);
#pragma unroll
            for (int k = 0; k < CPL; ++k) sout[(j * CPL + k) * (CPI + 1) + c] = make_float2(a.re[k], a.im[k]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < RC * CPI; i += kBlock) {
            const int jj = i / CPI, c = i - jj * CPI;
            if (c < ncells && it.c0 + jj < it.c1) {
                const float2 a = sout[jj * (CPI + 1) + c];
                const f32x2 v = {a.x, a.y};
                __builtin_nontemporal_store(
                    v, (f32x2 *)(grid + J.grid_off + (int64_t)(it.c0 + jj) * J.grid_bstride + cell0 + c));
            }
        }
        return;
    }
    // a long row: kLongGroups lane groups, group k summing the entries
    // r0 + k, r0 + k + kLongGroups, ... in order, then the group sums added in
    // group order; the same partition and order for every record width, so a
    // cell's value does not depend on which other coils share the table
    static_assert(CPP >= kLongGroups, "a long row needs kLongGroups lane groups");
    static_assert(RC * (CPI + 1) >= kLongGroups * RC, "group sums reuse the output stage");
    const int64_t cell = it.unit;
    const int r0 = J.mrow[cell] + it.e1 * kMatrixPart, r1 = min(J.mrow[cell + 1], r0 + kMatrixPart);
    Acc a;
#pragma unroll
    for (int k = 0; k < CPL; ++k) a.re[k] = a.im[k] = 0.0f;
    if (g < kLongGroups) {
        row_sum(mnz, true, r0 + g, r1, kLongGroups, a);
#pragma unroll
        for (int k = 0; k < CPL; ++k) sout[g * RC + j * CPL + k] = make_float2(a.re[k], a.im[k]);
    }
    __syncthreads();
    const int c = threadIdx.x;  // coil of the record
    if (c < RC && it.c0 + c < it.c1) {
        float2 s = sout[c];
#pragma unroll 1
        for (int k = 1; k < kLongGroups; ++k) {  // one read at a time: a full unroll holds 62 registers
            s.x += sout[k * RC + c].x;
            s.y += sout[k * RC + c].y;
        }
        if (it.slot < 0)
            grid[J.grid_off + (int64_t)(it.c0 + c) * J.grid_bstride + cell] = s;
        else
            parts[((int64_t)it.slot + it.e1) * kMatrixCoils + c] = s;
    }
}

// Cells of long rows summed in parts: cell = part 0 + part 1 + ... in part
// order (item: unit = cell, e1 = number of parts, slot = first record).
__global__ __launch_bounds__(kWave) void matrix_fold_kernel(const Job *__restrict__ jobs,
                                                            const Item *__restrict__ items,
                                                            const float2 *__restrict__ parts,
                                                            float2 *__restrict__ grid) {
    const Item it = items[blockIdx.x];
    const Job &J = jobs[it.job];
    const int c = threadIdx.x;
    if (c >= it.c1 - it.c0) return;
    float2 a = parts[(int64_t)it.slot * kMatrixCoils + c];
    for (int k = 1; k < it.e1; ++k) {
        const float2 p = parts[((int64_t)it.slot + k) * kMatrixCoils + c];
        a.x += p.x;
        a.y += p.y;
    }
    grid[J.grid_off + (int64_t)(it.c0 + c) * J.grid_bstride + it.unit] = a;
}

// Fold of split gridding tiles: out = part 0 + part 1 + ... in that order.
// Item: unit = tile, coils c0 .. c1 - 1, e0 = slot, e1 = parts, elements
// [part, slot) of the partial tile ((coil, z, y, x) order), lane by lane.
constexpr int kFoldBatch = 16;
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
    const int64_t ps = (int64_t)ncoil * cells;  // float2 elements per part
    const int Gx = J.G[0], Gy = J.G[1];
    const int start = it.part + (int)threadIdx.x;
    Idx4 c(start, kBlock, Tx, Ty, Tz);
    for (int i = start; i < it.slot; i += kBlock, c.next()) {
        if (c.x < ex && c.y < ey && c.z < ez) {
            const float2 *s2 = (const float2 *)(src + 2 * (int64_t)i);
            float2 a = s2[0];
            // up to kFoldBatch parts' loads in flight, added in part order
            for (int q = 1; q < nparts; q += kFoldBatch) {
                const int nb = min(kFoldBatch, nparts - q);
                float2 v[kFoldBatch];
#pragma unroll
                for (int u = 0; u < kFoldBatch; ++u)
                    if (u < nb) v[u] = s2[(int64_t)(q + u) * ps];
#pragma unroll
                for (int u = 0; u < kFoldBatch; ++u) {
                    if (u < nb) {
                        a.x += v[u].x;
                        a.y += v[u].y;
                    }
                }
            }
            grid[J.grid_off + (int64_t)(it.c0 + c.b) * J.grid_bstride +
                 ((int64_t)(oz + c.z) * Gy + (oy + c.y)) * Gx + (ox + c.x)] = a;
        }
    }
}

// Register-tile gridding (3D), the sums of one item's entries: lane (x, y)
// of the tile face (origin x0, y0, z0) adds, entry by entry, its cells' terms
// in the tile's 4 z planes for CW coils. Pointers are global-memory words
// (loads at wave-uniform addresses become scalar loads) and restrict: the
// kernel writes nothing they address. The loads of the next two entries are
// issued before the current entry's sums, then all are waited for.
using gfloat = const __attribute__((address_space(1))) float *;
template <int CW>
__device__ __forceinline__ void tiles_sum(gfloat __restrict__ tf, gfloat __restrict__ ef, gfloat __restrict__ sf,
                                          int e0, int e1, int x, int y, int x0, int y0, int z0, const int (&G)[3],
                                          const int (&W)[3], float (&ar)[kTilesT[2]][CW],
                                          float (&ai)[kTilesT[2]][CW]) {
    constexpr int TX = kTilesT[0], TY = kTilesT[1], TZ = kTilesT[2];
    // tile-local footprint origin along one axis: the origin is wrapped into
    // [0, G); a footprint that wraps onto the tile at 0 starts at origin - G
    // (grids hold at least two tiles per axis)
    auto local = [](int o, int t0, int T, int g) {
        const int l = o - t0;
        return l >= T ? l - g : l;
    };
    struct Stage {
        int smp;       // sample
        float z[4];    // z taps
        int o[3];      // footprint origin
        float wx, wy;  // this lane's x and y taps
        float v[2 * CW];
    };
    // taps table: 16 words per sample (x, y, z taps, origin); entries: 4
    // words, the sample's bits last; copy: CW complex per sample
    auto fetch_head = [&](int smp, Stage &g) {  // scalar loads: origin, z taps
        const int64_t b16 = 16 * (int64_t)smp;
        g.smp = smp;
#pragma unroll
        for (int k = 0; k < 4; ++k) g.z[k] = tf[b16 + 8 + k];
#pragma unroll
        for (int d = 0; d < 3; ++d) g.o[d] = __float_as_int(tf[b16 + 12 + d]);
    };
    auto fetch_body = [&](Stage &g) {  // the record (scalar), the lane's x and y taps (vector)
        const int64_t b16 = 16 * (int64_t)g.smp;
        const int lx = local(g.o[0], x0, TX, G[0]), ly = local(g.o[1], y0, TY, G[1]);
        g.wx = tf[b16 + ((x - lx) & 3)];
        g.wy = tf[b16 + 4 + ((y - ly) & 3)];
#pragma unroll
        for (int c = 0; c < 2 * CW; ++c) g.v[c] = sf[(int64_t)g.smp * 2 * CW + c];
    };
    auto sample_of = [&](int e) { return __float_as_int(ef[4 * (int64_t)e + 3]); };
    auto accumulate = [&](const Stage &g) {
        const int lx = local(g.o[0], x0, TX, G[0]), ly = local(g.o[1], y0, TY, G[1]);
        const int lz = local(g.o[2], z0, TZ, G[2]);
        if ((unsigned)(x - lx) < (unsigned)W[0] && (unsigned)(y - ly) < (unsigned)W[1]) {
#pragma unroll
            for (int p = 0; p < TZ; ++p) {
                const int k = p - lz;
                if ((unsigned)k < (unsigned)W[2]) {
                    const float wz = k == 0 ? g.z[0] : k == 1 ? g.z[1] : k == 2 ? g.z[2] : g.z[3];
                    const float w = (wz * g.wy) * g.wx;
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        ar[p][c] += w * g.v[2 * c];
                        ai[p][c] += w * g.v[2 * c + 1];
                    }
                }
            }
        }
    };
    // Three entries in flight: e complete (a), e + 1 with its origin and z
    // taps (b: its record and lane taps are loaded in this step), e + 2
    // (c: origin and z taps loaded in this step); entries past the end
    // reload the last one.
    const int elast = e1 - 1;
    Stage A, B, C;
    fetch_head(sample_of(e0), A);
    __builtin_amdgcn_s_waitcnt(0);
    fetch_body(A);
    fetch_head(sample_of(min(e0 + 1, elast)), B);
    int i2 = sample_of(min(e0 + 2, elast));
    __builtin_amdgcn_s_waitcnt(0);
    int e = e0;
    auto step = [&](const Stage &a, Stage &b, Stage &c) {
        fetch_body(b);
        fetch_head(i2, c);
        i2 = sample_of(min(e + 3, elast));
        accumulate(a);
        __builtin_amdgcn_s_waitcnt(0);
        return ++e < e1;
    };
    while (step(A, B, C) && step(B, C, A) && step(C, A, B)) {
    }
}

// Register-tile gridding (3D): one wave per item, four items per workgroup.
// Item: unit = tile of kTilesT cells, entries [e0, e1) of its bin (the
// samples whose footprints meet the tile, ascending), coils c0 .. c1 - 1 of
// chunk c0 / st_chunk of the coil-minor copy (CW coils per record, scaled
// and weighted), part / slot as for the LDS kernels. Lane (x, y) of the
// tile's 8x8 face sums its cells in the tile's 4 z planes in registers with
// the weight (wz wy) wx the other variants use: a cell's sum runs over its
// samples in ascending order, as the matrix rows do. The wave-uniform data
// of an entry (sample record, z taps, footprint origin) come through scalar
// loads, each lane's x and y taps through one vector load per axis.
template <int CW>
__global__ __launch_bounds__(kBlock) void grid_tiles_kernel(const Job *__restrict__ jobs,
                                                            const Item *__restrict__ items, int64_t nitems,
                                                            const float2 *__restrict__ st,
                                                            float2 *__restrict__ grid, float *__restrict__ scratch,
                                                            const int64_t *__restrict__ slot_base) {
    constexpr int TX = kTilesT[0], TY = kTilesT[1], TZ = kTilesT[2];
    static_assert(TX * TY == kWave, "one lane per (x, y) of the tile face");
    // items in list order, four per workgroup: the hardware deals workgroups
    // to the XCDs in turn, which spreads the hot tiles' parts (adjacent in
    // the list) over all of them; contiguous per-XCD ranges measured 2x
    // slower on a kooshball (its centre's parts landed on one XCD)
    const int64_t idx = (int64_t)blockIdx.x * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (idx >= nitems) return;
    const Item it = items[idx];
    const Job &J = jobs[it.job];
    const int lane = threadIdx.x % kWave, x = lane % TX, y = lane / TX;
    const int tile = it.unit;
    const int x0 = (tile % J.NT[0]) * TX, y0 = ((tile / J.NT[0]) % J.NT[1]) * TY, z0 = tile / (J.NT[0] * J.NT[1]) * TZ;
    float ar[TZ][CW], ai[TZ][CW];
#pragma unroll
    for (int p = 0; p < TZ; ++p)
#pragma unroll
        for (int c = 0; c < CW; ++c) ar[p][c] = ai[p][c] = 0.0f;
    if (it.e0 < it.e1) {
        const int G[3] = {J.G[0], J.G[1], J.G[2]}, W[3] = {J.W[0], J.W[1], J.W[2]};
        tiles_sum<CW>((gfloat)J.taps, (gfloat)J.entries,
                      (gfloat)(st + J.st_off + (int64_t)(it.c0 / J.st_chunk) * J.npts * CW), it.e0, it.e1, x, y,
                      x0, y0, z0, G, W, ar, ai);
    }
    const int nc = it.c1 - it.c0;
    if (it.slot < 0) {
        const int64_t Gx = J.G[0], Gy = J.G[1];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            if (c >= nc) break;
            float2 *g = grid + J.grid_off + (int64_t)(it.c0 + c) * J.grid_bstride + (int64_t)(y0 + y) * Gx + x0 + x;
#pragma unroll
            for (int p = 0; p < TZ; ++p) g[(int64_t)(z0 + p) * Gy * Gx] = make_float2(ar[p][c], ai[p][c]);
        }
    } else {  // partial of a split tile, in fold_kernel's (coil, z, y, x) order
        float *dst = scratch + slot_base[it.slot] + (int64_t)it.part * 2 * nc * (TX * TY * TZ);
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            if (c >= nc) break;
#pragma unroll
            for (int p = 0; p < TZ; ++p) {
                const int i = (c * TZ + p) * (TX * TY) + lane;
                *(float2 *)(dst + 2 * i) = make_float2(ar[p][c], ai[p][c]);
            }
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
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t step = (int64_t)gridDim.x * kBlock;
    Idx4 q(i0, step, pp.G[0], pp.G[1], pp.G[2]);
    for (int64_t i = i0; i < total; i += step, q.next()) {
        int mx = q.x + pp.N[0] / 2; if (mx >= pp.G[0]) mx -= pp.G[0];
        int my = q.y + pp.N[1] / 2; if (my >= pp.G[1]) my -= pp.G[1];
        int mz = q.z + pp.N[2] / 2; if (mz >= pp.G[2]) mz -= pp.G[2];
        float2 v = make_float2(0.0f, 0.0f);
        if (mx < pp.N[0] && my < pp.N[1] && mz < pp.N[2]) {
            v = img[q.b * nvol + ((int64_t)mz * pp.N[1] + my) * pp.N[0] + mx];
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
    const int64_t i0 = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const int64_t step = (int64_t)gridDim.x * kBlock;
    Idx4 q(i0, step, pp.N[0], pp.N[1], pp.N[2]);
    for (int64_t i = i0; i < total; i += step, q.next()) {
        const int mx = q.x, my = q.y, mz = q.z;
        int x = mx - pp.N[0] / 2; if (x < 0) x += pp.G[0];
        int y = my - pp.N[1] / 2; if (y < 0) y += pp.G[1];
        int z = mz - pp.N[2] / 2; if (z < 0) z += pp.G[2];
        float2 v = grid[q.b * gvol + ((int64_t)z * pp.G[1] + y) * pp.G[0] + x];
        img[i] = apodize(v, pp, mx, my, mz);
    }
}

unsigned xcd_blocks(int64_t n) { return (unsigned)((n + kXcds - 1) / kXcds * kXcds); }

unsigned stride_blocks(int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 256 * 64));
}

// ---- dispatch ----------------------------------------------------------------
using InterpFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, const float *);
using StagedFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, const float *, int);
using GridFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, const float *, float *,
                        const int64_t *, LowRank);
using InterpLrFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, const float *, LowRank);

template <int NDIM>
InterpLrFn pick_interp_lowrank(int wc) {
    switch (wc) {
        case 1: return interp_lowrank_kernel<NDIM, 1>;
        case 2: return interp_lowrank_kernel<NDIM, 2>;
        case 3: return interp_lowrank_kernel<NDIM, 3>;
        case 4: return interp_lowrank_kernel<NDIM, 4>;
        case 5: return interp_lowrank_kernel<NDIM, 5>;
        case 6: return interp_lowrank_kernel<NDIM, 6>;
        case 7: return interp_lowrank_kernel<NDIM, 7>;
        case 8: return interp_lowrank_kernel<NDIM, 8>;
        case 9: case 10: case 11: case 12: return interp_lowrank_kernel<NDIM, 12>;
        case 13: case 14: case 15: case 16: return interp_lowrank_kernel<NDIM, 16>;
        default: return nullptr;
    }
}

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
        case 9: case 10: case 11: case 12: return interp_kernel<NDIM, 12>;
        case 13: case 14: case 15: case 16: return interp_kernel<NDIM, 16>;
        default: return nullptr;
    }
}

template <int NDIM>
StagedFn pick_staged(int wc) {
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
GridFn pick_runs(int wc) {  // W^2 <= kRunsMaxTaps
    switch (wc) {
        case 1: return grid_runs_kernel<NDIM, 1>;
        case 2: return grid_runs_kernel<NDIM, 2>;
        case 3: return grid_runs_kernel<NDIM, 3>;
        case 4: return grid_runs_kernel<NDIM, 4>;
        case 5: return grid_runs_kernel<NDIM, 5>;
        case 6: return grid_runs_kernel<NDIM, 6>;
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

// kind = ndim * kKindStride + max width
int interp_kind(const std::vector<Job> &jobs) {
    int ndim = 0, wc = 1;
    for (const Job &j : jobs) {
        ndim = std::max(ndim, j.ndim);
        for (int d = 0; d < 3; ++d) wc = std::max(wc, j.W[d]);
    }
    return ndim * kKindStride + wc;
}

int launch_interp(const LaunchTables &t, const float2 *grid, float2 *samp, const float *weight,
                  hipStream_t stream) {
    if (t.n_interp == 0) return TN_OK;
    const int ndim = kind_ndim(t.kind_interp), wc = kind_width(t.kind_interp);
    if (t.interp_mode == kInterpStaged) {
        StagedFn fn = ndim == 2 ? pick_staged<2>(wc) : ndim == 3 ? pick_staged<3>(wc) : nullptr;
        if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d)", ndim, wc);
        hipLaunchKernelGGL(fn, dim3(xcd_blocks(t.n_interp)), dim3(kBlock), t.interp_lds, stream, t.d_jobs,
                           t.d_interp, t.n_interp, grid, samp, weight, t.interp_coils);
    } else {
        InterpFn fn = ndim == 2 ? pick_interp<2>(wc) : ndim == 3 ? pick_interp<3>(wc) : nullptr;
        if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d)", ndim, wc);
        hipLaunchKernelGGL(fn, dim3(xcd_blocks(t.n_interp)), dim3(kBlock), 0, stream, t.d_jobs,
                           t.d_interp, t.n_interp, grid, samp, weight);
    }
    TN_HIP(hipGetLastError());
    return TN_OK;
}

using MatrixFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, const float *, LowRank,
                          float2 *);

int launch_matrix_fold(const LaunchTables &t, float2 *grid, float *scratch, hipStream_t stream) {
    if (t.n_fold == 0) return TN_OK;
    hipLaunchKernelGGL(matrix_fold_kernel, dim3((unsigned)t.n_fold), dim3(kWave), 0, stream, t.d_jobs, t.d_fold,
                       (const float2 *)scratch, grid);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

template <bool LR>
MatrixFn pick_matrix(int coils) {
    if (coils <= 1) return grid_matrix_kernel<1, LR>;
    if (coils <= 2) return grid_matrix_kernel<2, LR>;
    if (coils <= 4) return grid_matrix_kernel<4, LR>;
    if (coils <= 8) return grid_matrix_kernel<8, LR>;
    return grid_matrix_kernel<kMatrixCoils, LR>;
}

int launch_grid(const LaunchTables &t, const float2 *samp, float2 *grid, const float *weight,
                hipStream_t stream, LowRank lr, float *scratch) {
    if (t.n_grid == 0) return TN_OK;
    if (scratch == nullptr) scratch = t.d_scratch;
    if (t.grid_mode == kGridMatrix && lr.ncoeff == 0 && t.st_cw > 0) {
        using TrFn = void (*)(const Job *, const Item *, const float2 *, const float *, float2 *);
        using MtFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, float2 *);
        TrFn tr = nullptr;
        MtFn mt = nullptr;
        const bool pair = t.st_cpl == 2;  // two coils per lane (the table's cells per item assume it)
        // entries in flight per lane: 4 in 2D, 8 in 3D; staged entries per
        // item: matrix_stage (each measured fastest)
        const bool d3 = kind_ndim(t.kind_grid) == 3;
        switch (t.st_cw) {
            case 2:
                tr = samp_transpose_kernel<1>;
                mt = d3 ? grid_matrix_t_kernel<1, 1, 8, kMatrixStage> : grid_matrix_t_kernel<1, 1, 4, kMatrixStage2d>;
                break;
            case 4:
                tr = samp_transpose_kernel<2>;
                mt = pair ? (d3 ? grid_matrix_t_kernel<1, 2, 8, kMatrixStage> : grid_matrix_t_kernel<1, 2, 4, kMatrixStage2d>)
                          : (d3 ? grid_matrix_t_kernel<2, 1, 8, kMatrixStage> : grid_matrix_t_kernel<2, 1, 4, kMatrixStage2d>);
                break;
            case 8:
                tr = samp_transpose_kernel<3>;
                mt = pair ? (d3 ? grid_matrix_t_kernel<2, 2, 8, kMatrixStage> : grid_matrix_t_kernel<2, 2, 4, kMatrixStage2d>)
                          : (d3 ? grid_matrix_t_kernel<3, 1, 8, kMatrixStage> : grid_matrix_t_kernel<3, 1, 4, kMatrixStage2d>);
                break;
            case 16:
                tr = samp_transpose_kernel<4>;  // 8 lanes x 2 coils
                mt = d3 ? grid_matrix_t_kernel<3, 2, 8, kMatrixStage> : grid_matrix_t_kernel<3, 2, 4, kMatrixStage2d>;
                break;
            default: return fail(TN_E_ARG, "matrix gridding: %d coils per record", t.st_cw);
        }
        if (t.st_cw == 16 ? !pair : (pair && t.st_cw < 4))
            return fail(TN_E_ARG, "matrix gridding: %d coils per lane for %d-coil records", t.st_cpl, t.st_cw);
        if (t.n_tr > 0) {
            hipLaunchKernelGGL(tr, dim3((unsigned)t.n_tr), dim3(kBlock), 0, stream, t.d_jobs, t.d_tr, samp, weight,
                               t.d_st);
            TN_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(mt, dim3(xcd_blocks(t.n_grid)), dim3(kBlock), 0, stream, t.d_jobs, t.d_grid, t.n_grid,
                           (const float2 *)t.d_st, grid, (float2 *)scratch);
        TN_HIP(hipGetLastError());
        return launch_matrix_fold(t, grid, scratch, stream);
    }
    if (t.grid_mode == kGridTiles && lr.ncoeff == 0 && t.st_cw > 0) {
        using TrFn = void (*)(const Job *, const Item *, const float2 *, const float *, float2 *);
        using TlFn = void (*)(const Job *, const Item *, int64_t, const float2 *, float2 *, float *, const int64_t *);
        TrFn tr = nullptr;
        TlFn tl = nullptr;
        switch (t.st_cw) {
            case 2: tr = samp_transpose_kernel<1>; tl = grid_tiles_kernel<2>; break;
            case 4: tr = samp_transpose_kernel<2>; tl = grid_tiles_kernel<4>; break;
            case 8: tr = samp_transpose_kernel<3>; tl = grid_tiles_kernel<8>; break;
            default: return fail(TN_E_ARG, "tile gridding: %d coils per record", t.st_cw);
        }
        if (kind_ndim(t.kind_grid) != 3 || kind_width(t.kind_grid) > kTilesMaxWidth)
            return fail(TN_E_ARG, "tile gridding: 3D tables with at most %d taps per axis", kTilesMaxWidth);
        if (t.n_tr > 0) {
            hipLaunchKernelGGL(tr, dim3((unsigned)t.n_tr), dim3(kBlock), 0, stream, t.d_jobs, t.d_tr, samp, weight,
                               t.d_st);
            TN_HIP(hipGetLastError());
        }
        const int64_t groups = (t.n_grid + kBlock / kWave - 1) / (kBlock / kWave);
        hipLaunchKernelGGL(tl, dim3((unsigned)groups), dim3(kBlock), 0, stream, t.d_jobs, t.d_grid, t.n_grid,
                           (const float2 *)t.d_st, grid, scratch, t.d_slot_base);
        TN_HIP(hipGetLastError());
        if (t.n_fold > 0) {
            hipLaunchKernelGGL(fold_kernel, dim3((unsigned)t.n_fold), dim3(kBlock), 0, stream, t.d_jobs,
                               t.d_fold, scratch, t.d_slot_base, grid);
            TN_HIP(hipGetLastError());
        }
        return TN_OK;
    }
    if (t.grid_mode == kGridMatrix) {
        if (t.grid_coils > kMatrixCoils) return fail(TN_E_ARG, "matrix gridding: %d coils per item", t.grid_coils);
        MatrixFn fn = lr.ncoeff > 0 ? pick_matrix<true>(t.grid_coils) : pick_matrix<false>(t.grid_coils);
        hipLaunchKernelGGL(fn, dim3(xcd_blocks(t.n_grid)), dim3(kBlock), 0, stream, t.d_jobs, t.d_grid, t.n_grid, samp,
                           grid, weight, lr, (float2 *)scratch);
        TN_HIP(hipGetLastError());
        return launch_matrix_fold(t, grid, scratch, stream);
    }
    const int ndim = kind_ndim(t.kind_grid), wc = kind_width(t.kind_grid);
    GridFn fn = nullptr;
    if (t.grid_mode == kGridRuns)
        fn = ndim == 2 ? pick_runs<2>(wc) : ndim == 3 ? pick_runs<3>(wc) : nullptr;
    else
        fn = ndim == 2 ? pick_grid<2>(wc) : ndim == 3 ? pick_grid<3>(wc) : nullptr;
    if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d, mode=%d)", ndim, wc, t.grid_mode);
    hipLaunchKernelGGL(fn, dim3(xcd_blocks(t.n_grid)), dim3(kBlock), t.grid_lds, stream, t.d_jobs,
                       t.d_grid, t.n_grid, samp, grid, weight, scratch, t.d_slot_base, lr);
    TN_HIP(hipGetLastError());
    if (t.n_fold > 0) {
        hipLaunchKernelGGL(fold_kernel, dim3((unsigned)t.n_fold), dim3(kBlock), 0, stream, t.d_jobs,
                           t.d_fold, scratch, t.d_slot_base, grid);
        TN_HIP(hipGetLastError());
    }
    return TN_OK;
}

// Fixed device workspace of the allocation-free gridding entry point: the
// partial records of long rows summed in parts (and the partial tiles of
// split tiles). One copy per device, held until the process ends; host code
// uses the per-device address from hipGetSymbolAddress, never the array's
// name.
static __device__ __align__(256) float scratch[kStaticScratchFloats];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int static_scratch(int device, float **out) {
    if (device < 0 || device >= kMaxDevices) return fail(TN_E_RANGE, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
    }
    *out = (float *)g_scratch_ptr[device];
    return TN_OK;
}

int launch_interp_lowrank(const LaunchTables &t, const float2 *grid, float2 *samp, const float *weight,
                          hipStream_t stream, LowRank lr) {
    if (t.n_interp == 0) return TN_OK;
    if (t.interp_mode != kInterpGather) return fail(TN_E_ARG, "low-rank interpolation needs gather tables");
    const int ndim = kind_ndim(t.kind_interp), wc = kind_width(t.kind_interp);
    InterpLrFn fn = ndim == 2 ? pick_interp_lowrank<2>(wc) : ndim == 3 ? pick_interp_lowrank<3>(wc) : nullptr;
    if (!fn) return fail(TN_E_ARG, "unsupported kernel (ndim=%d, width=%d)", ndim, wc);
    hipLaunchKernelGGL(fn, dim3(xcd_blocks(t.n_interp)), dim3(kBlock), 0, stream, t.d_jobs, t.d_interp,
                       t.n_interp, grid, samp, weight, lr);
    TN_HIP(hipGetLastError());
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
