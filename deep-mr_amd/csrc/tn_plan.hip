// Plan construction on the device: tile binning of every sample.
//
// Each sample is entered into every gridding tile its kernel footprint
// touches (owner-computes: a tile later accumulates only the taps that fall
// inside it, so every tile is written exactly once without global atomics).
// Entries are keyed by (frame, tile) and sorted with a stable radix sort, so
// inside a tile the samples keep ascending sample order: the gridding
// accumulation order is fixed and the adjoint is deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "tn_internal.h"
#include "tn_kb.h"

namespace tn {
namespace {

struct AxisGeom {
    int G, W, shift, T;
};

struct BinGeom {
    float beta[3];       // Kaiser-Bessel beta (x, y, z)
    int ndim;
    int64_t npts;
    int nframes;
    AxisGeom ax[3];      // (x, y, z)
    int NT[3];           // tiles per axis (x, y, z)
    int64_t ntiles;      // product of NT
};

// Distinct tiles touched along one axis by a footprint starting at `base`.
// A footprint that wraps around the grid edge can come back into a tile it
// already touched (when the last tile is narrower than W - 1), so the check
// is against every tile collected so far, not only the previous one.
__device__ __forceinline__ int axis_tiles(const AxisGeom &a, int base, int (&tl)[kMaxWidth]) {
    int n = 0;
    for (int k = 0; k < a.W; ++k) {
        const int t = wrap_add(base, k, a.G) / a.T;
        bool seen = false;
        for (int j = 0; j < n; ++j) seen |= (tl[j] == t);
        if (!seen) tl[n++] = t;
    }
    return n;
}

__device__ __forceinline__ float beta_of(const BinGeom &g, int d) { return g.beta[d]; }

__device__ __forceinline__ int axis_base(const AxisGeom &a, float g) {
    if (a.W == 1) return wrap_mod((int)g + a.shift, a.G);
    const float x0 = ceilf(g - 0.5f * (float)a.W);
    return wrap_mod((int)x0 + a.shift, a.G);
}

__global__ void bin_count_kernel(BinGeom geo, const float *__restrict__ coord,
                                 int32_t *__restrict__ counts) {
    const int64_t total = geo.npts * geo.nframes;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int c = 1;
        for (int d = 0; d < geo.ndim; ++d) {
            int tl[kMaxWidth];
            const int base = axis_base(geo.ax[d], coord[i * geo.ndim + d]);
            c *= axis_tiles(geo.ax[d], base, tl);
        }
        counts[i] = c;
    }
}

// Primary bins: every sample once, keyed by the tile holding its footprint
// origin (the staged interpolation loads that tile plus a W-1 halo).
__global__ void bin_primary_kernel(BinGeom geo, const float *__restrict__ coord,
                                   uint32_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int64_t total = geo.npts * geo.nframes;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frame = i / geo.npts;
        int t[3] = {0, 0, 0};
        for (int d = 0; d < geo.ndim; ++d)
            t[d] = axis_base(geo.ax[d], coord[i * geo.ndim + d]) / geo.ax[d].T;
        const int64_t tile = ((int64_t)t[2] * geo.NT[1] + t[1]) * geo.NT[0] + t[0];
        keys[i] = (uint32_t)(frame * geo.ntiles + tile);
        vals[i] = (int32_t)(i - frame * geo.npts);
    }
}

__global__ void bin_emit_kernel(BinGeom geo, const float *__restrict__ coord,
                                const int32_t *__restrict__ offsets,
                                uint32_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int64_t total = geo.npts * geo.nframes;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frame = i / geo.npts;
        const int32_t pt = (int32_t)(i - frame * geo.npts);
        int tl[3][kMaxWidth];
        int nt[3] = {1, 1, 1};
        for (int d = 0; d < 3; ++d) tl[d][0] = 0;
        for (int d = 0; d < geo.ndim; ++d) {
            const int base = axis_base(geo.ax[d], coord[i * geo.ndim + d]);
            nt[d] = axis_tiles(geo.ax[d], base, tl[d]);
        }
        int32_t pos = offsets[i];
        for (int a = 0; a < nt[2]; ++a)
            for (int b = 0; b < nt[1]; ++b)
                for (int c = 0; c < nt[0]; ++c) {
                    const int64_t tile =
                        ((int64_t)tl[2][a] * geo.NT[1] + tl[1][b]) * geo.NT[0] + tl[0][c];
                    keys[pos] = (uint32_t)(frame * geo.ntiles + tile);
                    vals[pos] = pt;
                    ++pos;
                }
    }
}

// bin_off[k] = first entry whose key >= k, for k in [0, nbins]; e == n
// stands for a virtual key nbins.
__global__ void bin_offsets_kernel(const uint32_t *__restrict__ keys, int64_t n, int64_t nbins,
                                   int32_t *__restrict__ bin_off) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e < n ? (int64_t)keys[e] : nbins;
        const int64_t kp = e > 0 ? (int64_t)keys[e - 1] : -1;
        for (int64_t b = kp + 1; b <= k; ++b) bin_off[b] = (int32_t)e;
    }
}

// Gridding matrix entries of every sample, in (sample, tap) order, tap x
// fastest: key = frame * cells + output cell, value = (sample within its
// frame, weight bits) packed low/high in 64 bits. The weight is the product
// of the per-axis Kaiser-Bessel taps (z * y first, then x), the same taps
// the other gridding variants evaluate.
__global__ void matrix_emit_kernel(BinGeom geo, const float *__restrict__ coord, int taps,
                                   uint32_t *__restrict__ keys, uint64_t *__restrict__ vals) {
    const int64_t total = geo.npts * geo.nframes;
    const int64_t ncell = (int64_t)geo.ax[0].G * geo.ax[1].G * geo.ax[2].G;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frame = i / geo.npts;
        const int64_t p = i - frame * geo.npts;
        int base[3] = {0, 0, 0};
        float w[3][kMaxWidth];
        for (int d = 0; d < 3; ++d) {
            if (d < geo.ndim) {
                axis_taps<kMaxWidth>(coord[i * geo.ndim + d], geo.ax[d].W, geo.ax[d].G, geo.ax[d].shift, beta_of(geo, d),
                             base[d], w[d]);
            } else {
                w[d][0] = 1.0f;
            }
        }
        const int Wx = geo.ax[0].W, Wy = geo.ax[1].W;
        const int64_t e0 = i * taps;
        for (int t = 0; t < taps; ++t) {
            const int kx = t % Wx, ky = (t / Wx) % Wy, kz = t / (Wx * Wy);
            const int64_t cell = ((int64_t)wrap_add(base[2], kz, geo.ax[2].G) * geo.ax[1].G +
                                  wrap_add(base[1], ky, geo.ax[1].G)) * geo.ax[0].G +
                                 wrap_add(base[0], kx, geo.ax[0].G);
            const float wt = (w[2][kz] * w[1][ky]) * w[0][kx];
            keys[e0 + t] = (uint32_t)(frame * ncell + cell);
            vals[e0 + t] = ((uint64_t)(uint32_t)__float_as_int(wt) << 32) | (uint64_t)(uint32_t)p;
        }
    }
}

// Per-sample taps of the register-tile gridding (3D, at most
// kTilesMaxWidth taps per axis): float4 x taps, y taps, z taps (zeros past
// the axis width) and the footprint origin (x, y, z) as int bits, the same
// taps and wrapped origin the other gridding variants evaluate.
__global__ void tiles_taps_kernel(BinGeom geo, const float *__restrict__ coord, float4 *__restrict__ taps) {
    const int64_t total = geo.npts * geo.nframes;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int base[3];
        float w[3][kTilesMaxWidth];
        for (int d = 0; d < 3; ++d)
            axis_taps<kTilesMaxWidth>(coord[i * 3 + d], geo.ax[d].W, geo.ax[d].G, geo.ax[d].shift, beta_of(geo, d),
                                      base[d], w[d]);
        for (int d = 0; d < 3; ++d) taps[4 * i + d] = make_float4(w[d][0], w[d][1], w[d][2], w[d][3]);
        taps[4 * i + 3] = make_float4(__int_as_float(base[0]), __int_as_float(base[1]), __int_as_float(base[2]), 0.0f);
    }
}

__global__ void bin_gather_kernel(BinGeom geo, const float *__restrict__ coord,
                                  const uint32_t *__restrict__ keys,
                                  const int32_t *__restrict__ vals, int64_t n,
                                  float4 *__restrict__ entries) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t frame = keys[e] / geo.ntiles;
        const int32_t pt = vals[e];
        const float *c = coord + (frame * geo.npts + pt) * geo.ndim;
        float4 v;
        v.x = c[0];
        v.y = geo.ndim > 1 ? c[1] : 0.0f;
        v.z = geo.ndim > 2 ? c[2] : 0.0f;
        v.w = __int_as_float(pt);
        entries[e] = v;
    }
}

// Reordering of gridding bins: inside each bin, entries are ordered by
// (rank of the sample among the bin's samples with the same footprint
// origin, origin, sample). Consecutive entries then mostly have distinct
// origins, so the lanes of one LDS atomic in the gridding kernel update
// distinct cells. The order is a pure function of the trajectory.
__device__ __forceinline__ uint32_t entry_origin(const BinGeom &geo, const float4 &en) {
    const float c[3] = {en.x, en.y, en.z};
    int b[3] = {0, 0, 0};
    for (int d = 0; d < geo.ndim; ++d) b[d] = axis_base(geo.ax[d], c[d]);
    const uint32_t gy = geo.ndim > 1 ? geo.ax[1].G : 1;
    return ((uint32_t)b[2] * gy + (uint32_t)b[1]) * (uint32_t)geo.ax[0].G + (uint32_t)b[0];
}

__global__ void origin_key_kernel(BinGeom geo, const uint32_t *__restrict__ bin_keys,
                                  const float4 *__restrict__ entries, int64_t n,
                                  uint64_t *__restrict__ key, int32_t *__restrict__ val) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        key[e] = ((uint64_t)bin_keys[e] << 32) | entry_origin(geo, entries[e]);
        val[e] = (int32_t)e;
    }
}

__global__ void group_head_kernel(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ head) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        head[e] = (e == 0 || key[e] != key[e - 1]) ? (int32_t)e : 0;
}

__global__ void rank_key_kernel(const uint64_t *__restrict__ key, const int32_t *__restrict__ start,
                                int64_t n, uint64_t *__restrict__ key2) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        key2[e] = (key[e] >> 32 << 32) | (uint32_t)(e - start[e]);
}

__global__ void permute_entries_kernel(const float4 *__restrict__ src, const int32_t *__restrict__ idx,
                                       int64_t n, float4 *__restrict__ dst) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = src[idx[e]];
}

struct MaxOp {
    __device__ __forceinline__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};

// Marks every grid cell touched by a footprint (one frame).
__global__ void footprint_mark_kernel(BinGeom geo, const float *__restrict__ coord,
                                      uint8_t *__restrict__ mask) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < geo.npts;
         i += (int64_t)gridDim.x * blockDim.x) {
        int base[3] = {0, 0, 0};
        int w[3] = {1, 1, 1};
        for (int d = 0; d < geo.ndim; ++d) {
            base[d] = axis_base(geo.ax[d], coord[i * geo.ndim + d]);
            w[d] = geo.ax[d].W;
        }
        const int Gx = geo.ax[0].G, Gy = geo.ndim > 1 ? geo.ax[1].G : 1;
        const int Gz = geo.ndim > 2 ? geo.ax[2].G : 1;
        for (int a = 0; a < w[2]; ++a) {
            const int z = geo.ndim > 2 ? wrap_add(base[2], a, Gz) : 0;
            for (int b = 0; b < w[1]; ++b) {
                const int y = geo.ndim > 1 ? wrap_add(base[1], b, Gy) : 0;
                for (int c = 0; c < w[0]; ++c) {
                    const int x = wrap_add(base[0], c, Gx);
                    mask[((int64_t)z * Gy + y) * Gx + x] = 1;
                }
            }
        }
    }
}

__global__ void count_mask_kernel(const uint8_t *__restrict__ mask, int64_t n,
                                  unsigned long long *__restrict__ out) {
    unsigned long long local = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        local += mask[i];
    __shared__ unsigned long long red[kBlock];
    red[threadIdx.x] = local;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

BinGeom make_geom(const tn_plan *plan, bool primary = false) {
    BinGeom g{};
    g.ndim = plan->p.ndim;
    g.npts = plan->p.npts;
    g.nframes = plan->p.nframes;
    for (int d = 0; d < 3; ++d) {
        const int zyx = 2 - d;
        g.ax[d].G = plan->p.grid[zyx];
        g.ax[d].W = plan->p.width[zyx];
        g.ax[d].shift = plan->p.shift[zyx];
        g.beta[d] = plan->p.beta[zyx];
        g.ax[d].T = primary ? plan->itile[zyx] : plan->tile[zyx];
        g.NT[d] = primary ? plan->nitiles[zyx] : plan->ntiles[zyx];
    }
    g.ntiles = primary ? plan->nitiles_total : plan->ntiles_total;
    return g;
}

unsigned grid_blocks(int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

// RAII holder for temporaries of the plan build.
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

namespace {

// Stable radix sort of (key, sample) pairs, then per-bin offsets and the
// sorted entry records (coordinates + sample index).
int order_by_origin(const BinGeom &geo, const uint32_t *bin_keys, int64_t n, float4 *entries,
                    bool spread, hipStream_t stream);

// Entry order inside each bin: sample order, footprint-origin order (runs
// of equal origin, gridding by register runs) or round-robin over origins
// (gridding by per-sample LDS atomics).
enum EntryOrder { kOrderSample = 0, kOrderOrigin = 1, kOrderSpread = 2 };

int sort_entries(const BinGeom &geo, const float *coord, void *keys, void *vals, int64_t n,
                 int32_t *bin_off, float4 **entries, hipStream_t stream, int order = kOrderSample) {
    const int64_t nbins = geo.nframes * geo.ntiles;
    DevBuf keys_sorted, vals_sorted, sort_tmp;
    TN_HIP(hipMalloc(&keys_sorted.p, sizeof(uint32_t) * n));
    TN_HIP(hipMalloc(&vals_sorted.p, sizeof(int32_t) * n));
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) < (uint64_t)nbins) ++end_bit;
    size_t sort_bytes = 0;
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)keys,
                                              (uint32_t *)keys_sorted.p, (int32_t *)vals,
                                              (int32_t *)vals_sorted.p, (int)n, 0, end_bit,
                                              stream));
    TN_HIP(hipMalloc(&sort_tmp.p, sort_bytes));
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, sort_bytes, (uint32_t *)keys,
                                              (uint32_t *)keys_sorted.p, (int32_t *)vals,
                                              (int32_t *)vals_sorted.p, (int)n, 0, end_bit,
                                              stream));
    hipLaunchKernelGGL(bin_offsets_kernel, dim3(grid_blocks(n + 1)), dim3(kBlock), 0, stream,
                       (const uint32_t *)keys_sorted.p, n, nbins, bin_off);
    TN_HIP(hipGetLastError());
    TN_HIP(hipMalloc(entries, sizeof(float4) * std::max<int64_t>(n, 1)));
    hipLaunchKernelGGL(bin_gather_kernel, dim3(grid_blocks(n)), dim3(kBlock), 0, stream, geo,
                       coord, (const uint32_t *)keys_sorted.p, (const int32_t *)vals_sorted.p, n,
                       *entries);
    TN_HIP(hipGetLastError());
    if (order != kOrderSample) {
        int rc = order_by_origin(geo, (const uint32_t *)keys_sorted.p, n, *entries, order == kOrderSpread,
                                 stream);
        if (rc) return rc;
    }
    TN_HIP(hipStreamSynchronize(stream));  // temporaries are freed on scope exit
    return TN_OK;
}

int order_by_origin(const BinGeom &geo, const uint32_t *bin_keys, int64_t n, float4 *entries,
                    bool spread, hipStream_t stream) {
    DevBuf k1, v1, k1s, v1s, head, start, k2, v2s, tmp, copy;
    TN_HIP(hipMalloc(&k1.p, sizeof(uint64_t) * n));
    TN_HIP(hipMalloc(&v1.p, sizeof(int32_t) * n));
    TN_HIP(hipMalloc(&k1s.p, sizeof(uint64_t) * n));
    TN_HIP(hipMalloc(&v1s.p, sizeof(int32_t) * n));
    TN_HIP(hipMalloc(&head.p, sizeof(int32_t) * n));
    TN_HIP(hipMalloc(&start.p, sizeof(int32_t) * n));
    TN_HIP(hipMalloc(&k2.p, sizeof(uint64_t) * n));
    TN_HIP(hipMalloc(&v2s.p, sizeof(int32_t) * n));
    TN_HIP(hipMalloc(&copy.p, sizeof(float4) * n));
    const unsigned nb = grid_blocks(n);
    hipLaunchKernelGGL(origin_key_kernel, dim3(nb), dim3(kBlock), 0, stream, geo, bin_keys, entries, n,
                       (uint64_t *)k1.p, (int32_t *)v1.p);
    TN_HIP(hipGetLastError());
    size_t bytes = 0, b2 = 0, b3 = 0;
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint64_t *)k1.p, (uint64_t *)k1s.p,
                                              (int32_t *)v1.p, (int32_t *)v1s.p, (int)n, 0, 64, stream));
    TN_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, b2, (int32_t *)head.p, (int32_t *)start.p, MaxOp(),
                                             (int)n, stream));
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b3, (uint64_t *)k2.p, (uint64_t *)k1s.p,
                                              (int32_t *)v1s.p, (int32_t *)v2s.p, (int)n, 0, 64, stream));
    TN_HIP(hipMalloc(&tmp.p, std::max(bytes, std::max(b2, b3))));
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, (uint64_t *)k1.p, (uint64_t *)k1s.p,
                                              (int32_t *)v1.p, (int32_t *)v1s.p, (int)n, 0, 64, stream));
    if (!spread) {  // (bin, origin, sample) order
        TN_HIP(hipMemcpyAsync(copy.p, entries, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
        hipLaunchKernelGGL(permute_entries_kernel, dim3(nb), dim3(kBlock), 0, stream, (const float4 *)copy.p,
                           (const int32_t *)v1s.p, n, entries);
        TN_HIP(hipGetLastError());
        TN_HIP(hipStreamSynchronize(stream));
        return TN_OK;
    }
    hipLaunchKernelGGL(group_head_kernel, dim3(nb), dim3(kBlock), 0, stream, (const uint64_t *)k1s.p, n,
                       (int32_t *)head.p);
    TN_HIP(hipGetLastError());
    TN_HIP(hipcub::DeviceScan::InclusiveScan(tmp.p, b2, (int32_t *)head.p, (int32_t *)start.p, MaxOp(),
                                             (int)n, stream));
    hipLaunchKernelGGL(rank_key_kernel, dim3(nb), dim3(kBlock), 0, stream, (const uint64_t *)k1s.p,
                       (const int32_t *)start.p, n, (uint64_t *)k2.p);
    TN_HIP(hipGetLastError());
    // stable: equal (bin, rank) keep the (bin, origin, sample) order
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, b3, (uint64_t *)k2.p, (uint64_t *)k1s.p,
                                              (int32_t *)v1s.p, (int32_t *)v2s.p, (int)n, 0, 64, stream));
    TN_HIP(hipMemcpyAsync(copy.p, entries, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(permute_entries_kernel, dim3(nb), dim3(kBlock), 0, stream, (const float4 *)copy.p,
                       (const int32_t *)v2s.p, n, entries);
    TN_HIP(hipGetLastError());
    TN_HIP(hipStreamSynchronize(stream));
    return TN_OK;
}

}  // namespace

int build_bins(tn_plan *plan, hipStream_t stream) {
    const BinGeom geo = make_geom(plan);
    const BinGeom pgeo = make_geom(plan, true);
    const int64_t total = geo.npts * geo.nframes;
    const int64_t nbins = geo.nframes * geo.ntiles;
    const int64_t npbins = pgeo.nframes * pgeo.ntiles;
    if (nbins >= (int64_t)UINT32_MAX || npbins >= (int64_t)UINT32_MAX)
        return fail(TN_E_RANGE, "too many tiles (%lld)", (long long)std::max(nbins, npbins));

    TN_HIP(hipMalloc(&plan->d_bin_off, sizeof(int32_t) * (nbins + 1)));
    TN_HIP(hipMalloc(&plan->d_pbin_off, sizeof(int32_t) * (npbins + 1)));
    plan->h_pbin_off.assign(npbins + 1, 0);
    plan->h_bin_off.assign(nbins + 1, 0);
    if (total == 0) {
        TN_HIP(hipMemsetAsync(plan->d_bin_off, 0, sizeof(int32_t) * (nbins + 1), stream));
        TN_HIP(hipMemsetAsync(plan->d_pbin_off, 0, sizeof(int32_t) * (npbins + 1), stream));
        TN_HIP(hipStreamSynchronize(stream));
        plan->nentries = 0;
        // an empty gridding matrix: all-zero row offsets, so every cell is
        // written as zero
        if (plan->grid_mode == kGridMatrix) return build_matrix(plan, stream);
        return TN_OK;
    }

    // ---- gridding bins: every tile a footprint touches --------------------
    if (plan->grid_mode == kGridMatrix) {  // gridding reads the matrix instead
        TN_HIP(hipMemsetAsync(plan->d_bin_off, 0, sizeof(int32_t) * (nbins + 1), stream));
        plan->nentries = 0;
        int rc = build_matrix(plan, stream);
        if (rc) return rc;
    } else {
        DevBuf counts, offsets, scan_tmp;
        TN_HIP(hipMalloc(&counts.p, sizeof(int32_t) * total));
        TN_HIP(hipMalloc(&offsets.p, sizeof(int32_t) * (total + 1)));
        hipLaunchKernelGGL(bin_count_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo,
                           plan->d_coord, (int32_t *)counts.p);
        TN_HIP(hipGetLastError());
        size_t tmp_bytes = 0;
        TN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (int32_t *)counts.p,
                                                (int32_t *)offsets.p, (int)total, stream));
        TN_HIP(hipMalloc(&scan_tmp.p, tmp_bytes));
        TN_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, tmp_bytes, (int32_t *)counts.p,
                                                (int32_t *)offsets.p, (int)total, stream));
        int32_t last_off = 0, last_cnt = 0;
        TN_HIP(hipMemcpyAsync(&last_off, (int32_t *)offsets.p + (total - 1), sizeof(int32_t),
                              hipMemcpyDeviceToHost, stream));
        TN_HIP(hipMemcpyAsync(&last_cnt, (int32_t *)counts.p + (total - 1), sizeof(int32_t),
                              hipMemcpyDeviceToHost, stream));
        TN_HIP(hipStreamSynchronize(stream));
        const int64_t n = (int64_t)last_off + last_cnt;
        if (n >= INT32_MAX) return fail(TN_E_RANGE, "too many tile entries (%lld)", (long long)n);
        plan->nentries = n;
        DevBuf keys, vals;
        TN_HIP(hipMalloc(&keys.p, sizeof(uint32_t) * n));
        TN_HIP(hipMalloc(&vals.p, sizeof(int32_t) * n));
        hipLaunchKernelGGL(bin_emit_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo,
                           plan->d_coord, (const int32_t *)offsets.p, (uint32_t *)keys.p,
                           (int32_t *)vals.p);
        TN_HIP(hipGetLastError());
        // register tiles: each tile's entries in sample order (the stable
        // sort's), the sums' fixed order
        int rc = sort_entries(geo, plan->d_coord, keys.p, vals.p, n, plan->d_bin_off,
                              &plan->d_entries, stream,
                              plan->grid_mode == kGridRuns    ? kOrderOrigin
                              : plan->grid_mode == kGridTiles ? kOrderSample
                                                              : kOrderSpread);
        if (rc) return rc;
        if (plan->grid_mode == kGridTiles) {
            TN_HIP(hipMalloc(&plan->d_taps, sizeof(float4) * 4 * total));
            hipLaunchKernelGGL(tiles_taps_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo,
                               plan->d_coord, plan->d_taps);
            TN_HIP(hipGetLastError());
        }
        plan->h_bin_off.assign(nbins + 1, 0);
        TN_HIP(hipMemcpyAsync(plan->h_bin_off.data(), plan->d_bin_off, sizeof(int32_t) * (nbins + 1),
                              hipMemcpyDeviceToHost, stream));
        TN_HIP(hipStreamSynchronize(stream));
    }
    // ---- interpolation bins: every sample once -----------------------------
    {
        DevBuf keys, vals;
        TN_HIP(hipMalloc(&keys.p, sizeof(uint32_t) * total));
        TN_HIP(hipMalloc(&vals.p, sizeof(int32_t) * total));
        hipLaunchKernelGGL(bin_primary_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream,
                           pgeo, plan->d_coord, (uint32_t *)keys.p, (int32_t *)vals.p);
        TN_HIP(hipGetLastError());
        // 3D: footprint-origin order inside each tile, so that the lanes of
        // a wave read neighbouring cells of the staged region (fewer LDS bank
        // conflicts; every sample's sum is the same whichever lane forms it);
        // 2D measures the same either way and keeps sample order
        int rc = sort_entries(pgeo, plan->d_coord, keys.p, vals.p, total, plan->d_pbin_off,
                              &plan->d_pentries, stream, pgeo.ndim == 3 ? kOrderOrigin : kOrderSample);
        if (rc) return rc;
        TN_HIP(hipMemcpyAsync(plan->h_pbin_off.data(), plan->d_pbin_off, sizeof(int32_t) * (npbins + 1),
                              hipMemcpyDeviceToHost, stream));
        TN_HIP(hipStreamSynchronize(stream));
    }
    return TN_OK;
}

// Sparse gridding matrix of every frame, CSR by output cell: the entries of
// every sample are sorted by (frame, cell) with a stable radix sort, so the
// samples of one cell stay in ascending order (a fixed summation order).
int build_matrix(tn_plan *plan, hipStream_t stream) {
    const BinGeom geo = make_geom(plan);
    const int64_t total = geo.npts * geo.nframes;
    const int64_t ncell = (int64_t)plan->p.grid[0] * plan->p.grid[1] * plan->p.grid[2];
    const int64_t nrows = ncell * geo.nframes;
    const int taps = plan->p.width[0] * plan->p.width[1] * plan->p.width[2];
    const int64_t n = total * taps;
    if (nrows >= (int64_t)UINT32_MAX) return fail(TN_E_RANGE, "gridding matrix: too many rows (%lld)", (long long)nrows);
    if (n >= INT32_MAX) return fail(TN_E_RANGE, "gridding matrix: too many entries (%lld)", (long long)n);
    if (geo.npts > ((int64_t)1 << 28)) return fail(TN_E_RANGE, "gridding matrix: too many samples per frame");
    TN_HIP(hipMalloc(&plan->d_mrow, sizeof(int32_t) * (nrows + 1)));
    TN_HIP(hipMalloc(&plan->d_mnz, sizeof(int2) * std::max<int64_t>(n, 1)));
    plan->mnnz = n;
    if (n == 0) {
        TN_HIP(hipMemsetAsync(plan->d_mrow, 0, sizeof(int32_t) * (nrows + 1), stream));
        TN_HIP(hipStreamSynchronize(stream));
        plan->h_mrow.assign(nrows + 1, 0);
        return TN_OK;
    }
    DevBuf keys, vals, keys_sorted, tmp;
    TN_HIP(hipMalloc(&keys.p, sizeof(uint32_t) * n));
    TN_HIP(hipMalloc(&vals.p, sizeof(uint64_t) * n));
    TN_HIP(hipMalloc(&keys_sorted.p, sizeof(uint32_t) * n));
    hipLaunchKernelGGL(matrix_emit_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo, plan->d_coord,
                       taps, (uint32_t *)keys.p, (uint64_t *)vals.p);
    TN_HIP(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) < (uint64_t)nrows) ++end_bit;
    size_t bytes = 0;
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t *)keys.p, (uint32_t *)keys_sorted.p,
                                              (uint64_t *)vals.p, (uint64_t *)plan->d_mnz, (int)n, 0, end_bit,
                                              stream));
    TN_HIP(hipMalloc(&tmp.p, bytes));
    TN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, (uint32_t *)keys.p, (uint32_t *)keys_sorted.p,
                                              (uint64_t *)vals.p, (uint64_t *)plan->d_mnz, (int)n, 0, end_bit,
                                              stream));
    hipLaunchKernelGGL(bin_offsets_kernel, dim3(grid_blocks(n + 1)), dim3(kBlock), 0, stream,
                       (const uint32_t *)keys_sorted.p, n, nrows, plan->d_mrow);
    TN_HIP(hipGetLastError());
    plan->h_mrow.resize(nrows + 1);
    TN_HIP(hipMemcpyAsync(plan->h_mrow.data(), plan->d_mrow, sizeof(int32_t) * (nrows + 1), hipMemcpyDeviceToHost,
                          stream));
    TN_HIP(hipStreamSynchronize(stream));  // temporaries are freed on scope exit
    return TN_OK;
}

int count_footprint(tn_plan *plan, hipStream_t stream) {
    const BinGeom geo = make_geom(plan);
    const int64_t gvol = (int64_t)plan->p.grid[0] * plan->p.grid[1] * plan->p.grid[2];
    DevBuf mask, counter;
    TN_HIP(hipMalloc(&mask.p, gvol));
    TN_HIP(hipMalloc(&counter.p, sizeof(unsigned long long)));
    TN_HIP(hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), stream));
    for (int f = 0; f < geo.nframes; ++f) {
        TN_HIP(hipMemsetAsync(mask.p, 0, gvol, stream));
        hipLaunchKernelGGL(footprint_mark_kernel, dim3(grid_blocks(geo.npts)), dim3(kBlock), 0,
                           stream, geo, plan->d_coord + (int64_t)f * geo.npts * geo.ndim,
                           (uint8_t *)mask.p);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL(count_mask_kernel, dim3(grid_blocks(gvol)), dim3(kBlock), 0, stream,
                           (const uint8_t *)mask.p, gvol, (unsigned long long *)counter.p);
        TN_HIP(hipGetLastError());
    }
    unsigned long long cells = 0;
    TN_HIP(hipMemcpyAsync(&cells, counter.p, sizeof(cells), hipMemcpyDeviceToHost, stream));
    TN_HIP(hipStreamSynchronize(stream));
    plan->footprint = (int64_t)cells;
    return TN_OK;
}

}  // namespace tn
