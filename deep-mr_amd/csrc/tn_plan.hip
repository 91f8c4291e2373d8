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

// ---- tile matrix (grid_mode 3) ----------------------------------------------
// Gridding of a tile sums, for every cell, weight * value over the samples of
// the tile's bin whose footprint covers the cell. The bin (samples in
// ascending order) is cut into batches of `tslots` slots; per batch and cell
// the matrix lists the entries (slot in the batch | tap code), in slot order.
// The tap code is the footprint tap per axis, tbits bits each (x lowest), so
// the gridding kernel forms the weight from the slot's per-axis taps.

struct TileGeom {
    BinGeom g;
    int tslots, sbits, tbits;
    int cells;  // tile cells T0 * T1 * T2 (edge tiles included in full)
};

__device__ __forceinline__ void tile_of_key(const BinGeom &g, uint32_t key, int64_t &frame, int (&t)[3]) {
    frame = key / g.ntiles;
    int64_t tile = key - frame * g.ntiles;
    t[0] = (int)(tile % g.NT[0]);
    tile /= g.NT[0];
    t[1] = (int)(tile % g.NT[1]);
    t[2] = (int)(tile / g.NT[1]);
}

// Batches per bin: ceil(bin size / tslots).
__global__ void batch_count_kernel(const int32_t *__restrict__ bin_off, int64_t nbins, int tslots,
                                   int32_t *__restrict__ nb) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nbins; k += (int64_t)gridDim.x * blockDim.x)
        nb[k] = (bin_off[k + 1] - bin_off[k] + tslots - 1) / tslots;
}

// Footprint taps of bin entry e that fall inside its tile (per axis: the
// taps k whose wrapped index lies in the tile), as (count, per-axis lists).
__device__ __forceinline__ int tile_taps(const TileGeom &tg, const float *coord, int64_t frame, int32_t pt,
                                         const int (&t)[3], int (&kk)[3][kMaxWidth], int (&nk)[3]) {
    const BinGeom &g = tg.g;
    int n = 1;
    for (int d = 0; d < 3; ++d) {
        nk[d] = 1;
        kk[d][0] = 0;
        if (d >= g.ndim) continue;
        const AxisGeom &a = g.ax[d];
        const int base = axis_base(a, coord[(frame * g.npts + pt) * g.ndim + d]);
        int m = 0;
        for (int k = 0; k < a.W; ++k)
            if (wrap_add(base, k, a.G) / a.T == t[d]) kk[d][m++] = k;
        nk[d] = m;
        n *= m;
    }
    return n;
}

__global__ void tile_tap_count_kernel(TileGeom tg, const float *__restrict__ coord, const uint32_t *__restrict__ keys,
                                      const int32_t *__restrict__ bins, int64_t n, int32_t *__restrict__ cnt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t frame;
        int t[3];
        tile_of_key(tg.g, keys[e], frame, t);
        int kk[3][kMaxWidth], nk[3];
        cnt[e] = tile_taps(tg, coord, frame, bins[e], t, kk, nk);
    }
}

__global__ void tile_emit_kernel(TileGeom tg, const float *__restrict__ coord, const uint32_t *__restrict__ keys,
                                 const int32_t *__restrict__ bins, const int32_t *__restrict__ bin_off,
                                 const int32_t *__restrict__ batch_off, const int32_t *__restrict__ eoff, int64_t n,
                                 uint32_t *__restrict__ okey, uint16_t *__restrict__ oval) {
    const BinGeom &g = tg.g;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t key = keys[e];
        int64_t frame;
        int t[3];
        tile_of_key(g, key, frame, t);
        int kk[3][kMaxWidth], nk[3];
        tile_taps(tg, coord, frame, bins[e], t, kk, nk);
        const int slot = (int)(e - bin_off[key]);
        const int64_t batch = (int64_t)batch_off[key] + slot / tg.tslots;
        const uint32_t s = (uint32_t)(slot % tg.tslots);
        int base[3] = {0, 0, 0};
        for (int d = 0; d < g.ndim; ++d)
            base[d] = axis_base(g.ax[d], coord[(frame * g.npts + bins[e]) * g.ndim + d]);
        const int T0 = g.ax[0].T, T1 = g.ndim > 1 ? g.ax[1].T : 1;
        int64_t pos = eoff[e];
        for (int a = 0; a < nk[2]; ++a) {
            const int kz = kk[2][a];
            const int lz = g.ndim > 2 ? wrap_add(base[2], kz, g.ax[2].G) - t[2] * g.ax[2].T : 0;
            for (int b = 0; b < nk[1]; ++b) {
                const int ky = kk[1][b];
                const int ly = g.ndim > 1 ? wrap_add(base[1], ky, g.ax[1].G) - t[1] * T1 : 0;
                for (int c = 0; c < nk[0]; ++c) {
                    const int kx = kk[0][c];
                    const int lx = wrap_add(base[0], kx, g.ax[0].G) - t[0] * T0;
                    const int cell = (lz * T1 + ly) * T0 + lx;
                    okey[pos] = (uint32_t)(batch * tg.cells + cell);
                    const uint32_t code = (uint32_t)kx | ((uint32_t)ky << tg.tbits) | ((uint32_t)kz << (2 * tg.tbits));
                    oval[pos] = (uint16_t)(s | (code << tg.sbits));
                    ++pos;
                }
            }
        }
    }
}

// Per batch: entry base and the cells + 1 row starts relative to it.
__global__ void tile_rows_kernel(const int32_t *__restrict__ rowstart, int64_t nbatch, int cells,
                                 int32_t *__restrict__ mbase, uint16_t *__restrict__ mrow) {
    const int64_t total = nbatch * (cells + 1);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / (cells + 1);
        const int c = (int)(i - b * (cells + 1));
        const int32_t b0 = rowstart[b * cells];
        mrow[i] = (uint16_t)(rowstart[b * cells + c] - b0);
        if (c == 0) mbase[b] = b0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) mbase[nbatch] = rowstart[nbatch * cells];
}

template <class T>
int device_scan(const int32_t *in, T *out, int64_t n, hipStream_t stream, int64_t *total) {
    // exclusive sum into out[0..n], out[n] = total
    DevBuf tmp;
    size_t bytes = 0;
    TN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, stream));
    TN_HIP(hipMalloc(&tmp.p, std::max<size_t>(bytes, 1)));
    TN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, stream));
    T last_out = 0;
    int32_t last_in = 0;
    if (n > 0) {
        TN_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(T), hipMemcpyDeviceToHost, stream));
        TN_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    }
    TN_HIP(hipStreamSynchronize(stream));
    *total = (int64_t)last_out + last_in;
    const T tot = (T)*total;
    TN_HIP(hipMemcpyAsync(out + n, &tot, sizeof(T), hipMemcpyHostToDevice, stream));
    TN_HIP(hipStreamSynchronize(stream));
    return TN_OK;
}

}  // namespace

// Slot / tap-code split of the 16-bit entries and the batch size of a plan.
void tile_matrix_params(const tn_plan *plan, int *tslots, int *sbits, int *tbits) {
    const tn_plan_params &p = plan->p;
    int tb = 0, in_tile = 1;
    for (int d = 3 - p.ndim; d < 3; ++d) {
        while ((1 << tb) < p.width[d]) ++tb;
        in_tile *= std::min(p.width[d], plan->tile[d]);
    }
    const int sb = 16 - p.ndim * tb;
    int s = std::min(kTileSlots, 1 << std::max(sb, 0));
    while (s > 1 && (int64_t)s * in_tile > 65535) s >>= 1;  // batch entries fit the 16-bit row starts
    *tslots = s;
    *sbits = sb;
    *tbits = tb;
}

int build_tile_matrix(tn_plan *plan, hipStream_t stream) {
    TileGeom tg{};
    tg.g = make_geom(plan);
    tile_matrix_params(plan, &tg.tslots, &tg.sbits, &tg.tbits);
    plan->tslots = tg.tslots;
    plan->sbits = tg.sbits;
    plan->tbits = tg.tbits;
    if (tg.sbits < 1) return fail(TN_E_ARG, "tile matrix: kernel width too large for 16-bit entries");
    tg.cells = plan->tile[0] * plan->tile[1] * plan->tile[2];
    const BinGeom &geo = tg.g;
    const int64_t total = geo.npts * geo.nframes;
    const int64_t nbins = geo.nframes * geo.ntiles;
    // ---- bins: every (tile, sample) pair, sorted by (frame, tile, sample)
    int64_t n = 0;
    DevBuf keys_sorted;
    if (total > 0) {
        DevBuf counts, offsets, keys, vals, sort_tmp;
        TN_HIP(hipMalloc(&counts.p, sizeof(int32_t) * total));
        TN_HIP(hipMalloc(&offsets.p, sizeof(int32_t) * (total + 1)));
        hipLaunchKernelGGL(bin_count_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo, plan->d_coord,
                           (int32_t *)counts.p);
        TN_HIP(hipGetLastError());
        int rc = device_scan((const int32_t *)counts.p, (int32_t *)offsets.p, total, stream, &n);
        if (rc) return rc;
        if (n >= INT32_MAX) return fail(TN_E_RANGE, "too many tile entries (%lld)", (long long)n);
        TN_HIP(hipMalloc(&keys.p, sizeof(uint32_t) * std::max<int64_t>(n, 1)));
        TN_HIP(hipMalloc(&vals.p, sizeof(int32_t) * std::max<int64_t>(n, 1)));
        TN_HIP(hipMalloc(&keys_sorted.p, sizeof(uint32_t) * std::max<int64_t>(n, 1)));
        TN_HIP(hipMalloc(&plan->d_bins, sizeof(int32_t) * std::max<int64_t>(n, 1)));
        hipLaunchKernelGGL(bin_emit_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream, geo, plan->d_coord,
                           (const int32_t *)offsets.p, (uint32_t *)keys.p, (int32_t *)vals.p);
        TN_HIP(hipGetLastError());
        int end_bit = 1;
        while (end_bit < 32 && ((uint64_t)1 << end_bit) < (uint64_t)nbins) ++end_bit;
        size_t bytes = 0;
        TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t *)keys.p, (uint32_t *)keys_sorted.p,
                                                  (int32_t *)vals.p, plan->d_bins, (int)n, 0, end_bit, stream));
        TN_HIP(hipMalloc(&sort_tmp.p, std::max<size_t>(bytes, 1)));
        TN_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, bytes, (uint32_t *)keys.p, (uint32_t *)keys_sorted.p,
                                                  (int32_t *)vals.p, plan->d_bins, (int)n, 0, end_bit, stream));
        hipLaunchKernelGGL(bin_offsets_kernel, dim3(grid_blocks(n + 1)), dim3(kBlock), 0, stream,
                           (const uint32_t *)keys_sorted.p, n, nbins, plan->d_bin_off);
        TN_HIP(hipGetLastError());
    } else {
        TN_HIP(hipMalloc(&plan->d_bins, sizeof(int32_t)));
        TN_HIP(hipMemsetAsync(plan->d_bin_off, 0, sizeof(int32_t) * (nbins + 1), stream));
    }
    plan->nentries = n;
    // ---- batches per bin
    TN_HIP(hipMalloc(&plan->d_batch_off, sizeof(int32_t) * (nbins + 1)));
    int64_t nbatch = 0;
    {
        DevBuf nb;
        TN_HIP(hipMalloc(&nb.p, sizeof(int32_t) * std::max<int64_t>(nbins, 1)));
        hipLaunchKernelGGL(batch_count_kernel, dim3(grid_blocks(nbins)), dim3(kBlock), 0, stream, plan->d_bin_off,
                           nbins, tg.tslots, (int32_t *)nb.p);
        TN_HIP(hipGetLastError());
        int rc = device_scan((const int32_t *)nb.p, plan->d_batch_off, nbins, stream, &nbatch);
        if (rc) return rc;
    }
    plan->nbatches = nbatch;
    if (nbatch * tg.cells >= (int64_t)UINT32_MAX)
        return fail(TN_E_RANGE, "tile matrix: too many rows (%lld batches)", (long long)nbatch);
    // ---- entries: (batch, cell) keys, slot | tap code values
    int64_t ne = 0;
    DevBuf okey_sorted;
    if (n > 0) {
        DevBuf cnt, eoff, okey, oval, sort_tmp;
        TN_HIP(hipMalloc(&cnt.p, sizeof(int32_t) * n));
        TN_HIP(hipMalloc(&eoff.p, sizeof(int32_t) * (n + 1)));
        hipLaunchKernelGGL(tile_tap_count_kernel, dim3(grid_blocks(n)), dim3(kBlock), 0, stream, tg, plan->d_coord,
                           (const uint32_t *)keys_sorted.p, plan->d_bins, n, (int32_t *)cnt.p);
        TN_HIP(hipGetLastError());
        int rc = device_scan((const int32_t *)cnt.p, (int32_t *)eoff.p, n, stream, &ne);
        if (rc) return rc;
        if (ne >= INT32_MAX) return fail(TN_E_RANGE, "tile matrix: too many entries (%lld)", (long long)ne);
        TN_HIP(hipMalloc(&okey.p, sizeof(uint32_t) * ne));
        TN_HIP(hipMalloc(&oval.p, sizeof(uint16_t) * ne));
        TN_HIP(hipMalloc(&okey_sorted.p, sizeof(uint32_t) * ne));
        TN_HIP(hipMalloc(&plan->d_ment, sizeof(uint16_t) * ne));
        hipLaunchKernelGGL(tile_emit_kernel, dim3(grid_blocks(n)), dim3(kBlock), 0, stream, tg, plan->d_coord,
                           (const uint32_t *)keys_sorted.p, plan->d_bins, plan->d_bin_off, plan->d_batch_off,
                           (const int32_t *)eoff.p, n, (uint32_t *)okey.p, (uint16_t *)oval.p);
        TN_HIP(hipGetLastError());
        const uint64_t nrows = (uint64_t)nbatch * tg.cells;
        int end_bit = 1;
        while (end_bit < 32 && ((uint64_t)1 << end_bit) < nrows) ++end_bit;
        size_t bytes = 0;
        TN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t *)okey.p, (uint32_t *)okey_sorted.p,
                                                  (uint16_t *)oval.p, plan->d_ment, (int)ne, 0, end_bit, stream));
        TN_HIP(hipMalloc(&sort_tmp.p, std::max<size_t>(bytes, 1)));
        TN_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, bytes, (uint32_t *)okey.p, (uint32_t *)okey_sorted.p,
                                                  (uint16_t *)oval.p, plan->d_ment, (int)ne, 0, end_bit, stream));
        TN_HIP(hipStreamSynchronize(stream));  // the sort's temporaries are freed on scope exit
    } else {
        TN_HIP(hipMalloc(&plan->d_ment, sizeof(uint16_t)));
    }
    plan->mnnz = ne;
    // ---- per-batch entry bases and row starts
    TN_HIP(hipMalloc(&plan->d_mbase, sizeof(int32_t) * (nbatch + 1)));
    TN_HIP(hipMalloc(&plan->d_mrow, sizeof(uint16_t) * std::max<int64_t>(nbatch * (tg.cells + 1), 1)));
    if (nbatch > 0) {
        DevBuf rowstart;
        const int64_t nrows = nbatch * tg.cells;
        TN_HIP(hipMalloc(&rowstart.p, sizeof(int32_t) * (nrows + 1)));
        hipLaunchKernelGGL(bin_offsets_kernel, dim3(grid_blocks(ne + 1)), dim3(kBlock), 0, stream,
                           (const uint32_t *)okey_sorted.p, ne, nrows, (int32_t *)rowstart.p);
        TN_HIP(hipGetLastError());
        hipLaunchKernelGGL(tile_rows_kernel, dim3(grid_blocks(nbatch * (tg.cells + 1))), dim3(kBlock), 0, stream,
                           (const int32_t *)rowstart.p, nbatch, tg.cells, plan->d_mbase, plan->d_mrow);
        TN_HIP(hipGetLastError());
        TN_HIP(hipStreamSynchronize(stream));
    } else {
        TN_HIP(hipMemsetAsync(plan->d_mbase, 0, sizeof(int32_t), stream));
    }
    plan->h_bin_off.assign(nbins + 1, 0);
    plan->h_batch_off.assign(nbins + 1, 0);
    TN_HIP(hipMemcpyAsync(plan->h_bin_off.data(), plan->d_bin_off, sizeof(int32_t) * (nbins + 1), hipMemcpyDeviceToHost,
                          stream));
    TN_HIP(hipMemcpyAsync(plan->h_batch_off.data(), plan->d_batch_off, sizeof(int32_t) * (nbins + 1),
                          hipMemcpyDeviceToHost, stream));
    TN_HIP(hipStreamSynchronize(stream));
    return TN_OK;
}

namespace {

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
        if (plan->grid_mode == kGridMatrix) return build_tile_matrix(plan, stream);  // empty bins, zero batches
        return TN_OK;
    }

    // ---- gridding bins: every tile a footprint touches --------------------
    if (plan->grid_mode == kGridMatrix) {
        int rc = build_tile_matrix(plan, stream);
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
        int rc = sort_entries(geo, plan->d_coord, keys.p, vals.p, n, plan->d_bin_off,
                              &plan->d_entries, stream,
                              plan->grid_mode == kGridRuns ? kOrderOrigin : kOrderSpread);
        if (rc) return rc;
        plan->h_bin_off.assign(nbins + 1, 0);
        TN_HIP(hipMemcpy(plan->h_bin_off.data(), plan->d_bin_off, sizeof(int32_t) * (nbins + 1),
                         hipMemcpyDeviceToHost));
    }
    // ---- interpolation bins: every sample once -----------------------------
    {
        DevBuf keys, vals;
        TN_HIP(hipMalloc(&keys.p, sizeof(uint32_t) * total));
        TN_HIP(hipMalloc(&vals.p, sizeof(int32_t) * total));
        hipLaunchKernelGGL(bin_primary_kernel, dim3(grid_blocks(total)), dim3(kBlock), 0, stream,
                           pgeo, plan->d_coord, (uint32_t *)keys.p, (int32_t *)vals.p);
        TN_HIP(hipGetLastError());
        int rc = sort_entries(pgeo, plan->d_coord, keys.p, vals.p, total, plan->d_pbin_off,
                              &plan->d_pentries, stream);
        if (rc) return rc;
        TN_HIP(hipMemcpy(plan->h_pbin_off.data(), plan->d_pbin_off, sizeof(int32_t) * (npbins + 1),
                         hipMemcpyDeviceToHost));
    }
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
