// Internal types shared by the tilenufft HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "tilenufft.h"
#include "tn_worklist.h"

namespace tn {

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWave = 64;
constexpr int kMaxWidth = 16;      // largest kernel width supported per axis
constexpr int kStagedMaxWidth = 8; // widest kernel of the staged interpolation and the tile gridding
                                   // variants; wider kernels gather and grid through the matrix
// launch "kind" of a table: ndim * kKindStride + widest kernel width
constexpr int kKindStride = 32;
constexpr int kind_ndim(int kind) { return kind / kKindStride; }
constexpr int kind_width(int kind) { return kind % kKindStride; }
constexpr int kInterpBlockPts = 256;  // samples per interpolation work item
constexpr size_t kGridLdsBudget = 80 * 1024;  // LDS per gridding workgroup (two per CU)
constexpr size_t kInterpLdsBudget = 64 * 1024;  // LDS per staged-interpolation workgroup
// samples per lane in the staged interpolation: 1 in 2D, 2 in 3D (each
// measured fastest on its workload; both appear in one table only alone)
constexpr int staged_ept(int ndim) { return ndim == 3 ? 2 : 1; }
constexpr int kLoadBatch = 4;  // global loads per lane issued before their LDS writes in staging loops
// default coils per region load of the staged interpolation, below what the
// LDS budget allows so that more workgroups fit on a CU (measured fastest:
// 2D 16x16 tiles, 8 coils 0.125 ms vs 16 coils 0.189 ms; 3D 8x8x16 tiles,
// 2 coils 1.33 ms vs 3 coils 1.61 ms)
constexpr int kInterpChunk2d = 8;
constexpr int kInterpChunk3d = 2;
constexpr int interp_segment(int ndim) { return 256 * staged_ept(ndim); }  // samples per staged work item
constexpr int kGridSplitDefault = 2048;  // max tile entries per gridding workgroup
constexpr int kGridBatch = 256;  // tile entries staged in LDS per gridding batch
constexpr int kGridPad = 4;      // accumulator row padding (floats)

// Gridding variants: one LDS atomic per tap and sample over entries spread
// across footprint origins, or per-lane register runs over origin-sorted
// entries (for at most kRunsMaxTaps taps per sample).
// Matrix: per-cell sums over the plan's sparse gridding matrix (CSR by grid
// cell), one lane per cell. Tiles (3D): one wave per gridding tile of
// kTilesT cells, its cells' sums in registers (lane = (x, y), one register
// set per z plane), over the tile's entries in sample order with per-sample
// taps from a plan table.
enum GridMode { kGridAuto = 0, kGridSpread = 1, kGridRuns = 2, kGridMatrix = 3, kGridTiles = 4 };
constexpr int kTilesT[3] = {8, 8, 4};  // register-tile gridding tile (x, y, z): 64 lanes x 4 planes
constexpr int kTilesMaxWidth = 4;      // taps per axis the register tiles hold
constexpr int kTilesCoils = 8;         // coils per register-tile item (4 planes x 8 coils x 2 floats)
constexpr int kTilesSplit = 384;       // default entries per register-tile item (kooshball 128^3, osf 2:
                                       // 3.30 ms against 3.52 at 256, 3.38 at 512, 3.86 at 1024)
constexpr int kMatrixLongRow = 128;  // default (2D): rows with more entries get a whole workgroup
constexpr int kMatrixLongRow3d = 256;  // default (3D): kooshball 128^3 at osf 2: 3.74 ms against 4.33 (128), 4.30 (512)
constexpr int kMatrixChunk = 16;     // default coils per item (record width of the coil-minor copy)
constexpr uint64_t kMatrixAutoBytes = 4ull << 30;  // largest gridding matrix the automatic mode builds
constexpr int kMatrixPasses = 4;     // cell passes per item of the coil-minor matrix gridding
constexpr int kLongGroups = 32;      // lane groups summing one long row of the coil-minor matrix gridding
constexpr int64_t kStaticScratchFloats = 8ll << 20;  // fixed gridding workspace: 32 MiB per device
constexpr int kMaxDevices = 64;
constexpr int kXcds = 8;          // accelerator complex dies (each with its own L2)
constexpr int kRunsMaxTaps = 36;  // register budget of the run sums of one z plane (W <= 6)
constexpr int kRunsBatch = 256;                 // entries per staged batch (= kBlock)
constexpr int kRunsMaxCoils = 64;               // coils per item: 4 waves x up to 16 lanes per group

// 32-bit words per staged entry of the register-run gridding: origin key,
// (tile-local tap-0 offset, wrap tap) per axis, W weights per axis.
__host__ __device__ constexpr int runs_record_words(int ndim, int wc) { return 1 + 2 * ndim + ndim * wc; }

// Dynamic LDS bytes of the gridding kernel for `ncoil` coils of a tile.
inline size_t grid_lds_bytes(int ncoil, int64_t plane, int wc, int mode = kGridSpread, int ndim = 3) {
    if (mode == kGridRuns)  // interleaved accumulators, entry records, sample values
        return sizeof(float) * (2 * (size_t)ncoil * plane + 4 + (size_t)runs_record_words(ndim, wc) * kRunsBatch +
                                2 * (size_t)(ncoil + 1) * kRunsBatch);
    return sizeof(float) * (2 * (size_t)ncoil * plane + 4 * (size_t)kGridBatch +
                            3 * (size_t)wc * kGridBatch + 2 * (size_t)ncoil * kGridBatch);
}

// Interpolation variants.
enum InterpMode { kInterpAuto = 0, kInterpGather = 1, kInterpStaged = 2 };

// One trajectory (frame) bound to its data placement for a launch. All
// per-axis arrays are (x, y, z) so that axis d matches coordinate column d.
struct alignas(16) Job {
    int32_t G[3];       // oversampled grid (x, y, z); z = 1 in 2D
    int32_t W[3];       // kernel width; 1 = Cartesian axis
    int32_t shift[3];   // output index = (tap + shift) mod G
    int32_t T[3];       // gridding tile
    int32_t NT[3];      // tiles per axis
    int32_t IT[3];      // staged-interpolation tile
    int32_t INT_[3];    // staged-interpolation tiles per axis
    int32_t st_chunk;   // tile gridding: coils per chunk of the coil-minor copy
    int32_t pad1;
    float beta[3];
    float inv_scale;
    int32_t ndim;
    int32_t nbatch;     // coils (batch) of this job
    int32_t mlong;      // matrix gridding: rows with more entries are summed by a whole workgroup
    int64_t npts;
    const float *coord;       // (npts, ndim) grid units
    const float4 *entries;    // (x, y, z, bits(sample)) sorted by tile
    const int32_t *bin_off;   // (ntiles + 1) offsets into entries
    const float4 *pentries;   // each sample once, sorted by the interpolation
                              // tile holding its footprint origin (3D: then
                              // by the origin cell)
    const int32_t *pbin_off;  // (interp tiles + 1) offsets into pentries
    const int32_t *mrow;      // matrix gridding: (cells + 1) row offsets of this frame
    const int2 *mnz;          // matrix gridding: (sample, weight bits) per entry
    int64_t grid_off;         // complex elements from the grid base
    int64_t grid_bstride;     // complex elements between coils
    int64_t samp_off;
    int64_t samp_bstride;
    int64_t w_off;            // offset into the weight base (-1: no weight)
    int64_t st_off;           // matrix / tile gridding: this job's coil-minor sample copy in the table's buffer
    const float4 *taps;       // tile gridding: per sample of this frame, (x taps), (y taps), (z taps),
                              // (footprint origin x, y, z as int bits, 0)
};

// Low-rank subspace arguments of a launch (ncoeff == 0: plain transform).
// Interpolation: coef = basis_adjoint (nframes, ncoeff). Gridding: coef =
// basis (ncoeff, nframes). The plan holds all frames as one trajectory of
// nframes * frame_npts samples.
struct LowRank {
    const float2 *coef = nullptr;
    int32_t ncoeff = 0;
    int32_t nframes = 1;
    int32_t frame_npts = 1;
};

struct LaunchTables {
    Job *d_jobs = nullptr;
    Item *d_interp = nullptr;
    Item *d_grid = nullptr;
    Item *d_fold = nullptr;
    int64_t *d_slot_base = nullptr;  // float offset of each fold slot in d_scratch
    float *d_scratch = nullptr;      // partial tiles of split gridding tiles
    int64_t n_interp = 0;
    int64_t n_grid = 0;
    int64_t n_fold = 0;
    int64_t scratch_floats = 0;
    int grid_coils = 1;        // coils per gridding workgroup (LDS bound)
    size_t grid_lds = 0;       // dynamic LDS bytes for the gridding kernel
    int kind_interp = 0;       // kernel variant selector
    int kind_grid = 0;
    int interp_mode = kInterpGather;
    size_t interp_lds = 0;     // dynamic LDS bytes for the staged interpolation
    int interp_coils = 1;      // coils per LDS region load (staged interpolation)
    int grid_mode = kGridSpread;  // gridding variant (matches the plans' entry order)
    // matrix gridding over a coil-minor copy of the samples: chunks of
    // st_cw coils, [chunk][sample][st_cw] per job, built by the transpose items
    Item *d_tr = nullptr;
    int64_t n_tr = 0;
    float2 *d_st = nullptr;
    int64_t st_elems = 0;
    int st_cw = 0;  // 0: the lane-per-cell matrix kernel reads the samples in place
    int st_cpl = 1; // coils per lane of the matrix kernel over the copy (1 or 2)
    int64_t device_bytes = 0;  // bytes of the device allocations above
};

}  // namespace tn

struct tn_plan {
    int device = 0;
    tn_plan_params p{};        // (z, y, x) as given; apod pointers cleared
    int32_t tile[3] = {1, 1, 1};   // (z, y, x)
    int32_t ntiles[3] = {1, 1, 1};
    int64_t ntiles_total = 1;
    int32_t itile[3] = {1, 1, 1};  // staged-interpolation tile (z, y, x)
    int32_t nitiles[3] = {1, 1, 1};
    int64_t nitiles_total = 1;
    float4 *d_pentries = nullptr;
    int32_t *d_pbin_off = nullptr;
    std::vector<int32_t> h_pbin_off;  // host copies (work-list construction)
    std::vector<int32_t> h_bin_off;
    std::vector<int32_t> h_mrow;      // host copy of d_mrow (work-list construction)
    // matrix-gridding runs per frame, keyed by the cells-per-item limit
    // (built once per plan; launch tables of any coil count reuse them; the
    // long-row threshold is the plan's grid_split, fixed for its lifetime)
    mutable std::map<int, std::vector<std::vector<tn::MatrixRun>>> mruns;
    float *d_coord = nullptr;
    float4 *d_entries = nullptr;
    int32_t *d_bin_off = nullptr;
    float *d_apod[3] = {nullptr, nullptr, nullptr};
    int64_t nentries = 0;
    int64_t footprint = -1;
    int grid_mode = tn::kGridSpread;  // effective gridding variant (entry order)
    int32_t *d_mrow = nullptr;  // matrix gridding: row offsets, (nframes * cells + 1)
    int2 *d_mnz = nullptr;      // matrix gridding: (sample, weight bits) sorted by (frame, cell)
    float2 *d_fft_tw[3] = {nullptr, nullptr, nullptr};  // pruned FFT twiddles exp(-2 pi i m / G), (x, y, z); null: none
    int64_t fft_tw_bytes = 0;
    int64_t mnnz = 0;
    float4 *d_taps = nullptr;   // tile gridding: 4 float4 per sample (taps per axis, origin)
    std::map<int64_t, tn::LaunchTables> cache;
    std::map<int64_t, tn::LaunchTables> cache_gather;  // gather-mode tables (low-rank interpolation)
    std::map<int64_t, tn::LaunchTables> cache_lowrank; // tables of the low-rank gridding (no coil-minor copy)
};

struct tn_batch {
    int device = 0;
    tn::LaunchTables tables;
    std::vector<tn_plan *> plans;
};

namespace tn {

// Error reporting (thread-local message + return code).
int fail(int code, const char *fmt, ...);
void clear_error();

// Fills job records for every frame of `plan` with the given placement.
void make_jobs(const tn_plan *plan, int64_t nbatch, int64_t grid_off,
               int64_t samp_off, int64_t w_off, std::vector<Job> &jobs);
void free_tables(LaunchTables &t);

// Kernel launchers (defined in the kernel translation units).
int launch_interp(const LaunchTables &t, const float2 *grid, float2 *samp,
                  const float *weight, hipStream_t stream);
int launch_grid(const LaunchTables &t, const float2 *samp, float2 *grid,
                const float *weight, hipStream_t stream, LowRank lr = LowRank{},
                float *scratch = nullptr);
// Device address of the fixed gridding workspace on `device` (looked up once
// per device with hipGetSymbolAddress; no allocation).
int static_scratch(int device, float **out);
int launch_interp_lowrank(const LaunchTables &t, const float2 *grid, float2 *samp,
                          const float *weight, hipStream_t stream, LowRank lr);
int launch_fold(const LaunchTables &t, float2 *grid, hipStream_t stream);
int launch_apod_pad(const tn_plan *plan, const float2 *image, float2 *grid,
                    int64_t nimg, hipStream_t stream);
int launch_crop_apod(const tn_plan *plan, const float2 *grid, float2 *image,
                     int64_t nimg, hipStream_t stream);
int build_bins(tn_plan *plan, hipStream_t stream);
// Whether register-tile gridding can serve a plan of these parameters: 3D,
// at most kTilesMaxWidth taps per axis, grids of whole tiles with at least
// two tiles per axis (a footprint then meets a tile at most once).
inline bool tiles_fit(const tn_plan_params &p) {
    if (p.ndim != 3) return false;
    for (int d = 0; d < 3; ++d) {
        const int T = kTilesT[2 - d];  // params are (z, y, x)
        if (p.width[d] > kTilesMaxWidth || p.grid[d] % T != 0 || p.grid[d] < 2 * T) return false;
    }
    return true;
}
int build_matrix(tn_plan *plan, hipStream_t stream);
// Pruned, fused oversampled-grid FFT stages (2D plans with supported lengths).
bool fft_supported(const tn_plan *plan);
int build_fft(tn_plan *plan, hipStream_t stream);
int launch_fft_forward(const tn_plan *plan, const float2 *image, float2 *grid, int64_t nimg, hipStream_t stream);
int launch_fft_adjoint(const tn_plan *plan, float2 *grid, float2 *image, int64_t nimg, hipStream_t stream);
int launch_fft_z_forward(const tn_plan *plan, const float2 *image, float2 *zout, int64_t nimg, hipStream_t stream);
int launch_fft_z_adjoint(const tn_plan *plan, const float2 *zin, float2 *image, int64_t nimg, hipStream_t stream);
int count_footprint(tn_plan *plan, hipStream_t stream);

// Variant selection shared by the table builder and the launchers.
int interp_kind(const std::vector<Job> &jobs);

}  // namespace tn

#define TN_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::tn::fail(TN_E_HIP, "%s failed: %s (%s:%d)", #call,       \
                              hipGetErrorString(e_), __FILE__, __LINE__);     \
    } while (0)
