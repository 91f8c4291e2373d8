// C ABI of tilenufft: plan lifetime, launch tables and entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tn_internal.h"

namespace tn {

namespace {
thread_local std::string g_error;
}

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

const char *last_error() { return g_error.c_str(); }

void make_jobs(const tn_plan *plan, int64_t nbatch, int64_t grid_off, int64_t samp_off,
               int64_t w_off, std::vector<Job> &jobs) {
    const tn_plan_params &p = plan->p;
    const int64_t gvol = (int64_t)p.grid[0] * p.grid[1] * p.grid[2];
    for (int f = 0; f < p.nframes; ++f) {
        Job j{};
        for (int d = 0; d < 3; ++d) {
            const int zyx = 2 - d;
            j.G[d] = p.grid[zyx];
            j.W[d] = p.width[zyx];
            j.shift[d] = p.shift[zyx];
            j.T[d] = plan->tile[zyx];
            j.NT[d] = plan->ntiles[zyx];
            j.IT[d] = plan->itile[zyx];
            j.INT_[d] = plan->nitiles[zyx];
            j.beta[d] = p.beta[zyx];
        }
        j.inv_scale = p.inv_scale;
        j.ndim = p.ndim;
        j.nbatch = (int32_t)nbatch;
        j.npts = p.npts;
        j.coord = plan->d_coord + (int64_t)f * p.npts * p.ndim;
        j.entries = plan->d_entries;
        j.bin_off = plan->d_bin_off + (int64_t)f * plan->ntiles_total;
        j.pentries = plan->d_pentries;
        j.pbin_off = plan->d_pbin_off + (int64_t)f * plan->nitiles_total;
        j.mrow = plan->d_mrow ? plan->d_mrow + (int64_t)f * gvol : nullptr;
        j.mnz = plan->d_mnz;
        j.taps = plan->d_taps ? plan->d_taps + (int64_t)f * p.npts * 4 : nullptr;
        j.mlong = p.grid_split == 0  ? (p.ndim == 3 ? kMatrixLongRow3d : kMatrixLongRow)
                  : p.grid_split < 0 ? matrix_stage(p.ndim)
                                     : std::min(p.grid_split, matrix_stage(p.ndim));
        // deepmr layout (batch, nframes, ...): frame f of coil b
        j.grid_off = grid_off + (int64_t)f * gvol;
        j.grid_bstride = (int64_t)p.nframes * gvol;
        j.samp_off = samp_off + (int64_t)f * p.npts;
        j.samp_bstride = (int64_t)p.nframes * p.npts;
        j.w_off = w_off < 0 ? -1 : w_off + (int64_t)f * p.npts;
        jobs.push_back(j);
    }
}

namespace {

int chunk_for(int requested, int64_t nbatch, int fallback) {
    int c = requested > 0 ? requested : fallback;
    return (int)std::max<int64_t>(1, std::min<int64_t>(c, nbatch));
}

}  // namespace

void free_tables(LaunchTables &t) {
    if (t.d_jobs) (void)hipFree(t.d_jobs);
    if (t.d_interp) (void)hipFree(t.d_interp);
    if (t.d_grid) (void)hipFree(t.d_grid);
    if (t.d_fold) (void)hipFree(t.d_fold);
    if (t.d_slot_base) (void)hipFree(t.d_slot_base);
    if (t.d_scratch) (void)hipFree(t.d_scratch);
    if (t.d_tr) (void)hipFree(t.d_tr);
    if (t.d_st) (void)hipFree(t.d_st);
    t = LaunchTables{};
}

// Launch tables go to the device on a library-owned non-blocking stream of
// the current device (default priority), waited for before the tables are
// used: a hipMemcpy would run on the null stream and first wait for all the
// work the caller has queued there.
static std::mutex g_upload_mu;
static hipStream_t g_upload[kMaxDevices] = {};

static int upload_stream(hipStream_t *out) {
    int dev = 0;
    TN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TN_E_RANGE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_upload_mu);
    if (!g_upload[dev]) TN_HIP(hipStreamCreateWithFlags(&g_upload[dev], hipStreamNonBlocking));
    *out = g_upload[dev];
    return TN_OK;
}

static int upload_tables(const std::vector<Job> &jobs, const std::vector<Item> &interp,
                         const std::vector<Item> &grid, const std::vector<Item> &fold,
                         const std::vector<int64_t> &slot_base, LaunchTables &t) {
    t.n_interp = (int64_t)interp.size();
    t.n_grid = (int64_t)grid.size();
    t.n_fold = (int64_t)fold.size();
    if (t.n_interp >= INT32_MAX || t.n_grid >= INT32_MAX)
        return fail(TN_E_RANGE, "work list too long");
    TN_HIP(hipMalloc(&t.d_jobs, sizeof(Job) * std::max<size_t>(jobs.size(), 1)));
    TN_HIP(hipMalloc(&t.d_interp, sizeof(Item) * std::max<size_t>(interp.size(), 1)));
    TN_HIP(hipMalloc(&t.d_grid, sizeof(Item) * std::max<size_t>(grid.size(), 1)));
    TN_HIP(hipMalloc(&t.d_fold, sizeof(Item) * std::max<size_t>(fold.size(), 1)));
    TN_HIP(hipMalloc(&t.d_slot_base, sizeof(int64_t) * std::max<size_t>(slot_base.size(), 1)));
    TN_HIP(hipMalloc(&t.d_scratch, sizeof(float) * std::max<int64_t>(t.scratch_floats, 1)));
    hipStream_t us = nullptr;
    int rc = upload_stream(&us);
    if (rc) return rc;
    if (!fold.empty())
        TN_HIP(hipMemcpyAsync(t.d_fold, fold.data(), sizeof(Item) * fold.size(), hipMemcpyHostToDevice, us));
    if (!slot_base.empty())
        TN_HIP(hipMemcpyAsync(t.d_slot_base, slot_base.data(), sizeof(int64_t) * slot_base.size(),
                              hipMemcpyHostToDevice, us));
    if (!jobs.empty())
        TN_HIP(hipMemcpyAsync(t.d_jobs, jobs.data(), sizeof(Job) * jobs.size(), hipMemcpyHostToDevice, us));
    if (!interp.empty())
        TN_HIP(hipMemcpyAsync(t.d_interp, interp.data(), sizeof(Item) * interp.size(), hipMemcpyHostToDevice,
                              us));
    if (!grid.empty())
        TN_HIP(hipMemcpyAsync(t.d_grid, grid.data(), sizeof(Item) * grid.size(), hipMemcpyHostToDevice, us));
    TN_HIP(hipStreamSynchronize(us));
    t.device_bytes = (int64_t)(sizeof(Job) * std::max<size_t>(jobs.size(), 1) +
                               sizeof(Item) * (std::max<size_t>(interp.size(), 1) + std::max<size_t>(grid.size(), 1) +
                                               std::max<size_t>(fold.size(), 1)) +
                               sizeof(int64_t) * std::max<size_t>(slot_base.size(), 1) +
                               sizeof(float) * std::max<int64_t>(t.scratch_floats, 1));
    return TN_OK;
}

// Matrix-gridding runs of frame f, cached in the plan per cells-per-item
// limit (matrix_runs_of).
static const std::vector<MatrixRun> &matrix_runs(const tn_plan *pl, int64_t f, int64_t ncell, int cpi, int mlong) {
    auto &per_frame = pl->mruns[cpi];
    if (per_frame.empty()) per_frame.resize(pl->p.nframes);
    std::vector<MatrixRun> &out = per_frame[f];
    if (out.empty() && ncell > 0)
        matrix_runs_of(pl->h_mrow.data() + f * ncell, ncell, cpi, mlong, out, matrix_stage(pl->p.ndim));
    return out;
}

// Work lists. Interpolation items are (job, sample block, coil range);
// gridding items are (job, tile, coil range). `plan_of[j]` is the plan job j
// came from (for its chunk preferences).
static int build_tables_for(const std::vector<Job> &jobs, const std::vector<const tn_plan *> &plan_of,
                            LaunchTables &t, bool force_gather = false, bool no_copy = false) {
    t.kind_interp = interp_kind(jobs);
    t.kind_grid = t.kind_interp;
    int64_t max_plane = 1, max_region = 1, max_rows = 1, max_cells = 1;
    for (const Job &j : jobs) {
        max_plane = std::max<int64_t>(max_plane, (int64_t)j.T[2] * j.T[1] * (j.T[0] + kGridPad) + kWave);
        max_region = std::max<int64_t>(max_region, (int64_t)(j.IT[0] + j.W[0] - 1) *
                                                       (j.IT[1] + j.W[1] - 1) * (j.IT[2] + j.W[2] - 1));
        max_rows = std::max<int64_t>(max_rows, (int64_t)(j.IT[1] + j.W[1] - 1) * (j.IT[2] + j.W[2] - 1));
        max_cells = std::max<int64_t>(max_cells, (int64_t)j.G[0] * j.G[1] * j.G[2]);
    }
    const int wc = kind_width(t.kind_grid);
    const int gmode = plan_of.empty() ? kGridSpread : plan_of[0]->grid_mode;
    for (const tn_plan *pl : plan_of)
        if (pl->grid_mode != gmode) return fail(TN_E_ARG, "batched plans must share the gridding variant");
    t.grid_mode = gmode;
    const int gndim = kind_ndim(t.kind_grid);
    // LDS sizing of the LDS gridding kernels (register tiles: of the LDS-atomic
    // kernel that serves their low-rank launches)
    const int lmode = gmode == kGridTiles ? kGridSpread : gmode;
    if (gmode != kGridMatrix && grid_lds_bytes(1, max_plane, wc, lmode, gndim) > kGridLdsBudget)
        return fail(TN_E_ARG, "gridding tile too large for LDS (%zu bytes for one coil)",
                    grid_lds_bytes(1, max_plane, wc, lmode, gndim));
    int lds_coils = 1;
    while (lds_coils < (gmode == kGridRuns ? kRunsMaxCoils : 64) &&
           grid_lds_bytes(lds_coils + 1, max_plane, wc, lmode, gndim) <= kGridLdsBudget)
        ++lds_coils;
    const size_t region_coil = (size_t)max_region * 2 * sizeof(float);
    // the staged interpolation's row table (one int per (z, y) row of a
    // region) follows the chunk's regions in LDS
    const size_t rows_bytes = ((size_t)max_rows * sizeof(int) + 15) / 16 * 16;
    const int region_coils =
        (int)std::max<size_t>(1, (kInterpLdsBudget - std::min(rows_bytes, kInterpLdsBudget)) / region_coil);

    int mode = plan_of.empty() ? kInterpGather : plan_of[0]->p.interp_mode;
    if (mode == kInterpAuto) mode = kInterpStaged;
    if (mode == kInterpStaged && region_coil + rows_bytes > kInterpLdsBudget) mode = kInterpGather;
    if (mode == kInterpStaged && max_cells >= INT32_MAX) mode = kInterpGather;  // 32-bit row offsets
    if (kind_width(t.kind_interp) > kStagedMaxWidth) mode = kInterpGather;  // wide kernels gather
    if (force_gather) mode = kInterpGather;
    t.interp_mode = mode;

    int max_gc = 1, max_ic = 1;
    int st_cw = 0;
    bool st_none = false;
    // matrix gridding: record width of the coil-minor copy (the widest chunk
    // of the table) and the cells per item the kernel assumes for it
    int mat_cpi = kBlock, mat_cpl = 1;
    if (gmode == kGridMatrix) {
        int cwmax = 0;
        bool inplace = false;
        for (size_t ji = 0; ji < jobs.size(); ++ji) {
            const int mc = std::min(chunk_for(plan_of[ji]->p.grid_chunk, jobs[ji].nbatch, kMatrixChunk), kMatrixCoils);
            if (mc < 2) inplace = true;
            cwmax = std::max(cwmax, mc <= 2 ? 2 : mc <= 4 ? 4 : mc <= 8 ? 8 : 16);
        }
        // coils per lane: two for 16-coil records, and for 4-8 coil records
        // in 2D (half the gather instructions); one for 4-8 coil records in
        // 3D, where twice the cells per wave measured slower (divergence)
        mat_cpl = (cwmax == 16 || (cwmax >= 4 && kind_ndim(t.kind_grid) == 2)) ? 2 : 1;
        if (!inplace && cwmax > 0) mat_cpi = std::min(kBlock, kBlock / (cwmax / mat_cpl) * kMatrixPasses);
    }
    std::vector<Item> interp, grid, fold, tr;
    std::vector<int64_t> slot_base;
    int64_t scratch = 0;
    for (size_t ji = 0; ji < jobs.size(); ++ji) {
        const Job &j = jobs[ji];
        const tn_plan *pl = plan_of[ji];
        // automatic coil chunk: as many as fit in LDS; for register-run gridding
        // one coil per wave (more coils per item cost occupancy, which hides
        // the latency of its LDS read-add-write chains)
        const int gc_auto = gmode == kGridRuns    ? std::min(lds_coils, kBlock / kWave)
                            : gmode == kGridTiles ? std::min(lds_coils, kTilesCoils)
                                                  : lds_coils;
        const int gc = std::min(chunk_for(pl->p.grid_chunk, j.nbatch, gc_auto),
                                gmode == kGridTiles ? std::min(lds_coils, kTilesCoils) : lds_coils);
        if (mode == kInterpStaged) {
            // one item per (tile, segment) covers every coil; the region is
            // loaded `ic` coils at a time
            const int ic_auto = std::min(region_coils, j.ndim == 3 ? kInterpChunk3d : kInterpChunk2d);
            const int ic = std::min(chunk_for(pl->p.interp_chunk, j.nbatch, ic_auto), region_coils);
            max_ic = std::max(max_ic, ic);
            const int64_t nt = (int64_t)j.INT_[0] * j.INT_[1] * j.INT_[2];
            const int64_t f = (j.pbin_off - pl->d_pbin_off) / std::max<int64_t>(pl->nitiles_total, 1);
            const int32_t *off = pl->h_pbin_off.data() + f * pl->nitiles_total;
            for (int64_t tl = 0; tl < nt; ++tl) {
                for (int32_t e0 = off[tl]; e0 < off[tl + 1]; e0 += interp_segment(j.ndim)) {
                    const int32_t e1 = std::min<int32_t>(off[tl + 1], e0 + interp_segment(j.ndim));
                    interp.push_back(Item{(int)ji, (int)tl, 0, j.nbatch, e0, e1, 0, 0});
                }
            }
        } else {
            const int ic = chunk_for(pl->p.interp_chunk, j.nbatch, j.nbatch >= 8 ? (int)((j.nbatch + 1) / 2) : j.nbatch);
            const int64_t nblk = (j.npts + kInterpBlockPts - 1) / kInterpBlockPts;
            for (int64_t b = 0; b < nblk; ++b)
                for (int c0 = 0; c0 < j.nbatch; c0 += ic)
                    interp.push_back(Item{(int)ji, (int)b, c0, std::min(j.nbatch, c0 + ic), 0, 0, 0, 0});
        }
        if (gmode == kGridMatrix) {
            // Items are runs of consecutive cells of one frame (unit = first
            // cell, e1 = cell count | entry count << 16, slot = first entry,
            // e0 = coil chunk): at most mat_cpi cells (what the coil-minor
            // kernel of the table's record width assumes; kBlock for the
            // lane-per-cell kernel) and at most kMatrixStage entries, so that
            // an item's entries (contiguous in the CSR order) fit the kernel's
            // LDS stage. Rows longer than mlong end the run and are summed by
            // workgroups of their own in parts of kMatrixPart entries (part 1,
            // unit = cell, e1 = part); rows of several parts leave partial
            // records folded in part order. Long rows are queued first.
            const int mc = std::min(chunk_for(pl->p.grid_chunk, j.nbatch, kMatrixChunk), kMatrixCoils);
            max_gc = std::max(max_gc, mc);
            const int cw = mc <= 2 ? 2 : mc <= 4 ? 4 : mc <= 8 ? 8 : 16;
            const int cpi = mat_cpi;
            const int64_t ncell = (int64_t)j.G[0] * j.G[1] * j.G[2];
            const int64_t f = (j.mrow - pl->d_mrow) / std::max<int64_t>(ncell, 1);
            const std::vector<MatrixRun> &runs = matrix_runs(pl, f, ncell, cpi, j.mlong);
            matrix_items((int)ji, runs, j.nbatch, mc, grid, fold, scratch);
            // coil-minor sample copy: chunks of mc coils padded to cw lanes
            // (not for tables whose gridding reads the samples in place:
            // low-rank gridding, or tables only used to interpolate)
            if (no_copy || force_gather) {
                st_none = true;
            } else if (mc >= 2) {
                st_cw = std::max(st_cw, cw);
                const int64_t pblk = (j.npts + kBlock - 1) / kBlock;
                for (int c0 = 0; c0 < j.nbatch; c0 += mc)
                    for (int64_t b = 0; b < pblk; ++b)
                        tr.push_back(Item{(int)ji, (int)b, c0, std::min(j.nbatch, c0 + mc), c0 / mc, 0, 0, -1});
            } else {
                st_none = true;  // a one-coil chunk: the lane-per-cell kernel reads in place
            }
            continue;
        }
        max_gc = std::max(max_gc, gc);
        if (gmode == kGridTiles) {
            // coil-minor sample copy in chunks of gc coils padded to cw
            // (scalar-loaded records); none for low-rank or gather tables,
            // whose gridding takes the LDS-atomic kernel over the same bins
            if (no_copy || force_gather) {
                st_none = true;
            } else {
                st_cw = std::max(st_cw, gc <= 2 ? 2 : gc <= 4 ? 4 : 8);
                const int64_t pblk = (j.npts + kBlock - 1) / kBlock;
                for (int c0 = 0; c0 < j.nbatch; c0 += gc)
                    for (int64_t b = 0; b < pblk; ++b)
                        tr.push_back(Item{(int)ji, (int)b, c0, std::min(j.nbatch, c0 + gc), c0 / gc, 0, 0, -1});
            }
        }
        const int64_t ntiles = (int64_t)j.NT[0] * j.NT[1] * j.NT[2];
        const int64_t fg = (j.bin_off - pl->d_bin_off) / std::max<int64_t>(pl->ntiles_total, 1);
        const int32_t *goff = pl->h_bin_off.data() + fg * pl->ntiles_total;
        const int split = pl->p.grid_split == 0 ? (gmode == kGridTiles ? kTilesSplit : kGridSplitDefault)
                          : pl->p.grid_split < 0 ? INT32_MAX : pl->p.grid_split;
        const int64_t cells = (int64_t)j.T[0] * j.T[1] * j.T[2];  // partial size per coil
        for (int64_t tl = 0; tl < ntiles; ++tl) {
            const int32_t b0 = goff[tl], b1 = goff[tl + 1];
            const int nparts = (int)std::max<int64_t>(1, ((int64_t)(b1 - b0) + split - 1) / split);
            for (int c0 = 0; c0 < j.nbatch; c0 += gc) {
                const int c1 = std::min(j.nbatch, c0 + gc);
                if (nparts == 1) {
                    grid.push_back(Item{(int)ji, (int)tl, c0, c1, b0, b1, 0, -1});
                    continue;
                }
                const int slot = (int)slot_base.size();
                slot_base.push_back(scratch);
                scratch += (int64_t)nparts * 2 * (c1 - c0) * cells;
                const int32_t per = (b1 - b0 + nparts - 1) / nparts;
                for (int q = 0; q < nparts; ++q) {
                    const int32_t e0 = b0 + q * per, e1 = std::min(b1, e0 + per);
                    grid.push_back(Item{(int)ji, (int)tl, c0, c1, e0, e1, q, slot});
                }
                tile_fold_items((int)ji, (int)tl, c0, c1, slot, nparts, (int64_t)(c1 - c0) * cells, kBlock, fold);
            }
        }
    }
    t.scratch_floats = scratch;
    t.grid_coils = max_gc;
    std::vector<Job> jobs_st(jobs);
    if ((gmode == kGridMatrix || gmode == kGridTiles) && st_cw > 0 && !st_none) {
        // every job gets its copy at st_cw coils per record
        int64_t off = 0;
        for (size_t ji = 0; ji < jobs.size(); ++ji) {
            const int mc = gmode == kGridTiles ? std::min(chunk_for(plan_of[ji]->p.grid_chunk, jobs[ji].nbatch,
                                                                    std::min(lds_coils, kTilesCoils)),
                                                          std::min(lds_coils, kTilesCoils))
                                               : std::min(chunk_for(plan_of[ji]->p.grid_chunk, jobs[ji].nbatch,
                                                                    kMatrixChunk),
                                                          kMatrixCoils);
            jobs_st[ji].st_off = off;
            jobs_st[ji].st_chunk = mc;
            off += (int64_t)((jobs[ji].nbatch + mc - 1) / mc) * jobs[ji].npts * st_cw;
        }
        t.st_cw = st_cw;
        t.st_cpl = mat_cpl;
        t.st_elems = off;
        t.n_tr = (int64_t)tr.size();
        TN_HIP(hipMalloc(&t.d_tr, sizeof(Item) * std::max<size_t>(tr.size(), 1)));
        if (!tr.empty()) {
            hipStream_t us = nullptr;
            int rc = upload_stream(&us);
            if (rc) return rc;
            TN_HIP(hipMemcpyAsync(t.d_tr, tr.data(), sizeof(Item) * tr.size(), hipMemcpyHostToDevice, us));
            TN_HIP(hipStreamSynchronize(us));
        }
        TN_HIP(hipMalloc(&t.d_st, sizeof(float2) * std::max<int64_t>(off, 1)));
    }
    const int64_t copy_bytes = t.d_st ? (int64_t)(sizeof(Item) * std::max<size_t>(tr.size(), 1) +
                                                  sizeof(float2) * std::max<int64_t>(t.st_elems, 1))
                                      : 0;
    t.grid_lds = gmode == kGridMatrix ? 0 : grid_lds_bytes(max_gc, max_plane, wc, lmode, gndim);
    t.interp_lds = mode == kInterpStaged ? region_coil * (size_t)max_ic + rows_bytes : 0;
    t.interp_coils = max_ic;
    const int rc = upload_tables(jobs_st, interp, grid, fold, slot_base, t);
    t.device_bytes += copy_bytes;
    return rc;
}

}  // namespace tn

using namespace tn;

namespace {

int validate(const tn_plan_params *p) {
    if (!p) return fail(TN_E_ARG, "params is NULL");
    if (p->ndim != 2 && p->ndim != 3) return fail(TN_E_ARG, "ndim must be 2 or 3 (got %d)", p->ndim);
    if (p->nframes < 1) return fail(TN_E_ARG, "nframes must be >= 1");
    if (p->npts < 0 || p->npts >= INT32_MAX) return fail(TN_E_RANGE, "npts out of range");
    for (int d = 0; d < 3; ++d) {
        const bool used = d >= 3 - p->ndim;
        if (!used) {
            if (p->grid[d] != 1 || p->shape[d] != 1 || p->width[d] != 1)
                return fail(TN_E_ARG, "unused axis %d must have grid = shape = width = 1", d);
            continue;
        }
        if (p->shape[d] < 1 || p->grid[d] < p->shape[d])
            return fail(TN_E_ARG, "axis %d: need 1 <= shape (%d) <= grid (%d)", d, p->shape[d], p->grid[d]);
        if (p->width[d] < 1 || p->width[d] > kMaxWidth || p->width[d] > p->grid[d])
            return fail(TN_E_ARG, "axis %d: width %d unsupported (1..%d, <= grid)", d, p->width[d], kMaxWidth);
        if (p->shift[d] < 0 || p->shift[d] >= p->grid[d])
            return fail(TN_E_ARG, "axis %d: shift %d outside [0, grid)", d, p->shift[d]);
        if (p->tile[d] < 0 || p->itile[d] < 0) return fail(TN_E_ARG, "axis %d: negative tile", d);
    }
    if ((int64_t)p->grid[0] * p->grid[1] * p->grid[2] >= INT32_MAX)
        return fail(TN_E_RANGE, "grid too large");
    if (p->interp_mode < 0 || p->interp_mode > 2) return fail(TN_E_ARG, "interp_mode must be 0, 1 or 2");
    if (p->grid_mode < 0 || p->grid_mode > 4) return fail(TN_E_ARG, "grid_mode must be 0, 1, 2, 3 or 4");
    return TN_OK;
}

void choose_tiles(tn_plan *pl) {
    const tn_plan_params &p = pl->p;
    int def[3];
    if (p.ndim == 2) {
        def[0] = 1; def[1] = 32; def[2] = 32;
    } else {
        def[0] = 16; def[1] = 16; def[2] = 16;
    }
    int idef[3];  // default staged-interpolation tiles (z, y, x): measured fastest with the default coil chunks
    if (p.ndim == 2) {
        idef[0] = 1; idef[1] = 16; idef[2] = 16;
    } else {
        idef[0] = 8; idef[1] = 8; idef[2] = 16;
    }
    pl->ntiles_total = 1;
    pl->nitiles_total = 1;
    for (int d = 0; d < 3; ++d) {
        int t = p.tile[d] > 0 ? p.tile[d] : def[d];
        t = std::min(t, p.grid[d]);
        pl->tile[d] = t;
        pl->ntiles[d] = (p.grid[d] + t - 1) / t;
        pl->ntiles_total *= pl->ntiles[d];
        int it = p.itile[d] > 0 ? p.itile[d] : idef[d];
        it = std::min(it, p.grid[d]);
        pl->itile[d] = it;
        pl->nitiles[d] = (p.grid[d] + it - 1) / it;
        pl->nitiles_total *= pl->nitiles[d];
    }
}

void destroy(tn_plan *pl) {
    if (!pl) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(pl->device);
    for (auto &kv : pl->cache) free_tables(kv.second);
    for (auto &kv : pl->cache_gather) free_tables(kv.second);
    for (auto &kv : pl->cache_lowrank) free_tables(kv.second);
    if (pl->d_coord) (void)hipFree(pl->d_coord);
    if (pl->d_entries) (void)hipFree(pl->d_entries);
    if (pl->d_bin_off) (void)hipFree(pl->d_bin_off);
    if (pl->d_pentries) (void)hipFree(pl->d_pentries);
    if (pl->d_pbin_off) (void)hipFree(pl->d_pbin_off);
    if (pl->d_mrow) (void)hipFree(pl->d_mrow);
    if (pl->d_mnz) (void)hipFree(pl->d_mnz);
    if (pl->d_taps) (void)hipFree(pl->d_taps);
    for (int a = 0; a < 3; ++a)
        if (pl->d_fft_tw[a]) (void)hipFree(pl->d_fft_tw[a]);
    for (int d = 0; d < 3; ++d)
        if (pl->d_apod[d]) (void)hipFree(pl->d_apod[d]);
    (void)hipSetDevice(cur);
    delete pl;
}

enum TableUse { kTablesPlain = 0, kTablesGather = 1, kTablesLowRank = 2 };

tn::LaunchTables *plan_tables(tn_plan *plan, int64_t nbatch, int *rc, int use = kTablesPlain) {
    auto &cache = use == kTablesGather ? plan->cache_gather : use == kTablesLowRank ? plan->cache_lowrank : plan->cache;
    auto it = cache.find(nbatch);
    if (it != cache.end()) {
        *rc = TN_OK;
        return &it->second;
    }
    std::vector<Job> jobs;
    make_jobs(plan, nbatch, 0, 0, 0, jobs);
    std::vector<const tn_plan *> owners(jobs.size(), plan);
    LaunchTables t;
    *rc = build_tables_for(jobs, owners, t, use == kTablesGather, use == kTablesLowRank);
    if (*rc != TN_OK) {
        free_tables(t);
        return nullptr;
    }
    return &(cache[nbatch] = t);
}

// Shared checks of the low-rank entry points: one-frame plan whose samples
// are `npts / frame_npts` frames of `frame_npts` samples.
int check_lowrank(const tn_plan *plan, int64_t nbatch, int32_t ncoeff, int32_t frame_npts,
                  const void *coef, LowRank &lr) {
    if (plan->p.nframes != 1)
        return fail(TN_E_ARG, "low-rank transforms take a plan of all frames as one trajectory");
    if (ncoeff < 1) return fail(TN_E_ARG, "ncoeff must be >= 1");
    if (frame_npts < 1 || plan->p.npts % frame_npts != 0)
        return fail(TN_E_ARG, "frame_npts (%d) must divide the plan's %lld samples", frame_npts,
                    (long long)plan->p.npts);
    if (!coef) return fail(TN_E_ARG, "basis is NULL");
    if (nbatch * (int64_t)ncoeff >= INT32_MAX) return fail(TN_E_RANGE, "nbatch * ncoeff too large");
    lr.coef = (const float2 *)coef;
    lr.ncoeff = ncoeff;
    lr.nframes = (int32_t)(plan->p.npts / frame_npts);
    lr.frame_npts = frame_npts;
    return TN_OK;
}

// Checks that the current device is the plan's device.
int check_device(int device) {
    int cur = -1;
    TN_HIP(hipGetDevice(&cur));
    if (cur != device)
        return fail(TN_E_ARG, "plan lives on device %d but device %d is current", device, cur);
    return TN_OK;
}

}  // namespace

extern "C" {

const char *tn_last_error(void) { return tn::last_error(); }

const char *tn_version(void) { return "tilenufft 0.1 (gfx950)"; }

int tn_plan_create(const tn_plan_params *params, const float *coord_dev, tn_stream stream,
                   tn_plan **out) {
    clear_error();
    if (!out) return fail(TN_E_ARG, "out is NULL");
    *out = nullptr;
    int rc = validate(params);
    if (rc) return rc;
    if (!coord_dev && params->npts > 0) return fail(TN_E_ARG, "coord is NULL");
    hipStream_t s = (hipStream_t)stream;
    tn_plan *pl = new tn_plan();
    pl->p = *params;
    for (int d = 0; d < 3; ++d) pl->p.apod[d] = nullptr;
    if (hipGetDevice(&pl->device) != hipSuccess) {
        delete pl;
        return fail(TN_E_HIP, "hipGetDevice failed");
    }
    choose_tiles(pl);
    {
        const int taps = params->width[1] * params->width[2];  // one z plane of the footprint
        const bool runs_ok = taps <= kRunsMaxTaps;
        if (params->grid_mode == kGridRuns && !runs_ok) {
            delete pl;
            return fail(TN_E_ARG, "grid_mode 2 supports at most %d taps per z plane (got %d)", kRunsMaxTaps, taps);
        }
        // automatic: the sparse gridding matrix when it fits kMatrixAutoBytes
        // (8 bytes per sample and tap), else register tiles (3D), else
        // register runs when the taps of a z plane fit the registers, else
        // per-sample LDS atomics
        const int64_t nnz = (int64_t)params->nframes * params->npts * params->width[0] * params->width[1] *
                            params->width[2];
        const int64_t nrows = (int64_t)params->nframes * params->grid[0] * params->grid[1] * params->grid[2];
        const bool matrix_ok = nnz < INT32_MAX && (uint64_t)nnz * sizeof(int2) <= kMatrixAutoBytes &&
                               nrows < (int64_t)UINT32_MAX && params->npts <= ((int64_t)1 << 28);
        // then register tiles where they fit (3D)
        const bool tiles_ok = tiles_fit(*params);
        if (params->grid_mode == kGridTiles && !tiles_ok) {
            delete pl;
            return fail(TN_E_ARG,
                        "grid_mode 4 (register tiles) needs a 3D plan with at most %d taps per axis and grid "
                        "lengths that are multiples of the %dx%dx%d tile with at least two tiles per axis",
                        kTilesMaxWidth, kTilesT[0], kTilesT[1], kTilesT[2]);
        }
        pl->grid_mode = params->grid_mode == kGridSpread   ? kGridSpread
                        : params->grid_mode == kGridMatrix ? kGridMatrix
                        : params->grid_mode == kGridRuns   ? kGridRuns
                        : params->grid_mode == kGridTiles  ? kGridTiles
                        : matrix_ok                        ? kGridMatrix
                        : tiles_ok                         ? kGridTiles
                        : runs_ok                          ? kGridRuns
                                                           : kGridSpread;
        if (pl->grid_mode == kGridTiles) {  // the tile the kernel is built for, whatever was asked
            pl->ntiles_total = 1;
            for (int d = 0; d < 3; ++d) {
                pl->tile[d] = kTilesT[2 - d];
                pl->ntiles[d] = params->grid[d] / pl->tile[d];
                pl->ntiles_total *= pl->ntiles[d];
            }
        }
        // kernels wider than kStagedMaxWidth grid only through the matrix
        const int wmax = std::max(params->width[0], std::max(params->width[1], params->width[2]));
        if (wmax > kStagedMaxWidth && pl->grid_mode != kGridMatrix) {
            delete pl;
            return fail(TN_E_ARG,
                        "width %d: kernels wider than %d grid through the gridding matrix (grid_mode 3), "
                        "which needs %lld entries here (at most %lld bytes)",
                        wmax, kStagedMaxWidth, (long long)nnz, (long long)kMatrixAutoBytes);
        }
    }
    const int64_t ncoord = (int64_t)params->nframes * params->npts * params->ndim;
    auto cleanup = [&](int code) {
        destroy(pl);
        return code;
    };
    if (hipMalloc(&pl->d_coord, sizeof(float) * std::max<int64_t>(ncoord, 1)) != hipSuccess)
        return cleanup(fail(TN_E_NOMEM, "hipMalloc of %lld coordinates failed", (long long)ncoord));
    if (ncoord > 0 &&
        hipMemcpyAsync(pl->d_coord, coord_dev, sizeof(float) * ncoord, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
        return cleanup(fail(TN_E_HIP, "coordinate copy failed"));
    for (int d = 0; d < 3; ++d) {
        if (!params->apod[d] || params->width[d] == 1) continue;
        const size_t n = (size_t)params->shape[d];
        if (hipMalloc(&pl->d_apod[d], sizeof(float) * n) != hipSuccess)
            return cleanup(fail(TN_E_NOMEM, "hipMalloc of apodization failed"));
        if (hipMemcpyAsync(pl->d_apod[d], params->apod[d], sizeof(float) * n, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return cleanup(fail(TN_E_HIP, "apodization upload failed"));
    }
    rc = build_bins(pl, s);
    if (rc) return cleanup(rc);
    rc = build_fft(pl, s);
    if (rc) return cleanup(rc);
    *out = pl;
    return TN_OK;
}

int tn_plan_destroy(tn_plan *plan) {
    clear_error();
    destroy(plan);
    return TN_OK;
}

int64_t tn_plan_device_bytes(const tn_plan *plan) {
    if (!plan) return 0;
    const tn_plan_params &p = plan->p;
    const int64_t total = (int64_t)p.nframes * p.npts;
    int64_t b = sizeof(float) * std::max<int64_t>(total * p.ndim, 1);                     // coordinates
    b += sizeof(float4) * std::max<int64_t>(plan->nentries, 1);                            // gridding bins
    b += sizeof(int32_t) * (plan->ntiles_total * p.nframes + 1);
    b += sizeof(float4) * std::max<int64_t>(total, 1);                                     // interpolation bins
    b += sizeof(int32_t) * (plan->nitiles_total * p.nframes + 1);
    if (plan->d_mrow)                                                                      // gridding matrix
        b += sizeof(int32_t) * ((int64_t)p.grid[0] * p.grid[1] * p.grid[2] * p.nframes + 1) +
             sizeof(int2) * std::max<int64_t>(plan->mnnz, 1);
    if (plan->d_taps) b += 4 * sizeof(float4) * total;                                    // tile gridding taps
    for (int d = 0; d < 3; ++d)
        if (plan->d_apod[d]) b += sizeof(float) * p.shape[d];
    b += plan->fft_tw_bytes;
    for (const auto &kv : plan->cache) b += kv.second.device_bytes;
    for (const auto &kv : plan->cache_gather) b += kv.second.device_bytes;
    for (const auto &kv : plan->cache_lowrank) b += kv.second.device_bytes;
    return b;
}

int64_t tn_plan_entries(const tn_plan *plan) { return plan ? plan->nentries : -1; }

int tn_plan_grid_mode(const tn_plan *plan) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    return plan->grid_mode;
}

int64_t tn_plan_footprint_cells(const tn_plan *plan_c) {
    clear_error();
    tn_plan *plan = const_cast<tn_plan *>(plan_c);
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (plan->footprint < 0) {
        int rc = check_device(plan->device);
        if (rc) return rc;
        rc = count_footprint(plan, nullptr);
        if (rc) return rc;
    }
    return plan->footprint;
}

int tn_plan_tile(const tn_plan *plan, int32_t tile_out[3]) {
    clear_error();
    if (!plan || !tile_out) return fail(TN_E_ARG, "NULL argument");
    for (int d = 0; d < 3; ++d) tile_out[d] = plan->tile[d];
    return TN_OK;
}

int tn_plan_prepare(tn_plan *plan, int64_t nbatch) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    int rc = check_device(plan->device);
    if (rc) return rc;
    plan_tables(plan, nbatch, &rc);
    return rc;
}

int tn_interp(tn_plan *plan, const void *grid, void *samp, int64_t nbatch, const float *weight,
              tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    if ((!grid || !samp) && plan->p.npts > 0) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    LaunchTables *t = plan_tables(plan, nbatch, &rc);
    if (!t) return rc;
    return launch_interp(*t, (const float2 *)grid, (float2 *)samp, weight, (hipStream_t)stream);
}

int tn_grid(tn_plan *plan, const void *samp, void *grid, int64_t nbatch, const float *weight,
            tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    if (!grid || (!samp && plan->p.npts > 0)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    LaunchTables *t = plan_tables(plan, nbatch, &rc);
    if (!t) return rc;
    return launch_grid(*t, (const float2 *)samp, (float2 *)grid, weight, (hipStream_t)stream);
}

int tn_grid_static(tn_plan *plan, const void *samp, void *grid, int64_t nbatch, const float *weight,
                   tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    if (!grid || (!samp && plan->p.npts > 0)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    auto it = plan->cache.find(nbatch);
    if (it == plan->cache.end())
        return fail(TN_E_ARG, "no launch tables for %lld coils: call tn_plan_prepare first", (long long)nbatch);
    const LaunchTables &t = it->second;
    if (t.scratch_floats > kStaticScratchFloats)
        return fail(TN_E_RANGE, "gridding needs %lld workspace floats, the fixed workspace holds %lld",
                    (long long)t.scratch_floats, (long long)kStaticScratchFloats);
    float *ws = nullptr;
    rc = static_scratch(plan->device, &ws);
    if (rc) return rc;
    return launch_grid(t, (const float2 *)samp, (float2 *)grid, weight, (hipStream_t)stream, LowRank{}, ws);
}

int tn_interp_lowrank(tn_plan *plan, const void *grid, void *samp, int64_t nbatch, int32_t ncoeff,
                      int32_t frame_npts, const void *basis_adjoint, const float *weight,
                      tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    if ((!grid || !samp) && plan->p.npts > 0) return fail(TN_E_ARG, "NULL data pointer");
    LowRank lr;
    int rc = check_lowrank(plan, nbatch, ncoeff, frame_npts, basis_adjoint, lr);
    if (rc) return rc;
    rc = check_device(plan->device);
    if (rc) return rc;
    LaunchTables *t = plan_tables(plan, nbatch, &rc, kTablesGather);
    if (!t) return rc;
    return launch_interp_lowrank(*t, (const float2 *)grid, (float2 *)samp, weight, (hipStream_t)stream, lr);
}

int tn_grid_lowrank(tn_plan *plan, const void *samp, void *grid, int64_t nbatch, int32_t ncoeff,
                    int32_t frame_npts, const void *basis, const float *weight, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nbatch < 1 || nbatch >= INT32_MAX) return fail(TN_E_ARG, "nbatch must be >= 1");
    if (!grid || (!samp && plan->p.npts > 0)) return fail(TN_E_ARG, "NULL data pointer");
    LowRank lr;
    int rc = check_lowrank(plan, nbatch, ncoeff, frame_npts, basis, lr);
    if (rc) return rc;
    rc = check_device(plan->device);
    if (rc) return rc;
    // accumulator coils are (coil, coeff) pairs
    LaunchTables *t = plan_tables(plan, nbatch * ncoeff, &rc, kTablesLowRank);
    if (!t) return rc;
    return launch_grid(*t, (const float2 *)samp, (float2 *)grid, weight, (hipStream_t)stream, lr);
}

int tn_apod_pad(tn_plan *plan, const void *image, void *grid, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !grid)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_apod_pad(plan, (const float2 *)image, (float2 *)grid, nimg, (hipStream_t)stream);
}

int tn_crop_apod(tn_plan *plan, const void *grid, void *image, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !grid)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_crop_apod(plan, (const float2 *)grid, (float2 *)image, nimg, (hipStream_t)stream);
}

int tn_plan_has_fft(const tn_plan *plan) { return plan && plan->d_fft_tw[0] ? 1 : 0; }

int tn_fft_forward(tn_plan *plan, const void *image, void *grid, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !grid)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_fft_forward(plan, (const float2 *)image, (float2 *)grid, nimg, (hipStream_t)stream);
}

int tn_fft_adjoint(tn_plan *plan, void *grid, void *image, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !grid)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_fft_adjoint(plan, (float2 *)grid, (float2 *)image, nimg, (hipStream_t)stream);
}

int tn_fft_z_forward(tn_plan *plan, const void *image, void *zout, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !zout)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_fft_z_forward(plan, (const float2 *)image, (float2 *)zout, nimg, (hipStream_t)stream);
}

int tn_fft_z_adjoint(tn_plan *plan, const void *zin, void *image, int64_t nimg, tn_stream stream) {
    clear_error();
    if (!plan) return fail(TN_E_ARG, "plan is NULL");
    if (nimg < 0) return fail(TN_E_ARG, "nimg must be >= 0");
    if (nimg > 0 && (!image || !zin)) return fail(TN_E_ARG, "NULL data pointer");
    int rc = check_device(plan->device);
    if (rc) return rc;
    return launch_fft_z_adjoint(plan, (const float2 *)zin, (float2 *)image, nimg, (hipStream_t)stream);
}

int tn_batch_create(tn_plan *const *plans, const int64_t *nbatch, const int64_t *grid_offset,
                    const int64_t *samp_offset, const int64_t *weight_offset, int32_t njobs,
                    tn_batch **out) {
    clear_error();
    if (!out || !plans || !nbatch || !grid_offset || !samp_offset)
        return fail(TN_E_ARG, "NULL argument");
    *out = nullptr;
    if (njobs < 1) return fail(TN_E_ARG, "njobs must be >= 1");
    std::vector<Job> jobs;
    std::vector<const tn_plan *> owners;
    int device = plans[0] ? plans[0]->device : -1;
    int ndim = plans[0] ? plans[0]->p.ndim : 0;
    for (int k = 0; k < njobs; ++k) {
        const tn_plan *pl = plans[k];
        if (!pl) return fail(TN_E_ARG, "plan %d is NULL", k);
        if (pl->device != device) return fail(TN_E_ARG, "plans live on different devices");
        if (pl->p.ndim != ndim) return fail(TN_E_ARG, "plans mix 2D and 3D");
        if (nbatch[k] < 1 || nbatch[k] >= INT32_MAX) return fail(TN_E_ARG, "nbatch[%d] < 1", k);
        if (grid_offset[k] < 0 || samp_offset[k] < 0) return fail(TN_E_ARG, "negative offset");
        const size_t before = jobs.size();
        make_jobs(pl, nbatch[k], grid_offset[k], samp_offset[k],
                  weight_offset ? weight_offset[k] : -1, jobs);
        owners.resize(jobs.size(), pl);
        (void)before;
    }
    int rc = check_device(device);
    if (rc) return rc;
    tn_batch *b = new tn_batch();
    b->device = device;
    b->plans.assign(plans, plans + njobs);
    rc = build_tables_for(jobs, owners, b->tables);
    if (rc) {
        free_tables(b->tables);
        delete b;
        return rc;
    }
    *out = b;
    return TN_OK;
}

int tn_batch_destroy(tn_batch *batch) {
    clear_error();
    if (!batch) return TN_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(batch->device);
    free_tables(batch->tables);
    (void)hipSetDevice(cur);
    delete batch;
    return TN_OK;
}

int tn_batch_interp(tn_batch *batch, const void *grid_base, void *samp_base,
                    const float *weight_base, tn_stream stream) {
    clear_error();
    if (!batch || !grid_base || !samp_base) return fail(TN_E_ARG, "NULL argument");
    int rc = check_device(batch->device);
    if (rc) return rc;
    return launch_interp(batch->tables, (const float2 *)grid_base, (float2 *)samp_base, weight_base,
                         (hipStream_t)stream);
}

int tn_batch_grid(tn_batch *batch, const void *samp_base, void *grid_base, const float *weight_base,
                  tn_stream stream) {
    clear_error();
    if (!batch || !grid_base || !samp_base) return fail(TN_E_ARG, "NULL argument");
    int rc = check_device(batch->device);
    if (rc) return rc;
    return launch_grid(batch->tables, (const float2 *)samp_base, (float2 *)grid_base, weight_base,
                       (hipStream_t)stream);
}

}  // extern "C"
