/*
 * tilenufft — MI355X (gfx950) Kaiser-Bessel NUFFT kernels behind a C ABI.
 *
 * Drop-in boundary for deepmr's NUFFT hot path. deepmr binds its kernels
 * from Python through Numba (src/deepmr/_utils/backend.py:9-61, zero-copy
 * __cuda_array_interface__); this library replaces those kernels and is bound
 * through ctypes (see deep-mr_amd/tilenufft/_lib.py). Every entry point
 * takes plain pointers and sizes, returns 0 on success or a negative
 * TN_E_* code, and records a thread-local message readable with
 * tn_last_error(). No exception crosses the ABI.
 *
 * Conventions
 *   - Complex data are complex64 (interleaved float re, im).
 *   - Grids are (z, y, x) row-major, x fastest; 2D grids use grid[0] == 1.
 *   - Coordinates are grid units after deepmr's _scale_coord
 *     (src/deepmr/fft/nufft.py:411-420), (npts, ndim) per frame, x first.
 *   - All device buffers passed to apply calls are owned by the caller; the
 *     apply calls allocate nothing and launch only on `stream`, so they can
 *     be captured into a HIP graph once the plan's launch tables exist
 *     (tn_plan_prepare).
 *   - A plan is immutable after creation except for its launch-table cache,
 *     which tn_plan_prepare fills; calls on one plan from several host
 *     threads must be serialised by the caller.
 */
#ifndef TILENUFFT_H
#define TILENUFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tn_plan tn_plan;
typedef struct tn_batch tn_batch;
typedef void *tn_stream; /* hipStream_t */

enum {
    TN_OK = 0,
    TN_E_ARG = -1,     /* invalid argument / shape                    */
    TN_E_HIP = -2,     /* HIP runtime error (message has the details)  */
    TN_E_NOMEM = -3,   /* device allocation failed                     */
    TN_E_RANGE = -4,   /* a size exceeds the supported range           */
};

/* Plan parameters. All per-axis arrays are (z, y, x); 2D uses index 0 = z
 * with grid[0] = shape[0] = width[0] = 1. */
typedef struct tn_plan_params {
    int32_t ndim;       /* 2 or 3                                               */
    int32_t nframes;    /* trajectories (deepmr's ncontrasts axis)              */
    int64_t npts;       /* samples per frame (nviews * nsamples)                */
    int32_t shape[3];   /* image size N                                         */
    int32_t grid[3];    /* oversampled grid G = ceil(osf * N)                   */
    int32_t width[3];   /* kernel width 1..16; 1 marks a Cartesian axis; widths
                           above 8 gather (interp_mode 1) and grid through
                           the matrix (grid_mode 3)                          */
    float beta[3];      /* Kaiser-Bessel beta                                   */
    int32_t shift[3];   /* output index = (tap index + shift) mod G; the fused
                           NUFFT uses G - G/2 (folds the centring shifts), the
                           plain interpolator 0                                 */
    int32_t tile[3];    /* gridding tile; 0 = automatic                         */
    int32_t itile[3];   /* staged-interpolation tile; 0 = automatic             */
    int32_t interp_mode;  /* 0 automatic, 1 gather through L1/L2, 2 staged in LDS */
    int32_t interp_chunk; /* coils per interpolation work item; 0 = automatic  */
    int32_t grid_chunk;   /* coils per gridding workgroup; 0 = automatic (LDS) */
    int32_t grid_split;   /* max samples per gridding workgroup before a hot tile
                             is split across workgroups (partials folded in a
                             fixed order); 0 = automatic, < 0 = never split   */
    int32_t grid_mode;    /* gridding variant: 0 automatic (3 when its matrix
                             fits in 4 GiB, else 4 where it fits, else 2
                             when W_y * W_x <= 36, else 1), 1 one LDS atomic per
                             tap and sample (entries spread over footprint
                             origins), 2 per-lane runs of origin-sorted
                             entries summed in registers, one z plane of the
                             footprint per pass (W_y * W_x <= 36), 3 per-cell
                             sums over a sparse gridding matrix built with
                             the plan (CSR by cell, 8 bytes per sample and
                             tap) reading coil-minor sample records;
                             grid_chunk = coils per record, <= 16 (0: 16),
                             grid_split = row length above which one cell
                             is summed by a workgroup of its own; 4 (3D,
                             widths <= 4, grid lengths multiples of the
                             8x8x4 (x, y, z) tile with two or more tiles
                             per axis) one wave per tile summing its cells
                             in registers over the tile's samples in
                             ascending order, per-sample taps from a plan
                             table (64 bytes per sample), coil-minor
                             records; grid_chunk <= 8 (0: 8); tile ignored */
    float inv_scale;    /* 1 / kernel scaling (src/deepmr/fft/_interp/plan.py:190) */
    const float *apod[3]; /* host apodization vectors of length shape[d], or
                             NULL for ones (src/deepmr/fft/nufft.py:423-447)   */
} tn_plan_params;

/* Builds a plan on the current HIP device: copies the coordinates
 * (device pointer, nframes*npts*ndim floats), bins every sample into the
 * gridding tiles its kernel footprint touches (stable, sorted by sample
 * index inside each tile) and uploads the apodization vectors.
 * Replaces plan_interpolator / _prepare_interpolator
 * (src/deepmr/fft/_interp/plan.py:14-143, :251-269). The per-axis tap
 * values deepmr tabulates (8 * W * d bytes per sample) are not stored:
 * interpolation evaluates the taps in registers at apply time. Gridding
 * mode 3 stores its matrix instead: 8 bytes per (sample, tap), i.e.
 * 8 * W^d bytes per sample, plus 4 bytes per grid cell and frame of row
 * offsets. */
int tn_plan_create(const tn_plan_params *params, const float *coord_dev,
                   tn_stream stream, tn_plan **out);
int tn_plan_destroy(tn_plan *plan);

/* Device bytes held by the plan (coordinates, bins, tables). */
int64_t tn_plan_device_bytes(const tn_plan *plan);
/* Number of (tile, sample) entries after duplication across tile edges. */
int64_t tn_plan_entries(const tn_plan *plan);
/* Size of the union of the kernel footprints (grid cells touched by at
 * least one sample), summed over frames. Used for algorithmic-byte counts. */
int64_t tn_plan_footprint_cells(const tn_plan *plan);
/* Effective gridding tile (z, y, x). */
int tn_plan_tile(const tn_plan *plan, int32_t tile_out[3]);
/* Effective gridding variant (1 to 4, with grid_mode 0 resolved), or a
 * negative status for a NULL plan. No counterpart in deepmr (introspection
 * for tuners and benchmarks). */
int tn_plan_grid_mode(const tn_plan *plan);

/* Builds (once) the device launch tables for `nbatch` coils so that later
 * apply calls with that nbatch allocate nothing (needed before graph
 * capture). Apply calls call it implicitly. */
int tn_plan_prepare(tn_plan *plan, int64_t nbatch);

/* Interpolation, grid -> samples.
 *   grid: (nbatch, nframes, *G) complex64, natural order when shift == 0
 *   samp: (nbatch, nframes, npts) complex64, every element written
 *   weight: NULL or (nframes, npts) float, multiplied into the output
 * out[b,f,p] = inv_scale * weight[f,p] * sum_taps KB * grid[b,f,taps].
 * Replaces _do_interpolation_cuda2/3 + the /scale of apply_interpolation
 * (src/deepmr/fft/_interp/interp.py:682-742, :120). */
int tn_interp(tn_plan *plan, const void *grid, void *samp, int64_t nbatch,
              const float *weight, tn_stream stream);

/* Gridding, samples -> grid (the transpose of tn_interp).
 *   samp: (nbatch, nframes, npts) complex64
 *   grid: (nbatch, nframes, *G) complex64, every element written (zeros
 *         where no sample contributes); tile-owned, deterministic.
 * Replaces _do_gridding_cuda2/3 + the /scale of apply_gridding
 * (src/deepmr/fft/_interp/grid.py:1073-1166, :125). */
int tn_grid(tn_plan *plan, const void *samp, void *grid, int64_t nbatch,
            const float *weight, tn_stream stream);

/* Gridding exactly as tn_grid, without allocating: no counterpart in
 * deepmr; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops). The launch tables for `nbatch` must
 * already exist (tn_plan_prepare, or an earlier tn_grid with this nbatch),
 * else TN_E_ARG. Partial sums (long matrix rows summed in parts, split
 * tiles) go to a fixed device workspace of 32 MiB per device instead of
 * the tables' own; TN_E_RANGE when a plan needs more. Each partial is
 * written before it is read, so nothing has to start at zero and the
 * previous call's contents do no harm. Every call on one device shares
 * that workspace, so calls that overlap on one device (concurrent streams
 * or host threads) race; calls ordered on one stream do not. */
int tn_grid_static(tn_plan *plan, const void *samp, void *grid, int64_t nbatch,
                   const float *weight, tn_stream stream);

/* Low-rank (subspace) interpolation. The plan must hold every frame's
 * samples as one trajectory (nframes = 1, npts = nframes_lr * frame_npts,
 * frame f = samples [f*frame_npts, (f+1)*frame_npts)).
 * grid (nbatch, ncoeff, *G) complex64 -> samp (nbatch, npts) complex64 with
 * samp[b, p] = sum_k basis_adjoint[f(p), k] * interp(grid[b, k])[p] / scale;
 * basis_adjoint is a device array (nframes_lr, ncoeff) complex64; weight as
 * in tn_interp. Replaces _interpolate_lowrank{2,3} / _interpolate_lowrank_cuda{2,3}
 * (src/deepmr/fft/_interp/interp.py:268-357, :559-649). */
int tn_interp_lowrank(tn_plan *plan, const void *grid, void *samp, int64_t nbatch,
                      int32_t ncoeff, int32_t frame_npts, const void *basis_adjoint,
                      const float *weight, tn_stream stream);

/* Low-rank (subspace) gridding, same plan convention as tn_interp_lowrank.
 * samp (nbatch, npts) complex64 -> grid (nbatch, ncoeff, *G) complex64 with
 * grid[b, k] = sum_f basis[k, f] * grid(samp[b, frame f]) / scale; basis is
 * a device array (ncoeff, nframes_lr) complex64. Every grid cell is written.
 * Replaces _gridding_lowrank{2,3} / _gridding_lowrank_cuda{2,3}_cplx
 * (src/deepmr/fft/_interp/grid.py:326-417, :837-931). */
int tn_grid_lowrank(tn_plan *plan, const void *samp, void *grid, int64_t nbatch,
                    int32_t ncoeff, int32_t frame_npts, const void *basis,
                    const float *weight, tn_stream stream);

/* Fused apodize + centred zero-pad + ifftshift: image (nimg, *N) ->
 * grid (nimg, *G) ready for an unnormalised forward FFT. Replaces
 * _apodize + resize + the ifftshift of fft.fft
 * (src/deepmr/fft/nufft.py:499-505, src/deepmr/_signal/resize.py:11-102,
 * src/deepmr/fft/fft.py:76-80). */
int tn_apod_pad(tn_plan *plan, const void *image, void *grid, int64_t nimg,
                tn_stream stream);

/* Fused fftshift + centred crop + apodize: grid (nimg, *G) after an
 * inverse FFT -> image (nimg, *N). Replaces fft.ifft's shifts + resize +
 * _apodize (src/deepmr/fft/nufft.py:587-593). */
int tn_crop_apod(tn_plan *plan, const void *grid, void *image, int64_t nimg,
                 tn_stream stream);

/* 1 when the plan has the pruned, fused oversampled FFT stages below, else
 * 0: 2D and 3D plans whose grid lengths are among the oversampled sizes
 * ceil(osf N) of N = 64..512 at osf 1.25, 1.5 and 2 (64, 80, 96, 128, 160,
 * 192, 240, 256, 288, 320, 384, 400, 480, 512, 576, 640, 768, 1024) and the
 * short lengths 16, 24, 32, 40, 48 (small grids, Cartesian slice axes).
 * For 3D plans tn_fft_forward / tn_fft_adjoint transform (y, x) planes
 * (nimg counts planes) and tn_fft_z_forward / tn_fft_z_adjoint the z axis:
 * forward = z stage into a (nimg, Gz, Ny, Nx) workspace, then the plane
 * stages on its nimg * Gz planes; adjoint = the plane stages into that
 * workspace, then the z stage. */
int tn_plan_has_fft(const tn_plan *plan);

/* Whole forward FFT stage of the NUFFT on the device: apodization, centred
 * zero padding, ifftshift and the unnormalised 2D DFT, image (nimg, Ny, Nx)
 * -> grid (nimg, Gy, Gx) in the layout tn_interp reads (every cell written).
 * Transforms only the Nx image columns along y, then every row along x
 * from its Nx data columns. Same result as tn_apod_pad followed by an
 * unnormalised FFT. Replaces _apodize + resize + fft.fft(norm=None)
 * (src/deepmr/fft/nufft.py:499-505). TN_E_ARG when tn_plan_has_fft is 0. */
int tn_fft_forward(tn_plan *plan, const void *image, void *grid, int64_t nimg,
                   tn_stream stream);

/* Whole adjoint FFT stage: inverse 2D DFT scaled by 1 / (Gx Gy), fftshift,
 * centred crop and apodization, grid (nimg, Gy, Gx) -> image (nimg, Ny, Nx).
 * The grid is used as scratch (overwritten). Same result as an inverse FFT
 * followed by tn_crop_apod. Replaces fft.ifft(norm=None) + resize + _apodize
 * (src/deepmr/fft/nufft.py:587-593). TN_E_ARG when tn_plan_has_fft is 0. */
int tn_fft_adjoint(tn_plan *plan, void *grid, void *image, int64_t nimg,
                   tn_stream stream);

/* z stage of a 3D plan's forward FFT: apodization along z, centred zero
 * padding and the unnormalised DFT along z, image (nimg, Nz, Ny, Nx) ->
 * zout (nimg, Gz, Ny, Nx) (every cell written). With the plane stages it
 * replaces _apodize + resize + fft.fft(norm=None) of a 3D plan
 * (src/deepmr/fft/nufft.py:499-505). TN_E_ARG without a z stage. */
int tn_fft_z_forward(tn_plan *plan, const void *image, void *zout, int64_t nimg,
                     tn_stream stream);

/* z stage of a 3D plan's adjoint FFT: inverse DFT along z scaled by 1 / Gz,
 * centred crop and apodization along z, zin (nimg, Gz, Ny, Nx) -> image
 * (nimg, Nz, Ny, Nx) (src/deepmr/fft/nufft.py:587-593, with the plane
 * stages). TN_E_ARG without a z stage. */
int tn_fft_z_adjoint(tn_plan *plan, const void *zin, void *image, int64_t nimg,
                     tn_stream stream);

/* ---- batched launches over several plans (ragged work queue) ----------
 * A batch binds `njobs` (plan, nbatch, grid offset, sample offset) tuples to
 * single launches: one interpolation launch and one gridding launch cover
 * every frame, coil and tile of every plan in the batch. Offsets are in
 * complex elements from the base pointers given at apply time; plan k's
 * data is laid out as in tn_interp / tn_grid starting at its offset.
 * No counterpart in deepmr (it launches once per call,
 * src/deepmr/fft/_interp/interp.py:682-742). */
int tn_batch_create(tn_plan *const *plans, const int64_t *nbatch,
                    const int64_t *grid_offset, const int64_t *samp_offset,
                    const int64_t *weight_offset, int32_t njobs, tn_batch **out);
int tn_batch_destroy(tn_batch *batch);
int tn_batch_interp(tn_batch *batch, const void *grid_base, void *samp_base,
                    const float *weight_base, tn_stream stream);
int tn_batch_grid(tn_batch *batch, const void *samp_base, void *grid_base,
                  const float *weight_base, tn_stream stream);

/* ---- CG-SENSE of a group of reconstructions sharing the image size --------
 * One record per reconstruction, in device memory. Every kernel handles one
 * reconstruction per workgroup row and reduces over its image in a fixed
 * order that depends only on npix, so a result does not depend on which
 * other reconstructions share the call. */
typedef struct tn_sense_job {
    const void *smaps;  /* (ncoils, npix) complex64 coil maps               */
    void *x, *r, *p;    /* (npix) complex64 CG iterate, residual, direction */
    void *ap;           /* (npix) complex64 coil-combined normal product     */
    float *scal;        /* 2 floats of CG state (rs, next rs)                */
    int64_t img_off;    /* first coil image in the coil-image arena (complex
                           elements); coil c at img_off + c * npix           */
    int32_t ncoils;
    int32_t mode;       /* 0: start, 1: CG iteration, 2: idle (untouched)    */
} tn_sense_job;

/* SENSE expansion of the CG direction, img[c] = smaps[c] * p, for every
 * record in mode 1. Replaces SenseOp.forward
 * (src/deepmr/linops/coil.py:11-98). */
int tn_sense_expand(const tn_sense_job *jobs, int64_t njobs, int64_t npix, void *img,
                    tn_stream stream);

/* SENSE combination ap = sum_c conj(smaps[c]) * img[c] (coils in order) for
 * every record in mode 0 or 1. Replaces SenseAdjointOp.forward
 * (src/deepmr/linops/coil.py:101-205). */
int tn_sense_combine(const tn_sense_job *jobs, int64_t njobs, int64_t npix, const void *img,
                     tn_stream stream);

/* ---- one differentiable CG iteration, for autograd callers ----------------
 * nsys independent systems of n complex64 elements, arrays (nsys, n)
 * contiguous in device memory; rho, rho1 (nsys) float32; ad (nsys, 2)
 * complex64. Replaces CGStep.forward (src/deepmr/optim/cg.py:203-212)
 * given q = A p (A Hermitian):
 *   d = <p, q>, alpha = rho / d, x1 = x + alpha p, r1 = r - alpha q,
 *   rho1 = <r1, r1>, p1 = r1 + (rho1 / rho) p,   ad = (alpha, d)
 * Three launches; `work` holds tn_cg_grad_work_floats(nsys, n) floats.
 * Every sum has a fixed order that depends only on n. */
int64_t tn_cg_grad_work_floats(int64_t nsys, int64_t n);
int tn_cg_iterate(const void *x, const void *r, const void *p, const void *q, const float *rho, void *x1, void *r1,
                  void *p1, float *rho1, void *ad, float *work, int64_t nsys, int64_t n, tn_stream stream);
/* Its vector-Jacobian product in torch's convention for complex gradients
 * (conjugate Wirtinger): from the gradients of x1, r1, p1 (each may be NULL
 * for zero) and rho1 (may be NULL), the gradients of r, p, q and rho (the
 * gradient of x is that of x1). Two launches, same `work` size. */
int tn_cg_iterate_backward(const void *p, const void *q, const void *r1, const float *rho, const float *rho1,
                           const void *ad, const void *gx1, const void *gr1, const void *gp1, const float *grho1,
                           void *gr, void *gp, void *gq, float *grho, float *work, int64_t nsys, int64_t n,
                           tn_stream stream);

/* Coil combination of SenseAdjointOp (src/deepmr/linops/coil.py:101-205) on
 * contiguous complex64 arrays: smaps (nsets, nb, ncoils, npix), y (nsets, nb,
 * ncoils, nk, npix) -> x (nb, nk, npix) with x[b, k, p] = sum over sets of
 * the sum over coils of conj(smaps[s, b, c, p]) * y[s, b, c, k, p] (coils
 * first, in order). One launch, no allocation. */
int tn_coil_combine(const void *smaps, const void *y, void *x, int64_t nsets, int64_t nb, int64_t ncoils,
                    int64_t nk, int64_t npix, tn_stream stream);

/* Floats of `work` that tn_cg_step needs for njobs records of npix pixels. */
int64_t tn_cg_work_floats(int64_t njobs, int64_t npix);

/* One conjugate-gradient update per record, given ap = A^H A p:
 *   mode 1: alpha = rs / <p, ap>, x += alpha p, r -= alpha ap,
 *           rs' = <r, r>, p = r + (rs' / rs) p, rs = rs'
 *   mode 0: x = 0, r = p = ap, rs = <ap, ap>   (start: ap = A^H y)
 * Replaces CGStep.forward (src/deepmr/optim/cg.py:203-212). Three launches,
 * no allocation. */
int tn_cg_step(const tn_sense_job *jobs, int64_t njobs, int64_t npix, float *work,
               tn_stream stream);

/* Thread-local description of the last error (empty string if none). */
const char *tn_last_error(void);
/* Library version string. */
const char *tn_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TILENUFFT_H */
