// Oversampled-grid FFT stages of the 2D NUFFT for gfx950, pruned to the
// lines that carry data and fused with the zero padding, the apodization
// and the crop (src/deepmr/fft/nufft.py:499-505 forward: apodize, resize,
// fft with norm=None; :587-593 adjoint: ifft, resize, apodize).
//
// Index maps: image pixel m of an axis sits at grid index (m - N/2) mod G
// (the centring shifts folded into the placement, as in the other stages),
// so a grid line along y holds data only at the Ny rows with
// (y + Ny/2) mod Gy < Ny until the y transform has run, and likewise along x.
//
//   forward  columns: image columns -> apodize -> DFT along y -> grid columns
//                     (only the Nx columns that carry data)
//            rows:    DFT along x of every grid row, reading only its Nx
//                     data columns
//   adjoint  rows:    inverse DFT along x of every grid row, writing only
//                     the Nx columns the crop keeps
//            columns: inverse DFT along y of those columns, keeping the Ny
//                     image rows, scaled by 1 / (Gx Gy), apodized
//
// Each line transform is a mixed-radix (8, 4, 2, 3, 5) Stockham FFT run by
// one wave: the radix-R butterflies of a pass are evaluated in registers
// with twiddles from a table computed in double precision on the host, and
// the passes exchange data through that wave's LDS line. All reads of a pass
// are issued before its writes, and a wave's LDS operations complete in
// order, so a line is transformed in place without barriers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "tn_internal.h"

namespace tn {
namespace {

// LDS column padding of a C-column tile, in float2: a column stride of
// 64 / C mod 16 float2 puts the 64 lanes of a row-major tile load (C columns
// x 64 / C rows) on 16 distinct bank pairs, four lanes each.
constexpr int fft_pad(int C) { return 64 / C; }

// The Stockham passes keep a line in LDS with one float2 of padding after
// every 8: the scattered writes of a radix-8 pass (lane j at 8 j + r, or
// in runs of 8 at stride 64) then fall on distinct banks instead of 8 or 16
// lanes per bank. A line of L elements takes fft_line(L) float2.
constexpr int fft_line(int L) { return L + L / 8; }
__device__ __forceinline__ int lds_at(int a) { return a + (a >> 3); }

// Radix of pass p of a length-L transform: radix 8 while possible, then 4,
// 2, 3 and 5; 0 past the last pass.
constexpr int fft_radix_at(int L, int p) {
    const int order[5] = {8, 4, 2, 3, 5};
    int i = 0;
    for (int o = 0; o < 5; ++o) {
        while (L > 1 && L % order[o] == 0) {
            if (i == p) return order[o];
            L /= order[o];
            ++i;
        }
    }
    return 0;
}

constexpr int fft_npass(int L) {
    int n = 0;
    while (fft_radix_at(L, n) != 0) ++n;
    return n;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (s i) with s = -1 (forward) or +1 (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_j(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// In-register DFTs of length R (sign -1 forward, +1 inverse).
template <bool INV>
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_j<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a1 = cadd(t1, t3);
    a2 = csub(t0, t2);
    a3 = csub(t1, t3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&v)[R]) {
    if constexpr (R == 2) {
        const float2 t = v[0];
        v[0] = cadd(t, v[1]);
        v[1] = csub(t, v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<INV>(e0, e1, e2, e3);
        dft4<INV>(o0, o1, o2, o3);
        const float h = 0.70710678118654752f;
        const float2 w1 = cscale(cadd(o1, mul_j<INV>(o1)), h);   // exp(s i pi/4) o1
        const float2 w2 = mul_j<INV>(o2);                        // exp(s i pi/2) o2
        const float2 w3 = cscale(csub(mul_j<INV>(o3), o3), h);   // exp(s 3i pi/4) o3
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, w1);
        v[5] = csub(e1, w1);
        v[2] = cadd(e2, w2);
        v[6] = csub(e2, w2);
        v[3] = cadd(e3, w3);
        v[7] = csub(e3, w3);
    } else if constexpr (R == 3) {
        const float c = -0.5f, sn = 0.86602540378443865f;
        const float2 t1 = cadd(v[1], v[2]), t2 = cscale(mul_j<INV>(csub(v[1], v[2])), sn);
        const float2 m = cadd(v[0], cscale(t1, c));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, t2);
        v[2] = csub(m, t2);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
        const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
        const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const float2 m1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
        const float2 m2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
        const float2 n1 = mul_j<INV>(cadd(cscale(t3, s1), cscale(t4, s2)));
        const float2 n2 = mul_j<INV>(csub(cscale(t3, s2), cscale(t4, s1)));
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
}

// Stockham passes P.. of NL length-L lines held by one wave (the lines are
// transformed together: their loads are independent and share the
// twiddles). Pass P, radix R, NS = product of the earlier radices:
// butterfly j < L/R reads elements j + r L/R, multiplies element r by
// W_{NS R}^{r (j mod NS)}, runs the R-point DFT and stores element r at
// (j / NS) NS R + (j mod NS) + r NS. Pass 0 reads line l through
// src(l, i), the last pass writes through dst(l, i, v), the others use
// buf[l] (the wave's LDS lines of fft_line(L) float2, addressed through
// lds_at; src / dst may address them too, unpadded).
template <int L, int NL, int P, int NS, bool INV, class Src, class Dst>
__device__ __forceinline__ void fft_passes(int lane, const float2 *__restrict__ tw, float2 *const *buf, Src &src,
                                           Dst &dst) {
    constexpr int NP = fft_npass(L);
    constexpr int R = fft_radix_at(L, P);
    constexpr int M = L / R;
    constexpr int Q = (M + kWave - 1) / kWave;
    constexpr int TS = L / (NS * R);  // twiddle table stride
    float2 v[NL][Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lane + q * kWave;
        if (M % kWave == 0 || j < M) {
            const int k = j % NS;
            float2 w[R];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    w[r] = tw[r * k * TS];
                    if (INV) w[r].y = -w[r].y;
                }
            }
#pragma unroll
            for (int l = 0; l < NL; ++l) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (P == 0)
                        v[l][q][r] = src(l, j + r * M);
                    else
                        v[l][q][r] = buf[l][lds_at(j + r * M)];
                }
            }
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[l][q][r] = cmul(v[l][q][r], w[r]);
                }
                dft<R, INV>(v[l][q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lane + q * kWave;
        if (M % kWave == 0 || j < M) {
            const int k = j % NS;
            const int d = (j / NS) * NS * R + k;
#pragma unroll
            for (int l = 0; l < NL; ++l) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (P == NP - 1)
                        dst(l, d + r * NS, v[l][q][r]);
                    else
                        buf[l][lds_at(d + r * NS)] = v[l][q][r];
                }
            }
        }
    }
    if constexpr (P + 1 < NP) fft_passes<L, NL, P + 1, NS * R, INV>(lane, tw, buf, src, dst);
}

template <int L, int NL, bool INV, class Src, class Dst>
__device__ __forceinline__ void fft_lines(int lane, const float2 *__restrict__ tw, float2 *const *buf, Src src,
                                          Dst dst) {
    fft_passes<L, NL, 0, 1, INV>(lane, tw, buf, src, dst);
}

// data index of pixel-space position i of an axis: kept iff (i + N/2) mod L < N
__device__ __forceinline__ bool kept(int i, int N, int L) {
    int m = i + N / 2;
    if (m >= L) m -= L;
    return m < N;
}

// Row transforms along x, in place, rows_per_wave(L) grid rows (length
// L = Gx) per wave, transformed together (independent loads in flight,
// shared twiddles; four rows up to L = 256, two up to 512, one above, where
// the LDS lines would cost occupancy: on 256^3 and 192^3 grids four rows cut
// the 3D stages by 3-5%, while four rows of 512 or eight of 256 were slower).
// Forward: reads only the Nx data columns (the others are zeros). Inverse:
// writes only the Nx columns the crop keeps.
constexpr int rows_per_wave(int L) { return L <= 256 ? 4 : L <= 512 ? 2 : 1; }

template <int L, bool INV>
__global__ __launch_bounds__(kBlock) void fft_rows_kernel(float2 *__restrict__ data, int64_t nlines, int Nx,
                                                          const float2 *__restrict__ tw) {
    constexpr int RL = rows_per_wave(L);
    __shared__ float2 lds[kBlock / kWave * RL][fft_line(L)];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t line0 = ((int64_t)blockIdx.x * (kBlock / kWave) + wave) * RL;
    if (line0 >= nlines) return;  // no barriers below
    // a row past the last one reads the wave's first row and writes nothing
    float2 *row[RL];
    bool ok[RL];
    float2 *buf[RL];
#pragma unroll
    for (int l = 0; l < RL; ++l) {
        ok[l] = line0 + l < nlines;
        row[l] = data + (ok[l] ? line0 + l : line0) * L;
        buf[l] = lds[wave * RL + l];
    }
    if (!INV) {
        fft_lines<L, RL, false>(
            lane, tw, buf, [&](int l, int i) { return kept(i, Nx, L) ? row[l][i] : make_float2(0.0f, 0.0f); },
            [&](int l, int i, float2 v) {
                if (ok[l]) row[l][i] = v;
            });
    } else {
        fft_lines<L, RL, true>(
            lane, tw, buf, [&](int l, int i) { return row[l][i]; },
            [&](int l, int i, float2 v) {
                if (ok[l] && kept(i, Nx, L)) row[l][i] = v;
            });
    }
}

// Block of C adjacent image columns of one image, dealt to the XCDs in
// contiguous ranges (neighbouring column groups share cache lines).
__device__ __forceinline__ bool column_block(int64_t nimg, int Nx, int C, int64_t &b, int &mx0) {
    const int ngrp = (Nx + C - 1) / C;
    const int64_t nblk = nimg * ngrp;
    const int64_t per = (nblk + kXcds - 1) / kXcds;
    const int64_t blk = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (blk >= nblk) return false;
    b = blk / ngrp;
    mx0 = (int)(blk - b * ngrp) * C;
    return true;
}

// Forward column stage: image columns mx0 .. mx0 + C - 1 of image b are
// apodized (x then y, as the separate pre-pass does), placed at their grid
// rows in LDS, transformed along y (length L = Gy; rows without data are
// zeros) and written to grid columns (mx - hx) mod Gx (hx = Nx/2 for the x
// axis of a grid; 0 where the columns are a flattened (y, x) plane that a
// later stage places).
template <int L, int C>
__global__ __launch_bounds__(kBlock) void fft_cols_fwd_kernel(const float2 *__restrict__ img,
                                                              float2 *__restrict__ grid, int64_t nimg, int Nx,
                                                              int Ny, int Gx, int hx, const float *__restrict__ apx,
                                                              const float *__restrict__ apy,
                                                              const float2 *__restrict__ tw) {
    constexpr int LP = fft_line(L) + fft_pad(C);
    extern __shared__ float2 lds[];  // C * LP
    int64_t b;
    int mx0;
    if (!column_block(nimg, Nx, C, b, mx0)) return;
    const int hy = Ny / 2;
    for (int i0 = threadIdx.x; i0 < Ny * C; i0 += kLoadBatch * kBlock) {  // loads first, then the LDS writes
        float2 v[kLoadBatch];
        float ax[kLoadBatch], ay[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            const int my = i / C, mx = mx0 + i % C;
            v[u] = make_float2(0.0f, 0.0f);
            ax[u] = ay[u] = 1.0f;
            if (i < Ny * C && mx < Nx) {
                v[u] = img[(b * Ny + my) * (int64_t)Nx + mx];
                if (apx) ax[u] = apx[mx];
                if (apy) ay[u] = apy[my];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            if (i >= Ny * C) continue;
            const int my = i / C, c = i % C;
            const float2 w = cscale(cscale(v[u], ax[u]), ay[u]);  // x then y, as the separate pre-pass
            int y = my - hy;
            if (y < 0) y += L;
            lds[c * LP + y] = w;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    {   // wave w transforms columns w, w + 4, ... together
        constexpr int NL = C / (kBlock / kWave);
        float2 *col[NL];
#pragma unroll
        for (int l = 0; l < NL; ++l) col[l] = lds + (wave + l * (kBlock / kWave)) * LP;
        fft_lines<L, NL, false>(
            lane, tw, col,
            [&](int l, int i) { return kept(i, Ny, L) ? col[l][i] : make_float2(0.0f, 0.0f); },
            [&](int l, int i, float2 v) { col[l][i] = v; });
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L * C; i += kBlock) {
        const int y = i / C, c = i % C;
        const int mx = mx0 + c;
        if (mx < Nx) {
            int x = mx - hx;
            if (x < 0) x += Gx;
            grid[(b * L + y) * (int64_t)Gx + x] = lds[c * LP + y];
        }
    }
}

// Adjoint column stage: grid columns (mx - hx) mod Gx for image columns
// mx0 .. mx0 + C - 1, inverse-transformed along y (L = Gy); the Ny image rows
// are scaled by 1 / (Gx Gy), apodized (x then y) and written to the image.
template <int L, int C>
__global__ __launch_bounds__(kBlock) void fft_cols_inv_kernel(const float2 *__restrict__ grid,
                                                              float2 *__restrict__ img, int64_t nimg, int Nx,
                                                              int Ny, int Gx, int hx, const float *__restrict__ apx,
                                                              const float *__restrict__ apy,
                                                              const float2 *__restrict__ tw, float scale) {
    constexpr int LP = fft_line(L) + fft_pad(C);
    extern __shared__ float2 lds[];  // C * LP
    int64_t b;
    int mx0;
    if (!column_block(nimg, Nx, C, b, mx0)) return;
    const int hy = Ny / 2;
    for (int i0 = threadIdx.x; i0 < L * C; i0 += kLoadBatch * kBlock) {  // loads first, then the LDS writes
        float2 v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            const int y = i / C, mx = mx0 + i % C;
            v[u] = make_float2(0.0f, 0.0f);
            if (i < L * C && mx < Nx) {
                int x = mx - hx;
                if (x < 0) x += Gx;
                v[u] = grid[(b * L + y) * (int64_t)Gx + x];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            if (i < L * C) lds[(i % C) * LP + i / C] = v[u];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    {   // wave w transforms columns w, w + 4, ... together
        constexpr int NL = C / (kBlock / kWave);
        float2 *col[NL];
#pragma unroll
        for (int l = 0; l < NL; ++l) col[l] = lds + (wave + l * (kBlock / kWave)) * LP;
        fft_lines<L, NL, true>(
            lane, tw, col, [&](int l, int i) { return col[l][i]; },
            [&](int l, int i, float2 v) { col[l][i] = v; });
    }
    __syncthreads();
    for (int i0 = threadIdx.x; i0 < Ny * C; i0 += kLoadBatch * kBlock) {  // apodization loads first
        float ax[kLoadBatch], ay[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            const int my = i / C, mx = mx0 + i % C;
            ax[u] = ay[u] = 1.0f;
            if (i < Ny * C && mx < Nx) {
                if (apx) ax[u] = apx[mx];
                if (apy) ay[u] = apy[my];
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const int i = i0 + u * kBlock;
            const int my = i / C, c = i % C, mx = mx0 + c;
            if (i < Ny * C && mx < Nx) {
                int y = my - hy;
                if (y < 0) y += L;
                const float2 v = cscale(cscale(cscale(lds[c * LP + y], scale), ax[u]), ay[u]);
                img[(b * Ny + my) * (int64_t)Nx + mx] = v;
            }
        }
    }
}

// Columns per column-stage workgroup. The strided row segments, not the
// transform, bound these stages; still, wider tiles (16 columns at L = 512,
// 66 KiB of LDS, two workgroups per CU) measured slower than 8 columns at
// four workgroups per CU.
constexpr int col_group(int L) { return L <= 256 ? 16 : L <= 512 ? 8 : 4; }
constexpr size_t col_lds_bytes(int L) {
    return sizeof(float2) * (size_t)col_group(L) * (fft_line(L) + fft_pad(col_group(L)));
}

using RowFn = void (*)(float2 *, int64_t, int, const float2 *);
using ColFwdFn = void (*)(const float2 *, float2 *, int64_t, int, int, int, int, const float *, const float *,
                          const float2 *);
using ColInvFn = void (*)(const float2 *, float2 *, int64_t, int, int, int, int, const float *, const float *,
                          const float2 *, float);

struct FftFns {
    int L = 0;
    RowFn row_fwd = nullptr, row_inv = nullptr;
    ColFwdFn col_fwd = nullptr;
    ColInvFn col_inv = nullptr;
    int C = 0;
    size_t lds = 0;  // dynamic LDS bytes of the column kernels
};

template <int L>
FftFns fns_for() {
    FftFns f;
    f.L = L;
    f.C = col_group(L);
    f.row_fwd = fft_rows_kernel<L, false>;
    f.row_inv = fft_rows_kernel<L, true>;
    f.col_fwd = fft_cols_fwd_kernel<L, col_group(L)>;
    f.col_inv = fft_cols_inv_kernel<L, col_group(L)>;
    f.lds = col_lds_bytes(L);
    return f;
}

// Transform lengths with kernels: the oversampled sizes ceil(osf N) of
// N = 64, 128, 192, 256, 320, 384, 512 at osf 1.25, 1.5 and 2 (all of the
// form 2^a 3^b 5^c), 64 .. 1024, and the short lengths 16 .. 48 of small
// grids and of Cartesian slice axes (osf 1).
FftFns fft_fns(int L) {
    switch (L) {
        case 16: return fns_for<16>();
        case 24: return fns_for<24>();
        case 32: return fns_for<32>();
        case 40: return fns_for<40>();
        case 48: return fns_for<48>();
        case 64: return fns_for<64>();
        case 80: return fns_for<80>();
        case 96: return fns_for<96>();
        case 128: return fns_for<128>();
        case 160: return fns_for<160>();
        case 192: return fns_for<192>();
        case 240: return fns_for<240>();
        case 256: return fns_for<256>();
        case 288: return fns_for<288>();
        case 320: return fns_for<320>();
        case 384: return fns_for<384>();
        case 400: return fns_for<400>();
        case 480: return fns_for<480>();
        case 512: return fns_for<512>();
        case 576: return fns_for<576>();
        case 640: return fns_for<640>();
        case 768: return fns_for<768>();
        case 1024: return fns_for<1024>();
        default: return FftFns{};
    }
}

unsigned col_blocks(int64_t nimg, int Nx, int C) {
    const int64_t n = nimg * ((Nx + C - 1) / C);
    return (unsigned)((n + kXcds - 1) / kXcds * kXcds);
}

}  // namespace

// 2D and 3D plans whose transformed grid lengths have kernels
bool fft_supported(const tn_plan *plan) {
    const tn_plan_params &p = plan->p;
    if (p.ndim != 2 && p.ndim != 3) return false;
    if (p.ndim == 3 && fft_fns(p.grid[0]).L == 0) return false;
    return fft_fns(p.grid[2]).L != 0 && fft_fns(p.grid[1]).L != 0;
}

int build_fft(tn_plan *plan, hipStream_t stream) {
    if (!fft_supported(plan)) return TN_OK;
    // dynamic LDS of the column tiles (up to 64 KiB with these column counts),
    // for the y columns and, in 3D, the z columns
    for (int a = 1; a < plan->p.ndim; ++a) {
        const FftFns fc = fft_fns(plan->p.grid[2 - a]);
        TN_HIP(hipFuncSetAttribute((const void *)fc.col_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fc.lds));
        TN_HIP(hipFuncSetAttribute((const void *)fc.col_inv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fc.lds));
    }
    for (int a = 0; a < plan->p.ndim; ++a) {  // 0: x (grid[2]), 1: y (grid[1]), 2: z (grid[0])
        const int L = plan->p.grid[2 - a];
        std::vector<float2> tw(L);
        const double pi = 3.14159265358979323846;
        for (int m = 0; m < L; ++m) {
            const double ang = -2.0 * pi * (double)m / (double)L;
            tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        TN_HIP(hipMalloc(&plan->d_fft_tw[a], sizeof(float2) * L));
        plan->fft_tw_bytes += (int64_t)sizeof(float2) * L;
        TN_HIP(hipMemcpyAsync(plan->d_fft_tw[a], tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice, stream));
        TN_HIP(hipStreamSynchronize(stream));  // before the host table goes out of scope
    }
    return TN_OK;
}

int launch_fft_forward(const tn_plan *plan, const float2 *image, float2 *grid, int64_t nimg, hipStream_t stream) {
    if (!plan->d_fft_tw[0]) return fail(TN_E_ARG, "the plan has no pruned FFT (2D and 3D plans with common oversampled grid lengths, 16..1024)");
    if (nimg == 0) return TN_OK;
    const int Nx = plan->p.shape[2], Ny = plan->p.shape[1];
    const int Gx = plan->p.grid[2], Gy = plan->p.grid[1];
    const FftFns fy = fft_fns(Gy), fx = fft_fns(Gx);
    hipLaunchKernelGGL(fy.col_fwd, dim3(col_blocks(nimg, Nx, fy.C)), dim3(kBlock), fy.lds, stream, image, grid, nimg, Nx,
                       Ny, Gx, Nx / 2, (const float *)plan->d_apod[2], (const float *)plan->d_apod[1],
                       (const float2 *)plan->d_fft_tw[1]);
    TN_HIP(hipGetLastError());
    const int64_t nlines = nimg * Gy;
    const int64_t rpb = (int64_t)rows_per_wave(Gx) * (kBlock / kWave);  // rows per block
    const int64_t rb = (nlines + rpb - 1) / rpb;
    hipLaunchKernelGGL(fx.row_fwd, dim3((unsigned)rb), dim3(kBlock), 0, stream, grid, nlines, Nx,
                       (const float2 *)plan->d_fft_tw[0]);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int launch_fft_adjoint(const tn_plan *plan, float2 *grid, float2 *image, int64_t nimg, hipStream_t stream) {
    if (!plan->d_fft_tw[0]) return fail(TN_E_ARG, "the plan has no pruned FFT (2D and 3D plans with common oversampled grid lengths, 16..1024)");
    if (nimg == 0) return TN_OK;
    const int Nx = plan->p.shape[2], Ny = plan->p.shape[1];
    const int Gx = plan->p.grid[2], Gy = plan->p.grid[1];
    const FftFns fy = fft_fns(Gy), fx = fft_fns(Gx);
    const int64_t nlines = nimg * Gy;
    const int64_t rpb = (int64_t)rows_per_wave(Gx) * (kBlock / kWave);  // rows per block
    const int64_t rb = (nlines + rpb - 1) / rpb;
    hipLaunchKernelGGL(fx.row_inv, dim3((unsigned)rb), dim3(kBlock), 0, stream, grid, nlines, Nx,
                       (const float2 *)plan->d_fft_tw[0]);
    TN_HIP(hipGetLastError());
    const float scale = (float)(1.0 / ((double)Gx * (double)Gy));
    hipLaunchKernelGGL(fy.col_inv, dim3(col_blocks(nimg, Nx, fy.C)), dim3(kBlock), fy.lds, stream, grid, image, nimg, Nx,
                       Ny, Gx, Nx / 2, (const float *)plan->d_apod[2], (const float *)plan->d_apod[1],
                       (const float2 *)plan->d_fft_tw[1], scale);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

// z stage of the 3D plans: each image volume (Nz, Ny, Nx) is a 2D image of
// Nz rows and Ny*Nx columns; the column kernels transform it along z (length
// Gz) with the z apodization, the z placement (mz - Nz/2) mod Gz (forward) or
// crop (adjoint) and, in the adjoint, the 1/Gz scale; the columns keep their
// flattened (y, x) order (hx = 0) for the plane stages.
int launch_fft_z_forward(const tn_plan *plan, const float2 *image, float2 *zout, int64_t nimg, hipStream_t stream) {
    if (plan->p.ndim != 3 || !plan->d_fft_tw[2]) return fail(TN_E_ARG, "the plan has no z stage (3D plans with pruned FFT stages)");
    if (nimg == 0) return TN_OK;
    const int cols = plan->p.shape[1] * plan->p.shape[2], Nz = plan->p.shape[0], Gz = plan->p.grid[0];
    const FftFns fz = fft_fns(Gz);
    hipLaunchKernelGGL(fz.col_fwd, dim3(col_blocks(nimg, cols, fz.C)), dim3(kBlock), fz.lds, stream, image, zout, nimg, cols,
                       Nz, cols, 0, (const float *)nullptr, (const float *)plan->d_apod[0],
                       (const float2 *)plan->d_fft_tw[2]);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

int launch_fft_z_adjoint(const tn_plan *plan, const float2 *zin, float2 *image, int64_t nimg, hipStream_t stream) {
    if (plan->p.ndim != 3 || !plan->d_fft_tw[2]) return fail(TN_E_ARG, "the plan has no z stage (3D plans with pruned FFT stages)");
    if (nimg == 0) return TN_OK;
    const int cols = plan->p.shape[1] * plan->p.shape[2], Nz = plan->p.shape[0], Gz = plan->p.grid[0];
    const FftFns fz = fft_fns(Gz);
    const float scale = (float)(1.0 / (double)Gz);
    hipLaunchKernelGGL(fz.col_inv, dim3(col_blocks(nimg, cols, fz.C)), dim3(kBlock), fz.lds, stream, zin, image, nimg, cols,
                       Nz, cols, 0, (const float *)nullptr, (const float *)plan->d_apod[0],
                       (const float2 *)plan->d_fft_tw[2], scale);
    TN_HIP(hipGetLastError());
    return TN_OK;
}

}  // namespace tn
