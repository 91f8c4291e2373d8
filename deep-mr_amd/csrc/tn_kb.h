// Kaiser-Bessel taps evaluated in registers (device code).
#pragma once

#include <hip/hip_runtime.h>

#include "tn_internal.h"

namespace tn {

// KB(u) = I0(beta * sqrt(1 - u^2)) for |u| <= 1, else 0, with I0 from
// Abramowitz & Stegun 9.8.1 (x < 3.75) and 9.8.2 (x >= 3.75), the same
// approximation deepmr uses (src/deepmr/fft/_interp/plan.py:216-248).
// Horner form in float32; KB(+-1) = I0(0) = 1 as in the reference.
__device__ __forceinline__ float kb_eval(float u, float beta) {
    const float r = 1.0f - u * u;
    if (r < 0.0f) return 0.0f;
    const float x = beta * sqrtf(r);
    if (x < 3.75f) {
        const float t = x * (1.0f / 3.75f);
        const float t2 = t * t;
        return 1.0f + t2 * (3.5156229f + t2 * (3.0899424f + t2 * (1.2067492f +
               t2 * (0.2659732f + t2 * (0.0360768f + t2 * 0.0045813f)))));
    }
    const float q = 3.75f * __builtin_amdgcn_rcpf(x);  // 1 / t
    const float poly = 0.39894228f + q * (0.01328592f + q * (0.00225319f +
                       q * (-0.00157565f + q * (0.00916281f + q * (-0.02057706f +
                       q * (0.02635537f + q * (-0.01647633f + q * 0.00392377f)))))));
    return poly * expf(x) * rsqrtf(x);
}

// Python-style (non-negative) modulo, valid for any int a and G > 0, so that
// every index stays inside the grid even for out-of-range coordinates.
__device__ __forceinline__ int wrap_mod(int a, int G) {
    int r = a % G;
    return r < 0 ? r + G : r;
}

// Taps along one axis for a sample at grid coordinate g.
//   W > 1: origin x0 = ceil(g - W/2), tap k at index (x0 + k + shift) mod G
//          with weight KB((x0 + k - g) / (W/2))   (plan.py:260-269)
//   W == 1 (Cartesian axis): index trunc(g) + shift, weight 1 (plan.py:118-121)
// `base` receives the (wrapped) index of tap 0; taps k > 0 follow at
// base + k with a wrap at G.
template <int WMAX>
__device__ __forceinline__ void axis_taps(float g, int W, int G, int shift, float beta,
                                          int &base, float (&w)[WMAX]) {
    if (W == 1) {
        base = wrap_mod((int)g + shift, G);
        w[0] = 1.0f;
#pragma unroll
        for (int k = 1; k < WMAX; ++k) w[k] = 0.0f;
        return;
    }
    const float half = 0.5f * (float)W;
    const float inv_half = 2.0f / (float)W;  // exact for the usual even widths
    const float x0 = ceilf(g - half);
    base = wrap_mod((int)x0 + shift, G);
#pragma unroll
    for (int k = 0; k < WMAX; ++k) {
        if (k < W) {
            const float u = ((x0 + (float)k) - g) * inv_half;
            w[k] = kb_eval(u, beta);
        } else {
            w[k] = 0.0f;
        }
    }
}

__device__ __forceinline__ int wrap_add(int base, int k, int G) {
    int i = base + k;
    return i >= G ? i - G : i;
}

}  // namespace tn
