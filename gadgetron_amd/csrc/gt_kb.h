// gt_kb.h -- Kaiser-Bessel gridding kernel, shared by host and device code.
//
// Same formulas as Gadgetron's KaiserKernel (toolboxes/nfft/ConvolutionKernel.hpp):
//   bessi0 rational approximant of I0 (float form, :39-68)
//   phi(u) = N_os * I0(beta * sqrt(1 - (2u/W)^2)) / W       (:85-94)
//   beta   = pi * sqrt(W^2/a^2 * (a - 1/2)^2 - 0.8)          (:287-303)
//   zero outside |u| <= W/2, tested per dimension            (:164-180)
// The D-dimensional weight is the product of the per-dimension weights, x
// first (:108-131), which is what makes the separable per-sample weight
// tables of the gridding kernels exact.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace gt {

__host__ __device__ inline float bessi0f(float x) {
    if (x == 0.0f) return 1.0f;
    const float z = x * x;
    const float num =
        (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z *
            (z * 0.210580722890567e-22f + 0.380715242345326e-19f) +
            0.479440257548300e-16f) + 0.435125971262668e-13f) +
            0.300931127112960e-10f) + 0.160224679395361e-7f) +
            0.654858370096785e-5f) + 0.202591084143397e-2f) +
            0.463076284721000e0f) + 0.754337328948189e2f) +
            0.830792541809429e4f) + 0.571661130563785e6f) +
            0.216415572361227e8f) + 0.356644482244025e9f) +
            0.144048298227235e10f);
    const float den = (z * (z * (z - 0.307646912682801e4f) + 0.347626332405882e7f) -
                       0.144048298227235e10f);
    return -num / den;
}

// One-dimensional Kaiser-Bessel weight for |u| <= W/2 (caller tests the radius).
__host__ __device__ inline float kb1(float u, float nos, float one_over_W, float beta) {
    const float t = 2.0f * u * one_over_W;
    const float tt = t * t;
    const float arg = beta * sqrtf(1.0f - tt);
    return nos * bessi0f(arg) * one_over_W;
}

// Derivative d phi / d u of kb1 (u >= 0, inside the radius).  Used by the
// trajectory gradient; differentiates the same rational approximant, so it is
// consistent with finite differences of kb1 itself.
__host__ __device__ inline float bessi0f_deriv(float x) {
    // d/dx [-num(z)/den(z)], z = x^2  ->  -2x (num' den - num den') / den^2
    const float z = x * x;
    const float c[15] = {0.144048298227235e10f, 0.356644482244025e9f, 0.216415572361227e8f,
                         0.571661130563785e6f,  0.830792541809429e4f, 0.754337328948189e2f,
                         0.463076284721000e0f,  0.202591084143397e-2f, 0.654858370096785e-5f,
                         0.160224679395361e-7f, 0.300931127112960e-10f, 0.435125971262668e-13f,
                         0.479440257548300e-16f, 0.380715242345326e-19f, 0.210580722890567e-22f};
    float num = c[14], dnum = 0.0f;
    for (int i = 13; i >= 0; i--) {
        dnum = dnum * z + num;
        num = num * z + c[i];
    }
    const float den = (z * (z * (z - 0.307646912682801e4f) + 0.347626332405882e7f) -
                       0.144048298227235e10f);
    const float dden = (3.0f * z - 2.0f * 0.307646912682801e4f) * z + 0.347626332405882e7f;
    return -2.0f * x * (dnum * den - num * dden) / (den * den);
}

__host__ __device__ inline float kb1_deriv(float u, float nos, float one_over_W, float beta) {
    const float t = 2.0f * u * one_over_W;
    const float s = sqrtf(fmaxf(1.0f - t * t, 0.0f));
    if (s <= 0.0f) return 0.0f;
    const float arg = beta * s;
    // d arg / du = beta * (-t / s) * 2 / W
    const float darg = -beta * t / s * 2.0f * one_over_W;
    return nos * one_over_W * bessi0f_deriv(arg) * darg;
}

// ConvolutionKernel.hpp:287-303, evaluated in float like the float reference.
// Double forms (ConvolutionKernel.hpp:9-37 and :70-82): the same rational
// approximant and kernel in double precision, for double-precision plans.
__host__ __device__ inline double bessi0d(double x) {
    if (x == 0.0) return 1.0;
    const double z = x * x;
    const double num =
        (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z * (z *
            (z * 0.210580722890567e-22 + 0.380715242345326e-19) +
            0.479440257548300e-16) + 0.435125971262668e-13) +
            0.300931127112960e-10) + 0.160224679395361e-7) +
            0.654858370096785e-5) + 0.202591084143397e-2) +
            0.463076284721000e0) + 0.754337328948189e2) +
            0.830792541809429e4) + 0.571661130563785e6) +
            0.216415572361227e8) + 0.356644482244025e9) +
            0.144048298227235e10);
    const double den = (z * (z * (z - 0.307646912682801e4) + 0.347626332405882e7) - 0.144048298227235e10);
    return -num / den;
}

__host__ __device__ inline double kb1d(double u, double nos, double one_over_W, double beta) {
    const double t = 2.0 * u * one_over_W;
    const double arg = beta * sqrt(1.0 - t * t);
    return nos * bessi0d(arg) * one_over_W;
}

inline double kb_beta_d(double W, double os_factor) {
    return M_PI * std::sqrt(W * W / (os_factor * os_factor) * (os_factor - 0.5) * (os_factor - 0.5) - 0.8);
}

inline float kb_beta(float W, float os_factor) {
    auto sqr = [](float x) { return x * x; };
    return float(M_PI) * std::sqrt(sqr(W) / sqr(os_factor) * sqr(os_factor - 0.5f) - 0.8f);
}

}  // namespace gt
