// Device source terms shared by the stage kernels (internal).
//
// The per-cell source of the PrecomputedSources bundle
// (Physics/Packages/precomputed_sources.jl:240-261) is positive - damping * N
// with
//   wind input   rate = raw Janssen rate (pressure_correlation.jl:109-136)
//                times the column's stress factor (precomputed_sources.jl:145-151),
//                split by split_growth_rate (Sources/Sources.jl:430-431);
//   dissipation  MeanSpectrumWhitecapping on the column's (m0, k_bar, sigma_bar)
//                (mean_spectrum.jl:117-135) or LocalSaturationDissipation
//                (local_saturation.jl:30-47), split by split_damping_rate
//                (Sources.jl:433-434).
// The per-cell wavenumber, direction, sigma and phase speed are the
// reference's hypot/atan/sqrt values, tabulated per bin on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace rpl {

template <typename T>
struct SourceParams {
    int on;           // any source term present
    int wind_input;   // RPL_WIND_INPUT_*
    int stress_cap;   // bundle stress cap
    int dissipation;  // RPL_DISSIPATION_*
    // per column, written by the wind kernel once per step
    const T *U10, *theta_u, *ustar;
    // per column, written by the state kernel on every stage input
    const T *factor, *m0, *kbar, *sbar;
    // per bin (nk, nphi), m fastest
    const T *k_cell, *theta_cell, *sigma_cell, *C_cell;
    T c0;       // (rho_air / rho_water) * (beta_max / von_karman^2)
    T z_alpha, p_in, alpha0, kvk, g;
    T cds, d1, d2;              // mean-spectrum whitecapping
    T ls_cds, ls_br, ls_p, ls_q; // local saturation
    T wc_rate, wc_thr, wc_sp;    // whitecapping: rate, saturation threshold and power
    const T *wc_factor;          // whitecapping: <kappa^p>_cell / kappa_ref^p per bin
    // learned surrogate: where surrogate != 0 and use_plain[col] has neither
    // bit 0 (outside the training envelope) nor bit 2 (failed audit), the
    // network term is rate[col + rate_ld * t] * out_scale[t] * N; it is the
    // whole source (surrogate_terms 0) or the wind input only, the
    // dissipation staying the plain bundle's (surrogate_terms 1).
    int surrogate, surrogate_terms;
    const float *rate, *out_scale;
    size_t rate_ld;
    const unsigned char *use_plain;
};

// Loads of a pointer parameter.  With AS1 the load goes through an
// address-space-1 (global) pointer: pointers read from memory (the ensemble's
// member descriptors) are generic to the compiler, which would otherwise emit
// flat loads with 64-bit address arithmetic.  Pointers that arrive as kernel
// arguments are loaded as they are (AS1 = false); the explicit cast costs the
// per-model kernels a few percent.
template <bool AS1, typename T>
__device__ __forceinline__ T gld(const T *p, size_t i)
{
    if (AS1)
        return ((const __attribute__((address_space(1))) T *)p)[i];
    return p[i];
}

template <typename T>
__device__ __forceinline__ void gp_store(T *p, size_t i, T v)
{
    ((__attribute__((address_space(1))) T *)p)[i] = v;
}

__device__ __forceinline__ float dcos(float x) { return cosf(x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float dsin(float x) { return sinf(x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double dpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }

// precomputed_sources.jl:113-124 (_jp_rate): zero outside the growth window.
template <typename T>
__device__ __forceinline__ T jp_rate(const SourceParams<T> &s, T k, T sigma, T C, T theta, T theta_u, T ustar)
{
#pragma clang fp contract(off)
    const T cos_tu = dcos(theta - theta_u);
    if (!(cos_tu > T(0)))
        return T(0);
    const T z1 = s.alpha0 * (ustar * ustar) / s.g;
    if (!(z1 > T(0)))
        return T(0);
    const T inv_age = ustar / C + s.z_alpha;
    const T Z = dlog(k * z1) + s.kvk / (inv_age * cos_tu);
    if (Z >= T(0))
        return T(0);
    const T cp = s.p_in == T(2) ? cos_tu * cos_tu : dpow(cos_tu, s.p_in);
    const T Z2 = Z * Z;
    return s.c0 * dexp(Z) * (Z2 * Z2) * (inv_age * inv_age) * cp * sigma;
}

// Source tendency of one cell: positive - damping * a.  col is the column
// index i + nx*j, t the bin index m + nk*n.
template <typename T, bool AS1 = false>
__device__ __forceinline__ T source_term(const SourceParams<T> &s, unsigned col, int t, T a)
{
#pragma clang fp contract(off)
    const bool net = s.surrogate && !(gld<AS1>(s.use_plain, col) & 5);
    T learned = T(0);
    if (net) {
        learned = (T)(gld<AS1>(s.rate, col + s.rate_ld * (size_t)t) * gld<AS1>(s.out_scale, t)) * a;
        if (s.surrogate_terms == 0)
            return learned;
    }
    T positive = T(0), damping = T(0);
    const T k = gld<AS1>(s.k_cell, t);
    if (s.wind_input && !net) {
        const T U10 = gld<AS1>(s.U10, col), ustar = gld<AS1>(s.ustar, col);
        if (U10 > T(0) && ustar > T(0) && k > T(0)) {
            const T raw = jp_rate<T>(s, k, gld<AS1>(s.sigma_cell, t), gld<AS1>(s.C_cell, t), gld<AS1>(s.theta_cell, t), gld<AS1>(s.theta_u, col), ustar);
            if (raw > T(0)) {
                const T rate = s.stress_cap ? raw * gld<AS1>(s.factor, col) : raw;
                if (rate >= T(0))
                    positive += rate * a;
                else
                    damping += -rate;
            }
        }
    }
    if (s.dissipation == 1) {
        const T m0 = gld<AS1>(s.m0, col), kb = gld<AS1>(s.kbar, col);
        if (m0 > T(0) && kb > T(0) && k > T(0)) {
            const T ab = m0 * (kb * kb);
            const T ratio = k / kb;
            const T bracket = s.d1 * ratio + s.d2 * (ratio * ratio);
            const T rate = -s.cds * (ab * ab) * gld<AS1>(s.sbar, col) * bracket;
            if (rate >= T(0))
                damping += rate;
            else
                positive += -rate * a;
        }
    } else if (s.dissipation == 2) {
        if (a > T(0) && k > T(0)) {
            const T sigma = gld<AS1>(s.sigma_cell, t);
            const T B = a * sigma * (k * k * k);
            T over = B / s.ls_br - T(1);
            over = over > T(0) ? over : T(0);
            if (over > T(0)) {
                over = over < T(1e6) ? over : T(1e6);
                const T rate = -s.ls_cds * dpow(sigma, s.ls_q) * dpow(over, s.ls_p);
                if (rate >= T(0))
                    damping += rate;
                else
                    positive += -rate * a;
            }
        }
    } else if (s.dissipation == 3) {
        // Sources.jl:588-598 with the column's zeroth moment from the state kernel.
        const T m0 = gld<AS1>(s.m0, col);
        T excess = m0 / s.wc_thr - T(1);
        excess = excess > T(0) ? excess : T(0);
        const T rate = s.wc_rate * dpow(excess, s.wc_sp) * gld<AS1>(s.wc_factor, t);
        if (rate >= T(0))
            damping += rate;
        else
            positive += -rate * a;
    }
    if (net)
        return (positive - damping * a) + learned;
    return positive - damping * a;
}

}  // namespace rpl
