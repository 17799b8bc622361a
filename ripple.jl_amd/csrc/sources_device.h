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

#include <type_traits>

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
    T ls_inv_br;                 // 1 / ls_br (the FP32 path multiplies by it)
    T wc_rate, wc_thr, wc_sp;    // whitecapping: rate, saturation threshold and power
    const T *wc_factor;          // whitecapping: <kappa^p>_cell / kappa_ref^p per bin
    // FP32 forms (null in FP64, which keeps the reference's operation order):
    // per bin cos and sin of theta, log k, k / sigma (= 1 / C), sigma k^3 and
    // sigma^q; per column cos and sin of theta_u and log z1 (wind kernel).
    const T *cth_cell, *sth_cell, *logk_cell, *invc_cell, *sk3_cell, *sq_cell;
    const T *cos_tu, *sin_tu, *log_z1;
    // learned surrogate: where surrogate != 0 and use_plain[col] has neither
    // bit 0 (outside the training envelope) nor bit 2 (failed audit), the
    // network term is rate[col + rate_ld * t] * out_scale[t] * N; it is the
    // whole source (surrogate_terms 0) or the wind input only, the
    // dissipation staying the plain bundle's (surrogate_terms 1).  The rates
    // are stored in FP16 (relative rounding 2^-11, well inside the network's
    // own error); one that overflows FP16 flags its column non-finite, which
    // sends the column to the plain bundle.
    int surrogate, surrogate_terms;
    const _Float16 *rate;
    const float *out_scale;
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

// A table entry read through the constant address space: at a wave-uniform
// index this is a scalar (s_load) load, which costs no vector memory or VALU
// instruction.  Only for tables that no kernel writes while it reads them.
template <typename T>
__device__ __forceinline__ T sld(const T *p, size_t i)
{
    return ((const __attribute__((address_space(4))) T *)p)[i];
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

// FP32 _jp_rate from tabulated factors: cos(theta - theta_u) by the
// angle-difference identity, log(k z1) = log k + log z1, u*/C = u* (k /
// sigma), and one reciprocal for kvk / (inv_age cos).  The same formula as
// jp_rate with FP32 rounding in different places (a few ulp); no range
// reduction, log or divide per cell.  log_z1 is +inf where z1 <= 0, which
// gives Z = +inf and the zero rate of the growth window.
__device__ __forceinline__ float jp_rate_f32(const SourceParams<float> &s, float cth, float sth, float logk,
                                             float invc, float sigma, float ctu, float stu, float lz1, float ustar)
{
#pragma clang fp contract(off)
    const float cos_tu = fmaf(cth, ctu, sth * stu);
    if (!(cos_tu > 0.0f))
        return 0.0f;
    const float inv_age = fmaf(ustar, invc, s.z_alpha);
    const float Z = (logk + lz1) + s.kvk * __builtin_amdgcn_rcpf(inv_age * cos_tu);
    if (!(Z < 0.0f))
        return 0.0f;
    const float cp = s.p_in == 2.0f ? cos_tu * cos_tu : powf(cos_tu, s.p_in);
    const float Z2 = Z * Z;
    // Z < 0 here, so exp(Z) lies in (0, 1): the hardware exponential
    // (__expf, a few ulp) keeps the FP32 rate well inside the FP32 path's
    // tolerance and costs a tenth of expf
    return s.c0 * __expf(Z) * (Z2 * Z2) * (inv_age * inv_age) * cp * sigma;
}

// The per-column inputs of the per-cell source (what the wind kernel and the
// state kernel wrote for column col), loaded once so that a kernel walking
// several bins of one column reads them once.
template <typename T>
struct SourceColumn {
    T U10, ustar, theta_u, factor, m0, kbar, sbar;
    T cos_tu, sin_tu, log_z1;  // FP32 wind-input factors
    bool net;                  // the network supplies this column's source
};

template <typename T, bool AS1 = false>
__device__ __forceinline__ SourceColumn<T> source_column(const SourceParams<T> &s, unsigned col)
{
    SourceColumn<T> c{};
    c.net = s.surrogate && !(gld<AS1>(s.use_plain, col) & 5);
    if (c.net && s.surrogate_terms == 0)
        return c;
    if (s.wind_input && !c.net) {
        c.U10 = gld<AS1>(s.U10, col);
        c.ustar = gld<AS1>(s.ustar, col);
        if constexpr (std::is_same<T, float>::value) {
            c.cos_tu = gld<AS1>(s.cos_tu, col);
            c.sin_tu = gld<AS1>(s.sin_tu, col);
            c.log_z1 = gld<AS1>(s.log_z1, col);
        } else {
            c.theta_u = gld<AS1>(s.theta_u, col);
        }
        if (s.stress_cap)
            c.factor = gld<AS1>(s.factor, col);
    }
    if (s.dissipation == 1) {
        c.m0 = gld<AS1>(s.m0, col);
        c.kbar = gld<AS1>(s.kbar, col);
        c.sbar = gld<AS1>(s.sbar, col);
    } else if (s.dissipation == 3) {
        c.m0 = gld<AS1>(s.m0, col);
    }
    return c;
}

// The per-bin inputs of the per-cell source (tables indexed by t = m + nk*n),
// loaded together so that a kernel can issue them with its stencil loads.
template <typename T>
struct SourceBin {
    T k, sigma, C, theta;                // reference-order inputs (FP64)
    T cth, sth, logk, invc, sk3, sq;     // FP32 factors
    T wcf, out_scale;
};

template <typename T, bool AS1 = false>
__device__ __forceinline__ SourceBin<T> source_bin(const SourceParams<T> &s, int t)
{
    // Every table the plain bundle uses: with a network attached, columns
    // outside its envelope still take the plain source.
    SourceBin<T> b{};
    if (s.surrogate)
        b.out_scale = sld(s.out_scale, t);
    b.k = sld(s.k_cell, t);
    if (s.wind_input) {
        if constexpr (std::is_same<T, float>::value) {
            b.cth = sld(s.cth_cell, t);
            b.sth = sld(s.sth_cell, t);
            b.logk = sld(s.logk_cell, t);
            b.invc = sld(s.invc_cell, t);
            b.sigma = sld(s.sigma_cell, t);
        } else {
            b.sigma = sld(s.sigma_cell, t);
            b.C = sld(s.C_cell, t);
            b.theta = sld(s.theta_cell, t);
        }
    }
    if (s.dissipation == 2) {
        if constexpr (std::is_same<T, float>::value) {
            b.sk3 = sld(s.sk3_cell, t);
            b.sq = sld(s.sq_cell, t);
        } else {
            b.sigma = sld(s.sigma_cell, t);
        }
    } else if (s.dissipation == 3) {
        b.wcf = sld(s.wc_factor, t);
    }
    return b;
}

// Source tendency of one cell of column col and bin t (inputs c and b):
// positive - damping * a.
template <typename T, bool AS1 = false>
__device__ __forceinline__ T source_term_cb(const SourceParams<T> &s, const SourceColumn<T> &c, const SourceBin<T> &b,
                                            unsigned col, int t, T a, T *raw_out = nullptr)
{
#pragma clang fp contract(off)
    const bool net = c.net;
    T learned = T(0);
    if (net) {
        learned = (T)((float)gld<AS1>(s.rate, col + s.rate_ld * (size_t)t) * b.out_scale) * a;
        if (s.surrogate_terms == 0)
            return learned;
    }
    T positive = T(0), damping = T(0);
    const T k = b.k;
    if (s.wind_input && !net) {
        const T U10 = c.U10, ustar = c.ustar;
        if (U10 > T(0) && ustar > T(0) && k > T(0)) {
            T raw;
            if constexpr (std::is_same<T, float>::value)
                raw = jp_rate_f32(s, b.cth, b.sth, b.logk, b.invc, b.sigma, c.cos_tu, c.sin_tu, c.log_z1, ustar);
            else
                raw = jp_rate<T>(s, k, b.sigma, b.C, b.theta, c.theta_u, ustar);
            if (raw_out)
                *raw_out = raw;
            if (raw > T(0)) {
                const T rate = s.stress_cap ? raw * c.factor : raw;
                if (rate >= T(0))
                    positive += rate * a;
                else
                    damping += -rate;
            }
        }
    }
    if (s.dissipation == 1) {
        const T m0 = c.m0, kb = c.kbar;
        if (m0 > T(0) && kb > T(0) && k > T(0)) {
            const T ab = m0 * (kb * kb);
            const T ratio = k / kb;
            const T bracket = s.d1 * ratio + s.d2 * (ratio * ratio);
            const T rate = -s.cds * (ab * ab) * c.sbar * bracket;
            if (rate >= T(0))
                damping += rate;
            else
                positive += -rate * a;
        }
    } else if (s.dissipation == 2) {
        if (a > T(0) && k > T(0)) {
            T B, sq;
            if constexpr (std::is_same<T, float>::value) {
                B = a * b.sk3;  // sigma k^3 and sigma^q tabulated
                sq = b.sq;
            } else {
                B = a * b.sigma * (k * k * k);
                sq = dpow(b.sigma, s.ls_q);
            }
            // FP32 multiplies by the tabulated 1 / B_r; FP64 divides as the reference
            T over = (std::is_same<T, float>::value ? B * s.ls_inv_br : B / s.ls_br) - T(1);
            over = over > T(0) ? over : T(0);
            if (over > T(0)) {
                over = over < T(1e6) ? over : T(1e6);
                const T op = std::is_same<T, float>::value && s.ls_p == T(2) ? over * over : dpow(over, s.ls_p);
                const T rate = -s.ls_cds * sq * op;
                if (rate >= T(0))
                    damping += rate;
                else
                    positive += -rate * a;
            }
        }
    } else if (s.dissipation == 3) {
        // Sources.jl:588-598 with the column's zeroth moment from the state kernel.
        T excess = c.m0 / s.wc_thr - T(1);
        excess = excess > T(0) ? excess : T(0);
        const T rate = s.wc_rate * dpow(excess, s.wc_sp) * b.wcf;
        if (rate >= T(0))
            damping += rate;
        else
            positive += -rate * a;
    }
    if (net)
        return (positive - damping * a) + learned;
    return positive - damping * a;
}

// Source kinds a stage kernel is compiled for: none, the common bundle of
// the wind input (any drag, with or without the stress cap) and
// local-saturation or no dissipation, without a network (SOURCE_JP_LS), the
// same bundle with a network attached (SOURCE_NET_LS, FP32), or every
// configuration through source_term_cb (SOURCE_GENERIC).
enum { SOURCE_NONE = 0, SOURCE_JP_LS = 1, SOURCE_GENERIC = 2, SOURCE_NET_LS = 3 };

// Whether a source configuration takes the SOURCE_JP_LS path.
template <typename T>
__host__ __device__ inline bool source_is_jp_ls(const SourceParams<T> &s)
{
    return s.on && !s.surrogate && (s.wind_input == 0 || s.wind_input == 1) &&
           (s.dissipation == 0 || s.dissipation == 2);
}

// Whether a source configuration takes the SOURCE_NET_LS path: the
// SOURCE_JP_LS bundle (the network's plain fallback) with a network.
template <typename T>
__host__ __device__ inline bool source_is_net_ls(const SourceParams<T> &s)
{
    return s.on && s.surrogate && (s.wind_input == 0 || s.wind_input == 1) &&
           (s.dissipation == 0 || s.dissipation == 2);
}

// The per-column inputs of SOURCE_JP_LS: u* (0 where U10 <= 0, which the
// bundle skips), cos and sin of the wind angle, log z1 and the stress factor
// (1 without the cap).  byte_off is the column's byte offset in an (nx, ny)
// FP32 plane.
struct JpColumn {
    float us, ctu, stu, lz1, fac;
    bool net;  // SOURCE_NET_LS: the network supplies this column's source
};

template <bool AS1 = false, bool NET = false>
__device__ __forceinline__ JpColumn jp_column(const SourceParams<float> &s, unsigned byte_off)
{
    JpColumn c{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, false};
    if constexpr (NET)
        c.net = !(gld<AS1>(s.use_plain, byte_off >> 2) & 5);
    if (!s.wind_input)
        return c;
    typedef const __attribute__((address_space(1))) char gchar;
    auto ld = [&](const float *base) {
        return *(const __attribute__((address_space(1))) float *)((gchar *)base + byte_off);
    };
    const float U10 = ld(s.U10);
    c.us = U10 > 0.0f ? ld(s.ustar) : 0.0f;
    c.ctu = ld(s.cos_tu);
    c.stu = ld(s.sin_tu);
    c.lz1 = ld(s.log_z1);
    if (s.stress_cap)
        c.fac = ld(s.factor);
    return c;
}

// source_term_cb for a SOURCE_JP_LS configuration, in its operation order
// (so the two paths give the same bits): the bundle's wind input times the
// stress factor plus local-saturation dissipation (precomputed_sources.jl:
// 196-218, local_saturation.jl:18-52).
__device__ __forceinline__ float source_jp_ls(const SourceParams<float> &s, const JpColumn &c,
                                              const SourceBin<float> &b, float a, float &raw_out)
{
#pragma clang fp contract(off)
    float positive = 0.0f, damping = 0.0f;
    const float k = b.k;
    raw_out = 0.0f;  // the raw wind-input rate (0 outside the growth window)
    if (s.wind_input && c.us > 0.0f && k > 0.0f) {
        const float raw = jp_rate_f32(s, b.cth, b.sth, b.logk, b.invc, b.sigma, c.ctu, c.stu, c.lz1, c.us);
        raw_out = raw;
        if (raw > 0.0f) {
            const float rate = s.stress_cap ? raw * c.fac : raw;
            if (rate >= 0.0f)
                positive += rate * a;
            else
                damping += -rate;
        }
    }
    if (s.dissipation == 2 && a > 0.0f && k > 0.0f) {
        const float B = a * b.sk3;
        float over = B * s.ls_inv_br - 1.0f;
        over = over > 0.0f ? over : 0.0f;
        if (over > 0.0f) {
            over = over < 1e6f ? over : 1e6f;
            const float op = s.ls_p == 2.0f ? over * over : powf(over, s.ls_p);
            const float rate = -s.ls_cds * b.sq * op;
            if (rate >= 0.0f)
                damping += rate;
            else
                positive += -rate * a;
        }
    }
    return positive - damping * a;
}

// source_term_cb for a SOURCE_NET_LS configuration, in its operation order
// (the same bits): on a network column the learned rate (read from the FP16
// rate plane by the caller) times out_scale and the action, plus, with
// surrogate_terms 1, the plain local-saturation dissipation; elsewhere the
// SOURCE_JP_LS bundle.  raw_out is the wind input's raw rate, 0 on network
// columns.
__device__ __forceinline__ float source_net_ls(const SourceParams<float> &s, const JpColumn &c,
                                               const SourceBin<float> &b, float net_rate, float a, float &raw_out)
{
#pragma clang fp contract(off)
    if (!c.net)
        return source_jp_ls(s, c, b, a, raw_out);
    raw_out = 0.0f;
    const float learned = (net_rate * b.out_scale) * a;
    if (s.surrogate_terms == 0)
        return learned;
    float positive = 0.0f, damping = 0.0f;
    const float k = b.k;
    if (s.dissipation == 2 && a > 0.0f && k > 0.0f) {
        const float B = a * b.sk3;
        float over = B * s.ls_inv_br - 1.0f;
        over = over > 0.0f ? over : 0.0f;
        if (over > 0.0f) {
            over = over < 1e6f ? over : 1e6f;
            const float op = s.ls_p == 2.0f ? over * over : powf(over, s.ls_p);
            const float rate = -s.ls_cds * b.sq * op;
            if (rate >= 0.0f)
                damping += rate;
            else
                positive += -rate * a;
        }
    }
    return (positive - damping * a) + learned;
}

// Source tendency of one cell: positive - damping * a.  col is the column
// index i + nx*j, t the bin index m + nk*n.
template <typename T, bool AS1 = false>
__device__ __forceinline__ T source_term(const SourceParams<T> &s, unsigned col, int t, T a)
{
    return source_term_cb<T, AS1>(s, source_column<T, AS1>(s, col), source_bin<T, AS1>(s, t), col, t, a);
}

}  // namespace rpl
