// Per-column source state: the wind of the step and the bundle state of each
// stage input (stress cap, mean-spectrum moments).
#include "rpl_internal.h"

#pragma clang fp contract(off)

namespace rpl {
namespace {

// Forcing/winds.jl:28-41 (two-point LinearStormTrack), 144-154
// (hurricane_speed for HollandHurricaneWind), 184-204 (wind_velocity,
// wind_speed, wind_angle); drag.jl:25-34 (friction_velocity).  FP64 per
// column, as the reference's host loop (precomputed_sources.jl:182-191).
template <typename T>
__global__ __launch_bounds__(256) void wind_columns_kernel(int nx, int ny, const double *__restrict__ xn,
                                                           const double *__restrict__ yn, double t, WindParams w,
                                                           T *U10, T *theta_u, T *ustar, T *cos_tu, T *sin_tu,
                                                           T *log_z1)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (ij >= (size_t)nx * ny)
        return;
    const int i = (int)(ij % nx), j = (int)(ij / nx);
    double speed, angle;
    if (w.field == RPL_WIND_UNIFORM) {
        speed = w.speed;
        angle = w.direction;
    } else {
        double cx, cy;
        if (t <= w.t0) {
            cx = w.x0;
            cy = w.y0;
        } else if (t >= w.t1) {
            cx = w.x1;
            cy = w.y1;
        } else {
            const double a = (t - w.t0) / (w.t1 - w.t0);
            cx = (1 - a) * w.x0 + a * w.x1;
            cy = (1 - a) * w.y0 + a * w.y1;
        }
        const double dx = xn[i] - cx, dy = yn[j] - cy;
        const double r = hypot(dx, dy);
        double vx, vy;
        if (r == 0) {
            vx = w.bx;
            vy = w.by;
        } else {
            const double shape_term = pow(w.rmax / r, w.shape);
            const double v = shape_term * exp(1 - shape_term);
            double sp = w.vmax * sqrt(v > 0 ? v : 0.0);
            if (r > w.radius)
                sp *= exp(-(r - w.radius) / w.decay);
            const double sg = (double)w.rotation;
            const double tx = -sg * dy / r, ty = sg * dx / r;
            const double rx = dx / r, ry = dy / r;
            vx = sp * (w.cos_inflow * tx - w.sin_inflow * rx) + w.bx;
            vy = sp * (w.cos_inflow * ty - w.sin_inflow * ry) + w.by;
        }
        speed = hypot(vx, vy);
        angle = atan2(vy, vx);
    }
    double cd;
    if (w.drag == RPL_DRAG_LINEAR) {
        cd = (0.8 + 0.065 * speed) * 1e-3;
    } else {
        cd = (8.058 + 0.967 * speed - 0.016 * (speed * speed)) * 1e-4;
        cd = cd > 4e-4 ? cd : 4e-4;
    }
    const double us = speed * sqrt(cd > 0 ? cd : 0.0);
    U10[ij] = (T)speed;
    theta_u[ij] = (T)angle;
    ustar[ij] = (T)us;
    if (cos_tu) {
        // Column factors of the FP32 wind input (jp_rate_f32); z1 from the
        // stored u*, as the per-cell form computes it.
        const double u = (double)(T)us, z1 = w.alpha0 * (u * u) / w.gravity;
        cos_tu[ij] = (T)cos(angle);
        sin_tu[ij] = (T)sin(angle);
        log_z1[ij] = z1 > 0 ? (T)log(z1) : (T)INFINITY;
    }
}

// precomputed_sources.jl:126-170 (stress factor) and :41-74 (mean-spectrum
// moments), thread per column, spectral loop n outer / m inner.  Consecutive
// threads read consecutive i of each (m, n) plane.
template <typename T>
__global__ __launch_bounds__(256) void source_state_kernel(int nx, int ny, int nk, int nphi, const T *__restrict__ N,
                                                           SourceParams<T> s, StateTables<T> tb, T *factor, T *m0o,
                                                           T *kbaro, T *sbaro)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (ij >= plane)
        return;
    // Where the network supplies the column's source, nothing (all terms) or
    // only the dissipation state (wind input from the network) is needed.
    const bool net = s.surrogate && !(s.use_plain[ij] & 3);
    if (net && s.surrogate_terms == 0)
        return;
    if (s.wind_input && s.stress_cap && !net) {
        const T U10 = s.U10[ij], ustar = s.ustar[ij];
        T f = T(1);
        if (U10 > T(0) && ustar > T(0)) {
            const T theta_u = s.theta_u[ij];
            const T tau_max = tb.rho_air * (ustar * ustar);
            T twx = T(0), twy = T(0);
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m) {
                    const T a = N[ij + plane * ((size_t)m + (size_t)nk * n)];
                    if (a > T(0)) {
                        const T k = tb.kappa[m];
                        if (k > T(0)) {
                            const T sigma = tb.sigma_k[m], C = sigma / k;
                            const T th = tb.phi[n];
                            const T rate = jp_rate<T>(s, k, sigma, C, th, theta_u, ustar);
                            if (rate > T(0)) {
                                const T mflux = rate * a / C * tb.w[m + nk * n] * tb.rho_w_g;
                                twx += mflux * tb.cos_phi[n];
                                twy += mflux * tb.sin_phi[n];
                            }
                        }
                    }
                }
            const T par = twx * dcos(theta_u) + twy * dsin(theta_u);
            const T eff = par > T(0) ? par : T(0);
            f = (eff > tb.cap_eps * tau_max && tau_max > T(0)) ? tb.cap_eps * tau_max / eff : T(1);
        }
        factor[ij] = f;
    }
    if (s.dissipation == 3) {
        // local_zeroth_moment (Sources.jl:1471-1480): every cell, n outer / m inner.
        T m0 = T(0);
        for (int n = 0; n < nphi; ++n)
            for (int m = 0; m < nk; ++m)
                m0 += N[ij + plane * ((size_t)m + (size_t)nk * n)] * tb.w[m + nk * n];
        m0o[ij] = m0;
    }
    if (s.dissipation == 1) {
        T m0 = T(0), kp = T(0), sp = T(0);
        for (int n = 0; n < nphi; ++n)
            for (int m = 0; m < nk; ++m) {
                const T a = N[ij + plane * ((size_t)m + (size_t)nk * n)];
                if (a > T(0)) {
                    const T k = tb.kappa[m];
                    if (k > T(0)) {
                        const T wt = tb.w[m + nk * n];
                        m0 += a * wt;
                        kp += a * wt * tb.k_pow[m];
                        sp += a * wt * tb.sigma_pow[m];
                    }
                }
            }
        m0o[ij] = m0;
        if (m0 > T(0)) {
            kbaro[ij] = dpow(kp / m0, tb.inv_p);
            sbaro[ij] = dpow(sp / m0, tb.inv_p);
        } else {
            kbaro[ij] = T(0);
            sbaro[ij] = T(0);
        }
    }
}

// FP32 form of the same state: one 256-thread workgroup per 64 consecutive
// columns, four bin groups per column (group g takes every fourth bin of the
// n-outer / m-inner order), partial sums combined in LDS in group order.
// Four times the parallelism of thread-per-column at the same coalesced
// reads, which matters for small planes.  `active` lets an ensemble launch
// skip members without a column state.
__device__ __forceinline__ void state_block_f32(int nx, int ny, int nk, int nphi, const float *__restrict__ N,
                                                const SourceParams<float> &s, const StateTables<float> &tb,
                                                float *factor, float *m0o, float *kbaro, float *sbaro, size_t col0)
{
    constexpr int G = 4;
    __shared__ float red[5][G][64];
    const size_t plane = (size_t)nx * ny;
    const int tx = threadIdx.x & 63, g = threadIdx.x >> 6;
    const size_t ij = col0 + tx;
    bool active = ij < plane;
    const bool net = active && s.surrogate && !(gld<true>(s.use_plain, ij) & 3);
    if (net && s.surrogate_terms == 0)
        active = false;  // the network supplies this column's whole source
    // with the network supplying the wind input, only the dissipation state is needed
    const bool cap = s.wind_input && s.stress_cap && !net, ms = s.dissipation == 1, wc = s.dissipation == 3;
    float twx = 0.0f, twy = 0.0f, m0 = 0.0f, kp = 0.0f, sp = 0.0f;
    float U10 = 0.0f, ustar = 0.0f, ctu = 0.0f, stu = 0.0f, lz1 = 0.0f;
    if (active && cap) {
        U10 = gld<true>(s.U10, ij);
        ustar = gld<true>(s.ustar, ij);
        ctu = gld<true>(s.cos_tu, ij);
        stu = gld<true>(s.sin_tu, ij);
        lz1 = gld<true>(s.log_z1, ij);
    }
    const bool wind = cap && U10 > 0.0f && ustar > 0.0f;
    if (active && (wind || ms || wc)) {
        // What bin t reads; the next bin's values are loaded before this
        // bin's arithmetic (the loop is latency-bound otherwise).
        struct BinIn {
            float a, k, w, invc, cph, sph, logk, sk, kpw, spw;
        };
        auto load = [&](int t, int m, int n) {
            BinIn v;
            v.a = N[ij + plane * (size_t)t];
            v.k = gld<true>(tb.kappa, m);
            v.w = gld<true>(tb.w, t);
            if (wind) {
                v.invc = gld<true>(tb.inv_c, m);
                v.cph = gld<true>(tb.cos_phi, n);
                v.sph = gld<true>(tb.sin_phi, n);
                v.logk = gld<true>(tb.log_kappa, m);
                v.sk = gld<true>(tb.sigma_k, m);
            }
            if (ms) {
                v.kpw = gld<true>(tb.k_pow, m);
                v.spw = gld<true>(tb.sigma_pow, m);
            }
            return v;
        };
        int m = g, n = 0;
        while (m >= nk) {
            m -= nk;
            ++n;
        }
        const int nbins = nk * nphi;
        BinIn cur{};
        if (g < nbins)
            cur = load(g, m, n);
        for (int t = g; t < nbins; t += G) {
            m += G;
            while (m >= nk) {
                m -= nk;
                ++n;
            }
            BinIn nxt = cur;
            if (t + G < nbins)
                nxt = load(t + G, m, n);
            const float a = cur.a, k = cur.k;
            if (wc)
                m0 += a * cur.w;  // zeroth moment over every cell
            if (a > 0.0f && k > 0.0f) {
                if (wind) {
                    const float rate = jp_rate_f32(s, cur.cph, cur.sph, cur.logk, cur.invc, cur.sk, ctu, stu, lz1,
                                                   ustar);
                    if (rate > 0.0f) {
                        const float mflux = rate * a * cur.invc * cur.w * tb.rho_w_g;
                        twx += mflux * cur.cph;
                        twy += mflux * cur.sph;
                    }
                }
                if (ms) {
                    const float wt = cur.w;
                    m0 += a * wt;
                    kp += a * wt * cur.kpw;
                    sp += a * wt * cur.spw;
                }
            }
            cur = nxt;
        }
    }
    red[0][g][tx] = twx;
    red[1][g][tx] = twy;
    red[2][g][tx] = m0;
    red[3][g][tx] = kp;
    red[4][g][tx] = sp;
    __syncthreads();
    if (g != 0 || !active)
        return;
    twx = red[0][0][tx];
    twy = red[1][0][tx];
    m0 = red[2][0][tx];
    kp = red[3][0][tx];
    sp = red[4][0][tx];
    for (int h = 1; h < G; ++h) {
        twx += red[0][h][tx];
        twy += red[1][h][tx];
        m0 += red[2][h][tx];
        kp += red[3][h][tx];
        sp += red[4][h][tx];
    }
    if (cap) {
        float f = 1.0f;
        if (wind) {
            const float tau_max = tb.rho_air * (ustar * ustar);
            const float par = twx * ctu + twy * stu;
            const float eff = par > 0.0f ? par : 0.0f;
            f = (eff > tb.cap_eps * tau_max && tau_max > 0.0f) ? tb.cap_eps * tau_max / eff : 1.0f;
        }
        gp_store(factor, ij, f);
    }
    if (wc)
        gp_store(m0o, ij, m0);
    if (ms) {
        gp_store(m0o, ij, m0);
        gp_store(kbaro, ij, m0 > 0.0f ? powf(kp / m0, tb.inv_p) : 0.0f);
        gp_store(sbaro, ij, m0 > 0.0f ? powf(sp / m0, tb.inv_p) : 0.0f);
    }
}

// The bundle's stress cap of one column (precomputed_sources.jl:126-170)
// from the per-kappa wave-supported stress sums of a stage kernel, summed
// over kappa in bin order.  Columns without wind have zero sums (factor 1).
__device__ __forceinline__ void stress_factor_column(size_t ij, size_t plane, int nk, const float *__restrict__ tx,
                                                     const float *__restrict__ ty, const SourceParams<float> &s,
                                                     const StateTables<float> &tb, float *factor)
{
    float twx = 0.0f, twy = 0.0f;
    for (int m = 0; m < nk; ++m) {
        twx += gld<true>(tx, ij + plane * (size_t)m);
        twy += gld<true>(ty, ij + plane * (size_t)m);
    }
    const float ustar = gld<true>(s.ustar, ij);
    const float ctu = gld<true>(s.cos_tu, ij), stu = gld<true>(s.sin_tu, ij);
    const float tau_max = tb.rho_air * (ustar * ustar);
    const float par = twx * ctu + twy * stu;
    const float eff = par > 0.0f ? par : 0.0f;
    const float f = (eff > tb.cap_eps * tau_max && tau_max > 0.0f) ? tb.cap_eps * tau_max / eff : 1.0f;
    gp_store(factor, ij, f);
}

__global__ __launch_bounds__(256) void stress_factor_kernel(int nx, int ny, int nk, const float *__restrict__ tx,
                                                            const float *__restrict__ ty, SourceParams<float> s,
                                                            StateTables<float> tb, float *factor)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (ij < plane)
        stress_factor_column(ij, plane, nk, tx, ty, s, tb, factor);
}

__global__ __launch_bounds__(256) void source_state_f32_kernel(int nx, int ny, int nk, int nphi,
                                                               const float *__restrict__ N, SourceParams<float> s,
                                                               StateTables<float> tb, float *factor, float *m0o,
                                                               float *kbaro, float *sbaro)
{
    state_block_f32(nx, ny, nk, nphi, N, s, tb, factor, m0o, kbaro, sbaro, (size_t)blockIdx.x * 64);
}

// All members' column states of one RK3 stage in one launch: the members'
// 64-column blocks are numbered member after member (EnsembleState::blk0).
__global__ __launch_bounds__(256) void ensemble_state_kernel(const EnsembleState *__restrict__ mem, int members)
{
    const long long b = blockIdx.x;
    int u = 0, v = members - 1;
    while (u < v) {
        const int c = (u + v + 1) >> 1;
        if (mem[c].blk0 <= b)
            u = c;
        else
            v = c - 1;
    }
    const EnsembleState &e = mem[__builtin_amdgcn_readfirstlane(u)];
    const size_t col0 = (size_t)__builtin_amdgcn_readfirstlane((int)(b - e.blk0)) * 64;
    if (e.tau_x) {  // the stress factor from the previous stage's sums
        const size_t plane = (size_t)e.nx * e.ny, ij = col0 + threadIdx.x;
        if (threadIdx.x < 64 && ij < plane)
            stress_factor_column(ij, plane, e.nk, e.tau_x, e.tau_y, e.s, e.tb, e.factor);
        return;
    }
    state_block_f32(e.nx, e.ny, e.nk, e.nphi, e.N, e.s, e.tb, e.factor, e.m0, e.kbar, e.sbar, col0);
}

}  // namespace

template <typename T>
hipError_t launch_wind_columns(int nx, int ny, const double *xn, const double *yn, double t, const WindParams &w,
                               T *U10, T *theta_u, T *ustar, T *cos_tu, T *sin_tu, T *log_z1, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    hipLaunchKernelGGL(wind_columns_kernel<T>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, nx, ny, xn, yn,
                       t, w, U10, theta_u, ustar, cos_tu, sin_tu, log_z1);
    return hipGetLastError();
}

namespace {
// Per-cell plain-bundle rates normalised per bin, r = S / (N * unit[t]): the
// learned surrogate's training targets.  Where N <= 0 the target is the rate
// of a vanishing action (evaluated at N = 1e-30), which is what the features
// of an empty bin cannot tell apart from a nearly empty one.  Bin-major like
// the action (r[ij + plane * t]).
template <typename T>
__global__ __launch_bounds__(256) void source_rates_kernel(int nx, int ny, const T *__restrict__ N,
                                                           SourceParams<T> s, const float *__restrict__ unit,
                                                           float *r)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (ij >= plane)
        return;
    const T a0 = N[ij + plane * t];
    const T a = a0 > T(1e-30) ? a0 : T(1e-30);
    const float v = (float)(source_term<T>(s, (unsigned)ij, t, a) / a) / unit[t];
    r[ij + plane * t] = v;
}
}  // namespace

template <typename T>
hipError_t launch_source_rates(int nx, int ny, int nbins, const T *N, const SourceParams<T> &sp, const float *unit,
                               float *r, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    hipLaunchKernelGGL(source_rates_kernel<T>, dim3((unsigned)((plane + 255) / 256), nbins), dim3(256), 0, s, nx, ny,
                       N, sp, unit, r);
    return hipGetLastError();
}

template hipError_t launch_source_rates<float>(int, int, int, const float *, const SourceParams<float> &,
                                               const float *, float *, hipStream_t);
template hipError_t launch_source_rates<double>(int, int, int, const double *, const SourceParams<double> &,
                                                const float *, float *, hipStream_t);

template <typename T>
hipError_t launch_source_state(int nx, int ny, int nk, int nphi, const T *N, const SourceParams<T> &sp,
                               const StateTables<T> &tb, T *factor, T *m0, T *kbar, T *sbar, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    if (sizeof(T) == 4) {
        hipLaunchKernelGGL(source_state_f32_kernel, dim3((unsigned)((plane + 63) / 64)), dim3(256), 0, s, nx, ny, nk,
                           nphi, (const float *)N, reinterpret_cast<const SourceParams<float> &>(sp),
                           reinterpret_cast<const StateTables<float> &>(tb), (float *)factor, (float *)m0,
                           (float *)kbar, (float *)sbar);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(source_state_kernel<T>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, nx, ny, nk,
                       nphi, N, sp, tb, factor, m0, kbar, sbar);
    return hipGetLastError();
}

template hipError_t launch_wind_columns<float>(int, int, const double *, const double *, double, const WindParams &,
                                              float *, float *, float *, float *, float *, float *, hipStream_t);
template hipError_t launch_wind_columns<double>(int, int, const double *, const double *, double, const WindParams &,
                                               double *, double *, double *, double *, double *, double *,
                                               hipStream_t);
hipError_t launch_stress_factor_from_sums(int nx, int ny, int nk, const float *tau_x, const float *tau_y,
                                          const SourceParams<float> &sp, const StateTables<float> &tb, float *factor,
                                          hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    hipLaunchKernelGGL(stress_factor_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, nx, ny, nk,
                       tau_x, tau_y, sp, tb, factor);
    return hipGetLastError();
}

hipError_t launch_ensemble_state(const EnsembleState *mem, int members, long long blocks, hipStream_t s)
{
    if (members < 1 || blocks < 1 || blocks > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(ensemble_state_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mem, members);
    return hipGetLastError();
}

template hipError_t launch_source_state<float>(int, int, int, int, const float *, const SourceParams<float> &,
                                               const StateTables<float> &, float *, float *, float *, float *,
                                               hipStream_t);
template hipError_t launch_source_state<double>(int, int, int, int, const double *, const SourceParams<double> &,
                                                const StateTables<double> &, double *, double *, double *, double *,
                                                hipStream_t);

}  // namespace rpl
