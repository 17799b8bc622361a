// Fused tendency + SSP-RK3 stage kernels in the reference's arithmetic order.
//
// This is the parity path: every floating-point operation of
// Coupling/wave_current_refraction.jl:41-151 and Models/intrinsic_transport.jl:14-68
// is evaluated in the same order and without FMA contraction, so in FP64 it
// reproduces the CPU restatement bit for bit.  One thread per cell-bin; the
// RK3 sub-update and the combine (Models/time_step.jl:67-105) are fused into
// the same thread, which removes the reference's separate add_scaled!/combine!
// launches and its three full-array copies per step.
#include "rpl_internal.h"

#pragma clang fp contract(off)

namespace rpl {
namespace {

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double value = 2.220446049250313e-16; };
template <> struct Eps<float> { static constexpr float value = 1.1920928955078125e-7f; };

// wave_current_refraction.jl:41-55.
template <typename T>
__device__ __forceinline__ T weno5(T fm2, T fm1, T f0, T fp1, T fp2)
{
    const T eps = T(10000) * Eps<T>::value;
    const T c13_12 = T(13) / T(12);
    const T c1_4 = T(1) / T(4);
    T v0 = (T(1) / T(3)) * fm2 - (T(7) / T(6)) * fm1 + (T(11) / T(6)) * f0;
    T v1 = -(T(1) / T(6)) * fm1 + (T(5) / T(6)) * f0 + (T(1) / T(3)) * fp1;
    T v2 = (T(1) / T(3)) * f0 + (T(5) / T(6)) * fp1 - (T(1) / T(6)) * fp2;
    T d0 = fm2 - T(2) * fm1 + f0, e0 = fm2 - T(4) * fm1 + T(3) * f0;
    T d1 = fm1 - T(2) * f0 + fp1, e1 = fm1 - fp1;
    T d2 = f0 - T(2) * fp1 + fp2, e2 = T(3) * f0 - T(4) * fp1 + fp2;
    T b0 = c13_12 * (d0 * d0) + c1_4 * (e0 * e0);
    T b1 = c13_12 * (d1 * d1) + c1_4 * (e1 * e1);
    T b2 = c13_12 * (d2 * d2) + c1_4 * (e2 * e2);
    T q0 = eps + b0, q1 = eps + b1, q2 = eps + b2;
    T a0 = (T(1) / T(10)) / (q0 * q0);
    T a1 = (T(6) / T(10)) / (q1 * q1);
    T a2 = (T(3) / T(10)) / (q2 * q2);
    T sa = a0 + a1 + a2;
    return (a0 * v0 + a1 * v1 + a2 * v2) / sa;
}

// wave_current_refraction.jl:57-64.  Only the upwind branch is evaluated:
// the reference's ifelse selects one of two computed values, so evaluating
// just the selected one gives the identical result.
template <typename T>
__device__ __forceinline__ T weno5_face(T im2, T im1, T i0, T ip1, T ip2, T ip3, T vel, bool has_stencil)
{
    bool positive = vel >= T(0);
    if (!has_stencil)
        return positive ? i0 : ip1;
    return positive ? weno5<T>(im2, im1, i0, ip1, ip2) : weno5<T>(ip3, ip2, ip1, i0, im1);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

template <typename T>
__device__ __forceinline__ T clip0(T x) { return x < T(0) ? T(0) : x; }

template <typename T, int TRANSPORT, int STAGE, bool SRC>
__global__ __launch_bounds__(256) void stage_reference_kernel(const T *__restrict__ Nk, const T *__restrict__ N0,
                                                              T *out, StageParams<T> p)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int mn = blockIdx.z;
    const int nx = p.nx, ny = p.ny, nk = p.nk, nphi = p.nphi;
    if (i >= nx || j >= ny)
        return;
    const int m = mn % nk, n = mn / nk;
    const size_t plane = (size_t)nx * ny;
    const size_t sk = plane, sp = plane * nk;  // strides of m and n
    const size_t c = (size_t)i + (size_t)nx * j + sk * m + sp * n;
    const T x0 = Nk[c];
    T L = T(0);

    if (TRANSPORT != RPL_TRANSPORT_NONE) {
        const size_t row = c - i;
        const size_t c3 = (size_t)i + (size_t)nx * j + plane * m;
        T ux, uy;
        if (p.cgm) {
            // transport.jl:72-83 at the column's depth: (cg kx) / k
            const T mg = p.cgm[(size_t)i + (size_t)nx * j + plane * m];
            const int t = m + nk * n;
            ux = (mg * p.cdx[t]) / p.cdk[t];
            uy = (mg * p.cdy[t]) / p.cdk[t];
        } else {
            ux = p.cgx[m + nk * n];
            uy = p.cgy[m + nk * n];
        }
        T ck = T(0), cp = T(0);
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            const T kap = p.kappa[m], cph = p.cosp[n], sph = p.sinp[n];
            const T Uxc = p.Ux[c3], Uyc = p.Uy[c3];
            const T Uxx = p.Uxx[c3], Uxy = p.Uxy[c3], Uyx = p.Uyx[c3], Uyy = p.Uyy[c3];
            ck = -kap * (cph * cph * Uxx + cph * sph * (Uyx + Uxy) + sph * sph * Uyy);
            cp = cph * sph * (Uxx - Uyy) + sph * sph * Uyx - cph * cph * Uxy;
            ux = ux + Uxc;
            uy = uy + Uyc;
        }
        // x faces (periodic).
        const T xm3 = Nk[row + wrap(i - 3, nx)], xm2 = Nk[row + wrap(i - 2, nx)], xm1 = Nk[row + wrap(i - 1, nx)];
        const T xp1 = Nk[row + wrap(i + 1, nx)], xp2 = Nk[row + wrap(i + 2, nx)], xp3 = Nk[row + wrap(i + 3, nx)];
        const T Nxp = weno5_face<T>(xm2, xm1, x0, xp1, xp2, xp3, ux, true);
        const T Nxm = weno5_face<T>(xm3, xm2, xm1, x0, xp1, xp2, ux, true);
        const T flux_x = ux * (Nxp - Nxm) * p.dx_inv;
        // y faces (periodic).
        const size_t col = c - (size_t)nx * j;
        const T ym3 = Nk[col + (size_t)nx * wrap(j - 3, ny)], ym2 = Nk[col + (size_t)nx * wrap(j - 2, ny)];
        const T ym1 = Nk[col + (size_t)nx * wrap(j - 1, ny)], yp1 = Nk[col + (size_t)nx * wrap(j + 1, ny)];
        const T yp2 = Nk[col + (size_t)nx * wrap(j + 2, ny)], yp3 = Nk[col + (size_t)nx * wrap(j + 3, ny)];
        const T Nyp = weno5_face<T>(ym2, ym1, x0, yp1, yp2, yp3, uy, true);
        const T Nym = weno5_face<T>(ym3, ym2, ym1, x0, yp1, yp2, uy, true);
        const T flux_y = uy * (Nyp - Nym) * p.dy_inv;
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            // kappa faces (bounded, no flux through the outer faces).
            const size_t kb = c - sk * m;
            const bool has_p = (m - 2 >= 0) && (m + 3 <= nk - 1);
            const bool has_m = (m - 3 >= 0) && (m + 2 <= nk - 1);
            const T km3 = Nk[kb + sk * clampi(m - 3, nk)], km2 = Nk[kb + sk * clampi(m - 2, nk)];
            const T km1 = Nk[kb + sk * clampi(m - 1, nk)], kp1 = Nk[kb + sk * clampi(m + 1, nk)];
            const T kp2 = Nk[kb + sk * clampi(m + 2, nk)], kp3 = Nk[kb + sk * clampi(m + 3, nk)];
            const T Nkp = weno5_face<T>(km2, km1, x0, kp1, kp2, kp3, ck, has_p);
            const T Nkm = weno5_face<T>(km3, km2, km1, x0, kp1, kp2, ck, has_m);
            const T fkp = (m == nk - 1) ? T(0) : ck * Nkp;
            const T fkm = (m == 0) ? T(0) : ck * Nkm;
            const T flux_k = (fkp - fkm) * p.dk_inv;
            // phi faces (periodic).
            const size_t pb = c - sp * n;
            const T pm3 = Nk[pb + sp * wrap(n - 3, nphi)], pm2 = Nk[pb + sp * wrap(n - 2, nphi)];
            const T pm1 = Nk[pb + sp * wrap(n - 1, nphi)], pp1 = Nk[pb + sp * wrap(n + 1, nphi)];
            const T pp2 = Nk[pb + sp * wrap(n + 2, nphi)], pp3 = Nk[pb + sp * wrap(n + 3, nphi)];
            const T Npp = weno5_face<T>(pm2, pm1, x0, pp1, pp2, pp3, cp, true);
            const T Npm = weno5_face<T>(pm3, pm2, pm1, x0, pp1, pp2, cp, true);
            const T flux_p = cp * (Npp - Npm) * p.dp_inv;
            L = -(flux_x + flux_y + flux_k + flux_p);
        } else {
            L = -(flux_x + flux_y);
        }
    }
    if (SRC)  // tendencies.jl:88-94, G += source_tendency
        L = L + source_term<T>(p.src, (unsigned)(i + nx * j), m + nk * n, x0);
    if (p.src_add)
        L = L + p.src_add[c];

    if (STAGE == STAGE_TENDENCY) {
        out[c] = L;
    } else if (STAGE == STAGE_1) {
        out[c] = clip0<T>(x0 + p.dt * L);
    } else {
        const T a = STAGE == STAGE_2 ? T(0.75) : T(1.0 / 3.0);
        const T b = STAGE == STAGE_2 ? T(0.25) : T(2.0 / 3.0);
        const T st = clip0<T>(x0 + p.dt * L);
        out[c] = clip0<T>(a * N0[c] + b * st);
    }
}

template <typename T, int TRANSPORT, bool SRC>
hipError_t launch_t(const StageParams<T> &p, int stage, const T *Nk, const T *N0, T *out, hipStream_t s)
{
    dim3 block(64, 4, 1);
    dim3 grid((p.nx + 63) / 64, (p.ny + 3) / 4, p.nk * p.nphi);
    switch (stage) {
    case STAGE_TENDENCY:
        hipLaunchKernelGGL((stage_reference_kernel<T, TRANSPORT, STAGE_TENDENCY, SRC>), grid, block, 0, s, Nk, N0, out, p);
        break;
    case STAGE_1:
        hipLaunchKernelGGL((stage_reference_kernel<T, TRANSPORT, STAGE_1, SRC>), grid, block, 0, s, Nk, N0, out, p);
        break;
    case STAGE_2:
        hipLaunchKernelGGL((stage_reference_kernel<T, TRANSPORT, STAGE_2, SRC>), grid, block, 0, s, Nk, N0, out, p);
        break;
    default:
        hipLaunchKernelGGL((stage_reference_kernel<T, TRANSPORT, STAGE_3, SRC>), grid, block, 0, s, Nk, N0, out, p);
        break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_any(const StageParams<T> &p, int transport, int stage, const T *Nk, const T *N0, T *out,
                      hipStream_t s)
{
    if (p.src.on) {
        if (transport == RPL_TRANSPORT_REFRACTION)
            return launch_t<T, RPL_TRANSPORT_REFRACTION, true>(p, stage, Nk, N0, out, s);
        if (transport == RPL_TRANSPORT_INTRINSIC)
            return launch_t<T, RPL_TRANSPORT_INTRINSIC, true>(p, stage, Nk, N0, out, s);
        return launch_t<T, RPL_TRANSPORT_NONE, true>(p, stage, Nk, N0, out, s);
    }
    if (transport == RPL_TRANSPORT_REFRACTION)
        return launch_t<T, RPL_TRANSPORT_REFRACTION, false>(p, stage, Nk, N0, out, s);
    if (transport == RPL_TRANSPORT_INTRINSIC)
        return launch_t<T, RPL_TRANSPORT_INTRINSIC, false>(p, stage, Nk, N0, out, s);
    return launch_t<T, RPL_TRANSPORT_NONE, false>(p, stage, Nk, N0, out, s);
}

}  // namespace

hipError_t launch_stage_reference(const StageParams<double> &p, int transport, int stage, const double *Nk,
                                  const double *N0, double *out, hipStream_t s)
{
    return launch_any<double>(p, transport, stage, Nk, N0, out, s);
}

hipError_t launch_stage_reference(const StageParams<float> &p, int transport, int stage, const float *Nk,
                                  const float *N0, float *out, hipStream_t s)
{
    return launch_any<float>(p, transport, stage, Nk, N0, out, s);
}

}  // namespace rpl
