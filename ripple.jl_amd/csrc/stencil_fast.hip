// FP32 fused tendency + SSP-RK3 stage kernel, restructured for gfx950.
//
// Same discretisation as Coupling/wave_current_refraction.jl:41-151 and
// Models/intrinsic_transport.jl:14-68 (WENO5-JS with the reference's FP32
// epsilon 1e4*eps(Float32), cell-centre velocity on both faces, first-order
// kappa fallback, zero outer kappa flux), with the arithmetic reorganised:
//   * only the upwind reconstruction of each face is evaluated;
//   * the reconstruction is written as v1 + w0 (v0 - v1) + w2 (v2 - v1), whose
//     two corrections are third differences of the data;
//   * the nonlinear weights are formed from the ratios q1/q0 and q1/q2
//     (q_k = eps + beta_k), so there is one reciprocal per ratio and one for
//     the final normalisation instead of four IEEE divides, and nothing
//     overflows for any finite action;
//   * the current gradients are centred differences of U formed in-kernel
//     (wave_current_refraction.jl:7-19) instead of four extra arrays.
// Results differ from the reference-order kernel by rounding only.
#include "rpl_internal.h"

namespace rpl {
namespace {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Upwind-left WENO5 value at face i+1/2 from (f[i-2] .. f[i+2]).
__device__ __forceinline__ float weno5_left(float fm2, float fm1, float f0, float fp1, float fp2)
{
    // Smoothness indicators scaled by 4 (and epsilon with them): the weights
    // depend only on ratios, so the common factor cancels.
    const float eps4 = 4.0f * (10000.0f * 1.1920928955078125e-7f);
    const float c13_3 = 13.0f / 3.0f;
    const float d0 = (fm2 + f0) - 2.0f * fm1;
    const float d1 = (fm1 + fp1) - 2.0f * f0;
    const float d2 = (f0 + fp2) - 2.0f * fp1;
    const float e0 = d0 + 2.0f * (f0 - fm1);
    const float e1 = fm1 - fp1;
    const float e2 = d2 + 2.0f * (f0 - fp1);
    const float q0 = fmaf(c13_3 * d0, d0, fmaf(e0, e0, eps4));
    const float q1 = fmaf(c13_3 * d1, d1, fmaf(e1, e1, eps4));
    const float q2 = fmaf(c13_3 * d2, d2, fmaf(e2, e2, eps4));
    const float r0 = q1 * frcp(q0), r2 = q1 * frcp(q2);
    const float s0 = r0 * r0, s2 = r2 * r2;
    // alpha0/alpha1 = s0/6, alpha2/alpha1 = s2/2;  v0-v1 = T0/3, v2-v1 = T1/6.
    const float T0 = d0 - d1;  // fm2 - 3 fm1 + 3 f0 - fp1
    const float T1 = d1 - d2;  // fm1 - 3 f0 + 3 fp1 - fp2
    const float v1 = (fmaf(5.0f, f0, 2.0f * fp1) - fm1) * (1.0f / 6.0f);
    const float num = fmaf(2.0f * s0, T0, 3.0f * s2 * T1);
    const float den = fmaf(6.0f, s0, fmaf(18.0f, s2, 36.0f));
    return fmaf(num, frcp(den), v1);
}

__device__ __forceinline__ float face(float im2, float im1, float i0, float ip1, float ip2, float ip3, float vel)
{
    return vel >= 0.0f ? weno5_left(im2, im1, i0, ip1, ip2) : weno5_left(ip3, ip2, ip1, i0, im1);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
__device__ __forceinline__ float clip0(float x) { return x < 0.0f ? 0.0f : x; }

template <int TRANSPORT, int STAGE>
__global__ __launch_bounds__(256) void stage_fast_kernel(const float *__restrict__ Nk, const float *__restrict__ N0,
                                                         float *out, StageParams<float> p)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int mn = blockIdx.z;
    const int nx = p.nx, ny = p.ny, nk = p.nk, nphi = p.nphi;
    if (i >= nx || j >= ny)
        return;
    const int m = mn % nk, n = mn / nk;
    const size_t plane = (size_t)nx * ny;
    const size_t sk = plane, sp = plane * nk;
    const size_t c = (size_t)i + (size_t)nx * j + sk * m + sp * n;
    const float x0 = Nk[c];
    float L = 0.0f;
    if (TRANSPORT != RPL_TRANSPORT_NONE) {
        const int ip1 = wrap(i + 1, nx), im1 = wrap(i - 1, nx);
        const int jp1 = wrap(j + 1, ny), jm1 = wrap(j - 1, ny);
        float ux = p.cgx[m + nk * n], uy = p.cgy[m + nk * n];
        float ck = 0.0f, cp = 0.0f;
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            const size_t b3 = plane * m;
            const float *Ux = p.Ux + b3, *Uy = p.Uy + b3;
            const float hx = 0.5f * p.dx_inv, hy = 0.5f * p.dy_inv;
            const float Uxx = (Ux[ip1 + (size_t)nx * j] - Ux[im1 + (size_t)nx * j]) * hx;
            const float Uxy = (Ux[i + (size_t)nx * jp1] - Ux[i + (size_t)nx * jm1]) * hy;
            const float Uyx = (Uy[ip1 + (size_t)nx * j] - Uy[im1 + (size_t)nx * j]) * hx;
            const float Uyy = (Uy[i + (size_t)nx * jp1] - Uy[i + (size_t)nx * jm1]) * hy;
            const float kap = p.kappa[m], cs = p.cosp[n], sn = p.sinp[n];
            const float cc = cs * cs, ss = sn * sn, csn = cs * sn;
            ck = -kap * (cc * Uxx + csn * (Uyx + Uxy) + ss * Uyy);
            cp = csn * (Uxx - Uyy) + ss * Uyx - cc * Uxy;
            ux += Ux[i + (size_t)nx * j];
            uy += Uy[i + (size_t)nx * j];
        }
        const size_t row = c - i;
        const float xm3 = Nk[row + wrap(i - 3, nx)], xm2 = Nk[row + wrap(i - 2, nx)], xm1 = Nk[row + im1];
        const float xp1 = Nk[row + ip1], xp2 = Nk[row + wrap(i + 2, nx)], xp3 = Nk[row + wrap(i + 3, nx)];
        const float fx = ux * (face(xm2, xm1, x0, xp1, xp2, xp3, ux) - face(xm3, xm2, xm1, x0, xp1, xp2, ux));
        const size_t col = c - (size_t)nx * j;
        const float ym3 = Nk[col + (size_t)nx * wrap(j - 3, ny)], ym2 = Nk[col + (size_t)nx * wrap(j - 2, ny)];
        const float ym1 = Nk[col + (size_t)nx * jm1], yp1 = Nk[col + (size_t)nx * jp1];
        const float yp2 = Nk[col + (size_t)nx * wrap(j + 2, ny)], yp3 = Nk[col + (size_t)nx * wrap(j + 3, ny)];
        const float fy = uy * (face(ym2, ym1, x0, yp1, yp2, yp3, uy) - face(ym3, ym2, ym1, x0, yp1, yp2, uy));
        L = -fmaf(fx, p.dx_inv, fy * p.dy_inv);
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            const size_t kb = c - sk * m;
            const float km3 = Nk[kb + sk * clampi(m - 3, nk)], km2 = Nk[kb + sk * clampi(m - 2, nk)];
            const float km1 = Nk[kb + sk * clampi(m - 1, nk)], kp1 = Nk[kb + sk * clampi(m + 1, nk)];
            const float kp2 = Nk[kb + sk * clampi(m + 2, nk)], kp3 = Nk[kb + sk * clampi(m + 3, nk)];
            const bool pos = ck >= 0.0f;
            float Nkp, Nkm;
            if ((m - 2 >= 0) && (m + 3 <= nk - 1))
                Nkp = face(km2, km1, x0, kp1, kp2, kp3, ck);
            else
                Nkp = pos ? x0 : kp1;
            if ((m - 3 >= 0) && (m + 2 <= nk - 1))
                Nkm = face(km3, km2, km1, x0, kp1, kp2, ck);
            else
                Nkm = pos ? km1 : x0;
            const float fkp = (m == nk - 1) ? 0.0f : ck * Nkp;
            const float fkm = (m == 0) ? 0.0f : ck * Nkm;
            const size_t pb = c - sp * n;
            const float pm3 = Nk[pb + sp * wrap(n - 3, nphi)], pm2 = Nk[pb + sp * wrap(n - 2, nphi)];
            const float pm1 = Nk[pb + sp * wrap(n - 1, nphi)], pp1 = Nk[pb + sp * wrap(n + 1, nphi)];
            const float pp2 = Nk[pb + sp * wrap(n + 2, nphi)], pp3 = Nk[pb + sp * wrap(n + 3, nphi)];
            const float fp = cp * (face(pm2, pm1, x0, pp1, pp2, pp3, cp) - face(pm3, pm2, pm1, x0, pp1, pp2, cp));
            L = L - fmaf(fkp - fkm, p.dk_inv, fp * p.dp_inv);
        }
    }
    if (p.src_add)
        L += p.src_add[c];
    if (STAGE == STAGE_TENDENCY) {
        out[c] = L;
    } else if (STAGE == STAGE_1) {
        out[c] = clip0(fmaf(p.dt, L, x0));
    } else {
        const float a = STAGE == STAGE_2 ? 0.75f : (float)(1.0 / 3.0);
        const float b = STAGE == STAGE_2 ? 0.25f : (float)(2.0 / 3.0);
        const float st = clip0(fmaf(p.dt, L, x0));
        out[c] = clip0(fmaf(a, N0[c], b * st));
    }
}

template <int TRANSPORT>
hipError_t launch_t(const StageParams<float> &p, int stage, const float *Nk, const float *N0, float *out,
                    hipStream_t s)
{
    dim3 block(64, 4, 1);
    dim3 grid((p.nx + 63) / 64, (p.ny + 3) / 4, p.nk * p.nphi);
    switch (stage) {
    case STAGE_TENDENCY:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_TENDENCY>), grid, block, 0, s, Nk, N0, out, p);
        break;
    case STAGE_1:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_1>), grid, block, 0, s, Nk, N0, out, p);
        break;
    case STAGE_2:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_2>), grid, block, 0, s, Nk, N0, out, p);
        break;
    default:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_3>), grid, block, 0, s, Nk, N0, out, p);
        break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_stage_fast(const StageParams<float> &p, int transport, int stage, const float *Nk, const float *N0,
                             float *out, hipStream_t s)
{
    if (transport == RPL_TRANSPORT_REFRACTION)
        return launch_t<RPL_TRANSPORT_REFRACTION>(p, stage, Nk, N0, out, s);
    if (transport == RPL_TRANSPORT_INTRINSIC)
        return launch_t<RPL_TRANSPORT_INTRINSIC>(p, stage, Nk, N0, out, s);
    return launch_t<RPL_TRANSPORT_NONE>(p, stage, Nk, N0, out, s);
}

}  // namespace rpl
