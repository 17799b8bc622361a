// FP32 fused tendency + SSP-RK3 stage kernel, restructured for gfx950.
//
// Same discretisation as Coupling/wave_current_refraction.jl:41-151 and
// Models/intrinsic_transport.jl:14-68 (WENO5-JS with the reference's FP32
// epsilon 1e4*eps(Float32), cell-centre velocity on both faces, first-order
// kappa fallback, zero outer kappa flux), with the arithmetic reorganised:
//   * only the upwind reconstruction of each face is evaluated, and the
//     upwind order is chosen per lane by selecting the stencil values, so the
//     wavefront never diverges on the velocity sign;
//   * the two faces of a cell share their second differences and smoothness
//     indicators (4 triplets instead of 6);
//   * each reconstruction is v1 + weighted third differences with one
//     reciprocal instead of four IEEE divides;
//   * the current gradients are centred differences of U formed in-kernel
//     (wave_current_refraction.jl:7-19) instead of four extra arrays;
//   * the refraction and intrinsic stages give each workgroup one (x, y)
//     tile of one kappa bin and walk its phi bins (phi_tile): what depends on
//     (i, j, kappa) only is formed once, the phi stencil is a register
//     window, and a phi face is carried from one bin to the next.
// Results differ from the reference-order kernel by rounding only.
#include "rpl_internal.h"

// Every fused multiply-add here is written out (fmaf); nothing else is
// contracted, so the per-model, phi-loop and ensemble instantiations of the
// same cell arithmetic give the same bits.
#pragma clang fp contract(off)

namespace rpl {
namespace {

// Element-type helpers.  The FP32 forms are the instructions the kernels
// were built on; FP64 takes the hardware reciprocal with one Newton step
// (e = 1 - x r, r + r e), which brings it to the last bit or two.
__device__ __forceinline__ float fm(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fm(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fab(float x) { return fabsf(x); }
__device__ __forceinline__ double fab(double x) { return fabs(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// WENO5 constants of the element type: 4 * epsilon with the reference's
// epsilon 1e4 eps(FT) (wave_current_refraction.jl:43; the smoothness
// indicators are carried times 4) and the largest denominator the product
// form may reach before the ratio form takes over.
template <typename T>
struct Weno;
template <>
struct Weno<float> {
    static constexpr float eps4 = 4.0f * (10000.0f * 1.1920928955078125e-7f);
    static constexpr float big = 3.0e38f;
};
template <>
struct Weno<double> {
    static constexpr double eps4 = 4.0 * (10000.0 * 2.220446049250313080847263336181640625e-16);
    static constexpr double big = 1.0e307;
};

// The two upwind-left WENO5 face values of one cell along one direction.
//
// The six values w0..w5 are the cell's neighbours ordered upwind-first (for a
// non-negative velocity w_k = f[c-3+k], otherwise w_k = f[c+3-k]), so that
// A = WENO5(w0..w4) is the upwind value at the cell's upstream face and
// B = WENO5(w1..w5) the value at its downstream face, and the reference's
// vel * (N_plus - N_minus) equals |vel| * (B - A) for either sign
// (wave_current_refraction.jl:57-64, 111-113).  The two reconstructions share
// their second differences and smoothness indicators.
//
// Each reconstruction is v1 + (A0 (v0 - v1) + A2 (v2 - v1)) / (A0 + A1 + A2),
// A_k = d_k / (eps + beta_k)^2 multiplied through by the three squared
// denominators, so one reciprocal replaces the reference's four divides.  The
// smoothness indicators are scaled by 4 together with epsilon, which leaves
// the weights unchanged, and every value is carried times 6 (the 1/6 is
// applied once to the flux).  If the product form overflows (FP32:
// differences of the action above ~1e4) the lane falls back to the ratio form.
//
// weno_face6 returns 6 * (value - v1) for the face with smoothness q0, q1, q2
// and third differences U0 = D_l - D_c, U1 = D_c - D_r.
template <typename T>
__device__ __forceinline__ T weno_face6(T q0, T q1, T q2, T U0, T U1)
{
    const T s0 = q0 * q0, s1 = q1 * q1, s2 = q2 * q2;
    // 6 (v0 - v1) = 2 U0, 6 (v2 - v1) = U1; weights 1 : 6 : 3 over s0, s1, s2.
    T num = s1 * fm(s2 * T(2), U0, T(3) * s0 * U1);
    T den = fm(s2, fm(T(6), s0, s1), T(3) * s0 * s1);
    if (!(den <= Weno<T>::big)) {
        const T r0 = q1 * frcp(q0), r2 = q1 * frcp(q2);
        const T t0 = r0 * r0, t2 = r2 * r2;
        num = fm(T(2) * t0, U0, T(3) * t2 * U1);
        den = fm(T(6), t0, fm(T(18), t2, T(36))) * (T(1) / T(6));
    }
    return num * frcp(den);
}

// 6 * (B - A) for the window w0..w5, and optionally A and B themselves.
template <bool VALUES, typename T>
__device__ __forceinline__ T weno_pair6(T w0, T w1, T w2, T w3, T w4, T w5, T *A, T *B)
{
    const T eps4 = Weno<T>::eps4;
    const T c13_3 = T(13) / T(3);
    const T g0 = w1 - w0, g1 = w2 - w1, g2 = w3 - w2, g3 = w4 - w3, g4 = w5 - w4;
    const T D1 = g1 - g0, D2 = g2 - g1, D3 = g3 - g2, D4 = g4 - g3;
    const T t1 = fm(c13_3 * D1, D1, eps4), t2 = fm(c13_3 * D2, D2, eps4);
    const T t3 = fm(c13_3 * D3, D3, eps4), t4 = fm(c13_3 * D4, D4, eps4);
    // beta0 of a face uses the right-edge form on its first triplet, beta1 the
    // centred form, beta2 the left-edge form on its last triplet.
    const T eA1 = fm(T(2), g1, D1), eA2 = fm(T(2), g2, D2);
    const T eC2 = g1 + g2, eC3 = g2 + g3;
    const T eB3 = fm(T(-2), g2, D3), eB4 = fm(T(-2), g3, D4);
    const T qA1 = fm(eA1, eA1, t1), qA2 = fm(eA2, eA2, t2);
    const T qC2 = fm(eC2, eC2, t2), qC3 = fm(eC3, eC3, t3);
    const T qB3 = fm(eB3, eB3, t3), qB4 = fm(eB4, eB4, t4);
    const T U1 = D1 - D2, U2 = D2 - D3, U3 = D3 - D4;
    const T fa = weno_face6(qA1, qC2, qB3, U1, U2);
    const T fb = weno_face6(qA2, qC3, qB4, U2, U3);
    // 6 v1 of each face: 5 w_c + 2 w_r - w_l.
    const T va = fm(T(5), w2, fm(T(2), w3, -w1));
    const T vb = fm(T(5), w3, fm(T(2), w4, -w2));
    if (VALUES) {
        *A = (va + fa) * (T(1) / T(6));
        *B = (vb + fb) * (T(1) / T(6));
    }
    return (vb - va) + (fb - fa);
}

template <typename T>
__device__ __forceinline__ void weno_pair(T w0, T w1, T w2, T w3, T w4, T w5, T &A, T &B)
{
    (void)weno_pair6<true>(w0, w1, w2, w3, w4, w5, &A, &B);
}

// One face of weno_pair6 on its own: for the window u0..u4 (upwind first) the
// face value is (v + f) / 6.  The operations are those weno_pair6 performs for
// its A face on w0..w4 (and for its B face on w1..w5), in the same order, so a
// face carried from one cell to its neighbour has the same bits as the pair
// evaluated at the neighbour.
template <typename T>
struct Face6 {
    T v, f;
};

template <typename T>
__device__ __forceinline__ Face6<T> weno_face_of5(T u0, T u1, T u2, T u3, T u4)
{
    const T eps4 = Weno<T>::eps4;
    const T c13_3 = T(13) / T(3);
    const T h0 = u1 - u0, h1 = u2 - u1, h2 = u3 - u2, h3 = u4 - u3;
    const T E1 = h1 - h0, E2 = h2 - h1, E3 = h3 - h2;
    const T s1 = fm(c13_3 * E1, E1, eps4), s2 = fm(c13_3 * E2, E2, eps4);
    const T s3 = fm(c13_3 * E3, E3, eps4);
    const T eA = fm(T(2), h1, E1), eC = h1 + h2, eB = fm(T(-2), h2, E3);
    const T qA = fm(eA, eA, s1), qC = fm(eC, eC, s2), qB = fm(eB, eB, s3);
    Face6<T> r;
    r.f = weno_face6(qA, qC, qB, E1 - E2, E2 - E3);
    r.v = fm(T(5), u2, fm(T(2), u3, -u1));
    return r;
}

// weno_pair6's two faces as (v, f) parts; (b.v - a.v) + (b.f - a.f) is its
// return value.
template <typename T>
__device__ __forceinline__ void weno_pair_parts(T w0, T w1, T w2, T w3, T w4, T w5, Face6<T> &a, Face6<T> &b)
{
    const T eps4 = Weno<T>::eps4;
    const T c13_3 = T(13) / T(3);
    const T g0 = w1 - w0, g1 = w2 - w1, g2 = w3 - w2, g3 = w4 - w3, g4 = w5 - w4;
    const T D1 = g1 - g0, D2 = g2 - g1, D3 = g3 - g2, D4 = g4 - g3;
    const T t1 = fm(c13_3 * D1, D1, eps4), t2 = fm(c13_3 * D2, D2, eps4);
    const T t3 = fm(c13_3 * D3, D3, eps4), t4 = fm(c13_3 * D4, D4, eps4);
    const T eA1 = fm(T(2), g1, D1), eA2 = fm(T(2), g2, D2);
    const T eC2 = g1 + g2, eC3 = g2 + g3;
    const T eB3 = fm(T(-2), g2, D3), eB4 = fm(T(-2), g3, D4);
    const T qA1 = fm(eA1, eA1, t1), qA2 = fm(eA2, eA2, t2);
    const T qC2 = fm(eC2, eC2, t2), qC3 = fm(eC3, eC3, t3);
    const T qB3 = fm(eB3, eB3, t3), qB4 = fm(eB4, eB4, t4);
    const T U1 = D1 - D2, U2 = D2 - D3, U3 = D3 - D4;
    a.f = weno_face6(qA1, qC2, qB3, U1, U2);
    b.f = weno_face6(qA2, qC3, qB4, U2, U3);
    a.v = fm(T(5), w2, fm(T(2), w3, -w1));
    b.v = fm(T(5), w3, fm(T(2), w4, -w2));
}

// |vel| * (N_plus - N_minus) for a direction with a full stencil on both faces.
// The 1/6 folds into the product with |vel|.
template <typename T>
__device__ __forceinline__ T upwind_difference(T fm3, T fm2, T fm1, T f0, T fp1, T fp2, T fp3, T vel)
{
    const bool pos = vel >= T(0);
    const T d6 = weno_pair6<false>(pos ? fm3 : fp3, pos ? fm2 : fp2, pos ? fm1 : fp1, f0, pos ? fp1 : fm1,
                                   pos ? fp2 : fm2, (T *)nullptr, (T *)nullptr);
    return fab(vel) * (T(1) / T(6)) * d6;
}

// The same for a velocity that is uniform across the wavefront (intrinsic
// transport: the group velocity of a bin): the upwind side is a branch taken
// by the whole wave instead of a per-lane selection of the six stencil
// values.  Same arithmetic, so the same bits.
template <typename T>
__device__ __forceinline__ T upwind_difference_uniform(T fm3, T fm2, T fm1, T f0, T fp1, T fp2, T fp3, T vel)
{
    T d6;
    if (vel >= T(0))
        d6 = weno_pair6<false>(fm3, fm2, fm1, f0, fp1, fp2, (T *)nullptr, (T *)nullptr);
    else
        d6 = weno_pair6<false>(fp3, fp2, fp1, f0, fm1, fm2, (T *)nullptr, (T *)nullptr);
    return fab(vel) * (T(1) / T(6)) * d6;
}

// Load at a 32-bit byte offset from a block-uniform base, so the address is a
// scalar base plus a 32-bit lane offset (global_load ... saddr form) and costs
// no 64-bit vector arithmetic.  Offsets stay below 4 GiB: a slab is nx*ny
// elements and the host rejects nx*ny >= 2^30 (2^29 for FP64 slabs).
// AS1: the base came from memory (ensemble descriptors) and is marked global
// so that the load stays a global (not flat) load.
template <bool AS1 = false, typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned byte_off)
{
    if (AS1) {
        typedef const __attribute__((address_space(1))) char gchar;
        typedef const __attribute__((address_space(1))) T gT;
        return *(gT *)((gchar *)base + byte_off);
    }
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

template <bool AS1 = false, typename T>
__device__ __forceinline__ void sto(T *base, unsigned byte_off, T v)
{
    if (AS1) {
        typedef __attribute__((address_space(1))) char gchar;
        typedef __attribute__((address_space(1))) T gT;
        *(gT *)((gchar *)base + byte_off) = v;
        return;
    }
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
template <typename T>
__device__ __forceinline__ T clip0(T x) { return x < T(0) ? T(0) : x; }

// Kappa flux difference of one cell (wave_current_refraction.jl:125-137):
// WENO5 where the six-point stencil fits, first-order upwind faces where it
// does not, zero flux through the outer faces.  m is uniform across the
// workgroup, so the edge path does not diverge.  fk[3] is unused (x0).
template <typename T>
__device__ __forceinline__ T kappa_difference(const T (&fk)[7], T x0, T ck, int m, int nk)
{
    if (m >= 3 && m + 3 <= nk - 1)
        return upwind_difference(fk[0], fk[1], fk[2], x0, fk[4], fk[5], fk[6], ck);
    const T km3 = fk[0], km2 = fk[1], km1 = fk[2], kp1 = fk[4], kp2 = fk[5], kp3 = fk[6];
    const bool pos = ck >= T(0);
    const T w2 = pos ? km1 : kp1, w4 = pos ? kp1 : km1;
    T A = w2, B = x0;  // first-order upstream / downstream faces
    const bool has_p = (m - 2 >= 0) && (m + 3 <= nk - 1);
    const bool has_m = (m - 3 >= 0) && (m + 2 <= nk - 1);
    if (has_p || has_m) {
        T Aw, Bw;
        weno_pair(pos ? km3 : kp3, pos ? km2 : kp2, w2, x0, w4, pos ? kp2 : km2, Aw, Bw);
        // For ck >= 0 the downstream face is the plus face, else the minus face.
        if (pos ? has_m : has_p)
            A = Aw;
        if (pos ? has_p : has_m)
            B = Bw;
    }
    const T plus = pos ? B : A, minus = pos ? A : B;
    const T fkp = (m == nk - 1) ? T(0) : ck * plus;
    const T fkm = (m == 0) ? T(0) : ck * minus;
    return fkp - fkm;
}

// Work order.  One 64 x STAGE_ROWS (i, j) tile of one (kappa, phi) bin per workgroup.
// Workgroups are dealt round-robin to the eight XCDs, so the linear block id
// is remapped to give every XCD one contiguous range of the logical order
// phi (fastest), kappa, tile row, tile column.  The kappa and phi neighbours
// a tile reads are then processed on the same XCD at about the same time and
// are served by that XCD's L2, and a tile's y-halo rows were just read by the
// tile before it.
struct Tile {
    int ib, jb, m, n;
};

// Linear work index of block b: XCD b % 8 takes the b / 8-th entry of its
// contiguous range of the total.
__device__ __forceinline__ int xcd_linear(int b, int total)
{
    const int xcd = b & 7, r = b >> 3;
    const int per = total >> 3, rem = total & 7;
    return xcd < rem ? xcd * (per + 1) + r : rem * (per + 1) + (xcd - rem) * per + r;
}

__device__ __forceinline__ Tile tile_of(int b, int total, int nbx, int nby, int nk, int nphi)
{
    const int L = xcd_linear(b, total);
    Tile t;
    t.n = L % nphi;
    int q = L / nphi;
    t.m = q % nk;
    q /= nk;
    t.jb = q % nby;
    t.ib = q / nby;
    (void)nbx;
    return t;
}

// One 64 x TY (i, j) tile of bin (m, n): tendency, sources and the SSP-RK3
// stage update for each cell.  Shared by the per-model kernel below and the
// ensemble work-list kernel; p may live in kernel arguments or in global
// memory (block-uniform either way).
template <int TRANSPORT, int STAGE, bool SRC, int TY, bool AS1>
__device__ __forceinline__ void stage_tile(const float *__restrict__ Nk_, const float *__restrict__ N0_, float *out_,
                                           const StageParams<float> &p, int ib, int jb, int m, int n)
{
    const float *__restrict__ Nk = Nk_;
    const float *__restrict__ N0 = N0_;
    float *out = out_;
    const int nx = p.nx, ny = p.ny, nk = p.nk, nphi = p.nphi;
    const int i0 = ib * 64, j0 = jb * TY;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i >= nx || j >= ny)
        return;
    const size_t plane = (size_t)nx * ny;
    // Block-uniform slab bases (scalar registers); per-lane byte offsets are 32-bit.
    const size_t bin = (size_t)m + (size_t)nk * n;
    const float *slab = Nk + plane * bin;
    const unsigned nxu = (unsigned)nx;
    const unsigned off = (unsigned)i + nxu * (unsigned)j;
    const unsigned bo = off << 2;
    const float x0 = ldo<AS1>(slab, bo);
    const bool x_inner = i0 >= 3 && i0 + 64 + 3 <= nx;  // block-uniform
    const bool y_inner = j0 >= 3 && j0 + TY + 3 <= ny;
    float L = 0.0f;
    if (TRANSPORT != RPL_TRANSPORT_NONE) {
        const unsigned rowb = (nxu * (unsigned)j) << 2, colb = (unsigned)i << 2;
        const unsigned sx = 4u, sy = nxu << 2;
        float ux = gld<AS1>(p.cgx, m + nk * n), uy = gld<AS1>(p.cgy, m + nk * n);
        float ck = 0.0f, cp = 0.0f;
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            const float *Ux = p.Ux + plane * m, *Uy = p.Uy + plane * m;
            const unsigned xp = x_inner ? bo + sx : rowb + ((unsigned)wrap(i + 1, nx) << 2);
            const unsigned xm = x_inner ? bo - sx : rowb + ((unsigned)wrap(i - 1, nx) << 2);
            const unsigned yp = y_inner ? bo + sy : colb + sy * (unsigned)wrap(j + 1, ny);
            const unsigned ym = y_inner ? bo - sy : colb + sy * (unsigned)wrap(j - 1, ny);
            const float hx = 0.5f * p.dx_inv, hy = 0.5f * p.dy_inv;
            const float Uxx = (ldo<AS1>(Ux, xp) - ldo<AS1>(Ux, xm)) * hx;
            const float Uxy = (ldo<AS1>(Ux, yp) - ldo<AS1>(Ux, ym)) * hy;
            const float Uyx = (ldo<AS1>(Uy, xp) - ldo<AS1>(Uy, xm)) * hx;
            const float Uyy = (ldo<AS1>(Uy, yp) - ldo<AS1>(Uy, ym)) * hy;
            const float kap = gld<AS1>(p.kappa, m), cs = gld<AS1>(p.cosp, n), sn = gld<AS1>(p.sinp, n);
            const float cc = cs * cs, ss = sn * sn, csn = cs * sn;
            ck = -kap * (cc * Uxx + csn * (Uyx + Uxy) + ss * Uyy);
            cp = csn * (Uxx - Uyy) + ss * Uyx - cc * Uxy;
            ux += ldo<AS1>(Ux, bo);
            uy += ldo<AS1>(Uy, bo);
        }
        float fx[7], fy[7];
        if (x_inner) {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                fx[d] = ldo<AS1>(slab, bo + (unsigned)(4 * (d - 3)));
        } else {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                fx[d] = ldo<AS1>(slab, rowb + ((unsigned)wrap(i + d - 3, nx) << 2));
        }
        if (y_inner) {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                fy[d] = ldo<AS1>(slab, bo + sy * (unsigned)(d - 3));
        } else {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                fy[d] = ldo<AS1>(slab, colb + sy * (unsigned)wrap(j + d - 3, ny));
        }
        const float dx = upwind_difference(fx[0], fx[1], fx[2], x0, fx[4], fx[5], fx[6], ux);
        const float dy = upwind_difference(fy[0], fy[1], fy[2], x0, fy[4], fy[5], fy[6], uy);
        L = -fmaf(dx, p.dx_inv, dy * p.dy_inv);
        if (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            float fp[7], fk[7];
#pragma unroll
            for (int d = 0; d < 7; ++d) {
                if (d == 3)
                    continue;
                const size_t pbin = (size_t)m + (size_t)nk * wrap(n + d - 3, nphi);
                const size_t kbin = (size_t)clampi(m + d - 3, nk) + (size_t)nk * n;
                fp[d] = ldo<AS1>(Nk + plane * pbin, bo);
                fk[d] = ldo<AS1>(Nk + plane * kbin, bo);
            }
            const float dp = upwind_difference(fp[0], fp[1], fp[2], x0, fp[4], fp[5], fp[6], cp);
            const float dk = kappa_difference(fk, x0, ck, m, nk);
            L = L - fmaf(dk, p.dk_inv, dp * p.dp_inv);
        }
    }
    if (SRC)  // tendencies.jl:88-94, G += source_tendency
        L += source_term<float, AS1>(p.src, off, m + nk * n, x0);
    if (p.src_add)
        L += ldo<AS1>(p.src_add + plane * bin, bo);
    float *dst = out + plane * bin;
    float r;
    if (STAGE == STAGE_TENDENCY) {
        r = L;
    } else if (STAGE == STAGE_1) {
        r = clip0(fmaf(p.dt, L, x0));
    } else {
        const float a = STAGE == STAGE_2 ? 0.75f : (float)(1.0 / 3.0);
        const float b = STAGE == STAGE_2 ? 0.25f : (float)(2.0 / 3.0);
        const float st = clip0(fmaf(p.dt, L, x0));
        r = clip0(fmaf(a, ldo<AS1>(N0 + plane * bin, bo), b * st));
    }
    sto<AS1>(dst, bo, r);
}

// Rows per workgroup tile of the per-model kernel (64 x STAGE_ROWS cells of
// one bin): taller tiles re-read fewer y-halo rows (a tile reads STAGE_ROWS + 6
// rows of its bin).
#ifndef RPL_STAGE_ROWS
#define RPL_STAGE_ROWS 4
#endif
constexpr int STAGE_ROWS = RPL_STAGE_ROWS;

// Work order of the transports without spectral coupling (intrinsic, none):
// every bin is an independent 2-D problem, so each XCD's contiguous range
// walks tile columns fastest, then tile rows, then bins.  The tiles in flight
// on one XCD then cover whole rows of one bin and share their x and y halos
// in that XCD's L2 (the refraction order, bins fastest, would fetch every
// tile's halo again).
__device__ __forceinline__ Tile tile_of_spatial(int b, int total, int nbx, int nby, int nk)
{
    const int L = xcd_linear(b, total);
    Tile t;
    t.ib = L % nbx;
    int q = L / nbx;
    t.jb = q % nby;
    q /= nby;
    t.m = q % nk;
    t.n = q / nk;
    return t;
}

// Per-bin stage of the intrinsic and no-transport models.
template <int TRANSPORT, int STAGE, bool SRC>
__global__ __launch_bounds__(64 * STAGE_ROWS) void stage_fast_kernel(const float *__restrict__ Nk,
                                                                     const float *__restrict__ N0, float *out,
                                                                     StageParams<float> p, int nbx, int nby)
{
    const Tile t = tile_of_spatial(blockIdx.x, nbx * nby * p.nk * p.nphi, nbx, nby, p.nk);
    stage_tile<TRANSPORT, STAGE, SRC, STAGE_ROWS, false>(Nk, N0, out, p, t.ib, t.jb, t.m, t.n);
}

// Stage with the phi bins looped inside the workgroup: one workgroup per
// 64 x TY (i, j) tile of one kappa bin, walking n = 0..nphi-1.
// Refraction: what depends on (i, j, m) only -- the Doppler velocity and its
// four centred gradients -- is loaded and formed once instead of once per
// bin, and the phi stencil is a register window that takes one new load per
// bin.  Intrinsic transport: x/y WENO only, with the column inputs of the
// source terms loaded once for the 36 bins.  The arithmetic of every cell is
// that of stage_tile, so the results are the same bit for bit.  Shared by the
// per-model kernel and the ensemble work-list kernel (AS1: p and the slab
// pointers came from memory).
template <typename T, int TRANSPORT, int STAGE, int SK, int TY, bool AS1, bool CGF = false>
__device__ __forceinline__ void phi_tile(const T *__restrict__ Nk, const T *__restrict__ N0, T *out,
                                         const StageParams<T> &p, int ib, int jb, int m)
{
    constexpr bool REF = TRANSPORT == RPL_TRANSPORT_REFRACTION;
    static_assert(REF || TRANSPORT == RPL_TRANSPORT_INTRINSIC, "phi_tile: refraction or intrinsic transport");
    const int nx = p.nx, ny = p.ny, nk = p.nk, nphi = p.nphi;
    const int i0 = ib * 64, j0 = jb * TY;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i >= nx || j >= ny)
        return;
    const size_t plane = (size_t)nx * ny;
    const unsigned nxu = (unsigned)nx;
    const unsigned off = (unsigned)i + nxu * (unsigned)j;
    constexpr unsigned ES = (unsigned)sizeof(T), SH = ES == 8 ? 3u : 2u;  // element bytes, log2
    const unsigned bo = off << SH;
    const bool x_inner = i0 >= 3 && i0 + 64 + 3 <= nx;  // block-uniform
    const bool y_inner = j0 >= 3 && j0 + TY + 3 <= ny;
    const unsigned rowb = (nxu * (unsigned)j) << SH, colb = (unsigned)i << SH;
    const unsigned sx = ES, sy = nxu << SH;
    T Uxx = T(0), Uxy = T(0), Uyx = T(0), Uyy = T(0), Uxc = T(0), Uyc = T(0), kap = T(0);
    if constexpr (REF) {
        // Doppler velocity and gradients at (i, j, m) (wave_current_refraction.jl:7-19).
        const T *Ux = p.Ux + plane * m, *Uy = p.Uy + plane * m;
        const unsigned xp = x_inner ? bo + sx : rowb + ((unsigned)wrap(i + 1, nx) << SH);
        const unsigned xm = x_inner ? bo - sx : rowb + ((unsigned)wrap(i - 1, nx) << SH);
        const unsigned yp = y_inner ? bo + sy : colb + sy * (unsigned)wrap(j + 1, ny);
        const unsigned ym = y_inner ? bo - sy : colb + sy * (unsigned)wrap(j - 1, ny);
        const T hx = T(0.5) * p.dx_inv, hy = T(0.5) * p.dy_inv;
        Uxx = (ldo<AS1>(Ux, xp) - ldo<AS1>(Ux, xm)) * hx;
        Uxy = (ldo<AS1>(Ux, yp) - ldo<AS1>(Ux, ym)) * hy;
        Uyx = (ldo<AS1>(Uy, xp) - ldo<AS1>(Uy, xm)) * hx;
        Uyy = (ldo<AS1>(Uy, yp) - ldo<AS1>(Uy, ym)) * hy;
        Uxc = ldo<AS1>(Ux, bo);
        Uyc = ldo<AS1>(Uy, bo);
        kap = gld<AS1>(p.kappa, m);
    }
    constexpr bool SRC = SK != SOURCE_NONE;
    // CGF (a depth per column): the cell's group speed at its column's depth,
    // the same for every phi bin of the walk; the velocity is that speed
    // times the bin's direction (cos, sin)
    T cgm = T(0);
    if constexpr (CGF)
        cgm = ldo<AS1>(p.cgm + plane * (size_t)m, bo);
    // the column's source inputs, loaded once for every bin of the tile
    SourceColumn<T> scol{};
    JpColumn jcol{};
    static_assert((SK != SOURCE_JP_LS && SK != SOURCE_NET_LS) || std::is_same<T, float>::value,
                  "the specialised source paths are FP32");
    if constexpr (SK == SOURCE_JP_LS)
        jcol = jp_column<AS1>(p.src, bo);
    else if constexpr (SK == SOURCE_NET_LS)
        jcol = jp_column<AS1, true>(p.src, bo);
    else if constexpr (SK == SOURCE_GENERIC)
        scol = source_column<T, AS1>(p.src, off);
    // What bin n reads: its x and y stencils, for refraction its kappa
    // stencil and the phi window's next value f(n + 4), otherwise the cell
    // itself, N0 and the added source at the cell, and the bin's source
    // tables.  The latency of these
    // loads is hidden by the other waves of the SIMD, so occupancy (registers
    // per lane) decides the speed: issuing bin n + 1's loads ahead of bin n's
    // arithmetic costs 20 more registers and runs slower.
    struct BinLoads {
        T fx[7], fy[7], fk[7], wnext, x0, n0, add;
        float nrate;  // SOURCE_NET_LS: the network's rate at the cell
        SourceBin<T> sb;
        T bux, buy, btc, bts, bcs, bsn;  // packed table: group velocity, stress weights, cos and sin of phi
    };
    // the FP32 per-model specialised source paths read one packed entry per bin
    constexpr bool PACKED = std::is_same<T, float>::value && !AS1 && !CGF &&
                            (SK == SOURCE_JP_LS || SK == SOURCE_NET_LS || SK == SOURCE_NONE);
    typedef float f32x16 __attribute__((ext_vector_type(16)));
    const unsigned cell_off = bo;
    auto load_bin = [&](int n, BinLoads &b, auto interior) {
        // Every load below is a scalar slab base plus this 32-bit lane offset.
        // The empty asm hides that the offset is the same in every bin;
        // otherwise the compiler hoists one 64-bit lane pointer per stencil
        // point out of the phi loop (26 more registers, one wave less per
        // SIMD).
        unsigned bo = cell_off;
        asm volatile("" : "+v"(bo));
        const T *slab = Nk + plane * ((size_t)m + (size_t)nk * n);
        if constexpr (decltype(interior)::value) {
            // tiles whose x and y stencils stay inside the slab: the stencil
            // points are the slab base moved by a uniform amount (scalar
            // arithmetic or the load's immediate offset), so the loop spends
            // no vector instructions on addresses
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fx[d] = ldo<AS1>(slab + (d - 3), bo);
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fy[d] = ldo<AS1>(slab + (ptrdiff_t)(d - 3) * nx, bo);
        } else {
        if (x_inner) {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fx[d] = ldo<AS1>(slab, bo + (unsigned)((int)ES * (d - 3)));
        } else {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fx[d] = ldo<AS1>(slab, rowb + ((unsigned)wrap(i + d - 3, nx) << SH));
        }
        if (y_inner) {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fy[d] = ldo<AS1>(slab, bo + sy * (unsigned)(d - 3));
        } else {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fy[d] = ldo<AS1>(slab, colb + sy * (unsigned)wrap(j + d - 3, ny));
        }
        }
        if constexpr (REF) {
#pragma unroll
            for (int d = 0; d < 7; ++d)
                if (d != 3)
                    b.fk[d] = ldo<AS1>(Nk + plane * ((size_t)clampi(m + d - 3, nk) + (size_t)nk * n), bo);
            // f(n + 4): n + 4 < 3 nphi for nphi >= 3 (it wraps twice round a
            // 3-bin circle), so two conditional subtractions suffice
            int nw = n + 4;
            nw = nw >= nphi ? nw - nphi : nw;
            nw = nw >= nphi ? nw - nphi : nw;
            b.wnext = ldo<AS1>(Nk + plane * ((size_t)m + (size_t)nk * nw), bo);
        } else {
            b.x0 = ldo<AS1>(slab, bo);
        }
        const size_t bin = (size_t)m + (size_t)nk * n;
        b.n0 = (STAGE == STAGE_2 || STAGE == STAGE_3) ? ldo<AS1>(N0 + plane * bin, bo) : T(0);
        b.add = p.src_add ? ldo<AS1>(p.src_add + plane * bin, bo) : T(0);
        if constexpr (PACKED) {
            const f32x16 e = sld(reinterpret_cast<const f32x16 *>(p.bt), bin);
            b.bux = e[BT_CGX];
            b.buy = e[BT_CGY];
            b.sb.k = e[BT_K];
            b.sb.cth = e[BT_CTH];
            b.sb.sth = e[BT_STH];
            b.sb.logk = e[BT_LOGK];
            b.sb.invc = e[BT_INVC];
            b.sb.sigma = e[BT_SIGMA];
            b.sb.sk3 = e[BT_SK3];
            b.sb.sq = e[BT_SQ];
            b.sb.out_scale = e[BT_OUTSCALE];
            b.btc = e[BT_TAUC];
            b.bts = e[BT_TAUS];
            b.bcs = e[BT_COSP];
            b.bsn = e[BT_SINP];
        } else if (SRC) {
            b.sb = source_bin<T, AS1>(p.src, (int)bin);
        }
        if constexpr (SK == SOURCE_NET_LS) {
            // the FP16 rate plane at this column: byte offset bo / 2
            typedef const __attribute__((address_space(1))) char gchar;
            b.nrate = (float)*(const __attribute__((address_space(1))) _Float16 *)(
                (gchar *)(p.src.rate + p.src.rate_ld * bin) + (bo >> 1));
        }
    };
    // phi window f(n - 3 .. n + 3) at this cell, periodic in phi (refraction)
    T w[7] = {};
    if constexpr (REF) {
#pragma unroll
        for (int d = 0; d < 7; ++d)
            w[d] = ldo<AS1>(Nk + plane * ((size_t)m + (size_t)nk * wrap(d - 3, nphi)), bo);
    }
    Face6<T> carry = {T(0), T(0)};  // phi face n - 1/2 (as (v, f) of weno_face_of5)
    bool cpos = true;            // its upwind side
    T twx = T(0), twy = T(0);  // wave-supported stress sums of the output (FP32 stress cap)
    // The phi walk, in two copies: tiles away from the x and y edges take
    // their stencils at fixed offsets, the others through the periodic wrap.
    auto walk = [&](auto interior) {
    for (int n = 0; n < nphi; ++n) {
        BinLoads cur;
        load_bin(n, cur, interior);
        const size_t bin = (size_t)m + (size_t)nk * n;
        const T x0 = REF ? w[3] : cur.x0;
        T ux, uy;
        if constexpr (CGF) {
            // the speed is nonnegative, so the upwind side is still the bin's
            ux = cgm * sld(p.cdx, (size_t)(m + nk * n));
            uy = cgm * sld(p.cdy, (size_t)(m + nk * n));
        } else if constexpr (PACKED) {
            ux = cur.bux;
            uy = cur.buy;
        } else if constexpr (REF) {
            if constexpr (!AS1) {
                ux = sld(p.cgx, (size_t)(m + nk * n));
                uy = sld(p.cgy, (size_t)(m + nk * n));
            } else {
                ux = gld<AS1>(p.cgx, m + nk * n);
                uy = gld<AS1>(p.cgy, m + nk * n);
            }
        } else {
            ux = sld(p.cgx, (size_t)(m + nk * n));  // the bin's group velocity: wave-uniform
            uy = sld(p.cgy, (size_t)(m + nk * n));
        }
        T ck = T(0), cp = T(0);
        if constexpr (REF) {
            T cs, sn;
            if constexpr (PACKED) {
                cs = cur.bcs;
                sn = cur.bsn;
            } else {
                if constexpr (!AS1) {
                    cs = sld(p.cosp, (size_t)n);
                    sn = sld(p.sinp, (size_t)n);
                } else {
                    cs = gld<AS1>(p.cosp, n);
                    sn = gld<AS1>(p.sinp, n);
                }
            }
            const T cc = cs * cs, ss = sn * sn, csn = cs * sn;
            ck = -kap * (cc * Uxx + csn * (Uyx + Uxy) + ss * Uyy);
            cp = csn * (Uxx - Uyy) + ss * Uyx - cc * Uxy;
            ux += Uxc;
            uy += Uyc;
        }
        const T *fx = cur.fx, *fy = cur.fy;
        T dx, dy;
        if constexpr (REF) {
            dx = upwind_difference(fx[0], fx[1], fx[2], x0, fx[4], fx[5], fx[6], ux);
            dy = upwind_difference(fy[0], fy[1], fy[2], x0, fy[4], fy[5], fy[6], uy);
        } else {
            dx = upwind_difference_uniform(fx[0], fx[1], fx[2], x0, fx[4], fx[5], fx[6], ux);
            dy = upwind_difference_uniform(fy[0], fy[1], fy[2], x0, fy[4], fy[5], fy[6], uy);
        }
        T L = -fm(dx, p.dx_inv, dy * p.dy_inv);
        if constexpr (REF) {
            // phi faces: the face n + 1/2 of this bin is the face n - 1/2 of
            // the next one.  When every lane keeps its upwind side from the
            // previous bin, only the new face is evaluated.
            const bool ppos = cp >= T(0);
            T d6;
            if (n > 0 && __all(ppos == cpos)) {
                const Face6<T> nf = weno_face_of5(ppos ? w[1] : w[6], ppos ? w[2] : w[5], ppos ? x0 : w[4],
                                               ppos ? w[4] : x0, ppos ? w[5] : w[2]);
                d6 = ppos ? (nf.v - carry.v) + (nf.f - carry.f) : (carry.v - nf.v) + (carry.f - nf.f);
                carry = nf;
            } else {
                Face6<T> a, b;
                weno_pair_parts(ppos ? w[0] : w[6], ppos ? w[1] : w[5], ppos ? w[2] : w[4], x0,
                                ppos ? w[4] : w[2], ppos ? w[5] : w[1], a, b);
                d6 = (b.v - a.v) + (b.f - a.f);
                carry = ppos ? b : a;
            }
            cpos = ppos;
            const T dp = fab(cp) * (T(1) / T(6)) * d6;
            const T dk = kappa_difference(cur.fk, x0, ck, m, nk);
            L = L - fm(dk, p.dk_inv, dp * p.dp_inv);
        }
        // tendencies.jl:88-94, G += source_tendency
        T raw = T(0);
        if constexpr (SK == SOURCE_JP_LS)
            L += source_jp_ls(p.src, jcol, cur.sb, x0, raw);
        else if constexpr (SK == SOURCE_NET_LS)
            L += source_net_ls(p.src, jcol, cur.sb, cur.nrate, x0, raw);
        else if constexpr (SK == SOURCE_GENERIC)
            L += source_term_cb<T, AS1>(p.src, scol, cur.sb, off, m + nk * n, x0, &raw);
        if (p.src_add)
            L += cur.add;
        T r;
        if (STAGE == STAGE_TENDENCY) {
            r = L;
        } else if (STAGE == STAGE_1) {
            r = clip0(fm(p.dt, L, x0));
        } else {
            const T a = STAGE == STAGE_2 ? T(0.75) : (T)(1.0 / 3.0);
            const T b = STAGE == STAGE_2 ? T(0.25) : (T)(2.0 / 3.0);
            const T st = clip0(fm(p.dt, L, x0));
            r = clip0(fm(a, cur.n0, b * st));
        }
        sto<AS1>(out + plane * bin, bo, r);
        if constexpr (SK != SOURCE_NONE) {
            // the output's wave-supported stress for the next stage's cap:
            // raw is this bin's rate at this column's wind, r the next input
            // (both source paths give raw the same bits)
            if (p.tau_x) {
                const T mf = raw * r;
                if constexpr (PACKED) {
                    twx = fm(mf, cur.btc, twx);
                    twy = fm(mf, cur.bts, twy);
                } else {
                    twx = fm(mf, sld(p.tau_c, bin), twx);
                    twy = fm(mf, sld(p.tau_s, bin), twy);
                }
            }
        }
        if constexpr (REF) {
#pragma unroll
            for (int d = 0; d < 6; ++d)
                w[d] = w[d + 1];
            w[6] = cur.wnext;
        }
    }
    };
    if constexpr (REF) {
        walk(std::false_type{});  // a second copy of the refraction walk spills (measured 24 % slower)
    } else {
        if (x_inner && y_inner)
            walk(std::true_type{});
        else
            walk(std::false_type{});
    }
    if constexpr (SK != SOURCE_NONE) {
        if (p.tau_x) {
            sto<AS1>(p.tau_x + plane * (size_t)m, bo, twx);
            sto<AS1>(p.tau_y + plane * (size_t)m, bo, twy);
        }
    }
}

// Per-model phi-loop stage.  Work order (XCD-aware, each XCD one contiguous
// range): refraction kappa fastest, then tile row and column, so the kappa
// neighbours a tile reads are processed together on one XCD; intrinsic
// transport has no spectral coupling and walks tile columns fastest, then
// tile rows, then kappa, so the tiles in flight share their x and y halos.
// Occupancy floors (waves per SIMD): intrinsic transport seven (72
// registers, one spilled value: 1.71 -> 1.61 ms per launch on the 512^2
// hurricane with sources); refraction without sources six (80 registers,
// one spilled value: 2.216 -> 2.184 ms per launch on the 512^2 sheared
// eddies); refraction with sources five (six spilled values; config-5
// members one after another: 137.8 -> 134.7 ms per step).  Higher floors
// spill more and run slower.
#ifndef RPL_F64_MIN_WAVES
#define RPL_F64_MIN_WAVES 2
#endif
template <typename T, int TRANSPORT, int STAGE, int SK, bool CGF>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? RPL_F64_MIN_WAVES
                                                  : (TRANSPORT == RPL_TRANSPORT_INTRINSIC ? 7 : (SK != SOURCE_NONE ? 5 : 6)),
                                   8)))
void stage_phi_kernel(const T *__restrict__ Nk, const T *__restrict__ N0, T *out, StageParams<T> p,
                      int nbx, int nby)
{
    Tile t;
    if constexpr (TRANSPORT == RPL_TRANSPORT_REFRACTION)
        t = tile_of(blockIdx.x, nbx * nby * p.nk, nbx, nby, p.nk, 1);
    else
        t = tile_of_spatial(blockIdx.x, nbx * nby * p.nk, nbx, nby, p.nk);
    phi_tile<T, TRANSPORT, STAGE, SK, 4, false, CGF>(Nk, N0, out, p, t.ib, t.jb, t.m);
}

// The per-bin values of the FP32 specialised source paths packed 16 floats
// per bin (BT_* fields; absent tables give 0).
__global__ __launch_bounds__(256) void pack_bins_kernel(float *bt, int nbins, int nk, const float *cgx,
                                                        const float *cgy, const float *cosp, const float *sinp,
                                                        SourceParams<float> s, const float *tau_c,
                                                        const float *tau_s)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nbins)
        return;
    auto at = [&](const float *a) { return a ? a[t] : 0.0f; };
    float *e = bt + (size_t)BT_FIELDS * t;
    e[BT_CGX] = at(cgx);
    e[BT_CGY] = at(cgy);
    e[BT_K] = s.on ? at(s.k_cell) : 0.0f;
    const bool wi = s.on && s.wind_input, ls = s.on && s.dissipation == 2;
    e[BT_CTH] = wi ? at(s.cth_cell) : 0.0f;
    e[BT_STH] = wi ? at(s.sth_cell) : 0.0f;
    e[BT_LOGK] = wi ? at(s.logk_cell) : 0.0f;
    e[BT_INVC] = wi ? at(s.invc_cell) : 0.0f;
    e[BT_SIGMA] = wi ? at(s.sigma_cell) : 0.0f;
    e[BT_SK3] = ls ? at(s.sk3_cell) : 0.0f;
    e[BT_SQ] = ls ? at(s.sq_cell) : 0.0f;
    e[BT_TAUC] = at(tau_c);
    e[BT_TAUS] = at(tau_s);
    e[BT_OUTSCALE] = s.on && s.surrogate ? at(s.out_scale) : 0.0f;
    e[BT_COSP] = cosp ? cosp[t / nk] : 0.0f;
    e[BT_SINP] = sinp ? sinp[t / nk] : 0.0f;
    e[BT_FIELDS - 1] = 0.0f;
}

template <typename T, int TRANSPORT, int SK, bool CGF = false>
hipError_t launch_phi(const StageParams<T> &p, int stage, const T *Nk, const T *N0, T *out,
                      hipStream_t s)
{
    const int nbx = (p.nx + 63) / 64, nby = (p.ny + 3) / 4;
    const long long total = (long long)nbx * nby * p.nk;
    if (total > 0x7fffffffLL)
        return hipErrorInvalidValue;
    dim3 block(64, 4, 1), grid((unsigned)total, 1, 1);
    switch (stage) {
    case STAGE_TENDENCY:
        hipLaunchKernelGGL((stage_phi_kernel<T, TRANSPORT, STAGE_TENDENCY, SK, CGF>), grid, block, 0, s, Nk, N0, out,
                           p, nbx, nby);
        break;
    case STAGE_1:
        hipLaunchKernelGGL((stage_phi_kernel<T, TRANSPORT, STAGE_1, SK, CGF>), grid, block, 0, s, Nk, N0, out, p, nbx,
                           nby);
        break;
    case STAGE_2:
        hipLaunchKernelGGL((stage_phi_kernel<T, TRANSPORT, STAGE_2, SK, CGF>), grid, block, 0, s, Nk, N0, out, p, nbx,
                           nby);
        break;
    default:
        hipLaunchKernelGGL((stage_phi_kernel<T, TRANSPORT, STAGE_3, SK, CGF>), grid, block, 0, s, Nk, N0, out, p, nbx,
                           nby);
        break;
    }
    return hipGetLastError();
}

// Ensemble work list: every member's tiles (64 x TY cells of one bin) are
// numbered consecutively, member after member (EnsembleMember::tile0 is the
// prefix sum).  The grid is persistent.  Workgroups are dealt round-robin to
// the eight XCDs, so XCD x takes the contiguous range [x T / 8, (x+1) T / 8)
// of the T tiles and its workgroups sweep it together, each taking every
// (workgroups per XCD)-th tile.  This kernel serves the transports without
// spectral coupling (refraction members go to ensemble_phi_kernel); within a
// member the order is tile column fastest, then tile row, kappa, phi (as in
// tile_of_spatial), so the workgroups of one XCD work on neighbouring tiles
// of one bin at the same time and share their x and y halos in that XCD's
// L2.  The member is looked up (binary search over tile0) only when a
// workgroup's next tile lies past the member it last worked on, starting from
// a table that gives the member of every 2^ENSEMBLE_CHUNK_SHIFT-th tile.
template <int TRANSPORT, int STAGE, bool SRC, int TY>
__global__ __launch_bounds__(64 * TY) void ensemble_stage_kernel(const EnsembleMember *__restrict__ mem, int members,
                                                                 long long total,
                                                                 const int *__restrict__ chunk_member)
{
    const int xcd = blockIdx.x & 7;
    const long long stride = (long long)(gridDim.x >> 3);
    const long long lo = total * xcd / 8, hi = total * (xcd + 1) / 8;
    int a = 0;
    long long a_begin = 0, a_end = 0;
    for (long long L = lo + (blockIdx.x >> 3); L < hi; L += stride) {
        if (L >= a_end) {
            // the member holding the first tile of L's chunk, then forward to L's member
            int u = chunk_member[L >> ENSEMBLE_CHUNK_SHIFT];
            while (u + 1 < members && mem[u + 1].tile0 <= L)
                ++u;
            a = __builtin_amdgcn_readfirstlane(u);
            a_begin = mem[a].tile0;
            a_end = a + 1 < members ? mem[a + 1].tile0 : total;
        }
        const EnsembleMember &e = mem[a];
        // The tile coordinates are uniform across the workgroup; saying so keeps
        // the member's parameters and the slab bases in scalar registers.
        int q = __builtin_amdgcn_readfirstlane((int)(L - a_begin));
        const int ib = q % e.nbx;
        q /= e.nbx;
        const int jb = q % e.nby;
        q /= e.nby;
        const int m = q % e.p.nk, n = q / e.p.nk;
        stage_tile<TRANSPORT, STAGE, SRC, TY, true>(e.Nk, e.N0, e.out, e.p, ib, jb, m, n);
    }
}

// Ensemble refraction and intrinsic stages: the same work list with one
// entry per (64 x TY tile, kappa bin) of a member, in the per-model kernel's
// order (refraction: kappa fastest, then tile row and tile column;
// intrinsic: tile column fastest, then tile row and kappa); each workgroup
// walks the phi bins of its entry as the per-model kernel does (phi_tile),
// so a member's results are those of the member stepped alone, bit for bit.
// At least five waves per SIMD: the refraction-with-sources instantiation
// needs a few registers more than the per-model kernel (descriptor-table
// addressing) and would drop to four; held at five it spills a few values
// and runs faster (measured on 32 config-5 members: 157.5 -> 149.9 ms per
// step).
template <int TRANSPORT, int STAGE, int SK, int TY>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(5, 8))) void ensemble_phi_kernel(
    const EnsembleMember *__restrict__ mem, int members, long long total, const int *__restrict__ chunk_member)
{
    const int xcd = blockIdx.x & 7;
    const long long stride = (long long)(gridDim.x >> 3);
    const long long lo = total * xcd / 8, hi = total * (xcd + 1) / 8;
    int a = 0;
    long long a_begin = 0, a_end = 0;
    for (long long L = lo + (blockIdx.x >> 3); L < hi; L += stride) {
        if (L >= a_end) {
            int u = chunk_member[L >> ENSEMBLE_CHUNK_SHIFT];
            while (u + 1 < members && mem[u + 1].tile0 <= L)
                ++u;
            a = __builtin_amdgcn_readfirstlane(u);
            a_begin = mem[a].tile0;
            a_end = a + 1 < members ? mem[a + 1].tile0 : total;
        }
        const EnsembleMember &e = mem[a];
        int q = __builtin_amdgcn_readfirstlane((int)(L - a_begin));
        int m, jb, ib;
        if constexpr (TRANSPORT == RPL_TRANSPORT_REFRACTION) {
            m = q % e.p.nk;
            q /= e.p.nk;
            jb = q % e.nby;
            ib = q / e.nby;
        } else {
            ib = q % e.nbx;
            q /= e.nbx;
            jb = q % e.nby;
            m = q / e.nby;
        }
        phi_tile<float, TRANSPORT, STAGE, SK, TY, true>(e.Nk, e.N0, e.out, e.p, ib, jb, m);
    }
}

template <int TRANSPORT, bool SRC>
hipError_t launch_t(const StageParams<float> &p, int stage, const float *Nk, const float *N0, float *out,
                    hipStream_t s)
{
    const int nbx = (p.nx + 63) / 64, nby = (p.ny + STAGE_ROWS - 1) / STAGE_ROWS;
    const long long total = (long long)nbx * nby * p.nk * p.nphi;
    if (total > 0x7fffffffLL)
        return hipErrorInvalidValue;
    dim3 block(64, STAGE_ROWS, 1);
    dim3 grid((unsigned)total, 1, 1);
    switch (stage) {
    case STAGE_TENDENCY:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_TENDENCY, SRC>), grid, block, 0, s, Nk, N0, out, p, nbx, nby);
        break;
    case STAGE_1:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_1, SRC>), grid, block, 0, s, Nk, N0, out, p, nbx, nby);
        break;
    case STAGE_2:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_2, SRC>), grid, block, 0, s, Nk, N0, out, p, nbx, nby);
        break;
    default:
        hipLaunchKernelGGL((stage_fast_kernel<TRANSPORT, STAGE_3, SRC>), grid, block, 0, s, Nk, N0, out, p, nbx, nby);
        break;
    }
    return hipGetLastError();
}

template <int TRANSPORT, int STAGE, int SK>
hipError_t launch_ens_ty(const EnsembleMember *mem, int members, long long total, const int *chunk_member, int ty,
                         int blocks, hipStream_t s)
{
    if constexpr (TRANSPORT == RPL_TRANSPORT_REFRACTION || TRANSPORT == RPL_TRANSPORT_INTRINSIC) {
        switch (ty) {
        case 2:
            hipLaunchKernelGGL((ensemble_phi_kernel<TRANSPORT, STAGE, SK, 2>), dim3(blocks), dim3(64, 2), 0, s, mem,
                               members, total, chunk_member);
            break;
        case 8:
            hipLaunchKernelGGL((ensemble_phi_kernel<TRANSPORT, STAGE, SK, 8>), dim3(blocks), dim3(64, 8), 0, s, mem,
                               members, total, chunk_member);
            break;
        default:
            hipLaunchKernelGGL((ensemble_phi_kernel<TRANSPORT, STAGE, SK, 4>), dim3(blocks), dim3(64, 4), 0, s, mem,
                               members, total, chunk_member);
            break;
        }
        return hipGetLastError();
    }
    switch (ty) {
    case 2:
        hipLaunchKernelGGL((ensemble_stage_kernel<TRANSPORT, STAGE, SK != SOURCE_NONE, 2>), dim3(blocks), dim3(64, 2), 0, s, mem,
                           members, total, chunk_member);
        break;
    case 8:
        hipLaunchKernelGGL((ensemble_stage_kernel<TRANSPORT, STAGE, SK != SOURCE_NONE, 8>), dim3(blocks), dim3(64, 8), 0, s, mem,
                           members, total, chunk_member);
        break;
    default:
        hipLaunchKernelGGL((ensemble_stage_kernel<TRANSPORT, STAGE, SK != SOURCE_NONE, 4>), dim3(blocks), dim3(64, 4), 0, s, mem,
                           members, total, chunk_member);
        break;
    }
    return hipGetLastError();
}

template <int TRANSPORT, int SK>
hipError_t launch_ens_stage(const EnsembleMember *mem, int members, long long total, const int *chunk_member,
                            int stage, int ty, int blocks, hipStream_t s)
{
    switch (stage) {
    case STAGE_1:
        return launch_ens_ty<TRANSPORT, STAGE_1, SK>(mem, members, total, chunk_member, ty, blocks, s);
    case STAGE_2:
        return launch_ens_ty<TRANSPORT, STAGE_2, SK>(mem, members, total, chunk_member, ty, blocks, s);
    default:
        return launch_ens_ty<TRANSPORT, STAGE_3, SK>(mem, members, total, chunk_member, ty, blocks, s);
    }
}

}  // namespace

hipError_t launch_ensemble_stage(const EnsembleMember *mem, int members, long long total, const int *chunk_member,
                                 int transport, int src_kind, int stage, int ty, int blocks, hipStream_t s)
{
    if (members < 1 || total < 1 || !chunk_member || blocks < 8 || blocks % 8 || (ty != 2 && ty != 4 && ty != 8) ||
        stage < STAGE_1 || stage > STAGE_3 || src_kind < SOURCE_NONE || src_kind > SOURCE_GENERIC)
        return hipErrorInvalidValue;
    const int *cm = chunk_member;
    if (transport == RPL_TRANSPORT_REFRACTION) {
        if (src_kind == SOURCE_JP_LS)
            return launch_ens_stage<RPL_TRANSPORT_REFRACTION, SOURCE_JP_LS>(mem, members, total, cm, stage, ty, blocks, s);
        return src_kind == SOURCE_GENERIC
                   ? launch_ens_stage<RPL_TRANSPORT_REFRACTION, SOURCE_GENERIC>(mem, members, total, cm, stage, ty, blocks, s)
                   : launch_ens_stage<RPL_TRANSPORT_REFRACTION, SOURCE_NONE>(mem, members, total, cm, stage, ty, blocks, s);
    }
    if (transport == RPL_TRANSPORT_INTRINSIC) {
        if (src_kind == SOURCE_JP_LS)
            return launch_ens_stage<RPL_TRANSPORT_INTRINSIC, SOURCE_JP_LS>(mem, members, total, cm, stage, ty, blocks, s);
        return src_kind == SOURCE_GENERIC
                   ? launch_ens_stage<RPL_TRANSPORT_INTRINSIC, SOURCE_GENERIC>(mem, members, total, cm, stage, ty, blocks, s)
                   : launch_ens_stage<RPL_TRANSPORT_INTRINSIC, SOURCE_NONE>(mem, members, total, cm, stage, ty, blocks, s);
    }
    return hipErrorInvalidValue;
}

// Fills the packed bin table of a phi-walk launch (refraction or intrinsic
// transport; the kernels that read it are the per-model ones without the
// generic source path).
static void pack_bins(const StageParams<float> &p, int transport, hipStream_t s)
{
    if (!p.bt || transport == RPL_TRANSPORT_NONE)
        return;
    const int nbins = p.nk * p.nphi;
    hipLaunchKernelGGL(pack_bins_kernel, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, s,
                       const_cast<float *>(p.bt), nbins, p.nk, p.cgx, p.cgy, p.cosp, p.sinp, p.src, p.tau_c,
                       p.tau_s);
}

hipError_t launch_stage_fast(const StageParams<float> &p, int transport, int stage, const float *Nk, const float *N0,
                             float *out, hipStream_t s)
{
    if (p.cgm && transport != RPL_TRANSPORT_NONE) {
        // a depth per column: generic source path
        if (transport == RPL_TRANSPORT_REFRACTION)
            return p.src.on ? launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_GENERIC, true>(p, stage, Nk, N0, out, s)
                            : launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_NONE, true>(p, stage, Nk, N0, out, s);
        return p.src.on ? launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_GENERIC, true>(p, stage, Nk, N0, out, s)
                        : launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_NONE, true>(p, stage, Nk, N0, out, s);
    }
    if (p.src.on) {
        // the specialised bundle paths (same bits as the generic one)
        const bool jp = source_is_jp_ls(p.src), net = source_is_net_ls(p.src);
        if (jp || net)
            pack_bins(p, transport, s);
        if (transport == RPL_TRANSPORT_REFRACTION)
            return jp    ? launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_JP_LS>(p, stage, Nk, N0, out, s)
                   : net ? launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_NET_LS>(p, stage, Nk, N0, out, s)
                         : launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_GENERIC>(p, stage, Nk, N0, out, s);
        if (transport == RPL_TRANSPORT_INTRINSIC)
            return jp    ? launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_JP_LS>(p, stage, Nk, N0, out, s)
                   : net ? launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_NET_LS>(p, stage, Nk, N0, out, s)
                         : launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_GENERIC>(p, stage, Nk, N0, out, s);
        return launch_t<RPL_TRANSPORT_NONE, true>(p, stage, Nk, N0, out, s);
    }
    pack_bins(p, transport, s);
    if (transport == RPL_TRANSPORT_REFRACTION)
        return launch_phi<float, RPL_TRANSPORT_REFRACTION, SOURCE_NONE>(p, stage, Nk, N0, out, s);
    if (transport == RPL_TRANSPORT_INTRINSIC)
        return launch_phi<float, RPL_TRANSPORT_INTRINSIC, SOURCE_NONE>(p, stage, Nk, N0, out, s);
    return launch_t<RPL_TRANSPORT_NONE, false>(p, stage, Nk, N0, out, s);
}

hipError_t launch_stage_fast(const StageParams<double> &p, int transport, int stage, const double *Nk,
                             const double *N0, double *out, hipStream_t s)
{
    // 32-bit byte offsets within a slab: FP64 slabs below 2^29 cells
    if ((size_t)p.nx * p.ny >= ((size_t)1 << 29))
        return hipErrorInvalidValue;
    if (p.cgm) {
        if (transport == RPL_TRANSPORT_REFRACTION)
            return p.src.on ? launch_phi<double, RPL_TRANSPORT_REFRACTION, SOURCE_GENERIC, true>(p, stage, Nk, N0, out, s)
                            : launch_phi<double, RPL_TRANSPORT_REFRACTION, SOURCE_NONE, true>(p, stage, Nk, N0, out, s);
        if (transport == RPL_TRANSPORT_INTRINSIC)
            return p.src.on ? launch_phi<double, RPL_TRANSPORT_INTRINSIC, SOURCE_GENERIC, true>(p, stage, Nk, N0, out, s)
                            : launch_phi<double, RPL_TRANSPORT_INTRINSIC, SOURCE_NONE, true>(p, stage, Nk, N0, out, s);
        return hipErrorInvalidValue;
    }
    if (transport == RPL_TRANSPORT_REFRACTION)
        return p.src.on ? launch_phi<double, RPL_TRANSPORT_REFRACTION, SOURCE_GENERIC>(p, stage, Nk, N0, out, s)
                        : launch_phi<double, RPL_TRANSPORT_REFRACTION, SOURCE_NONE>(p, stage, Nk, N0, out, s);
    if (transport == RPL_TRANSPORT_INTRINSIC)
        return p.src.on ? launch_phi<double, RPL_TRANSPORT_INTRINSIC, SOURCE_GENERIC>(p, stage, Nk, N0, out, s)
                        : launch_phi<double, RPL_TRANSPORT_INTRINSIC, SOURCE_NONE>(p, stage, Nk, N0, out, s);
    return hipErrorInvalidValue;
}

}  // namespace rpl
