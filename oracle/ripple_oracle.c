/*
 * ripple oracle -- CPU restatement of the Ripple.jl wave-action time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product library (ripple.jl_amd/) never links, loads or
 * calls it, and has no CPU fallback.
 *
 * It restates, in plain C and in the reference's operation order, the
 * reference functions cited next to each definition (paths relative to the
 * Ripple.jl source tree).  Built with -ffp-contract=off (Julia forms no FMA).
 * FP64 is the oracle precision; an FP32 instantiation exists to compare the
 * FP32 device path op-for-op where that is meaningful.
 *
 * Pinning: each function is checked in tests/test_oracle.py against the
 * reference's own known answers (test/coordinate_grids/finite_volume_
 * integration.jl, test/diagnostics/moments.jl, test/coupling/q_transform.jl,
 * test/integration/transport.jl) and against closed forms (WENO5 exactness on
 * polynomial data, analytic derivatives, RK3 amplification factor, discrete
 * conservation).  Sources and winds carry their own notes in sources_impl.h.
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Coupling/wave_current_refraction.jl:66-67 (_periodic, _clamp_idx), 0-based. */
static inline int orc_periodic(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
static inline int orc_clamp(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

#define REAL double
#define REAL_EPS DBL_EPSILON
#define SFX(name) name##_f64
/* Group velocities per cell (a spatially varying depth, transport.jl:
 * 105-109 intrinsic_velocity_component of a 4-D table) when set, per (kappa,
 * phi) bin otherwise.  Set by the front end before each transport call. */
static int g_cg_cells = 0;
int orc_set_cg_cells(int on)
{
    g_cg_cells = on != 0;
    return 0;
}

/* Pseudomomentum overlap tables and Q weights per column (a spatially
 * varying depth: pseudomomentum.jl:72-87 overlap[i, j, t, s], :356/:380
 * q_depth_at(depth, i, j)) when set: O[ij + nx ny (t + nk s)] and
 * W[ij + nx ny (k + nz m)]; one table for every column otherwise. */
static int g_pm_cols = 0;
int orc_set_pm_columns(int on)
{
    g_pm_cols = on != 0;
    return 0;
}

#include "oracle_impl.h"
#undef REAL
#undef REAL_EPS
#undef SFX

#define REAL float
#define REAL_EPS FLT_EPSILON
#define SFX(name) name##_f32
#include "oracle_impl.h"
#undef REAL
#undef REAL_EPS
#undef SFX

#include "sources_impl.h"

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Spectral geometry (FP64, the reference's coordinate float type).          */
/* ------------------------------------------------------------------------ */

/* CoordinateGrids/spectral_grids.jl:4-27: faces at midpoints, end faces
 * mirrored about the end centres; lower face clamped when lower_bounded. */
void orc_centers_to_faces(int n, const double *c, double *f, int lower_bounded, double lower)
{
    if (n == 1) {
        f[0] = c[0] - 0.5;
        f[1] = c[0] + 0.5;
    } else {
        for (int i = 1; i < n; ++i)
            f[i] = (c[i - 1] + c[i]) / 2;
        f[0] = c[0] - (f[1] - c[0]);
        f[n] = c[n - 1] + (c[n - 1] - f[n - 1]);
    }
    if (lower_bounded && f[0] < lower)
        f[0] = lower;
}

/* CoordinateGrids/polar_wavevector_grid.jl:34-37: uniform phi faces starting
 * half a spacing below the first centre. */
void orc_phi_faces(int n, const double *phi_c, double *f)
{
    double dphi = n == 1 ? 2 * M_PI : 2 * M_PI / n;
    for (int i = 0; i <= n; ++i)
        f[i] = phi_c[0] - dphi / 2 + i * dphi;
}

/* polar_wavevector_grid.jl:44-47: w[m,n] = (kf[m+1]^2 - kf[m]^2)/2 * dphi[n]. */
void orc_cell_measures(int nk, int nphi, const double *kf, const double *pf, double *w)
{
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m) {
            double radial = (kf[m + 1] * kf[m + 1] - kf[m] * kf[m]) / 2;
            w[m + (size_t)nk * n] = radial * (pf[n + 1] - pf[n]);
        }
}

/* polar_wavevector_grid.jl:73-115: exact first moments and the trace of the
 * second moments, (kf2^4 - kf1^4)/4 * dphi. */
void orc_moment_measures(int nk, int nphi, const double *kf, const double *pf,
                         double *mu_x, double *mu_y, double *trace)
{
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m) {
            double k1 = kf[m], k2 = kf[m + 1], p1 = pf[n], p2 = pf[n + 1];
            double r1 = (k2 * k2 * k2 - k1 * k1 * k1) / 3;
            size_t t = (size_t)m + (size_t)nk * n;
            mu_x[t] = r1 * (sin(p2) - sin(p1));
            mu_y[t] = r1 * (cos(p1) - cos(p2));
            double r2 = (pow(k2, 4) - pow(k1, 4)) / 4;
            double dphi = p2 - p1;
            double s2 = sin(2 * p2) - sin(2 * p1);
            double cc = dphi / 2 + s2 / 4, ss = dphi / 2 - s2 / 4;
            trace[t] = r2 * cc + r2 * ss;
        }
}

/* Diagnostics/bulk_statistics.jl:23-26 (_deep_water_frequency_cell on polar
 * and frequency-direction grids) = sqrt(g) * spectral_radial_power_measure(
 * grid, m, n, 1//2), with CoordinateGrids/finite_volume_integration.jl:24-27
 * (radial_power_annular_measure): (phi2 - phi1) * (k2^(p+2) - k1^(p+2)) /
 * (p + 2) at p = 1/2.  Julia evaluates 1//2 + 2 as the exact rational 5//2,
 * so the exponent and divisor are the double 2.5. */
void orc_energy_measures(int nk, int nphi, const double *kf, const double *pf, double gravity, double *e)
{
    const double sg = sqrt(gravity);
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m) {
            double k1 = kf[m], k2 = kf[m + 1], p1 = pf[n], p2 = pf[n + 1];
            double annular = (p2 - p1) * (pow(k2, 2.5) - pow(k1, 2.5)) / 2.5;
            e[(size_t)m + (size_t)nk * n] = sg * annular;
        }
}

/* Models/transport.jl:61-70: deep-water group velocity averaged over the
 * spectral cell; :72-83: finite depth, point-sampled at (kappa_m, phi_n). */
void orc_group_velocity_tables(int nk, int nphi, const double *kc, const double *kf,
                               const double *pc, const double *pf, const double *w,
                               int finite_depth, double depth, double gravity,
                               double *cgx, double *cgy)
{
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m) {
            size_t t = (size_t)m + (size_t)nk * n;
            if (!finite_depth) {
                double k1 = kf[m], k2 = kf[m + 1], p1 = pf[n], p2 = pf[n + 1];
                double radial = (pow(k2, 1.5) - pow(k1, 1.5)) / 1.5;
                double ax = sin(p2) - sin(p1), ay = cos(p1) - cos(p2);
                double scale = sqrt(gravity) / (2 * w[t]);
                cgx[t] = scale * radial * ax;
                cgy[t] = scale * radial * ay;
            } else {
                double k = kc[m];
                double kx = k * cos(pc[n]), ky = k * sin(pc[n]);
                if (k == 0) {
                    cgx[t] = 0;
                    cgy[t] = 0;
                    continue;
                }
                double mu = k * depth;
                double phase = sqrt(gravity * tanh(mu) / k);
                double factor = (1 + 2 * mu / sinh(2 * mu)) / 2;
                double cg = factor * phase;
                cgx[t] = cg * kx / k;
                cgy[t] = cg * ky / k;
            }
        }
}

/* Models/transport.jl:111-139 with a spatially varying depth: every column's
 * point-sampled finite-depth group velocity (transport.jl:72-83) at its own
 * depth[i + nx j], tables (nx, ny, nk, nphi), i fastest. */
void orc_group_velocity_field(int nx, int ny, int nk, int nphi, const double *kc, const double *pc,
                              const double *depth, double gravity, double *cgx, double *cgy)
{
    const size_t plane = (size_t)nx * ny;
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m)
            for (size_t ij = 0; ij < plane; ++ij) {
                size_t o = ij + plane * ((size_t)m + (size_t)nk * n);
                double k = kc[m];
                double kx = k * cos(pc[n]), ky = k * sin(pc[n]);
                if (k == 0) {
                    cgx[o] = 0;
                    cgy[o] = 0;
                    continue;
                }
                double mu = k * depth[ij];
                double phase = sqrt(gravity * tanh(mu) / k);
                double factor = (1 + 2 * mu / sinh(2 * mu)) / 2;
                double cg = factor * phase;
                cgx[o] = cg * kx / k;
                cgy[o] = cg * ky / k;
            }
}

/* ------------------------------------------------------------------------ */
/* Q-transform (Coupling/q_kernel.jl:60-92, doppler_velocity.jl:45-62).       */
/* ------------------------------------------------------------------------ */

/* q_kernel.jl:77-85 (q_cdf_kernel): overflow-safe CDF of the vertical
 * projection kernel, s itself when |kappa*d| < small_mu. */
double orc_q_cdf(double kappa, double z, double depth, double small_mu)
{
    double mu = kappa * depth;
    double sgn = mu < 0 ? -1.0 : 1.0;
    double mus = fabs(mu) < small_mu ? sgn * small_mu : mu;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    double a = 2 * mus * s;
    double b = 2 * mus;
    double fin = (exp(a - b) - exp(-a - b)) / (1 - exp(-2 * b));
    return fabs(mu) < small_mu ? s : fin;
}

/* q_kernel.jl:87-92: W[k,m] = CDF(zhigh) - CDF(zlow) for cell k, stored
 * k-fastest (Julia weights[k, m]). */
void orc_q_weights(int nz, const double *zf, int nk, const double *kappa, double depth, double *W)
{
    for (int m = 0; m < nk; ++m)
        for (int k = 0; k < nz; ++k) {
            double z1 = zf[k], z2 = zf[k + 1];
            double lo = z1 < z2 ? z1 : z2, hi = z1 < z2 ? z2 : z1;
            W[k + (size_t)nz * m] = orc_q_cdf(kappa[m], hi, depth, 1e-6) - orc_q_cdf(kappa[m], lo, depth, 1e-6);
        }
}

/* The weights of every column at its own depth (q_depth_at(depth, i, j),
 * q_transform.jl:124-125): Wc[ij + nx ny (k + nz m)]. */
void orc_q_weights_columns(int ncol, int nz, const double *zf, int nk, const double *kappa, const double *depth,
                           double *Wc)
{
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)ncol; ++ij)
        for (int m = 0; m < nk; ++m)
            for (int k = 0; k < nz; ++k) {
                double z1 = zf[k], z2 = zf[k + 1];
                double lo = z1 < z2 ? z1 : z2, hi = z1 < z2 ? z2 : z1;
                Wc[(size_t)ij + (size_t)ncol * ((size_t)k + (size_t)nz * m)] =
                    orc_q_cdf(kappa[m], hi, depth[ij], 1e-6) - orc_q_cdf(kappa[m], lo, depth[ij], 1e-6);
            }
}

/* q_kernel.jl:30-33 (q_cell_integral): CDF(zhigh) - CDF(zlow). */
double orc_q_cell_integral(double kappa, double z1, double z2, double depth)
{
    double lo = z1 < z2 ? z1 : z2, hi = z1 < z2 ? z2 : z1;
    return orc_q_cdf(kappa, hi, depth, 1e-6) - orc_q_cdf(kappa, lo, depth, 1e-6);
}

/* q_kernel.jl:35-53 (q_cdf_kappa_derivative): d CDF / d kappa, with the
 * small-mu branch depth * 4 mu s (s^2 - 1) / 3. */
double orc_q_cdf_dkappa(double kappa, double z, double depth, double small_mu)
{
    double mu = kappa * depth;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    if (fabs(mu) < small_mu)
        return depth * (4 * mu * s * (s * s - 1) / 3);
    double a = exp(-2 * mu * (1 - s));
    double b = exp(-2 * mu * (1 + s));
    double e = exp(-4 * mu);
    double num = a - b, den = 1 - e;
    double dnum = -2 * (1 - s) * a + 2 * (1 + s) * b;
    double dden = 4 * e;
    double dF = (dnum * den - num * dden) / (den * den);
    return depth * dF;
}

/* q_kernel.jl:55-59 (q_cell_integral_kappa_derivative). */
double orc_q_cell_integral_dkappa(double kappa, double z1, double z2, double depth)
{
    double lo = z1 < z2 ? z1 : z2, hi = z1 < z2 ? z2 : z1;
    return orc_q_cdf_dkappa(kappa, hi, depth, 1e-6) - orc_q_cdf_dkappa(kappa, lo, depth, 1e-6);
}

/* Coupling/pseudomomentum.jl:103-122 (pseudomomentum_overlap) for a constant
 * depth (pseudomomentum.jl:89-101):
 *   O[t,s]  = sum_k W(kappa_t, k) W(kappa_s, k) / |dz_k|
 *   dO[t,s] = sum_k dW/dkappa(kappa_t, k) W(kappa_s, k) / |dz_k|
 * stored target fastest. */
void orc_pm_overlap(int nz, const double *zf, int nk, const double *kappa, double depth, double *O, double *dO)
{
    for (int s = 0; s < nk; ++s)
        for (int t = 0; t < nk; ++t) {
            double total = 0, dtotal = 0;
            for (int k = 0; k < nz; ++k) {
                double z1 = zf[k], z2 = zf[k + 1];
                double dz = fabs(z2 - z1);
                double sw = orc_q_cell_integral(kappa[s], z1, z2, depth);
                double tw = orc_q_cell_integral(kappa[t], z1, z2, depth);
                double td = orc_q_cell_integral_dkappa(kappa[t], z1, z2, depth);
                total += tw * sw / dz;
                dtotal += td * sw / dz;
            }
            O[(size_t)t + (size_t)nk * s] = total;
            dO[(size_t)t + (size_t)nk * s] = dtotal;
        }
}

/* pseudomomentum.jl:72-87: the overlap tables of every column at its own
 * depth (pseudomomentum_overlap, :103-122), O[ij + ncol (t + nk s)]. */
void orc_pm_overlap_columns(int ncol, int nz, const double *zf, int nk, const double *kappa, const double *depth,
                            double *O, double *dO)
{
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)ncol; ++ij)
        for (int s = 0; s < nk; ++s)
            for (int t = 0; t < nk; ++t) {
                double total = 0, dtotal = 0;
                for (int k = 0; k < nz; ++k) {
                    double z1 = zf[k], z2 = zf[k + 1];
                    double dz = fabs(z2 - z1);
                    double sw = orc_q_cell_integral(kappa[s], z1, z2, depth[ij]);
                    double tw = orc_q_cell_integral(kappa[t], z1, z2, depth[ij]);
                    double td = orc_q_cell_integral_dkappa(kappa[t], z1, z2, depth[ij]);
                    total += tw * sw / dz;
                    dtotal += td * sw / dz;
                }
                O[(size_t)ij + (size_t)ncol * ((size_t)t + (size_t)nk * s)] = total;
                dO[(size_t)ij + (size_t)ncol * ((size_t)t + (size_t)nk * s)] = dtotal;
            }
}

/* doppler_velocity.jl:45-62: U[i,j,m] = sum_k uL[i,j,k] * W[k,m], sequential
 * accumulation over k. */
void orc_doppler(int nx, int ny, int nz, const double *uL, const double *vL, int nk,
                 const double *W, double *Ux, double *Uy)
{
#pragma omp parallel for schedule(static)
    for (long jm = 0; jm < (long)ny * nk; ++jm) {
        int j = (int)(jm % ny), m = (int)(jm / ny);
        for (int i = 0; i < nx; ++i) {
            double ax = 0, ay = 0;
            for (int k = 0; k < nz; ++k) {
                double q = W[k + (size_t)nz * m];
                size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                ax += uL[c] * q;
                ay += vL[c] * q;
            }
            size_t o = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * m);
            Ux[o] = ax;
            Uy[o] = ay;
        }
    }
}

/* doppler_velocity.jl:45-62 with a depth per column: the weights of column
 * ij are Wc[ij + nx ny (k + nz m)]; sequential accumulation over k. */
void orc_doppler_columns(int nx, int ny, int nz, const double *uL, const double *vL, int nk, const double *Wc,
                         double *Ux, double *Uy)
{
    const size_t plane = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (long jm = 0; jm < (long)ny * nk; ++jm) {
        int j = (int)(jm % ny), m = (int)(jm / ny);
        for (int i = 0; i < nx; ++i) {
            const size_t ij = (size_t)i + (size_t)nx * j;
            double ax = 0, ay = 0;
            for (int k = 0; k < nz; ++k) {
                double q = Wc[ij + plane * ((size_t)k + (size_t)nz * m)];
                size_t c = ij + plane * k;
                ax += uL[c] * q;
                ay += vL[c] * q;
            }
            Ux[ij + plane * m] = ax;
            Uy[ij + plane * m] = ay;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Diagnostics needing exact sums.                                           */
/* ------------------------------------------------------------------------ */

/* Diagnostics/bulk_statistics.jl:87-91 + diagnostic_operations.jl:495-503:
 * sum_ij (sum_mn N w) * dx * dy.  The per-column sum follows the reference
 * order; the global sum is carried in double-double (TwoSum), i.e. exact to
 * well below 1e-20 relative at these sizes. */
static void two_sum(double a, double b, double *s, double *e)
{
    double x = a + b, bv = x - a, av = x - bv;
    *s = x;
    *e = (a - av) + (b - bv);
}

double orc_total_weighted_f64(int nx, int ny, int nk, int nphi, const double *N, const double *w,
                              double dx, double dy)
{
    double hi = 0, lo = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double acc = 0;
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m)
                    acc += N[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * ((size_t)m + (size_t)nk * n))] *
                           w[(size_t)m + (size_t)nk * n];
            double v = acc * dx * dy, s, e;
            two_sum(hi, v, &s, &e);
            hi = s;
            lo += e;
        }
    return hi + lo;
}

double orc_total_weighted_f32(int nx, int ny, int nk, int nphi, const float *N, const double *w,
                              double dx, double dy)
{
    double hi = 0, lo = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double acc = 0;
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m)
                    acc += (double)N[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * ((size_t)m + (size_t)nk * n))] *
                           w[(size_t)m + (size_t)nk * n];
            double v = acc * dx * dy, s, e;
            two_sum(hi, v, &s, &e);
            hi = s;
            lo += e;
        }
    return hi + lo;
}
