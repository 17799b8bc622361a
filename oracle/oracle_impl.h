/*
 * ripple oracle -- precision-generic part.  TEST INFRASTRUCTURE ONLY.
 *
 * This header is included twice by ripple_oracle.c, once with REAL=double and
 * once with REAL=float.  SFX(name) appends _f64 / _f32.  Every function is a
 * CPU restatement of the Ripple.jl reference arithmetic, in the reference's
 * operation order, compiled with -ffp-contract=off so no FMA is formed (Julia
 * does not contract either).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may call it.
 *
 * Layouts (all i-fastest, interior only -- the reference kernels never read
 * halos, they wrap indices explicitly):
 *   N, G       : [n][m][j][i]     (Nx, Ny, Nkappa, Nphi)
 *   U*, grads  : [m][j][i]        (Nx, Ny, Nkappa)
 *   tables     : [n][m]           (Nkappa, Nphi)  (Julia column-major [m, n])
 */

#define IDX4(i, j, m, n) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * ((size_t)(m) + (size_t)nk * (size_t)(n))))
#define IDX3(i, j, m) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(m)))

/* Coupling/wave_current_refraction.jl:41-55 (weno5_reconstruct). */
static inline REAL SFX(weno5_reconstruct)(REAL fm2, REAL fm1, REAL f0, REAL fp1, REAL fp2)
{
    const REAL eps = (REAL)10000 * REAL_EPS;
    const REAL c13_12 = (REAL)13 / (REAL)12;
    const REAL c1_4 = (REAL)1 / (REAL)4;
    REAL v0 = ((REAL)1 / (REAL)3) * fm2 - ((REAL)7 / (REAL)6) * fm1 + ((REAL)11 / (REAL)6) * f0;
    REAL v1 = -((REAL)1 / (REAL)6) * fm1 + ((REAL)5 / (REAL)6) * f0 + ((REAL)1 / (REAL)3) * fp1;
    REAL v2 = ((REAL)1 / (REAL)3) * f0 + ((REAL)5 / (REAL)6) * fp1 - ((REAL)1 / (REAL)6) * fp2;
    REAL d0 = fm2 - (REAL)2 * fm1 + f0, e0 = fm2 - (REAL)4 * fm1 + (REAL)3 * f0;
    REAL d1 = fm1 - (REAL)2 * f0 + fp1, e1 = fm1 - fp1;
    REAL d2 = f0 - (REAL)2 * fp1 + fp2, e2 = (REAL)3 * f0 - (REAL)4 * fp1 + fp2;
    REAL b0 = c13_12 * (d0 * d0) + c1_4 * (e0 * e0);
    REAL b1 = c13_12 * (d1 * d1) + c1_4 * (e1 * e1);
    REAL b2 = c13_12 * (d2 * d2) + c1_4 * (e2 * e2);
    REAL q0 = eps + b0, q1 = eps + b1, q2 = eps + b2;
    REAL a0 = ((REAL)1 / (REAL)10) / (q0 * q0);
    REAL a1 = ((REAL)6 / (REAL)10) / (q1 * q1);
    REAL a2 = ((REAL)3 / (REAL)10) / (q2 * q2);
    REAL sa = a0 + a1 + a2;
    return (a0 * v0 + a1 * v1 + a2 * v2) / sa;
}

/* Coupling/wave_current_refraction.jl:57-64 (weno5_face_iphalf): value at the
 * i+1/2 face, upwinded by the sign of the cell-centre velocity; vel >= 0 is
 * "positive" (zero counts as positive). */
static inline REAL SFX(weno5_face)(REAL im2, REAL im1, REAL i0, REAL ip1, REAL ip2, REAL ip3,
                                    REAL vel, int has_stencil)
{
    int positive = vel >= (REAL)0;
    if (!has_stencil)
        return positive ? i0 : ip1;
    return positive ? SFX(weno5_reconstruct)(im2, im1, i0, ip1, ip2)
                    : SFX(weno5_reconstruct)(ip3, ip2, ip1, i0, im1);
}


/* Coupling/wave_current_refraction.jl:7-19 (_current_gradients!): centred
 * periodic differences times (1/dx)/2. */
void SFX(orc_current_gradients)(int nx, int ny, int nk, const REAL *Ux, const REAL *Uy,
                                REAL dx_inv, REAL dy_inv,
                                REAL *Ux_x, REAL *Ux_y, REAL *Uy_x, REAL *Uy_y)
{
    const REAL hx = dx_inv / (REAL)2, hy = dy_inv / (REAL)2;
    for (int m = 0; m < nk; ++m)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int ip = i == nx - 1 ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
                int jp = j == ny - 1 ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
                size_t c = IDX3(i, j, m);
                Ux_x[c] = (Ux[IDX3(ip, j, m)] - Ux[IDX3(im, j, m)]) * hx;
                Ux_y[c] = (Ux[IDX3(i, jp, m)] - Ux[IDX3(i, jm, m)]) * hy;
                Uy_x[c] = (Uy[IDX3(ip, j, m)] - Uy[IDX3(im, j, m)]) * hx;
                Uy_y[c] = (Uy[IDX3(i, jp, m)] - Uy[IDX3(i, jm, m)]) * hy;
            }
}

/* Coupling/wave_current_refraction.jl:74-151 (_wave_current_refraction_tendency!).
 * G = -(flux_x + flux_y + flux_kappa + flux_phi); x, y, phi periodic WENO5;
 * kappa WENO5 only where the 6-point stencil fits, else first-order upwind,
 * zero flux through the two outer kappa faces.  Cell-centre velocity on both
 * faces. */
void SFX(orc_refraction_tendency)(int nx, int ny, int nk, int nphi, const REAL *N, REAL *G,
                                  const REAL *Ux, const REAL *Uy,
                                  const REAL *Ux_x, const REAL *Ux_y, const REAL *Uy_x, const REAL *Uy_y,
                                  const REAL *kappa_c, const REAL *cgx, const REAL *cgy,
                                  const REAL *cos_t, const REAL *sin_t,
                                  REAL dx_inv, REAL dy_inv, REAL dk_inv, REAL dp_inv)
{
    long total = (long)nphi * nk * ny;
#pragma omp parallel for schedule(static)
    for (long jmn = 0; jmn < total; ++jmn) {
        int j = (int)(jmn % ny);
        int m = (int)((jmn / ny) % nk);
        int n = (int)(jmn / ((long)ny * nk));
        const REAL kap = kappa_c[m];
        const REAL cph = cos_t[n], sph = sin_t[n];
        for (int i = 0; i < nx; ++i) {
            size_t c3 = IDX3(i, j, m);
            REAL Uxc = Ux[c3], Uyc = Uy[c3];
            REAL Uxx = Ux_x[c3], Uxy = Ux_y[c3], Uyx = Uy_x[c3], Uyy = Uy_y[c3];
            REAL ck = -kap * (cph * cph * Uxx + cph * sph * (Uyx + Uxy) + sph * sph * Uyy);
            REAL cp = cph * sph * (Uxx - Uyy) + sph * sph * Uyx - cph * cph * Uxy;
            REAL ux = (g_cg_cells ? cgx[IDX4(i, j, m, n)] : cgx[m + (size_t)nk * n]) + Uxc;
            REAL uy = (g_cg_cells ? cgy[IDX4(i, j, m, n)] : cgy[m + (size_t)nk * n]) + Uyc;

            REAL xm3 = N[IDX4(orc_periodic(i - 3, nx), j, m, n)], xm2 = N[IDX4(orc_periodic(i - 2, nx), j, m, n)];
            REAL xm1 = N[IDX4(orc_periodic(i - 1, nx), j, m, n)], x0 = N[IDX4(i, j, m, n)];
            REAL xp1 = N[IDX4(orc_periodic(i + 1, nx), j, m, n)], xp2 = N[IDX4(orc_periodic(i + 2, nx), j, m, n)];
            REAL xp3 = N[IDX4(orc_periodic(i + 3, nx), j, m, n)];
            REAL Nxp = SFX(weno5_face)(xm2, xm1, x0, xp1, xp2, xp3, ux, 1);
            REAL Nxm = SFX(weno5_face)(xm3, xm2, xm1, x0, xp1, xp2, ux, 1);
            REAL flux_x = ux * (Nxp - Nxm) * dx_inv;

            REAL ym3 = N[IDX4(i, orc_periodic(j - 3, ny), m, n)], ym2 = N[IDX4(i, orc_periodic(j - 2, ny), m, n)];
            REAL ym1 = N[IDX4(i, orc_periodic(j - 1, ny), m, n)], y0 = x0;
            REAL yp1 = N[IDX4(i, orc_periodic(j + 1, ny), m, n)], yp2 = N[IDX4(i, orc_periodic(j + 2, ny), m, n)];
            REAL yp3 = N[IDX4(i, orc_periodic(j + 3, ny), m, n)];
            REAL Nyp = SFX(weno5_face)(ym2, ym1, y0, yp1, yp2, yp3, uy, 1);
            REAL Nym = SFX(weno5_face)(ym3, ym2, ym1, y0, yp1, yp2, uy, 1);
            REAL flux_y = uy * (Nyp - Nym) * dy_inv;

            int has_p = (m - 2 >= 0) && (m + 3 <= nk - 1);
            int has_m = (m - 3 >= 0) && (m + 2 <= nk - 1);
            REAL km3 = N[IDX4(i, j, orc_clamp(m - 3, nk), n)], km2 = N[IDX4(i, j, orc_clamp(m - 2, nk), n)];
            REAL km1 = N[IDX4(i, j, orc_clamp(m - 1, nk), n)], k0 = x0;
            REAL kp1 = N[IDX4(i, j, orc_clamp(m + 1, nk), n)], kp2 = N[IDX4(i, j, orc_clamp(m + 2, nk), n)];
            REAL kp3 = N[IDX4(i, j, orc_clamp(m + 3, nk), n)];
            REAL Nkp = SFX(weno5_face)(km2, km1, k0, kp1, kp2, kp3, ck, has_p);
            REAL Nkm = SFX(weno5_face)(km3, km2, km1, k0, kp1, kp2, ck, has_m);
            REAL fkp = (m == nk - 1) ? (REAL)0 : ck * Nkp;
            REAL fkm = (m == 0) ? (REAL)0 : ck * Nkm;
            REAL flux_k = (fkp - fkm) * dk_inv;

            REAL pm3 = N[IDX4(i, j, m, orc_periodic(n - 3, nphi))], pm2 = N[IDX4(i, j, m, orc_periodic(n - 2, nphi))];
            REAL pm1 = N[IDX4(i, j, m, orc_periodic(n - 1, nphi))], p0 = x0;
            REAL pp1 = N[IDX4(i, j, m, orc_periodic(n + 1, nphi))], pp2 = N[IDX4(i, j, m, orc_periodic(n + 2, nphi))];
            REAL pp3 = N[IDX4(i, j, m, orc_periodic(n + 3, nphi))];
            REAL Npp = SFX(weno5_face)(pm2, pm1, p0, pp1, pp2, pp3, cp, 1);
            REAL Npm = SFX(weno5_face)(pm3, pm2, pm1, p0, pp1, pp2, cp, 1);
            REAL flux_p = cp * (Npp - Npm) * dp_inv;

            G[IDX4(i, j, m, n)] = -(flux_x + flux_y + flux_k + flux_p);
        }
    }
}

/* Models/intrinsic_transport.jl:14-68 (_intrinsic_transport_kernel!), periodic
 * x and y (the only topology the reference dispatches to it,
 * Models/tendencies.jl:43-50): x/y WENO5 at the bin's intrinsic c_g. */
void SFX(orc_intrinsic_tendency)(int nx, int ny, int nk, int nphi, const REAL *N, REAL *G,
                                 const REAL *cgx, const REAL *cgy, REAL dx_inv, REAL dy_inv)
{
    long total = (long)nphi * nk * ny;
#pragma omp parallel for schedule(static)
    for (long jmn = 0; jmn < total; ++jmn) {
        int j = (int)(jmn % ny);
        int m = (int)((jmn / ny) % nk);
        int n = (int)(jmn / ((long)ny * nk));
        const REAL ux_bin = cgx[m + (size_t)nk * n], uy_bin = cgy[m + (size_t)nk * n];
        for (int i = 0; i < nx; ++i) {
            const REAL ux = g_cg_cells ? cgx[IDX4(i, j, m, n)] : ux_bin;
            const REAL uy = g_cg_cells ? cgy[IDX4(i, j, m, n)] : uy_bin;
            REAL xm3 = N[IDX4(orc_periodic(i - 3, nx), j, m, n)], xm2 = N[IDX4(orc_periodic(i - 2, nx), j, m, n)];
            REAL xm1 = N[IDX4(orc_periodic(i - 1, nx), j, m, n)], x0 = N[IDX4(i, j, m, n)];
            REAL xp1 = N[IDX4(orc_periodic(i + 1, nx), j, m, n)], xp2 = N[IDX4(orc_periodic(i + 2, nx), j, m, n)];
            REAL xp3 = N[IDX4(orc_periodic(i + 3, nx), j, m, n)];
            REAL Nxp = SFX(weno5_face)(xm2, xm1, x0, xp1, xp2, xp3, ux, 1);
            REAL Nxm = SFX(weno5_face)(xm3, xm2, xm1, x0, xp1, xp2, ux, 1);
            REAL flux_x = ux * (Nxp - Nxm) * dx_inv;
            REAL ym3 = N[IDX4(i, orc_periodic(j - 3, ny), m, n)], ym2 = N[IDX4(i, orc_periodic(j - 2, ny), m, n)];
            REAL ym1 = N[IDX4(i, orc_periodic(j - 1, ny), m, n)];
            REAL yp1 = N[IDX4(i, orc_periodic(j + 1, ny), m, n)], yp2 = N[IDX4(i, orc_periodic(j + 2, ny), m, n)];
            REAL yp3 = N[IDX4(i, orc_periodic(j + 3, ny), m, n)];
            REAL Nyp = SFX(weno5_face)(ym2, ym1, x0, yp1, yp2, yp3, uy, 1);
            REAL Nym = SFX(weno5_face)(ym3, ym2, ym1, x0, yp1, yp2, uy, 1);
            REAL flux_y = uy * (Nyp - Nym) * dy_inv;
            G[IDX4(i, j, m, n)] = -(flux_x + flux_y);
        }
    }
}

/* Models/time_step.jl:67-74 (_add_scaled_kernel!): N = max(0, N + dt*G).
 * Julia's max(0, NaN) is NaN, so NaN is passed through. */
static inline REAL SFX(clip0)(REAL x) { return x < (REAL)0 ? (REAL)0 : x; }

void SFX(orc_add_scaled)(size_t count, REAL *N, const REAL *G, REAL dt)
{
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)count; ++c)
        N[c] = SFX(clip0)(N[c] + dt * G[c]);
}

/* Models/time_step.jl:98-105 (_combine_kernel!): dest = max(0, a*A + b*B). */
void SFX(orc_combine)(size_t count, REAL *dest, REAL a, const REAL *A, REAL b, const REAL *B)
{
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)count; ++c)
        dest[c] = SFX(clip0)(a * A[c] + b * B[c]);
}

/* Per-column reductions, Diagnostics/diagnostic_operations.jl:71-79 (sum),
 * :110-122 (ratio, fallback when den == 0), :153-165 (angle).  Loop order n
 * outer, m inner, as the reference. */
void SFX(orc_column_sums)(int nx, int ny, int nk, int nphi, const REAL *N,
                          const double *w, const double *mu_x, const double *mu_y, const double *trace,
                          double *m0, double *kappa_rms, double *mean_dir)
{
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)nx * ny; ++ij) {
        int i = (int)(ij % nx), j = (int)(ij / nx);
        double acc = 0, num = 0, den = 0, sx = 0, sy = 0;
        for (int n = 0; n < nphi; ++n)
            for (int m = 0; m < nk; ++m) {
                double v = (double)N[IDX4(i, j, m, n)];
                size_t t = (size_t)m + (size_t)nk * n;
                acc += v * w[t];
                num += v * trace[t];
                den += v * w[t];
                sx += v * mu_x[t];
                sy += v * mu_y[t];
            }
        m0[ij] = acc;
        kappa_rms[ij] = sqrt(den == 0 ? 0.0 : num / den);
        mean_dir[ij] = atan2(sy, sx);
    }
}

/* Coupling/pseudomomentum.jl:316-344 (_compute_pseudomomentum_doppler_velocity_kernel!):
 *   U[i,j,t] = sum_{n, s} O[t,s] * N[i,j,s,n] * kx[s,n]
 * (and Uy with ky; dU/dkappa with the derivative overlap dO when non-NULL),
 * accumulated n outer, s inner, each term formed as (O * N) * k.  O and dO
 * are Nk x Nk, target fastest (Julia overlap[target_m, source_m]); kx, ky
 * are the first-moment measures [n][m]. */
void SFX(orc_pm_doppler)(int nx, int ny, int nk, int nphi, const REAL *N, const REAL *O, const REAL *dO,
                         const REAL *kx, const REAL *ky, REAL *Ux, REAL *Uy, REAL *dUx, REAL *dUy)
{
#pragma omp parallel for schedule(static)
    for (long jt = 0; jt < (long)ny * nk; ++jt) {
        int j = (int)(jt % ny), t = (int)(jt / ny);
        for (int i = 0; i < nx; ++i) {
            REAL ax = 0, ay = 0, dax = 0, day = 0;
            for (int n = 0; n < nphi; ++n)
                for (int s = 0; s < nk; ++s) {
                    REAL a = N[IDX4(i, j, s, n)];
                    size_t st = (size_t)s + (size_t)nk * n;
                    size_t ot = g_pm_cols ? IDX3(i, j, 0) + (size_t)nx * ny * ((size_t)t + (size_t)nk * s)
                                          : (size_t)t + (size_t)nk * s;
                    REAL q = O[ot];
                    ax += q * a * kx[st];
                    ay += q * a * ky[st];
                    if (dO) {
                        REAL dq = dO[ot];
                        dax += dq * a * kx[st];
                        day += dq * a * ky[st];
                    }
                }
            Ux[IDX3(i, j, t)] = ax;
            Uy[IDX3(i, j, t)] = ay;
            if (dO) {
                dUx[IDX3(i, j, t)] = dax;
                dUy[IDX3(i, j, t)] = day;
            }
        }
    }
}

/* Coupling/pseudomomentum.jl:375-400 (_compute_pseudomomentum_cells_kernel!):
 *   p[i,j,k] = scale_k * sum_{n, m} W[k,m] * N[i,j,m,n] * kx[m,n]
 * with W the Q cell weights (nz x Nk, k fastest), n outer, m inner, and
 * scale_k = 1 / |z_{k+1} - z_k| for cell averages, 1 for cell integrals. */
void SFX(orc_pm_cells)(int nx, int ny, int nk, int nphi, int nz, const REAL *N, const REAL *W, const double *zf,
                       const REAL *kx, const REAL *ky, int cell_average, REAL *px, REAL *py)
{
#pragma omp parallel for schedule(static)
    for (long jk = 0; jk < (long)ny * nz; ++jk) {
        int j = (int)(jk % ny), k = (int)(jk / ny);
        REAL scale = cell_average ? (REAL)1 / (REAL)fabs(zf[k + 1] - zf[k]) : (REAL)1;
        for (int i = 0; i < nx; ++i) {
            REAL ax = 0, ay = 0;
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m) {
                    REAL q = g_pm_cols ? W[IDX3(i, j, 0) + (size_t)nx * ny * ((size_t)k + (size_t)nz * m)]
                                       : W[(size_t)k + (size_t)nz * m];
                    REAL a = N[IDX4(i, j, m, n)];
                    size_t mn = (size_t)m + (size_t)nk * n;
                    ax += q * a * kx[mn];
                    ay += q * a * ky[mn];
                }
            size_t o = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
            px[o] = ax * scale;
            py[o] = ay * scale;
        }
    }
}

/* Models/time_step.jl:119-170 (time_step!) for :RK3, :LSRK3 and
 * :ForwardEuler, driving the tendency of Models/tendencies.jl:18-109:
 *   mode 0: no transport and no sources (G = 0, tendencies.jl:60-62)
 *   mode 1: fused intrinsic transport (tendencies.jl:28-41)
 *   mode 2: fused Doppler transport + refraction (tendencies.jl:82-101)
 *   mode 3: mode 2 with the self-coupled pseudomomentum current
 *           (tendencies.jl:103-107): before every tendency evaluation U is
 *           recomputed from the stage input (current_coupling.jl:140-143)
 *           through the overlap pm_O and the measures pm_kx, pm_ky, into
 *           Ux, Uy, which must then be writable.
 * stepper 0 = RK3, 1 = LSRK3 (identical arithmetic, time_step.jl:107-115),
 * 2 = ForwardEuler.  In mode 2 the current gradients are formed once: the
 * reference recomputes them every call from the same static Ux, Uy, giving
 * the same values; in mode 3 they follow U every evaluation.  extra_tendency,
 * when non-NULL, is called after the transport tendency with (N, G, user) and
 * adds source terms (tendencies.jl:88-94). */
typedef void (*SFX(orc_tendency_hook))(const REAL *N, REAL *G, void *user);

static void SFX(transport_tendency)(int nx, int ny, int nk, int nphi, int mode, const REAL *N, REAL *G,
                                    const REAL *Ux, const REAL *Uy, REAL *gr, int fresh_gradients,
                                    const REAL *pm_O, const REAL *pm_kx, const REAL *pm_ky,
                                    const REAL *kappa_c, const REAL *cgx, const REAL *cgy,
                                    const REAL *cos_t, const REAL *sin_t,
                                    REAL dx_inv, REAL dy_inv, REAL dk_inv, REAL dp_inv)
{
    size_t cnt = (size_t)nx * ny * nk * nphi, c3 = (size_t)nx * ny * nk;
    if (mode == 3)
        SFX(orc_pm_doppler)(nx, ny, nk, nphi, N, pm_O, NULL, pm_kx, pm_ky, (REAL *)Ux, (REAL *)Uy, NULL, NULL);
    if (mode >= 2) {
        if (fresh_gradients || mode == 3)
            SFX(orc_current_gradients)(nx, ny, nk, Ux, Uy, dx_inv, dy_inv, gr, gr + c3, gr + 2 * c3, gr + 3 * c3);
        SFX(orc_refraction_tendency)(nx, ny, nk, nphi, N, G, Ux, Uy, gr, gr + c3, gr + 2 * c3, gr + 3 * c3,
                                     kappa_c, cgx, cgy, cos_t, sin_t, dx_inv, dy_inv, dk_inv, dp_inv);
    } else if (mode == 1) {
        SFX(orc_intrinsic_tendency)(nx, ny, nk, nphi, N, G, cgx, cgy, dx_inv, dy_inv);
    } else {
        memset(G, 0, cnt * sizeof(REAL));
    }
}

int SFX(orc_run)(int nx, int ny, int nk, int nphi, int mode, int stepper, REAL *N,
                 REAL *Ux, REAL *Uy, const REAL *kappa_c,
                 const REAL *cgx, const REAL *cgy, const REAL *cos_t, const REAL *sin_t,
                 REAL dx_inv, REAL dy_inv, REAL dk_inv, REAL dp_inv, double dt, int nsteps,
                 SFX(orc_tendency_hook) extra_tendency, void *user,
                 const REAL *pm_O, const REAL *pm_kx, const REAL *pm_ky)
{
    size_t cnt = (size_t)nx * ny * nk * nphi, c3 = (size_t)nx * ny * nk;
    REAL *G = (REAL *)malloc(cnt * sizeof(REAL));
    REAL *N0 = (REAL *)malloc(cnt * sizeof(REAL));
    REAL *S = (REAL *)malloc(cnt * sizeof(REAL));
    REAL *gr = mode >= 2 ? (REAL *)malloc(4 * c3 * sizeof(REAL)) : NULL;
    if (!G || !N0 || !S || (mode >= 2 && !gr)) {
        free(G); free(N0); free(S); free(gr);
        return -1;
    }
    if (mode == 2)
        SFX(orc_current_gradients)(nx, ny, nk, Ux, Uy, dx_inv, dy_inv, gr, gr + c3, gr + 2 * c3, gr + 3 * c3);
    const REAL dtf = (REAL)dt;
    for (int step = 0; step < nsteps; ++step) {
        int nstage = stepper == 2 ? 1 : 3;
        if (nstage == 3)
            memcpy(N0, N, cnt * sizeof(REAL));
        for (int stage = 0; stage < nstage; ++stage) {
            SFX(transport_tendency)(nx, ny, nk, nphi, mode, N, G, Ux, Uy, gr, 0, pm_O, pm_kx, pm_ky, kappa_c, cgx,
                                    cgy, cos_t, sin_t, dx_inv, dy_inv, dk_inv, dp_inv);
            if (extra_tendency)
                extra_tendency(N, G, user);
            if (stage == 0) {
                SFX(orc_add_scaled)(cnt, N, G, dtf);
            } else {
                memcpy(S, N, cnt * sizeof(REAL));
                SFX(orc_add_scaled)(cnt, S, G, dtf);
                if (stage == 1)
                    SFX(orc_combine)(cnt, N, (REAL)0.75, N0, (REAL)0.25, S);
                else
                    SFX(orc_combine)(cnt, N, (REAL)(1.0 / 3.0), N0, (REAL)(2.0 / 3.0), S);
            }
        }
    }
    free(G); free(N0); free(S); free(gr);
    return 0;
}

/* One tendency evaluation (the parity hook); in mode 3 Ux, Uy receive the
 * coupling's velocity of N. */
void SFX(orc_tendency)(int nx, int ny, int nk, int nphi, int mode, const REAL *N, REAL *G,
                       REAL *Ux, REAL *Uy, const REAL *kappa_c,
                       const REAL *cgx, const REAL *cgy, const REAL *cos_t, const REAL *sin_t,
                       REAL dx_inv, REAL dy_inv, REAL dk_inv, REAL dp_inv,
                       const REAL *pm_O, const REAL *pm_kx, const REAL *pm_ky)
{
    size_t c3 = (size_t)nx * ny * nk;
    REAL *gr = NULL;
    if (mode >= 2 && !(gr = (REAL *)malloc(4 * c3 * sizeof(REAL))))
        return;
    SFX(transport_tendency)(nx, ny, nk, nphi, mode, N, G, Ux, Uy, gr, 1, pm_O, pm_kx, pm_ky, kappa_c, cgx, cgy,
                            cos_t, sin_t, dx_inv, dy_inv, dk_inv, dp_inv);
    free(gr);
}

/* Models/propagation_smoothing.jl:117-157 (ensure_spatial_averaging_tables!):
 * deep-water group speed per kappa bin, cg[m] = 1/2 sqrt(g / kappa[m]) with
 * g = 9.81 in the element type, and half the difference of its neighbours,
 * dcg[m] = |cg[m-1] - cg[m+1]| / 2 (the bin itself at the two ends). */
void SFX(orc_gse_tables)(int nk, const REAL *kappa, REAL *cg, REAL *dcg)
{
    const REAL g = (REAL)9.81, half = (REAL)1 / (REAL)2;
    for (int m = 0; m < nk; ++m)
        cg[m] = half * (REAL)sqrt(g / kappa[m]);
    for (int m = 0; m < nk; ++m) {
        REAL lo = m == 0 ? cg[0] : half * (REAL)sqrt(g / kappa[m - 1]);
        REAL hi = m == nk - 1 ? cg[nk - 1] : half * (REAL)sqrt(g / kappa[m + 1]);
        REAL d = lo - hi;
        dcg[m] = (d < 0 ? -d : d) / (REAL)2;
    }
}

/* Wrap of a 1-based index onto 1..n: one period either way, as the
 * reference's _gse_periodic (propagation_smoothing.jl:53), then further
 * periods for displacements the reference would read out of range. */
static inline int SFX(gse_wrap)(int i, int n)
{
    i = i < 1 ? i + n : (i > n ? i - n : i);
    if (i < 1 || i > n)
        i = ((i - 1) % n + n) % n + 1;
    return i;
}

/* The reference's _gse_clamp (propagation_smoothing.jl:56): a 1-based index
 * held to 1..n on a Bounded axis. */
static inline int SFX(gse_clamp)(int i, int n)
{
    return i < 1 ? 1 : (i > n ? n : i);
}

/* Models/propagation_smoothing.jl:56-110 (_spatial_averaging_kernel!): for
 * every cell the mean of four bilinear interpolations at the corners
 * (s1 Ls e_s + s2 Ln e_n), s1, s2 = -1, +1 (s1 outer), of a rectangle with
 * half-widths Ls = as dcg dt along the bin's direction and Ln = an cg dphi dt
 * across it, in cell units.  Indices are formed 1-based as in the reference
 * so the fractional weights round the same way; each corner index wraps on
 * a periodic axis (px, py nonzero) and is clamped to the edge cell on a
 * bounded one (propagation_smoothing.jl:97-100). */
void SFX(orc_spatial_averaging)(int nx, int ny, int nk, int nphi, const REAL *Nin, REAL *Nout,
                                const REAL *cg, const REAL *dcg, const REAL *cos_t, const REAL *sin_t,
                                REAL dphi, REAL as, REAL an, REAL dt, REAL dx_inv, REAL dy_inv, int px, int py)
{
#pragma omp parallel for schedule(static)
    for (long mn = 0; mn < (long)nk * nphi; ++mn) {
        int m = (int)(mn % nk), n = (int)(mn / nk);
        REAL Ls = as * dcg[m] * dt;
        REAL Ln = an * cg[m] * dphi * dt;
        REAL c = cos_t[n], s = sin_t[n];
        REAL Ls_x = Ls * c * dx_inv, Ls_y = Ls * s * dy_inv;
        REAL Ln_x = -Ln * s * dx_inv, Ln_y = Ln * c * dy_inv;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                REAL acc = 0;
                for (int s1 = -1; s1 <= 1; s1 += 2)
                    for (int s2 = -1; s2 <= 1; s2 += 2) {
                        REAL di = (REAL)s1 * Ls_x + (REAL)s2 * Ln_x;
                        REAL dj = (REAL)s1 * Ls_y + (REAL)s2 * Ln_y;
                        REAL ir = (REAL)(i + 1) + di, jr = (REAL)(j + 1) + dj;
                        int i0 = (int)floor(ir), j0 = (int)floor(jr);
                        REAL wi = ir - (REAL)i0, wj = jr - (REAL)j0;
                        int i0c = (px ? SFX(gse_wrap)(i0, nx) : SFX(gse_clamp)(i0, nx)) - 1;
                        int i1c = (px ? SFX(gse_wrap)(i0 + 1, nx) : SFX(gse_clamp)(i0 + 1, nx)) - 1;
                        int j0c = (py ? SFX(gse_wrap)(j0, ny) : SFX(gse_clamp)(j0, ny)) - 1;
                        int j1c = (py ? SFX(gse_wrap)(j0 + 1, ny) : SFX(gse_clamp)(j0 + 1, ny)) - 1;
                        REAL f00 = Nin[IDX4(i0c, j0c, m, n)], f10 = Nin[IDX4(i1c, j0c, m, n)];
                        REAL f01 = Nin[IDX4(i0c, j1c, m, n)], f11 = Nin[IDX4(i1c, j1c, m, n)];
                        REAL one = 1;
                        acc += (one - wi) * (one - wj) * f00 + wi * (one - wj) * f10 + (one - wi) * wj * f01 +
                               wi * wj * f11;
                    }
                Nout[IDX4(i, j, m, n)] = acc / (REAL)4;
            }
    }
}

/* Exposed for known-answer tests of the reconstruction itself. */
REAL SFX(orc_weno5)(REAL fm2, REAL fm1, REAL f0, REAL fp1, REAL fp2)
{
    return SFX(weno5_reconstruct)(fm2, fm1, f0, fp1, fp2);
}
