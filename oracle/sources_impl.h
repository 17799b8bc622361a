/*
 * ripple oracle -- source terms and winds (FP64).  TEST INFRASTRUCTURE ONLY.
 *
 * Restates the PrecomputedSources bundle path of the reference
 * (Physics/Packages/precomputed_sources.jl:23-283) for its members
 *   wind input  PressureCorrelationInput  (Physics/WindInput/pressure_correlation.jl:24-156)
 *               with BulkWindDrag          (Physics/shared/drag.jl:21-34)
 *   dissipation MeanSpectrumWhitecapping   (Physics/Dissipation/mean_spectrum.jl:20-140)
 *               LocalSaturationDissipation (Physics/Dissipation/local_saturation.jl:18-52)
 *               WhitecappingDissipation    (Sources/Sources.jl:105-122, 583-598; constant rate)
 * and the winds UniformWind (a number + direction) and HollandHurricaneWind
 * on a two-point LinearStormTrack (Forcing/winds.jl:4-41, 89-198).
 *
 * The source tendency G += positive - damping * N is evaluated on the stage
 * input state with the wind at the step's clock time
 * (Models/tendencies.jl:88-94; the clock advances once per step).
 */

typedef struct orc_sources {
    int wind_input;   /* 0 none, 1 PressureCorrelationInput */
    int drag;         /* 0 BulkWindDrag(:linear), 1 BulkWindDrag(:capped) */
    int wind_field;   /* 0 uniform (speed, direction), 1 Holland hurricane */
    int stress_cap;   /* 1: bundle stress cap (precomputed_sources.jl:126-170) */
    int dissipation;  /* 0 none, 1 MeanSpectrumWhitecapping, 2 LocalSaturationDissipation, 3 WhitecappingDissipation */
    int rotation;     /* +1 counter-clockwise, -1 clockwise */
    double wind_speed, wind_direction;
    double beta_max, z_alpha, p_in, alpha0, rho_air, rho_water, von_karman, gravity, cap_epsilon;
    double vmax, rmax, radius, shape, inflow_angle, background_x, background_y;
    double track_t0, track_t1, track_x0, track_y0, track_x1, track_y1;
    double cds, delta1, delta2, p_mean;
    double ls_cds, ls_br, ls_p, ls_sigma_power;
    double wc_rate, wc_threshold, wc_saturation_power, wc_wavenumber_power, wc_reference_wavenumber;
    int track_n;  /* > 0: LinearStormTrack of track_n points, replacing the two-point track */
    const double *track_t, *track_x, *track_y;
    int nonlinear;  /* 0 none, 1 DiscreteInteractionApproximation(SymmetricQuadruplet) */
    double dia_coefficient, dia_frequency_offset, dia_gravity;
} orc_sources;

/* drag.jl:25-34. */
static double orc_friction_velocity(const orc_sources *s, double U10)
{
    double cd;
    if (s->drag == 0) {
        cd = (0.8 + 0.065 * U10) * 1e-3;
    } else {
        cd = (8.058 + 0.967 * U10 - 0.016 * (U10 * U10)) * 1e-4;
        cd = cd > 4e-4 ? cd : 4e-4;
    }
    return U10 * sqrt(cd > 0 ? cd : 0.0);
}

/* winds.jl:28-41 (searchsortedfirst for the upper point). */
void orc_track_at(int n, const double *T, const double *X, const double *Y, double t, double *cx, double *cy)
{
    if (t <= T[0]) {
        *cx = X[0];
        *cy = Y[0];
        return;
    }
    if (t >= T[n - 1]) {
        *cx = X[n - 1];
        *cy = Y[n - 1];
        return;
    }
    int upper = 0;
    while (T[upper] < t)
        ++upper;
    const int lower = upper - 1;
    const double w = (t - T[lower]) / (T[upper] - T[lower]);
    *cx = (1 - w) * X[lower] + w * X[upper];
    *cy = (1 - w) * Y[lower] + w * Y[upper];
}

static void orc_track(const orc_sources *s, double t, double *cx, double *cy)
{
    if (s->track_n > 0) {
        orc_track_at(s->track_n, s->track_t, s->track_x, s->track_y, t, cx, cy);
        return;
    }
    if (t <= s->track_t0) {
        *cx = s->track_x0;
        *cy = s->track_y0;
    } else if (t >= s->track_t1) {
        *cx = s->track_x1;
        *cy = s->track_y1;
    } else {
        double w = (t - s->track_t0) / (s->track_t1 - s->track_t0);
        *cx = (1 - w) * s->track_x0 + w * s->track_x1;
        *cy = (1 - w) * s->track_y0 + w * s->track_y1;
    }
}

/* winds.jl:144-154 (hurricane_speed) and 184-204 (wind_velocity, speed, angle). */
void orc_wind_at(const orc_sources *s, double x, double y, double t, double *speed, double *angle)
{
    if (s->wind_field == 0) {
        *speed = s->wind_speed;
        *angle = s->wind_direction;
        return;
    }
    double cx, cy;
    orc_track(s, t, &cx, &cy);
    double dx = x - cx, dy = y - cy;
    double r = hypot(dx, dy);
    double vx, vy;
    if (r == 0) {
        vx = s->background_x;
        vy = s->background_y;
    } else {
        double sp = 0;
        double scaled = s->rmax / r;
        double shape_term = pow(scaled, s->shape);
        double v = shape_term * exp(1 - shape_term);
        sp = s->vmax * sqrt(v > 0 ? v : 0.0);
        if (r > s->radius) {
            double eps_r = nextafter(s->radius, INFINITY) - s->radius;
            double decay = s->radius - s->rmax;
            decay = decay > eps_r ? decay : eps_r;
            sp *= exp(-(r - s->radius) / decay);
        }
        double sg = (double)s->rotation;
        double tx = -sg * dy / r, ty = sg * dx / r;
        double rx = dx / r, ry = dy / r;
        vx = sp * (cos(s->inflow_angle) * tx - sin(s->inflow_angle) * rx) + s->background_x;
        vy = sp * (cos(s->inflow_angle) * ty - sin(s->inflow_angle) * ry) + s->background_y;
    }
    *speed = hypot(vx, vy);
    *angle = atan2(vy, vx);
}

/* precomputed_sources.jl:113-124 (_jp_rate) and pressure_correlation.jl:109-136:
 * raw Janssen growth rate, zero outside the growth window. */
static double orc_jp_rate(const orc_sources *s, double k, double sigma, double C, double theta, double theta_u,
                          double u_star)
{
    double cos_tu = cos(theta - theta_u);
    if (!(cos_tu > 0))
        return 0.0;
    double z1 = s->alpha0 * (u_star * u_star) / s->gravity;
    if (!(z1 > 0))
        return 0.0;
    double inv_age = u_star / C + s->z_alpha;
    double Z = log(k * z1) + s->von_karman / (inv_age * cos_tu);
    if (Z >= 0)
        return 0.0;
    double cp = s->p_in == 2.0 ? cos_tu * cos_tu : pow(cos_tu, s->p_in);
    double Z2 = Z * Z;
    return (s->rho_air / s->rho_water) * (s->beta_max / (s->von_karman * s->von_karman)) * exp(Z) * (Z2 * Z2) *
           (inv_age * inv_age) * cp * sigma;
}

/* ------------------------------------------------------------------------ */
/* DiscreteInteractionApproximation{SymmetricQuadruplet}                     */
/* (Physics/NonlinearInteractions/symmetric_quadruplet.jl:30-238).            */
/* ------------------------------------------------------------------------ */

/* symmetric_quadruplet.jl:38-49: the direction offsets of the two receivers
 * from momentum conservation at lambda. */
void orc_dia_offsets(double lambda, double *dtheta_plus, double *dtheta_minus)
{
    double a = (1 + lambda) * (1 + lambda), b = (1 - lambda) * (1 - lambda);
    double cp = (a * a - b * b + 4) / (4 * a), cm = (b * b - a * a + 4) / (4 * b);
    cp = cp < -1 ? -1 : (cp > 1 ? 1 : cp);
    cm = cm < -1 ? -1 : (cm > 1 ? 1 : cm);
    *dtheta_plus = acos(cp);
    *dtheta_minus = acos(cm);
}

/* symmetric_quadruplet.jl:54-77 (_bilinear_lookup), 0-based bins out. */
static int orc_dia_lookup(double s, double th, double smin, double smax, double log_xfr, double phimin, double dphi,
                          int nk, int nphi, int *mlo, int *mhi, int *nlo, int *nhi, double w[4])
{
    if (s < smin || s > smax)
        return 0;
    double mr = 1.0 + log(s / smin) / log_xfr;
    long t = (long)mr;
    int ml = (int)(t < 1 ? 1 : (t > nk - 1 ? nk - 1 : t));
    double wf = mr - ml;
    double twopi = 2 * M_PI;
    double r = fmod(th - phimin, twopi);
    if (r != 0 && (r < 0) != (twopi < 0))
        r += twopi;
    double nr = r / dphi;
    long nraw = (long)nr;
    long nm = nraw % nphi;
    if (nm < 0)
        nm += nphi;
    int nl = (int)nm + 1;
    double wt = nr - (double)nraw;
    int nh = nl == nphi ? 1 : nl + 1;
    w[0] = (1 - wf) * (1 - wt);
    w[1] = wf * (1 - wt);
    w[2] = (1 - wf) * wt;
    w[3] = wf * wt;
    *mlo = ml - 1;
    *mhi = ml;
    *nlo = nl - 1;
    *nhi = nh - 1;
    return 1;
}

static inline double orc_dia_energy_factor(double sigma, double g) { return 4 * M_PI * (sigma * sigma) / g; }

/* symmetric_quadruplet.jl:82-208: the transfer T[i,j,m,n] of every column,
 * donors visited n outer, m inner, orientation +1 then -1 (the order in
 * which a sequential backend runs the kernel's atomic updates).  f_a^11 and
 * the constants DAL1..3, C/g^4 use libm pow (Julia's integer powers may
 * differ in the last bit). */
void orc_dia_transfer(int nx, int ny, int nk, int nphi, const double *N, const double *kappa, const double *phi,
                      double coefficient, double lambda, double g, double *T)
{
    double dtp, dtm;
    orc_dia_offsets(lambda, &dtp, &dtm);
    const double DAL1 = 1 / pow(1 + lambda, 4), DAL2 = 1 / pow(1 - lambda, 4), DAL3 = 2 * DAL1 * DAL2;
    const double Ceff = coefficient / pow(g, 4);
    const double smin = sqrt(g * kappa[0]), smax = sqrt(g * kappa[nk - 1]);
    const double log_xfr = log(sqrt(g * kappa[1]) / sqrt(g * kappa[0]));
    const double phimin = phi[0], dphi = 2 * M_PI / nphi;
    const size_t plane = (size_t)nx * ny;
    memset(T, 0, plane * nk * nphi * sizeof(double));
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)plane; ++ij) {
        const double *Nc = N + ij;
        double *Tc = T + ij;
#define ORC_N(m, n) Nc[plane * ((size_t)(m) + (size_t)nk * (n))]
#define ORC_T(m, n) Tc[plane * ((size_t)(m) + (size_t)nk * (n))]
        for (int na = 0; na < nphi; ++na)
            for (int ma = 0; ma < nk; ++ma) {
                double Na = ORC_N(ma, na);
                if (!(Na > 0))
                    continue;
                double sa_ = sqrt(g * kappa[ma]);
                double fa = sa_ / (2 * M_PI);
                double tha = phi[na];
                double Efa = orc_dia_energy_factor(sa_, g);
                double Ea = Efa * Na;
                double sp = (1 + lambda) * sa_, sm = (1 - lambda) * sa_;
                for (int o = 1; o >= -1; o -= 2) {
                    double thp = tha + o * dtp, thm = tha - o * dtm;
                    int mlp, mhp, nlp, nhp, mlm, mhm, nlm, nhm;
                    double wp[4], wm[4];
                    int inp = orc_dia_lookup(sp, thp, smin, smax, log_xfr, phimin, dphi, nk, nphi, &mlp, &mhp, &nlp,
                                             &nhp, wp);
                    int inm = orc_dia_lookup(sm, thm, smin, smax, log_xfr, phimin, dphi, nk, nphi, &mlm, &mhm, &nlm,
                                             &nhm, wm);
                    if (!inp && !inm)
                        continue;
                    double Ep = 0, Em = 0, Efp1 = 0, Efp2 = 0, Efm1 = 0, Efm2 = 0;
                    if (inp) {
                        Efp1 = orc_dia_energy_factor(sqrt(g * kappa[mlp]), g);
                        Efp2 = orc_dia_energy_factor(sqrt(g * kappa[mhp]), g);
                        double a0 = ORC_N(mlp, nlp), a1 = ORC_N(mhp, nlp), a2 = ORC_N(mlp, nhp), a3 = ORC_N(mhp, nhp);
                        Ep = wp[0] * Efp1 * (a0 > 0 ? a0 : 0) + wp[1] * Efp2 * (a1 > 0 ? a1 : 0) +
                             wp[2] * Efp1 * (a2 > 0 ? a2 : 0) + wp[3] * Efp2 * (a3 > 0 ? a3 : 0);
                    }
                    if (inm) {
                        Efm1 = orc_dia_energy_factor(sqrt(g * kappa[mlm]), g);
                        Efm2 = orc_dia_energy_factor(sqrt(g * kappa[mhm]), g);
                        double a0 = ORC_N(mlm, nlm), a1 = ORC_N(mhm, nlm), a2 = ORC_N(mlm, nhm), a3 = ORC_N(mhm, nhm);
                        Em = wm[0] * Efm1 * (a0 > 0 ? a0 : 0) + wm[1] * Efm2 * (a1 > 0 ? a1 : 0) +
                             wm[2] * Efm1 * (a2 > 0 ? a2 : 0) + wm[3] * Efm2 * (a3 > 0 ? a3 : 0);
                    }
                    double s = Ea * (Ep * DAL1 + Em * DAL2) - Ep * Em * DAL3;
                    double TE = Ceff * pow(fa, 11) * s / 2;
                    double donor = (inp && inm) ? 2.0 : 1.0;
                    ORC_T(ma, na) -= donor * TE / Efa;
                    if (inp) {
                        ORC_T(mlp, nlp) += wp[0] * TE / Efp1;
                        ORC_T(mhp, nlp) += wp[1] * TE / Efp2;
                        ORC_T(mlp, nhp) += wp[2] * TE / Efp1;
                        ORC_T(mhp, nhp) += wp[3] * TE / Efp2;
                    }
                    if (inm) {
                        ORC_T(mlm, nlm) += wm[0] * TE / Efm1;
                        ORC_T(mhm, nlm) += wm[1] * TE / Efm2;
                        ORC_T(mlm, nhm) += wm[2] * TE / Efm1;
                        ORC_T(mhm, nhm) += wm[3] * TE / Efm2;
                    }
                }
            }
#undef ORC_N
#undef ORC_T
    }
}

/* symmetric_quadruplet.jl:210-222 (source_split): T >= 0 is a positive
 * source, T < 0 a damping rate min(-T / max(N, eps), 1). */
static inline void orc_dia_split(double T, double Na, double *positive, double *damping)
{
    if (T >= 0) {
        *positive += T;
    } else {
        double d = -T / (Na > DBL_EPSILON ? Na : DBL_EPSILON);
        *damping += d < 1 ? d : 1;
    }
}

typedef struct orc_source_ctx {
    const orc_sources *src;
    int nx, ny, nk, nphi;
    const double *x_nodes, *y_nodes;  /* nx, ny */
    const double *kappa, *phi;        /* nk, nphi centres */
    const double *w;                  /* (nk, nphi) cell measures */
    double time;                      /* clock time of the step */
    const double *kappa_faces, *phi_faces;  /* nk + 1, nphi + 1 */
} orc_source_ctx;

/* radial_power_source_factor (Sources.jl:569-574): the cell average of
 * kappa^p, radial_power_annular_measure / cell measure
 * (finite_volume_integration.jl:24-33, 71-73), over kappa_ref^p. */
static double orc_radial_power_factor(const orc_source_ctx *c, int m, int n, double p, double kref)
{
    if (p == 0)
        return 1.0;
    const double k1 = c->kappa_faces[m], k2 = c->kappa_faces[m + 1];
    const double measure = (c->phi_faces[n + 1] - c->phi_faces[n]) * (pow(k2, p + 2) - pow(k1, p + 2)) / (p + 2);
    return measure / c->w[m + (size_t)c->nk * n] / pow(kref, p);
}

/* Adds positive - damping * N for every cell (tendencies.jl:88-94), with the
 * bundle state of precomputed_sources.jl:41-101, 126-218 formed first from
 * the same N. */
void orc_sources_hook_f64(const double *N, double *G, void *user)
{
    const orc_source_ctx *c = (const orc_source_ctx *)user;
    const orc_sources *s = c->src;
    const int nx = c->nx, ny = c->ny, nk = c->nk, nphi = c->nphi;
    const double g = s->gravity;
    /* The nonlinear state of the bundle (precomputed_sources.jl:227-233). */
    double *Tdia = NULL;
    if (s->nonlinear == 1) {
        Tdia = (double *)malloc((size_t)nx * ny * nk * nphi * sizeof(double));
        if (!Tdia)
            return;
        orc_dia_transfer(nx, ny, nk, nphi, N, c->kappa, c->phi, s->dia_coefficient, s->dia_frequency_offset,
                         s->dia_gravity, Tdia);
    }
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)nx * ny; ++ij) {
        int i = (int)(ij % nx), j = (int)(ij / nx);
        double U10 = 0, theta_u = 0, u_star = 0, factor = 1.0;
        if (s->wind_input) {
            orc_wind_at(s, c->x_nodes[i], c->y_nodes[j], c->time, &U10, &theta_u);
            u_star = orc_friction_velocity(s, U10);
        }
        /* Stress cap (precomputed_sources.jl:126-170). */
        if (s->wind_input && s->stress_cap) {
            if (U10 > 0 && u_star > 0) {
                double tau_max = s->rho_air * (u_star * u_star);
                double twx = 0, twy = 0;
                double rho_w_g = s->rho_water * s->gravity;
                for (int n = 0; n < nphi; ++n)
                    for (int m = 0; m < nk; ++m) {
                        double a = N[ij + (size_t)nx * ny * ((size_t)m + (size_t)nk * n)];
                        if (!(a > 0))
                            continue;
                        double k = c->kappa[m];
                        if (!(k > 0))
                            continue;
                        double sigma = sqrt(g * k), C = sigma / k, th = c->phi[n];
                        double rate = orc_jp_rate(s, k, sigma, C, th, theta_u, u_star);
                        if (rate > 0) {
                            double mflux = rate * a / C * c->w[m + (size_t)nk * n] * rho_w_g;
                            twx += mflux * cos(th);
                            twy += mflux * sin(th);
                        }
                    }
                double par = twx * cos(theta_u) + twy * sin(theta_u);
                double eff = par > 0 ? par : 0.0;
                factor = (eff > s->cap_epsilon * tau_max && tau_max > 0) ? s->cap_epsilon * tau_max / eff : 1.0;
            }
        }
        /* Mean-spectrum moments (precomputed_sources.jl:41-74). */
        double m0 = 0, kbar = 0, sbar = 0;
        if (s->dissipation == 1) {
            double kp = 0, sp = 0;
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m) {
                    double a = N[ij + (size_t)nx * ny * ((size_t)m + (size_t)nk * n)];
                    if (!(a > 0))
                        continue;
                    double wt = c->w[m + (size_t)nk * n], k = c->kappa[m];
                    if (!(k > 0))
                        continue;
                    double sigma = sqrt(g * k);
                    m0 += a * wt;
                    kp += a * wt * pow(k, s->p_mean);
                    sp += a * wt * pow(sigma, s->p_mean);
                }
            if (m0 > 0) {
                kbar = pow(kp / m0, 1 / s->p_mean);
                sbar = pow(sp / m0, 1 / s->p_mean);
            }
        }
        /* local_zeroth_moment (Sources.jl:1471-1480) for WhitecappingDissipation. */
        double m0w = 0;
        if (s->dissipation == 3)
            for (int n = 0; n < nphi; ++n)
                for (int m = 0; m < nk; ++m)
                    m0w += N[ij + (size_t)nx * ny * ((size_t)m + (size_t)nk * n)] * c->w[m + (size_t)nk * n];
        /* Per cell (source_split of each member, bundle sum). */
        for (int n = 0; n < nphi; ++n)
            for (int m = 0; m < nk; ++m) {
                size_t cidx = ij + (size_t)nx * ny * ((size_t)m + (size_t)nk * n);
                double a = N[cidx];
                double kx = c->kappa[m] * cos(c->phi[n]), ky = c->kappa[m] * sin(c->phi[n]);
                double k = hypot(kx, ky);
                double positive = 0, damping = 0;
                if (s->wind_input && U10 > 0 && u_star > 0 && k > 0) {
                    double sigma = sqrt(s->gravity * k), C = sigma / k, th = atan2(ky, kx);
                    double raw = orc_jp_rate(s, k, sigma, C, th, theta_u, u_star);
                    if (raw > 0) {
                        double rate = s->stress_cap ? raw * factor : raw;
                        /* split_growth_rate (Sources.jl:430-431). */
                        if (rate >= 0)
                            positive += rate * a;
                        else
                            damping += -rate;
                    }
                }
                if (s->dissipation == 1 && m0 > 0 && kbar > 0 && k > 0) {
                    double alpha_bar = m0 * (kbar * kbar);
                    double ratio = k / kbar;
                    double bracket = s->delta1 * ratio + s->delta2 * (ratio * ratio);
                    double rate = -s->cds * (alpha_bar * alpha_bar) * sbar * bracket;
                    /* split_damping_rate (Sources.jl:433-434). */
                    if (rate >= 0)
                        damping += rate;
                    else
                        positive += -rate * a;
                } else if (s->dissipation == 2 && a > 0 && k > 0) {
                    double sigma = sqrt(s->gravity * k);
                    double B = a * sigma * (k * k * k);
                    double over = B / s->ls_br - 1;
                    over = over > 0 ? over : 0.0;
                    if (over > 0) {
                        over = over < 1e6 ? over : 1e6;
                        double rate = -s->ls_cds * pow(sigma, s->ls_sigma_power) * pow(over, s->ls_p);
                        if (rate >= 0)
                            damping += rate;
                        else
                            positive += -rate * a;
                    }
                } else if (s->dissipation == 3) {
                    /* Sources.jl:588-597. */
                    double excess = m0w / s->wc_threshold - 1;
                    excess = excess > 0 ? excess : 0.0;
                    double rate = s->wc_rate * pow(excess, s->wc_saturation_power) *
                                  orc_radial_power_factor(c, m, n, s->wc_wavenumber_power, s->wc_reference_wavenumber);
                    if (rate >= 0)
                        damping += rate;
                    else
                        positive += -rate * a;
                }
                if (Tdia)  /* precomputed_sources.jl:256-259 */
                    orc_dia_split(Tdia[cidx], a, &positive, &damping);
                G[cidx] += positive - damping * a;
            }
    }
    free(Tdia);
}

/* Per-column wind fields at time t, for tests of the device wind kernel. */
void orc_wind_columns(const orc_sources *s, int nx, int ny, const double *x_nodes, const double *y_nodes, double t,
                      double *U10, double *theta_u, double *u_star)
{
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t o = (size_t)i + (size_t)nx * j;
            orc_wind_at(s, x_nodes[i], y_nodes[j], t, &U10[o], &theta_u[o]);
            u_star[o] = orc_friction_velocity(s, U10[o]);
        }
}

/* DiscreteInteractionApproximation{ListedQuadruplets} (Sources/Sources.jl:
 * 374-428, 1332-1385): for every cell and every listed interaction (donors
 * d1, d2, receivers r1, r2; 0-based (m, n) bins),
 *   flux = rate max(N_d1, 0)^p max(N_d2, 0)^p,
 *   receivers: positive += flux / w,  donors: damping += flux / (w N) (N > 0),
 * in list order, then G += positive - damping N.  The rate is the
 * interaction's constant, or with rate_field set its column's value
 * (source_value of an array or a function of position, Sources.jl:
 * 1450-1467), rate_field[col + nx ny q]. */
typedef struct orc_listed_ctx {
    int nx, ny, nk, nphi;
    int count;
    const int *bins;      /* 8 per interaction: d1m d1n d2m d2n r1m r1n r2m r2n */
    const double *rates;  /* per interaction */
    double power;
    const double *w;      /* (nk, nphi) cell measures */
    const double *rate_field;  /* NULL, or (nx ny, count), column fastest */
} orc_listed_ctx;

void orc_listed_hook_f64(const double *N, double *G, void *user)
{
    const orc_listed_ctx *c = (const orc_listed_ctx *)user;
    const size_t plane = (size_t)c->nx * c->ny;
    const int nk = c->nk;
#pragma omp parallel for schedule(static)
    for (long ij = 0; ij < (long)plane; ++ij)
        for (int n = 0; n < c->nphi; ++n)
            for (int m = 0; m < nk; ++m) {
                const size_t cell = ij + plane * ((size_t)m + (size_t)nk * n);
                const double weight = c->w[m + (size_t)nk * n];
                const double a = N[cell];
                double positive = 0, damping = 0;
                for (int q = 0; q < c->count; ++q) {
                    const int *b = c->bins + 8 * (size_t)q;
                    double a1 = N[ij + plane * ((size_t)b[0] + (size_t)nk * b[1])];
                    double a2 = N[ij + plane * ((size_t)b[2] + (size_t)nk * b[3])];
                    a1 = a1 > 0 ? a1 : 0;
                    a2 = a2 > 0 ? a2 : 0;
                    const double rate = c->rate_field ? c->rate_field[ij + plane * (size_t)q] : c->rates[q];
                    const double flux = rate * pow(a1, c->power) * pow(a2, c->power);
                    if (flux == 0)
                        continue;
                    if (m == b[4] && n == b[5])
                        positive += flux / weight;
                    if (m == b[6] && n == b[7])
                        positive += flux / weight;
                    if (m == b[0] && n == b[1] && a > 0)
                        damping += flux / (weight * a);
                    if (m == b[2] && n == b[3] && a > 0)
                        damping += flux / (weight * a);
                }
                G[cell] += positive - damping * a;
            }
}
