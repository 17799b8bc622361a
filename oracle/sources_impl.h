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
                G[cidx] += positive - damping * a;
            }
    }
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
