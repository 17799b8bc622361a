// Host side of the C ABI declared in include/ripple_mi355x.h: the context,
// geometry tables, stage-buffer rotation and dispatch onto the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mlp_internal.h"
#include "rpl_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define RPL_HIP(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(RPL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

}  // namespace

int rpl::set_error(int code, const char *msg) { return fail(code, msg); }

struct rpl_ctx {
    rpl_config cfg{};
    int device = 0;
    // Bumped by every call that changes what a step launches (geometry,
    // current, sources, storm track, nonlinear terms, smoothing, surrogate),
    // so an ensemble knows when its captured graph is stale.
    uint64_t generation = 0;
    hipStream_t stream = nullptr;
    size_t cnt = 0;    // interior cell-bins
    size_t plane = 0;  // nx * ny
    size_t esize = 0;  // sizeof element

    void *A = nullptr, *B = nullptr, *C = nullptr;  // action (A) and stage buffers
    void *Ux = nullptr, *Uy = nullptr, *Uxx = nullptr, *Uxy = nullptr, *Uyx = nullptr, *Uyy = nullptr;
    void *kappa = nullptr, *cgx = nullptr, *cgy = nullptr, *cosp = nullptr, *sinp = nullptr;
    double *w = nullptr, *trace = nullptr, *mux = nullptr, *muy = nullptr, *energy = nullptr;
    double *diag = nullptr;      // 3 * plane
    double *partials = nullptr;  // 2 * blocks
    void *scratch_in = nullptr;  // FP64 staging for host uploads

    double dx = 0, dy = 0, dk = 0, dp = 0;
    bool has_geometry = false, has_current = false;
    std::vector<double> h_mux, h_muy;  // first-moment measures (Nk, Nphi), kept for the pseudomomentum tables

    // Doppler velocity kappa derivatives (allocated by the first coupling that sets them)
    void *dUx = nullptr, *dUy = nullptr;
    bool has_dU = false;
    // self-coupled pseudomomentum current (rpl_set_pseudomomentum_coupling)
    bool pm = false;
    void *pm_O = nullptr, *pm_dO = nullptr, *pm_kx = nullptr, *pm_ky = nullptr;
    // DIA nonlinear member of the sources: column-independent tables and the
    // per-stage source plane the stage kernels add (StageParams::src_add)
    bool dia = false;
    void *dia_flags = nullptr, *dia_bins = nullptr, *dia_wef = nullptr, *dia_wdiv = nullptr, *dia_efa = nullptr,
         *dia_inv_efa = nullptr, *dia_cfa = nullptr, *dia_add = nullptr, *dia_gat_off = nullptr,
         *dia_gat_q = nullptr, *dia_gat_w = nullptr;
    double dia_dal[3] = {0, 0, 0};
    // ListedQuadruplets (rpl_set_quadruplets)
    bool listed = false;
    int32_t listed_n = 0;
    double listed_power = 1;
    void *listed_bins = nullptr, *listed_rates = nullptr, *listed_w = nullptr;
    // per-column rates (rpl_set_quadruplet_rate_field): (nx ny, listed_n)
    void *listed_field = nullptr;
    bool listed_has_field = false;
    // SpatialAveraging after every step (rpl_set_propagation_smoothing)
    bool gse = false;
    double gse_as = 0.5, gse_an = 0.5, gse_cg_max = 0, gse_dcg_max = 0;
    // the grid's x and y topology (rpl_set_topology): periodic or bounded
    bool periodic_x = true, periodic_y = true;
    // a depth per column (rpl_set_depth_field): the host values, the group
    // speed per (column, kappa bin) and the bins' direction tables (kx, ky, k
    // for the reference-order kernel; kx / k, ky / k for the phi walk)
    bool has_depth_field = false;
    bool pm_percol = false;      // the pseudomomentum overlap tables are per column
    size_t pm_O_elems = 0;       // the allocated size of pm_O and pm_dO
    double gravity = 9.81;  // rpl_set_geometry's
    std::vector<double> h_depth;
    void *cgm = nullptr, *cg_vec = nullptr, *cg_dir = nullptr;
    // non-finite check after every nf_interval-th step (rpl_set_nonfinite_check):
    // a device counter and the record {first iteration, count}
    int32_t nf_interval = 0;
    void *nf_buf = nullptr;
    void *gse_cg = nullptr, *gse_dcg = nullptr;
    double time = 0;
    int64_t iteration = 0;

    // source terms
    bool has_sources = false;
    rpl_sources src{};
    rpl::WindParams wind{};
    double *xn = nullptr, *yn = nullptr;                       // cell centres (FP64)
    void *U10 = nullptr, *theta_u = nullptr, *ustar = nullptr;  // per column
    void *factor = nullptr, *cm0 = nullptr, *kbar = nullptr, *sbar = nullptr;
    // FP32 stress cap fused into the stage kernels: per (column, kappa bin)
    // wave-supported stress sums of a stage's output, the per-bin weights,
    // and the buffer those sums belong to (null: none pending)
    void *tau_x = nullptr, *tau_y = nullptr, *tau_c = nullptr, *tau_s = nullptr;
    void *bin_tab = nullptr;  // FP32: the per-bin tables packed for the specialised source paths
    const void *tau_for = nullptr;
    void *k_cell = nullptr, *theta_cell = nullptr, *sigma_cell = nullptr, *C_cell = nullptr;  // per bin
    void *t_phi = nullptr, *t_w = nullptr, *t_sigma_k = nullptr, *t_cos = nullptr, *t_sin = nullptr;
    void *t_kpow = nullptr, *t_spow = nullptr, *t_wc = nullptr;
    // FP32 wind input and saturation factors (FP32 contexts only)
    void *f_cth = nullptr, *f_sth = nullptr, *f_logk = nullptr, *f_invc = nullptr, *f_sk3 = nullptr,
         *f_sq = nullptr;                                              // per bin
    void *f_logkap = nullptr, *f_invck = nullptr;                      // per kappa bin
    void *f_ctu = nullptr, *f_stu = nullptr, *f_lz1 = nullptr;         // per column
    std::vector<double> h_kc, h_pc, h_kf, h_pf;  // geometry kept for source tables
    std::vector<double> h_sigma;                 // per-bin intrinsic frequency of the source tables
    std::vector<double> track_t, track_x, track_y;  // storm track (rpl_set_storm_track), or empty

    // learned surrogate (rpl_attach_surrogate)
    rpl_mlp *net = nullptr;
    rpl_surrogate scfg{};
    bool s_active = false;  // the stage takes the network's source where allowed
    bool s_due = true;      // refresh 1: the next source preparation infers (set by each step's wind)
    rpl::mlp::bf16 *sX = nullptr, *sA = nullptr, *sB = nullptr;
    _Float16 *rate = nullptr;  // the network's rates, [out_pad][s_ld], FP16
    float *out_scale = nullptr, *s_err = nullptr, *bin_hi = nullptr, *s_amax = nullptr;
    unsigned char *s_flags = nullptr;  // per column: envelope flags of the current stage
    unsigned char *use_plain = nullptr;
    unsigned *s_counts = nullptr;
    size_t s_ld = 0, s_chunk = 0;
    int64_t s_columns = 0;

    // event timing: phase 0 = stage kernel launches, 1 = surrogate inference
    // (features and network forward of one stage input), 2 = pseudomomentum
    // coupling update (contraction and, on the reference path, gradients),
    // 3 = propagation smoothing, 4 = nonlinear (DIA) source, 5 = the source
    // bundle's per-column state of one stage input
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<int> event_phase;
    size_t events_used = 0;
    double timing_ms[6] = {0, 0, 0, 0, 0, 0};
    int64_t timing_launches[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

void free_ptr(void *p)
{
    if (p)
        (void)hipFree(p);
}

int alloc(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes == 0 ? 1 : bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(RPL_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return RPL_OK;
}

bool is_f64(const rpl_ctx *c) { return c->cfg.dtype == RPL_F64; }

// Uploads n doubles to a device array of the context's element type.
int upload_real(rpl_ctx *c, const double *host, void *dst, size_t n)
{
    if (is_f64(c)) {
        RPL_HIP(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<float> tmp(n);
        for (size_t k = 0; k < n; ++k)
            tmp[k] = (float)host[k];
        RPL_HIP(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        RPL_HIP(hipStreamSynchronize(c->stream));
    }
    return RPL_OK;
}

int upload_f64(rpl_ctx *c, const std::vector<double> &host, double *dst)
{
    RPL_HIP(hipMemcpyAsync(dst, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RPL_HIP(hipStreamSynchronize(c->stream));
    return RPL_OK;
}

bool phi_walk_kernel(const rpl_ctx *c);

template <typename T>
rpl::StageParams<T> stage_params(const rpl_ctx *c, double dt)
{
    rpl::StageParams<T> p{};
    p.nx = c->cfg.nx;
    p.ny = c->cfg.ny;
    p.nk = c->cfg.nkappa;
    p.nphi = c->cfg.nphi;
    p.Ux = (const T *)c->Ux;
    p.Uy = (const T *)c->Uy;
    p.Uxx = (const T *)c->Uxx;
    p.Uxy = (const T *)c->Uxy;
    p.Uyx = (const T *)c->Uyx;
    p.Uyy = (const T *)c->Uyy;
    p.kappa = (const T *)c->kappa;
    p.cgx = (const T *)c->cgx;
    p.cgy = (const T *)c->cgy;
    p.cosp = (const T *)c->cosp;
    if constexpr (std::is_same<T, float>::value)
        p.bt = (const float *)c->bin_tab;
    p.sinp = (const T *)c->sinp;
    if (c->has_depth_field) {
        const size_t tab = (size_t)c->cfg.nkappa * c->cfg.nphi;
        p.cgm = (const T *)c->cgm;
        if (phi_walk_kernel(c)) {
            p.cdx = (const T *)c->cg_dir;
            p.cdy = (const T *)c->cg_dir + tab;
        } else {
            p.cdx = (const T *)c->cg_vec;
            p.cdy = (const T *)c->cg_vec + tab;
            p.cdk = (const T *)c->cg_vec + 2 * tab;
        }
    }
    // 1 / first spacing, formed in the element type (wave_current_refraction.jl:213).
    p.dx_inv = T(1) / (T)c->dx;
    p.dy_inv = T(1) / (T)c->dy;
    p.dk_inv = T(1) / (T)c->dk;
    p.dp_inv = T(1) / (T)c->dp;
    p.dt = (T)dt;
    p.src_add = (c->dia || c->listed) ? (const T *)c->dia_add : nullptr;
    p.src = rpl::SourceParams<T>{};
    if (c->has_sources && (c->src.wind_input || c->src.dissipation)) {
        const rpl_sources &s = c->src;
        rpl::SourceParams<T> &q = p.src;
        q.on = 1;
        q.wind_input = s.wind_input;
        q.stress_cap = s.stress_cap;
        q.dissipation = s.dissipation;
        q.U10 = (const T *)c->U10;
        q.theta_u = (const T *)c->theta_u;
        q.ustar = (const T *)c->ustar;
        q.factor = (const T *)c->factor;
        q.m0 = (const T *)c->cm0;
        q.kbar = (const T *)c->kbar;
        q.sbar = (const T *)c->sbar;
        q.k_cell = (const T *)c->k_cell;
        q.theta_cell = (const T *)c->theta_cell;
        q.sigma_cell = (const T *)c->sigma_cell;
        q.C_cell = (const T *)c->C_cell;
        q.c0 = (T)((s.rho_air / s.rho_water) * (s.beta_max / (s.von_karman * s.von_karman)));
        q.z_alpha = (T)s.z_alpha;
        q.p_in = (T)s.p_in;
        q.alpha0 = (T)s.alpha0;
        q.kvk = (T)s.von_karman;
        q.g = (T)s.gravity;
        q.cds = (T)s.cds;
        q.d1 = (T)s.delta1;
        q.d2 = (T)s.delta2;
        q.ls_cds = (T)s.ls_cds;
        q.ls_br = (T)s.ls_br;
        q.ls_inv_br = (T)(1.0 / s.ls_br);
        q.ls_p = (T)s.ls_p;
        q.ls_q = (T)s.ls_sigma_power;
        q.wc_rate = (T)s.wc_rate;
        q.wc_thr = (T)s.wc_threshold;
        q.wc_sp = (T)s.wc_saturation_power;
        q.wc_factor = (const T *)c->t_wc;
        if constexpr (std::is_same<T, float>::value) {
            q.cth_cell = (const T *)c->f_cth;
            q.sth_cell = (const T *)c->f_sth;
            q.logk_cell = (const T *)c->f_logk;
            q.invc_cell = (const T *)c->f_invc;
            q.sk3_cell = (const T *)c->f_sk3;
            q.sq_cell = (const T *)c->f_sq;
            q.cos_tu = (const T *)c->f_ctu;
            q.sin_tu = (const T *)c->f_stu;
            q.log_z1 = (const T *)c->f_lz1;
        }
        if constexpr (std::is_same<T, float>::value) {
            p.tau_c = (const T *)c->tau_c;
            p.tau_s = (const T *)c->tau_s;
        }
        if (c->s_active) {
            q.surrogate = 1;
            q.surrogate_terms = c->scfg.terms;
            q.rate = c->rate;
            q.out_scale = c->out_scale;
            q.rate_ld = c->s_ld;
            q.use_plain = c->use_plain;
        }
    }
    return p;
}

template <typename T>
rpl::StateTables<T> state_tables(const rpl_ctx *c)
{
    rpl::StateTables<T> t{};
    t.kappa = (const T *)c->kappa;
    t.phi = (const T *)c->t_phi;
    t.w = (const T *)c->t_w;
    t.sigma_k = (const T *)c->t_sigma_k;
    t.cos_phi = (const T *)c->t_cos;
    t.sin_phi = (const T *)c->t_sin;
    t.k_pow = (const T *)c->t_kpow;
    t.sigma_pow = (const T *)c->t_spow;
    t.log_kappa = (const T *)c->f_logkap;
    t.inv_c = (const T *)c->f_invck;
    t.rho_air = (T)c->src.rho_air;
    t.rho_w_g = (T)(c->src.rho_water * c->src.gravity);
    t.cap_eps = (T)c->src.cap_epsilon;
    t.inv_p = (T)(1 / c->src.p_mean);
    return t;
}

// Wind of the step at the clock time (constant over the three stages).
int launch_wind(rpl_ctx *c)
{
    c->s_due = true;  // a new step (or a tendency evaluation) begins
    if (!c->has_sources || !c->src.wind_input)
        return RPL_OK;
    hipError_t e;
    const int nx = c->cfg.nx, ny = c->cfg.ny;
    if (!c->track_t.empty()) {
        // LinearStormTrack (winds.jl:28-41) of any length: the segment holding
        // the clock time; the wind kernel interpolates (and clamps) on it.
        const std::vector<double> &T = c->track_t;
        const size_t n = T.size();
        size_t lo = 0, hi = 0;
        if (n > 1) {
            const size_t up = std::upper_bound(T.begin(), T.end(), c->time) - T.begin();  // first T > time
            hi = std::min(std::max<size_t>(up, 1), n - 1);
            lo = hi - 1;
        }
        c->wind.t0 = T[lo];
        c->wind.t1 = T[hi];
        c->wind.x0 = c->track_x[lo];
        c->wind.y0 = c->track_y[lo];
        c->wind.x1 = c->track_x[hi];
        c->wind.y1 = c->track_y[hi];
    }
    if (is_f64(c))
        e = rpl::launch_wind_columns<double>(nx, ny, c->xn, c->yn, c->time, c->wind, (double *)c->U10,
                                             (double *)c->theta_u, (double *)c->ustar, nullptr, nullptr, nullptr,
                                             c->stream);
    else
        e = rpl::launch_wind_columns<float>(nx, ny, c->xn, c->yn, c->time, c->wind, (float *)c->U10,
                                            (float *)c->theta_u, (float *)c->ustar, (float *)c->f_ctu,
                                            (float *)c->f_stu, (float *)c->f_lz1, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("wind kernel: ") + hipGetErrorString(e));
    return RPL_OK;
}

int record_begin(rpl_ctx *c, int phase);
int record_end(rpl_ctx *c);
bool use_fast(const rpl_ctx *c);

// Whether the stress cap of the next RK3 stage comes from the wave-supported
// stress sums the stage kernels write for their output (FP32 fast kernels,
// wind input with the cap, local-saturation or no dissipation, no network:
// the stage kernel evaluates every cell's raw wind-input rate anyway), instead
// of a second pass over the stage input.  The first stage of every step still
// runs the column-state kernel: the wind changes between steps.
bool fuse_cap(const rpl_ctx *c)
{
    return !is_f64(c) && use_fast(c) && c->has_sources && c->src.wind_input && c->src.stress_cap && !c->s_active &&
           (c->src.dissipation == RPL_DISSIPATION_NONE || c->src.dissipation == RPL_DISSIPATION_LOCAL_SATURATION) &&
           (c->cfg.transport == RPL_TRANSPORT_INTRINSIC || c->cfg.transport == RPL_TRANSPORT_REFRACTION) &&
           c->cfg.timestepper != RPL_FORWARD_EULER && c->tau_x && c->tau_c;
}

// Bundle state of one stage input (prepare_sources, precomputed_sources.jl:230-238).
int launch_state(rpl_ctx *c, const void *Nk)
{
    if (!c->has_sources)
        return RPL_OK;
    const bool needs = (c->src.wind_input && c->src.stress_cap) || c->src.dissipation == RPL_DISSIPATION_MEAN_SPECTRUM ||
                       c->src.dissipation == RPL_DISSIPATION_WHITECAPPING;
    if (!needs)
        return RPL_OK;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    int rc = record_begin(c, 5);
    if (rc != RPL_OK)
        return rc;
    hipError_t e;
    if (c->tau_for == Nk && fuse_cap(c)) {
        // the previous stage kernel summed this input's wave-supported stress
        c->tau_for = nullptr;
        e = rpl::launch_stress_factor_from_sums(nx, ny, nk, (const float *)c->tau_x, (const float *)c->tau_y,
                                                stage_params<float>(c, 0.0).src, state_tables<float>(c),
                                                (float *)c->factor, c->stream);
        if (e != hipSuccess)
            return fail(RPL_ERR_HIP, std::string("stress factor kernel: ") + hipGetErrorString(e));
        return record_end(c);
    }
    c->tau_for = nullptr;
    if (is_f64(c)) {
        auto p = stage_params<double>(c, 0.0);
        e = rpl::launch_source_state<double>(nx, ny, nk, np, (const double *)Nk, p.src, state_tables<double>(c),
                                             (double *)c->factor, (double *)c->cm0, (double *)c->kbar,
                                             (double *)c->sbar, c->stream);
    } else {
        auto p = stage_params<float>(c, 0.0);
        e = rpl::launch_source_state<float>(nx, ny, nk, np, (const float *)Nk, p.src, state_tables<float>(c),
                                            (float *)c->factor, (float *)c->cm0, (float *)c->kbar, (float *)c->sbar,
                                            c->stream);
    }
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("source state kernel: ") + hipGetErrorString(e));
    return record_end(c);
}

// Network inference of one stage input: features and forward pass chunk by
// chunk into the rate planes, then the per-column envelope / audit flags.
template <typename T>
int surrogate_rates(rpl_ctx *c, const T *Nk)
{
    const int nbins = c->cfg.nkappa * c->cfg.nphi;
    const rpl_surrogate &g = c->scfg;
    rpl::mlp::FeatureScales fs{g.floor, g.shift, g.scale, g.wind_scale, g.depth_feature};
    hipError_t e = rpl::mlp::envelope_clear(c->plane, c->s_amax, c->s_flags, c->stream);
    for (size_t col0 = 0; col0 < c->plane && e == hipSuccess; col0 += c->s_chunk) {
        const size_t rows = std::min(c->s_chunk, (c->plane - col0 + 63) / 64 * 64);
        e = rpl::mlp::features<T>(Nk, c->plane, col0, rows, nbins, c->net->dims[0], (const T *)c->U10,
                                  (const T *)c->theta_u, fs, c->sX, c->bin_hi, g.action_hi, c->s_amax, c->s_flags,
                                  c->stream);
        if (e == hipSuccess)
            e = rpl::mlp::forward_rows(c->net, c->sX, (int)rows, c->sA, c->sB, nullptr, c->s_ld, c->s_flags + col0,
                                       c->stream, c->rate + col0);
    }
    if (e == hipSuccess) {
        rpl::mlp::Envelope env{g.wind_lo, g.wind_hi, g.action_hi, g.audit_stride};
        e = rpl::mlp::envelope<T>(c->plane, (const T *)c->U10, c->s_flags, env, c->use_plain, c->s_counts,
                                  c->stream);
    }
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("surrogate inference: ") + hipGetErrorString(e));
    c->s_columns += (int64_t)c->plane;
    return RPL_OK;
}

template <typename T>
int surrogate_audit(rpl_ctx *c, const T *Nk)
{
    auto p = stage_params<T>(c, 0.0);
    p.src.surrogate = 0;  // the audit compares against the plain bundle
    if (c->scfg.terms == 1)
        p.src.dissipation = RPL_DISSIPATION_NONE;  // ... the part the network stands in for
    hipError_t e = rpl::mlp::audit<T>(Nk, c->plane, c->cfg.nkappa * c->cfg.nphi, c->rate, c->s_ld, c->out_scale,
                                      c->use_plain, c->scfg.audit_stride, p.src, c->scfg.audit_bound,
                                      c->s_counts, c->s_err, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("surrogate audit: ") + hipGetErrorString(e));
    return RPL_OK;
}

int record_begin(rpl_ctx *c, int phase);
int record_end(rpl_ctx *c);

// Source preparation of one stage input: surrogate inference (if attached),
// the bundle's column state (only on plain and audit columns when the
// surrogate runs), then the audit.
// The DIA source of the stage input into the additive plane.
template <typename T>
int launch_dia_t(rpl_ctx *c, const T *Nk)
{
    rpl::DiaTables<T> tb{(const int *)c->dia_flags, (const int *)c->dia_bins,  (const T *)c->dia_wef,
                         (const T *)c->dia_wdiv,       (const T *)c->dia_efa,  (const T *)c->dia_inv_efa,
                         (const T *)c->dia_cfa,        (T)c->dia_dal[0],       (T)c->dia_dal[1],
                         (T)c->dia_dal[2],             (const int *)c->dia_gat_off, (const int *)c->dia_gat_q,
                         (const T *)c->dia_gat_w};
    hipError_t e = rpl::launch_dia_source<T>(c->cfg.nx, c->cfg.ny, c->cfg.nkappa, c->cfg.nphi, Nk, tb,
                                             (T *)c->dia_add, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("symmetric quadruplet: ") + hipGetErrorString(e));
    return RPL_OK;
}

// The non-finite check of the action after a step (rpl_set_nonfinite_check).
int launch_nonfinite(rpl_ctx *c)
{
    auto *counter = (unsigned long long *)c->nf_buf;
    auto *record = (long long *)c->nf_buf + 1;
    hipError_t e = is_f64(c) ? rpl::launch_nonfinite_check<double>(c->cnt, (const double *)c->A, counter, record,
                                                                    c->iteration, c->stream)
                             : rpl::launch_nonfinite_check<float>(c->cnt, (const float *)c->A, counter, record,
                                                                   c->iteration, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("non-finite check: ") + hipGetErrorString(e));
    return RPL_OK;
}

int launch_dia(rpl_ctx *c, const void *Nk)
{
    if (c->listed) {
        int rc = record_begin(c, 4);
        if (rc)
            return rc;
        hipError_t e;
        const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
        if (is_f64(c))
            e = rpl::launch_listed_source<double>(nx, ny, nk, np, c->listed_n, (const int *)c->listed_bins,
                                                  (const double *)c->listed_rates,
                                                  c->listed_has_field ? (const double *)c->listed_field : nullptr,
                                                  c->listed_power,
                                                  (const double *)c->listed_w, (const double *)Nk,
                                                  (double *)c->dia_add, c->stream);
        else
            e = rpl::launch_listed_source<float>(nx, ny, nk, np, c->listed_n, (const int *)c->listed_bins,
                                                 (const float *)c->listed_rates,
                                                 c->listed_has_field ? (const float *)c->listed_field : nullptr,
                                                 (float)c->listed_power,
                                                 (const float *)c->listed_w, (const float *)Nk, (float *)c->dia_add,
                                                 c->stream);
        if (e != hipSuccess)
            return fail(RPL_ERR_HIP, std::string("listed quadruplets: ") + hipGetErrorString(e));
        return record_end(c);
    }
    if (!c->dia)
        return RPL_OK;
    int rc = record_begin(c, 4);
    if (rc == RPL_OK)
        rc = is_f64(c) ? launch_dia_t<double>(c, (const double *)Nk) : launch_dia_t<float>(c, (const float *)Nk);
    return rc == RPL_OK ? record_end(c) : rc;
}

int launch_sources(rpl_ctx *c, const void *Nk)
{
    int rc0 = launch_dia(c, Nk);
    if (rc0)
        return rc0;
    if (!c->s_active)
        return launch_state(c, Nk);
    int rc = RPL_OK;
    if (c->scfg.refresh == 0 || c->s_due) {
        rc = record_begin(c, 1);
        if (rc == RPL_OK)
            rc = is_f64(c) ? surrogate_rates<double>(c, (const double *)Nk)
                           : surrogate_rates<float>(c, (const float *)Nk);
        if (rc == RPL_OK)
            rc = record_end(c);
        c->s_due = false;
    }
    if (rc || (rc = launch_state(c, Nk)))
        return rc;
    if (c->scfg.audit_stride > 0)
        rc = is_f64(c) ? surrogate_audit<double>(c, (const double *)Nk) : surrogate_audit<float>(c, (const float *)Nk);
    return rc;
}

bool use_fast(const rpl_ctx *c)
{
    if (is_f64(c))
        return false;
    if (c->cfg.kernel == RPL_KERNEL_REFERENCE)
        return false;
    return true;
}

// Whether launch_stage takes the phi-walk kernel (launch_stage_fast) rather
// than the reference-order one.
bool phi_walk_kernel(const rpl_ctx *c)
{
    if (is_f64(c))
        return c->cfg.kernel == RPL_KERNEL_FAST && c->cfg.transport != RPL_TRANSPORT_NONE &&
               c->plane < ((size_t)1 << 29);
    return use_fast(c);
}

int record_begin(rpl_ctx *c, int phase)
{
    if (!c->timing)
        return RPL_OK;
    if (c->events_used == c->events.size()) {
        hipEvent_t a, b;
        RPL_HIP(hipEventCreate(&a));
        RPL_HIP(hipEventCreate(&b));
        c->events.emplace_back(a, b);
        c->event_phase.push_back(0);
    }
    c->event_phase[c->events_used] = phase;
    RPL_HIP(hipEventRecord(c->events[c->events_used].first, c->stream));
    return RPL_OK;
}

int fold_timing(rpl_ctx *c)
{
    if (c->events_used == 0)
        return RPL_OK;
    RPL_HIP(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < c->events_used; ++k) {
        float ms = 0;
        RPL_HIP(hipEventElapsedTime(&ms, c->events[k].first, c->events[k].second));
        c->timing_ms[c->event_phase[k]] += ms;
        c->timing_launches[c->event_phase[k]] += 1;
    }
    c->events_used = 0;
    return RPL_OK;
}

int record_end(rpl_ctx *c)
{
    if (!c->timing)
        return RPL_OK;
    RPL_HIP(hipEventRecord(c->events[c->events_used].second, c->stream));
    c->events_used++;
    if (c->events_used >= 3072)
        return fold_timing(c);
    return RPL_OK;
}

// Pseudomomentum contractions take the reference order unless the context
// asked for the fast kernels (FP32 AUTO or FAST, FP64 FAST).
bool pm_ordered(const rpl_ctx *c)
{
    return !(c->cfg.kernel == RPL_KERNEL_FAST || (!is_f64(c) && c->cfg.kernel == RPL_KERNEL_AUTO));
}

// update_coupling! of the self-coupled current (current_coupling.jl:140-143,
// run first by compute_tendencies!, tendencies.jl:103-107): U and dU/dkappa of
// the stage input Nk, then the gradients the reference-order stencil reads
// (the fast stencil forms its own from U).
int launch_coupling(rpl_ctx *c, const void *Nk)
{
    if (!c->pm)
        return RPL_OK;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    const bool ordered = pm_ordered(c);
    int rc = record_begin(c, 2);
    if (rc != RPL_OK)
        return rc;
    hipError_t e;
    if (is_f64(c)) {
        e = rpl::launch_pm_contract<double>(ordered, nx, ny, nk, np, nk, (const double *)Nk,
                                            (const double *)c->pm_O,
                                            (const double *)c->pm_dO, (const double *)c->pm_kx,
                                            (const double *)c->pm_ky, nullptr, (double *)c->Ux, (double *)c->Uy,
                                            (double *)c->dUx, (double *)c->dUy, c->stream, c->pm_percol);
        if (e == hipSuccess && !use_fast(c))
            e = rpl::launch_gradients<double>(nx, ny, nk, (const double *)c->Ux, (const double *)c->Uy, 1.0 / c->dx,
                                              1.0 / c->dy, (double *)c->Uxx, (double *)c->Uxy, (double *)c->Uyx,
                                              (double *)c->Uyy, c->stream);
    } else {
        e = rpl::launch_pm_contract<float>(ordered, nx, ny, nk, np, nk, (const float *)Nk, (const float *)c->pm_O, (const float *)c->pm_dO, (const float *)c->pm_kx,
                                           (const float *)c->pm_ky, nullptr, (float *)c->Ux, (float *)c->Uy,
                                           (float *)c->dUx, (float *)c->dUy, c->stream, c->pm_percol);
        if (e == hipSuccess && !use_fast(c))
            e = rpl::launch_gradients<float>(nx, ny, nk, (const float *)c->Ux, (const float *)c->Uy,
                                             1.0f / (float)c->dx, 1.0f / (float)c->dy, (float *)c->Uxx,
                                             (float *)c->Uxy, (float *)c->Uyx, (float *)c->Uyy, c->stream);
    }
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("pseudomomentum coupling: ") + hipGetErrorString(e));
    return record_end(c);
}

// Everything a tendency evaluation needs before the stage kernel: the
// coupling's velocity and the source state of the stage input.
int prepare_stage(rpl_ctx *c, const void *Nk)
{
    int rc = launch_coupling(c, Nk);
    return rc ? rc : launch_sources(c, Nk);
}

int launch_stage(rpl_ctx *c, double dt, int stage, const void *Nk, const void *N0, void *out)
{
    int rc = record_begin(c, 0);
    if (rc != RPL_OK)
        return rc;
    hipError_t e;
    const int tr = c->cfg.transport;
    if (is_f64(c)) {
        auto p = stage_params<double>(c, dt);
        // FP64 under RPL_KERNEL_FAST: the phi-loop kernel (rounding-level
        // differences from the reference order); slabs of 2^29 cells and
        // more stay on the reference-order kernel (32-bit byte offsets)
        if (phi_walk_kernel(c))
            e = rpl::launch_stage_fast(p, tr, stage, (const double *)Nk, (const double *)N0, (double *)out,
                                       c->stream);
        else
            e = rpl::launch_stage_reference(p, tr, stage, (const double *)Nk, (const double *)N0, (double *)out,
                                            c->stream);
    } else {
        auto p = stage_params<float>(c, dt);
        // stages 1 and 2 sum their output's wave-supported stress for the next stage's cap
        c->tau_for = nullptr;
        if ((stage == rpl::STAGE_1 || stage == rpl::STAGE_2) && fuse_cap(c)) {
            p.tau_x = (float *)c->tau_x;
            p.tau_y = (float *)c->tau_y;
            c->tau_for = out;
        }
        if (use_fast(c))
            e = rpl::launch_stage_fast(p, tr, stage, (const float *)Nk, (const float *)N0, (float *)out, c->stream);
        else
            e = rpl::launch_stage_reference(p, tr, stage, (const float *)Nk, (const float *)N0, (float *)out,
                                            c->stream);
    }
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("stage kernel launch: ") + hipGetErrorString(e));
    return record_end(c);
}

// apply_propagation_smoothing! (Models/propagation_smoothing.jl:160-199):
// the averaged field goes to B, which then becomes the action (the
// reference copies it back; a buffer swap does the same here).
int launch_smoothing(rpl_ctx *c, double dt)
{
    if (!c->gse)
        return RPL_OK;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    // The kernels index the corners with 32-bit integers: keep displacements
    // far inside that range.
    const double reach = (c->gse_as * c->gse_dcg_max + c->gse_an * c->gse_cg_max * std::fabs(c->dp)) * dt *
                         std::max(1 / c->dx, 1 / c->dy);
    if (!(reach < 1048576.0))
        return fail(RPL_ERR_ARG, "propagation smoothing would move cells by more than 2^20 cells in one step");
    int rc = record_begin(c, 3);
    if (rc != RPL_OK)
        return rc;
    hipError_t e;
    if (is_f64(c))
        e = rpl::launch_spatial_averaging<double>(false, nx, ny, nk, np, (const double *)c->A, (double *)c->B,
                                                  (const double *)c->gse_cg, (const double *)c->gse_dcg,
                                                  (const double *)c->cosp, (const double *)c->sinp, c->dp,
                                                  c->gse_as, c->gse_an, dt, 1 / c->dx, 1 / c->dy, c->periodic_x,
                                                  c->periodic_y, c->stream);
    else
        e = rpl::launch_spatial_averaging<float>(use_fast(c), nx, ny, nk, np, (const float *)c->A, (float *)c->B,
                                                 (const float *)c->gse_cg, (const float *)c->gse_dcg,
                                                 (const float *)c->cosp, (const float *)c->sinp, (float)c->dp,
                                                 (float)c->gse_as, (float)c->gse_an, (float)dt,
                                                 1.0f / (float)c->dx, 1.0f / (float)c->dy, c->periodic_x,
                                                 c->periodic_y, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("spatial averaging: ") + hipGetErrorString(e));
    std::swap(c->A, c->B);
    return record_end(c);
}

int check_ready(rpl_ctx *c)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (c->cfg.transport == RPL_TRANSPORT_REFRACTION && !c->has_current)
        return fail(RPL_ERR_STATE, "refraction transport needs rpl_set_current or rpl_set_doppler first");
    RPL_HIP(hipSetDevice(c->device));
    return RPL_OK;
}

}  // namespace

namespace {
void free_surrogate(rpl_ctx *c)
{
    for (void *p : {(void *)c->sX, (void *)c->sA, (void *)c->sB, (void *)c->rate, (void *)c->out_scale,
                    (void *)c->s_err, (void *)c->use_plain, (void *)c->s_counts, (void *)c->bin_hi,
                    (void *)c->s_amax, (void *)c->s_flags})
        free_ptr(p);
    c->sX = c->sA = c->sB = nullptr;
    c->rate = nullptr;
    c->out_scale = c->s_err = c->bin_hi = c->s_amax = nullptr;
    c->s_flags = nullptr;
    c->use_plain = nullptr;
    c->s_counts = nullptr;
    c->net = nullptr;
    c->s_active = false;
}

template <typename T>
int training_pairs(rpl_ctx *c, const rpl_surrogate *cfg, float *features, float *targets)
{
    const int nbins = c->cfg.nkappa * c->cfg.nphi, in_dim = nbins + 3, in_pad = (in_dim + 63) / 64 * 64;
    const size_t chunk = std::min<size_t>((c->plane + 63) / 64 * 64, 16384);
    rpl::mlp::bf16 *X = nullptr;
    float *r = nullptr, *unit = nullptr;
    int rc = RPL_OK;
    if ((rc = alloc((void **)&X, chunk * in_pad * 2)) || (rc = alloc((void **)&r, c->plane * nbins * 4)) ||
        (rc = alloc((void **)&unit, nbins * 4))) {
        free_ptr(X);
        free_ptr(r);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (features) {
        rpl::mlp::FeatureScales fs{cfg->floor, cfg->shift, cfg->scale, cfg->wind_scale, cfg->depth_feature};
        std::vector<rpl::mlp::bf16> h(chunk * in_pad);
        for (size_t col0 = 0; col0 < c->plane && e == hipSuccess; col0 += chunk) {
            const size_t rows = std::min(chunk, (c->plane - col0 + 63) / 64 * 64);
            e = rpl::mlp::features<T>((const T *)c->A, c->plane, col0, rows, nbins, in_pad, (const T *)c->U10,
                                      (const T *)c->theta_u, fs, X, nullptr, 0.0f, nullptr, nullptr, c->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess)
                e = hipMemcpy(h.data(), X, rows * in_pad * 2, hipMemcpyDeviceToHost);
            const size_t valid = std::min(rows, c->plane - col0);
            for (size_t q = 0; e == hipSuccess && q < valid; ++q)
                for (int f = 0; f < in_dim; ++f) {
                    const uint32_t bits = (uint32_t)h[q * in_pad + f] << 16;
                    float v;
                    std::memcpy(&v, &bits, 4);
                    features[(col0 + q) * in_dim + f] = v;
                }
        }
    }
    if (e == hipSuccess && targets) {
        std::vector<float> u(nbins);
        for (int t = 0; t < nbins; ++t)
            u[t] = (float)(c->h_sigma[t] * cfg->rate_unit);
        e = hipMemcpy(unit, u.data(), nbins * 4, hipMemcpyHostToDevice);
        auto p = stage_params<T>(c, 0.0);
        if (cfg->terms == 1)
            p.src.dissipation = RPL_DISSIPATION_NONE;  // targets of the wind input alone
        if (e == hipSuccess)
            e = rpl::launch_source_rates<T>(c->cfg.nx, c->cfg.ny, nbins, (const T *)c->A, p.src, unit, r, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        std::vector<float> h(c->plane * nbins);
        if (e == hipSuccess)
            e = hipMemcpy(h.data(), r, h.size() * 4, hipMemcpyDeviceToHost);
        for (int t = 0; e == hipSuccess && t < nbins; ++t)
            for (size_t col = 0; col < c->plane; ++col)
                targets[col * nbins + t] = h[col + c->plane * t];
    }
    free_ptr(X);
    free_ptr(r);
    free_ptr(unit);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("surrogate training pairs: ") + hipGetErrorString(e));
    return RPL_OK;
}
}  // namespace

struct rpl_ensemble {
    std::vector<rpl_ctx *> members;      // borrowed
    std::vector<hipStream_t> own_stream; // each member's stream, restored on destroy
    int device = 0;
    hipStream_t stream = nullptr;
    int transport = 0;
    bool src = false;
    int src_kind = 0;  // rpl::SOURCE_* of the members' stage launches
    int ty = 4, blocks = 0, cus = 0;
    rpl::EnsembleMember *desc = nullptr;  // 3 stages x members, device
    long long total = 0;
    rpl::EnsembleState *sdesc = nullptr;  // 3 stages x members with a column state, device
    int *chunk_member = nullptr;          // member of every 2^ENSEMBLE_CHUNK_SHIFT-th tile, device
    size_t chunk_cap = 0;
    int smembers = 0;
    long long sblocks = 0;
    // Members with at least `split` tiles of 64 x 4 cells x one kappa bin
    // run their own stage kernel (as rpl_time_step does) instead of joining
    // the work list: 0 the default threshold, negative never.
    long long split = 0;
    std::vector<rpl_ctx *> own;            // those members ...
    std::vector<int> own_branch;           // ... and the branch stream each runs on
    hipStream_t branch[4] = {};            // branch streams (default priority)
    hipEvent_t fork = nullptr, join[4] = {};
    std::vector<rpl_ctx *> list;           // the members in the work list
    int lmembers = 0;
    double desc_dt = -1;
    int desc_ty = 0;
    bool use_graph = false;
    hipGraphExec_t exec = nullptr;
    double graph_dt = -1;
    int graph_ty = 0, graph_blocks = 0;
    long long graph_split = 0;
    std::vector<uint64_t> graph_gen;  // the members' generations at capture
};

namespace {

// Stage buffers of an RK3 step: stage 1 A -> B, stage 2 (B, A) -> C, stage 3 (C, A) -> A.
void stage_buffers(rpl_ctx *c, int stage, const void **Nk, const void **N0, void **out)
{
    *N0 = c->A;
    if (stage == rpl::STAGE_1) {
        *Nk = c->A;
        *out = c->B;
    } else if (stage == rpl::STAGE_2) {
        *Nk = c->B;
        *out = c->C;
    } else {
        *Nk = c->C;
        *out = c->A;
    }
}

// Default grid: one workgroup per tile (0 = resolved against the tile count
// of the current launch shape when launching).
int default_blocks(const rpl_ensemble *) { return 0; }

// Default size (tiles of 64 x 4 cells x one kappa bin) from which a member
// runs its own stage kernel: two rounds of workgroups over 256 CUs at the
// stage kernel's occupancy, where the per-model launch fills the chip.
constexpr long long ENSEMBLE_OWN_TILES = 4096;
// Branch streams the large members' step chains are spread over (the HIP
// runtime's default of four hardware queues per process).
constexpr int ENSEMBLE_BRANCHES = 4;
static_assert(sizeof(rpl_ensemble::branch) / sizeof(hipStream_t) == ENSEMBLE_BRANCHES, "branch streams");

int launch_blocks(const rpl_ensemble *e)
{
    if (e->blocks > 0)
        return e->blocks;
    const long long g = (e->total + 7) / 8 * 8;
    return (int)std::max<long long>(8, std::min<long long>(g, 0x7ffffff8LL));
}


// Tiles of 64 x 4 cells x one kappa bin of a member: the size by which a
// member is sent to its own stage kernel.
long long member_tiles(const rpl_ctx *c)
{
    return (long long)((c->cfg.nx + 63) / 64) * ((c->cfg.ny + 3) / 4) * c->cfg.nkappa;
}

bool own_kernel(const rpl_ensemble *e, const rpl_ctx *c)
{
    const long long thr = e->split == 0 ? ENSEMBLE_OWN_TILES : e->split;
    return thr > 0 && member_tiles(c) >= thr;
}

int ensemble_descriptors(rpl_ensemble *e, double dt)
{
    // Work-list members first (k < M), then the members with a kernel of their own.
    std::vector<rpl_ctx *> order;
    for (rpl_ctx *c : e->members)
        if (!own_kernel(e, c))
            order.push_back(c);
    const int M = (int)order.size();
    for (rpl_ctx *c : e->members)
        if (own_kernel(e, c))
            order.push_back(c);
    // The large members' step chains go to the branch streams, largest
    // first onto the least loaded one.
    e->own.assign(order.begin() + M, order.end());
    std::sort(e->own.begin(), e->own.end(),
              [](const rpl_ctx *a, const rpl_ctx *b) { return member_tiles(a) > member_tiles(b); });
    e->own_branch.assign(e->own.size(), 0);
    long long load[ENSEMBLE_BRANCHES] = {};
    for (size_t k = 0; k < e->own.size(); ++k) {
        const int b = (int)(std::min_element(load, load + ENSEMBLE_BRANCHES) - load);
        e->own_branch[k] = b;
        load[b] += member_tiles(e->own[k]);
    }
    order.resize(M);
    std::vector<rpl::EnsembleMember> h(3 * (size_t)M);
    long long t0 = 0;
    for (int k = 0; k < M; ++k) {
        rpl_ctx *c = order[k];
        const int nbx = (c->cfg.nx + 63) / 64, nby = (c->cfg.ny + e->ty - 1) / e->ty;
        for (int st = 0; st < 3; ++st) {
            rpl::EnsembleMember &d = h[(size_t)st * M + k];
            d.p = stage_params<float>(c, dt);
            if (st < 2 && fuse_cap(c)) {  // as launch_stage: stages 1 and 2 sum their output's stress
                d.p.tau_x = (float *)c->tau_x;
                d.p.tau_y = (float *)c->tau_y;
            }
            const void *Nk, *N0;
            void *out;
            stage_buffers(c, st + 1, &Nk, &N0, &out);
            d.Nk = (const float *)Nk;
            d.N0 = (const float *)N0;
            d.out = (float *)out;
            d.nbx = nbx;
            d.nby = nby;
            d.tile0 = t0;
        }
        // Refraction and intrinsic entries walk the phi bins inside the
        // workgroup: one per (tile, kappa bin); without transport, one per
        // (tile, bin).
        const bool phi_loop = e->transport == RPL_TRANSPORT_REFRACTION || e->transport == RPL_TRANSPORT_INTRINSIC;
        t0 += (long long)nbx * nby * c->cfg.nkappa * (phi_loop ? 1 : c->cfg.nphi);
    }
    e->total = t0;
    e->lmembers = M;
    e->list = order;
    // One source kind for the whole launch: the specialised bundle path when
    // every member qualifies (the two paths give the same bits).
    e->src_kind = rpl::SOURCE_NONE;
    if (e->src) {
        e->src_kind = rpl::SOURCE_JP_LS;
        for (int k = 0; k < M; ++k)
            if (!rpl::source_is_jp_ls(h[k].p.src))
                e->src_kind = rpl::SOURCE_GENERIC;
    }
    if (M)
        RPL_HIP(hipMemcpy(e->desc, h.data(), h.size() * sizeof(rpl::EnsembleMember), hipMemcpyHostToDevice));
    const size_t nchunks = (size_t)((t0 + (1LL << rpl::ENSEMBLE_CHUNK_SHIFT) - 1) >> rpl::ENSEMBLE_CHUNK_SHIFT);
    std::vector<int> cm(nchunks);
    for (size_t q = 0, k = 0; q < nchunks; ++q) {
        const long long tile = (long long)q << rpl::ENSEMBLE_CHUNK_SHIFT;
        while (k + 1 < (size_t)M && h[k + 1].tile0 <= tile)
            ++k;
        cm[q] = (int)k;
    }
    if (nchunks > e->chunk_cap) {
        if (e->chunk_member)
            (void)hipFree(e->chunk_member);
        e->chunk_member = nullptr;
        e->chunk_cap = 0;
        RPL_HIP(hipMalloc((void **)&e->chunk_member, nchunks * sizeof(int)));
        e->chunk_cap = nchunks;
    }
    if (nchunks)
        RPL_HIP(hipMemcpy(e->chunk_member, cm.data(), nchunks * sizeof(int), hipMemcpyHostToDevice));
    // Column states (stress cap, mean-spectrum moments) of the work-list members that need one.
    std::vector<rpl_ctx *> sm;
    for (rpl_ctx *c : order)
        if (c->has_sources && ((c->src.wind_input && c->src.stress_cap) ||
                               c->src.dissipation == RPL_DISSIPATION_MEAN_SPECTRUM ||
                               c->src.dissipation == RPL_DISSIPATION_WHITECAPPING))
            sm.push_back(c);
    const int S = (int)sm.size();
    std::vector<rpl::EnsembleState> hs(3 * (size_t)S);
    long long b0 = 0;
    for (int k = 0; k < S; ++k) {
        rpl_ctx *c = sm[k];
        for (int st = 0; st < 3; ++st) {
            rpl::EnsembleState &d = hs[(size_t)st * S + k];
            d.nx = c->cfg.nx;
            d.ny = c->cfg.ny;
            d.nk = c->cfg.nkappa;
            d.nphi = c->cfg.nphi;
            const void *Nk, *N0;
            void *out;
            stage_buffers(c, st + 1, &Nk, &N0, &out);
            d.N = (const float *)Nk;
            d.s = stage_params<float>(c, 0.0).src;
            d.tb = state_tables<float>(c);
            d.factor = (float *)c->factor;
            d.m0 = (float *)c->cm0;
            d.kbar = (float *)c->kbar;
            d.sbar = (float *)c->sbar;
            d.tau_x = d.tau_y = nullptr;
            if (st > 0 && fuse_cap(c)) {  // stages 2 and 3: the factor from the sums
                d.tau_x = (const float *)c->tau_x;
                d.tau_y = (const float *)c->tau_y;
            }
            d.blk0 = b0;
        }
        b0 += (long long)((c->plane + 63) / 64);
    }
    e->smembers = S;
    e->sblocks = b0;
    if (S)
        RPL_HIP(hipMemcpy(e->sdesc, hs.data(), hs.size() * sizeof(rpl::EnsembleState), hipMemcpyHostToDevice));
    e->desc_dt = dt;
    e->desc_ty = e->ty;
    return RPL_OK;
}

// Stages first..last of the work-list members on the ensemble stream: per
// stage the network inference of surrogate members, the column states in one
// launch, the surrogate audits, then one launch over the work list.
int ensemble_list_stages(rpl_ensemble *e, int first, int last)
{
    const int M = e->lmembers;
    if (M == 0)
        return RPL_OK;
    for (int st = first; st <= last; ++st) {
        if (e->src) {
            for (rpl_ctx *c : e->list) {
                if (!c->s_active || (c->scfg.refresh == 1 && st != rpl::STAGE_1))
                    continue;  // refresh 1: the first stage's rates serve the step
                const void *Nk, *N0;
                void *out;
                stage_buffers(c, st, &Nk, &N0, &out);
                int rc = surrogate_rates<float>(c, (const float *)Nk);
                if (rc)
                    return rc;
            }
            if (e->smembers) {
                hipError_t err = rpl::launch_ensemble_state(e->sdesc + (size_t)(st - 1) * e->smembers, e->smembers,
                                                            e->sblocks, e->stream);
                if (err != hipSuccess)
                    return fail(RPL_ERR_HIP, std::string("ensemble state launch: ") + hipGetErrorString(err));
            }
            for (rpl_ctx *c : e->list) {
                if (!c->s_active || c->scfg.audit_stride <= 0)
                    continue;
                const void *Nk, *N0;
                void *out;
                stage_buffers(c, st, &Nk, &N0, &out);
                int rc = surrogate_audit<float>(c, (const float *)Nk);
                if (rc)
                    return rc;
            }
        }
        hipError_t err = rpl::launch_ensemble_stage(e->desc + (size_t)(st - 1) * M, M, e->total, e->chunk_member,
                                                    e->transport, e->src_kind, st, e->ty, launch_blocks(e), e->stream);
        if (err != hipSuccess)
            return fail(RPL_ERR_HIP, std::string("ensemble stage launch: ") + hipGetErrorString(err));
    }
    return RPL_OK;
}

// Stages first..last of a large member on its branch stream: the launches of
// rpl_time_step (source preparation, stage kernel), so the member advances
// exactly as when stepped alone.
int ensemble_own_stages(rpl_ensemble *e, size_t k, int first, int last, double dt)
{
    rpl_ctx *c = e->own[k];
    c->stream = e->branch[e->own_branch[k]];
    int rc = RPL_OK;
    for (int st = first; st <= last && rc == RPL_OK; ++st) {
        const void *Nk, *N0;
        void *out;
        stage_buffers(c, st, &Nk, &N0, &out);
        if ((rc = prepare_stage(c, Nk)) == RPL_OK)
            rc = launch_stage(c, dt, st, Nk, N0, out);
    }
    c->stream = e->stream;
    return rc;
}

// Stages first..last of every member (no wind: that is evaluated at the
// clock time before each step).  The large members' chains fork from the
// ensemble stream onto the branch streams and join it at the end, so they
// overlap one another and the work-list launches.
int ensemble_run(rpl_ensemble *e, int first, int last, double dt)
{
    const bool fork = !e->own.empty();
    if (fork) {
        RPL_HIP(hipEventRecord(e->fork, e->stream));
        for (int b = 0; b < ENSEMBLE_BRANCHES; ++b)
            RPL_HIP(hipStreamWaitEvent(e->branch[b], e->fork, 0));
    }
    int rc = RPL_OK;
    for (size_t k = 0; k < e->own.size() && rc == RPL_OK; ++k)
        rc = ensemble_own_stages(e, k, first, last, dt);
    if (rc == RPL_OK)
        rc = ensemble_list_stages(e, first, last);
    if (fork) {
        // join even after an error, so a capture in progress can end
        for (int b = 0; b < ENSEMBLE_BRANCHES; ++b) {
            hipError_t x = hipEventRecord(e->join[b], e->branch[b]);
            if (x == hipSuccess)
                x = hipStreamWaitEvent(e->stream, e->join[b], 0);
            if (x != hipSuccess && rc == RPL_OK)
                rc = fail(RPL_ERR_HIP, std::string("ensemble join: ") + hipGetErrorString(x));
        }
    }
    return rc;
}

int ensemble_stages(rpl_ensemble *e, double dt) { return ensemble_run(e, rpl::STAGE_1, rpl::STAGE_3, dt); }

// The rules a context must meet to step in an ensemble with `first`
// (checked at creation and again before every ensemble step, since a
// member's configuration may change in between).
int ensemble_member_check(rpl_ctx *c, rpl_ctx *first)
{
    if (c->cfg.dtype != RPL_F32 || !use_fast(c))
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members must be FP32 contexts on the fast kernel");
    if (c->cfg.timestepper == RPL_FORWARD_EULER)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members must use SSP-RK3");
    if (c->cfg.transport != first->cfg.transport || c->cfg.transport == RPL_TRANSPORT_NONE)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members need the same transport (intrinsic or refraction)");
    if (c->has_sources != first->has_sources)
        return fail(RPL_ERR_UNSUPPORTED, "either every member or no member has source terms");
    if (c->pm)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members take prescribed currents only (no pseudomomentum coupling)");
    if (c->gse)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members run without propagation smoothing");
    if (c->dia || c->listed)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members run without the nonlinear (DIA) term");
    if (c->has_depth_field)
        return fail(RPL_ERR_UNSUPPORTED, "ensemble members take a constant or infinite depth");
    if (c->device != first->device)
        return fail(RPL_ERR_ARG, "ensemble members must live on one device");
    return RPL_OK;
}

void destroy_branches(rpl_ensemble *e)
{
    for (int b = 0; b < ENSEMBLE_BRANCHES; ++b) {
        if (e->branch[b])
            (void)hipStreamDestroy(e->branch[b]);
        if (e->join[b])
            (void)hipEventDestroy(e->join[b]);
        e->branch[b] = nullptr;
        e->join[b] = nullptr;
    }
    if (e->fork)
        (void)hipEventDestroy(e->fork);
    e->fork = nullptr;
}

void drop_graph(rpl_ensemble *e)
{
    if (e->exec)
        (void)hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
    e->graph_dt = -1;
    e->graph_gen.clear();
}

}  // namespace

extern "C" {

const char *rpl_last_error(void) { return g_last_error.c_str(); }

int rpl_create(rpl_ctx **out, const rpl_config *cfg)
{
    if (!out || !cfg)
        return fail(RPL_ERR_ARG, "null argument");
    *out = nullptr;
    const rpl_config &k = *cfg;
    if (k.nx < 3 || k.ny < 3)
        return fail(RPL_ERR_ARG, "nx and ny must be >= 3 (periodic WENO5 stencils reach 3 cells)");
    if (k.nphi < 3)
        return fail(RPL_ERR_ARG, "nphi must be >= 3 (periodic WENO5 stencil in phi)");
    if (k.nkappa < 1)
        return fail(RPL_ERR_ARG, "nkappa must be >= 1");
    if (k.dtype != RPL_F32 && k.dtype != RPL_F64)
        return fail(RPL_ERR_ARG, "dtype must be RPL_F32 or RPL_F64");
    if (k.timestepper != RPL_RK3 && k.timestepper != RPL_LSRK3 && k.timestepper != RPL_FORWARD_EULER)
        return fail(RPL_ERR_ARG, "unsupported timestepper");
    if (k.transport < RPL_TRANSPORT_NONE || k.transport > RPL_TRANSPORT_REFRACTION)
        return fail(RPL_ERR_ARG, "unsupported transport");
    if (k.kernel < RPL_KERNEL_AUTO || k.kernel > RPL_KERNEL_FAST)
        return fail(RPL_ERR_ARG, "unsupported kernel selector");
    if ((long long)k.nkappa * k.nphi > 65535)
        return fail(RPL_ERR_UNSUPPORTED, "nkappa * nphi must be <= 65535");
    if ((long long)k.nx * k.ny >= (1LL << 30))
        return fail(RPL_ERR_UNSUPPORTED, "nx * ny must be < 2^30 (32-bit in-slab byte offsets)");

    rpl_ctx *c = new rpl_ctx();
    c->cfg = k;
    c->plane = (size_t)k.nx * k.ny;
    c->cnt = c->plane * k.nkappa * k.nphi;
    c->esize = k.dtype == RPL_F64 ? sizeof(double) : sizeof(float);
    int rc = RPL_OK;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(RPL_ERR_HIP, std::string("stream creation: ") + hipGetErrorString(e));
    }
    const size_t p3 = c->plane * k.nkappa, tab = (size_t)k.nkappa * k.nphi;
    void **bufs3[] = {&c->Ux, &c->Uy, &c->Uxx, &c->Uxy, &c->Uyx, &c->Uyy};
    if ((rc = alloc(&c->A, c->cnt * c->esize)) || (rc = alloc(&c->B, c->cnt * c->esize)) ||
        (rc = alloc(&c->C, c->cnt * c->esize)))
        goto error;
    for (void **b : bufs3)
        if ((rc = alloc(b, p3 * c->esize)))
            goto error;
    if ((rc = alloc(&c->kappa, k.nkappa * c->esize)) || (rc = alloc(&c->cgx, tab * c->esize)) ||
        (rc = alloc(&c->cgy, tab * c->esize)) || (rc = alloc(&c->cosp, k.nphi * c->esize)) ||
        (rc = alloc(&c->sinp, k.nphi * c->esize)))
        goto error;
    if (!is_f64(c) && (rc = alloc(&c->bin_tab, tab * rpl::BT_FIELDS * sizeof(float))))
        goto error;
    if ((rc = alloc((void **)&c->w, tab * 8)) || (rc = alloc((void **)&c->trace, tab * 8)) ||
        (rc = alloc((void **)&c->mux, tab * 8)) || (rc = alloc((void **)&c->muy, tab * 8)) ||
        (rc = alloc((void **)&c->energy, tab * 8)))
        goto error;
    if ((rc = alloc((void **)&c->partials, 2 * 1024 * 8)))
        goto error;
    e = hipMemsetAsync(c->A, 0, c->cnt * c->esize, c->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(c->Ux, 0, p3 * c->esize, c->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(c->Uy, 0, p3 * c->esize, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        rc = fail(RPL_ERR_HIP, std::string("initial clear: ") + hipGetErrorString(e));
        goto error;
    }
    *out = c;
    return RPL_OK;
error:
    rpl_destroy(c);
    return rc;
}

int rpl_destroy(rpl_ctx *c)
{
    if (!c)
        return RPL_OK;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    for (void *p : {c->A, c->B, c->C, c->Ux, c->Uy, c->Uxx, c->Uxy, c->Uyx, c->Uyy, c->kappa, c->cgx, c->cgy,
                    c->cosp, c->sinp, (void *)c->w, (void *)c->trace, (void *)c->mux, (void *)c->muy,
                    (void *)c->energy, (void *)c->diag, (void *)c->partials, c->scratch_in, (void *)c->xn,
                    (void *)c->yn, c->U10, c->theta_u, c->ustar, c->factor, c->cm0, c->kbar, c->sbar, c->k_cell,
                    c->theta_cell, c->sigma_cell, c->C_cell, c->t_phi, c->t_w, c->t_sigma_k, c->t_cos, c->t_sin,
                    c->t_kpow, c->t_spow, c->t_wc, c->f_cth, c->f_sth, c->f_logk, c->f_invc, c->f_sk3, c->f_sq,
                    c->tau_x, c->tau_y, c->tau_c, c->tau_s, c->bin_tab,
                    c->f_logkap, c->f_invck, c->f_ctu, c->f_stu, c->f_lz1, (void *)c->sX, (void *)c->sA, (void *)c->sB, (void *)c->rate,
                    (void *)c->out_scale, (void *)c->s_err, (void *)c->use_plain, (void *)c->s_counts,
                    (void *)c->bin_hi, (void *)c->s_amax, (void *)c->s_flags, c->dUx, c->dUy, c->pm_O,
                    c->pm_dO, c->pm_kx, c->pm_ky, c->gse_cg, c->gse_dcg, c->dia_flags, c->dia_bins, c->dia_wef,
                    c->dia_wdiv, c->dia_efa, c->dia_inv_efa, c->dia_cfa, c->dia_add, c->dia_gat_off, c->dia_gat_q, c->dia_gat_w, c->listed_bins,
                    c->listed_rates, c->listed_w, c->listed_field, c->nf_buf, c->cgm, c->cg_vec, c->cg_dir})
        free_ptr(p);
    for (auto &ev : c->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return RPL_OK;
}

int rpl_set_geometry(rpl_ctx *c, double dx, double dy, const double *kappa_c, const double *kappa_f,
                     const double *phi_c, const double *phi_f, int32_t depth_kind, double depth, double gravity)
{
    if (!c || !kappa_c || !kappa_f || !phi_c || !phi_f)
        return fail(RPL_ERR_ARG, "null argument");
    ++c->generation;
    if (!(dx > 0) || !(dy > 0))
        return fail(RPL_ERR_ARG, "dx and dy must be positive");
    if (!(gravity > 0))
        return fail(RPL_ERR_ARG, "gravity must be positive");
    if (depth_kind != RPL_DEPTH_INFINITE && depth_kind != RPL_DEPTH_CONSTANT)
        return fail(RPL_ERR_ARG, "depth_kind must be RPL_DEPTH_INFINITE or RPL_DEPTH_CONSTANT");
    if (depth_kind == RPL_DEPTH_CONSTANT && !(depth > 0))
        return fail(RPL_ERR_ARG, "finite-depth dispersion requires positive depth");  // transport.jl:77
    const int nk = c->cfg.nkappa, np = c->cfg.nphi;
    for (int m = 0; m < nk; ++m)
        if (!(kappa_f[m + 1] > kappa_f[m]) || !(kappa_f[m] <= kappa_c[m] && kappa_c[m] <= kappa_f[m + 1]))
            return fail(RPL_ERR_ARG, "kappa faces must increase and bracket the centres");
    for (int n = 0; n < np; ++n)
        if (!(phi_f[n + 1] > phi_f[n]))
            return fail(RPL_ERR_ARG, "phi faces must increase");
    RPL_HIP(hipSetDevice(c->device));
    const size_t tab = (size_t)nk * np;
    std::vector<double> w(tab), tr(tab), mx(tab), my(tab), en(tab), cgx(tab), cgy(tab), cs(np), sn(np);
    const double sqrt_g = std::sqrt(gravity);
    for (int n = 0; n < np; ++n) {
        cs[n] = std::cos(phi_c[n]);
        sn[n] = std::sin(phi_c[n]);
        for (int m = 0; m < nk; ++m) {
            const size_t t = (size_t)m + (size_t)nk * n;
            const double k1 = kappa_f[m], k2 = kappa_f[m + 1], p1 = phi_f[n], p2 = phi_f[n + 1];
            // polar_wavevector_grid.jl:44-47: exact polar cell measure.
            w[t] = ((k2 * k2 - k1 * k1) / 2) * (p2 - p1);
            // polar_wavevector_grid.jl:73-115: first moments and trace of the second.
            const double r1 = (k2 * k2 * k2 - k1 * k1 * k1) / 3;
            mx[t] = r1 * (std::sin(p2) - std::sin(p1));
            my[t] = r1 * (std::cos(p1) - std::cos(p2));
            const double r2 = (std::pow(k2, 4) - std::pow(k1, 4)) / 4;
            const double s2 = std::sin(2 * p2) - std::sin(2 * p1);
            tr[t] = r2 * ((p2 - p1) / 2 + s2 / 4) + r2 * ((p2 - p1) / 2 - s2 / 4);
            // bulk_statistics.jl:24-26 + finite_volume_integration.jl:24-27 (power 1/2).
            en[t] = sqrt_g * ((p2 - p1) * (std::pow(k2, 2.5) - std::pow(k1, 2.5)) / 2.5);
            if (depth_kind == RPL_DEPTH_INFINITE) {
                // transport.jl:61-70: cell-averaged deep-water group velocity.
                const double radial = (std::pow(k2, 1.5) - std::pow(k1, 1.5)) / 1.5;
                const double scale = sqrt_g / (2 * w[t]);
                cgx[t] = scale * radial * (std::sin(p2) - std::sin(p1));
                cgy[t] = scale * radial * (std::cos(p1) - std::cos(p2));
            } else {
                // transport.jl:72-83: finite depth, point-sampled.
                const double kk = kappa_c[m];
                const double kx = kk * std::cos(phi_c[n]), ky = kk * std::sin(phi_c[n]);
                if (kk == 0) {
                    cgx[t] = cgy[t] = 0;
                } else {
                    const double mu = kk * depth;
                    const double phase = std::sqrt(gravity * std::tanh(mu) / kk);
                    const double factor = (1 + 2 * mu / std::sinh(2 * mu)) / 2;
                    const double cg = factor * phase;
                    cgx[t] = cg * kx / kk;
                    cgy[t] = cg * ky / kk;
                }
            }
        }
    }
    std::vector<double> kc(kappa_c, kappa_c + nk);
    c->h_kc = kc;
    c->h_kf.assign(kappa_f, kappa_f + nk + 1);
    c->h_pc.assign(phi_c, phi_c + np);
    c->h_pf.assign(phi_f, phi_f + np + 1);
    c->h_mux = mx;
    c->h_muy = my;
    c->has_sources = false;  // source tables depend on the geometry
    c->dia = false;
    c->listed = false;
    if (c->pm) {             // so do the pseudomomentum tables
        c->pm = false;
        c->has_current = false;
    }
    c->gse = false;          // and the smoothing tables
    c->has_depth_field = false;  // and a depth field's group speeds
    int rc;
    if ((rc = upload_real(c, kc.data(), c->kappa, nk)) || (rc = upload_real(c, cgx.data(), c->cgx, tab)) ||
        (rc = upload_real(c, cgy.data(), c->cgy, tab)) || (rc = upload_real(c, cs.data(), c->cosp, np)) ||
        (rc = upload_real(c, sn.data(), c->sinp, np)))
        return rc;
    if ((rc = upload_f64(c, w, c->w)) || (rc = upload_f64(c, tr, c->trace)) || (rc = upload_f64(c, mx, c->mux)) ||
        (rc = upload_f64(c, my, c->muy)) || (rc = upload_f64(c, en, c->energy)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->dx = dx;
    c->dy = dy;
    c->gravity = gravity;
    c->dk = kappa_f[1] - kappa_f[0];  // first(coordinate_spacings(cgrid, 1))
    c->dp = phi_f[1] - phi_f[0];
    c->has_geometry = true;
    return RPL_OK;
}

static int compute_gradients(rpl_ctx *c)
{
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa;
    hipError_t e;
    if (is_f64(c))
        e = rpl::launch_gradients<double>(nx, ny, nk, (const double *)c->Ux, (const double *)c->Uy, 1.0 / c->dx,
                                          1.0 / c->dy, (double *)c->Uxx, (double *)c->Uxy, (double *)c->Uyx,
                                          (double *)c->Uyy, c->stream);
    else
        e = rpl::launch_gradients<float>(nx, ny, nk, (const float *)c->Ux, (const float *)c->Uy,
                                         1.0f / (float)c->dx, 1.0f / (float)c->dy, (float *)c->Uxx,
                                         (float *)c->Uxy, (float *)c->Uyx, (float *)c->Uyy, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("gradients: ") + hipGetErrorString(e));
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->has_current = true;
    return RPL_OK;
}

// q_kernel.jl:77-85 (q_cdf_kernel).
static double q_cdf(double kappa, double z, double depth)
{
    const double small_mu = 1e-6;
    const double mu = kappa * depth;
    const double mus = std::fabs(mu) < small_mu ? (mu < 0 ? -small_mu : small_mu) : mu;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    const double a = 2 * mus * s, b = 2 * mus;
    const double fin = (std::exp(a - b) - std::exp(-a - b)) / (1 - std::exp(-2 * b));
    return std::fabs(mu) < small_mu ? s : fin;
}

// q_kernel.jl:30-33 (q_cell_integral): CDF(zhigh) - CDF(zlow).
static double q_cell(double kappa, double z1, double z2, double depth)
{
    return q_cdf(kappa, std::max(z1, z2), depth) - q_cdf(kappa, std::min(z1, z2), depth);
}

// q_kernel.jl:35-53 (q_cdf_kappa_derivative): d CDF / d kappa of
// sinh(2 mu s) / sinh(2 mu), mu = kappa d, in the reference's form; below
// |mu| = 1e-6 its series depth * 4 mu s (s^2 - 1) / 3.
static double q_cdf_dkappa(double kappa, double z, double depth)
{
    const double mu = kappa * depth;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    if (std::fabs(mu) < 1e-6)
        return depth * (4 * mu * s * (s * s - 1) / 3);
    const double a = std::exp(-2 * mu * (1 - s)), b = std::exp(-2 * mu * (1 + s)), e = std::exp(-4 * mu);
    const double num = a - b, den = 1 - e;
    const double dnum = -2 * (1 - s) * a + 2 * (1 + s) * b, dden = 4 * e;
    const double dF = (dnum * den - num * dden) / (den * den);
    return depth * dF;
}

// q_kernel.jl:55-59 (q_cell_integral_kappa_derivative).
static double q_cell_dkappa(double kappa, double z1, double z2, double depth)
{
    return q_cdf_dkappa(kappa, std::max(z1, z2), depth) - q_cdf_dkappa(kappa, std::min(z1, z2), depth);
}

static std::vector<double> host_kappa(rpl_ctx *c) { return c->h_kc; }

static int ensure_dU(rpl_ctx *c)
{
    const size_t p3 = c->plane * c->cfg.nkappa;
    int rc;
    if (!c->dUx && (rc = alloc(&c->dUx, p3 * c->esize)))
        return rc;
    if (!c->dUy && (rc = alloc(&c->dUy, p3 * c->esize)))
        return rc;
    return RPL_OK;
}

int rpl_set_depth_field(rpl_ctx *c, const double *depth)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    const size_t plane = c->plane;
    if (!depth) {
        c->has_depth_field = false;
        c->h_depth.clear();
        c->has_current = false;  // the coupling's Q depth came from the field
        c->pm = false;
        return RPL_OK;
    }
    for (size_t k = 0; k < plane; ++k)
        if (!(depth[k] > 0) || !std::isfinite(depth[k]))
            return fail(RPL_ERR_ARG, "depth values must be positive");  // lagrangian_velocities.jl:123-124
    const int nk = c->cfg.nkappa, np = c->cfg.nphi;
    const size_t tab = (size_t)nk * np;
    const double g = c->gravity;
    // transport.jl:72-83 at each column's depth: cg = (1 + 2 mu / sinh 2mu) / 2
    // * sqrt(g tanh(mu) / k), mu = k d; the velocity is (cg kx) / k.
    std::vector<double> mag(plane * nk);
    for (int m = 0; m < nk; ++m) {
        const double k = c->h_kc[m];
        for (size_t ij = 0; ij < plane; ++ij) {
            if (k == 0) {
                mag[ij + plane * m] = 0;
                continue;
            }
            const double mu = k * depth[ij];
            const double phase = std::sqrt(g * std::tanh(mu) / k);
            const double factor = (1 + 2 * mu / std::sinh(2 * mu)) / 2;
            mag[ij + plane * m] = factor * phase;
        }
    }
    std::vector<double> vec(3 * tab), dir(2 * tab);
    for (int n = 0; n < np; ++n)
        for (int m = 0; m < nk; ++m) {
            const size_t t = (size_t)m + (size_t)nk * n;
            const double k = c->h_kc[m];
            const double kx = k * std::cos(c->h_pc[n]), ky = k * std::sin(c->h_pc[n]);
            vec[t] = k == 0 ? 0 : kx;
            vec[tab + t] = k == 0 ? 0 : ky;
            vec[2 * tab + t] = k == 0 ? 1 : k;  // a zero wavenumber has zero velocity
            dir[t] = k == 0 ? 0 : kx / k;
            dir[tab + t] = k == 0 ? 0 : ky / k;
        }
    RPL_HIP(hipSetDevice(c->device));
    int rc;
    if ((!c->cgm && (rc = alloc(&c->cgm, plane * nk * c->esize))) ||
        (!c->cg_vec && (rc = alloc(&c->cg_vec, 3 * tab * c->esize))) ||
        (!c->cg_dir && (rc = alloc(&c->cg_dir, 2 * tab * c->esize))))
        return rc;
    c->has_depth_field = false;  // until the tables are on the device
    if ((rc = upload_real(c, mag.data(), c->cgm, plane * nk)) || (rc = upload_real(c, vec.data(), c->cg_vec, 3 * tab)) ||
        (rc = upload_real(c, dir.data(), c->cg_dir, 2 * tab)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->h_depth.assign(depth, depth + plane);
    c->has_depth_field = true;
    c->has_current = false;  // a current set before takes the field's Q depth when set again
    c->pm = false;
    return RPL_OK;
}

int rpl_set_current(rpl_ctx *c, int32_t nz, const double *z_faces, const double *u, const double *v, double q_depth)
{
    if (!c || !z_faces || !u || !v)
        return fail(RPL_ERR_ARG, "null argument");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (nz < 1)
        return fail(RPL_ERR_ARG, "nz must be >= 1");
    if (!(q_depth > 0))
        return fail(RPL_ERR_ARG, "depth must be positive");  // q_kernel.jl:19
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa;
    RPL_HIP(hipSetDevice(c->device));
    // The kappa centres in the context's element type, as the reference's
    // coupling.kappa (current_coupling.jl:80).
    std::vector<double> kc = host_kappa(c);
    if (!is_f64(c))
        for (double &k : kc)
            k = (float)k;
    // q_kernel.jl:87-92: W[k,m] = CDF(zhigh) - CDF(zlow), k fastest, and its
    // kappa derivative (doppler_velocity.jl:25-43, 64-82).
    std::vector<double> W((size_t)nz * nk), dWk((size_t)nz * nk);
    for (int m = 0; m < nk; ++m)
        for (int k = 0; k < nz; ++k) {
            W[k + (size_t)nz * m] = q_cell(kc[m], z_faces[k], z_faces[k + 1], q_depth);
            dWk[k + (size_t)nz * m] = q_cell_dkappa(kc[m], z_faces[k], z_faces[k + 1], q_depth);
        }
    const size_t nu = c->plane * nz;
    void *du = nullptr, *dv = nullptr, *dW = nullptr, *dD = nullptr;
    int rc;
    if ((rc = ensure_dU(c)) || (rc = alloc(&du, nu * c->esize)) || (rc = alloc(&dv, nu * c->esize)) ||
        (rc = alloc(&dW, W.size() * c->esize)) || (rc = alloc(&dD, W.size() * c->esize))) {
        free_ptr(du);
        free_ptr(dv);
        free_ptr(dW);
        return rc;
    }
    rc = upload_real(c, u, du, nu);
    if (!rc)
        rc = upload_real(c, v, dv, nu);
    if (!rc)
        rc = upload_real(c, W.data(), dW, W.size());
    if (!rc)
        rc = upload_real(c, dWk.data(), dD, dWk.size());
    if (!rc && c->has_depth_field) {
        // q_projection_depth of a depth field (lagrangian_velocities.jl:192-193):
        // every column's weights at its own depth, formed on the device
        void *dzf = nullptr, *dkc = nullptr, *ddep = nullptr;
        if ((rc = alloc(&dzf, (nz + 1) * sizeof(double))) || (rc = alloc(&dkc, nk * sizeof(double))) ||
            (rc = alloc(&ddep, c->plane * sizeof(double)))) {
            free_ptr(dzf);
            free_ptr(dkc);
        } else {
            RPL_HIP(hipMemcpy(dzf, z_faces, (nz + 1) * sizeof(double), hipMemcpyHostToDevice));
            RPL_HIP(hipMemcpy(dkc, kc.data(), nk * sizeof(double), hipMemcpyHostToDevice));
            RPL_HIP(hipMemcpy(ddep, c->h_depth.data(), c->plane * sizeof(double), hipMemcpyHostToDevice));
            hipError_t e = hipSuccess;
            for (int d = 0; d < 2 && e == hipSuccess; ++d) {
                if (is_f64(c))
                    e = rpl::launch_doppler_columns<double>(nx, ny, nz, nk, (const double *)du, (const double *)dv,
                                                            (const double *)dzf, (const double *)dkc,
                                                            (const double *)ddep, d == 1,
                                                            (double *)(d ? c->dUx : c->Ux),
                                                            (double *)(d ? c->dUy : c->Uy), c->stream);
                else
                    e = rpl::launch_doppler_columns<float>(nx, ny, nz, nk, (const float *)du, (const float *)dv,
                                                           (const double *)dzf, (const double *)dkc,
                                                           (const double *)ddep, d == 1, (float *)(d ? c->dUx : c->Ux),
                                                           (float *)(d ? c->dUy : c->Uy), c->stream);
            }
            if (e == hipSuccess)
                e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess)
                rc = fail(RPL_ERR_HIP, std::string("doppler (depth per column): ") + hipGetErrorString(e));
            free_ptr(dzf);
            free_ptr(dkc);
            free_ptr(ddep);
        }
    } else if (!rc) {
        hipError_t e;
        if (is_f64(c) && c->cfg.kernel == RPL_KERNEL_FAST)
            e = rpl::launch_doppler_mfma<double>(nx, ny, nz, nk, (const double *)du, (const double *)dv,
                                                 (const double *)dW, (double *)c->Ux, (double *)c->Uy, c->stream);
        else if (is_f64(c))
            e = rpl::launch_doppler(nx, ny, nz, nk, (const double *)du, (const double *)dv, (const double *)dW,
                                    (double *)c->Ux, (double *)c->Uy, c->stream);
        else
            e = rpl::launch_doppler(nx, ny, nz, nk, (const float *)du, (const float *)dv, (const float *)dW,
                                    (float *)c->Ux, (float *)c->Uy, c->stream);
        // dU/dkappa: the same contraction with the derivative weights.
        if (e == hipSuccess && is_f64(c) && c->cfg.kernel == RPL_KERNEL_FAST)
            e = rpl::launch_doppler_mfma<double>(nx, ny, nz, nk, (const double *)du, (const double *)dv,
                                                 (const double *)dD, (double *)c->dUx, (double *)c->dUy, c->stream);
        else if (e == hipSuccess && is_f64(c))
            e = rpl::launch_doppler(nx, ny, nz, nk, (const double *)du, (const double *)dv, (const double *)dD,
                                    (double *)c->dUx, (double *)c->dUy, c->stream);
        else if (e == hipSuccess)
            e = rpl::launch_doppler(nx, ny, nz, nk, (const float *)du, (const float *)dv, (const float *)dD,
                                    (float *)c->dUx, (float *)c->dUy, c->stream);
        if (e != hipSuccess)
            rc = fail(RPL_ERR_HIP, std::string("doppler: ") + hipGetErrorString(e));
        else if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
            rc = fail(RPL_ERR_HIP, std::string("doppler: ") + hipGetErrorString(e));
    }
    free_ptr(du);
    free_ptr(dv);
    free_ptr(dW);
    free_ptr(dD);
    if (rc)
        return rc;
    c->pm = false;
    c->has_dU = true;
    return compute_gradients(c);
}

int rpl_set_pseudomomentum_coupling(rpl_ctx *c, int32_t nz, const double *z_faces, double q_depth)
{
    if (!c || !z_faces)
        return fail(RPL_ERR_ARG, "null argument");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (c->cfg.transport != RPL_TRANSPORT_REFRACTION)
        return fail(RPL_ERR_STATE, "the pseudomomentum coupling drives the refraction transport (RPL_TRANSPORT_REFRACTION)");
    if (nz < 1)
        return fail(RPL_ERR_ARG, "nz must be >= 1");
    if (!c->has_depth_field && !(q_depth > 0))
        return fail(RPL_ERR_ARG, "depth must be positive");  // q_kernel.jl:19
    for (int k = 0; k < nz; ++k)
        if (!(z_faces[k] != z_faces[k + 1]))
            return fail(RPL_ERR_ARG, "vertical cells must have nonzero thickness");
    const int nk = c->cfg.nkappa;
    RPL_HIP(hipSetDevice(c->device));
    std::vector<double> kc = host_kappa(c);
    if (!is_f64(c))
        for (double &k : kc)
            k = (float)k;
    // pseudomomentum.jl:89-122: overlap O[t,s] = sum_k W_t W_s / |dz| and its
    // derivative in the target kappa, target fastest.
    const size_t o_elems = (c->has_depth_field ? c->plane : 1) * (size_t)nk * nk;
    std::vector<double> O(c->has_depth_field ? 0 : (size_t)nk * nk), dO(O.size());
    for (int src = 0; src < nk && !c->has_depth_field; ++src)
        for (int t = 0; t < nk; ++t) {
            double total = 0, dtotal = 0;
            for (int k = 0; k < nz; ++k) {
                const double z1 = z_faces[k], z2 = z_faces[k + 1], dz = std::fabs(z2 - z1);
                const double sw = q_cell(kc[src], z1, z2, q_depth);
                const double tw = q_cell(kc[t], z1, z2, q_depth);
                const double td = q_cell_dkappa(kc[t], z1, z2, q_depth);
                total += tw * sw / dz;
                dtotal += td * sw / dz;
            }
            O[t + (size_t)nk * src] = total;
            dO[t + (size_t)nk * src] = dtotal;
        }
    const size_t tab = (size_t)nk * c->cfg.nphi, p3 = c->plane * nk;
    int rc;
    if ((rc = ensure_dU(c)))
        return rc;
    if (c->pm_O_elems < o_elems) {  // per-column tables need more room
        free_ptr(c->pm_O);
        free_ptr(c->pm_dO);
        c->pm_O = c->pm_dO = nullptr;
        c->pm_O_elems = 0;
    }
    for (void **b : {&c->pm_O, &c->pm_dO})
        if (!*b && (rc = alloc(b, o_elems * c->esize)))
            return rc;
    c->pm_O_elems = std::max(c->pm_O_elems, o_elems);
    for (void **b : {&c->pm_kx, &c->pm_ky})
        if (!*b && (rc = alloc(b, tab * c->esize)))
            return rc;
    c->pm = false;  // until the tables are on the device
    if ((rc = upload_real(c, c->h_mux.data(), c->pm_kx, tab)) || (rc = upload_real(c, c->h_muy.data(), c->pm_ky, tab)))
        return rc;
    if (c->has_depth_field) {
        // pseudomomentum.jl:72-87: the tables of every column at its depth
        void *dzf = nullptr, *dkc = nullptr, *ddep = nullptr;
        if ((rc = alloc(&dzf, (nz + 1) * sizeof(double))) || (rc = alloc(&dkc, nk * sizeof(double))) ||
            (rc = alloc(&ddep, c->plane * sizeof(double)))) {
            free_ptr(dzf);
            free_ptr(dkc);
            return rc;
        }
        hipError_t e = hipMemcpy(dzf, z_faces, (nz + 1) * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(dkc, kc.data(), nk * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(ddep, c->h_depth.data(), c->plane * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = is_f64(c) ? rpl::launch_pm_overlap_columns<double>(c->plane, nz, nk, (const double *)dzf,
                                                                   (const double *)dkc, (const double *)ddep,
                                                                   (double *)c->pm_O, (double *)c->pm_dO, c->stream)
                          : rpl::launch_pm_overlap_columns<float>(c->plane, nz, nk, (const double *)dzf,
                                                                  (const double *)dkc, (const double *)ddep,
                                                                  (float *)c->pm_O, (float *)c->pm_dO, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        free_ptr(dzf);
        free_ptr(dkc);
        free_ptr(ddep);
        if (e != hipSuccess)
            return fail(RPL_ERR_HIP, std::string("pseudomomentum overlap tables: ") + hipGetErrorString(e));
    } else if ((rc = upload_real(c, O.data(), c->pm_O, O.size())) ||
               (rc = upload_real(c, dO.data(), c->pm_dO, dO.size()))) {
        return rc;
    }
    c->pm_percol = c->has_depth_field;
    // A new coupling holds zero velocities until its first update
    // (current_coupling.jl:104-109).
    for (void *b : {c->Ux, c->Uy, c->dUx, c->dUy})
        RPL_HIP(hipMemsetAsync(b, 0, p3 * c->esize, c->stream));
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->pm = true;
    c->has_dU = true;
    c->has_current = true;
    return RPL_OK;
}

int rpl_update_coupling(rpl_ctx *c)
{
    int rc = check_ready(c);
    if (rc)
        return rc;
    if ((rc = launch_coupling(c, c->A)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    return RPL_OK;
}

int rpl_set_topology(rpl_ctx *c, int32_t periodic_x, int32_t periodic_y)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if ((periodic_x != 0 && periodic_x != 1) || (periodic_y != 0 && periodic_y != 1))
        return fail(RPL_ERR_ARG, "periodic_x and periodic_y must be 0 (Bounded) or 1 (Periodic)");
    if (c->cfg.transport == RPL_TRANSPORT_INTRINSIC && !(periodic_x && periodic_y))
        return fail(RPL_ERR_UNSUPPORTED, "the fused intrinsic transport kernel runs on Periodic x Periodic grids "
                                         "only (the reference uses Oceananigans' per-bin advection otherwise)");
    ++c->generation;
    c->periodic_x = periodic_x != 0;
    c->periodic_y = periodic_y != 0;
    return RPL_OK;
}

int rpl_set_propagation_smoothing(rpl_ctx *c, int32_t enable, double alpha_s, double alpha_n)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (!enable) {
        c->gse = false;
        return RPL_OK;
    }
    if (!(alpha_s >= 0) || !(alpha_n >= 0) || !std::isfinite(alpha_s) || !std::isfinite(alpha_n))
        return fail(RPL_ERR_ARG, "SpatialAveraging alphas must be finite and non-negative");
    const int nk = c->cfg.nkappa;
    RPL_HIP(hipSetDevice(c->device));
    // propagation_smoothing.jl:117-135 in the element type: cg = 1/2 sqrt(g/kappa)
    // with g = 9.81, dcg = |cg[m-1] - cg[m+1]| / 2 (the bin itself at the ends).
    std::vector<double> cg(nk), dcg(nk);
    if (is_f64(c)) {
        auto speed = [&](int m) { return 1.0 / 2.0 * std::sqrt(9.81 / c->h_kc[m]); };
        for (int m = 0; m < nk; ++m)
            cg[m] = speed(m);
        for (int m = 0; m < nk; ++m)
            dcg[m] = std::fabs((m == 0 ? cg[0] : speed(m - 1)) - (m == nk - 1 ? cg[nk - 1] : speed(m + 1))) / 2.0;
    } else {
        auto speed = [&](int m) { return 1.0f / 2.0f * std::sqrt(9.81f / (float)c->h_kc[m]); };
        for (int m = 0; m < nk; ++m)
            cg[m] = speed(m);
        for (int m = 0; m < nk; ++m)
            dcg[m] = std::fabs((m == 0 ? (float)cg[0] : speed(m - 1)) - (m == nk - 1 ? (float)cg[nk - 1] : speed(m + 1))) /
                     2.0f;
    }
    int rc;
    for (void **b : {&c->gse_cg, &c->gse_dcg})
        if (!*b && (rc = alloc(b, nk * c->esize)))
            return rc;
    if ((rc = upload_real(c, cg.data(), c->gse_cg, nk)) || (rc = upload_real(c, dcg.data(), c->gse_dcg, nk)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->gse_as = alpha_s;
    c->gse_an = alpha_n;
    c->gse_cg_max = *std::max_element(cg.begin(), cg.end());
    c->gse_dcg_max = *std::max_element(dcg.begin(), dcg.end());
    c->gse = true;
    return RPL_OK;
}

int rpl_get_doppler_kappa_derivative(rpl_ctx *c, double *dUx, double *dUy)
{
    if (!c || !dUx || !dUy)
        return fail(RPL_ERR_ARG, "null argument");
    if (!c->has_dU)
        return fail(RPL_ERR_STATE, "no coupling has set dU/dkappa (rpl_set_current or rpl_set_pseudomomentum_coupling)");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    const size_t p3 = c->plane * c->cfg.nkappa;
    if (is_f64(c)) {
        RPL_HIP(hipMemcpy(dUx, c->dUx, p3 * 8, hipMemcpyDeviceToHost));
        RPL_HIP(hipMemcpy(dUy, c->dUy, p3 * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(p3);
        RPL_HIP(hipMemcpy(t.data(), c->dUx, p3 * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < p3; ++k)
            dUx[k] = t[k];
        RPL_HIP(hipMemcpy(t.data(), c->dUy, p3 * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < p3; ++k)
            dUy[k] = t[k];
    }
    return RPL_OK;
}

// The pseudomomentum cell contraction of a device action field N (the
// action itself, or its tendency), scaled by coefficient when it is not 1.
static int pseudomomentum_cells(rpl_ctx *c, const void *Nsrc, int32_t nz, const double *z_faces, double q_depth,
                                int32_t cell_average, double coefficient, double *px, double *py);

int rpl_pseudomomentum(rpl_ctx *c, int32_t nz, const double *z_faces, double q_depth, int32_t cell_average,
                       double *px, double *py)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null argument");
    return pseudomomentum_cells(c, c->A, nz, z_faces, q_depth, cell_average, 1.0, px, py);
}

int rpl_pseudomomentum_tendency(rpl_ctx *c, const void *old_N, double dt, int32_t nz, const double *z_faces,
                                double q_depth, double coefficient, double *ptx, double *pty)
{
    if (!c || !old_N)
        return fail(RPL_ERR_ARG, "null argument");
    if (!(dt > 0))
        return fail(RPL_ERR_ARG, "pseudomomentum tendency timestep must be positive");  // pseudomomentum.jl:225
    RPL_HIP(hipSetDevice(c->device));
    // The stage buffers are free between steps: old action into B, tendency into C.
    RPL_HIP(hipMemcpyAsync(c->B, old_N, c->cnt * c->esize, hipMemcpyHostToDevice, c->stream));
    hipError_t e = is_f64(c) ? rpl::launch_action_tendency<double>(c->cnt, (const double *)c->A, (const double *)c->B,
                                                                    dt, (double *)c->C, c->stream)
                             : rpl::launch_action_tendency<float>(c->cnt, (const float *)c->A, (const float *)c->B,
                                                                   (float)dt, (float *)c->C, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("action tendency: ") + hipGetErrorString(e));
    return pseudomomentum_cells(c, c->C, nz, z_faces, q_depth, 1, coefficient, ptx, pty);
}

static int pseudomomentum_cells(rpl_ctx *c, const void *Nsrc, int32_t nz, const double *z_faces, double q_depth,
                                int32_t cell_average, double coefficient, double *px, double *py)
{
    if (!c || !z_faces || !px || !py)
        return fail(RPL_ERR_ARG, "null argument");
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (nz < 1)
        return fail(RPL_ERR_ARG, "nz must be >= 1");
    // q_depth = 0 with a depth field: every column at its own depth
    const bool percol = q_depth == 0 && c->has_depth_field;
    if (!percol && !(q_depth > 0))
        return fail(RPL_ERR_ARG, "depth must be positive");
    for (int k = 0; k < nz; ++k)
        if (!(z_faces[k] != z_faces[k + 1]))
            return fail(RPL_ERR_ARG, "vertical cells must have nonzero thickness");
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    RPL_HIP(hipSetDevice(c->device));
    std::vector<double> kc = host_kappa(c);
    if (!is_f64(c))
        for (double &k : kc)
            k = (float)k;
    // pseudomomentum.jl:375-400: W[k,m] (k fastest) and 1/|dz_k|.
    std::vector<double> W(percol ? 0 : (size_t)nz * nk), sc(nz);
    for (int m = 0; m < nk && !percol; ++m)
        for (int k = 0; k < nz; ++k)
            W[k + (size_t)nz * m] = q_cell(kc[m], z_faces[k], z_faces[k + 1], q_depth);
    for (int k = 0; k < nz; ++k)
        sc[k] = 1 / std::fabs(z_faces[k + 1] - z_faces[k]);
    const size_t tab = (size_t)nk * np, out = c->plane * nz;
    void *dW = nullptr, *dS = nullptr, *dkx = nullptr, *dky = nullptr, *dpx = nullptr, *dpy = nullptr;
    int rc = RPL_OK;
    if ((rc = alloc(&dW, (percol ? c->plane * nz * nk : W.size()) * c->esize)) ||
        (rc = alloc(&dS, sc.size() * c->esize)) ||
        (rc = alloc(&dkx, tab * c->esize)) || (rc = alloc(&dky, tab * c->esize)) ||
        (rc = alloc(&dpx, out * c->esize)) || (rc = alloc(&dpy, out * c->esize)))
        goto done;
    if ((!percol && (rc = upload_real(c, W.data(), dW, W.size()))) || (rc = upload_real(c, sc.data(), dS, sc.size())) ||
        (rc = upload_real(c, c->h_mux.data(), dkx, tab)) || (rc = upload_real(c, c->h_muy.data(), dky, tab)))
        goto done;
    if (percol) {
        // q_depth_at(depth, i, j) (pseudomomentum.jl:375-400): each column's weights
        void *dzf = nullptr, *dkc = nullptr, *ddep = nullptr;
        hipError_t e = hipMalloc(&dzf, (nz + 1) * sizeof(double));
        if (e == hipSuccess)
            e = hipMalloc(&dkc, nk * sizeof(double));
        if (e == hipSuccess)
            e = hipMalloc(&ddep, c->plane * sizeof(double));
        if (e == hipSuccess)
            e = hipMemcpy(dzf, z_faces, (nz + 1) * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(dkc, kc.data(), nk * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(ddep, c->h_depth.data(), c->plane * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = is_f64(c) ? rpl::launch_q_weights_columns<double>(c->plane, nz, nk, (const double *)dzf,
                                                                  (const double *)dkc, (const double *)ddep,
                                                                  (double *)dW, c->stream)
                          : rpl::launch_q_weights_columns<float>(c->plane, nz, nk, (const double *)dzf,
                                                                 (const double *)dkc, (const double *)ddep,
                                                                 (float *)dW, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        (void)hipFree(dzf);
        (void)hipFree(dkc);
        (void)hipFree(ddep);
        if (e != hipSuccess) {
            rc = fail(RPL_ERR_HIP, std::string("pseudomomentum weights: ") + hipGetErrorString(e));
            goto done;
        }
    }
    {
        hipError_t e;
        const bool ordered = pm_ordered(c);
        if (is_f64(c))
            e = rpl::launch_pm_contract<double>(ordered, nx, ny, nk, np, nz, (const double *)Nsrc,
                                                (const double *)dW,
                                                nullptr, (const double *)dkx, (const double *)dky,
                                                cell_average ? (const double *)dS : nullptr, (double *)dpx,
                                                (double *)dpy, nullptr, nullptr, c->stream, percol);
        else
            e = rpl::launch_pm_contract<float>(ordered, nx, ny, nk, np, nz, (const float *)Nsrc,
                                               (const float *)dW,
                                               nullptr, (const float *)dkx, (const float *)dky,
                                               cell_average ? (const float *)dS : nullptr, (float *)dpx,
                                               (float *)dpy, nullptr, nullptr, c->stream, percol);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            rc = fail(RPL_ERR_HIP, std::string("pseudomomentum: ") + hipGetErrorString(e));
            goto done;
        }
        // _scale_field_pair_kernel! (pseudomomentum.jl:430-436) in the element
        // type, when a coefficient other than one is asked for.
        if (is_f64(c)) {
            if ((e = hipMemcpy(px, dpx, out * 8, hipMemcpyDeviceToHost)) == hipSuccess)
                e = hipMemcpy(py, dpy, out * 8, hipMemcpyDeviceToHost);
            if (coefficient != 1.0)
                for (size_t k = 0; k < out; ++k) {
                    px[k] *= coefficient;
                    py[k] *= coefficient;
                }
        } else {
            const float cf = (float)coefficient;
            std::vector<float> t(out);
            if ((e = hipMemcpy(t.data(), dpx, out * 4, hipMemcpyDeviceToHost)) == hipSuccess)
                for (size_t k = 0; k < out; ++k)
                    px[k] = coefficient != 1.0 ? t[k] * cf : t[k];
            if (e == hipSuccess && (e = hipMemcpy(t.data(), dpy, out * 4, hipMemcpyDeviceToHost)) == hipSuccess)
                for (size_t k = 0; k < out; ++k)
                    py[k] = coefficient != 1.0 ? t[k] * cf : t[k];
        }
        if (e != hipSuccess)
            rc = fail(RPL_ERR_HIP, std::string("pseudomomentum copy: ") + hipGetErrorString(e));
    }
done:
    for (void *p : {dW, dS, dkx, dky, dpx, dpy})
        free_ptr(p);
    return rc;
}

int rpl_set_doppler(rpl_ctx *c, const double *Ux, const double *Uy)
{
    if (!c || !Ux || !Uy)
        return fail(RPL_ERR_ARG, "null argument");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    RPL_HIP(hipSetDevice(c->device));
    const size_t p3 = c->plane * c->cfg.nkappa;
    int rc;
    if ((rc = upload_real(c, Ux, c->Ux, p3)) || (rc = upload_real(c, Uy, c->Uy, p3)))
        return rc;
    c->pm = false;      // the velocities are now prescribed ...
    c->has_dU = false;  // ... with no kappa derivative
    return compute_gradients(c);
}

int rpl_get_doppler(rpl_ctx *c, double *Ux, double *Uy)
{
    if (!c || !Ux || !Uy)
        return fail(RPL_ERR_ARG, "null argument");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    const size_t p3 = c->plane * c->cfg.nkappa;
    if (is_f64(c)) {
        RPL_HIP(hipMemcpy(Ux, c->Ux, p3 * 8, hipMemcpyDeviceToHost));
        RPL_HIP(hipMemcpy(Uy, c->Uy, p3 * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> t(p3);
        RPL_HIP(hipMemcpy(t.data(), c->Ux, p3 * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < p3; ++k)
            Ux[k] = t[k];
        RPL_HIP(hipMemcpy(t.data(), c->Uy, p3 * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < p3; ++k)
            Uy[k] = t[k];
    }
    return RPL_OK;
}

int rpl_set_action(rpl_ctx *c, const void *N)
{
    if (!c || !N)
        return fail(RPL_ERR_ARG, "null argument");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipMemcpyAsync(c->A, N, c->cnt * c->esize, hipMemcpyHostToDevice, c->stream));
    RPL_HIP(hipStreamSynchronize(c->stream));
    return RPL_OK;
}

int rpl_get_action(rpl_ctx *c, void *N)
{
    if (!c || !N)
        return fail(RPL_ERR_ARG, "null argument");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    RPL_HIP(hipMemcpy(N, c->A, c->cnt * c->esize, hipMemcpyDeviceToHost));
    return RPL_OK;
}

int rpl_tendency(rpl_ctx *c, void *G)
{
    int rc = check_ready(c);
    if (rc)
        return rc;
    if (!G)
        return fail(RPL_ERR_ARG, "null argument");
    if ((rc = launch_wind(c)) || (rc = prepare_stage(c, c->A)) ||
        (rc = launch_stage(c, 0.0, rpl::STAGE_TENDENCY, c->A, c->A, c->B)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    RPL_HIP(hipMemcpy(G, c->B, c->cnt * c->esize, hipMemcpyDeviceToHost));
    return RPL_OK;
}

int rpl_time_step(rpl_ctx *c, double dt, int32_t nsteps)
{
    int rc = check_ready(c);
    if (rc)
        return rc;
    if (!(dt > 0))
        return fail(RPL_ERR_ARG, "time step must be positive");  // time_step.jl:120
    if (nsteps < 0)
        return fail(RPL_ERR_ARG, "nsteps must be >= 0");
    for (int s = 0; s < nsteps; ++s) {
        if ((rc = launch_wind(c)))
            return rc;
        if (c->cfg.timestepper == RPL_FORWARD_EULER) {
            if ((rc = prepare_stage(c, c->A)) || (rc = launch_stage(c, dt, rpl::STAGE_1, c->A, c->A, c->B)))
                return rc;
            std::swap(c->A, c->B);
        } else {
            // SSP-RK3 over three buffers; A holds N0 and receives N3 in place.
            if ((rc = prepare_stage(c, c->A)) || (rc = launch_stage(c, dt, rpl::STAGE_1, c->A, c->A, c->B)) ||
                (rc = prepare_stage(c, c->B)) || (rc = launch_stage(c, dt, rpl::STAGE_2, c->B, c->A, c->C)) ||
                (rc = prepare_stage(c, c->C)) || (rc = launch_stage(c, dt, rpl::STAGE_3, c->C, c->A, c->A)))
                return rc;
        }
        if ((rc = launch_smoothing(c, dt)))  // apply_propagation_smoothing!, time_step.jl:167
            return rc;
        c->time += dt;  // tick!(clock, dt), once per step
        c->iteration += 1;
        if (c->nf_interval > 0 && c->iteration % c->nf_interval == 0 && (rc = launch_nonfinite(c)))
            return rc;
    }
    return RPL_OK;
}

int rpl_set_nonfinite_check(rpl_ctx *c, int32_t interval)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (interval < 0)
        return fail(RPL_ERR_ARG, "the non-finite check interval must be >= 0");
    RPL_HIP(hipSetDevice(c->device));
    int rc;
    if (interval > 0 && !c->nf_buf && (rc = alloc(&c->nf_buf, 3 * sizeof(long long))))
        return rc;
    if (c->nf_buf) {
        const long long init[3] = {0, -1, 0};  // counter, first iteration, count
        RPL_HIP(hipStreamSynchronize(c->stream));
        RPL_HIP(hipMemcpy(c->nf_buf, init, sizeof init, hipMemcpyHostToDevice));
    }
    c->nf_interval = interval;
    return RPL_OK;
}

int rpl_get_nonfinite(rpl_ctx *c, int64_t *first_iteration, int64_t *count)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    long long rec[3] = {0, -1, 0};
    if (c->nf_buf) {
        RPL_HIP(hipSetDevice(c->device));
        RPL_HIP(hipStreamSynchronize(c->stream));
        RPL_HIP(hipMemcpy(rec, c->nf_buf, sizeof rec, hipMemcpyDeviceToHost));
    }
    if (first_iteration)
        *first_iteration = rec[1];
    if (count)
        *count = rec[2];
    return RPL_OK;
}

int rpl_synchronize(rpl_ctx *c)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    return RPL_OK;
}

int rpl_get_clock(rpl_ctx *c, double *time, int64_t *iteration)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (time)
        *time = c->time;
    if (iteration)
        *iteration = c->iteration;
    return RPL_OK;
}

int rpl_set_clock(rpl_ctx *c, double time, int64_t iteration)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    c->time = time;
    c->iteration = iteration;
    return RPL_OK;
}

int rpl_diagnostics(rpl_ctx *c, double *m0, double *kappa_rms, double *mean_dir)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    RPL_HIP(hipSetDevice(c->device));
    int rc;
    if (!c->diag && (rc = alloc((void **)&c->diag, 3 * c->plane * 8)))
        return rc;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    double *d0 = c->diag, *d1 = c->diag + c->plane, *d2 = c->diag + 2 * c->plane;
    hipError_t e;
    if (is_f64(c))
        e = rpl::launch_column_diagnostics<double>(nx, ny, nk, np, (const double *)c->A, c->w, c->trace, c->mux,
                                                   c->muy, d0, d1, d2, c->stream);
    else
        e = rpl::launch_column_diagnostics<float>(nx, ny, nk, np, (const float *)c->A, c->w, c->trace, c->mux,
                                                  c->muy, d0, d1, d2, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("diagnostics: ") + hipGetErrorString(e));
    RPL_HIP(hipStreamSynchronize(c->stream));
    if (m0)
        RPL_HIP(hipMemcpy(m0, d0, c->plane * 8, hipMemcpyDeviceToHost));
    if (kappa_rms)
        RPL_HIP(hipMemcpy(kappa_rms, d1, c->plane * 8, hipMemcpyDeviceToHost));
    if (mean_dir)
        RPL_HIP(hipMemcpy(mean_dir, d2, c->plane * 8, hipMemcpyDeviceToHost));
    return RPL_OK;
}

static int weighted_total(rpl_ctx *c, const double *measure, double *result)
{
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    const int blocks = rpl::weighted_total_blocks(nx, ny);
    const double area = c->dx * c->dy;
    hipError_t e;
    if (is_f64(c))
        e = rpl::launch_weighted_total<double>(nx, ny, nk, np, (const double *)c->A, measure, area, c->partials,
                                               blocks, c->stream);
    else
        e = rpl::launch_weighted_total<float>(nx, ny, nk, np, (const float *)c->A, measure, area, c->partials,
                                              blocks, c->stream);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("weighted total: ") + hipGetErrorString(e));
    std::vector<double> part(2 * (size_t)blocks);
    RPL_HIP(hipStreamSynchronize(c->stream));
    RPL_HIP(hipMemcpy(part.data(), c->partials, part.size() * 8, hipMemcpyDeviceToHost));
    double hi = 0, lo = 0;
    for (int b = 0; b < blocks; ++b) {
        const double x = part[2 * b];
        const double s = hi + x, bv = s - hi, av = s - bv;
        lo += (hi - av) + (x - bv) + part[2 * b + 1];
        hi = s;
    }
    *result = hi + lo;
    return RPL_OK;
}

int rpl_conservation(rpl_ctx *c, double *total_action, double *total_energy)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    RPL_HIP(hipSetDevice(c->device));
    int rc;
    if (total_action && (rc = weighted_total(c, c->w, total_action)))
        return rc;
    if (total_energy && (rc = weighted_total(c, c->energy, total_energy)))
        return rc;
    return RPL_OK;
}

int rpl_weighted_total_noalloc(const void *N, int32_t dtype, int32_t nx, int32_t ny, int32_t nkappa, int32_t nphi,
                               const double *measure, double area, double *out, void *stream)
{
    if (!N || !measure || !out)
        return fail(RPL_ERR_ARG, "null argument");
    if (dtype != RPL_F32 && dtype != RPL_F64)
        return fail(RPL_ERR_ARG, "dtype must be RPL_F32 or RPL_F64");
    if (nx < 1 || ny < 1 || nkappa < 1 || nphi < 1)
        return fail(RPL_ERR_ARG, "nx, ny, nkappa and nphi must be positive");
    if ((int64_t)nx * ny >= ((int64_t)1 << 30) || (int64_t)nkappa * nphi >= ((int64_t)1 << 30))
        return fail(RPL_ERR_ARG, "rpl_weighted_total_noalloc: nx * ny and nkappa * nphi must each stay below 2^30");
    if (!std::isfinite(area))
        return fail(RPL_ERR_ARG, "area must be finite");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = dtype == RPL_F64
                       ? rpl::launch_weighted_total_noalloc<double>(nx, ny, nkappa, nphi, (const double *)N, measure,
                                                                    area, out, s)
                       : rpl::launch_weighted_total_noalloc<float>(nx, ny, nkappa, nphi, (const float *)N, measure,
                                                                   area, out, s);
    if (e != hipSuccess)
        return fail(RPL_ERR_HIP, std::string("weighted total (no allocation): ") + hipGetErrorString(e));
    return RPL_OK;
}

// symmetric_quadruplet.jl:54-77 (_bilinear_lookup), 0-based bins.
static bool dia_lookup(double s, double th, double smin, double smax, double log_xfr, double phimin, double dphi,
                       int nk, int nphi, int bins[4], double w[4], int *mlo, int *mhi)
{
    if (s < smin || s > smax)
        return false;
    const double mr = 1.0 + std::log(s / smin) / log_xfr;
    const long t = (long)mr;  // unsafe_trunc
    const int ml = (int)std::min<long>(std::max<long>(t, 1), nk - 1);
    const double wf = mr - ml;
    const double twopi = 2 * M_PI;
    double r = std::fmod(th - phimin, twopi);  // Julia mod: the sign of the divisor
    if (r != 0 && (r < 0) != (twopi < 0))
        r += twopi;
    const double nr = r / dphi;
    const long nraw = (long)nr;
    long nm = nraw % nphi;
    if (nm < 0)
        nm += nphi;
    const int nl = (int)nm;                        // 0-based n_lo
    const int nh = nl + 1 == nphi ? 0 : nl + 1;    // 0-based n_hi
    const double wt = nr - (double)nraw;
    w[0] = (1 - wf) * (1 - wt);
    w[1] = wf * (1 - wt);
    w[2] = (1 - wf) * wt;
    w[3] = wf * wt;
    *mlo = ml - 1;
    *mhi = ml;
    bins[0] = *mlo + nk * nl;
    bins[1] = *mhi + nk * nl;
    bins[2] = *mlo + nk * nh;
    bins[3] = *mhi + nk * nh;
    return true;
}

// The column-independent DIA tables (symmetric_quadruplet.jl:38-49, 82-208).
static int build_dia_tables(rpl_ctx *c, const rpl_sources &s)
{
    c->dia = false;
    if (s.nonlinear != RPL_NONLINEAR_SYMMETRIC_QUADRUPLET)
        return RPL_OK;
    const int nk = c->cfg.nkappa, np = c->cfg.nphi, nb = nk * np;
    const double lam = s.dia_frequency_offset, g = s.dia_gravity;
    const double a = (1 + lam) * (1 + lam), b = (1 - lam) * (1 - lam);
    const double dtp = std::acos(std::min(1.0, std::max(-1.0, (a * a - b * b + 4) / (4 * a))));
    const double dtm = std::acos(std::min(1.0, std::max(-1.0, (b * b - a * a + 4) / (4 * b))));
    const double ceff = s.dia_coefficient / std::pow(g, 4);
    std::vector<double> kc = c->h_kc;
    if (!is_f64(c))
        for (double &k : kc)
            k = (float)k;
    auto energy_factor = [&](double sigma) { return 4 * M_PI * (sigma * sigma) / g; };
    const double smin = std::sqrt(g * kc[0]), smax = std::sqrt(g * kc[nk - 1]);
    const double log_xfr = std::log(std::sqrt(g * kc[1]) / std::sqrt(g * kc[0]));
    const double phimin = c->h_pc[0], dphi = 2 * M_PI / np;
    std::vector<int> flags(2 * (size_t)nb, 0), bins(16 * (size_t)nb, 0);
    std::vector<double> wef(16 * (size_t)nb, 0), w(16 * (size_t)nb, 0), ef(16 * (size_t)nb, 1), efa(nb), cfa(nb),
        inv_efa(nb), wdiv(16 * (size_t)nb, 0);
    for (int n = 0; n < np; ++n)
        for (int m = 0; m < nk; ++m) {
            const int d = m + nk * n;
            const double sa = std::sqrt(g * kc[m]);
            efa[d] = energy_factor(sa);
            inv_efa[d] = 1 / efa[d];
            cfa[d] = ceff * std::pow(sa / (2 * M_PI), 11);
            const double sp = (1 + lam) * sa, sm = (1 - lam) * sa, tha = c->h_pc[n];
            for (int o = 0; o < 2; ++o) {
                const int sign = o == 0 ? 1 : -1, q = 2 * d + o;
                int mlo, mhi;
                double wq[4];
                if (dia_lookup(sp, tha + sign * dtp, smin, smax, log_xfr, phimin, dphi, nk, np, &bins[8 * (size_t)q], wq, &mlo,
                               &mhi)) {
                    flags[q] |= 1;
                    const double e1 = energy_factor(std::sqrt(g * kc[mlo])), e2 = energy_factor(std::sqrt(g * kc[mhi]));
                    const double efk[4] = {e1, e2, e1, e2};
                    for (int k = 0; k < 4; ++k) {
                        wef[8 * (size_t)q + k] = wq[k] * efk[k];
                        w[8 * (size_t)q + k] = wq[k];
                        ef[8 * (size_t)q + k] = efk[k];
                    }
                }
                if (dia_lookup(sm, tha - sign * dtm, smin, smax, log_xfr, phimin, dphi, nk, np, &bins[8 * (size_t)q + 4], wq,
                               &mlo, &mhi)) {
                    flags[q] |= 2;
                    const double e1 = energy_factor(std::sqrt(g * kc[mlo])), e2 = energy_factor(std::sqrt(g * kc[mhi]));
                    const double efk[4] = {e1, e2, e1, e2};
                    for (int k = 0; k < 4; ++k) {
                        wef[8 * (size_t)q + 4 + k] = wq[k] * efk[k];
                        w[8 * (size_t)q + 4 + k] = wq[k];
                        ef[8 * (size_t)q + 4 + k] = efk[k];
                    }
                }
            }
        }
    int rc;
    for (void **p : {&c->dia_flags})
        if (!*p && (rc = alloc(p, 2 * (size_t)nb * sizeof(int))))
            return rc;
    for (void **p : {&c->dia_bins})
        if (!*p && (rc = alloc(p, 16 * (size_t)nb * sizeof(int))))
            return rc;
    for (size_t k = 0; k < wdiv.size(); ++k)
        wdiv[k] = w[k] / ef[k];
    for (void **p : {&c->dia_wef, &c->dia_wdiv})
        if (!*p && (rc = alloc(p, 16 * (size_t)nb * c->esize)))
            return rc;
    for (void **p : {&c->dia_efa, &c->dia_inv_efa, &c->dia_cfa})
        if (!*p && (rc = alloc(p, (size_t)nb * c->esize)))
            return rc;
    if (!c->dia_add && (rc = alloc(&c->dia_add, c->cnt * c->esize)))
        return rc;
    RPL_HIP(hipMemcpy(c->dia_flags, flags.data(), flags.size() * sizeof(int), hipMemcpyHostToDevice));
    RPL_HIP(hipMemcpy(c->dia_bins, bins.data(), bins.size() * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = upload_real(c, wef.data(), c->dia_wef, wef.size())) ||
        (rc = upload_real(c, wdiv.data(), c->dia_wdiv, wdiv.size())) || (rc = upload_real(c, efa.data(), c->dia_efa, nb)) ||
        (rc = upload_real(c, inv_efa.data(), c->dia_inv_efa, nb)) || (rc = upload_real(c, cfa.data(), c->dia_cfa, nb)))
        return rc;
    // The scatter transposed, per receiving bin: every receiver corner and
    // the donor's own loss (twice when both receivers are in range).
    std::vector<std::vector<std::pair<int, double>>> into(nb);
    for (int d = 0; d < nb; ++d)
        for (int o = 0; o < 2; ++o) {
            const int q = 2 * d + o, fl = flags[q];
            if (!fl)
                continue;
            for (int k = (fl & 1) ? 0 : 4; k < ((fl & 2) ? 8 : 4); ++k)
                into[bins[8 * (size_t)q + k]].push_back({q, wdiv[8 * (size_t)q + k]});
            into[d].push_back({q, -(fl == 3 ? 2.0 : 1.0) * inv_efa[d]});
        }
    std::vector<int> gat_off(nb + 1, 0), gat_q;
    std::vector<double> gat_w;
    for (int r = 0; r < nb; ++r) {
        for (auto &pr : into[r]) {
            gat_q.push_back(pr.first);
            gat_w.push_back(pr.second);
        }
        gat_off[r + 1] = (int)gat_q.size();
    }
    for (void **p : {&c->dia_gat_off, &c->dia_gat_q, &c->dia_gat_w}) {
        free_ptr(*p);
        *p = nullptr;
    }
    const size_t nnz = std::max<size_t>(gat_q.size(), 1);
    if ((rc = alloc(&c->dia_gat_off, gat_off.size() * sizeof(int))) || (rc = alloc(&c->dia_gat_q, nnz * sizeof(int))) ||
        (rc = alloc(&c->dia_gat_w, nnz * c->esize)))
        return rc;
    RPL_HIP(hipMemcpy(c->dia_gat_off, gat_off.data(), gat_off.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!gat_q.empty()) {
        RPL_HIP(hipMemcpy(c->dia_gat_q, gat_q.data(), gat_q.size() * sizeof(int), hipMemcpyHostToDevice));
        if ((rc = upload_real(c, gat_w.data(), c->dia_gat_w, gat_w.size())))
            return rc;
    }
    c->dia_dal[0] = 1 / std::pow(1 + lam, 4);
    c->dia_dal[1] = 1 / std::pow(1 - lam, 4);
    c->dia_dal[2] = 2 * c->dia_dal[0] * c->dia_dal[1];
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->dia = true;
    return RPL_OK;
}

int rpl_set_sources(rpl_ctx *c, const rpl_sources *src, const double *x_nodes, const double *y_nodes)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (!src) {
        c->has_sources = false;
        c->dia = false;
        return RPL_OK;
    }
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    const rpl_sources &s = *src;
    if (s.wind_input != RPL_WIND_INPUT_NONE && s.wind_input != RPL_WIND_INPUT_PRESSURE_CORRELATION)
        return fail(RPL_ERR_ARG, "unsupported wind input");
    if (s.dissipation < RPL_DISSIPATION_NONE || s.dissipation > RPL_DISSIPATION_WHITECAPPING)
        return fail(RPL_ERR_ARG, "unsupported dissipation");
    if (s.wind_input) {
        if (!x_nodes || !y_nodes)
            return fail(RPL_ERR_ARG, "wind input needs the cell-centre coordinates");
        if (s.drag != RPL_DRAG_LINEAR && s.drag != RPL_DRAG_CAPPED)
            return fail(RPL_ERR_ARG, "unsupported drag");
        if (s.wind_field != RPL_WIND_UNIFORM && s.wind_field != RPL_WIND_HOLLAND)
            return fail(RPL_ERR_ARG, "unsupported wind field");
        if (!(s.gravity > 0) || !(s.rho_water > 0) || !(s.von_karman > 0))
            return fail(RPL_ERR_ARG, "wind input constants must be positive");
        if (s.wind_field == RPL_WIND_HOLLAND) {
            // winds.jl:108-111
            if (!(s.vmax >= 0))
                return fail(RPL_ERR_ARG, "Holland hurricane vmax must be nonnegative");
            if (!(s.rmax > 0))
                return fail(RPL_ERR_ARG, "Holland hurricane rmax must be positive");
            if (!(s.radius > s.rmax))
                return fail(RPL_ERR_ARG, "Holland hurricane radius must exceed rmax");
            if (!(s.shape > 0))
                return fail(RPL_ERR_ARG, "Holland hurricane shape_parameter must be positive");
            if (!(s.track_t1 > s.track_t0))
                return fail(RPL_ERR_ARG, "storm-track times must be strictly increasing");
        }
    }
    if (s.nonlinear != RPL_NONLINEAR_NONE && s.nonlinear != RPL_NONLINEAR_SYMMETRIC_QUADRUPLET)
        return fail(RPL_ERR_ARG, "unsupported nonlinear term");
    if (s.nonlinear == RPL_NONLINEAR_SYMMETRIC_QUADRUPLET) {
        if (c->listed)
            return fail(RPL_ERR_STATE, "listed quadruplets are set: one nonlinear term at a time");
        if (c->cfg.nkappa < 2)
            return fail(RPL_ERR_ARG, "the symmetric quadruplet needs at least two frequency bins");
        if (!(s.dia_frequency_offset > 0 && s.dia_frequency_offset < 1))
            return fail(RPL_ERR_ARG, "frequency_offset must lie in (0, 1)");
        if (!(s.dia_gravity > 0) || !std::isfinite(s.dia_coefficient))
            return fail(RPL_ERR_ARG, "symmetric-quadruplet gravity must be positive and its coefficient finite");
        // The DIA workgroup holds a column's action and both orientations'
        // transfers (3 arrays of nkappa * nphi elements) in 64 KiB of LDS.
        const int nbins = c->cfg.nkappa * c->cfg.nphi;
        const int cols = is_f64(c) ? rpl::dia_columns_per_workgroup<double>(nbins)
                                   : rpl::dia_columns_per_workgroup<float>(nbins);
        if (cols < 1)
            return fail(RPL_ERR_UNSUPPORTED, "the symmetric quadruplet holds a column's action and two transfers "
                                             "(3 * nkappa * nphi elements) in 64 KiB of LDS: too many bins");
    }
    if (s.dissipation == RPL_DISSIPATION_MEAN_SPECTRUM && !(s.p_mean != 0))
        return fail(RPL_ERR_ARG, "mean-spectrum exponent p must be nonzero");
    if (s.dissipation == RPL_DISSIPATION_WHITECAPPING) {
        // Sources.jl:589-590, 570-571
        if (!(s.wc_threshold > 0))
            return fail(RPL_ERR_ARG, "whitecapping saturation threshold must be positive");
        if (!(s.wc_reference_wavenumber > 0))
            return fail(RPL_ERR_ARG, "whitecapping reference wavenumber must be positive");
        if (!(s.wc_wavenumber_power >= 0))
            return fail(RPL_ERR_ARG, "wavenumber power must be nonnegative");
    }
    RPL_HIP(hipSetDevice(c->device));
    const int nx = c->cfg.nx, ny = c->cfg.ny, nk = c->cfg.nkappa, np = c->cfg.nphi;
    const size_t tab = (size_t)nk * np, plane = c->plane;
    int rc;
    void **col_bufs[] = {&c->U10, &c->theta_u, &c->ustar, &c->factor, &c->cm0, &c->kbar, &c->sbar};
    for (void **b : col_bufs)
        if (!*b && (rc = alloc(b, plane * c->esize)))
            return rc;
    void **tab_bufs[] = {&c->k_cell, &c->theta_cell, &c->sigma_cell, &c->C_cell, &c->t_w, &c->t_wc};
    for (void **b : tab_bufs)
        if (!*b && (rc = alloc(b, tab * c->esize)))
            return rc;
    void **k_bufs[] = {&c->t_sigma_k, &c->t_kpow, &c->t_spow};
    for (void **b : k_bufs)
        if (!*b && (rc = alloc(b, nk * c->esize)))
            return rc;
    void **p_bufs[] = {&c->t_phi, &c->t_cos, &c->t_sin};
    for (void **b : p_bufs)
        if (!*b && (rc = alloc(b, np * c->esize)))
            return rc;
    if (!is_f64(c)) {
        void **f_col[] = {&c->f_ctu, &c->f_stu, &c->f_lz1};
        for (void **b : f_col)
            if (!*b && (rc = alloc(b, plane * sizeof(float))))
                return rc;
        void **f_tab[] = {&c->f_cth, &c->f_sth, &c->f_logk, &c->f_invc, &c->f_sk3, &c->f_sq};
        for (void **b : f_tab)
            if (!*b && (rc = alloc(b, tab * sizeof(float))))
                return rc;
        void **f_k[] = {&c->f_logkap, &c->f_invck};
        for (void **b : f_k)
            if (!*b && (rc = alloc(b, nk * sizeof(float))))
                return rc;
    }
    if (!c->xn && (rc = alloc((void **)&c->xn, nx * 8)))
        return rc;
    if (!c->yn && (rc = alloc((void **)&c->yn, ny * 8)))
        return rc;
    // Per-bin values of the per-cell path: k = hypot(kx, ky), theta = atan(ky, kx)
    // from (kx, ky) = kappa (cos phi, sin phi) (pressure_correlation.jl:118-123).
    const double g = s.gravity;
    std::vector<double> kcell(tab), thcell(tab), scell(tab), Ccell(tab), w(tab), wcf(tab, 0.0);
    std::vector<double> sk(nk), kpow(nk), spow(nk), cphi(np), sphi(np);
    for (int n = 0; n < np; ++n) {
        cphi[n] = std::cos(c->h_pc[n]);
        sphi[n] = std::sin(c->h_pc[n]);
        for (int m = 0; m < nk; ++m) {
            const size_t t = (size_t)m + (size_t)nk * n;
            const double kx = c->h_kc[m] * std::cos(c->h_pc[n]), ky = c->h_kc[m] * std::sin(c->h_pc[n]);
            const double k = std::hypot(kx, ky);
            kcell[t] = k;
            thcell[t] = std::atan2(ky, kx);
            scell[t] = std::sqrt(g * k);
            Ccell[t] = k > 0 ? scell[t] / k : 0.0;
            const double k1 = c->h_kf[m], k2 = c->h_kf[m + 1];
            w[t] = ((k2 * k2 - k1 * k1) / 2) * (c->h_pf[n + 1] - c->h_pf[n]);
            if (s.dissipation == RPL_DISSIPATION_WHITECAPPING) {
                // radial_power_source_factor (Sources.jl:569-574): the cell average
                // of kappa^p (finite_volume_integration.jl:24-33, 71-73) over kappa_ref^p.
                const double pw = s.wc_wavenumber_power;
                if (pw == 0) {
                    wcf[t] = 1.0;
                } else {
                    const double measure = (c->h_pf[n + 1] - c->h_pf[n]) * (std::pow(k2, pw + 2) - std::pow(k1, pw + 2)) /
                                           (pw + 2);
                    wcf[t] = measure / w[t] / std::pow(s.wc_reference_wavenumber, pw);
                }
            }
        }
    }
    for (int m = 0; m < nk; ++m) {
        sk[m] = std::sqrt(g * c->h_kc[m]);
        kpow[m] = s.dissipation == RPL_DISSIPATION_MEAN_SPECTRUM ? std::pow(c->h_kc[m], s.p_mean) : 0.0;
        spow[m] = s.dissipation == RPL_DISSIPATION_MEAN_SPECTRUM ? std::pow(sk[m], s.p_mean) : 0.0;
    }
    c->h_sigma = scell;
    if (!is_f64(c)) {
        // FP32 factors: cos/sin theta, log k, k / sigma, sigma k^3, sigma^q per
        // bin; log kappa and kappa / sigma per kappa bin (state kernel).
        std::vector<double> fc(tab), fs(tab), flk(tab), fic(tab), fsk3(tab), fsq(tab), flkap(nk), fick(nk);
        for (size_t t = 0; t < tab; ++t) {
            const double k = kcell[t], sg = scell[t];
            fc[t] = std::cos(thcell[t]);
            fs[t] = std::sin(thcell[t]);
            flk[t] = k > 0 ? std::log(k) : 0.0;
            fic[t] = sg > 0 ? k / sg : 0.0;
            fsk3[t] = sg * (k * k * k);
            fsq[t] = std::pow(sg, s.ls_sigma_power);
        }
        for (int m = 0; m < nk; ++m) {
            flkap[m] = c->h_kc[m] > 0 ? std::log(c->h_kc[m]) : 0.0;
            fick[m] = sk[m] > 0 ? c->h_kc[m] / sk[m] : 0.0;
        }
        if (s.wind_input && s.stress_cap) {
            // stress-cap weights of the fused path: the state kernel's
            // (k / sigma) w rho_w g per bin times cos and sin of the direction
            std::vector<double> tc(tab), ts(tab);
            for (size_t t = 0; t < tab; ++t) {
                const double wt = fic[t] * w[t] * (s.rho_water * s.gravity);
                tc[t] = wt * fc[t];
                ts[t] = wt * fs[t];
            }
            void **tt[] = {&c->tau_c, &c->tau_s};
            for (void **b : tt)
                if (!*b && (rc = alloc(b, tab * sizeof(float))))
                    return rc;
            void **tp[] = {&c->tau_x, &c->tau_y};
            for (void **b : tp)
                if (!*b && (rc = alloc(b, plane * (size_t)nk * sizeof(float))))
                    return rc;
            if ((rc = upload_real(c, tc.data(), c->tau_c, tab)) || (rc = upload_real(c, ts.data(), c->tau_s, tab)))
                return rc;
        }
        if ((rc = upload_real(c, fc.data(), c->f_cth, tab)) || (rc = upload_real(c, fs.data(), c->f_sth, tab)) ||
            (rc = upload_real(c, flk.data(), c->f_logk, tab)) || (rc = upload_real(c, fic.data(), c->f_invc, tab)) ||
            (rc = upload_real(c, fsk3.data(), c->f_sk3, tab)) || (rc = upload_real(c, fsq.data(), c->f_sq, tab)) ||
            (rc = upload_real(c, flkap.data(), c->f_logkap, nk)) || (rc = upload_real(c, fick.data(), c->f_invck, nk)))
            return rc;
    }
    if ((rc = upload_real(c, kcell.data(), c->k_cell, tab)) || (rc = upload_real(c, thcell.data(), c->theta_cell, tab)) ||
        (rc = upload_real(c, scell.data(), c->sigma_cell, tab)) || (rc = upload_real(c, Ccell.data(), c->C_cell, tab)) ||
        (rc = upload_real(c, w.data(), c->t_w, tab)) || (rc = upload_real(c, wcf.data(), c->t_wc, tab)) ||
        (rc = upload_real(c, sk.data(), c->t_sigma_k, nk)) ||
        (rc = upload_real(c, kpow.data(), c->t_kpow, nk)) || (rc = upload_real(c, spow.data(), c->t_spow, nk)) ||
        (rc = upload_real(c, c->h_pc.data(), c->t_phi, np)) || (rc = upload_real(c, cphi.data(), c->t_cos, np)) ||
        (rc = upload_real(c, sphi.data(), c->t_sin, np)))
        return rc;
    if (s.wind_input) {
        RPL_HIP(hipMemcpy(c->xn, x_nodes, nx * 8, hipMemcpyHostToDevice));
        RPL_HIP(hipMemcpy(c->yn, y_nodes, ny * 8, hipMemcpyHostToDevice));
    }
    // Stress factor defaults to 1 where no cap is computed.
    std::vector<double> ones(plane, 1.0);
    if ((rc = upload_real(c, ones.data(), c->factor, plane)))
        return rc;
    rpl::WindParams &wp = c->wind;
    wp = rpl::WindParams{};
    wp.field = s.wind_field;
    wp.alpha0 = s.alpha0;
    wp.gravity = s.gravity;
    wp.drag = s.drag;
    wp.rotation = s.rotation >= 0 ? 1 : -1;
    wp.speed = s.wind_speed;
    wp.direction = s.wind_direction;
    wp.vmax = s.vmax;
    wp.rmax = s.rmax;
    wp.radius = s.radius;
    wp.shape = s.shape;
    // winds.jl:150: max(radius - rmax, eps(radius)).
    const double eps_r = std::nextafter(s.radius, INFINITY) - s.radius;
    wp.decay = std::max(s.radius - s.rmax, eps_r);
    wp.bx = s.background_x;
    wp.by = s.background_y;
    wp.cos_inflow = std::cos(s.inflow_angle);
    wp.sin_inflow = std::sin(s.inflow_angle);
    wp.t0 = s.track_t0;
    wp.t1 = s.track_t1;
    wp.x0 = s.track_x0;
    wp.y0 = s.track_y0;
    wp.x1 = s.track_x1;
    wp.y1 = s.track_y1;
    if ((rc = build_dia_tables(c, s)))
        return rc;
    c->src = s;
    c->has_sources = s.wind_input != RPL_WIND_INPUT_NONE || s.dissipation != RPL_DISSIPATION_NONE;
    RPL_HIP(hipStreamSynchronize(c->stream));
    return RPL_OK;
}

int rpl_set_quadruplets(rpl_ctx *c, int32_t n, const int32_t *bins, const double *rates, const double *frequency,
                        double power, double resonance_tolerance)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (!c->has_geometry)
        return fail(RPL_ERR_STATE, "rpl_set_geometry has not been called");
    if (n <= 0) {
        c->listed = false;
        return RPL_OK;
    }
    if (!bins || !rates || !frequency)
        return fail(RPL_ERR_ARG, "null argument");
    if (c->dia)
        return fail(RPL_ERR_STATE, "the symmetric quadruplet is set: one nonlinear term at a time");
    if (!(power > 0))
        return fail(RPL_ERR_ARG, "DIA transfer power must be positive");  // Sources.jl:1334
    if (!(resonance_tolerance >= 0))
        return fail(RPL_ERR_ARG, "quadruplet resonance tolerance must be nonnegative");  // Sources.jl:1297
    const int nk = c->cfg.nkappa, np = c->cfg.nphi;
    auto freq = [&](int m) { return frequency[m]; };  // diagnostics/moments.jl:15
    for (int q = 0; q < n; ++q) {
        const int32_t *b = bins + 8 * (size_t)q;
        for (int k = 0; k < 8; k += 2)
            if (b[k] < 0 || b[k] >= nk || b[k + 1] < 0 || b[k + 1] >= np)
                return fail(RPL_ERR_ARG, "quadruplet transfer index outside the spectral grid");  // Sources.jl:1274-1285
        // Sources.jl:1287-1292: the donor and receiver bin pairs differ.
        std::pair<int, int> d1{b[0], b[1]}, d2{b[2], b[3]}, r1{b[4], b[5]}, r2{b[6], b[7]};
        if (std::minmax(d1, d2) == std::minmax(r1, r2))
            return fail(RPL_ERR_ARG, "quadruplet transfer donor and receiver bin pairs must differ");
        // Sources.jl:1296-1329: resonance in frequency and wavevector.
        const double fd = freq(b[0]) + freq(b[2]), fr = freq(b[4]) + freq(b[6]);
        const double fscale = std::max({std::fabs(fd), std::fabs(fr), 2.220446049250313e-16});
        if (!(std::fabs(fr - fd) <= resonance_tolerance * fscale))
            return fail(RPL_ERR_ARG, "quadruplet transfer requires receiver frequencies to match donor frequencies within tolerance");
        double kx[4], ky[4];
        for (int k = 0; k < 4; ++k) {
            kx[k] = c->h_kc[b[2 * k]] * std::cos(c->h_pc[b[2 * k + 1]]);
            ky[k] = c->h_kc[b[2 * k]] * std::sin(c->h_pc[b[2 * k + 1]]);
        }
        const double dkx = kx[0] + kx[1], dky = ky[0] + ky[1], rkx = kx[2] + kx[3], rky = ky[2] + ky[3];
        double vscale = std::max(std::hypot(dkx, dky), std::hypot(rkx, rky));
        for (int k = 0; k < 4; ++k)
            vscale = std::max({vscale, std::fabs(kx[k]), std::fabs(ky[k])});
        vscale = std::max(vscale, 2.220446049250313e-16);
        if (!(std::hypot(rkx - dkx, rky - dky) <= resonance_tolerance * vscale))
            return fail(RPL_ERR_ARG, "quadruplet transfer requires receiver wavevectors to match donor wavevectors within tolerance");
        if (!(rates[q] >= 0))
            return fail(RPL_ERR_ARG, "DIA transfer rates must be nonnegative");  // Sources.jl:1347
    }
    const size_t tab = (size_t)nk * np;
    std::vector<double> w(tab);
    for (int pn = 0; pn < np; ++pn)
        for (int m = 0; m < nk; ++m) {
            const double k1 = c->h_kf[m], k2 = c->h_kf[m + 1];
            w[m + (size_t)nk * pn] = ((k2 * k2 - k1 * k1) / 2) * (c->h_pf[pn + 1] - c->h_pf[pn]);
            if (!(w[m + (size_t)nk * pn] > 0))
                return fail(RPL_ERR_ARG, "DIA transfer requires positive spectral cell measures");
        }
    if ((size_t)nk * np > 65535)
        return fail(RPL_ERR_UNSUPPORTED, "listed quadruplets support at most 65535 spectral bins");
    RPL_HIP(hipSetDevice(c->device));
    c->listed = false;  // until the new list is on the device
    c->listed_has_field = false;  // a new list takes its constant rates
    free_ptr(c->listed_field);    // and may have another length
    c->listed_field = nullptr;
    free_ptr(c->listed_bins);
    free_ptr(c->listed_rates);
    c->listed_bins = c->listed_rates = nullptr;
    int rc;
    if ((rc = alloc(&c->listed_bins, 8 * (size_t)n * sizeof(int32_t))) ||
        (rc = alloc(&c->listed_rates, (size_t)n * c->esize)) ||
        (!c->listed_w && (rc = alloc(&c->listed_w, tab * c->esize))) ||
        (!c->dia_add && (rc = alloc(&c->dia_add, c->cnt * c->esize))))
        return rc;
    RPL_HIP(hipMemcpy(c->listed_bins, bins, 8 * (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
    if ((rc = upload_real(c, rates, c->listed_rates, n)) || (rc = upload_real(c, w.data(), c->listed_w, tab)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->listed_n = n;
    c->listed_power = power;
    c->listed = true;
    return RPL_OK;
}

int rpl_set_quadruplet_rate_field(rpl_ctx *c, const double *rates)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (!rates) {
        c->listed_has_field = false;
        return RPL_OK;
    }
    if (!c->listed)
        return fail(RPL_ERR_STATE, "rpl_set_quadruplets has not set a list of interactions");
    const size_t n = (size_t)c->cfg.nx * c->cfg.ny * c->listed_n;
    for (size_t k = 0; k < n; ++k)
        if (!(rates[k] >= 0))
            return fail(RPL_ERR_ARG, "DIA transfer rates must be nonnegative");  // Sources.jl:1347
    RPL_HIP(hipSetDevice(c->device));
    int rc;
    if (!c->listed_field && (rc = alloc(&c->listed_field, n * c->esize)))
        return rc;
    c->listed_has_field = false;  // until the field is on the device
    if ((rc = upload_real(c, rates, c->listed_field, n)))
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    c->listed_has_field = true;
    return RPL_OK;
}

int rpl_set_storm_track(rpl_ctx *c, int32_t n, const double *times, const double *x, const double *y)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    if (n == 0) {
        c->track_t.clear();
        c->track_x.clear();
        c->track_y.clear();
        return RPL_OK;
    }
    if (n < 0 || !times || !x || !y)
        return fail(RPL_ERR_ARG, "storm-track times and centers must be non-empty");  // winds.jl:14-15
    for (int k = 1; k < n; ++k)
        if (!(times[k] > times[k - 1]))
            return fail(RPL_ERR_ARG, "storm-track times must be strictly increasing");  // winds.jl:16-19
    c->track_t.assign(times, times + n);
    c->track_x.assign(x, x + n);
    c->track_y.assign(y, y + n);
    return RPL_OK;
}

int rpl_get_wind(rpl_ctx *c, double *U10, double *theta_u, double *u_star)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (!c->has_sources || !c->src.wind_input)
        return fail(RPL_ERR_STATE, "no wind input installed");
    RPL_HIP(hipSetDevice(c->device));
    int rc = launch_wind(c);
    if (rc)
        return rc;
    RPL_HIP(hipStreamSynchronize(c->stream));
    double *outs[] = {U10, theta_u, u_star};
    void *ins[] = {c->U10, c->theta_u, c->ustar};
    for (int q = 0; q < 3; ++q) {
        if (!outs[q])
            continue;
        if (is_f64(c)) {
            RPL_HIP(hipMemcpy(outs[q], ins[q], c->plane * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(c->plane);
            RPL_HIP(hipMemcpy(t.data(), ins[q], c->plane * 4, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < c->plane; ++k)
                outs[q][k] = t[k];
        }
    }
    return RPL_OK;
}

int rpl_enable_kernel_timing(rpl_ctx *c, int32_t enable)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    c->timing = enable != 0;
    return RPL_OK;
}

int rpl_get_kernel_timing(rpl_ctx *c, double *total_ms, int64_t *launches)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[0];
    if (launches)
        *launches = c->timing_launches[0];
    c->timing_ms[0] = 0;
    c->timing_launches[0] = 0;
    return RPL_OK;
}

int rpl_get_coupling_timing(rpl_ctx *c, double *total_ms, int64_t *updates)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[2];
    if (updates)
        *updates = c->timing_launches[2];
    c->timing_ms[2] = 0;
    c->timing_launches[2] = 0;
    return RPL_OK;
}

int rpl_get_nonlinear_timing(rpl_ctx *c, double *total_ms, int64_t *launches)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[4];
    if (launches)
        *launches = c->timing_launches[4];
    c->timing_ms[4] = 0;
    c->timing_launches[4] = 0;
    return RPL_OK;
}

int rpl_get_smoothing_timing(rpl_ctx *c, double *total_ms, int64_t *passes)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[3];
    if (passes)
        *passes = c->timing_launches[3];
    c->timing_ms[3] = 0;
    c->timing_launches[3] = 0;
    return RPL_OK;
}

int rpl_get_source_timing(rpl_ctx *c, double *total_ms, int64_t *launches)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    RPL_HIP(hipSetDevice(c->device));
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[5];
    if (launches)
        *launches = c->timing_launches[5];
    c->timing_ms[5] = 0;
    c->timing_launches[5] = 0;
    return RPL_OK;
}

int rpl_get_surrogate_timing(rpl_ctx *c, double *total_ms, int64_t *evaluations)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    int rc = fold_timing(c);
    if (rc)
        return rc;
    if (total_ms)
        *total_ms = c->timing_ms[1];
    if (evaluations)
        *evaluations = c->timing_launches[1];
    c->timing_ms[1] = 0;
    c->timing_launches[1] = 0;
    return RPL_OK;
}

int rpl_attach_surrogate(rpl_ctx *c, rpl_mlp *net, const rpl_surrogate *cfg, const float *bin_action_hi)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    ++c->generation;
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    if (!net) {
        free_surrogate(c);
        return RPL_OK;
    }
    if (!cfg)
        return fail(RPL_ERR_ARG, "null surrogate configuration");
    if (!c->has_sources || !c->src.wind_input)
        return fail(RPL_ERR_STATE, "the surrogate needs rpl_set_sources with a wind input (its plain fallback)");
    const int nbins = c->cfg.nkappa * c->cfg.nphi;
    if (net->in_dim != nbins + 3 || net->out_dim != nbins)
        return fail(RPL_ERR_ARG, "network must map nkappa*nphi + 3 features to nkappa*nphi rates");
    if (net->device != c->device)
        return fail(RPL_ERR_ARG, "network and context live on different devices");
    if (!(cfg->rate_unit > 0) || cfg->audit_stride < 0 || !(cfg->floor > 0))
        return fail(RPL_ERR_ARG, "need rate_unit > 0, floor > 0, audit_stride >= 0");
    if (cfg->terms != 0 && cfg->terms != 1)
        return fail(RPL_ERR_ARG, "terms must be 0 (all sources) or 1 (wind input)");
    if (cfg->refresh != 0 && cfg->refresh != 1)
        return fail(RPL_ERR_ARG, "refresh must be 0 (every stage) or 1 (once per step)");
    free_surrogate(c);
    const int in_pad = net->dims[0], out_pad = net->dims.back(), hid = net->hidden;
    c->s_ld = (c->plane + 63) / 64 * 64;
    c->s_chunk = std::min<size_t>(c->s_ld, 65536);  // measured: 2^18 and 2^20 columns per chunk are 2-3 % slower
    int rc;
    if ((rc = alloc((void **)&c->sX, c->s_chunk * in_pad * 2)) || (rc = alloc((void **)&c->sA, c->s_chunk * hid * 2)) ||
        (rc = alloc((void **)&c->sB, c->s_chunk * hid * 2)) ||
        (rc = alloc((void **)&c->rate, c->s_ld * out_pad * sizeof(_Float16))) ||
        (rc = alloc((void **)&c->out_scale, out_pad * sizeof(float))) || (rc = alloc((void **)&c->s_err, 16)) ||
        (rc = alloc((void **)&c->use_plain, c->s_ld)) || (rc = alloc((void **)&c->s_counts, 16)) ||
        (rc = alloc((void **)&c->s_amax, c->s_ld * sizeof(float))) || (rc = alloc((void **)&c->s_flags, c->s_ld))) {
        free_surrogate(c);
        return rc;
    }
    std::vector<float> scale(out_pad, 1.0f);
    for (int t = 0; t < nbins; ++t)
        scale[t] = (float)(c->h_sigma[t] * cfg->rate_unit);
    RPL_HIP(hipMemcpy(c->out_scale, scale.data(), out_pad * sizeof(float), hipMemcpyHostToDevice));
    RPL_HIP(hipMemset(c->s_counts, 0, 16));
    RPL_HIP(hipMemset(c->s_err, 0, 16));
    RPL_HIP(hipMemset(c->use_plain, 1, c->s_ld));
    if (bin_action_hi) {
        if ((rc = alloc((void **)&c->bin_hi, nbins * sizeof(float)))) {
            free_surrogate(c);
            return rc;
        }
        RPL_HIP(hipMemcpy(c->bin_hi, bin_action_hi, nbins * sizeof(float), hipMemcpyHostToDevice));
    }
    c->net = net;
    c->scfg = *cfg;
    c->s_active = true;
    c->s_due = true;
    c->s_columns = 0;
    return RPL_OK;
}

int rpl_surrogate_stats(rpl_ctx *c, int64_t *columns, int64_t *plain, int64_t *audited, int64_t *failed,
                        double *err_max)
{
    if (!c)
        return fail(RPL_ERR_ARG, "null context");
    if (!c->s_active)
        return fail(RPL_ERR_STATE, "no surrogate attached");
    RPL_HIP(hipSetDevice(c->device));
    RPL_HIP(hipStreamSynchronize(c->stream));
    unsigned h[4] = {0, 0, 0, 0};
    float e = 0;
    RPL_HIP(hipMemcpy(h, c->s_counts, 16, hipMemcpyDeviceToHost));
    RPL_HIP(hipMemcpy(&e, c->s_err, 4, hipMemcpyDeviceToHost));
    if (columns)
        *columns = c->s_columns;
    if (plain)
        *plain = h[0];
    if (audited)
        *audited = h[1];
    if (failed)
        *failed = h[2];
    if (err_max)
        *err_max = e;
    RPL_HIP(hipMemset(c->s_counts, 0, 16));
    RPL_HIP(hipMemset(c->s_err, 0, 16));
    c->s_columns = 0;
    return RPL_OK;
}


int rpl_surrogate_training_pairs(rpl_ctx *c, const rpl_surrogate *cfg, float *features, float *targets)
{
    int rc = check_ready(c);
    if (rc)
        return rc;
    if (!cfg)
        return fail(RPL_ERR_ARG, "null surrogate configuration");
    if (!c->has_sources || !c->src.wind_input)
        return fail(RPL_ERR_STATE, "training pairs need rpl_set_sources with a wind input");
    if (!(cfg->rate_unit > 0) || !(cfg->floor > 0))
        return fail(RPL_ERR_ARG, "need rate_unit > 0 and floor > 0");
    const bool was = c->s_active;
    c->s_active = false;  // targets come from the plain bundle on every column
    if ((rc = launch_wind(c)) == RPL_OK && (rc = launch_state(c, c->A)) == RPL_OK)
        rc = is_f64(c) ? training_pairs<double>(c, cfg, features, targets)
                       : training_pairs<float>(c, cfg, features, targets);
    c->s_active = was;
    return rc;
}

int rpl_ensemble_create(rpl_ensemble **out, rpl_ctx *const *members, int32_t count)
{
    if (!out || !members || count < 1)
        return fail(RPL_ERR_ARG, "need an output pointer and at least one member");
    *out = nullptr;
    rpl_ctx *first = members[0];
    for (int k = 0; k < count; ++k) {
        rpl_ctx *c = members[k];
        if (!c)
            return fail(RPL_ERR_ARG, "null member");
        int rc = check_ready(c);
        if (rc)
            return rc;
        for (int q = 0; q < k; ++q)
            if (members[q] == c)
                return fail(RPL_ERR_ARG, "a context appears twice in the ensemble");
        if ((rc = ensemble_member_check(c, first)))
            return rc;
    }
    rpl_ensemble *e = new rpl_ensemble();
    e->device = first->device;
    e->transport = first->cfg.transport;
    e->src = first->has_sources;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, e->device) != hipSuccess ||
        hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(RPL_ERR_HIP, "ensemble stream creation");
    }
    e->cus = prop.multiProcessorCount;
    e->blocks = default_blocks(e);
    bool ok = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
    for (int b = 0; b < ENSEMBLE_BRANCHES && ok; ++b)
        ok = hipStreamCreateWithFlags(&e->branch[b], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&e->join[b], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        destroy_branches(e);
        (void)hipStreamDestroy(e->stream);
        delete e;
        return fail(RPL_ERR_HIP, "ensemble branch streams");
    }
    if (hipMalloc((void **)&e->desc, 3 * (size_t)count * sizeof(rpl::EnsembleMember)) != hipSuccess ||
        hipMalloc((void **)&e->sdesc, 3 * (size_t)count * sizeof(rpl::EnsembleState)) != hipSuccess) {
        if (e->desc)
            (void)hipFree(e->desc);
        destroy_branches(e);
        (void)hipStreamDestroy(e->stream);
        delete e;
        return fail(RPL_ERR_NOMEM, "ensemble descriptors");
    }
    for (int k = 0; k < count; ++k) {
        rpl_ctx *c = members[k];
        (void)hipStreamSynchronize(c->stream);
        (void)fold_timing(c);
        c->timing = false;
        e->members.push_back(c);
        e->own_stream.push_back(c->stream);
        c->stream = e->stream;  // every member's work is ordered on the ensemble stream
    }
    *out = e;
    return RPL_OK;
}

int rpl_ensemble_destroy(rpl_ensemble *e)
{
    if (!e)
        return RPL_OK;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (size_t k = 0; k < e->members.size(); ++k)
        e->members[k]->stream = e->own_stream[k];
    if (e->desc)
        (void)hipFree(e->desc);
    if (e->sdesc)
        (void)hipFree(e->sdesc);
    if (e->chunk_member)
        (void)hipFree(e->chunk_member);
    for (int b = 0; b < ENSEMBLE_BRANCHES; ++b)
        if (e->branch[b])
            (void)hipStreamSynchronize(e->branch[b]);
    destroy_branches(e);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return RPL_OK;
}

int rpl_ensemble_set_launch(rpl_ensemble *e, int32_t tile_rows, int32_t blocks)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    if (tile_rows != 0 && tile_rows != 2 && tile_rows != 4 && tile_rows != 8)
        return fail(RPL_ERR_ARG, "tile_rows must be 2, 4 or 8 (0: keep)");
    if (blocks < 0 || blocks % 8)
        return fail(RPL_ERR_ARG, "blocks must be a multiple of 8 (0: default)");
    if (tile_rows)
        e->ty = tile_rows;
    e->blocks = blocks ? blocks : default_blocks(e);
    return RPL_OK;
}

int rpl_ensemble_get_launch(rpl_ensemble *e, int32_t *tile_rows, int32_t *blocks, int64_t *tiles)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    if (tile_rows)
        *tile_rows = e->ty;
    if (blocks)
        *blocks = e->blocks;  // 0: one workgroup per tile
    if (tiles) {
        long long t = 0;
        for (rpl_ctx *c : e->members) {
            if (e->transport == RPL_TRANSPORT_INTRINSIC || e->transport == RPL_TRANSPORT_REFRACTION)
                t += (long long)((c->cfg.nx + 63) / 64) * ((c->cfg.ny + e->ty - 1) / e->ty) * c->cfg.nkappa;
            else
                t += (long long)((c->cfg.nx + 63) / 64) * ((c->cfg.ny + e->ty - 1) / e->ty) * c->cfg.nkappa *
                     c->cfg.nphi;
        }
        *tiles = t;
    }
    return RPL_OK;
}

int rpl_ensemble_set_split(rpl_ensemble *e, int64_t min_tiles)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    e->split = min_tiles;
    e->desc_dt = -1;
    return RPL_OK;
}

int rpl_ensemble_get_split(rpl_ensemble *e, int64_t *min_tiles, int32_t *own_members)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    if (min_tiles)
        *min_tiles = e->split == 0 ? ENSEMBLE_OWN_TILES : e->split;
    if (own_members) {
        int n = 0;
        for (rpl_ctx *c : e->members)
            n += own_kernel(e, c) ? 1 : 0;
        *own_members = n;
    }
    return RPL_OK;
}

int rpl_ensemble_use_graph(rpl_ensemble *e, int32_t enable)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    RPL_HIP(hipSetDevice(e->device));
    RPL_HIP(hipStreamSynchronize(e->stream));
    drop_graph(e);
    e->use_graph = enable != 0;
    return RPL_OK;
}

int rpl_ensemble_step(rpl_ensemble *e, double dt, int32_t nsteps)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    if (!(dt > 0))
        return fail(RPL_ERR_ARG, "time step must be positive");
    if (nsteps < 0)
        return fail(RPL_ERR_ARG, "nsteps must be >= 0");
    RPL_HIP(hipSetDevice(e->device));
    int rc;
    for (rpl_ctx *c : e->members)
        if ((rc = check_ready(c)) || (rc = ensemble_member_check(c, e->members[0])))
            return rc;
    // Parameters may change between calls (sources, surrogate attachment);
    // the descriptors are small, so they are rebuilt on every call.
    RPL_HIP(hipStreamSynchronize(e->stream));
    if ((rc = ensemble_descriptors(e, dt)))
        return rc;
    std::vector<uint64_t> gen;
    for (rpl_ctx *c : e->members)
        gen.push_back(c->generation);
    if (e->use_graph && (e->exec == nullptr || e->graph_dt != dt || e->graph_ty != e->ty ||
                         e->graph_blocks != e->blocks || e->graph_split != e->split || e->graph_gen != gen)) {
        drop_graph(e);
        // the captured stages stand for a whole step: members whose network
        // infers once per step record that inference at their first stage
        for (rpl_ctx *c : e->members)
            c->s_due = true;
        hipGraph_t g = nullptr;
        RPL_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        rc = ensemble_stages(e, dt);
        hipError_t err = hipStreamEndCapture(e->stream, &g);
        if (rc)
            return rc;
        if (err != hipSuccess)
            return fail(RPL_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(err));
        err = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            e->exec = nullptr;
            return fail(RPL_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(err));
        }
        e->graph_dt = dt;
        e->graph_ty = e->ty;
        e->graph_blocks = e->blocks;
        e->graph_split = e->split;
        e->graph_gen = gen;
    }
    for (int s = 0; s < nsteps; ++s) {
        for (rpl_ctx *c : e->members)
            if ((rc = launch_wind(c)))
                return rc;
        if (e->use_graph) {
            RPL_HIP(hipGraphLaunch(e->exec, e->stream));
        } else if ((rc = ensemble_stages(e, dt))) {
            return rc;
        }
        for (rpl_ctx *c : e->members) {
            c->time += dt;
            c->iteration += 1;
        }
    }
    return RPL_OK;
}

int rpl_ensemble_synchronize(rpl_ensemble *e)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    RPL_HIP(hipSetDevice(e->device));
    RPL_HIP(hipStreamSynchronize(e->stream));
    return RPL_OK;
}

int rpl_ensemble_autotune(rpl_ensemble *e, double dt, int32_t reps, int32_t *tile_rows, int32_t *blocks,
                          double *best_ms)
{
    if (!e)
        return fail(RPL_ERR_ARG, "null ensemble");
    if (!(dt > 0) || reps < 1)
        return fail(RPL_ERR_ARG, "need dt > 0 and reps >= 1");
    RPL_HIP(hipSetDevice(e->device));
    RPL_HIP(hipStreamSynchronize(e->stream));
    drop_graph(e);
    int rc;
    // Stage 1 reads A and writes the scratch buffer B, so timing it leaves
    // every member's state unchanged.  Sources are prepared once on A.
    for (rpl_ctx *c : e->members)
        if ((rc = launch_wind(c)) || (e->src && (rc = launch_sources(c, c->A))))
            return rc;
    hipEvent_t a, b;
    RPL_HIP(hipEventCreate(&a));
    RPL_HIP(hipEventCreate(&b));
    // Mean time of `reps` runs of `fn` (after one warm-up) on the ensemble stream, in ms.
    auto timed = [&](auto fn, float *ms) {
        int r = fn();
        if (r == RPL_OK && hipEventRecord(a, e->stream) != hipSuccess)
            r = fail(RPL_ERR_HIP, "autotune: event record");
        for (int q = 0; q < reps && r == RPL_OK; ++q)
            r = fn();
        if (r == RPL_OK && (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                            hipEventElapsedTime(ms, a, b) != hipSuccess))
            r = fail(RPL_ERR_HIP, "autotune: event timing");
        *ms /= reps;
        return r;
    };
    double best = 1e300;
    int best_ty = e->ty, best_blocks = e->blocks;
    long long best_split = e->split;
    // -1: every member in the work list; 1: every member on its own chain
    const long long splits[5] = {-1, 16384, ENSEMBLE_OWN_TILES, 1024, 1};
    const int tys[3] = {2, 4, 8};
    const int mult[5] = {2, 4, 8, 16, 0};  // workgroups per CU; 0: one workgroup per tile
    rc = RPL_OK;
    for (long long split : splits) {
        e->split = split;
        // the work list's shape first, timing its launch alone ...
        int lty = e->ty, lblocks = e->blocks;
        if ((rc = ensemble_descriptors(e, dt)))
            break;
        if (e->lmembers) {
            float lbest = 1e30f;
            for (int ty : tys) {
                e->ty = ty;
                if ((rc = ensemble_descriptors(e, dt)))
                    break;
                for (int mu : mult) {
                    long long g = mu ? (long long)e->cus * mu : (e->total + 7) / 8 * 8;
                    g = std::max<long long>(8, std::min<long long>(g, 0x7ffffff8LL) / 8 * 8);
                    float ms = 0;
                    rc = timed(
                        [&]() {
                            hipError_t x = rpl::launch_ensemble_stage(e->desc, e->lmembers, e->total,
                                                                      e->chunk_member, e->transport, e->src_kind,
                                                                      rpl::STAGE_1, ty, (int)g, e->stream);
                            return x == hipSuccess ? RPL_OK
                                                   : fail(RPL_ERR_HIP, std::string("autotune: ") +
                                                                           hipGetErrorString(x));
                        },
                        &ms);
                    if (rc)
                        break;
                    if (ms < lbest) {
                        lbest = ms;
                        lty = ty;
                        lblocks = mu ? (int)g : 0;
                    }
                }
                if (rc)
                    break;
            }
            if (rc)
                break;
            e->ty = lty;
            e->blocks = lblocks;
            if ((rc = ensemble_descriptors(e, dt)))
                break;
        }
        // ... then the whole first stage: the large members' chains on the
        // branch streams beside the work list
        float ms = 0;
        if ((rc = timed([&]() { return ensemble_run(e, rpl::STAGE_1, rpl::STAGE_1, dt); }, &ms)))
            break;
        if (ms < best) {
            best = ms;
            best_ty = lty;
            best_blocks = lblocks;
            best_split = split;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (rpl_ctx *c : e->members)
        c->tau_for = nullptr;
    if (rc)
        return rc;
    e->split = best_split;
    e->ty = best_ty;
    e->blocks = best_blocks;
    e->desc_dt = -1;
    if (tile_rows)
        *tile_rows = best_ty;
    if (blocks)
        *blocks = best_blocks;
    if (best_ms)
        *best_ms = best;
    return RPL_OK;
}

void *rpl_stream(rpl_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"
