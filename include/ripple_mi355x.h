/*
 * ripple_mi355x.h -- C ABI of the MI355X (gfx950) wave-action time step.
 *
 * The reference (Ripple.jl) has no C ABI: its device boundary is a
 * KernelAbstractions launch inside Julia methods.  Each entry point below
 * replaces the reference interface cited beside it (paths relative to the
 * Ripple.jl source tree).  A host layer (the Julia module in
 * ripple.jl_amd/julia/, the Python module ripple.jl_amd/ripple_amd/) keeps
 * the reference's SpectralWaveModel / set! / time_step! / Simulation surface
 * and calls these functions.
 *
 * Conventions
 *   - Every function returns RPL_OK (0) or a negative RPL_ERR_* code; the
 *     message of the calling thread's last failure is rpl_last_error().
 *   - Host pointers are borrowed for the duration of the call only.
 *   - All device memory is owned by the context.  A context is used from one
 *     host thread at a time; several contexts may run concurrently.
 *   - Array layout at the boundary is the interior of the reference's
 *     ProductField, i fastest: N[i + nx*(j + ny*(m + nkappa*n))] for
 *     N[i,j,m,n] (ProductFields/product_field.jl:166-186 without the halo).
 *     (nx, ny, nkappa) arrays are U[i + nx*(j + ny*m)]; (nx, ny, nz) current
 *     arrays are u[i + nx*(j + ny*k)].
 *   - Work is enqueued on the context's HIP stream; functions that return
 *     data to the host synchronize that stream first.
 */
#ifndef RIPPLE_MI355X_H
#define RIPPLE_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPL_OK 0
#define RPL_ERR_ARG (-1)         /* invalid argument (reference: ArgumentError) */
#define RPL_ERR_STATE (-2)       /* call out of order (e.g. step before geometry) */
#define RPL_ERR_HIP (-3)         /* HIP runtime failure */
#define RPL_ERR_UNSUPPORTED (-4) /* configuration outside this build's scope */
#define RPL_ERR_NOMEM (-5)       /* device allocation failed */

enum { RPL_F32 = 0, RPL_F64 = 1 };

/* Models/time_step.jl:117-166: :RK3, :LSRK3 (same arithmetic), :ForwardEuler. */
enum { RPL_RK3 = 0, RPL_LSRK3 = 1, RPL_FORWARD_EULER = 2 };

/* Models/tendencies.jl:18-109 dispatch, resolved by the host layer:
 *   NONE       horizontal_advection = nothing, no coupling (G = 0 + sources)
 *   INTRINSIC  fused intrinsic transport (intrinsic_transport.jl:14-68)
 *   REFRACTION fused Doppler transport + refraction (wave_current_refraction.jl:74-151) */
enum { RPL_TRANSPORT_NONE = 0, RPL_TRANSPORT_INTRINSIC = 1, RPL_TRANSPORT_REFRACTION = 2 };

/* Intrinsic group velocity (Models/transport.jl:53-95). */
enum { RPL_DEPTH_INFINITE = 0, RPL_DEPTH_CONSTANT = 1 };

/* Implementation selector.  REFERENCE evaluates the reference formulas in
 * their order (the FP64 parity path, bit-identical to the CPU restatement);
 * FAST is the restructured stencil (one upwind reconstruction per face,
 * shared smoothness indicators, the phi bins walked inside the workgroup),
 * in FP32 and, for refraction and intrinsic transport, in FP64, where it
 * agrees with the reference order to rounding.  The Doppler contraction of
 * rpl_set_current runs on the matrix cores (MFMA) for FP32 always and for
 * FP64 under FAST; AUTO is FAST for FP32 and REFERENCE for FP64. */
enum { RPL_KERNEL_AUTO = 0, RPL_KERNEL_REFERENCE = 1, RPL_KERNEL_FAST = 2 };

typedef struct rpl_config {
    int32_t nx, ny;        /* physical interior cells (periodic x and y) */
    int32_t nkappa, nphi;  /* spectral bins */
    int32_t dtype;         /* RPL_F32 | RPL_F64: the action's element type */
    int32_t timestepper;   /* RPL_RK3 | RPL_LSRK3 | RPL_FORWARD_EULER */
    int32_t transport;     /* RPL_TRANSPORT_* */
    int32_t kernel;        /* RPL_KERNEL_* */
} rpl_config;

typedef struct rpl_ctx rpl_ctx;

/* SpectralWaveModel construction (Models/spectral_wave_model.jl:155-211):
 * allocates the action, its RK3 stage buffers and coupling caches on the
 * current HIP device.  Requires nx, ny, nphi >= 3 (the reference's periodic
 * wraps, wave_current_refraction.jl:66, reach 3 cells) and nkappa >= 1. */
int rpl_create(rpl_ctx **out, const rpl_config *cfg);
int rpl_destroy(rpl_ctx *ctx);

/* Spectral and physical geometry: PolarWaveVectorGrid / FrequencyDirectionGrid
 * centres and faces (CoordinateGrids/polar_wavevector_grid.jl:18-56,
 * frequency_direction_grid.jl:20-69; for frequency grids pass the derived
 * kappa), first-cell spacings dx, dy (wave_current_refraction.jl:200-203) and
 * depth.  Builds the cell measures and the intrinsic group-velocity tables
 * (Models/transport.jl:111-139) and the cos/sin tables
 * (wave_current_refraction.jl:154-175). */
int rpl_set_geometry(rpl_ctx *ctx, double dx, double dy,
                     const double *kappa_c, const double *kappa_f,
                     const double *phi_c, const double *phi_f,
                     int32_t depth_kind, double depth, double gravity);

/* A spatially varying depth (Oceananigans Field or function on the
 * horizontal grid as SpectralWaveModel's depth; the reference's
 * is_spatially_varying_depth, Models/transport.jl:102-139): depth holds
 * nx ny positive values, i fastest.  The group velocity becomes the
 * point-sampled finite-depth one of every column at its own depth
 * (transport.jl:72-83, 111-139; intrinsic_velocity_component of the 4-D
 * table, :105-109), and a prescribed current set afterwards takes each
 * column's depth as its Q depth (q_depth_at, q_transform.jl:124-125;
 * q_projection_depth, lagrangian_velocities.jl:192-193).  Call after
 * rpl_set_geometry and before rpl_set_current (a current set before is
 * dropped); NULL returns to rpl_set_geometry's depth.  Non-positive values:
 * RPL_ERR_ARG (lagrangian_velocities.jl:123-124).  A pseudomomentum coupling
 * set afterwards builds one pair of overlap tables per column
 * (pseudomomentum.jl:72-87), and rpl_pseudomomentum /
 * rpl_pseudomomentum_tendency with q_depth = 0 take each column's depth.
 * Ensembles take a constant depth (RPL_ERR_UNSUPPORTED). */
int rpl_set_depth_field(rpl_ctx *ctx, const double *depth);

/* Prescribed Lagrangian-mean current (Coupling/current_coupling.jl:73-91,
 * 127-138): builds the Q weights W[k,m] = CDF(z_{k+1}) - CDF(z_k)
 * (q_kernel.jl:77-92) and runs the Doppler contraction
 * U[i,j,m] = sum_k u[i,j,k] W[k,m] (doppler_velocity.jl:5-23,45-62) and the
 * current gradients (wave_current_refraction.jl:7-39) on the device.
 * u, v: (nx, ny, nz) in FP64; z_faces: nz+1 values; q_depth > 0. */
int rpl_set_current(rpl_ctx *ctx, int32_t nz, const double *z_faces,
                    const double *u, const double *v, double q_depth);

/* Direct access to the Doppler velocity caches Ux, Uy (nx, ny, nkappa), FP64
 * at the boundary (the reference's coupling.Ux / coupling.Uy). */
int rpl_set_doppler(rpl_ctx *ctx, const double *Ux, const double *Uy);
int rpl_get_doppler(rpl_ctx *ctx, double *Ux, double *Uy);

/* Self-coupled pseudomomentum current: CWCMPseudomomentumCoupling
 * (Coupling/current_coupling.jl:93-121), built by build_coupling for
 * PseudomomentumVelocities (lagrangian_velocities.jl:294-311).  Builds the
 * overlap O[t,s] = sum_k W_t,k W_s,k / |dz_k| and its target-kappa derivative
 * (pseudomomentum.jl:68-122, constant depth) and clears U and dU/dkappa.  From
 * then on every tendency evaluation first recomputes, from its own input N,
 *   U[i,j,t] = sum_{n,s} O[t,s] N[i,j,s,n] mu_x[s,n]  (Uy with mu_y)
 * and dU/dkappa with the derivative overlap (compute_tendencies!,
 * Models/tendencies.jl:103-107; pseudomomentum.jl:124-140, 316-344), then the
 * current gradients and the refraction stage.  Needs RPL_TRANSPORT_REFRACTION.
 * rpl_set_current / rpl_set_doppler return the context to a prescribed
 * current.  Ensembles reject members with this coupling. */
int rpl_set_pseudomomentum_coupling(rpl_ctx *ctx, int32_t nz, const double *z_faces, double q_depth);

/* SpatialAveraging(; alpha_s, alpha_n) (Models/propagation_smoothing.jl:22-51),
 * the Tolman (2002) garden-sprinkler alleviation that
 * SpectralWaveModel(...; propagation_smoothing) applies after every time step
 * (time_step.jl:167; propagation_smoothing.jl:160-199): each (kappa, phi)
 * bin's x-y field becomes the mean of four bilinear interpolations at the
 * corners of a rectangle with half-widths alpha_s |dcg| dt along the bin's
 * direction and alpha_n cg dphi dt across it (cg = sqrt(9.81/kappa)/2,
 * propagation_smoothing.jl:117-135).  enable = 0 turns it off.  Needs
 * rpl_set_geometry first; a new geometry turns it off.  Ensembles reject
 * members with smoothing. */
int rpl_set_propagation_smoothing(rpl_ctx *ctx, int32_t enable, double alpha_s, double alpha_n);

/* The x and y topology of the model's RectilinearGrid (Oceananigans'
 * topology(grid)): 1 = Periodic, 0 = Bounded; a new context is periodic in
 * both.  Replaces the periodic_x / periodic_y flags the reference's kernels
 * read from the grid (propagation_smoothing.jl:175-177,
 * intrinsic_transport.jl:109-111).  On a bounded axis SpatialAveraging
 * clamps its corner indices to the edge cells (propagation_smoothing.jl:
 * 56, 97-100).  The refraction kernel wraps x and y on either topology, as
 * the reference's does (wave_current_refraction.jl:7-19, 97-111); the fused
 * intrinsic kernel runs on Periodic x Periodic grids only (tendencies.jl:
 * 45-50; the reference takes Oceananigans' per-bin advection otherwise), so
 * a bounded axis with intrinsic transport returns RPL_ERR_UNSUPPORTED. */
int rpl_set_topology(rpl_ctx *ctx, int32_t periodic_x, int32_t periodic_y);

/* update_coupling!(model) (current_coupling.jl:123-148): recomputes the
 * pseudomomentum coupling's U and dU/dkappa from the current action; a no-op
 * for a prescribed current (static). */
int rpl_update_coupling(rpl_ctx *ctx);

/* coupling.dUxdkappa / coupling.dUydkappa (nx, ny, nkappa), FP64 at the
 * boundary: for a prescribed current sum_k u W'[k,m] with the kappa
 * derivative of the Q weights (doppler_velocity.jl:25-43, 64-82); for the
 * pseudomomentum coupling the contraction with the derivative overlap.
 * RPL_ERR_STATE after rpl_set_doppler (no derivative known). */
int rpl_get_doppler_kappa_derivative(rpl_ctx *ctx, double *dUx, double *dUy);

/* compute_pseudomomentum_cell_averages / _cell_integrals
 * (pseudomomentum.jl:162-210, 375-400) of the current action on the vertical
 * cells z_faces[0..nz]:
 *   p[i,j,k] = scale_k sum_{n,m} W[k,m] N[i,j,m,n] mu_x[m,n]  (py with mu_y)
 * with scale_k = 1/|dz_k| when cell_average is nonzero, else 1.  px, py:
 * (nx, ny, nz) FP64, i fastest. */
int rpl_pseudomomentum(rpl_ctx *ctx, int32_t nz, const double *z_faces, double q_depth, int32_t cell_average,
                       double *px, double *py);

/* compute_pseudomomentum_tendency_cell_averages! and
 * cwcm_momentum_tendency_fields! (pseudomomentum.jl:219-283, 402-436): the
 * cell-averaged pseudomomentum of the action tendency (N - old_N) / dt,
 * with N the current action and old_N a host array of the context's element
 * type and layout, times coefficient (1 for the tendency fields, -1 for the
 * CWCM momentum forcing).  dt > 0, else RPL_ERR_ARG (pseudomomentum.jl:225).
 * Uses the stage buffers: call between steps.  ptx, pty: (nx, ny, nz) FP64. */
int rpl_pseudomomentum_tendency(rpl_ctx *ctx, const void *old_N, double dt, int32_t nz, const double *z_faces,
                                double q_depth, double coefficient, double *ptx, double *pty);

/* set!(model; N=array) (ProductFields/set.jl:20-30, 44-48) and interior(N)
 * (product_field.jl:221-230).  Element type is the context dtype. */
int rpl_set_action(rpl_ctx *ctx, const void *N);
int rpl_get_action(rpl_ctx *ctx, void *N);

/* compute_tendencies!(model) (Models/tendencies.jl:109): one evaluation of
 * G = L(N) on the current action, copied to the host (the parity hook). */
int rpl_tendency(rpl_ctx *ctx, void *G);

/* time_step!(model, dt) repeated nsteps times (Models/time_step.jl:119-170):
 * dt > 0, else RPL_ERR_ARG (time_step.jl:120).  Enqueued, not waited for. */
int rpl_time_step(rpl_ctx *ctx, double dt, int32_t nsteps);
int rpl_synchronize(rpl_ctx *ctx);

/* model.clock (Oceananigans Clock): time and iteration, advanced once per
 * step (time_step.jl:168). */
int rpl_get_clock(rpl_ctx *ctx, double *time, int64_t *iteration);
int rpl_set_clock(rpl_ctx *ctx, double time, int64_t iteration);

/* A non-finite check of the action after every interval-th step (a flag,
 * not a repair): one pass over N counts its NaN and infinite values on the
 * step's stream, and the first checked iteration that found any is kept
 * with its count.  interval = 0 turns it off; setting it clears the record.
 * Steps of ensemble members (rpl_ensemble_step) are not checked.  No
 * counterpart in the reference, which clips negative action
 * (time_step.jl:72,103) and detects nothing; Oceananigans' NaNChecker
 * callback is the nearest analogue. */
int rpl_set_nonfinite_check(rpl_ctx *ctx, int32_t interval);
/* The record: first_iteration = -1 (and count = 0) if no check has found a
 * non-finite value since rpl_set_nonfinite_check.  Waits for the stream. */
int rpl_get_nonfinite(rpl_ctx *ctx, int64_t *first_iteration, int64_t *count);

/* m0, root_mean_square_wavenumber and mean_direction
 * (Diagnostics/moments.jl:33-123, diagnostic_operations.jl:56-165): three
 * (nx, ny) FP64 arrays; any pointer may be NULL. */
int rpl_diagnostics(rpl_ctx *ctx, double *m0, double *kappa_rms, double *mean_dir);

/* total_action (Diagnostics/bulk_statistics.jl:87-91) and
 * total_deep_water_energy (bulk_statistics.jl:93-99), both accumulated on
 * the device in FP64 with compensated (double-double) summation. */
int rpl_conservation(rpl_ctx *ctx, double *total_action, double *total_energy);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the compensated FP64 total of rpl_conservation,
 *   *out = sum_ij area * sum_mn N[i,j,m,n] * measure[m,n]
 * (total_action with the cell measures w, total_deep_water_energy with the
 * energy measures), over a caller-owned device action N (nx*ny*nkappa*nphi
 * elements of dtype, i fastest) and caller-owned device measures (nkappa*nphi
 * FP64, m fastest), written to one device double `out`, enqueued on `stream`
 * (a hipStream_t; NULL for the null stream) of the current device without
 * synchronizing.  Nothing is allocated: the per-workgroup partial sums go to
 * a library-held device array of 1024 (hi, lo) pairs (16 KiB per GPU, held
 * until the process ends; the launch grid never exceeds it).  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (other
 * streams or threads) race; calls in order on one stream do not.  Limits:
 * nx * ny < 2^30 and nkappa * nphi < 2^30 (RPL_ERR_ARG beyond). */
int rpl_weighted_total_noalloc(const void *N, int32_t dtype, int32_t nx, int32_t ny, int32_t nkappa, int32_t nphi,
                               const double *measure, double area, double *out, void *stream);

/* Source terms of the PrecomputedSources bundle
 * (Physics/Packages/precomputed_sources.jl:23-283), evaluated on the device on
 * every stage input and added to the transport tendency
 * (Models/tendencies.jl:88-94):
 *   wind input  PressureCorrelationInput (Physics/WindInput/pressure_correlation.jl:24-156)
 *               with BulkWindDrag(:linear | :capped) (Physics/shared/drag.jl:21-34)
 *               and, when stress_cap = 1, the bundle's wave-supported-stress
 *               cap (precomputed_sources.jl:126-170);
 *   dissipation MeanSpectrumWhitecapping (Physics/Dissipation/mean_spectrum.jl:20-140),
 *               LocalSaturationDissipation (Physics/Dissipation/local_saturation.jl:18-52)
 *               or WhitecappingDissipation (Sources/Sources.jl:105-122, 583-598).
 * Winds: a uniform speed and direction (the reference's `wind::Number` with
 * `direction`), or HollandHurricaneWind on a two-point LinearStormTrack
 * (Forcing/winds.jl:4-41, 89-198), evaluated per column at the clock time of
 * each step. */
enum { RPL_WIND_INPUT_NONE = 0, RPL_WIND_INPUT_PRESSURE_CORRELATION = 1 };
enum { RPL_DRAG_LINEAR = 0, RPL_DRAG_CAPPED = 1 };
enum { RPL_WIND_UNIFORM = 0, RPL_WIND_HOLLAND = 1 };
enum {
    RPL_DISSIPATION_NONE = 0,
    RPL_DISSIPATION_MEAN_SPECTRUM = 1,
    RPL_DISSIPATION_LOCAL_SATURATION = 2,
    RPL_DISSIPATION_WHITECAPPING = 3
};

typedef struct rpl_sources {
    int32_t wind_input, drag, wind_field, stress_cap, dissipation, rotation; /* rotation: +1 ccw, -1 cw */
    double wind_speed, wind_direction;                                       /* RPL_WIND_UNIFORM */
    double beta_max, z_alpha, p_in, alpha0, rho_air, rho_water, von_karman, gravity, cap_epsilon;
    double vmax, rmax, radius, shape, inflow_angle, background_x, background_y; /* RPL_WIND_HOLLAND */
    double track_t0, track_t1, track_x0, track_y0, track_x1, track_y1;
    double cds, delta1, delta2, p_mean;          /* MeanSpectrumWhitecapping */
    double ls_cds, ls_br, ls_p, ls_sigma_power;  /* LocalSaturationDissipation */
    /* WhitecappingDissipation (Sources/Sources.jl:105-122, 583-598) with a
     * constant rate: lambda = rate * max(m0 / threshold - 1, 0)^saturation_power
     * * <kappa^wavenumber_power>_cell / reference_wavenumber^wavenumber_power. */
    double wc_rate, wc_threshold, wc_saturation_power, wc_wavenumber_power, wc_reference_wavenumber;
    /* DiscreteInteractionApproximation(SymmetricQuadruplet(; coefficient,
     * frequency_offset, gravity)) as the bundle's nonlinear member
     * (Physics/NonlinearInteractions/symmetric_quadruplet.jl:30-238,
     * precomputed_sources.jl:227-233, 256-259): RPL_NONLINEAR_*.  Meant for
     * FrequencyDirectionGrid spectra (geometric frequencies, uniform phi
     * from phi[0]), as the reference's bilinear lookup assumes. */
    int32_t nonlinear;
    double dia_coefficient, dia_frequency_offset, dia_gravity;
} rpl_sources;

enum { RPL_NONLINEAR_NONE = 0, RPL_NONLINEAR_SYMMETRIC_QUADRUPLET = 1 };

/* Installs (or, with NULL, removes) the source terms.  x_nodes / y_nodes are
 * the nx / ny cell-centre coordinates the winds are evaluated at. */
int rpl_set_sources(rpl_ctx *ctx, const rpl_sources *src, const double *x_nodes, const double *y_nodes);

/* LinearStormTrack of n >= 1 points (Forcing/winds.jl:4-41): the Holland
 * centre follows the piecewise-linear track (clamped to the end points)
 * instead of the two track points of rpl_sources.  n = 0 removes it.  Times
 * must increase strictly (RPL_ERR_ARG otherwise, as the reference's
 * ArgumentError). */
/* DiscreteInteractionApproximation(interactions; power, resonance_tolerance)
 * with the ListedQuadruplets method (Sources/Sources.jl:374-428, 1268-1385):
 * n explicit interactions, bins[8 q .. 8 q + 7] = donor1 (m, n), donor2,
 * receiver1, receiver2 (0-based), constant rates[q] >= 0; frequency holds
 * the nkappa frequency centers of the FrequencyDirectionGrid.  Checked as the
 * reference checks them (indices, distinct donor and receiver pairs,
 * frequency and wavevector resonance within the tolerance, positive power).
 * The source enters every tendency evaluation; n = 0 removes it.  One
 * nonlinear term at a time (not with rpl_sources.nonlinear). */
int rpl_set_quadruplets(rpl_ctx *ctx, int32_t n, const int32_t *bins, const double *rates,
                        const double *frequency, double power, double resonance_tolerance);

/* Rates of the listed interactions given per column: the reference's
 * source_value (Sources/Sources.jl:1450-1467) of a QuadrupletTransferInteraction
 * rate that is an (Nx, Ny) array (value[i, j]) or a function of position and
 * time (value(x, y, t), value(x, y), value(t), evaluated at the cell centre
 * and model.clock.time, which stays at the step's start for all of its
 * stages).  rates holds nx ny values per interaction, column i + nx j
 * fastest, interaction q at offset q nx ny, all >= 0 (RPL_ERR_ARG
 * otherwise, as the reference's ArgumentError); the host evaluates a
 * time-dependent rate and calls this before each step.  NULL returns to the
 * constant rates of rpl_set_quadruplets; a new list drops the field.  Needs
 * rpl_set_quadruplets first. */
int rpl_set_quadruplet_rate_field(rpl_ctx *ctx, const double *rates);

int rpl_set_storm_track(rpl_ctx *ctx, int32_t n, const double *times, const double *x, const double *y);

/* Per-column wind state of the last step: U10, wind angle, friction velocity,
 * (nx, ny) FP64 each (the reference's _compute_wind_input_state host loop). */
int rpl_get_wind(rpl_ctx *ctx, double *U10, double *theta_u, double *u_star);

/* Learned source surrogate.  No counterpart in the reference: a small MLP
 * (in_dim -> hidden x hidden_layers -> out_dim, GELU) stands in for the
 * per-column source evaluation of the PrecomputedSources bundle above and is
 * trained on that bundle's own tendencies.  BF16 weights and activations on
 * the matrix cores with FP32 accumulation; FP32 master weights and Adam.
 * Parameters (get/set, gradients) are the concatenation over layers of the
 * padded W_l [out_pad_l x in_pad_l] (row-major) and b_l [out_pad_l]; padded
 * dims are multiples of 64 (rpl_mlp_dims).  X rows are in_dim features, Y / T
 * rows out_dim values, all FP32 on the host. */
typedef struct rpl_mlp rpl_mlp;
int rpl_mlp_create(rpl_mlp **out, int32_t in_dim, int32_t hidden, int32_t hidden_layers, int32_t out_dim,
                   int32_t batch_capacity, uint64_t seed);
int rpl_mlp_destroy(rpl_mlp *net);
int rpl_mlp_dims(rpl_mlp *net, int32_t *in_pad, int32_t *out_pad, int64_t *nparams);
int rpl_mlp_set_params(rpl_mlp *net, const float *params);
int rpl_mlp_get_params(rpl_mlp *net, float *params);
/* Inference: the kernels the RK3 stage runs (hidden activations rounded to
 * bf16 in the GEMM epilogue, the output layer computed transposed). */
int rpl_mlp_forward(rpl_mlp *net, int32_t batch, const float *X, float *Y);
/* Loss = mean over the batch and out_dim of (Y - T)^2; grads of the loss. */
int rpl_mlp_gradients(rpl_mlp *net, int32_t batch, const float *X, const float *T, float *loss, float *grads);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch. */
int rpl_mlp_train_step(rpl_mlp *net, int32_t batch, const float *X, const float *T, float lr, float *loss);
/* `steps` Adam steps on a device-resident training set: X (rows x in_dim)
 * and T (rows x out_dim), FP32 row-major DEVICE pointers on the network's
 * device.  Each step draws `batch` rows uniformly with replacement (a
 * counter-based hash of seed and the network's step count), so no host
 * traffic is involved; the learning rate decays geometrically from lr0 to
 * lr1 over the call.  losses (host, `steps` values, may be null) receives
 * each step's mean squared error (reading it synchronizes every step). */
int rpl_mlp_train_device(rpl_mlp *net, int64_t rows, const float *X, const float *T, int32_t batch, int32_t steps,
                         float lr0, float lr1, uint64_t seed, float *losses);

/* Surrogate use in the time step: on every stage the network maps each
 * column's features to per-bin rates r (source = r * rate_unit * sigma_bin * N).
 * Columns outside the training envelope (U10 outside [wind_lo, wind_hi], an
 * action above action_hi or above the bin's bin_action_hi) or with an
 * output that is not finite in FP16 (the stage reads the rates in FP16) use
 * the plain bundle,
 * which must be installed with rpl_set_sources (with a wind input).  Every
 * audit_stride-th column (0: none) is audited: its network source is compared
 * with the plain bundle's, and where the relative L2 error exceeds
 * audit_bound the column takes the plain source.  Features: (log(N + floor)
 * - shift) * scale per bin, U10 (cos, sin) * wind_scale, depth_feature.
 * net = NULL detaches.  The network must outlive its attachment. */
typedef struct rpl_surrogate {
    float floor, shift, scale, wind_scale, depth_feature;
    float rate_unit;
    float wind_lo, wind_hi, action_hi;
    int32_t audit_stride;
    float audit_bound;
    /* 0: the network supplies the whole source; 1: the wind input only (with
     * its stress cap), the dissipation staying the plain bundle's. */
    int32_t terms;
    /* 0: the rates are inferred for every RK3 stage input (default); 1: once
     * per step, for the step's first stage, and held through its later
     * stages (the source stays rate * N of each stage input; columns on the
     * plain bundle are still evaluated per stage).  An approximation that
     * trades accuracy for a third of the inference. */
    int32_t refresh;
} rpl_surrogate;
/* bin_action_hi (nkappa*nphi values, may be NULL): the per-bin largest
 * action of the training set; a column with any bin above it is outside the
 * envelope. */
int rpl_attach_surrogate(rpl_ctx *ctx, rpl_mlp *net, const rpl_surrogate *cfg, const float *bin_action_hi);
/* Columns evaluated, of which by the plain bundle, audited, audit failures,
 * and the largest audited error since the last call (then reset). */
int rpl_surrogate_stats(rpl_ctx *ctx, int64_t *columns, int64_t *plain, int64_t *audited, int64_t *failed,
                        double *err_max);
/* The surrogate features of the current action and clock, (nx*ny) rows of
 * nbins + 3 FP32 values (bf16-representable: what the network sees), and the
 * plain bundle's per-cell rates r = source / (N sigma_bin rate_unit) as
 * network targets (for N <= 1e-30, the rate at N = 1e-30; with cfg->terms 1,
 * of the wind input alone), (nx*ny) rows of nbins.  Either output
 * may be NULL.  Column order is i fastest; nbins = nkappa * nphi, m fastest. */
int rpl_surrogate_training_pairs(rpl_ctx *ctx, const rpl_surrogate *cfg, float *features, float *targets);

/* Per-kernel timing of the fused stage kernel: when enabled, HIP events on
 * the context stream bracket every stage launch; the accumulated elapsed
 * milliseconds and launch count are read (and reset) with
 * rpl_get_kernel_timing, which synchronizes. */
int rpl_enable_kernel_timing(rpl_ctx *ctx, int32_t enable);
int rpl_get_kernel_timing(rpl_ctx *ctx, double *total_ms, int64_t *launches);
/* The same for the surrogate inference of each stage input (features, the
 * network forward pass over all chunks, and the envelope flags), under the
 * same enable switch.  No counterpart in the reference. */
int rpl_get_surrogate_timing(rpl_ctx *ctx, double *total_ms, int64_t *evaluations);

/* Device time of the pseudomomentum coupling updates (one per tendency
 * evaluation) since the last call, under the same enable switch.  No
 * counterpart in the reference. */
int rpl_get_coupling_timing(rpl_ctx *ctx, double *total_ms, int64_t *updates);

/* Device time of the propagation-smoothing passes (one per step) since the
 * last call, under the same enable switch.  No counterpart in the reference. */
int rpl_get_smoothing_timing(rpl_ctx *ctx, double *total_ms, int64_t *passes);

/* Device time of the source bundle's per-column state launches (one per
 * tendency evaluation: the stress cap's wave-supported stress and the
 * dissipation's column moments, precomputed_sources.jl:41-218) since the last
 * call, under the same enable switch.  No counterpart in the reference. */
int rpl_get_source_timing(rpl_ctx *ctx, double *total_ms, int64_t *launches);

/* Device time of the nonlinear (DIA) source launches (one per tendency
 * evaluation) since the last call, under the same enable switch.  No
 * counterpart in the reference. */
int rpl_get_nonlinear_timing(rpl_ctx *ctx, double *total_ms, int64_t *launches);

/* Ensemble batching.  No counterpart in the reference, which runs one model
 * per process.  Members are ordinary contexts -- FP32 on the fast kernel,
 * SSP-RK3, one transport kind (intrinsic or refraction) and either all with
 * or all without source terms -- whose sizes and spectral grids may differ.
 * rpl_ensemble_step advances every member together: per RK3 stage, the
 * members' source preparation (column state, surrogate inference), then the
 * stage kernel of each large member (as rpl_time_step launches it) and ONE
 * persistent-kernel launch over a work list of the other members' tiles (64 x
 * tile_rows cells of one bin).  While in an ensemble, a member's work runs on
 * the ensemble's stream; members must outlive the ensemble, and are used
 * through the ensemble only (their own calls still work and are ordered on
 * that stream).  Each member's clock advances as with rpl_time_step. */
typedef struct rpl_ensemble rpl_ensemble;
int rpl_ensemble_create(rpl_ensemble **out, rpl_ctx *const *members, int32_t count);
int rpl_ensemble_destroy(rpl_ensemble *ens);
int rpl_ensemble_step(rpl_ensemble *ens, double dt, int32_t nsteps);
int rpl_ensemble_synchronize(rpl_ensemble *ens);
/* Launch shape: tile_rows 2, 4 or 8 (0 keeps the current one) and the grid
 * size in workgroups (a multiple of 8: a persistent grid sweeping the work
 * list; 0: one workgroup per tile, the default). */
int rpl_ensemble_set_launch(rpl_ensemble *ens, int32_t tile_rows, int32_t blocks);
int rpl_ensemble_get_launch(rpl_ensemble *ens, int32_t *tile_rows, int32_t *blocks, int64_t *tiles);
/* Which members are large: those with at least min_tiles tiles of 64 x 4
 * cells x one kappa bin (0: the default, 4096; negative: none, every member
 * joins the work list).  The getter returns the threshold in force and how
 * many members run their own stage kernel. */
int rpl_ensemble_set_split(rpl_ensemble *ens, int64_t min_tiles);
int rpl_ensemble_get_split(rpl_ensemble *ens, int64_t *min_tiles, int32_t *own_members);
/* Times stage 1 (which leaves the members' state unchanged) for the splits
 * {none, 16384, 4096, 1024, 1 (every member on its own chain)} and, for the
 * work list, tile_rows in {2, 4, 8}
 * and grids of {2, 4, 8, 16} workgroups per CU or one workgroup per tile,
 * reps launches each; keeps the fastest and returns its shape (blocks 0: one
 * workgroup per tile; the split through rpl_ensemble_get_split). */
int rpl_ensemble_autotune(rpl_ensemble *ens, double dt, int32_t reps, int32_t *tile_rows, int32_t *blocks,
                          double *best_ms);
/* Capture the three stages of a step (source preparation + persistent
 * launches) in a hipGraph and replay it per step; the wind, which depends on
 * the clock, is launched before each replay.  Recaptured when dt or the launch
 * shape changes; call again after changing a member's sources or surrogate. */
int rpl_ensemble_use_graph(rpl_ensemble *ens, int32_t enable);

/* The context's hipStream_t (for callers that order their own work after it). */
void *rpl_stream(rpl_ctx *ctx);

/* Message of this thread's last failed call ("" if none). */
const char *rpl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RIPPLE_MI355X_H */
