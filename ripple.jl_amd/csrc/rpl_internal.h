// Internal declarations shared by the host API (rpl_capi.cpp) and the HIP
// kernel translation units.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/ripple_mi355x.h"
#include "sources_device.h"

namespace rpl {

// Records msg as the thread's last error (rpl_last_error) and returns code.
int set_error(int code, const char *msg);

// Everything the fused tendency + RK3 stage kernels read besides the action.
// T is the action's element type.  Tables are (nkappa, nphi), m fastest.
template <typename T>
struct StageParams {
    int nx, ny, nk, nphi;
    const T *Ux, *Uy;                 // Doppler velocity caches (nx, ny, nk)
    const T *Uxx, *Uxy, *Uyx, *Uyy;   // current gradients (nx, ny, nk)
    const T *kappa;                   // kappa centres (nk)
    const T *cgx, *cgy;               // intrinsic group velocity (nk, nphi)
    // A spatially varying depth (transport.jl:102-139): when cgm is set the
    // group velocity of cell (i, j, m, n) is the point-sampled finite-depth
    // speed cgm[i + nx j + nx ny m] of the column's depth times the bin's
    // direction: (cgm cdx) / cdk in the reference-order kernel (cdx, cdy =
    // kappa cos, kappa sin; cdk = kappa) and cgm cdx in the others (cdx, cdy
    // = cos, sin).  (nk, nphi) tables.
    const T *cgm, *cdx, *cdy, *cdk;
    const T *cosp, *sinp;             // cos/sin of phi centres (nphi)
    T dx_inv, dy_inv, dk_inv, dp_inv; // 1 / first spacing
    T dt;                             // time step in T (time_step.jl:16-20)
    const T *src_add;                 // optional additive tendency (nx,ny,nk,nphi) or null
    SourceParams<T> src;              // source terms (src.on == 0: none)
    // Wave-supported stress of the stage's OUTPUT for the next stage's
    // stress cap (FP32 wind-input bundle): when tau_x is set, the kernel
    // writes per (column, kappa bin) the sums over phi of raw * N_out *
    // tau_c[t] (and tau_s) into tau_x / tau_y, (nx, ny, nk) planes.
    T *tau_x, *tau_y;
    const T *tau_c, *tau_s;           // (k / sigma) w rho_w g (cos theta, sin theta) per bin
    // FP32 per-model phi-walk kernels without the generic source path: the
    // per-bin values above packed 16 floats per bin (BT_* below), filled
    // from the separate tables before every stage launch, so that a bin's
    // values are one scalar load and one base address (null: not used).
    const float *bt;
};

// Fields of a packed bin-table entry.
enum { BT_CGX = 0, BT_CGY, BT_K, BT_CTH, BT_STH, BT_LOGK, BT_INVC, BT_SIGMA, BT_SK3, BT_SQ, BT_TAUC, BT_TAUS,
       BT_OUTSCALE, BT_COSP, BT_SINP, BT_FIELDS = 16 };

// One ensemble member's stage launch in the work-list kernel: its stage
// parameters, buffers and work-list numbering (64 x TY tiles, nby =
// ceil(ny / TY); one entry per tile and kappa bin for refraction transport,
// which walks the phi bins inside the workgroup, one per tile and (kappa,
// phi) bin otherwise; tile0 = entries of all earlier members).
struct EnsembleMember {
    StageParams<float> p;
    const float *Nk, *N0;
    float *out;
    int nbx, nby;
    long long tile0;
};

// Wind of one step (Forcing/winds.jl), FP64.
struct WindParams {
    int field, drag, rotation;
    double speed, direction;                               // uniform
    double vmax, rmax, radius, shape, decay, bx, by;       // Holland
    double cos_inflow, sin_inflow;
    double t0, t1, x0, y0, x1, y1;                         // two-point storm track
    double alpha0, gravity;                                // z1 = alpha0 u*^2 / g of the wind input
};

// Tables of the per-column state kernel (precomputed_sources.jl:41-74, 126-170).
template <typename T>
struct StateTables {
    const T *kappa, *phi, *w;           // centres, cell measures (nk, nphi)
    const T *sigma_k, *cos_phi, *sin_phi;
    const T *k_pow, *sigma_pow;         // kappa^p, sigma^p (mean-spectrum moments)
    const T *log_kappa, *inv_c;         // FP32 wind input: log kappa, kappa / sigma (per kappa bin)
    T rho_air, rho_w_g, cap_eps, inv_p;
};

// One member's column-state launch in an ensemble (FP32): inputs, bundle
// parameters, outputs, and its first 64-column block in the work list.
// tau_x set: the stress factor comes from the previous stage's
// wave-supported stress sums (stress_factor_from_sums) instead of a pass
// over N.
struct EnsembleState {
    int nx, ny, nk, nphi;
    const float *N;
    SourceParams<float> s;
    StateTables<float> tb;
    float *factor, *m0, *kbar, *sbar;
    const float *tau_x, *tau_y;
    long long blk0;
};

// The stress factor of every column from the per-kappa wave-supported stress
// sums a stage kernel wrote for its output (StageParams::tau_x / tau_y),
// summed over kappa in bin order: the bundle cap (precomputed_sources.jl:
// 126-170) without a second pass over N.
hipError_t launch_stress_factor_from_sums(int nx, int ny, int nk, const float *tau_x, const float *tau_y,
                                          const SourceParams<float> &s, const StateTables<float> &tb, float *factor,
                                          hipStream_t st);


template <typename T>
hipError_t launch_wind_columns(int nx, int ny, const double *xn, const double *yn, double t, const WindParams &w,
                               T *U10, T *theta_u, T *ustar, T *cos_tu, T *sin_tu, T *log_z1, hipStream_t s);
template <typename T>
hipError_t launch_source_rates(int nx, int ny, int nbins, const T *N, const SourceParams<T> &sp, const float *unit,
                               float *r, hipStream_t s);
template <typename T>
hipError_t launch_source_state(int nx, int ny, int nk, int nphi, const T *N, const SourceParams<T> &sp,
                               const StateTables<T> &tb, T *factor, T *m0, T *kbar, T *sbar, hipStream_t s);

// Stage selector of the fused stage kernels (time_step.jl:138-152):
//   0  out = L(Nk)                                   (tendency only)
//   1  out = max(0, Nk + dt L(Nk))                    (RK3 stage 1, Euler)
//   2  out = max(0, 3/4 N0 + 1/4 max(0, Nk + dt L))   (RK3 stage 2)
//   3  out = max(0, 1/3 N0 + 2/3 max(0, Nk + dt L))   (RK3 stage 3)
// N0 is read only at the output cell, so stage 3 may write in place over N0.
enum { STAGE_TENDENCY = 0, STAGE_1 = 1, STAGE_2 = 2, STAGE_3 = 3 };

hipError_t launch_stage_reference(const StageParams<double> &p, int transport, int stage,
                                  const double *Nk, const double *N0, double *out, hipStream_t s);
hipError_t launch_stage_reference(const StageParams<float> &p, int transport, int stage,
                                  const float *Nk, const float *N0, float *out, hipStream_t s);
// Column states of every member in one launch (`blocks` 64-column blocks in all).
hipError_t launch_ensemble_state(const EnsembleState *mem, int members, long long blocks, hipStream_t s);
// One launch of stage `stage` (STAGE_1..3) over every member's tiles;
// `blocks` (a multiple of 8) workgroups of 64 x ty threads; chunk_member[c]
// is the member holding tile c << ENSEMBLE_CHUNK_SHIFT.
constexpr int ENSEMBLE_CHUNK_SHIFT = 8;

hipError_t launch_ensemble_stage(const EnsembleMember *mem, int members, long long total, const int *chunk_member,
                                 int transport, int src_kind, int stage, int ty, int blocks, hipStream_t s);
hipError_t launch_stage_fast(const StageParams<float> &p, int transport, int stage,
                             const float *Nk, const float *N0, float *out, hipStream_t s);
// The phi-loop stage in FP64 (refraction and intrinsic transport): the FP32
// kernel's restructured arithmetic, not the reference operation order.
hipError_t launch_stage_fast(const StageParams<double> &p, int transport, int stage, const double *Nk,
                             const double *N0, double *out, hipStream_t s);

// Coupling: U[i,j,m] = sum_k u[i,j,k] W[k,m] (W is (nz, nk), k fastest) and the
// centred periodic gradients of U.
hipError_t launch_doppler(int nx, int ny, int nz, int nk, const double *u, const double *v,
                          const double *W, double *Ux, double *Uy, hipStream_t s);
hipError_t launch_doppler(int nx, int ny, int nz, int nk, const float *u, const float *v,
                          const float *W, float *Ux, float *Uy, hipStream_t s);
template <typename T>
hipError_t launch_doppler_mfma(int nx, int ny, int nz, int nk, const T *u, const T *v, const T *W, T *Ux, T *Uy,
                               hipStream_t s);
// Doppler velocity of a prescribed current with a depth per column
// (doppler_velocity.jl:45-62 with q_depth_at(depth, i, j), q_transform.jl:
// 124-125): each column's Q weights (or their kappa derivative) formed on
// the device at its depth, kc the kappa centres, zf the nz + 1 faces.
template <typename T>
hipError_t launch_doppler_columns(int nx, int ny, int nz, int nk, const T *u, const T *v, const double *zf,
                                  const double *kc, const double *depth, bool derivative, T *Ux, T *Uy,
                                  hipStream_t s);
template <typename T>
hipError_t launch_gradients(int nx, int ny, int nk, const T *Ux, const T *Uy, T dx_inv, T dy_inv,
                            T *Uxx, T *Uxy, T *Uyx, T *Uyy, hipStream_t s);

// Pseudomomentum contractions X[i,j,o] = scale_o sum_{n,s} A[o,s] N[i,j,s,n] kx[s,n]
// (Y with ky; dX, dY with B when B is non-null; scale may be null).  ordered
// keeps the reference's accumulation order; otherwise the sums go through the
// column's directional moments.  The factored kernel holds the moments of 64
// columns and the matrices in LDS (pm_factored_lds bytes, at most
// PM_MOMENT_LDS_MAX); shapes beyond that take the ordered kernel.
constexpr size_t PM_MOMENT_LDS_MAX = 64 * 1024;
inline size_t pm_factored_lds(int nk, int nout, bool with_b, size_t esize)
{
    return (2 * (size_t)nk * 64 + (size_t)nk * nout * (with_b ? 2 : 1)) * esize;
}
// per_column: A and B hold one matrix per column, A[ij + nx ny (o + nout s)]
// (a depth per column); the ordered kernel then serves every shape.
template <typename T>
hipError_t launch_pm_contract(bool ordered, int nx, int ny, int nk, int nphi, int nout, const T *N, const T *A,
                              const T *B, const T *kx, const T *ky, const T *scale, T *X, T *Y, T *dX, T *dY,
                              hipStream_t s, bool per_column = false);
// The pseudomomentum overlap tables and Q weights of every column at its own
// depth (FP64 weights, cast to T): O, dO[ij + plane (t + nk s)],
// W[ij + plane (k + nz m)]; zf, kc and depth on the device.
template <typename T>
hipError_t launch_pm_overlap_columns(size_t plane, int nz, int nk, const double *zf, const double *kc,
                                     const double *depth, T *O, T *dO, hipStream_t s);
template <typename T>
hipError_t launch_q_weights_columns(size_t plane, int nz, int nk, const double *zf, const double *kc,
                                    const double *depth, T *W, hipStream_t s);

// out = (Nnew - Nold) / dt elementwise (pseudomomentum.jl:412).
template <typename T>
hipError_t launch_action_tendency(size_t n, const T *Nnew, const T *Nold, T dt, T *out, hipStream_t s);

// SpatialAveraging (Models/propagation_smoothing.jl:56-110): Nout = mean of
// four bilinear corner interpolations of Nin per (kappa, phi) bin, each axis
// periodic (px, py) or clamped at its edges; cg, dcg per kappa bin, cos/sin
// of the phi centres.  fast (FP32 only) forms the corner weights once per bin
// instead of per cell.
template <typename T>
hipError_t launch_spatial_averaging(bool fast, int nx, int ny, int nk, int nphi, const T *Nin, T *Nout, const T *cg,
                                    const T *dcg, const T *cosp, const T *sinp, T dphi, T as, T an, T dt, T dx_inv,
                                    T dy_inv, bool px, bool py, hipStream_t s);

// DiscreteInteractionApproximation{SymmetricQuadruplet}: per donor bin d and
// orientation o (entry q = 2 d + o, o = 0 for +1, 1 for -1) the receivers'
// corner bins (four for sigma+, four for sigma-), weight times energy factor
// (the gather) and weight over energy factor (the scatter), and flags (bit 0:
// sigma+ on the grid, bit 1: sigma- on the grid); per donor its energy factor,
// its reciprocal and C/g^4 f^11.  Column-independent; built on the host.
template <typename T>
struct DiaTables {
    const int *flags;        // 2 nbins
    const int *bins;         // 16 nbins
    const T *wef, *wdiv;     // 16 nbins each
    const T *ef_a, *inv_ef_a, *cf_a;  // nbins each
    T dal1, dal2, dal3;
    // The scatter transposed: bin r's transfer is the sum over
    // gat_q[gat_off[r] .. gat_off[r + 1]) of TE[q] gat_w (receiver corners
    // and the donor's own loss), TE[q = 2 d + o] the donor d's transfer of
    // orientation o.
    const int *gat_off, *gat_q;
    const T *gat_w;
};
template <typename T>
int dia_columns_per_workgroup(int nbins);
// ListedQuadruplets source: bins 8 per interaction (d1m d1n d2m d2n r1m r1n
// r2m r2n, 0-based), constant rates[q], or per-column rates rate_field
// (nx ny, count), column fastest, when rate_field is set; w the (nk, nphi)
// cell measures.
template <typename T>
hipError_t launch_listed_source(int nx, int ny, int nk, int nphi, int count, const int *bins, const T *rates,
                                const T *rate_field, T power, const T *w, const T *N, T *out, hipStream_t s);
// out = the DIA source (T or -min(-T/max(N, eps), 1) N) of the action N.
template <typename T>
hipError_t launch_dia_source(int nx, int ny, int nk, int nphi, const T *N, const DiaTables<T> &tb, T *out,
                             hipStream_t s);

// Diagnostics: per-column sums with FP64 accumulation; compensated totals.
template <typename T>
hipError_t launch_column_diagnostics(int nx, int ny, int nk, int nphi, const T *N,
                                     const double *w, const double *trace, const double *mux,
                                     const double *muy, double *m0, double *krms, double *mdir,
                                     hipStream_t s);
// Writes 2*nblocks doubles (hi, lo pairs) of the weighted total sum_ij (sum_mn N w) * area.
template <typename T>
hipError_t launch_weighted_total(int nx, int ny, int nk, int nphi, const T *N, const double *w,
                                 double area, double *partials, int nblocks, hipStream_t s);
// The same total into one device double `out`, with the workgroup partials
// in the library's static device scratch (no allocation, no host sync).
template <typename T>
hipError_t launch_weighted_total_noalloc(int nx, int ny, int nk, int nphi, const T *N, const double *w, double area,
                                         double *out, hipStream_t s);
int weighted_total_blocks(int nx, int ny);
// Counts the non-finite values of N (n elements, 16-byte aligned) and, when
// there are any and record[0] < 0, sets record = {iteration, count};
// counter (one device word) is left at zero.
template <typename T>
hipError_t launch_nonfinite_check(size_t n, const T *N, unsigned long long *counter, long long *record,
                                  long long iteration, hipStream_t s);

template <typename T>
hipError_t launch_cast_from_f64(const double *src, T *dst, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_cast_to_f64(const T *src, double *dst, size_t n, hipStream_t s);

}  // namespace rpl
