// Nonlinear four-wave transfer by the discrete interaction approximation
// with one symmetric quadruplet (the reference's
// DiscreteInteractionApproximation{SymmetricQuadruplet},
// Physics/NonlinearInteractions/symmetric_quadruplet.jl:30-238).
//
// Each donor bin (m, n) of a column interacts with two receivers at
// (1 +- lambda) sigma and directions theta +- offsets, read and updated
// through bilinear corners.  Where the receivers fall depends only on the
// spectral grid, so the host tabulates, per donor and orientation, the
// corner bins and weights (DiaTables) and the transposed list of what each
// bin receives; a workgroup then holds the action and the transfer of a few
// adjacent columns in LDS, computes every donor's transfer from its
// receivers' corners, and each bin gathers what its donors send it (no
// atomics).  The last phase turns the transfer into the source the stage
// adds, T where T >= 0 and -min(-T / max(N, eps), 1) N otherwise
// (symmetric_quadruplet.jl:210-222).
#include "rpl_internal.h"

#pragma clang fp contract(off)

namespace rpl {
namespace {

__device__ __forceinline__ long long dia_xcd_order(long long b, long long total)
{
    const long long xcd = b & 7, r = b >> 3, per = total >> 3, rem = total & 7;
    return xcd < rem ? xcd * (per + 1) + r : rem * (per + 1) + (xcd - rem) * per + r;
}

template <typename T>
__device__ __forceinline__ T eps_of();
template <>
__device__ __forceinline__ float eps_of<float>() { return 1.1920928955078125e-7f; }
template <>
__device__ __forceinline__ double eps_of<double>() { return 2.220446049250313e-16; }

// Threads per DIA workgroup: 8 wavefronts share one LDS tile (measured
// 10 % faster than 4 at 256^2 x 40 x 36).
constexpr int DIA_THREADS = 512;

template <typename T>
__global__ __launch_bounds__(DIA_THREADS) void dia_source_kernel(size_t plane, int nbins, int cols, long long tiles,
                                                         const T *__restrict__ N, DiaTables<T> tb, T *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char dia_lds[];
    T *Nl = reinterpret_cast<T *>(dia_lds);
    T *TEl = Nl + (size_t)nbins * cols;  // [2 nbins][column]
    const size_t ij0 = (size_t)dia_xcd_order(blockIdx.x, tiles) * cols;
    const int ncol = (int)min((size_t)cols, plane - ij0);
    const int items = nbins * cols;
    // The columns' action, bin-major ([bin][column]).
    for (int e = threadIdx.x; e < items; e += DIA_THREADS) {
        const int bin = e / cols, cc = e - bin * cols;
        Nl[e] = cc < ncol ? N[ij0 + cc + plane * (size_t)bin] : T(0);
    }
    __syncthreads();
    // symmetric_quadruplet.jl:92-173 for every (donor, column) pair: the
    // transfer of both orientations, gathered from the receivers' corners.
    for (int e = threadIdx.x; e < items; e += DIA_THREADS) {
        const int d = e / cols, cc = e - d * cols;
        const T Na = Nl[e];
        T TE[2] = {T(0), T(0)};
        if (Na > T(0)) {
            const T Ea = tb.ef_a[d] * Na;
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const int q = 2 * d + o;
                const int fl = tb.flags[q];
                const int *bp = tb.bins + 8 * (size_t)q;
                const T *wef = tb.wef + 8 * (size_t)q;
                T Ep = T(0), Em = T(0);
                if (fl & 1)
                    Ep = wef[0] * fmax(Nl[bp[0] * cols + cc], T(0)) + wef[1] * fmax(Nl[bp[1] * cols + cc], T(0)) +
                         wef[2] * fmax(Nl[bp[2] * cols + cc], T(0)) + wef[3] * fmax(Nl[bp[3] * cols + cc], T(0));
                if (fl & 2)
                    Em = wef[4] * fmax(Nl[bp[4] * cols + cc], T(0)) + wef[5] * fmax(Nl[bp[5] * cols + cc], T(0)) +
                         wef[6] * fmax(Nl[bp[6] * cols + cc], T(0)) + wef[7] * fmax(Nl[bp[7] * cols + cc], T(0));
                const T sa = Ea * (Ep * tb.dal1 + Em * tb.dal2) - Ep * Em * tb.dal3;
                TE[o] = fl ? tb.cf_a[d] * sa * T(0.5) : T(0);
            }
        }
        TEl[(2 * d) * cols + cc] = TE[0];
        TEl[(2 * d + 1) * cols + cc] = TE[1];
    }
    __syncthreads();
    // Each bin gathers what its donors send it (the scatter transposed on the
    // host; no atomics), then the split of symmetric_quadruplet.jl:210-222.
    // The action changes use tabulated w / E and 1 / E_a instead of the
    // reference's divisions (rounding-level difference).
    for (int e = threadIdx.x; e < items; e += DIA_THREADS) {
        const int bin = e / cols, cc = e - bin * cols;
        if (cc >= ncol)
            continue;
        T t = T(0);
        const int c1 = tb.gat_off[bin + 1];
        for (int c = tb.gat_off[bin]; c < c1; ++c)
            t = t + TEl[tb.gat_q[c] * cols + cc] * tb.gat_w[c];
        const T a = Nl[e];
        T src;
        if (t >= T(0)) {
            src = t;
        } else {
            const T rate = -t / fmax(a, eps_of<T>());
            src = -(fmin(rate, T(1)) * a);
        }
        out[ij0 + cc + plane * (size_t)bin] = src;
    }
}

// The same transfer with the columns as a short vector: a workgroup takes
// COLS adjacent columns (4 in FP32, 2 in FP64: 16 bytes per bin), and thread
// t owns bins and donors t, t + 512, ... (DPT of them) in every phase, for
// all COLS columns at once: one table load per donor instead of one per
// (donor, column), one 16-byte LDS read per corner.  The transfer of the
// thread's donors stays in registers until every corner read of the action
// is done, then overwrites the action in LDS ([2 nbins][COLS]), so the tile
// needs 2 nbins COLS elements of LDS instead of 3; the thread's own action
// stays in registers for the last phase.  Every value follows the operation
// order of dia_source_kernel, so both kernels give the same bits.
template <typename T, int COLS, int DPT>
__global__ __launch_bounds__(DIA_THREADS) void dia_source_vec_kernel(size_t plane, int nbins, long long tiles,
                                                             const T *__restrict__ N, DiaTables<T> tb,
                                                             T *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char dia_lds[];
    T *L = reinterpret_cast<T *>(dia_lds);
    const size_t ij0 = (size_t)dia_xcd_order(blockIdx.x, tiles) * COLS;
    const int ncol = (int)min((size_t)COLS, plane - ij0);
    const int tid = threadIdx.x;
    typedef T VecT __attribute__((ext_vector_type(COLS)));
    // whole 16-byte groups: every bin's COLS values are one aligned vector
    const bool whole = ncol == COLS && plane % COLS == 0;
    T own[DPT][COLS];
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int bin = tid + DIA_THREADS * k;
        if (bin < nbins) {
            if (whole) {
                const VecT v = *reinterpret_cast<const VecT *>(N + ij0 + plane * (size_t)bin);
#pragma unroll
                for (int c = 0; c < COLS; ++c)
                    own[k][c] = v[c];
            } else {
#pragma unroll
                for (int c = 0; c < COLS; ++c)
                    own[k][c] = c < ncol ? N[ij0 + c + plane * (size_t)bin] : T(0);
            }
            VecT v;
#pragma unroll
            for (int c = 0; c < COLS; ++c)
                v[c] = own[k][c];
            *reinterpret_cast<VecT *>(L + bin * COLS) = v;
        }
    }
    __syncthreads();
    T te[DPT][2][COLS];
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid + DIA_THREADS * k;
#pragma unroll
        for (int o = 0; o < 2; ++o)
#pragma unroll
            for (int c = 0; c < COLS; ++c)
                te[k][o][c] = T(0);
        if (d >= nbins)
            continue;
        bool any = false;
#pragma unroll
        for (int c = 0; c < COLS; ++c)
            any = any || own[k][c] > T(0);
        if (!any)
            continue;
        const T efa = tb.ef_a[d], cfa = tb.cf_a[d];
        typedef int Int4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const int q = 2 * d + o;
            const int fl = tb.flags[q];
            const Int4 *bq = reinterpret_cast<const Int4 *>(tb.bins + 8 * (size_t)q);
            const T *wef = tb.wef + 8 * (size_t)q;
            T Ep[COLS], Em[COLS];
#pragma unroll
            for (int c = 0; c < COLS; ++c)
                Ep[c] = Em[c] = T(0);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!(fl & (1 << h)))
                    continue;
                const Int4 b4 = bq[h];
                T acc[COLS];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const T w = wef[4 * h + j];
                    const VecT row = *reinterpret_cast<const VecT *>(L + b4[j] * COLS);
#pragma unroll
                    for (int c = 0; c < COLS; ++c) {
                        const T v = w * fmax(row[c], T(0));
                        acc[c] = j == 0 ? v : acc[c] + v;
                    }
                }
#pragma unroll
                for (int c = 0; c < COLS; ++c) {
                    if (h == 0)
                        Ep[c] = acc[c];
                    else
                        Em[c] = acc[c];
                }
            }
#pragma unroll
            for (int c = 0; c < COLS; ++c) {
                const T Ea = efa * own[k][c];
                const T sa = Ea * (Ep[c] * tb.dal1 + Em[c] * tb.dal2) - Ep[c] * Em[c] * tb.dal3;
                te[k][o][c] = (own[k][c] > T(0) && fl) ? cfa * sa * T(0.5) : T(0);
            }
        }
    }
    __syncthreads();  // every corner read of the action is done
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid + DIA_THREADS * k;
        if (d < nbins)
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                VecT v;
#pragma unroll
                for (int c = 0; c < COLS; ++c)
                    v[c] = te[k][o][c];
                *reinterpret_cast<VecT *>(L + (2 * d + o) * COLS) = v;
            }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int bin = tid + DIA_THREADS * k;
        if (bin >= nbins)
            continue;
        T t[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c)
            t[c] = T(0);
        const int c1 = tb.gat_off[bin + 1];
        // the list loads of four entries are issued before their sums
        int e = tb.gat_off[bin];
        for (; e + 4 <= c1; e += 4) {
            int gq[4];
            T gw[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                gq[u] = tb.gat_q[e + u];
                gw[u] = tb.gat_w[e + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const VecT row = *reinterpret_cast<const VecT *>(L + gq[u] * COLS);
#pragma unroll
                for (int c = 0; c < COLS; ++c)
                    t[c] = t[c] + row[c] * gw[u];
            }
        }
        for (; e < c1; ++e) {
            const T w = tb.gat_w[e];
            const VecT row = *reinterpret_cast<const VecT *>(L + tb.gat_q[e] * COLS);
#pragma unroll
            for (int c = 0; c < COLS; ++c)
                t[c] = t[c] + row[c] * w;
        }
        VecT sv;
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
            const T a = own[k][c];
            T src;
            if (t[c] >= T(0)) {
                src = t[c];
            } else {
                const T rate = -t[c] / fmax(a, eps_of<T>());
                src = -(fmin(rate, T(1)) * a);
            }
            sv[c] = src;
        }
        if (whole) {
            *reinterpret_cast<VecT *>(out + ij0 + plane * (size_t)bin) = sv;
        } else {
#pragma unroll
            for (int c = 0; c < COLS; ++c)
                if (c < ncol)
                    out[ij0 + c + plane * (size_t)bin] = sv[c];
        }
    }
}

// ListedQuadruplets (Sources/Sources.jl:1332-1385): one thread per cell
// walks the listed interactions in order and sums, as the reference,
// flux / w where the cell is a receiver and flux / (w N) (N > 0) where it is
// a donor, flux = rate max(N_d1, 0)^p max(N_d2, 0)^p; the source is
// positive - damping N.  The list is uniform across the grid (scalar loads);
// the donor values come from the thread's own column.  Each interaction's
// rate is a constant (rates[q]) or, with rate_field set, its column's value
// rate_field[col + plane q] (a rate given per column or as a function of
// position and time, evaluated by the host at the step's clock time).
template <typename T>
__global__ __launch_bounds__(256) void listed_source_kernel(size_t plane, int nk, int nphi, int count,
                                                            const int *__restrict__ bins,
                                                            const T *__restrict__ rates,
                                                            const T *__restrict__ rate_field, T power,
                                                            bool unit_power, const T *__restrict__ w,
                                                            const T *__restrict__ N, T *__restrict__ out)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int mn = blockIdx.y;
    if (ij >= plane)
        return;
    const int m = mn % nk, n = mn / nk;
    const size_t cell = ij + plane * (size_t)mn;
    const T weight = w[mn];
    const T a = N[cell];
    T positive = T(0), damping = T(0);
    for (int q = 0; q < count; ++q) {
        const int *b = bins + 8 * (size_t)q;
        T a1 = N[ij + plane * ((size_t)b[0] + (size_t)nk * b[1])];
        T a2 = N[ij + plane * ((size_t)b[2] + (size_t)nk * b[3])];
        a1 = a1 > T(0) ? a1 : T(0);
        a2 = a2 > T(0) ? a2 : T(0);
        // source_value (Sources.jl:1450-1467): a constant, or the column's value
        const T rate = rate_field ? rate_field[ij + plane * (size_t)q] : rates[q];
        const T flux = unit_power ? rate * a1 * a2 : rate * pow(a1, power) * pow(a2, power);
        if (flux == T(0))
            continue;
        if (m == b[4] && n == b[5])
            positive += flux / weight;
        if (m == b[6] && n == b[7])
            positive += flux / weight;
        if (m == b[0] && n == b[1] && a > T(0))
            damping += flux / (weight * a);
        if (m == b[2] && n == b[3] && a > T(0))
            damping += flux / (weight * a);
    }
    out[cell] = positive - damping * a;
}

}  // namespace

template <typename T>
int dia_columns_per_workgroup(int nbins)
{
    const size_t per_col = 3 * (size_t)nbins * sizeof(T);  // action and two transfers
    const size_t cols = (64 * 1024) / per_col;
    return (int)std::min<size_t>(cols, 8);
}

template <typename T, int COLS, int DPT>
hipError_t launch_dia_vec(size_t plane, int nbins, const T *N, const DiaTables<T> &tb, T *out, hipStream_t s)
{
    const long long tiles = (long long)((plane + COLS - 1) / COLS);
    if (tiles > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const size_t lds = 2 * (size_t)nbins * COLS * sizeof(T);
    hipLaunchKernelGGL((dia_source_vec_kernel<T, COLS, DPT>), dim3((unsigned)tiles), dim3(DIA_THREADS), lds, s, plane,
                       nbins, tiles, N, tb, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dia_source(int nx, int ny, int nk, int nphi, const T *N, const DiaTables<T> &tb, T *out,
                             hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    const int nbins = nk * nphi;
    // Grids up to 2048 bins: the column-vector kernel (16 bytes per bin and
    // column group, 2 nbins of them within 64 KiB of LDS).
    constexpr int VC = 16 / sizeof(T);
    if (nbins <= DIA_THREADS)
        return launch_dia_vec<T, VC, 1>(plane, nbins, N, tb, out, s);
    if (nbins <= 2 * DIA_THREADS)
        return launch_dia_vec<T, VC, 2>(plane, nbins, N, tb, out, s);
    if (nbins <= 3 * DIA_THREADS)
        return launch_dia_vec<T, VC, 3>(plane, nbins, N, tb, out, s);
    if (nbins <= 4 * DIA_THREADS)
        return launch_dia_vec<T, VC, 4>(plane, nbins, N, tb, out, s);
    const int cols = dia_columns_per_workgroup<T>(nbins);
    if (cols < 1)
        return hipErrorInvalidValue;
    const long long tiles = (long long)((plane + cols - 1) / cols);
    if (tiles > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const size_t lds = 3 * (size_t)nbins * cols * sizeof(T);
    hipLaunchKernelGGL(dia_source_kernel<T>, dim3((unsigned)tiles), dim3(DIA_THREADS), lds, s, plane, nbins, cols, tiles, N,
                       tb, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_listed_source(int nx, int ny, int nk, int nphi, int count, const int *bins, const T *rates,
                                const T *rate_field, T power, const T *w, const T *N, T *out, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    dim3 grid((unsigned)((plane + 255) / 256), (unsigned)(nk * nphi));
    hipLaunchKernelGGL(listed_source_kernel<T>, grid, dim3(256), 0, s, plane, nk, nphi, count, bins, rates,
                       rate_field, power, power == T(1), w, N, out);
    return hipGetLastError();
}

template hipError_t launch_listed_source<float>(int, int, int, int, int, const int *, const float *, const float *,
                                                float, const float *, const float *, float *, hipStream_t);
template hipError_t launch_listed_source<double>(int, int, int, int, int, const int *, const double *, const double *,
                                                 double, const double *, const double *, double *, hipStream_t);
template int dia_columns_per_workgroup<float>(int);
template int dia_columns_per_workgroup<double>(int);
template hipError_t launch_dia_source<float>(int, int, int, int, const float *, const DiaTables<float> &, float *,
                                             hipStream_t);
template hipError_t launch_dia_source<double>(int, int, int, int, const double *, const DiaTables<double> &,
                                              double *, hipStream_t);

}  // namespace rpl
