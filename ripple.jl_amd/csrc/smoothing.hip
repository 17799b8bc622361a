// Garden-sprinkler alleviation by spatial averaging (Tolman 2002), the
// reference's SpatialAveraging (Models/propagation_smoothing.jl:22-206),
// periodic or clamped at the edges of bounded axes:
// after every time step each (kappa, phi) bin's x-y field is replaced by the
// mean of four bilinear interpolations at the corners of a rectangle along
// and across the bin's propagation direction.  The displacement is the same
// for every cell of a bin, so a workgroup (a 64 x 4 tile of one bin) reads a
// shifted copy of its own rows, coalesced, and the corner stencils of its
// rows overlap.  The tiles of one bin are numbered consecutively (x
// fastest) and each XCD takes one contiguous eighth of the list, so the
// tiles that share rows run together on one XCD and meet in its L2.
#include "rpl_internal.h"

#pragma clang fp contract(off)

namespace rpl {
namespace {

__device__ __forceinline__ double floor_t(double x) { return floor(x); }
__device__ __forceinline__ float floor_t(float x) { return floorf(x); }

// 1-based wrap onto 1..n: one period either way as the reference's
// _gse_periodic (propagation_smoothing.jl:53), further periods for
// displacements beyond the domain.
__device__ __forceinline__ int gse_wrap(int i, int n)
{
    i = i < 1 ? i + n : (i > n ? i - n : i);
    if (i < 1 || i > n)
        i = ((i - 1) % n + n) % n + 1;
    return i;
}

// The reference's _gse_clamp (propagation_smoothing.jl:56): a 1-based index
// held to 1..n on a Bounded axis.
__device__ __forceinline__ int gse_clamp(int i, int n) { return i < 1 ? 1 : (i > n ? n : i); }

// Workgroup b runs on XCD b % 8: XCD x gets the x-th contiguous share of
// [0, total), the first total % 8 shares one longer.
__device__ __forceinline__ long long xcd_order(long long b, long long total)
{
    const long long xcd = b & 7, r = b >> 3, per = total >> 3, rem = total & 7;
    return xcd < rem ? xcd * (per + 1) + r : rem * (per + 1) + (xcd - rem) * per + r;
}

// propagation_smoothing.jl:56-110 in the reference's operation order;
// indices are formed 1-based as there, so the fractional weights round the
// same way.  A corner index wraps on a periodic axis (px, py) and is clamped
// to the edge cell on a bounded one (propagation_smoothing.jl:97-100).
template <typename T>
__global__ __launch_bounds__(256) void spatial_averaging_kernel(int nx, int ny, int nk, int nbx, int nby,
                                                                long long total, const T *__restrict__ Nin,
                                                                T *__restrict__ Nout, const T *__restrict__ cg,
                                                                const T *__restrict__ dcg, const T *__restrict__ cosp,
                                                                const T *__restrict__ sinp, T dphi, T as, T an, T dt,
                                                                T dx_inv, T dy_inv, int px, int py)
{
    const long long t = xcd_order(blockIdx.x, total);
    const int tiles = nbx * nby;
    const int mn = (int)(t / tiles), r = (int)(t - (long long)mn * tiles);
    const int tj = r / nbx, ti = r - tj * nbx;
    const int i = ti * 64 + (int)threadIdx.x, j = tj * 4 + (int)threadIdx.y;
    if (i >= nx || j >= ny)
        return;
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)i + (size_t)nx * j;
    const int m = mn % nk, n = mn / nk;
    const T Ls = as * dcg[m] * dt;
    const T Ln = an * cg[m] * dphi * dt;
    const T c = cosp[n], s = sinp[n];
    const T Ls_x = Ls * c * dx_inv, Ls_y = Ls * s * dy_inv;
    const T Ln_x = -Ln * s * dx_inv, Ln_y = Ln * c * dy_inv;
    const T *slab = Nin + plane * (size_t)mn;
    T acc = T(0);
#pragma unroll
    for (int s1 = -1; s1 <= 1; s1 += 2)
#pragma unroll
        for (int s2 = -1; s2 <= 1; s2 += 2) {
            const T di = (T)s1 * Ls_x + (T)s2 * Ln_x;
            const T dj = (T)s1 * Ls_y + (T)s2 * Ln_y;
            const T ir = (T)(i + 1) + di, jr = (T)(j + 1) + dj;
            const int i0 = (int)floor_t(ir), j0 = (int)floor_t(jr);
            const T wi = ir - (T)i0, wj = jr - (T)j0;
            const int i0c = (px ? gse_wrap(i0, nx) : gse_clamp(i0, nx)) - 1;
            const int i1c = (px ? gse_wrap(i0 + 1, nx) : gse_clamp(i0 + 1, nx)) - 1;
            const int j0c = (py ? gse_wrap(j0, ny) : gse_clamp(j0, ny)) - 1;
            const int j1c = (py ? gse_wrap(j0 + 1, ny) : gse_clamp(j0 + 1, ny)) - 1;
            const T f00 = slab[i0c + (size_t)nx * j0c], f10 = slab[i1c + (size_t)nx * j0c];
            const T f01 = slab[i0c + (size_t)nx * j1c], f11 = slab[i1c + (size_t)nx * j1c];
            const T one = T(1);
            acc += (one - wi) * (one - wj) * f00 + wi * (one - wj) * f10 + (one - wi) * wj * f01 + wi * wj * f11;
        }
    Nout[plane * (size_t)mn + ij] = acc / T(4);
}

// 0-based periodic index for any integer.
__device__ __forceinline__ int wrap0(int i, int n)
{
    i = i < 0 ? i + n : (i >= n ? i - n : i);
    if (i < 0 || i >= n)
        i = (i % n + n) % n;
    return i;
}

// 0-based index held to 0..n-1 (bounded axis).
__device__ __forceinline__ int clamp0(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// FP32 production form.  In exact arithmetic the corner displacement of a
// cell splits into the bin's integer offset floor(d) and the fraction
// d - floor(d), the same for every cell of the bin, so the corner weights are
// formed once per thread.  Each corner value is interpolated along x within
// two rows and then between the rows; a thread walks ROWS consecutive rows,
// so the lower row of one cell's corner is the upper row of the next cell's:
// two loads per corner and cell.  (The reference forms i + d per cell and
// weights the four points directly; the results differ by rounding.)  A
// workgroup takes a 64 x 4 ROWS tile, each thread one column.  On a bounded
// axis (px or py zero) the corner indices are clamped to the edge cells
// instead of wrapped, as in the reference form.
template <int ROWS>
__global__ __launch_bounds__(256) void spatial_averaging_fast_kernel(int nx, int ny, int nk, int nbx, int nby,
                                                                     long long total, const float *__restrict__ Nin,
                                                                     float *__restrict__ Nout,
                                                                     const float *__restrict__ cg,
                                                                     const float *__restrict__ dcg,
                                                                     const float *__restrict__ cosp,
                                                                     const float *__restrict__ sinp, float dphi,
                                                                     float as, float an, float dt, float dx_inv,
                                                                     float dy_inv, int px, int py)
{
    const long long t = xcd_order(blockIdx.x, total);
    const int tiles = nbx * nby;
    const int mn = (int)(t / tiles), r = (int)(t - (long long)mn * tiles);
    const int tj = r / nbx, ti = r - tj * nbx;
    const int i = ti * 64 + (int)threadIdx.x;
    const int j0 = tj * 4 * ROWS + (int)threadIdx.y * ROWS;
    if (i >= nx || j0 >= ny)
        return;
    const int m = mn % nk, n = mn / nk;
    const float Ls = as * dcg[m] * dt;
    const float Ln = an * cg[m] * dphi * dt;
    const float c = cosp[n], s = sinp[n];
    const float Ls_x = Ls * c * dx_inv, Ls_y = Ls * s * dy_inv;
    const float Ln_x = -Ln * s * dx_inv, Ln_y = Ln * c * dy_inv;
    int ia[4], ib[4], jo[4];
    float wx[4], wy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float s1 = (q & 2) ? 1.0f : -1.0f, s2 = (q & 1) ? 1.0f : -1.0f;
        const float di = s1 * Ls_x + s2 * Ln_x, dj = s1 * Ls_y + s2 * Ln_y;
        const float fi = floorf(di), fj = floorf(dj);
        wx[q] = di - fi;
        wy[q] = dj - fj;
        ia[q] = px ? wrap0(i + (int)fi, nx) : clamp0(i + (int)fi, nx);
        ib[q] = px ? wrap0(i + (int)fi + 1, nx) : clamp0(i + (int)fi + 1, nx);
        jo[q] = (int)fj;
    }
    const size_t plane = (size_t)nx * ny;
    const float *slab = Nin + plane * (size_t)mn;
    float *dst = Nout + plane * (size_t)mn;
    // Element offset of the next row each corner reads (rows wrap at ny, or
    // stop at the edge rows on a bounded y axis; nx * ny < 2^30, so 32 bits
    // hold it), and h[q], corner q's x-interpolation in row j + jo[q] of the
    // current cell.
    const unsigned nxu = (unsigned)nx, planeu = (unsigned)plane;
    const int jend = min(j0 + ROWS, ny);
    float h[4];
    if (py) {
        unsigned ro[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ro[q] = nxu * (unsigned)wrap0(j0 + jo[q], ny);
            const float *row = slab + ro[q];
            h[q] = row[ia[q]] + wx[q] * (row[ib[q]] - row[ia[q]]);
            ro[q] += nxu;
            ro[q] = ro[q] == planeu ? 0u : ro[q];
        }
        for (int j = j0; j < jend; ++j) {
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float *row = slab + ro[q];
                const float hn = row[ia[q]] + wx[q] * (row[ib[q]] - row[ia[q]]);
                acc += h[q] + wy[q] * (hn - h[q]);
                h[q] = hn;
                ro[q] += nxu;
                ro[q] = ro[q] == planeu ? 0u : ro[q];
            }
            dst[(size_t)i + (size_t)nx * j] = acc * 0.25f;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float *row = slab + nxu * (unsigned)clamp0(j0 + jo[q], ny);
            h[q] = row[ia[q]] + wx[q] * (row[ib[q]] - row[ia[q]]);
        }
        for (int j = j0; j < jend; ++j) {
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float *row = slab + nxu * (unsigned)clamp0(j + 1 + jo[q], ny);
                const float hn = row[ia[q]] + wx[q] * (row[ib[q]] - row[ia[q]]);
                acc += h[q] + wy[q] * (hn - h[q]);
                h[q] = hn;
            }
            dst[(size_t)i + (size_t)nx * j] = acc * 0.25f;
        }
    }
}

}  // namespace

template <typename T>
hipError_t launch_spatial_averaging(bool fast, int nx, int ny, int nk, int nphi, const T *Nin, T *Nout, const T *cg,
                                    const T *dcg, const T *cosp, const T *sinp, T dphi, T as, T an, T dt, T dx_inv,
                                    T dy_inv, bool px, bool py, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (fast) {
            constexpr int ROWS = 32;
            const int nbx = (nx + 63) / 64, nby = (ny + 4 * ROWS - 1) / (4 * ROWS);
            const long long total = (long long)nbx * nby * nk * nphi;
            if (total > 0x7fffffffLL)
                return hipErrorInvalidValue;
            hipLaunchKernelGGL(spatial_averaging_fast_kernel<ROWS>, dim3((unsigned)total), dim3(64, 4), 0, s, nx, ny,
                               nk, nbx, nby, total, Nin, Nout, cg, dcg, cosp, sinp, dphi, as, an, dt, dx_inv, dy_inv,
                               (int)px, (int)py);
            return hipGetLastError();
        }
    }
    const int nbx = (nx + 63) / 64, nby = (ny + 3) / 4;
    const long long total = (long long)nbx * nby * nk * nphi;
    if (total > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(spatial_averaging_kernel<T>, dim3((unsigned)total), dim3(64, 4), 0, s, nx, ny, nk, nbx, nby,
                       total, Nin, Nout, cg, dcg, cosp, sinp, dphi, as, an, dt, dx_inv, dy_inv, (int)px, (int)py);
    return hipGetLastError();
}

template hipError_t launch_spatial_averaging<float>(bool, int, int, int, int, const float *, float *, const float *,
                                                    const float *, const float *, const float *, float, float, float,
                                                    float, float, float, bool, bool, hipStream_t);
template hipError_t launch_spatial_averaging<double>(bool, int, int, int, int, const double *, double *, const double *,
                                                     const double *, const double *, const double *, double, double,
                                                     double, double, double, double, bool, bool, hipStream_t);

}  // namespace rpl
