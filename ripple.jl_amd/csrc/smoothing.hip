// Garden-sprinkler alleviation by spatial averaging (Tolman 2002), the
// reference's SpatialAveraging (Models/propagation_smoothing.jl:22-206):
// after every time step each (kappa, phi) bin's x-y field is replaced by the
// mean of four bilinear interpolations at the corners of a rectangle along
// and across the bin's propagation direction.  The displacement is the same
// for every cell of a bin, so a workgroup of 256 consecutive cells of one
// bin reads a shifted copy of its own rows: coalesced, and the four corner
// stencils hit the same L2 lines.
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

// propagation_smoothing.jl:56-110 on a periodic x-y grid, in the
// reference's operation order; indices are formed 1-based as there, so the
// fractional weights round the same way.
template <typename T>
__global__ __launch_bounds__(256) void spatial_averaging_kernel(int nx, int ny, int nk, const T *__restrict__ Nin,
                                                                T *__restrict__ Nout, const T *__restrict__ cg,
                                                                const T *__restrict__ dcg, const T *__restrict__ cosp,
                                                                const T *__restrict__ sinp, T dphi, T as, T an, T dt,
                                                                T dx_inv, T dy_inv)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (ij >= plane)
        return;
    const int mn = blockIdx.y;
    const int m = mn % nk, n = mn / nk;
    const int i = (int)(ij % (size_t)nx), j = (int)(ij / (size_t)nx);
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
            const int i0c = gse_wrap(i0, nx) - 1, i1c = gse_wrap(i0 + 1, nx) - 1;
            const int j0c = gse_wrap(j0, ny) - 1, j1c = gse_wrap(j0 + 1, ny) - 1;
            const T f00 = slab[i0c + (size_t)nx * j0c], f10 = slab[i1c + (size_t)nx * j0c];
            const T f01 = slab[i0c + (size_t)nx * j1c], f11 = slab[i1c + (size_t)nx * j1c];
            const T one = T(1);
            acc += (one - wi) * (one - wj) * f00 + wi * (one - wj) * f10 + (one - wi) * wj * f01 + wi * wj * f11;
        }
    Nout[plane * (size_t)mn + ij] = acc / T(4);
}

}  // namespace

template <typename T>
hipError_t launch_spatial_averaging(int nx, int ny, int nk, int nphi, const T *Nin, T *Nout, const T *cg,
                                    const T *dcg, const T *cosp, const T *sinp, T dphi, T as, T an, T dt, T dx_inv,
                                    T dy_inv, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    dim3 grid((unsigned)((plane + 255) / 256), (unsigned)(nk * nphi));
    hipLaunchKernelGGL(spatial_averaging_kernel<T>, grid, dim3(256), 0, s, nx, ny, nk, Nin, Nout, cg, dcg, cosp,
                       sinp, dphi, as, an, dt, dx_inv, dy_inv);
    return hipGetLastError();
}

template hipError_t launch_spatial_averaging<float>(int, int, int, int, const float *, float *, const float *,
                                                    const float *, const float *, const float *, float, float, float,
                                                    float, float, float, hipStream_t);
template hipError_t launch_spatial_averaging<double>(int, int, int, int, const double *, double *, const double *,
                                                     const double *, const double *, const double *, double, double,
                                                     double, double, double, double, hipStream_t);

}  // namespace rpl
