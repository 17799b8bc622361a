// Coupling precompute: the Q-transform Doppler contraction and the current
// gradients.  Both run once per set_current (the current is prescribed and
// static, Coupling/current_coupling.jl:127-138).
#include "rpl_internal.h"

#pragma clang fp contract(off)

namespace rpl {
namespace {

// doppler_velocity.jl:45-62 in its accumulation order: U[i,j,m] = sum over k
// (ascending) of u[i,j,k] * W[k,m].  Thread per (i, j, m); consecutive
// threads take consecutive i, so u[., ., k] loads are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void doppler_sequential_kernel(int nx, int ny, int nz, int nk,
                                                                 const T *__restrict__ u, const T *__restrict__ v,
                                                                 const T *__restrict__ W, T *Ux, T *Uy)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int m = blockIdx.y;
    if (ij >= plane)
        return;
    T ax = T(0), ay = T(0);
    for (int k = 0; k < nz; ++k) {
        const T q = W[k + (size_t)nz * m];
        ax += u[ij + plane * k] * q;
        ay += v[ij + plane * k] * q;
    }
    Ux[ij + plane * m] = ax;
    Uy[ij + plane * m] = ay;
}

// wave_current_refraction.jl:7-19.
template <typename T>
__global__ __launch_bounds__(256) void gradients_kernel(int nx, int ny, int nk, const T *__restrict__ Ux,
                                                        const T *__restrict__ Uy, T hx, T hy, T *Uxx, T *Uxy,
                                                        T *Uyx, T *Uyy)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int m = blockIdx.y;
    if (ij >= plane)
        return;
    const int i = (int)(ij % nx), j = (int)(ij / nx);
    const int ip = i == nx - 1 ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
    const int jp = j == ny - 1 ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
    const size_t b = plane * m;
    const size_t c = b + ij;
    Uxx[c] = (Ux[b + ip + (size_t)nx * j] - Ux[b + im + (size_t)nx * j]) * hx;
    Uxy[c] = (Ux[b + i + (size_t)nx * jp] - Ux[b + i + (size_t)nx * jm]) * hy;
    Uyx[c] = (Uy[b + ip + (size_t)nx * j] - Uy[b + im + (size_t)nx * j]) * hx;
    Uyy[c] = (Uy[b + i + (size_t)nx * jp] - Uy[b + i + (size_t)nx * jm]) * hy;
}

template <typename T>
__global__ __launch_bounds__(256) void cast_from_f64_kernel(const double *__restrict__ src, T *dst, size_t n)
{
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n)
        dst[k] = (T)src[k];
}

template <typename T>
__global__ __launch_bounds__(256) void cast_to_f64_kernel(const T *__restrict__ src, double *dst, size_t n)
{
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n)
        dst[k] = (double)src[k];
}

}  // namespace

hipError_t launch_doppler(int nx, int ny, int nz, int nk, const double *u, const double *v, const double *W,
                          double *Ux, double *Uy, hipStream_t s)
{
    size_t plane = (size_t)nx * ny;
    dim3 grid((unsigned)((plane + 255) / 256), nk);
    hipLaunchKernelGGL(doppler_sequential_kernel<double>, grid, dim3(256), 0, s, nx, ny, nz, nk, u, v, W, Ux, Uy);
    return hipGetLastError();
}

hipError_t launch_doppler(int nx, int ny, int nz, int nk, const float *u, const float *v, const float *W,
                          float *Ux, float *Uy, hipStream_t s)
{
    size_t plane = (size_t)nx * ny;
    dim3 grid((unsigned)((plane + 255) / 256), nk);
    hipLaunchKernelGGL(doppler_sequential_kernel<float>, grid, dim3(256), 0, s, nx, ny, nz, nk, u, v, W, Ux, Uy);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_gradients(int nx, int ny, int nk, const T *Ux, const T *Uy, T dx_inv, T dy_inv, T *Uxx, T *Uxy,
                            T *Uyx, T *Uyy, hipStream_t s)
{
    size_t plane = (size_t)nx * ny;
    dim3 grid((unsigned)((plane + 255) / 256), nk);
    hipLaunchKernelGGL(gradients_kernel<T>, grid, dim3(256), 0, s, nx, ny, nk, Ux, Uy, dx_inv / T(2), dy_inv / T(2),
                       Uxx, Uxy, Uyx, Uyy);
    return hipGetLastError();
}

template hipError_t launch_gradients<double>(int, int, int, const double *, const double *, double, double, double *,
                                             double *, double *, double *, hipStream_t);
template hipError_t launch_gradients<float>(int, int, int, const float *, const float *, float, float, float *,
                                            float *, float *, float *, hipStream_t);

template <typename T>
hipError_t launch_cast_from_f64(const double *src, T *dst, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(cast_from_f64_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cast_to_f64(const T *src, double *dst, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(cast_to_f64_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

template hipError_t launch_cast_from_f64<float>(const double *, float *, size_t, hipStream_t);
template hipError_t launch_cast_from_f64<double>(const double *, double *, size_t, hipStream_t);
template hipError_t launch_cast_to_f64<float>(const float *, double *, size_t, hipStream_t);
template hipError_t launch_cast_to_f64<double>(const double *, double *, size_t, hipStream_t);

}  // namespace rpl
