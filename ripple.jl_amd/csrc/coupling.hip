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

// The same contraction on the matrix cores: U[ij, m] = sum_k u[ij, k] W[k, m]
// is an [nx*ny x nz] x [nz x nk] product (M = nx*ny rows, K = nz, N = nk).
// One wavefront owns a 16 x 16 tile of U (16 columns ij, 16 bins m) for both
// u and v, and walks K in steps of 4 with v_mfma_*_16x16x4 (FP32 inputs with
// FP32 accumulation, or FP64).  Operand layout of the 16x16x4 instructions
// (measured on gfx950): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// accumulator register r of lane l is D[4 (l / 16) + r][l % 16] for the FP32
// instruction and D[(l / 16) + 4 r][l % 16] for the FP64 one.  Out-of-range
// rows, columns and k contribute zeros; stores are guarded.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16x16x4(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16x16x4(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T, typename V>
__global__ __launch_bounds__(256) void doppler_mfma_kernel(int nx, int ny, int nz, int nk, const T *__restrict__ u,
                                                           const T *__restrict__ v, const T *__restrict__ W, T *Ux,
                                                           T *Uy)
{
    const size_t M = (size_t)nx * ny;
    const int lane = threadIdx.x & 63;
    const size_t tile_row = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int tile_col = blockIdx.y * 16;
    if (tile_row >= M)
        return;  // whole wavefront leaves together
    const int r16 = lane & 15, q = lane >> 4;
    const size_t arow = tile_row + r16;
    const int bcol = tile_col + r16;
    V accu = {0, 0, 0, 0}, accv = {0, 0, 0, 0};
    for (int k0 = 0; k0 < nz; k0 += 4) {
        const int k = k0 + q;
        const bool ka = k < nz;
        const T au = (ka && arow < M) ? u[arow + M * k] : T(0);
        const T av = (ka && arow < M) ? v[arow + M * k] : T(0);
        const T b = (ka && bcol < nk) ? W[k + (size_t)nz * bcol] : T(0);
        accu = mfma16x16x4(au, b, accu);
        accv = mfma16x16x4(av, b, accv);
    }
    const int col = tile_col + r16;
    if (col < nk) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const size_t row = sizeof(T) == 4 ? tile_row + 4 * q + r : tile_row + q + 4 * r;
            if (row < M) {
                Ux[row + M * col] = accu[r];
                Uy[row + M * col] = accv[r];
            }
        }
    }
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
    return launch_doppler_mfma(nx, ny, nz, nk, u, v, W, Ux, Uy, s);
}

template <typename T>
hipError_t launch_doppler_mfma(int nx, int ny, int nz, int nk, const T *u, const T *v, const T *W, T *Ux, T *Uy,
                               hipStream_t s)
{
    const size_t M = (size_t)nx * ny;
    const size_t tiles = (M + 15) / 16;
    dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)((nk + 15) / 16));
    if (sizeof(T) == 4)
        hipLaunchKernelGGL((doppler_mfma_kernel<float, f32x4>), grid, dim3(256), 0, s, nx, ny, nz, nk,
                           (const float *)u, (const float *)v, (const float *)W, (float *)Ux, (float *)Uy);
    else
        hipLaunchKernelGGL((doppler_mfma_kernel<double, f64x4>), grid, dim3(256), 0, s, nx, ny, nz, nk,
                           (const double *)u, (const double *)v, (const double *)W, (double *)Ux, (double *)Uy);
    return hipGetLastError();
}

template hipError_t launch_doppler_mfma<float>(int, int, int, int, const float *, const float *, const float *, float *,
                                               float *, hipStream_t);
template hipError_t launch_doppler_mfma<double>(int, int, int, int, const double *, const double *, const double *,
                                                double *, double *, hipStream_t);

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
