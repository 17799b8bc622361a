// Coupling: the Q-transform Doppler contraction and the current gradients
// (once per set_current: the prescribed current is static,
// Coupling/current_coupling.jl:127-138), and the pseudomomentum contractions
// of the action (before every tendency evaluation for the self-coupled
// current, current_coupling.jl:140-143, and for the pseudomomentum fields).
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

// q_kernel.jl:60-92 (q_cdf_kernel) in FP64 on the device: the CDF of the
// vertical projection kernel, sinh(2 mu s) / sinh(2 mu) written with
// exponentials that cannot overflow, s itself below |mu| = 1e-6; and
// q_kernel.jl:35-53, its kappa derivative.  The host forms the same
// expressions for the tables of a constant depth.
__device__ __forceinline__ double q_cdf_dev(double kappa, double z, double depth)
{
    const double small_mu = 1e-6;
    const double mu = kappa * depth;
    const double mus = fabs(mu) < small_mu ? (mu < 0 ? -small_mu : small_mu) : mu;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    const double a = 2 * mus * s, b = 2 * mus;
    const double fin = (exp(a - b) - exp(-a - b)) / (1 - exp(-2 * b));
    return fabs(mu) < small_mu ? s : fin;
}

__device__ __forceinline__ double q_cdf_dkappa_dev(double kappa, double z, double depth)
{
    const double mu = kappa * depth;
    double s = (z + depth) / depth;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    if (fabs(mu) < 1e-6)
        return depth * (4 * mu * s * (s * s - 1) / 3);
    const double a = exp(-2 * mu * (1 - s)), b = exp(-2 * mu * (1 + s)), e = exp(-4 * mu);
    const double num = a - b, den = 1 - e;
    const double dnum = -2 * (1 - s) * a + 2 * (1 + s) * b, dden = 4 * e;
    const double dF = (dnum * den - num * dden) / (den * den);
    return depth * dF;
}

// doppler_velocity.jl:45-62 with a depth per column (q_depth_at(depth, i,
// j), q_transform.jl:124-125): thread per (i, j, m) forms its column's
// weights W[k, m] = CDF(z_high) - CDF(z_low) (or their kappa derivative,
// doppler_velocity.jl:64-82) at depth[i + nx j] and accumulates over k in
// ascending order, as the reference.
template <typename T>
__global__ __launch_bounds__(256) void doppler_columns_kernel(size_t plane, int nz, const T *__restrict__ u,
                                                              const T *__restrict__ v,
                                                              const double *__restrict__ zf,
                                                              const double *__restrict__ kc,
                                                              const double *__restrict__ depth, bool derivative,
                                                              T *Ux, T *Uy)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int m = blockIdx.y;
    if (ij >= plane)
        return;
    const double d = depth[ij], k = kc[m];
    T ax = T(0), ay = T(0);
    for (int q = 0; q < nz; ++q) {
        const double z1 = zf[q], z2 = zf[q + 1];
        const double hi = z1 < z2 ? z2 : z1, lo = z1 < z2 ? z1 : z2;
        const double w = derivative ? q_cdf_dkappa_dev(k, hi, d) - q_cdf_dkappa_dev(k, lo, d)
                                    : q_cdf_dev(k, hi, d) - q_cdf_dev(k, lo, d);
        const T wt = (T)w;
        ax += u[ij + plane * q] * wt;
        ay += v[ij + plane * q] * wt;
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

// Pseudomomentum contractions of the action column (i, j):
//   X[i,j,o] = scale_o * sum_{n, s} A[o,s] N[i,j,s,n] kx[s,n]     (Y with ky)
// and, with a second matrix B, dX, dY the same with B in place of A.  A and B
// are (nout x nk), o fastest.  With A the overlap O[t,s] (and B its kappa
// derivative) this is the self-coupled Doppler velocity
// (pseudomomentum.jl:316-344); with A[k,m] the Q cell weights W[k,m] and
// scale_k = 1/|dz_k| (or none) it is the pseudomomentum cell average (or
// integral) (pseudomomentum.jl:375-400).
//
// Reference order: per output, n outer, s inner, each term (A * N) * k, a
// final product by scale_o when given.  One thread per (i, j) and a chunk of
// PM_CHUNK outputs; consecutive threads take consecutive (i, j), so every
// N[., ., s, n] load is coalesced and A, B, kx, ky are wave-uniform.
constexpr int PM_CHUNK = 8;
constexpr int PM_OUT = 8;

template <typename T>
__global__ __launch_bounds__(256) void pm_ordered_kernel(size_t plane, int nk, int nphi, int nout,
                                                         const T *__restrict__ N, const T *__restrict__ A,
                                                         const T *__restrict__ B, const T *__restrict__ kx,
                                                         const T *__restrict__ ky, const T *__restrict__ scale, T *X,
                                                         T *Y, T *dX, T *dY, bool per_column)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int o0 = blockIdx.y * PM_CHUNK;
    if (ij >= plane)
        return;
    T ax[PM_CHUNK], ay[PM_CHUNK], bx[PM_CHUNK], by[PM_CHUNK];
#pragma unroll
    for (int c = 0; c < PM_CHUNK; ++c)
        ax[c] = ay[c] = bx[c] = by[c] = T(0);
    for (int n = 0; n < nphi; ++n)
        for (int s = 0; s < nk; ++s) {
            const size_t sn = (size_t)s + (size_t)nk * n;
            const T a = N[ij + plane * sn];
            const T kxv = kx[sn], kyv = ky[sn];
#pragma unroll
            for (int c = 0; c < PM_CHUNK; ++c) {
                const int o = o0 + c;
                if (o < nout) {
                    // one matrix, or the column's own (a depth per column,
                    // pseudomomentum.jl:72-87): A[ij + plane (o + nout s)]
                    const size_t os = (size_t)o + (size_t)nout * s;
                    const size_t ai = per_column ? ij + plane * os : os;
                    const T q = A[ai];
                    ax[c] += q * a * kxv;
                    ay[c] += q * a * kyv;
                    if (B) {
                        const T dq = B[ai];
                        bx[c] += dq * a * kxv;
                        by[c] += dq * a * kyv;
                    }
                }
            }
        }
#pragma unroll
    for (int c = 0; c < PM_CHUNK; ++c) {
        const int o = o0 + c;
        if (o < nout) {
            const size_t out = ij + plane * o;
            X[out] = scale ? ax[c] * scale[o] : ax[c];
            Y[out] = scale ? ay[c] * scale[o] : ay[c];
            if (B) {
                dX[out] = bx[c];
                dY[out] = by[c];
            }
        }
    }
}

// The same contractions factored through the column's directional moments,
// M_x[s] = sum_n N[s,n] kx[s,n] (and M_y), then X[o] = scale_o sum_s A[o,s]
// M_x[s]: N is read once and the work per cell-bin is two products, so the
// launch runs at the HBM rate of reading N.  A workgroup holds 64 columns;
// its four wavefronts split the source bins s for the moments, which meet in
// LDS (2 nk x 64 values) next to the matrices, and then split the outputs o.
// The sums are regrouped, so the bits differ from the reference order by
// rounding only.
template <typename T>
__global__ __launch_bounds__(256) void pm_moment_kernel(size_t plane, int nk, int nphi, int nout,
                                                        const T *__restrict__ N, const T *__restrict__ A,
                                                        const T *__restrict__ B,
                                                        const T *__restrict__ kx, const T *__restrict__ ky,
                                                        const T *__restrict__ scale, T *X, T *Y, T *dX, T *dY)
{
    extern __shared__ unsigned char pm_lds[];
    T *Mx = reinterpret_cast<T *>(pm_lds);
    T *My = Mx + (size_t)nk * 64;
    // The matrices, A then B, [s][o] as in global memory.
    T *Ash = My + (size_t)nk * 64;
    T *Bsh = Ash + (size_t)nk * nout;
    // The wavefront index, held in a scalar register so that the source bins
    // and outputs each wavefront walks are scalar (uniform) values.
    const int col = threadIdx.x & 63, grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t ij = (size_t)blockIdx.x * 64 + col;
    const bool in = ij < plane;
    const size_t sstride = plane * nk;
    for (int e = threadIdx.x; e < nk * nout; e += 256) {
        Ash[e] = A[e];
        if (B)
            Bsh[e] = B[e];
    }
    for (int s = grp; s < nk; s += 4) {
        T mx = T(0), my = T(0);
        if (in) {
            const T *Ns = N + ij + plane * s;
#pragma unroll 6
            for (int n = 0; n < nphi; ++n) {
                const T a = Ns[sstride * n];
                const size_t sn = (size_t)s + (size_t)nk * n;
                mx += a * kx[sn];
                my += a * ky[sn];
            }
        }
        Mx[s * 64 + col] = mx;
        My[s * 64 + col] = my;
    }
    __syncthreads();
    if (!in)
        return;
    // Each wavefront takes PM_OUT consecutive outputs at a time; the matrix
    // entries it reads are the same for all its lanes (LDS broadcasts).
    for (int o0 = grp * PM_OUT; o0 < nout; o0 += 4 * PM_OUT) {
        T x[PM_OUT], y[PM_OUT], dx[PM_OUT], dy[PM_OUT];
#pragma unroll
        for (int c = 0; c < PM_OUT; ++c)
            x[c] = y[c] = dx[c] = dy[c] = T(0);
        for (int s = 0; s < nk; ++s) {
            const T mx = Mx[s * 64 + col], my = My[s * 64 + col];
            const T *As = Ash + (size_t)nout * s + o0, *Bs = Bsh + (size_t)nout * s + o0;
#pragma unroll
            for (int c = 0; c < PM_OUT; ++c) {
                if (o0 + c < nout) {  // uniform
                    const T q = As[c];  // one LDS address for the whole wavefront
                    x[c] += q * mx;
                    y[c] += q * my;
                    if (B) {
                        const T dq = Bs[c];
                        dx[c] += dq * mx;
                        dy[c] += dq * my;
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < PM_OUT; ++c) {
            const int o = o0 + c;
            if (o < nout) {
                const size_t out = ij + plane * o;
                X[out] = scale ? x[c] * scale[o] : x[c];
                Y[out] = scale ? y[c] * scale[o] : y[c];
                if (B) {
                    dX[out] = dx[c];
                    dY[out] = dy[c];
                }
            }
        }
    }
}

// pseudomomentum.jl:412: the action tendency (new - old) / dt per cell.
template <typename T>
__global__ __launch_bounds__(256) void action_tendency_kernel(size_t n, const T *__restrict__ Nnew,
                                                              const T *__restrict__ Nold, T dt, T *out)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n)
        out[k] = (Nnew[k] - Nold[k]) / dt;
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

template <typename T>
hipError_t launch_doppler_columns(int nx, int ny, int nz, int nk, const T *u, const T *v, const double *zf,
                                  const double *kc, const double *depth, bool derivative, T *Ux, T *Uy, hipStream_t s)
{
    const size_t plane = (size_t)nx * ny;
    if (nk > 65535 || (plane + 255) / 256 > 0x7fffffffULL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(doppler_columns_kernel<T>, dim3((unsigned)((plane + 255) / 256), (unsigned)nk), dim3(256), 0,
                       s, plane, nz, u, v, zf, kc, depth, derivative, Ux, Uy);
    return hipGetLastError();
}

template hipError_t launch_doppler_columns<double>(int, int, int, int, const double *, const double *,
                                                   const double *, const double *, const double *, bool, double *,
                                                   double *, hipStream_t);
template hipError_t launch_doppler_columns<float>(int, int, int, int, const float *, const float *, const double *,
                                                  const double *, const double *, bool, float *, float *,
                                                  hipStream_t);
template hipError_t launch_gradients<double>(int, int, int, const double *, const double *, double, double, double *,
                                             double *, double *, double *, hipStream_t);
template hipError_t launch_gradients<float>(int, int, int, const float *, const float *, float, float, float *,
                                            float *, float *, float *, hipStream_t);

template <typename T>
hipError_t launch_pm_contract(bool ordered, int nx, int ny, int nk, int nphi, int nout, const T *N, const T *A,
                              const T *B, const T *kx, const T *ky, const T *scale, T *X, T *Y, T *dX, T *dY,
                              hipStream_t s, bool per_column)
{
    const size_t plane = (size_t)nx * ny;
    const size_t lds = pm_factored_lds(nk, nout, B != nullptr, sizeof(T));
    if (ordered || per_column || lds > PM_MOMENT_LDS_MAX) {
        dim3 grid((unsigned)((plane + 255) / 256), (unsigned)((nout + PM_CHUNK - 1) / PM_CHUNK));
        hipLaunchKernelGGL(pm_ordered_kernel<T>, grid, dim3(256), 0, s, plane, nk, nphi, nout, N, A, B, kx, ky,
                           scale, X, Y, dX, dY, per_column);
    } else {
        hipLaunchKernelGGL(pm_moment_kernel<T>, dim3((unsigned)((plane + 63) / 64)), dim3(256), lds, s, plane, nk,
                           nphi, nout, N, A, B, kx, ky, scale, X, Y, dX, dY);
    }
    return hipGetLastError();
}

template hipError_t launch_pm_contract<float>(bool, int, int, int, int, int, const float *, const float *,
                                              const float *, const float *, const float *, const float *, float *,
                                              float *, float *, float *, hipStream_t, bool);
template hipError_t launch_pm_contract<double>(bool, int, int, int, int, int, const double *, const double *,
                                               const double *, const double *, const double *, const double *,
                                               double *, double *, double *, double *, hipStream_t, bool);

// pseudomomentum.jl:72-122 with a depth per column: the overlap tables of
// column ij at its own depth, O[ij + plane (t + nk s)] = sum over z cells of
// W_t W_s / |dz| and dO with dW_t/dkappa, ascending in z as the reference's
// pseudomomentum_overlap; the weights in FP64 (q_cdf_dev).
template <typename T>
__global__ __launch_bounds__(256) void pm_overlap_columns_kernel(size_t plane, int nz, int nk,
                                                                 const double *__restrict__ zf,
                                                                 const double *__restrict__ kc,
                                                                 const double *__restrict__ depth, T *O, T *dO)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int ts = blockIdx.y, t = ts % nk, src = ts / nk;
    if (ij >= plane)
        return;
    const double d = depth[ij];
    double total = 0, dtotal = 0;
    for (int k = 0; k < nz; ++k) {
        const double z1 = zf[k], z2 = zf[k + 1];
        const double hi = z1 < z2 ? z2 : z1, lo = z1 < z2 ? z1 : z2, dz = fabs(z2 - z1);
        const double sw = q_cdf_dev(kc[src], hi, d) - q_cdf_dev(kc[src], lo, d);
        const double tw = q_cdf_dev(kc[t], hi, d) - q_cdf_dev(kc[t], lo, d);
        const double td = q_cdf_dkappa_dev(kc[t], hi, d) - q_cdf_dkappa_dev(kc[t], lo, d);
        total += tw * sw / dz;
        dtotal += td * sw / dz;
    }
    O[ij + plane * (size_t)ts] = (T)total;
    dO[ij + plane * (size_t)ts] = (T)dtotal;
}

// The Q weights of every column at its own depth, W[ij + plane (k + nz m)]
// (q_depth_at in the pseudomomentum cells, pseudomomentum.jl:375-400).
template <typename T>
__global__ __launch_bounds__(256) void q_weights_columns_kernel(size_t plane, int nz, int nk,
                                                                const double *__restrict__ zf,
                                                                const double *__restrict__ kc,
                                                                const double *__restrict__ depth, T *W)
{
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int km = blockIdx.y, k = km % nz, m = km / nz;
    if (ij >= plane)
        return;
    const double z1 = zf[k], z2 = zf[k + 1];
    const double hi = z1 < z2 ? z2 : z1, lo = z1 < z2 ? z1 : z2;
    W[ij + plane * (size_t)km] = (T)(q_cdf_dev(kc[m], hi, depth[ij]) - q_cdf_dev(kc[m], lo, depth[ij]));
}

template <typename T>
hipError_t launch_pm_overlap_columns(size_t plane, int nz, int nk, const double *zf, const double *kc,
                                     const double *depth, T *O, T *dO, hipStream_t s)
{
    if ((long long)nk * nk > 65535)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(pm_overlap_columns_kernel<T>, dim3((unsigned)((plane + 255) / 256), (unsigned)(nk * nk)),
                       dim3(256), 0, s, plane, nz, nk, zf, kc, depth, O, dO);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_q_weights_columns(size_t plane, int nz, int nk, const double *zf, const double *kc,
                                    const double *depth, T *W, hipStream_t s)
{
    if ((long long)nz * nk > 65535)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(q_weights_columns_kernel<T>, dim3((unsigned)((plane + 255) / 256), (unsigned)(nz * nk)),
                       dim3(256), 0, s, plane, nz, nk, zf, kc, depth, W);
    return hipGetLastError();
}

template hipError_t launch_pm_overlap_columns<float>(size_t, int, int, const double *, const double *,
                                                     const double *, float *, float *, hipStream_t);
template hipError_t launch_pm_overlap_columns<double>(size_t, int, int, const double *, const double *,
                                                      const double *, double *, double *, hipStream_t);
template hipError_t launch_q_weights_columns<float>(size_t, int, int, const double *, const double *,
                                                    const double *, float *, hipStream_t);
template hipError_t launch_q_weights_columns<double>(size_t, int, int, const double *, const double *,
                                                     const double *, double *, hipStream_t);

template <typename T>
hipError_t launch_action_tendency(size_t n, const T *Nnew, const T *Nold, T dt, T *out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(action_tendency_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, Nnew, Nold,
                       dt, out);
    return hipGetLastError();
}

template hipError_t launch_action_tendency<float>(size_t, const float *, const float *, float, float *, hipStream_t);
template hipError_t launch_action_tendency<double>(size_t, const double *, const double *, double, double *,
                                                   hipStream_t);

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
