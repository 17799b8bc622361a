// Spectral diagnostics: per-column weighted sums (m0, kappa_rms, mean
// direction) and the compensated global totals used for conservation checks.
#include "rpl_internal.h"

#include <algorithm>

#include <mutex>

#pragma clang fp contract(off)

namespace rpl {
namespace {

// Diagnostics/diagnostic_operations.jl:71-79 (sum), :110-122 (ratio with
// fallback 0 when the denominator is zero), :153-165 (atan(sy, sx)).  One
// thread per column, spectral loop n outer / m inner as the reference, FP64
// accumulation.  Consecutive threads take consecutive i, so each (m, n) plane
// is read with unit stride.
template <typename T>
__global__ __launch_bounds__(256) void column_diagnostics_kernel(int nx, int ny, int nk, int nphi,
                                                                 const T *__restrict__ N,
                                                                 const double *__restrict__ w,
                                                                 const double *__restrict__ trace,
                                                                 const double *__restrict__ mux,
                                                                 const double *__restrict__ muy, double *m0,
                                                                 double *krms, double *mdir)
{
    const size_t plane = (size_t)nx * ny;
    const size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (ij >= plane)
        return;
    double acc = 0, num = 0, sx = 0, sy = 0;
    for (int n = 0; n < nphi; ++n)
        for (int m = 0; m < nk; ++m) {
            const int t = m + nk * n;
            const double v = (double)N[ij + plane * t];
            acc += v * w[t];
            num += v * trace[t];
            sx += v * mux[t];
            sy += v * muy[t];
        }
    if (m0)
        m0[ij] = acc;
    if (krms)
        krms[ij] = sqrt(acc == 0.0 ? 0.0 : num / acc);
    if (mdir)
        mdir[ij] = atan2(sy, sx);
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    s = x;
    e = (a - av) + (b - bv);
}

// Adds (bh, bl) into (ah, al) keeping the rounding error of the high parts.
__device__ __forceinline__ void dd_add(double &ah, double &al, double bh, double bl)
{
    double s, e;
    two_sum(ah, bh, s, e);
    ah = s;
    al = al + bl + e;
}

// bulk_statistics.jl:87-99: sum over columns of (sum_mn N w) * dx * dy, with
// a double-double accumulator per thread and a double-double tree in LDS.
template <typename T>
__global__ __launch_bounds__(256) void weighted_total_kernel(int nx, int ny, int nk, int nphi,
                                                             const T *__restrict__ N,
                                                             const double *__restrict__ w, double area,
                                                             double *partials)
{
    __shared__ double sh[256], sl[256];
    const size_t plane = (size_t)nx * ny;
    double hi = 0, lo = 0;
    for (size_t ij = (size_t)blockIdx.x * 256 + threadIdx.x; ij < plane; ij += (size_t)gridDim.x * 256) {
        double acc = 0;
        for (int n = 0; n < nphi; ++n)
            for (int m = 0; m < nk; ++m) {
                const int t = m + nk * n;
                acc += (double)N[ij + plane * t] * w[t];
            }
        dd_add(hi, lo, acc * area, 0.0);
    }
    sh[threadIdx.x] = hi;
    sl[threadIdx.x] = lo;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double h = sh[threadIdx.x], l = sl[threadIdx.x];
            dd_add(h, l, sh[threadIdx.x + s], sl[threadIdx.x + s]);
            sh[threadIdx.x] = h;
            sl[threadIdx.x] = l;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0];
        partials[2 * blockIdx.x + 1] = sl[0];
    }
}

// Folds the per-workgroup (hi, lo) pairs in workgroup order into one
// double, the same operations in the same order as the host-side fold of
// rpl_conservation, so the two totals are bit-identical.
__global__ __launch_bounds__(64) void fold_partials_kernel(const double *__restrict__ partials, int nblocks,
                                                           double *__restrict__ out)
{
    if (threadIdx.x != 0)
        return;
    double hi = 0, lo = 0;
    for (int b = 0; b < nblocks; ++b) {
        const double x = partials[2 * b];
        const double s = hi + x, bv = s - hi, av = s - bv;
        lo += (hi - av) + (x - bv) + partials[2 * b + 1];
        hi = s;
    }
    out[0] = hi + lo;
}

}  // namespace

// Workspace of the allocation-free total: one (hi, lo) pair per workgroup of
// weighted_total_kernel, whose grid is capped at NOALLOC_MAX_PARTIALS
// workgroups (16 KiB on every GPU the library runs on, held until the process
// ends).  Host code never passes this name to a kernel: it takes the
// device's address from hipGetSymbolAddress once per device.
constexpr int NOALLOC_MAX_PARTIALS = 1024;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_PARTIALS];

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
static hipError_t scratch_address(double **out)
{
    constexpr int MAX_DEVICES = 64;
    static std::mutex lock;
    static double *addr[MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(lock);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        addr[dev] = (double *)p;
    }
    *out = addr[dev];
    return hipSuccess;
}

int weighted_total_blocks(int nx, int ny)
{
    size_t plane = (size_t)nx * ny;
    size_t b = (plane + 255) / 256;
    return (int)(b < 1024 ? b : 1024);
}

template <typename T>
hipError_t launch_column_diagnostics(int nx, int ny, int nk, int nphi, const T *N, const double *w,
                                     const double *trace, const double *mux, const double *muy, double *m0,
                                     double *krms, double *mdir, hipStream_t s)
{
    size_t plane = (size_t)nx * ny;
    hipLaunchKernelGGL(column_diagnostics_kernel<T>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, nx, ny,
                       nk, nphi, N, w, trace, mux, muy, m0, krms, mdir);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_weighted_total(int nx, int ny, int nk, int nphi, const T *N, const double *w, double area,
                                 double *partials, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(weighted_total_kernel<T>, dim3(nblocks), dim3(256), 0, s, nx, ny, nk, nphi, N, w, area,
                       partials);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_weighted_total_noalloc(int nx, int ny, int nk, int nphi, const T *N, const double *w, double area,
                                         double *out, hipStream_t s)
{
    const int nblocks = weighted_total_blocks(nx, ny);
    if (nblocks > NOALLOC_MAX_PARTIALS)
        return hipErrorInvalidValue;
    double *partials = nullptr;
    hipError_t e = scratch_address(&partials);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(weighted_total_kernel<T>, dim3(nblocks), dim3(256), 0, s, nx, ny, nk, nphi, N, w, area,
                       partials);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    hipLaunchKernelGGL(fold_partials_kernel, dim3(1), dim3(64), 0, s, (const double *)partials, nblocks, out);
    return hipGetLastError();
}

// Non-finite check of the action (a flag, not a repair): one pass over N
// counts its NaN and infinite values into counter[0] (16-byte loads, one
// atomic per workgroup that found any), then one thread records the first
// checked iteration with a nonzero count, record = {iteration, count}, and
// clears the counter for the next check.  Both kernels run on the step's
// stream; the host reads the record when asked.
template <typename T>
__global__ __launch_bounds__(256) void nonfinite_count_kernel(size_t n, const T *__restrict__ N,
                                                              unsigned long long *counter)
{
    constexpr int V = 16 / sizeof(T);
    unsigned long long local = 0;
    const size_t nv = n / V, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nv; k += stride) {
        typedef T Vec __attribute__((ext_vector_type(V)));
        const Vec v = *reinterpret_cast<const Vec *>(N + k * V);
#pragma unroll
        for (int q = 0; q < V; ++q)
            local += !isfinite(v[q]);
    }
    for (size_t k = nv * V + (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        local += !isfinite(N[k]);
    __shared__ unsigned long long part[256 / 64];
    for (int off = 32; off > 0; off >>= 1)
        local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < 256 / 64; ++w)
            tot += part[w];
        if (tot)
            atomicAdd(counter, tot);
    }
}

__global__ void nonfinite_commit_kernel(unsigned long long *counter, long long *record, long long iteration)
{
    const unsigned long long c = *counter;
    if (c && record[0] < 0) {
        record[0] = iteration;
        record[1] = (long long)c;
    }
    *counter = 0;
}

template <typename T>
hipError_t launch_nonfinite_check(size_t n, const T *N, unsigned long long *counter, long long *record,
                                  long long iteration, hipStream_t s)
{
    if ((reinterpret_cast<uintptr_t>(N) & 15) != 0)
        return hipErrorInvalidValue;
    const size_t want = (n / (16 / sizeof(T)) + 255) / 256;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(want, 8192));
    hipLaunchKernelGGL(nonfinite_count_kernel<T>, dim3(blocks), dim3(256), 0, s, n, N, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(nonfinite_commit_kernel, dim3(1), dim3(1), 0, s, counter, record, iteration);
    return hipGetLastError();
}

template hipError_t launch_nonfinite_check<float>(size_t, const float *, unsigned long long *, long long *, long long,
                                                  hipStream_t);
template hipError_t launch_nonfinite_check<double>(size_t, const double *, unsigned long long *, long long *,
                                                   long long, hipStream_t);
template hipError_t launch_weighted_total_noalloc<float>(int, int, int, int, const float *, const double *, double,
                                                         double *, hipStream_t);
template hipError_t launch_weighted_total_noalloc<double>(int, int, int, int, const double *, const double *, double,
                                                          double *, hipStream_t);
template hipError_t launch_column_diagnostics<float>(int, int, int, int, const float *, const double *,
                                                     const double *, const double *, const double *, double *,
                                                     double *, double *, hipStream_t);
template hipError_t launch_column_diagnostics<double>(int, int, int, int, const double *, const double *,
                                                      const double *, const double *, const double *, double *,
                                                      double *, double *, hipStream_t);
template hipError_t launch_weighted_total<float>(int, int, int, int, const float *, const double *, double, double *,
                                                 int, hipStream_t);
template hipError_t launch_weighted_total<double>(int, int, int, int, const double *, const double *, double,
                                                  double *, int, hipStream_t);

}  // namespace rpl
