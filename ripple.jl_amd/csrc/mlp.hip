// Learned source surrogate: BF16 matrix-core GEMMs with FP32 accumulation and
// fused epilogues, GELU, bias reductions, Adam on FP32 master weights, and the
// column feature / fallback kernels that put the network into the RK3 stage.
//
// The reference has no learned component; this network stands in for the
// per-column source evaluation of the PrecomputedSources bundle
// (Physics/Packages/precomputed_sources.jl:240-261) and is trained on that
// bundle's own tendencies.
#include "mlp_internal.h"

#include <cstdint>

namespace rpl {
namespace mlp {
namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float(((unsigned)v) << 16); }

// Round to nearest even; NaN stays NaN.
__device__ __forceinline__ bf16 f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return (bf16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

// Standard normal CDF Phi(z) = erfc(-z / sqrt 2) / 2, with erfc(x), x >= 0,
// from the Chebyshev fit of Press et al., Numerical Recipes in C (2nd ed.,
// section 6.2, erfcc: fractional error below 1.2e-7): erfc(x) = t exp(-x^2 +
// P(t)), t = 1 / (1 + x / 2).  Evaluated with the hardware reciprocal and
// exp2 it stays within 2.5e-5 relative in FP32 over the whole line, also in
// the tail z << 0 where 1 + erf(z / sqrt 2) cancels; it is cheaper than
// erff, which the GEMM epilogues evaluate once per activation.
__device__ __forceinline__ float phi_cdf(float z)
{
    const float x = fabsf(z) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));
    float p = 0.17087277f;
    p = fmaf(p, t, -0.82215223f);
    p = fmaf(p, t, 1.48851587f);
    p = fmaf(p, t, -1.13520398f);
    p = fmaf(p, t, 0.27886807f);
    p = fmaf(p, t, -0.18628806f);
    p = fmaf(p, t, 0.09678418f);
    p = fmaf(p, t, 0.37409196f);
    p = fmaf(p, t, 1.00002368f);
    p = fmaf(p, t, -1.26551223f);
    const float h = 0.5f * t * __builtin_amdgcn_exp2f(fmaf(-x, x, p) * 1.4426950408889634f);  // erfc(x) / 2
    return z >= 0.0f ? 1.0f - h : h;
}
__device__ __forceinline__ float gelu(float z) { return z * phi_cdf(z); }
__device__ __forceinline__ float gelu_grad(float z)
{
    const float pdf = 0.39894228040143268f * __builtin_amdgcn_exp2f(-0.72134752044448170f * z * z);
    return phi_cdf(z) + z * pdf;
}

// C[m, n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn],
// bf16 operands, FP32 accumulation on v_mfma_f32_16x16x16bf16_1k.  Block tile
// 64 x 64, four wavefronts of 32 x 32 (2 x 2 MFMA tiles), K in chunks of 32
// staged through LDS as As[m][k] and Bs[n][k] (k contiguous, so each lane's
// four-k operand is one 8-byte LDS read).  M, N multiples of 64, K of 32.
// Operand layout (measured on gfx950): lane l, element e holds
// A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16]; accumulator
// register r holds D[4 (l / 16) + r][l % 16].
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 16;  // rows padded to 96 bytes (16-byte aligned)

template <bool K_CONTIG>
__device__ __forceinline__ void stage_tile(bf16 (*S)[LDK], const bf16 *__restrict__ G, size_t s_row, size_t s_k,
                                           int tid)
{
    if (K_CONTIG) {
        // 64 rows x 32 k: each thread moves 8 consecutive k (16 bytes).
        const int r = tid >> 2, k = (tid & 3) * 8;
        const uint4 v = *reinterpret_cast<const uint4 *>(G + (size_t)r * s_row + k);
        *reinterpret_cast<uint4 *>(&S[r][k]) = v;
    } else {
        // row index contiguous in memory: each thread moves 8 consecutive rows at one k.
        const int k = tid >> 3, r = (tid & 7) * 8;
        const uint4 v = *reinterpret_cast<const uint4 *>(G + (size_t)k * s_k + r);
        const bf16 *e = reinterpret_cast<const bf16 *>(&v);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            S[r + q][k] = e[q];
    }
}

template <int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs &g, int m, int n, float c)
{
    const size_t o = (size_t)m * g.scm + (size_t)n * g.scn;
    if (EPI == EPI_STORE_F32) {
        g.Cf[o] = c;
    } else if (EPI == EPI_BIAS_GELU) {
        const float z = c + g.bias[n];
        g.Cf[o] = z;
        g.Cb[o] = f2bf(gelu(z));
    } else if (EPI == EPI_BIAS_GELU_INFER) {
        g.Cb[o] = f2bf(gelu(c + g.bias[n]));
    } else if (EPI == EPI_BIAS_F32) {
        g.Cf[o] = c + g.bias[n];
    } else if (EPI == EPI_BIAS_ROW_F32) {
        const float v = c + g.bias[m];
        g.Cf[o] = v;
        if (g.nonfinite && !(fabsf(v) < 3.0e38f))
            g.nonfinite[n] = 2;
    } else if (EPI == EPI_BIAS_ROW_F16) {
        const _Float16 h = (_Float16)(c + g.bias[m]);
        g.Ch[o] = h;
        if (g.nonfinite && !(fabsf((float)h) <= 65504.0f))
            g.nonfinite[n] = 2;
    } else {  // EPI_GELU_BACKWARD: dZ = dA * gelu'(Z)
        g.Cb[o] = f2bf(c * gelu_grad(g.Zin[o]));
    }
}

template <bool A_KC, bool B_KC, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g)
{
    __shared__ bf16 As[BM][LDK];
    __shared__ bf16 Bs[BN][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const bf16 *Ab = g.A + (size_t)m0 * g.sam;
    const bf16 *Bb = g.B + (size_t)n0 * g.sbn;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, q = lane >> 4;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        __syncthreads();
        stage_tile<A_KC>(As, Ab + (size_t)k0 * g.sak, g.sam, g.sak, tid);
        stage_tile<B_KC>(Bs, Bb + (size_t)k0 * g.sbk, g.sbn, g.sbk, tid);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            s16x4 af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
                af[a] = *reinterpret_cast<const s16x4 *>(&As[wm + a * 16 + r16][kk + 4 * q]);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bfr[b] = *reinterpret_cast<const s16x4 *>(&Bs[wn + b * 16 + r16][kk + 4 * q]);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
    }
    // Epilogue.
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + a * 16 + 4 * q + r;
                const int n = n0 + wn + b * 16 + r16;
                epilogue<EPI>(g, m, n, acc[a][b][r]);
            }
}

// Large-tile GEMM for MFMA-bound shapes: block tile 128 x 128, K in steps of
// 64, v_mfma_f32_16x16x32_bf16 (the gfx950 double-rate bf16 instruction).
// Eight wavefronts, each a 64 x 32 block = 4 x 2 MFMA tiles (measured: the
// K loop is bound by the latency of its global loads, and eight waves per
// workgroup -- four per SIMD -- hide more of it than four waves of 64 x 64
// each: 7 % faster inference).  Both tiles
// are staged through LDS as [row][k] (k contiguous, rows padded to 144 bytes
// so the 16-byte fragment reads of one wavefront spread over the banks) in
// two buffers: the next K step is fetched into registers while the current
// one is multiplied.  Operand layout (measured on gfx950): lane l supplies
// A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16], e = 0..7, and
// accumulator register r holds D[4 (l / 16) + r][l % 16].  Needs M and N
// multiples of 128 and K of 64.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int WM = 128, WN = 128, WK = 64, WLD = WK + 8;



// Block order of the wide GEMM: one-dimensional grid; workgroups are dealt
// round-robin to the eight XCDs, so XCD x takes the contiguous range
// [x T / 8, (x+1) T / 8) of the T tiles, and within it the tiles walk the
// smaller tile dimension fastest.  The tiles that share a slice of the large
// operand (the 65,536-column activations) then run together on one XCD and
// read it once through that XCD's L2; the small operand (the weights) stays
// in every L2.  (A two-dimensional grid with M fastest re-read the
// activations from HBM once per tile of the other dimension.)
__device__ __forceinline__ void wide_tile(int b, int mt, int nt, int &m0, int &n0)
{
    const int total = mt * nt;
    const int xcd = b & 7, r = b >> 3;
    const int per = total >> 3, rem = total & 7;
    const int L = xcd < rem ? xcd * (per + 1) + r : rem * (per + 1) + (xcd - rem) * per + r;
    if (nt <= mt) {
        n0 = (L % nt) * WN;
        m0 = (L / nt) * WM;
    } else {
        m0 = (L % mt) * WM;
        n0 = (L / mt) * WN;
    }
}


template <bool K_CONTIG>
__device__ __forceinline__ void fetch_tile(u32x4 (&r)[2], const bf16 *__restrict__ G, size_t s_row, size_t s_k,
                                            int tid)
{
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int idx = tid + 512 * q;
        if (K_CONTIG) {
            const int row = idx >> 3, k = (idx & 7) * 8;
            r[q] = *reinterpret_cast<const u32x4 *>(G + (size_t)row * s_row + k);
        } else {
            const int k = idx >> 4, row = (idx & 15) * 8;
            r[q] = *reinterpret_cast<const u32x4 *>(G + (size_t)k * s_k + row);
        }
    }
}

template <bool K_CONTIG>
__device__ __forceinline__ void store_tile(bf16 (*S)[WLD], const u32x4 (&r)[2], int tid)
{
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int idx = tid + 512 * q;
        if (K_CONTIG) {
            const int row = idx >> 3, k = (idx & 7) * 8;
            *reinterpret_cast<u32x4 *>(&S[row][k]) = r[q];
        } else {
            const int k = idx >> 4, row = (idx & 15) * 8;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                S[row + 2 * u][k] = (bf16)(r[q][u] & 0xffffu);
                S[row + 2 * u + 1][k] = (bf16)(r[q][u] >> 16);
            }
        }
    }
}

template <bool A_KC, bool B_KC, int EPI>
__global__ __launch_bounds__(512) void gemm_wide_kernel(GemmArgs g)
{
    __shared__ bf16 As[2][WM][WLD];
    __shared__ bf16 Bs[2][WN][WLD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int m0, n0;
    wide_tile(blockIdx.x, g.M / WM, g.N / WN, m0, n0);
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
    const bf16 *Ab = g.A + (size_t)m0 * g.sam;
    const bf16 *Bb = g.B + (size_t)n0 * g.sbn;
    f32x4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, q8 = (lane >> 4) * 8;
    u32x4 ra[2], rb[2];
    fetch_tile<A_KC>(ra, Ab, g.sam, g.sak, tid);
    fetch_tile<B_KC>(rb, Bb, g.sbn, g.sbk, tid);
    store_tile<A_KC>(As[0], ra, tid);
    store_tile<B_KC>(Bs[0], rb, tid);
    __syncthreads();
    const int nk = g.K / WK;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            fetch_tile<A_KC>(ra, Ab + (size_t)(kt + 1) * WK * g.sak, g.sam, g.sak, tid);
            fetch_tile<B_KC>(rb, Bb + (size_t)(kt + 1) * WK * g.sbk, g.sbn, g.sbk, tid);
        }
#pragma unroll
        for (int ks = 0; ks < WK; ks += 32) {
            bf16x8 af[4], bfr[2];
#pragma unroll
            for (int a = 0; a < 4; ++a)
                af[a] = *reinterpret_cast<const bf16x8 *>(&As[cur][wm + a * 16 + r16][ks + q8]);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bfr[b] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn + b * 16 + r16][ks + q8]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        if (more) {
            store_tile<A_KC>(As[cur ^ 1], ra, tid);
            store_tile<B_KC>(Bs[cur ^ 1], rb, tid);
        }
        __syncthreads();
    }
    if constexpr (EPI == EPI_BIAS_GELU_INFER) {
        if (g.scn == 1 && (g.scm & 7) == 0 && ((size_t)g.Cb & 15) == 0) {
            // Each wave stages its 64 x 32 bf16 results in LDS (As and Bs are
            // free after the last barrier of the K loop; four waves in each)
            // and writes them as 16-byte row chunks, instead of 32 two-byte
            // stores per lane.
            float bn[2];
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bn[b] = g.bias[n0 + wn + b * 16 + r16];
            bf16(*T)[40] = reinterpret_cast<bf16(*)[40]>(wave < 4 ? &As[0][0][0] : &Bs[0][0][0]) + (wave & 3) * 64;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        T[a * 16 + 4 * (lane >> 4) + r][b * 16 + r16] = f2bf(gelu(acc[a][b][r] + bn[b]));
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = lane + 64 * k, row = c >> 2, col = (c & 3) * 8;
                const u32x4 v = *reinterpret_cast<const u32x4 *>(&T[row][col]);
                *reinterpret_cast<u32x4 *>(g.Cb + (size_t)(m0 + wm + row) * g.scm + (n0 + wn + col)) = v;
            }
            return;
        }
    }
    if constexpr (EPI == EPI_BIAS_ROW_F32) {
        // the row bias of the thread's 16 rows, loaded once
        float bm[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                bm[a][r] = g.bias[m0 + wm + a * 16 + 4 * (lane >> 4) + r];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
                    const int n = n0 + wn + b * 16 + r16;
                    const float v = acc[a][b][r] + bm[a][r];
                    g.Cf[(size_t)m * g.scm + (size_t)n * g.scn] = v;
                    if (g.nonfinite && !(fabsf(v) < 3.0e38f))
                        g.nonfinite[n] = 2;
                }
        return;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + b * 16 + r16;
                epilogue<EPI>(g, m, n, acc[a][b][r]);
            }
}

// The same 128 x 128 block tile and eight 64 x 32 wavefront blocks for
// operands that are both contiguous along k (the inference GEMMs and the
// training forward), with the K steps moved from global memory straight into
// LDS by global_load_lds_dwordx4: no staging registers and no LDS store
// instructions, so the next step's loads run under the current step's MFMAs
// with one barrier per step.  A load instruction fills 1 KB of LDS, lane l at
// byte 16 l, i.e. eight 128-byte rows of one 64-deep K step; each lane picks
// the global 16-byte chunk that belongs at its LDS position.  Within a row
// chunk c is stored at position c ^ ((row >> 1) & 7): the rows of a
// fragment read, all at one chunk, then spread over the banks.  Measured
// with rocprofv3 (SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE): no conflicts in
// the K loop, against 0.5 with c ^ (row & 7) ^ ((row >> 3) & 1) and 0.7
// unswizzled (1.7x slower).  Measured at the surrogate's shapes (65,536
// columns, 512 outputs; rate layer 1152 x 65,536) against the
// register-staged kernel above: K = 1216 150 -> 108 us, K = 512 72 -> 61 us,
// rate layer 191 -> 180 us.  Not kept: all of a step's fragment reads
// before its MFMAs (no change; with a scheduling barrier between them 33 %
// slower), raised wave priority around the MFMAs (no change), a third
// activation buffer (80 KB per workgroup: the compiler then waits for every
// load in flight before the fragment reads, or spills when unrolled).
// Needs M and N multiples of 128, K of 64, 16-byte aligned operands and row
// strides that are multiples of 8 elements.
typedef __attribute__((address_space(1))) const void *global_ptr;
typedef __attribute__((address_space(3))) void *lds_ptr;

__device__ __forceinline__ int chunk_pos(int c, int row) { return c ^ ((row >> 1) & 7); }

// 128 rows x 64 k of G (row stride s) into the 16 KB step buffer S: two
// load instructions per wavefront.
__device__ __forceinline__ void dma_step(char *S, const bf16 *__restrict__ G, size_t s, int wave, int lane)
{
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r0 = (wave * 2 + i) * 8;
        const int row = r0 + (lane >> 3);
        const bf16 *src = G + (size_t)row * s + chunk_pos(lane & 7, row) * 8;
        __builtin_amdgcn_global_load_lds((global_ptr)src, (lds_ptr)(S + r0 * 128), 16, 0, 0);
    }
}

template <int EPI>
__global__ __launch_bounds__(512) void gemm_dma_kernel(GemmArgs g)
{
    // Two step buffers per operand as separate arrays, so that the compiler
    // sees that the loads into one do not feed the reads of the other.
    __shared__ __attribute__((aligned(1024))) char As0[WM * 128];
    __shared__ __attribute__((aligned(1024))) char As1[WM * 128];
    __shared__ __attribute__((aligned(1024))) char Bs0[WN * 128];
    __shared__ __attribute__((aligned(1024))) char Bs1[WN * 128];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int m0, n0;
    wide_tile(blockIdx.x, g.M / WM, g.N / WN, m0, n0);
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
    const bf16 *Ab = g.A + (size_t)m0 * g.sam;
    const bf16 *Bb = g.B + (size_t)n0 * g.sbn;
    f32x4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Fragment of a 16-row tile at k = 32 h + 8 (l / 16): row l % 16, chunk 4 h + l / 16.
    const int r16 = lane & 15;
    const int off[2] = {r16 * 128 + chunk_pos(lane >> 4, r16) * 16, r16 * 128 + chunk_pos(4 + (lane >> 4), r16) * 16};
    auto multiply = [&](const char *As, const char *Bs) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 af[4], bfr[2];
#pragma unroll
            for (int a = 0; a < 4; ++a)
                af[a] = *reinterpret_cast<const bf16x8 *>(As + (wm + a * 16) * 128 + off[h]);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bfr[b] = *reinterpret_cast<const bf16x8 *>(Bs + (wn + b * 16) * 128 + off[h]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
    };
    const int nk = g.K / WK;
    dma_step(As0, Ab, g.sam, wave, lane);
    dma_step(Bs0, Bb, g.sbn, wave, lane);
    // Two K steps per trip, one per buffer pair.  Before each step every
    // wavefront waits for its own loads (vmcnt 0) and the barrier makes all
    // of them visible; the loads issued after the barrier refill the pair
    // that every wavefront finished reading before it.
    for (int kt = 0; kt < nk; kt += 2) {
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        __syncthreads();
        if (kt + 1 < nk) {
            dma_step(As1, Ab + (size_t)(kt + 1) * WK, g.sam, wave, lane);
            dma_step(Bs1, Bb + (size_t)(kt + 1) * WK, g.sbn, wave, lane);
        }
        multiply(As0, Bs0);
        if (kt + 1 >= nk)
            break;
        __builtin_amdgcn_s_waitcnt(0x0f70);
        __syncthreads();
        if (kt + 2 < nk) {
            dma_step(As0, Ab + (size_t)(kt + 2) * WK, g.sam, wave, lane);
            dma_step(Bs0, Bb + (size_t)(kt + 2) * WK, g.sbn, wave, lane);
        }
        multiply(As1, Bs1);
    }
    if constexpr (EPI == EPI_BIAS_GELU_INFER) {
        if (g.scn == 1 && (g.scm & 7) == 0 && ((size_t)g.Cb & 15) == 0) {
            // Each wavefront stages its 64 x 32 bf16 results (4 KB; four
            // wavefronts in As0, four in Bs0, free after the barrier) with
            // the row's four 16-byte chunks rotated by row, and writes them
            // as 16-byte row chunks.
            __syncthreads();
            float bn[2];
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bn[b] = g.bias[n0 + wn + b * 16 + r16];
            char *R = (wave < 4 ? As0 : Bs0) + (wave & 3) * 4096;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = a * 16 + 4 * (lane >> 4) + r, col = b * 16 + r16;
                        *reinterpret_cast<bf16 *>(R + row * 64 + (((col >> 3) ^ (row & 3)) * 16) + (col & 7) * 2) =
                            f2bf(gelu(acc[a][b][r] + bn[b]));
                    }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = lane + 64 * k, row = c >> 2, j = c & 3;
                const u32x4 v = *reinterpret_cast<const u32x4 *>(R + row * 64 + ((j ^ (row & 3)) * 16));
                *reinterpret_cast<u32x4 *>(g.Cb + (size_t)(m0 + wm + row) * g.scm + (n0 + wn + j * 8)) = v;
            }
            return;
        }
    }
    if constexpr (EPI == EPI_BIAS_ROW_F32) {
        float bm[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                bm[a][r] = g.bias[m0 + wm + a * 16 + 4 * (lane >> 4) + r];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
                    const int n = n0 + wn + b * 16 + r16;
                    const float v = acc[a][b][r] + bm[a][r];
                    g.Cf[(size_t)m * g.scm + (size_t)n * g.scn] = v;
                    if (g.nonfinite && !(fabsf(v) < 3.0e38f))
                        g.nonfinite[n] = 2;
                }
        return;
    }
    if constexpr (EPI == EPI_BIAS_ROW_F16) {
        if (g.scn == 1 && (g.scm & 7) == 0 && ((size_t)g.Ch & 15) == 0) {
            // The rate planes in FP16: each wavefront stages its 64 x 32
            // results (4 KB, as the GELU epilogue) and writes 64-byte row
            // segments in 16-byte chunks.  A value that FP16 cannot hold
            // flags its column.
            __syncthreads();
            char *R = (wave < 4 ? As0 : Bs0) + (wave & 3) * 4096;
            bool bad[2] = {false, false};
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = a * 16 + 4 * (lane >> 4) + r;
                    const float bm = g.bias[m0 + wm + row];
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const int col = b * 16 + r16;
                        const _Float16 h = (_Float16)(acc[a][b][r] + bm);
                        bad[b] = bad[b] || !(fabsf((float)h) <= 65504.0f);
                        *reinterpret_cast<_Float16 *>(R + row * 64 + (((col >> 3) ^ (row & 3)) * 16) + (col & 7) * 2) =
                            h;
                    }
                }
            if (g.nonfinite) {
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    if (bad[b])
                        g.nonfinite[n0 + wn + b * 16 + r16] = 2;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = lane + 64 * k, row = c >> 2, j = c & 3;
                const u32x4 v = *reinterpret_cast<const u32x4 *>(R + row * 64 + ((j ^ (row & 3)) * 16));
                *reinterpret_cast<u32x4 *>(g.Ch + (size_t)(m0 + wm + row) * g.scm + (n0 + wn + j * 8)) = v;
            }
            return;
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + b * 16 + r16;
                epilogue<EPI>(g, m, n, acc[a][b][r]);
            }
}

__global__ __launch_bounds__(256) void column_sum_kernel(const bf16 *__restrict__ D, int M, int N, float *out)
{
    // out[n] = sum over m of D[m*N + n] (bias gradient).  Thread per n, rows in
    // four interleaved partial sums for accuracy and ILP.
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N)
        return;
    float s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int m = 0;
    for (; m + 3 < M; m += 4) {
        s0 += bf2f(D[(size_t)m * N + n]);
        s1 += bf2f(D[(size_t)(m + 1) * N + n]);
        s2 += bf2f(D[(size_t)(m + 2) * N + n]);
        s3 += bf2f(D[(size_t)(m + 3) * N + n]);
    }
    for (; m < M; ++m)
        s0 += bf2f(D[(size_t)m * N + n]);
    out[n] = (s0 + s1) + (s2 + s3);
}

// Mean-squared error over the valid rows and output features, and its
// gradient dY = 2 (Y - T) / count in bf16 (zero on padded rows / columns).
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ Y, const float *__restrict__ T, int rows,
                                                  int rows_valid, int cols, int cols_valid, float inv_count,
                                                  bf16 *dY, float *loss_partial)
{
    __shared__ float red[256];
    const size_t total = (size_t)rows * cols;
    float s = 0;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < total; k += (size_t)gridDim.x * 256) {
        const int m = (int)(k / cols), n = (int)(k % cols);
        float d = 0;
        if (m < rows_valid && n < cols_valid) {
            d = Y[k] - T[k];
            s += d * d;
        }
        dY[k] = f2bf(2.0f * d * inv_count);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        loss_partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on FP32 master parameters, bias-corrected, writing
// the bf16 copies the GEMMs read.
__global__ __launch_bounds__(256) void adam_kernel(float *w, const float *__restrict__ g, float *mom, float *vel,
                                                   bf16 *wb, size_t n, float lr, float b1, float b2, float eps,
                                                   float c1, float c2)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n)
        return;
    const float gk = g[k];
    const float mk = b1 * mom[k] + (1 - b1) * gk;
    const float vk = b2 * vel[k] + (1 - b2) * gk * gk;
    mom[k] = mk;
    vel[k] = vk;
    const float upd = lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    const float nw = w[k] - upd;
    w[k] = nw;
    if (wb)
        wb[k] = f2bf(nw);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float *__restrict__ src, bf16 *dst, size_t n)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n)
        dst[k] = f2bf(src[k]);
}

// One mini-batch row per workgroup from a device-resident training set: row
// r < batch takes set row splitmix64(seed + step * 2^32 + r) mod rows (uniform
// draws with replacement), converted to bf16 features and FP32 targets with
// zero padding; rows past the batch are zero.
__global__ __launch_bounds__(256) void gather_rows_kernel(const float *__restrict__ X, const float *__restrict__ T,
                                                          unsigned long long rows, int in_dim, int in_pad,
                                                          int out_dim, int out_pad, int batch,
                                                          unsigned long long key, bf16 *__restrict__ Xb,
                                                          float *__restrict__ Tb)
{
    const int r = blockIdx.x;
    unsigned long long src = 0;
    const bool live = r < batch;
    if (live) {
        unsigned long long z = key + (unsigned long long)r + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        src = (z ^ (z >> 31)) % rows;
    }
    const float *xs = X + src * (unsigned long long)in_dim;
    const float *ts = T + src * (unsigned long long)out_dim;
    for (int c = threadIdx.x; c < in_pad; c += 256)
        Xb[(size_t)r * in_pad + c] = f2bf(live && c < in_dim ? xs[c] : 0.0f);
    for (int c = threadIdx.x; c < out_pad; c += 256)
        Tb[(size_t)r * out_pad + c] = live && c < out_dim ? ts[c] : 0.0f;
}

// Four consecutive values of an action plane as floats: one 16-byte load
// (FP32) or two (FP64); `at` must be 4-element aligned.
__device__ __forceinline__ void load4(const float *at, float (&a)[4])
{
    const float4 t = *reinterpret_cast<const float4 *>(at);
    a[0] = t.x, a[1] = t.y, a[2] = t.z, a[3] = t.w;
}
__device__ __forceinline__ void load4(const double *at, float (&a)[4])
{
    const double2 t0 = *reinterpret_cast<const double2 *>(at), t1 = *reinterpret_cast<const double2 *>(at + 2);
    a[0] = (float)t0.x, a[1] = (float)t0.y, a[2] = (float)t1.x, a[3] = (float)t1.y;
}

// Surrogate features of one 256-column x 64-feature tile per workgroup:
// x = (log(N + floor) - shift) * scale for every bin, then U10x/ref,
// U10y/ref, depth/ref, zero padding.  Lane l of wavefront w reads columns
// 4 l .. 4 l + 3 of bins w, w + 4, ..., w + 60 (one 16-byte load per bin
// and lane, all sixteen issued before any is used), and the bf16 features
// go through an LDS tile [bin][column] (rows of 258 values: the transposed
// reads of eight lanes per column fall on distinct banks) to 16-byte stores
// of each column's 64 features.  Measured at 65,536 columns x 1152 bins:
// 156 -> 105 us per chunk (3.0 -> 4.4 TB/s) against one column per lane
// with four-byte loads, output bit-identical.  With `flags` set, the tile
// also folds its part of the envelope: the column's largest action (atomic
// max of the non-negative bit patterns into amax) and bit 0 of flags where a
// bin exceeds its training maximum bin_hi (if given) or the overall
// action_hi.  Columns past the plane get zero features.
template <typename T>
__global__ __launch_bounds__(256) void features_kernel(const T *__restrict__ N, size_t plane, size_t col0,
                                                       size_t ncols, int nbins, int in_pad, bool vec,
                                                       const T *__restrict__ U10, const T *__restrict__ theta,
                                                       FeatureScales fs, bf16 *X, const float *__restrict__ bin_hi,
                                                       float action_hi, float *amax, unsigned char *flags)
{
    constexpr int RW = 258;
    __shared__ bf16 tile[64 * RW];
    const size_t c0 = col0 + (size_t)blockIdx.x * 256, cend = col0 + ncols;
    const int b0 = blockIdx.y * 64;
    const int cq = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t colq = c0 + 4 * cq;
    float mx[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool over[4] = {false, false, false, false};
    auto feature = [&](float a) {
        return (__builtin_amdgcn_logf(fmaxf(a, 0.0f) + fs.floor) * 0.693147180559945f - fs.shift) * fs.scale;
    };
    auto put = [&](int r, const float (&v)[4]) {
        unsigned *d = reinterpret_cast<unsigned *>(tile + r * RW + 4 * cq);
        d[0] = (unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16);
        d[1] = (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16);
    };
    if (vec && colq + 4 <= plane && b0 + 64 <= nbins) {
        float a[16][4];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            load4(N + colq + plane * (size_t)(b0 + w + 4 * q), a[q]);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = w + 4 * q;
            const float hi = bin_hi ? bin_hi[b0 + r] : 0.0f;
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = feature(a[q][u]);
                mx[u] = fmaxf(mx[u], a[q][u]);
                if (bin_hi)
                    over[u] = over[u] || !(a[q][u] <= hi);
            }
            put(r, v);
        }
    } else {
        for (int q = 0; q < 16; ++q) {
            const int r = w + 4 * q, bin = b0 + r;
            float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const size_t col = colq + u;
                if (col >= plane)
                    continue;
                if (bin < nbins) {
                    const float a = (float)N[col + plane * (size_t)bin];
                    v[u] = feature(a);
                    mx[u] = fmaxf(mx[u], a);
                    if (bin_hi)
                        over[u] = over[u] || !(a <= bin_hi[bin]);
                } else if (bin == nbins) {
                    v[u] = (float)U10[col] * cosf((float)theta[col]) * fs.wind_scale;
                } else if (bin == nbins + 1) {
                    v[u] = (float)U10[col] * sinf((float)theta[col]) * fs.wind_scale;
                } else if (bin == nbins + 2) {
                    v[u] = fs.depth_feature;
                }
            }
            put(r, v);
        }
    }
    __syncthreads();
    // entry q of the 2048 takes column q / 8 and features 8 (q % 8) .. + 7
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int q = threadIdx.x + 256 * k, c = q >> 3, h = q & 7;
        const size_t cl = c0 + c;
        unsigned wv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            wv[e] = (unsigned)tile[(8 * h + 2 * e) * RW + c] | ((unsigned)tile[(8 * h + 2 * e + 1) * RW + c] << 16);
        if (cl < cend)
            *reinterpret_cast<uint4 *>(X + (cl - col0) * (size_t)in_pad + b0 + 8 * h) =
                make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
    if (flags) {
        // the four wavefronts' column maxima through the tile, a bin above
        // its maximum in the sign bit (the maxima are non-negative)
        float(*red)[256] = reinterpret_cast<float(*)[256]>(tile);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u)
            red[w][4 * cq + u] = __uint_as_float(__float_as_uint(mx[u]) | (over[u] ? 0x80000000u : 0u));
        __syncthreads();
        const int c = threadIdx.x;
        const size_t col = c0 + c;
        if (col < plane && col < cend) {
            float m = 0.0f;
            bool o = false;
#pragma unroll
            for (int ww = 0; ww < 4; ++ww) {
                const unsigned r = __float_as_uint(red[ww][c]);
                o = o || (r >> 31);
                m = fmaxf(m, __uint_as_float(r & 0x7fffffffu));
            }
            if (m > 0.0f)
                atomicMax(reinterpret_cast<unsigned *>(amax) + col, __float_as_uint(m));  // m >= 0: bit order = value order
            if (o || !(m <= action_hi))
                flags[col] = 1;
        }
    }
}

// Clears the per-column envelope accumulators before a stage's features.
__global__ __launch_bounds__(256) void envelope_clear_kernel(size_t plane, float *amax, unsigned char *flags)
{
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (col < plane) {
        amax[col] = 0.0f;
        flags[col] = 0;
    }
}

// Envelope and audit decision per column: the plain bundle is used where the
// wind is outside the training range, where the features found an action
// outside it (flags bit 0), and where the network produced a non-finite rate
// (flags bit 1, set by the output layer's epilogue); every audit_stride-th
// column is flagged for the audit.
template <typename T>
__global__ __launch_bounds__(256) void envelope_kernel(size_t plane, const T *__restrict__ U10,
                                                       const unsigned char *__restrict__ flags, Envelope env,
                                                       unsigned char *use_plain, unsigned *counts)
{
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= plane)
        return;
    const float u = (float)U10[col];
    const bool plain = !(u >= env.wind_lo && u <= env.wind_hi) || flags[col] != 0;
    const bool audit = env.audit_stride > 0 && (col % (size_t)env.audit_stride) == 0;
    use_plain[col] = (unsigned char)((plain ? 1 : 0) | (audit ? 2 : 0));
    if (plain)
        atomicAdd(&counts[0], 1u);
    if (audit)
        atomicAdd(&counts[1], 1u);
}

// Audit: on audit columns compare the network's source with the plain
// bundle's (both on the same state, sp with the surrogate switched off);
// columns whose relative L2 error exceeds the bound are counted and marked
// (bit 2) so that the stage takes the plain source there.
template <typename T>
__global__ __launch_bounds__(64) void audit_kernel(const T *__restrict__ N, size_t plane, int nbins,
                                                   const _Float16 *__restrict__ rate, size_t ld,
                                                   const float *__restrict__ out_scale,
                                                   unsigned char *use_plain, int stride, SourceParams<T> sp,
                                                   float bound, unsigned *counts, float *err_max)
{
    const size_t col = (size_t)blockIdx.x * (size_t)stride;
    if (col >= plane || (use_plain[col] & 3) != 2)
        return;
    float num = 0, den = 0;
    for (int b = threadIdx.x; b < nbins; b += 64) {
        const T a = N[col + plane * b];
        const float sp_plain = (float)source_term<T>(sp, (unsigned)col, b, a);
        const float sp_net = (float)rate[col + ld * b] * out_scale[b] * (float)a;
        num += (sp_net - sp_plain) * (sp_net - sp_plain);
        den += sp_plain * sp_plain;
    }
    for (int off = 32; off > 0; off >>= 1) {
        num += __shfl_down(num, off);
        den += __shfl_down(den, off);
    }
    if (threadIdx.x == 0) {
        const float e = den > 0 ? sqrtf(num / den) : (num > 0 ? 1.0f : 0.0f);
        if (!(e <= bound)) {
            atomicAdd(&counts[2], 1u);
            use_plain[col] |= 4;
        }
        atomicMax(reinterpret_cast<unsigned *>(err_max), __float_as_uint(e));  // e >= 0: bit order = value order
    }
}

}  // namespace

template <bool A_KC, bool B_KC, int EPI>
static void gemm_launch(const GemmArgs &g, hipStream_t s)
{
    const bool wide = g.M % WM == 0 && g.N % WN == 0 && g.K % WK == 0;
    const bool dma_ok = A_KC && B_KC && ((size_t)g.A & 15) == 0 && ((size_t)g.B & 15) == 0 && (g.sam & 7) == 0 &&
                        (g.sbn & 7) == 0;
    if (wide && dma_ok)
        hipLaunchKernelGGL((gemm_dma_kernel<EPI>), dim3((g.M / WM) * (g.N / WN)), dim3(512), 0, s, g);
    else if (wide)
        hipLaunchKernelGGL((gemm_wide_kernel<A_KC, B_KC, EPI>), dim3((g.M / WM) * (g.N / WN)), dim3(512), 0, s, g);
    else
        hipLaunchKernelGGL((gemm_kernel<A_KC, B_KC, EPI>), dim3(g.M / BM, g.N / BN), dim3(256), 0, s, g);
}

template <bool A_KC, bool B_KC>
static hipError_t gemm_dispatch(const GemmArgs &g, int epi, hipStream_t s)
{
    switch (epi) {
    case EPI_STORE_F32:
        gemm_launch<A_KC, B_KC, EPI_STORE_F32>(g, s);
        break;
    case EPI_BIAS_GELU:
        gemm_launch<A_KC, B_KC, EPI_BIAS_GELU>(g, s);
        break;
    case EPI_BIAS_F32:
        gemm_launch<A_KC, B_KC, EPI_BIAS_F32>(g, s);
        break;
    case EPI_BIAS_GELU_INFER:
        gemm_launch<A_KC, B_KC, EPI_BIAS_GELU_INFER>(g, s);
        break;
    case EPI_BIAS_ROW_F32:
        gemm_launch<A_KC, B_KC, EPI_BIAS_ROW_F32>(g, s);
        break;
    case EPI_BIAS_ROW_F16:
        gemm_launch<A_KC, B_KC, EPI_BIAS_ROW_F16>(g, s);
        break;
    default:
        gemm_launch<A_KC, B_KC, EPI_GELU_BACKWARD>(g, s);
        break;
    }
    return hipGetLastError();
}

hipError_t gemm(const GemmArgs &g, int epi, hipStream_t s)
{
    if (g.M % BM || g.N % BN || g.K % BK)
        return hipErrorInvalidValue;
    const bool akc = g.sak == 1, bkc = g.sbk == 1;
    // The non-k-contiguous operand must be contiguous along its row index.
    if ((!akc && g.sam != 1) || (!bkc && g.sbn != 1))
        return hipErrorInvalidValue;
    if (akc && bkc)
        return gemm_dispatch<true, true>(g, epi, s);
    if (akc)
        return gemm_dispatch<true, false>(g, epi, s);
    if (bkc)
        return gemm_dispatch<false, true>(g, epi, s);
    return gemm_dispatch<false, false>(g, epi, s);
}

hipError_t column_sum(const bf16 *D, int M, int N, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(column_sum_kernel, dim3((N + 255) / 256), dim3(256), 0, s, D, M, N, out);
    return hipGetLastError();
}

hipError_t mse(const float *Y, const float *T, int rows, int rows_valid, int cols, int cols_valid, bf16 *dY,
               float *partials, int blocks, hipStream_t s)
{
    const float inv = 1.0f / ((float)rows_valid * (float)cols_valid);
    hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, s, Y, T, rows, rows_valid, cols, cols_valid, inv, dY,
                       partials);
    return hipGetLastError();
}

hipError_t adam(float *w, const float *g, float *m, float *v, bf16 *wb, size_t n, float lr, float b1, float b2,
                float eps, int step, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, g, m, v, wb, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t gather_rows(const float *X, const float *T, unsigned long long rows, int in_dim, int in_pad, int out_dim,
                       int out_pad, int batch, int rows_pad, unsigned long long seed, unsigned long long step,
                       bf16 *Xb, float *Tb, hipStream_t s)
{
    if (rows == 0 || batch > rows_pad)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)rows_pad), dim3(256), 0, s, X, T, rows, in_dim, in_pad,
                       out_dim, out_pad, batch, seed + (step << 32), Xb, Tb);
    return hipGetLastError();
}

hipError_t to_bf16(const float *src, bf16 *dst, size_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

template <typename T>
hipError_t features(const T *N, size_t plane, size_t col0, size_t ncols, int nbins, int in_pad, const T *U10,
                    const T *theta, const FeatureScales &fs, bf16 *X, const float *bin_hi, float action_hi,
                    float *amax, unsigned char *flags, hipStream_t s)
{
    if (ncols % 64 || in_pad % 64)
        return hipErrorInvalidValue;
    const bool vec = plane % 4 == 0 && col0 % 4 == 0 && ((size_t)N & 15) == 0;
    hipLaunchKernelGGL(features_kernel<T>, dim3((unsigned)((ncols + 255) / 256), (unsigned)(in_pad / 64)), dim3(256), 0,
                       s, N, plane, col0, ncols, nbins, in_pad, vec, U10, theta, fs, X, bin_hi, action_hi, amax,
                       flags);
    return hipGetLastError();
}

hipError_t envelope_clear(size_t plane, float *amax, unsigned char *flags, hipStream_t s)
{
    hipLaunchKernelGGL(envelope_clear_kernel, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, plane, amax,
                       flags);
    return hipGetLastError();
}

template <typename T>
hipError_t envelope(size_t plane, const T *U10, const unsigned char *flags, const Envelope &env,
                    unsigned char *use_plain, unsigned *counts, hipStream_t s)
{
    hipLaunchKernelGGL(envelope_kernel<T>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, plane, U10, flags,
                       env, use_plain, counts);
    return hipGetLastError();
}

template <typename T>
hipError_t audit(const T *N, size_t plane, int nbins, const _Float16 *rate, size_t ld, const float *out_scale,
                 unsigned char *use_plain, int stride, const SourceParams<T> &sp, float bound, unsigned *counts,
                 float *err_max, hipStream_t s)
{
    if (stride < 1)
        return hipSuccess;
    hipLaunchKernelGGL(audit_kernel<T>, dim3((unsigned)((plane + stride - 1) / stride)), dim3(64), 0, s, N, plane,
                       nbins, rate, ld, out_scale, use_plain, stride, sp, bound, counts, err_max);
    return hipGetLastError();
}

template hipError_t features<float>(const float *, size_t, size_t, size_t, int, int, const float *, const float *,
                                    const FeatureScales &, bf16 *, const float *, float, float *, unsigned char *,
                                    hipStream_t);
template hipError_t features<double>(const double *, size_t, size_t, size_t, int, int, const double *,
                                     const double *, const FeatureScales &, bf16 *, const float *, float, float *,
                                     unsigned char *, hipStream_t);
template hipError_t envelope<float>(size_t, const float *, const unsigned char *, const Envelope &, unsigned char *,
                                    unsigned *, hipStream_t);
template hipError_t envelope<double>(size_t, const double *, const unsigned char *, const Envelope &,
                                     unsigned char *, unsigned *, hipStream_t);
template hipError_t audit<float>(const float *, size_t, int, const _Float16 *, size_t, const float *,
                                 unsigned char *, int, const SourceParams<float> &, float, unsigned *, float *,
                                 hipStream_t);
template hipError_t audit<double>(const double *, size_t, int, const _Float16 *, size_t, const float *,
                                  unsigned char *, int, const SourceParams<double> &, float, unsigned *, float *,
                                  hipStream_t);

}  // namespace mlp
}  // namespace rpl
