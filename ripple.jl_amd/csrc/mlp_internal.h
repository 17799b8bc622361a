// Internal declarations of the learned source surrogate (not part of the ABI).
#pragma once

#include "rpl_internal.h"

#include <vector>

namespace rpl {
namespace mlp {

typedef unsigned short bf16;

enum {
    EPI_STORE_F32 = 0,       // C = product (FP32)
    EPI_BIAS_GELU = 1,       // Z = product + bias[n] (FP32), A = gelu(Z) (bf16)
    EPI_BIAS_F32 = 2,        // C = product + bias[n] (FP32)
    EPI_GELU_BACKWARD = 3,   // dZ = product * gelu'(Zin) (bf16)
    EPI_BIAS_GELU_INFER = 4, // A = gelu(product + bias[n]) (bf16)
    EPI_BIAS_ROW_F32 = 5,    // C = product + bias[m] (FP32)
    EPI_BIAS_ROW_F16 = 6     // C = product + bias[m] (FP16, to Ch)
};

// C[m, n] = sum_k A(m, k) B(k, n); strides in elements.  M, N multiples of 64,
// K of 32; each operand is contiguous along k or along its row index.
struct GemmArgs {
    int M, N, K;
    const bf16 *A;
    size_t sam, sak;
    const bf16 *B;
    size_t sbk, sbn;
    float *Cf;         // FP32 output (pre-activation or plain product)
    bf16 *Cb;          // bf16 output (activation or activation gradient)
    size_t scm, scn;   // output strides
    const float *bias; // per n (per m for EPI_BIAS_ROW_F32)
    const float *Zin;  // pre-activation read by the GELU backward epilogue (strides scm, scn)
    unsigned char *nonfinite;  // EPI_BIAS_ROW_F32/F16: set to 2 at column n where an output is not finite (or null)
    _Float16 *Ch;              // FP16 output (EPI_BIAS_ROW_F16)
};

struct FeatureScales {
    float floor, shift, scale;  // x = (log(N + floor) - shift) * scale
    float wind_scale;           // U10 components times this
    float depth_feature;        // depth / reference depth
};

struct Envelope {
    float wind_lo, wind_hi;  // training wind range
    float action_hi;         // largest action seen in training
    int audit_stride;        // every audit_stride-th column is audited (0: none)
};

hipError_t gemm(const GemmArgs &g, int epi, hipStream_t s);
hipError_t column_sum(const bf16 *D, int M, int N, float *out, hipStream_t s);
hipError_t mse(const float *Y, const float *T, int rows, int rows_valid, int cols, int cols_valid, bf16 *dY,
               float *partials, int blocks, hipStream_t s);
hipError_t adam(float *w, const float *g, float *m, float *v, bf16 *wb, size_t n, float lr, float b1, float b2,
                float eps, int step, hipStream_t s);
hipError_t to_bf16(const float *src, bf16 *dst, size_t n, hipStream_t s);
// Mini-batch of `batch` rows drawn on the device from a training set of
// `rows` rows (X: rows x in_dim, T: rows x out_dim, FP32) into Xb (rows_pad x
// in_pad, bf16) and Tb (rows_pad x out_pad, FP32), zero padded.
hipError_t gather_rows(const float *X, const float *T, unsigned long long rows, int in_dim, int in_pad, int out_dim,
                       int out_pad, int batch, int rows_pad, unsigned long long seed, unsigned long long step,
                       bf16 *Xb, float *Tb, hipStream_t s);
// Features of columns [col0, col0 + ncols) (ncols a multiple of 64) into X
// rows 0..ncols-1; columns at or past `plane` get zero features.  With amax
// and flags (both cleared by envelope_clear first), also the column's largest
// action and flag bit 0 where an action leaves the envelope (bin_hi per bin,
// may be null; action_hi overall).
template <typename T>
hipError_t features(const T *N, size_t plane, size_t col0, size_t ncols, int nbins, int in_pad, const T *U10,
                    const T *theta, const FeatureScales &fs, bf16 *X, const float *bin_hi, float action_hi,
                    float *amax, unsigned char *flags, hipStream_t s);
hipError_t envelope_clear(size_t plane, float *amax, unsigned char *flags, hipStream_t s);
// use_plain per column from the wind range and the flags (bit 0: action
// outside the envelope; bit 1: non-finite network output), plus the audit grid.
template <typename T>
hipError_t envelope(size_t plane, const T *U10, const unsigned char *flags, const Envelope &env,
                    unsigned char *use_plain, unsigned *counts, hipStream_t s);
template <typename T>
hipError_t audit(const T *N, size_t plane, int nbins, const _Float16 *rate, size_t ld, const float *out_scale,
                 unsigned char *use_plain, int stride, const SourceParams<T> &sp, float bound, unsigned *counts,
                 float *err_max, hipStream_t s);

}  // namespace mlp
}  // namespace rpl

// Device-resident network: FP32 master parameters with bf16 copies, Adam
// moments, and the activations of one batch.  Layer l maps dims[l] -> dims[l+1]
// (all multiples of 64); W_l is [dims[l+1] x dims[l]] row-major.
struct rpl_mlp_impl {
    int device = 0;
    hipStream_t stream = nullptr;
    int in_dim = 0, out_dim = 0, hidden = 0, hidden_layers = 0;
    std::vector<int> dims;  // padded: in_pad, hidden..., out_pad
    int mcap = 0;           // batch capacity (multiple of 64)
    int step = 0;
    std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<rpl::mlp::bf16 *> Wb;
    rpl::mlp::bf16 *X = nullptr;
    std::vector<float *> Z;                // pre-activations of hidden layers
    std::vector<rpl::mlp::bf16 *> A;       // activations of hidden layers
    float *Y = nullptr, *T = nullptr, *Xf = nullptr, *partials = nullptr;
    rpl::mlp::bf16 *dY = nullptr, *dZa = nullptr, *dZb = nullptr;
};

struct rpl_mlp : rpl_mlp_impl {};

namespace rpl {
namespace mlp {
// Inference of `rows` (multiple of 64) feature rows X [rows x in_pad]: hidden
// layers with the bf16 inference epilogue into act_a / act_b, output layer
// computed transposed (W X^T) so that out[row + ld * o] is stored along rows:
// in FP32 to out, or (out null) in FP16 to out_h.
hipError_t forward_rows(rpl_mlp *net, const bf16 *X, int rows, bf16 *act_a, bf16 *act_b, float *out, size_t ld,
                        unsigned char *nonfinite, hipStream_t s, _Float16 *out_h = nullptr);
}  // namespace mlp
}  // namespace rpl
