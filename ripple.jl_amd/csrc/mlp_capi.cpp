// Host side of the learned-surrogate network: parameter storage, the training
// step (forward, MSE, backward, Adam) and the C ABI for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mlp_internal.h"

using rpl::mlp::bf16;

namespace {

int set_err(int code, const std::string &m) { return rpl::set_error(code, m.c_str()); }

#define MLP_HIP(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_err(RPL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

int round64(int x) { return (x + 63) / 64 * 64; }

template <typename P>
int dalloc(P **p, size_t count)
{
    hipError_t e = hipMalloc((void **)p, count * sizeof(P) + 16);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(RPL_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return RPL_OK;
}

size_t layer_w(const rpl_mlp *n, int l) { return (size_t)n->dims[l + 1] * n->dims[l]; }
int nlayers(const rpl_mlp *n) { return (int)n->dims.size() - 1; }

int64_t param_count(const rpl_mlp *n)
{
    int64_t c = 0;
    for (int l = 0; l < nlayers(n); ++l)
        c += (int64_t)layer_w(n, l) + n->dims[l + 1];
    return c;
}

// splitmix64: the parameter initialiser's generator.
uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int upload_params(rpl_mlp *n, const float *p)
{
    size_t off = 0;
    for (int l = 0; l < nlayers(n); ++l) {
        const size_t nw = layer_w(n, l), nb = n->dims[l + 1];
        MLP_HIP(hipMemcpy(n->W[l], p + off, nw * 4, hipMemcpyHostToDevice));
        off += nw;
        MLP_HIP(hipMemcpy(n->b[l], p + off, nb * 4, hipMemcpyHostToDevice));
        off += nb;
        MLP_HIP(rpl::mlp::to_bf16(n->W[l], n->Wb[l], nw, n->stream));
    }
    MLP_HIP(hipStreamSynchronize(n->stream));
    return RPL_OK;
}

// Pads a host batch into the FP32 staging buffer and converts to bf16 X.
int load_inputs(rpl_mlp *n, int batch, const float *X)
{
    const int in_pad = n->dims[0], rows = round64(batch);
    std::vector<float> h((size_t)rows * in_pad, 0.0f);
    for (int r = 0; r < batch; ++r)
        std::memcpy(&h[(size_t)r * in_pad], X + (size_t)r * n->in_dim, n->in_dim * 4);
    MLP_HIP(hipMemcpy(n->Xf, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    MLP_HIP(rpl::mlp::to_bf16(n->Xf, n->X, h.size(), n->stream));
    return RPL_OK;
}

int load_targets(rpl_mlp *n, int batch, const float *T)
{
    const int out_pad = n->dims.back(), rows = round64(batch);
    std::vector<float> h((size_t)rows * out_pad, 0.0f);
    for (int r = 0; r < batch; ++r)
        std::memcpy(&h[(size_t)r * out_pad], T + (size_t)r * n->out_dim, n->out_dim * 4);
    MLP_HIP(hipMemcpy(n->T, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    return RPL_OK;
}

rpl::mlp::GemmArgs fwd_args(const rpl_mlp *n, int l, const bf16 *in, int rows)
{
    rpl::mlp::GemmArgs g{};
    g.M = rows;
    g.N = n->dims[l + 1];
    g.K = n->dims[l];
    g.A = in;
    g.sam = (size_t)n->dims[l];
    g.sak = 1;
    g.B = n->Wb[l];
    g.sbk = 1;
    g.sbn = (size_t)n->dims[l];
    g.bias = n->b[l];
    return g;
}

// Training forward: keeps Z and A of every hidden layer, Y in FP32.
int forward_train(rpl_mlp *n, int rows)
{
    const int L = nlayers(n);
    const bf16 *in = n->X;
    for (int l = 0; l < L; ++l) {
        auto g = fwd_args(n, l, in, rows);
        if (l < L - 1) {
            g.Cf = n->Z[l];
            g.Cb = n->A[l];
            g.scm = (size_t)g.N;
            g.scn = 1;
            MLP_HIP(rpl::mlp::gemm(g, rpl::mlp::EPI_BIAS_GELU, n->stream));
            in = n->A[l];
        } else {
            g.Cf = n->Y;
            g.scm = (size_t)g.N;
            g.scn = 1;
            MLP_HIP(rpl::mlp::gemm(g, rpl::mlp::EPI_BIAS_F32, n->stream));
        }
    }
    return RPL_OK;
}

int backward(rpl_mlp *n, int rows, int batch, float *loss)
{
    const int L = nlayers(n), out_pad = n->dims.back();
    const int blocks = 256;
    MLP_HIP(rpl::mlp::mse(n->Y, n->T, rows, batch, out_pad, n->out_dim, n->dY, n->partials, blocks, n->stream));
    const bf16 *dZ = n->dY;
    bf16 *bufs[2] = {n->dZa, n->dZb};
    for (int l = L - 1; l >= 0; --l) {
        const int dout = n->dims[l + 1], din = n->dims[l];
        const bf16 *in = l == 0 ? n->X : n->A[l - 1];
        // gW[n_, k] = sum_m dZ[m, n_] in[m, k]
        rpl::mlp::GemmArgs g{};
        g.M = dout;
        g.N = din;
        g.K = rows;
        g.A = dZ;
        g.sam = 1;
        g.sak = (size_t)dout;
        g.B = in;
        g.sbk = (size_t)din;
        g.sbn = 1;
        g.Cf = n->gW[l];
        g.scm = (size_t)din;
        g.scn = 1;
        MLP_HIP(rpl::mlp::gemm(g, rpl::mlp::EPI_STORE_F32, n->stream));
        MLP_HIP(rpl::mlp::column_sum(dZ, rows, dout, n->gb[l], n->stream));
        if (l > 0) {
            // dZ_{l-1} = (dZ W_l) * gelu'(Z_{l-1})
            rpl::mlp::GemmArgs h{};
            h.M = rows;
            h.N = din;
            h.K = dout;
            h.A = dZ;
            h.sam = (size_t)dout;
            h.sak = 1;
            h.B = n->Wb[l];
            h.sbk = (size_t)din;
            h.sbn = 1;
            h.Cb = bufs[l & 1];
            h.scm = (size_t)din;
            h.scn = 1;
            h.Zin = n->Z[l - 1];
            MLP_HIP(rpl::mlp::gemm(h, rpl::mlp::EPI_GELU_BACKWARD, n->stream));
            dZ = bufs[l & 1];
        }
    }
    if (!loss)
        return RPL_OK;
    std::vector<float> part(blocks);
    MLP_HIP(hipStreamSynchronize(n->stream));
    MLP_HIP(hipMemcpy(part.data(), n->partials, blocks * 4, hipMemcpyDeviceToHost));
    double s = 0;
    for (float v : part)
        s += v;
    *loss = (float)(s / ((double)batch * n->out_dim));
    return RPL_OK;
}

int check_batch(rpl_mlp *n, int batch)
{
    if (!n)
        return set_err(RPL_ERR_ARG, "null network");
    if (batch < 1 || round64(batch) > n->mcap)
        return set_err(RPL_ERR_ARG, "batch must be in [1, batch_capacity]");
    MLP_HIP(hipSetDevice(n->device));
    return RPL_OK;
}

}  // namespace

namespace rpl {
namespace mlp {

hipError_t forward_rows(rpl_mlp *n, const bf16 *X, int rows, bf16 *act_a, bf16 *act_b, float *out, size_t ld,
                        unsigned char *nonfinite, hipStream_t s, _Float16 *out_h)
{
    const int L = nlayers(n);
    const bf16 *in = X;
    bf16 *acts[2] = {act_a, act_b};
    for (int l = 0; l < L - 1; ++l) {
        auto g = fwd_args(n, l, in, rows);
        g.Cb = acts[l & 1];
        g.scm = (size_t)g.N;
        g.scn = 1;
        hipError_t e = gemm(g, EPI_BIAS_GELU_INFER, s);
        if (e != hipSuccess)
            return e;
        in = acts[l & 1];
    }
    // out^T = W_L in^T + b_L: rows of the product are output units.
    GemmArgs g{};
    g.M = n->dims[L];
    g.N = rows;
    g.K = n->dims[L - 1];
    g.A = n->Wb[L - 1];
    g.sam = (size_t)g.K;
    g.sak = 1;
    g.B = in;
    g.sbk = 1;
    g.sbn = (size_t)g.K;
    g.Cf = out;
    g.Ch = out_h;
    g.scm = ld;
    g.scn = 1;
    g.bias = n->b[L - 1];
    g.nonfinite = nonfinite;
    return gemm(g, out ? EPI_BIAS_ROW_F32 : EPI_BIAS_ROW_F16, s);
}

}  // namespace mlp
}  // namespace rpl

extern "C" {

int rpl_mlp_create(rpl_mlp **out, int32_t in_dim, int32_t hidden, int32_t hidden_layers, int32_t out_dim,
                   int32_t batch_capacity, uint64_t seed)
{
    if (!out)
        return set_err(RPL_ERR_ARG, "null argument");
    *out = nullptr;
    if (in_dim < 1 || out_dim < 1 || hidden < 64 || hidden % 64 || hidden_layers < 1 || batch_capacity < 1)
        return set_err(RPL_ERR_ARG, "need in_dim, out_dim >= 1, hidden a positive multiple of 64, >= 1 hidden layer");
    rpl_mlp *n = new rpl_mlp();
    n->in_dim = in_dim;
    n->out_dim = out_dim;
    n->hidden = hidden;
    n->hidden_layers = hidden_layers;
    n->dims.push_back(round64(in_dim));
    for (int h = 0; h < hidden_layers; ++h)
        n->dims.push_back(hidden);
    n->dims.push_back(round64(out_dim));
    n->mcap = round64(batch_capacity);
    int rc = RPL_OK;
    if (hipGetDevice(&n->device) != hipSuccess || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return set_err(RPL_ERR_HIP, "no HIP device / stream");
    }
    const int L = nlayers(n);
    const int maxd = *std::max_element(n->dims.begin(), n->dims.end());
    n->W.resize(L);
    n->b.resize(L);
    n->gW.resize(L);
    n->gb.resize(L);
    n->mW.resize(L);
    n->vW.resize(L);
    n->mb.resize(L);
    n->vb.resize(L);
    n->Wb.resize(L);
    for (int l = 0; l < L; ++l) {
        const size_t nw = layer_w(n, l), nb = n->dims[l + 1];
        if ((rc = dalloc(&n->W[l], nw)) || (rc = dalloc(&n->gW[l], nw)) || (rc = dalloc(&n->mW[l], nw)) ||
            (rc = dalloc(&n->vW[l], nw)) || (rc = dalloc(&n->Wb[l], nw)) || (rc = dalloc(&n->b[l], nb)) ||
            (rc = dalloc(&n->gb[l], nb)) || (rc = dalloc(&n->mb[l], nb)) || (rc = dalloc(&n->vb[l], nb)))
            break;
        (void)hipMemset(n->mW[l], 0, nw * 4);
        (void)hipMemset(n->vW[l], 0, nw * 4);
        (void)hipMemset(n->mb[l], 0, nb * 4);
        (void)hipMemset(n->vb[l], 0, nb * 4);
    }
    n->Z.resize(hidden_layers);
    n->A.resize(hidden_layers);
    if (!rc)
        for (int h = 0; h < hidden_layers && !rc; ++h)
            if ((rc = dalloc(&n->Z[h], (size_t)n->mcap * hidden)) == RPL_OK)
                rc = dalloc(&n->A[h], (size_t)n->mcap * hidden);
    if (!rc && ((rc = dalloc(&n->X, (size_t)n->mcap * n->dims[0])) ||
                (rc = dalloc(&n->Xf, (size_t)n->mcap * n->dims[0])) ||
                (rc = dalloc(&n->Y, (size_t)n->mcap * n->dims.back())) ||
                (rc = dalloc(&n->T, (size_t)n->mcap * n->dims.back())) ||
                (rc = dalloc(&n->dY, (size_t)n->mcap * n->dims.back())) ||
                (rc = dalloc(&n->dZa, (size_t)n->mcap * maxd)) || (rc = dalloc(&n->dZb, (size_t)n->mcap * maxd)) ||
                (rc = dalloc(&n->partials, 256))))
        ;
    if (rc) {
        rpl_mlp_destroy(n);
        return rc;
    }
    // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights on the real (unpadded)
    // block, zero padding and zero biases.
    std::vector<float> p(param_count(n), 0.0f);
    uint64_t s = seed ^ 0x5DEECE66Dull;
    size_t off = 0;
    for (int l = 0; l < L; ++l) {
        const int din = n->dims[l], dout = n->dims[l + 1];
        const int real_in = l == 0 ? in_dim : hidden, real_out = l == L - 1 ? out_dim : hidden;
        const float bound = 1.0f / std::sqrt((float)real_in);
        for (int o = 0; o < real_out; ++o)
            for (int i = 0; i < real_in; ++i) {
                const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                p[off + (size_t)o * din + i] = (float)((2 * u - 1) * bound);
            }
        off += (size_t)dout * din + dout;
    }
    if ((rc = upload_params(n, p.data()))) {
        rpl_mlp_destroy(n);
        return rc;
    }
    *out = n;
    return RPL_OK;
}

int rpl_mlp_destroy(rpl_mlp *n)
{
    if (!n)
        return RPL_OK;
    (void)hipSetDevice(n->device);
    if (n->stream)
        (void)hipStreamSynchronize(n->stream);
    for (auto *v : {&n->W, &n->b, &n->gW, &n->gb, &n->mW, &n->vW, &n->mb, &n->vb, &n->Z})
        for (float *p : *v)
            if (p)
                (void)hipFree(p);
    for (auto *p : n->Wb)
        if (p)
            (void)hipFree(p);
    for (auto *p : n->A)
        if (p)
            (void)hipFree(p);
    for (void *p : {(void *)n->X, (void *)n->Xf, (void *)n->Y, (void *)n->T, (void *)n->dY, (void *)n->dZa,
                    (void *)n->dZb, (void *)n->partials})
        if (p)
            (void)hipFree(p);
    if (n->stream)
        (void)hipStreamDestroy(n->stream);
    delete n;
    return RPL_OK;
}

int rpl_mlp_dims(rpl_mlp *n, int32_t *in_pad, int32_t *out_pad, int64_t *nparams)
{
    if (!n)
        return set_err(RPL_ERR_ARG, "null network");
    if (in_pad)
        *in_pad = n->dims[0];
    if (out_pad)
        *out_pad = n->dims.back();
    if (nparams)
        *nparams = param_count(n);
    return RPL_OK;
}

int rpl_mlp_set_params(rpl_mlp *n, const float *params)
{
    if (!n || !params)
        return set_err(RPL_ERR_ARG, "null argument");
    MLP_HIP(hipSetDevice(n->device));
    return upload_params(n, params);
}

int rpl_mlp_get_params(rpl_mlp *n, float *params)
{
    if (!n || !params)
        return set_err(RPL_ERR_ARG, "null argument");
    MLP_HIP(hipSetDevice(n->device));
    MLP_HIP(hipStreamSynchronize(n->stream));
    size_t off = 0;
    for (int l = 0; l < nlayers(n); ++l) {
        const size_t nw = layer_w(n, l), nb = n->dims[l + 1];
        MLP_HIP(hipMemcpy(params + off, n->W[l], nw * 4, hipMemcpyDeviceToHost));
        off += nw;
        MLP_HIP(hipMemcpy(params + off, n->b[l], nb * 4, hipMemcpyDeviceToHost));
        off += nb;
    }
    return RPL_OK;
}

int rpl_mlp_forward(rpl_mlp *n, int32_t batch, const float *X, float *Y)
{
    int rc = check_batch(n, batch);
    if (rc)
        return rc;
    if (!X || !Y)
        return set_err(RPL_ERR_ARG, "null argument");
    // The inference path of the RK3 stage (forward_rows): activations
    // straight to bf16, the output transposed ([out_pad][rows]).
    const int rows = round64(batch), out_pad = n->dims.back();
    if ((rc = load_inputs(n, batch, X)))
        return rc;
    bf16 *act_b = n->A[std::min(1, n->hidden_layers - 1)];
    MLP_HIP(rpl::mlp::forward_rows(n, n->X, rows, n->A[0], act_b, n->Y, (size_t)rows, nullptr, n->stream));
    MLP_HIP(hipStreamSynchronize(n->stream));
    std::vector<float> h((size_t)rows * out_pad);
    MLP_HIP(hipMemcpy(h.data(), n->Y, h.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < batch; ++r)
        for (int o = 0; o < n->out_dim; ++o)
            Y[(size_t)r * n->out_dim + o] = h[(size_t)o * rows + r];
    return RPL_OK;
}

int rpl_mlp_gradients(rpl_mlp *n, int32_t batch, const float *X, const float *T, float *loss, float *grads)
{
    int rc = check_batch(n, batch);
    if (rc)
        return rc;
    if (!X || !T)
        return set_err(RPL_ERR_ARG, "null argument");
    const int rows = round64(batch);
    if ((rc = load_inputs(n, batch, X)) || (rc = load_targets(n, batch, T)) || (rc = forward_train(n, rows)) ||
        (rc = backward(n, rows, batch, loss)))
        return rc;
    MLP_HIP(hipStreamSynchronize(n->stream));  // the stream does not order with the copies below
    if (grads) {
        size_t off = 0;
        for (int l = 0; l < nlayers(n); ++l) {
            const size_t nw = layer_w(n, l), nb = n->dims[l + 1];
            MLP_HIP(hipMemcpy(grads + off, n->gW[l], nw * 4, hipMemcpyDeviceToHost));
            off += nw;
            MLP_HIP(hipMemcpy(grads + off, n->gb[l], nb * 4, hipMemcpyDeviceToHost));
            off += nb;
        }
    }
    return RPL_OK;
}

int rpl_mlp_train_device(rpl_mlp *n, int64_t rows, const float *X, const float *T, int32_t batch, int32_t steps,
                         float lr0, float lr1, uint64_t seed, float *losses)
{
    int rc = check_batch(n, batch);
    if (rc)
        return rc;
    if (!X || !T || rows < 1 || steps < 0 || !(lr0 > 0) || !(lr1 > 0))
        return set_err(RPL_ERR_ARG, "need device rows, rows >= 1, steps >= 0 and positive learning rates");
    hipPointerAttribute_t ax, at;
    if (hipPointerGetAttributes(&ax, X) != hipSuccess || hipPointerGetAttributes(&at, T) != hipSuccess ||
        ax.type != hipMemoryTypeDevice || at.type != hipMemoryTypeDevice || ax.device != n->device ||
        at.device != n->device) {
        (void)hipGetLastError();
        return set_err(RPL_ERR_ARG, "training rows must be device memory on the network's device");
    }
    const int rp = round64(batch);
    for (int k = 0; k < steps; ++k) {
        const float lr = lr0 * (float)std::pow((double)lr1 / lr0, steps > 1 ? (double)k / (steps - 1) : 0.0);
        MLP_HIP(rpl::mlp::gather_rows(X, T, (unsigned long long)rows, n->in_dim, n->dims[0], n->out_dim,
                                      n->dims.back(), batch, rp, seed, (unsigned long long)n->step, n->X, n->T,
                                      n->stream));
        if ((rc = forward_train(n, rp)) || (rc = backward(n, rp, batch, losses ? losses + k : nullptr)))
            return rc;
        n->step += 1;
        for (int l = 0; l < nlayers(n); ++l) {
            MLP_HIP(rpl::mlp::adam(n->W[l], n->gW[l], n->mW[l], n->vW[l], n->Wb[l], layer_w(n, l), lr, 0.9f, 0.999f,
                                   1e-8f, n->step, n->stream));
            MLP_HIP(rpl::mlp::adam(n->b[l], n->gb[l], n->mb[l], n->vb[l], nullptr, n->dims[l + 1], lr, 0.9f, 0.999f,
                                   1e-8f, n->step, n->stream));
        }
    }
    MLP_HIP(hipStreamSynchronize(n->stream));
    return RPL_OK;
}

int rpl_mlp_train_step(rpl_mlp *n, int32_t batch, const float *X, const float *T, float lr, float *loss)
{
    int rc = check_batch(n, batch);
    if (rc)
        return rc;
    if (!X || !T || !(lr > 0))
        return set_err(RPL_ERR_ARG, "null argument or non-positive learning rate");
    const int rows = round64(batch);
    if ((rc = load_inputs(n, batch, X)) || (rc = load_targets(n, batch, T)) || (rc = forward_train(n, rows)) ||
        (rc = backward(n, rows, batch, loss)))
        return rc;
    n->step += 1;
    for (int l = 0; l < nlayers(n); ++l) {
        MLP_HIP(rpl::mlp::adam(n->W[l], n->gW[l], n->mW[l], n->vW[l], n->Wb[l], layer_w(n, l), lr, 0.9f, 0.999f, 1e-8f,
                               n->step, n->stream));
        MLP_HIP(rpl::mlp::adam(n->b[l], n->gb[l], n->mb[l], n->vb[l], nullptr, n->dims[l + 1], lr, 0.9f, 0.999f,
                               1e-8f, n->step, n->stream));
    }
    MLP_HIP(hipStreamSynchronize(n->stream));
    return RPL_OK;
}

}  // extern "C"
