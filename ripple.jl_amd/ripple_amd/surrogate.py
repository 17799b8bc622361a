"""Learned per-column source surrogate (no counterpart in the reference).

A small MLP stands in for the per-column evaluation of the PrecomputedSources
bundle (Physics/Packages/precomputed_sources.jl:240-261): it maps a column's
log-scaled action spectrum, the 10 m wind vector and the depth to per-bin
growth rates r, with source = r * rate_unit * sigma_bin * N.  It is trained
on the GPU by the library's own BF16 matrix-core forward/backward kernels
(FP32 master weights, Adam) on (features, rate) pairs that the plain bundle
produces along the model's own runs; the plain bundle stays installed as the
fallback outside the training envelope and on failed audit columns.
"""
import ctypes
from dataclasses import dataclass, fields

import numpy as np

from . import _lib


class SourceNetwork:
    """Device-resident MLP in_dim -> hidden x hidden_layers -> out_dim (GELU).

    Parameters are the concatenation over layers of W_l (out_pad_l x in_pad_l,
    row-major, zero padding) and b_l (out_pad_l), FP32.
    """

    def __init__(self, in_dim, out_dim, hidden=512, hidden_layers=3, batch_capacity=4096, seed=0):
        lib = _lib.load()
        p = ctypes.c_void_p()
        _lib.check(lib.rpl_mlp_create(ctypes.byref(p), in_dim, hidden, hidden_layers, out_dim, batch_capacity,
                                      ctypes.c_uint64(seed)))
        self._lib, self.ptr = lib, p
        self.in_dim, self.out_dim, self.hidden, self.hidden_layers = in_dim, out_dim, hidden, hidden_layers
        self.batch_capacity = batch_capacity
        ip, op, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _lib.check(lib.rpl_mlp_dims(p, ctypes.byref(ip), ctypes.byref(op), ctypes.byref(n)))
        self.in_pad, self.out_pad, self.nparams = ip.value, op.value, n.value
        self.dims = [self.in_pad] + [hidden] * hidden_layers + [self.out_pad]

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self._lib.rpl_mlp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _lib.check(self._lib.rpl_mlp_get_params(self.ptr, _lib.fptr(out)))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {flat.size}")
        _lib.check(self._lib.rpl_mlp_set_params(self.ptr, _lib.fptr(flat)))

    def layers(self, flat=None):
        """[(W, b)] views of a flat parameter vector (padded shapes)."""
        flat = self.params() if flat is None else flat
        out, off = [], 0
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            W = flat[off:off + din * dout].reshape(dout, din)
            off += din * dout
            b = flat[off:off + dout]
            off += dout
            out.append((W, b))
        return out

    def _rows(self, X, width):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != width:
            raise ValueError(f"expected rows of {width} values, got shape {X.shape}")
        if not 1 <= X.shape[0] <= self.batch_capacity:
            raise ValueError(f"batch must be in [1, {self.batch_capacity}]")
        return X

    def forward(self, X):
        X = self._rows(X, self.in_dim)
        Y = np.empty((X.shape[0], self.out_dim), dtype=np.float32)
        _lib.check(self._lib.rpl_mlp_forward(self.ptr, X.shape[0], _lib.fptr(X), _lib.fptr(Y)))
        return Y

    def gradients(self, X, T):
        """(loss, flat gradient) of mean((Y - T)^2) over the batch."""
        X = self._rows(X, self.in_dim)
        T = self._rows(T, self.out_dim)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        _lib.check(self._lib.rpl_mlp_gradients(self.ptr, X.shape[0], _lib.fptr(X), _lib.fptr(T), _lib.fptr(loss),
                                               _lib.fptr(g)))
        return float(loss[0]), g

    def train_step(self, X, T, lr):
        X = self._rows(X, self.in_dim)
        T = self._rows(T, self.out_dim)
        loss = np.zeros(1, dtype=np.float32)
        _lib.check(self._lib.rpl_mlp_train_step(self.ptr, X.shape[0], _lib.fptr(X), _lib.fptr(T), float(lr),
                                                _lib.fptr(loss)))
        return float(loss[0])


@dataclass
class SurrogateConfig:
    """Feature scaling, output unit, training envelope and audit policy."""
    floor: float = 1e-12          # x = (log(N + floor) - shift) * scale
    shift: float = 0.0
    scale: float = 0.1
    wind_scale: float = 1.0 / 20.0
    depth_feature: float = 1.0
    rate_unit: float = 1e-5       # source = r * rate_unit * sigma * N
    wind_lo: float = 0.0          # envelope: U10 in [wind_lo, wind_hi]
    wind_hi: float = 1e9
    action_hi: float = 3.0e38     # envelope: every bin's action <= action_hi
    audit_stride: int = 0         # every audit_stride-th column is audited (0: none)
    audit_bound: float = 0.1      # relative L2 source error above which an audit column falls back
    terms: int = 0                # 0: the network supplies every source term; 1: the wind input only
    refresh: int = 0              # 0: rates inferred every RK3 stage; 1: once per step, held through its stages

    def to_struct(self):
        s = _lib.rpl_surrogate()
        for f in fields(self):
            setattr(s, f.name, getattr(self, f.name))
        return s


def training_pairs(model, cfg):
    """Features (Nx*Ny, nbins + 3) and plain-bundle rate targets (Nx*Ny, nbins)
    of the model's current action and clock; rows are columns, i fastest."""
    Nx, Ny, Nk, Np = model.size
    nb = Nk * Np
    X = np.empty((Nx * Ny, nb + 3), dtype=np.float32)
    T = np.empty((Nx * Ny, nb), dtype=np.float32)
    s = cfg.to_struct()
    _lib.check(_lib.load().rpl_surrogate_training_pairs(model._ctx.ptr, ctypes.byref(s), _lib.fptr(X),
                                                        _lib.fptr(T)))
    return X, T


def attach_surrogate(model, net, cfg, bin_action_hi=None):
    """Let the model's time step take its sources from `net` (None detaches).

    bin_action_hi: optional per-bin largest training action (nkappa*nphi,
    m fastest); a column with any bin above it uses the plain bundle."""
    lib = _lib.load()
    if net is None:
        _lib.check(lib.rpl_attach_surrogate(model._ctx.ptr, None, None, None))
        model._surrogate = None
        return
    s = cfg.to_struct()
    hi = None
    if bin_action_hi is not None:
        hi = np.ascontiguousarray(bin_action_hi, dtype=np.float32).ravel()
        Nx, Ny, Nk, Np = model.size
        if hi.size != Nk * Np:
            raise ValueError(f"bin_action_hi needs {Nk * Np} values, got {hi.size}")
    _lib.check(lib.rpl_attach_surrogate(model._ctx.ptr, net.ptr, ctypes.byref(s), _lib.fptr(hi)))
    model._surrogate = (net, cfg)  # keeps the network alive while attached


def surrogate_stats(model):
    """Counters since the last call: columns evaluated, plain-fallback columns,
    audited columns, failed audits, largest audited relative error."""
    v = [ctypes.c_int64() for _ in range(4)]
    e = ctypes.c_double()
    _lib.check(_lib.load().rpl_surrogate_stats(model._ctx.ptr, *[ctypes.byref(x) for x in v], ctypes.byref(e)))
    return {"columns": v[0].value, "plain": v[1].value, "audited": v[2].value, "failed": v[3].value,
            "err_max": e.value}


def train(net, X, T, *, steps, batch=2048, lr=1e-3, seed=0, lr_final=None, log_every=0):
    """Adam on random mini-batches of (X, T) rows; returns the loss history.

    The learning rate decays geometrically from lr to lr_final (default lr/10).
    """
    rng = np.random.default_rng(seed)
    batch = min(batch, net.batch_capacity, X.shape[0])
    lr_final = lr / 10 if lr_final is None else lr_final
    losses = []
    for k in range(steps):
        idx = rng.integers(0, X.shape[0], size=batch)
        rate = lr * (lr_final / lr) ** (k / max(steps - 1, 1))
        losses.append(net.train_step(X[idx], T[idx], rate))
        if log_every and (k % log_every == 0 or k == steps - 1):
            print(f"step {k}: loss {losses[-1]:.4e}", flush=True)
    return np.array(losses)


def collect_pairs(model, dt, steps, every, cfg, columns=None, seed=0, return_bin_max=False):
    """Advance `model` with its plain sources for `steps` steps, taking
    training pairs every `every` steps (and at the start); `columns` caps the
    number of columns kept per snapshot (random subset).  Returns (X, T), and
    with return_bin_max also the per-bin largest action over the snapshots
    (nkappa*nphi, m fastest), for attach_surrogate's bin_action_hi."""
    from .model import time_step_
    rng = np.random.default_rng(seed)
    Xs, Ts = [], []
    Nx, Ny, Nk, Np = model.size
    bin_max = np.zeros(Nk * Np)
    for k in range(steps + 1):
        if k % every == 0:
            X, T = training_pairs(model, cfg)
            if return_bin_max:
                a = np.asarray(model.action, dtype=np.float64).reshape(Nx * Ny, Nk * Np, order="F")
                bin_max = np.maximum(bin_max, a.max(axis=0))
            if columns is not None and columns < X.shape[0]:
                idx = rng.choice(X.shape[0], size=columns, replace=False)
                X, T = X[idx], T[idx]
            Xs.append(X)
            Ts.append(T)
        if k < steps:
            time_step_(model, dt)
    if return_bin_max:
        return np.concatenate(Xs), np.concatenate(Ts), bin_max
    return np.concatenate(Xs), np.concatenate(Ts)


def envelope_from_pairs(X, cfg, margin=0.02):
    """Set cfg's training envelope (wind range, largest action) from the
    feature rows the network was trained on; returns cfg.  The features are
    bf16, so the action bound is recovered to within bf16 rounding of the
    log (see collect_pairs' return_bin_max for exact per-bin maxima)."""
    nb = X.shape[1] - 3
    x = X[:, :nb].astype(np.float64)
    amax = float(np.max(np.exp(np.max(x) / cfg.scale + cfg.shift) - cfg.floor))
    u = np.hypot(X[:, nb].astype(np.float64), X[:, nb + 1].astype(np.float64)) / cfg.wind_scale
    cfg.wind_lo = float(max(0.0, u.min() * (1 - margin)))
    cfg.wind_hi = float(u.max() * (1 + margin))
    cfg.action_hi = amax * (1 + margin)
    return cfg


def _device_rows(A, width, name, rows):
    """(device address, rows) of a training array: a contiguous FP32 torch GPU
    tensor of rows x width, or a raw device address (int or c_void_p) with
    `rows` given."""
    if isinstance(A, (int, ctypes.c_void_p)):
        if rows is None or rows < 1:
            raise ValueError(f"{name} given as a device address needs rows >= 1")
        return ctypes.c_void_p(A.value if isinstance(A, ctypes.c_void_p) else A), int(rows)
    import torch  # device memory for training sets
    if not isinstance(A, torch.Tensor) or not A.is_cuda or A.dtype != torch.float32 or A.dim() != 2 \
            or A.shape[1] != width or not A.is_contiguous():
        raise ValueError(f"{name} must be a contiguous FP32 GPU tensor of rows x {width}")
    torch.cuda.synchronize(A.device)  # the rows are written on torch's stream
    return ctypes.c_void_p(A.data_ptr()), int(A.shape[0])


def train_device(net, X, T, *, steps, batch=4096, lr=1e-3, lr_final=None, seed=0, log_every=0, rows=None):
    """Adam on mini-batches drawn on the device (rpl_mlp_train_device) from a
    training set held in GPU memory: X (rows x in_dim) and T (rows x out_dim),
    FP32 row-major, either contiguous torch GPU tensors or device addresses
    with `rows` given.  The learning rate decays geometrically from lr to
    lr_final (default lr/10) over all steps.  Returns the loss history."""
    xp, xr = _device_rows(X, net.in_dim, "X", rows)
    tp, tr = _device_rows(T, net.out_dim, "T", rows)
    if xr != tr:
        raise ValueError("X and T need the same number of rows")
    batch = min(batch, net.batch_capacity)
    lr_final = lr / 10 if lr_final is None else lr_final
    losses = np.zeros(steps, dtype=np.float32)
    chunk = log_every if log_every else max(steps, 1)
    span = max(steps - 1, 1)
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        lr0 = lr * (lr_final / lr) ** (done / span)
        lr1 = lr * (lr_final / lr) ** ((done + n - 1) / span)
        part = losses[done:done + n]
        _lib.check(net._lib.rpl_mlp_train_device(net.ptr, xr, xp, tp, batch, n, lr0, lr1, ctypes.c_uint64(seed),
                                                 _lib.fptr(part)))
        done += n
        if log_every:
            print(f"step {done}: loss {float(np.mean(part)):.4e} (mean of the last {n}), lr {lr1:.2e}", flush=True)
    return losses


def _standardisation(T):
    T = np.asarray(T, dtype=np.float64)
    mt, st = T.mean(axis=0), T.std(axis=0)
    return mt, np.maximum(st, 1e-3 * max(float(np.sqrt(np.mean(T * T))), 1e-30))


def _fold_output_affine(net, mt, st):
    """Last layer <- st * (last layer) + mt, so the network emits unstandardised rates."""
    flat = net.params().astype(np.float64)
    layers = net.layers(flat)
    WL, bL = layers[-1]
    n_out = net.out_dim
    WL[:n_out, :] *= st[:, None]
    bL[:n_out] = st * bL[:n_out] + mt
    net.set_params(flat.astype(np.float32))


def fit_device(net, X, T, *, steps, batch=4096, lr=1e-3, lr_final=None, seed=0, log_every=0, chunk=1 << 18):
    """fit() on a GPU-resident training set (torch tensors as train_device
    takes them).  T is standardised IN PLACE (per output bin, mean and std
    accumulated in FP64 over row chunks) and the affine map is folded into
    the last layer afterwards.  Returns the loss history (standardised units)."""
    import torch
    _device_rows(T, net.out_dim, "T", None)
    rows = T.shape[0]
    s1 = torch.zeros(net.out_dim, dtype=torch.float64, device=T.device)
    s2 = torch.zeros_like(s1)
    for a in range(0, rows, chunk):
        c = T[a:a + chunk].double()
        s1 += c.sum(0)
        s2 += (c * c).sum(0)
    mt = s1 / rows
    var = torch.clamp(s2 / rows - mt * mt, min=0.0)
    rms = float(torch.sqrt(torch.sum(s2) / (rows * net.out_dim)))
    st = torch.clamp(torch.sqrt(var), min=1e-3 * max(rms, 1e-30))
    for a in range(0, rows, chunk):
        T[a:a + chunk].sub_(mt.float()).div_(st.float())
    losses = train_device(net, X, T, steps=steps, batch=batch, lr=lr, lr_final=lr_final, seed=seed,
                          log_every=log_every)
    _fold_output_affine(net, mt.cpu().numpy(), st.cpu().numpy())
    return losses


def _bf16_bits(a):
    """Round-to-nearest-even bf16 bit patterns of FP32 values (NaN kept NaN)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def save_network(path, net, cfg, bin_action_hi=None, **meta):
    """Write a trained network for inference: weights as bf16 (the precision
    the inference GEMMs read them in, so nothing is lost), biases FP32, the
    surrogate configuration and optional per-bin action bounds, plus free-form
    metadata (JSON-serialisable), to an .npz file."""
    import json
    arrays = {}
    for l, (W, b) in enumerate(net.layers()):
        arrays[f"W{l}"] = _bf16_bits(W)
        arrays[f"b{l}"] = np.asarray(b, dtype=np.float32)
    head = {"in_dim": net.in_dim, "out_dim": net.out_dim, "hidden": net.hidden,
            "hidden_layers": net.hidden_layers, "config": {f.name: getattr(cfg, f.name) for f in fields(cfg)},
            "meta": meta}
    arrays["header"] = np.frombuffer(json.dumps(head).encode(), dtype=np.uint8)
    if bin_action_hi is not None:
        arrays["bin_action_hi"] = np.asarray(bin_action_hi, dtype=np.float32)
    np.savez_compressed(path, **arrays)


def load_network(path, batch_capacity=64):
    """(network, SurrogateConfig, bin_action_hi or None, metadata) from save_network's file."""
    import json
    with np.load(path, allow_pickle=False) as z:
        head = json.loads(bytes(z["header"]).decode())
        net = SourceNetwork(head["in_dim"], head["out_dim"], hidden=head["hidden"],
                            hidden_layers=head["hidden_layers"], batch_capacity=batch_capacity)
        parts = []
        for l in range(len(net.dims) - 1):
            W = (z[f"W{l}"].astype(np.uint32) << 16).view(np.float32)
            if W.shape != (net.dims[l + 1], net.dims[l]):
                net.close()
                raise ValueError(f"layer {l} has shape {W.shape}, expected {(net.dims[l + 1], net.dims[l])}")
            parts += [W.ravel(), z[f"b{l}"].astype(np.float32)]
        net.set_params(np.concatenate(parts))
        hi = z["bin_action_hi"].astype(np.float64) if "bin_action_hi" in z.files else None
    return net, SurrogateConfig(**head["config"]), hi, head["meta"]


def fit(net, X, T, *, steps, batch=2048, lr=1e-3, lr_final=None, seed=0, log_every=0):
    """Train `net` to map feature rows X to target rows T (both as returned by
    `training_pairs`), with every output bin standardised over the training
    set, then fold that affine map into the last layer (FP32 bias, the scale
    multiplies the weights) so that the network emits rates in the
    configuration's rate_unit.  Inputs are not standardised per feature:
    folding 1/std into bf16 first-layer weights would cancel catastrophically
    against the folded mean; the feature scaling of the configuration keeps
    them O(1).  Returns the loss history (in standardised units)."""
    mt, st = _standardisation(T)
    Tn = ((np.asarray(T, dtype=np.float64) - mt) / st).astype(np.float32)
    losses = train(net, np.asarray(X, dtype=np.float32), Tn, steps=steps, batch=batch, lr=lr, lr_final=lr_final,
                   seed=seed, log_every=log_every)
    _fold_output_affine(net, mt, st)
    return losses
