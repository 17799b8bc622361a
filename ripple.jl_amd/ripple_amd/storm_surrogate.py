"""The translating hurricane's source surrogate (no counterpart in the
reference): collection of (column, plain-bundle rate) pairs from randomised
storms and uniform-wind seas on the hurricane's 32 kappa x 36 phi grid, GPU
training of the 1155 -> 512 x 3 -> 1152 network on them, and the trained
network shipped with the package (NETWORK_FILE).

    python -m ripple_amd.storm_surrogate --out <file.npz>

retrains it (about 2 M pairs; a few minutes on one MI355X).  The pairs are
the PrecomputedSources(PressureCorrelationInput, LocalSaturationDissipation)
rates of precomputed_sources.jl:240-261 along the model's own runs, so the
network stands in for the whole bundle (terms = 0).
"""
import argparse
import math
import os
import sys
import time

import numpy as np

from . import surrogate as sg
from . import workloads

NETWORK_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "storm_surrogate_k32_p36.npz")


def _model(wl):
    from . import grids
    from .model import SpectralWaveModel, set_
    s = wl.spectral_grid
    spec = grids.PolarWaveVectorGrid(np.float32, kappa=s.kappa, phi=s.phi, kappa_faces=s.kappa_faces,
                                     phi_faces=s.phi_faces)
    m = SpectralWaveModel(wl.grid, spec, velocities=None, **wl.model_kwargs)
    set_(m, N=wl.N0)
    return m


def collect(storms=8, seas=16, every=8, cfg=None, log=print):
    """Plain-bundle pairs of `storms` hurricane_training_storm runs (128^2,
    120 steps of 60 s) and `seas` uniform_wind_sea runs (64^2, 60 steps),
    every `every` steps from the start, every column.  Returns GPU tensors X
    (rows x 1155) and T (rows x 1152), the per-bin largest action seen, and a
    list of the runs."""
    import torch
    from .model import time_step_
    cfg = cfg or sg.SurrogateConfig(rate_unit=1.0, scale=0.05, floor=1e-24)
    Xs, Ts, runs = [], [], []
    bin_max = None
    wls = [("storm", k) for k in range(storms)] + [("sea", k) for k in range(seas)]
    for kind, k in wls:
        wl = workloads.hurricane_training_storm(k) if kind == "storm" else workloads.uniform_wind_sea(k)
        m = _model(wl)
        Nx, Ny, Nk, Np = m.size
        if bin_max is None:
            bin_max = np.zeros(Nk * Np)
        t0, rows = time.perf_counter(), 0
        for step in range(wl.nsteps + 1):
            if step % every == 0:
                X, T = sg.training_pairs(m, cfg)
                a = np.asarray(m.action, dtype=np.float64).reshape(Nx * Ny, Nk * Np, order="F")
                bin_max = np.maximum(bin_max, a.max(axis=0))
                Xs.append(torch.from_numpy(X).cuda())
                Ts.append(torch.from_numpy(T).cuda())
                rows += X.shape[0]
            if step < wl.nsteps:
                time_step_(m, wl.dt)
        del m
        runs.append(wl.name)
        log(f"{wl.name}: {rows} pairs in {time.perf_counter() - t0:.1f} s")
    X = torch.cat(Xs)
    del Xs
    T = torch.cat(Ts)
    del Ts
    return X, T, bin_max, runs


def train(out, storms=8, seas=16, every=8, steps=40000, batch=4096, lr=1e-3, lr_final=1e-5, seed=7, terms=0,
          log=print):
    """Collect, train and save the network (save_network) to `out`; returns
    the network, its configuration and the training summary.  terms 0: the
    network supplies every source term; 1: the wind input only (the
    dissipation stays the plain bundle's)."""
    import torch
    cfg = sg.SurrogateConfig(rate_unit=1.0, scale=0.05, floor=1e-24, terms=terms)
    X, T, bin_max, runs = collect(storms, seas, every, cfg, log)
    nb = T.shape[1]
    # rate unit: the rms of the non-zero targets, so the network's outputs are O(1)
    ss, cnt = 0.0, 0
    for a in range(0, T.shape[0], 1 << 16):  # row chunks: one mask over the whole set exceeds 2^31 elements
        c = T[a:a + (1 << 16)]
        ss += float(torch.sum(c.double() ** 2))
        cnt += int(torch.count_nonzero(c))
    unit = math.sqrt(ss / max(cnt, 1))
    T.div_(unit)
    cfg.rate_unit = unit
    u = torch.hypot(X[:, nb], X[:, nb + 1]).double() / cfg.wind_scale
    cfg.wind_lo = 0.0
    cfg.wind_hi = float(u.max()) * 1.02
    del u
    bin_hi = bin_max * 1.05
    cfg.action_hi = float(bin_hi.max())
    # hold out every 64th row to report the fit on rows the network never saw
    rows = X.shape[0]
    hold = torch.arange(0, rows, 64, device=X.device)
    keep = torch.ones(rows, dtype=torch.bool, device=X.device)
    keep[hold] = False
    Xh, Th = X[hold].cpu().numpy(), T[hold].cpu().numpy()
    X, T = X[keep].contiguous(), T[keep].contiguous()
    del keep
    log(f"{X.shape[0]} training pairs, {Xh.shape[0]} held out; rate unit {unit:.4e}; "
        f"U10 <= {cfg.wind_hi:.1f} m/s; action <= {cfg.action_hi:.3e}")
    net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=batch, seed=seed)
    t0 = time.perf_counter()
    losses = sg.fit_device(net, X, T, steps=steps, batch=batch, lr=lr, lr_final=lr_final, seed=seed,
                           log_every=max(steps // 40, 1))
    t_train = time.perf_counter() - t0
    del X, T
    torch.cuda.empty_cache()
    Y = np.concatenate([net.forward(Xh[a:a + batch]) for a in range(0, Xh.shape[0], batch)])
    err = float(np.linalg.norm(Y.astype(np.float64) - Th) / np.linalg.norm(Th.astype(np.float64)))
    summary = {"pairs": int(rows), "held_out": int(Xh.shape[0]), "runs": runs, "terms": terms, "steps": steps,
               "batch": batch,
               "lr": lr, "lr_final": lr_final, "seed": seed, "train_s": t_train,
               "loss_first": float(losses[0]), "loss_last_1000": float(np.mean(losses[-1000:])),
               "held_out_rate_rel_l2": err}
    log(f"trained in {t_train:.1f} s; held-out rate rel-L2 {err:.4e}")
    sg.save_network(out, net, cfg, bin_action_hi=bin_hi, **summary)
    return net, cfg, summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=NETWORK_FILE)
    ap.add_argument("--storms", type=int, default=8)
    ap.add_argument("--seas", type=int, default=16)
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--terms", type=int, default=0, choices=(0, 1))
    a = ap.parse_args(argv)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    train(a.out, a.storms, a.seas, a.every, a.steps, a.batch, a.lr, a.lr_final, a.seed, a.terms,
          log=lambda s: print(s, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
