"""Learned source surrogate on the device: the BF16 matrix-core forward and
backward passes against the CPU oracle (oracle/mlp_oracle.py, itself pinned
by finite differences), the Adam step, training, the plain-bundle fallback
paths, and what a trained network gains over running without sources."""
import math
import sys

import numpy as np
import pytest

import mlp_oracle as mo
import ripple_amd as ra
from ripple_amd import grids, workloads
from ripple_amd import surrogate as sg
from hip_buffers import HipBuffers

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _net_and_batch(seed=0, in_dim=100, out_dim=70, hidden=64, layers=2, batch=100):
    net = sg.SourceNetwork(in_dim, out_dim, hidden=hidden, hidden_layers=layers, batch_capacity=max(256, batch),
                           seed=seed)
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(batch, in_dim)).astype(np.float32)
    T = (0.3 * rng.normal(size=(batch, out_dim))).astype(np.float32)
    return net, X, T


def _oracle_inputs(net, X, T):
    layers = mo.unpack(net.params(), net.dims)
    rows = (X.shape[0] + 63) // 64 * 64
    return layers, mo.pad_rows(X, net.in_pad, rows), mo.pad_rows(T, net.out_pad, rows)


def test_forward_matches_oracle():
    say("mlp forward vs oracle")
    net, X, T = _net_and_batch()
    layers, Xp, _ = _oracle_inputs(net, X, T)
    Y = net.forward(X)
    Ye, _ = mo.forward(layers, Xp, emulate=True)
    Yx, _ = mo.forward(layers, Xp)
    e_emu = rel_l2(Y, Ye[:X.shape[0], :net.out_dim])
    e_exact = rel_l2(Y, Yx[:X.shape[0], :net.out_dim])
    say(f"forward rel-L2 vs bf16-emulating oracle {e_emu:.2e}, vs FP64 oracle {e_exact:.2e}")
    assert e_emu < 2e-3
    assert e_exact < 3e-2
    net.close()


def test_gradients_match_oracle():
    say("mlp gradients vs oracle")
    net, X, T = _net_and_batch(1)
    layers, Xp, Tp = _oracle_inputs(net, X, T)
    loss, g = net.gradients(X, T)
    Le, ge = mo.backward(layers, Xp, Tp, X.shape[0], net.out_dim, emulate=True)
    Lx, gx = mo.backward(layers, Xp, Tp, X.shape[0], net.out_dim)
    ge_flat = mo.pack(ge)
    gx_flat = mo.pack(gx)
    say(f"loss {loss:.6e} oracle {Le:.6e}; grad rel-L2 emu {rel_l2(g, ge_flat):.2e} exact {rel_l2(g, gx_flat):.2e}")
    assert abs(loss - Le) / Le < 1e-4
    assert rel_l2(g, ge_flat) < 5e-3
    assert rel_l2(g, gx_flat) < 5e-2
    # per-layer, so that a wrong layer cannot hide behind a large one
    off = 0
    for (W, b), (gW, gb) in zip(layers, ge):
        nW = W.size
        assert rel_l2(g[off:off + nW], gW.ravel()) < 1e-2
        off += nW
        assert rel_l2(g[off:off + b.size], gb) < 1e-2
        off += b.size
    net.close()


def test_large_tile_gemm_forward_and_gradients_match_oracle():
    """Dims that are multiples of 128 take the 128 x 128 MFMA 16x16x32 GEMM
    (forward, weight and input gradients, all operand orientations)."""
    say("mlp large-tile gemm vs oracle")
    net, X, T = _net_and_batch(3, in_dim=200, out_dim=250, hidden=256, layers=2, batch=256)
    assert net.dims == [256, 256, 256, 256]
    layers, Xp, Tp = _oracle_inputs(net, X, T)
    Y = net.forward(X)
    Ye, _ = mo.forward(layers, Xp, emulate=True)
    e_f = rel_l2(Y, Ye[:X.shape[0], :net.out_dim])
    loss, g = net.gradients(X, T)
    Le, ge = mo.backward(layers, Xp, Tp, X.shape[0], net.out_dim, emulate=True)
    e_g = rel_l2(g, mo.pack(ge))
    say(f"large-tile: forward rel-L2 {e_f:.2e}, gradient rel-L2 {e_g:.2e}")
    assert e_f < 2e-3 and e_g < 5e-3 and abs(loss - Le) / Le < 1e-4
    off = 0
    for (W, b), (gW, gb) in zip(layers, ge):
        assert rel_l2(g[off:off + W.size], gW.ravel()) < 1e-2
        off += W.size + b.size
    net.close()


def test_inference_kernels_match_oracle():
    """rpl_mlp_forward runs the kernels of the RK3 stage's inference: 256 rows
    take the 128 x 128 GEMM in every layer (the output layer transposed, 256
    units x 256 rows), 192 rows the 64 x 64 one where 128 does not divide.
    Both against the oracle, and the rows they share agree to FP32 rounding."""
    say("mlp inference kernels vs oracle")
    net, X, T = _net_and_batch(8, in_dim=200, out_dim=250, hidden=512, layers=3, batch=256)
    assert net.dims == [256, 512, 512, 512, 256]
    layers, Xp, _ = _oracle_inputs(net, X, T)
    Ye, _ = mo.forward(layers, Xp, emulate=True)
    Y = net.forward(X)
    Y192 = net.forward(X[:192])
    e = rel_l2(Y, Ye[:256, :net.out_dim])
    e192 = rel_l2(Y192, Ye[:192, :net.out_dim])
    d = rel_l2(Y[:192], Y192)
    say(f"forward rel-L2 vs oracle: 256 rows {e:.2e}, 192 rows {e192:.2e}; between kernels {d:.2e}")
    assert e < 2e-3 and e192 < 2e-3 and d < 1e-3
    net.close()


def test_adam_step_matches_oracle():
    say("mlp adam step")
    net, X, T = _net_and_batch(2)
    p0 = net.params()
    _, g = net.gradients(X, T)
    lr = 1e-3
    net.train_step(X, T, lr)
    p1 = net.params()
    ref, _, _ = mo.adam(p0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, lr)
    assert np.max(np.abs(p1 - ref)) < 1e-3 * lr
    # padded weights stay zero
    layers = net.layers(p1)
    W0, _ = layers[0]
    assert not np.any(W0[:, net.in_dim:])
    assert not np.any(layers[-1][0][net.out_dim:, :])
    net.close()


def test_training_reduces_loss():
    say("mlp training")
    rng = np.random.default_rng(5)
    net = sg.SourceNetwork(48, 40, hidden=128, hidden_layers=2, batch_capacity=512, seed=3)
    X = rng.uniform(-1, 1, size=(4096, 48)).astype(np.float32)
    A = rng.normal(size=(48, 40)) / 7
    T = np.tanh(X @ A).astype(np.float32) + 0.5 * np.sin(X[:, :40] * 2).astype(np.float32)
    l0, _ = net.gradients(X[:512], T[:512])
    losses = sg.train(net, X, T, steps=400, batch=512, lr=2e-3)
    l1, _ = net.gradients(X[:512], T[:512])
    say(f"training loss {l0:.3e} -> {l1:.3e}")
    assert l1 < l0 / 10
    assert np.all(np.isfinite(losses))
    net.close()


def _drawn_rows(seed, step, batch, rows):
    """The rows rpl_mlp_train_device draws for one step: splitmix64 of
    seed + step * 2^32 + r + golden-ratio increment, modulo rows."""
    m64 = (1 << 64) - 1
    out = []
    for r in range(batch):
        z = (seed + (step << 32) + r + 0x9E3779B97F4A7C15) & m64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m64
        out.append((z ^ (z >> 31)) % rows)
    return np.array(out)


def test_device_training_equals_host_steps_on_the_drawn_rows():
    """rpl_mlp_train_device draws its batches on the device; two steps of it
    give the same parameters, bitwise, as two host train_step calls on the
    rows the hash draws (same bf16 conversion, padding and Adam)."""
    say("mlp device training vs host steps")
    rng = np.random.default_rng(9)
    rows, batch = 1000, 100
    X = rng.normal(size=(rows, 70)).astype(np.float32)
    T = (0.3 * rng.normal(size=(rows, 50))).astype(np.float32)
    a = sg.SourceNetwork(70, 50, hidden=64, hidden_layers=2, batch_capacity=128, seed=4)
    b = sg.SourceNetwork(70, 50, hidden=64, hidden_layers=2, batch_capacity=128, seed=4)
    seed = 12345
    hip = HipBuffers()
    Xd, Td = hip.upload(X), hip.upload(T)
    losses = sg.train_device(a, Xd, Td, rows=rows, steps=2, batch=batch, lr=2e-3, lr_final=1e-3, seed=seed)
    hip.free()
    host = []
    for step, lr in ((0, 2e-3), (1, 1e-3)):
        idx = _drawn_rows(seed, step, batch, rows)
        host.append(b.train_step(X[idx], T[idx], lr))
    pa, pb = a.params(), b.params()
    say(f"losses device {losses} host {host}; max |dp| {np.max(np.abs(pa - pb)):.3e}")
    assert np.array_equal(pa, pb)
    assert np.allclose(losses, host, rtol=1e-6)
    a.close()
    b.close()


def test_device_training_reduces_loss_and_checks_its_rows():
    import ctypes
    say("mlp device training")
    rng = np.random.default_rng(5)
    net = sg.SourceNetwork(48, 40, hidden=128, hidden_layers=2, batch_capacity=512, seed=3)
    X = rng.uniform(-1, 1, size=(4096, 48)).astype(np.float32)
    A = rng.normal(size=(48, 40)) / 7
    T = np.tanh(X @ A).astype(np.float32) + 0.5 * np.sin(X[:, :40] * 2).astype(np.float32)
    l0, _ = net.gradients(X[:512], T[:512])
    hip = HipBuffers()
    losses = sg.train_device(net, hip.upload(X), hip.upload(T), rows=X.shape[0], steps=400, batch=512, lr=2e-3,
                             log_every=100)
    hip.free()
    l1, _ = net.gradients(X[:512], T[:512])
    say(f"device training loss {l0:.3e} -> {l1:.3e}")
    assert l1 < l0 / 10 and np.all(np.isfinite(losses))
    # host memory is refused before any launch
    rc = net._lib.rpl_mlp_train_device(net.ptr, X.shape[0], ctypes.c_void_p(X.ctypes.data),
                                       ctypes.c_void_p(T.ctypes.data), 64, 1, 1e-3, 1e-3, ctypes.c_uint64(0), None)
    assert rc == -1  # RPL_ERR_ARG
    with pytest.raises(ValueError):
        sg.train_device(net, 1 << 40, 1 << 41, steps=1)  # an address without its row count
    net.close()


def test_saved_network_infers_identically(tmp_path):
    """save_network keeps bf16 weights (what the inference GEMMs read) and
    FP32 biases: the reloaded network's outputs are bitwise the original's."""
    say("mlp save / load")
    net, X, T = _net_and_batch(6)
    for _ in range(3):
        net.train_step(X, T, 1e-3)
    cfg = sg.SurrogateConfig(rate_unit=3e-6, scale=0.07, terms=1, wind_hi=33.0)
    hi = np.linspace(1.0, 2.0, 70)
    path = tmp_path / "net.npz"
    sg.save_network(path, net, cfg, bin_action_hi=hi, note="test")
    net2, cfg2, hi2, meta = sg.load_network(path, batch_capacity=256)
    assert np.array_equal(net.forward(X), net2.forward(X))
    assert cfg2 == cfg and meta == {"note": "test"}
    assert np.array_equal(hi2, hi.astype(np.float32))
    net.close()
    net2.close()


def test_shipped_storm_network_tracks_the_plain_bundle():
    """The trained hurricane network shipped with the package, on a 128^2
    translating hurricane (a coarser relative of the 1024^2 case it was
    trained for) over 30 steps of 60 s: the significant wave height stays
    within 5 % (relative L2) of the plain-source run, the domain-mean m0
    within 2 %, while dropping the sources moves Hs by tens of percent."""
    from ripple_amd import storm_surrogate, diagnostics as dg
    say("shipped storm surrogate")
    net, cfg, hi, meta = sg.load_network(storm_surrogate.NETWORK_FILE)
    wl = workloads.translating_hurricane(nx=128, ny=128, nz=8)
    runs = {}
    for label in ("plain", "surrogate", "none"):
        m = _model(wl, np.float32, sources=label != "none")
        if label == "surrogate":
            sg.attach_surrogate(m, net, cfg, bin_action_hi=hi)
        ra.time_step_(m, wl.dt, 30)
        runs[label] = (np.asarray(dg.significant_wave_height(m), np.float64),
                       float(np.mean(np.asarray(dg.m0(m), np.float64))))
        if label == "surrogate":
            st = sg.surrogate_stats(m)
            sg.attach_surrogate(m, None, None)
    e = rel_l2(runs["surrogate"][0], runs["plain"][0])
    e_none = rel_l2(runs["none"][0], runs["plain"][0])
    e_m0 = abs(runs["surrogate"][1] / runs["plain"][1] - 1)
    say(f"meta {dict((k, meta[k]) for k in ('pairs', 'terms', 'held_out_rate_rel_l2'))}; stats {st}; "
        f"Hs rel-L2 {e:.3e} (no sources {e_none:.3e}); m0 {e_m0:.3e}")
    assert st["plain"] < 0.1 * st["columns"]
    assert e < 0.05 and e_m0 < 0.02 and e_none > 0.2
    net.close()


def _storm(float_type, vmax=50.0, n=32, nkappa=12, nphi=12):
    wl = workloads.translating_hurricane(float_type=float_type, nx=n, ny=n, nz=4, nkappa=nkappa, nphi=nphi,
                                         dt=60.0)
    s = wl.model_kwargs["sources"].wind_input.wind
    s.vmax = vmax
    s.rmax, s.radius = 20e3, 150e3
    # a smaller box (200 km) so that the storm covers much of it
    L = 200e3
    g = wl.grid
    grid = grids.RectilinearGrid(float_type=float_type, size=(n, n, g.Nz), x=(0, L), y=(-L / 2, L / 2),
                                 z=g.z_faces, topology=(grids.PERIODIC, grids.PERIODIC, grids.BOUNDED))
    s.track.centers = [(60e3, 0.0), (60e3 + 5.0 * s.track.times[1], 0.0)]
    wl.grid = grid
    wl.N0 = workloads.jonswap_action(grid, wl.spectral_grid, 0.5, 0.15)
    return wl


def _model(wl, float_type, sources=True):
    kw = dict(wl.model_kwargs)
    if not sources:
        kw.pop("sources")
    sgr = wl.spectral_grid
    spec = grids.PolarWaveVectorGrid(float_type, kappa=sgr.kappa, phi=sgr.phi, kappa_faces=sgr.kappa_faces,
                                     phi_faces=sgr.phi_faces)
    m = ra.SpectralWaveModel(wl.grid, spec, velocities=None, **kw)
    ra.set_(m, N=wl.N0)
    return m


@pytest.mark.parametrize("float_type", [np.float32, np.float64])
def test_fallbacks_reproduce_the_plain_bundle(float_type):
    """Envelope fallback on every column, and failed audits on every column,
    both give the plain-source run; the counters say so.  FP64 is exact.  In
    FP32 the plain run takes the stress cap of RK3 stages 2 and 3 from the
    sums its stage kernels form (another summation order than the
    column-state kernel the network path runs), so the two agree to FP32
    rounding of the cap factor."""
    say(f"surrogate fallbacks {np.dtype(float_type).name}")
    wl = _storm(float_type)
    nb = wl.spectral_grid.kappa.size * wl.spectral_grid.phi.size
    plane = wl.grid.Nx * wl.grid.Ny
    ref = _model(wl, float_type)
    ra.time_step_(ref, wl.dt, 3)
    net = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=2, batch_capacity=256, seed=7)
    # a network whose every rate exceeds the FP16 range (the stage reads the
    # rates in FP16): each column is flagged non-finite and takes the plain bundle
    big = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=2, batch_capacity=256, seed=7)
    flat = big.params()
    out_pad = big.layers(flat)[-1][1].size
    flat[-out_pad:] = 1.0e5
    big.set_params(flat)
    for name, cfg in (("envelope", sg.SurrogateConfig(wind_lo=1e6, wind_hi=2e6)),
                      ("audit", sg.SurrogateConfig(audit_stride=1, audit_bound=0.0)),
                      ("fp16 overflow", sg.SurrogateConfig())):
        m = _model(wl, float_type)
        sg.attach_surrogate(m, big if name == "fp16 overflow" else net, cfg)
        ra.time_step_(m, wl.dt, 3)
        st = sg.surrogate_stats(m)
        say(f"{name}: {st}")
        assert st["columns"] == 9 * plane
        if name in ("envelope", "fp16 overflow"):
            assert st["plain"] == 9 * plane and st["audited"] == 0
        else:
            assert st["audited"] == 9 * plane and st["plain"] == 0
            assert st["failed"] >= 9 * plane - 9 * np.count_nonzero(np.all(wl.N0 == 0, axis=(2, 3)))
        if float_type == np.float64:
            assert np.array_equal(m.action, ref.action), rel_l2(m.action, ref.action)
        else:
            assert rel_l2(m.action, ref.action) <= 1e-6, rel_l2(m.action, ref.action)
        sg.attach_surrogate(m, None, None)
    net.close()
    big.close()


def test_stage_source_is_the_network_and_targets_are_the_bundle():
    """The source the stage adds on network columns equals the network's
    output on the training-pair features (scaled as documented), and the
    training targets reproduce the plain bundle's source."""
    say("surrogate consistency")
    ft = np.float64  # FP64 state: the difference of two tendencies isolates the source exactly
    wl = _storm(ft, 45.0, nkappa=8, nphi=8)
    nb = 64
    m_plain, m_none = _model(wl, ft), _model(wl, ft, sources=False)
    ra.time_step_(m_plain, wl.dt, 2)
    ra.set_(m_none, N=m_plain.action.astype(np.float64))
    ra._lib.check(m_none._ctx.lib.rpl_set_clock(m_none._ctx.ptr, m_plain.clock.time, 2))
    cfg = sg.SurrogateConfig(rate_unit=1e-4)
    X, T = sg.training_pairs(m_plain, cfg)
    G_plain, G_none = m_plain.tendency().astype(np.float64), m_none.tendency().astype(np.float64)
    N = m_plain.action.astype(np.float64)
    Nx, Ny = N.shape[:2]
    cols = N.reshape(Nx * Ny, nb, order="F")
    sigma = np.sqrt(9.81 * wl.spectral_grid.kappa)[:, None] * np.ones(wl.spectral_grid.phi.size)[None, :]
    sig = sigma.ravel(order="F")
    S_plain = (G_plain - G_none).reshape(Nx * Ny, nb, order="F")
    S_from_T = T * cfg.rate_unit * sig[None, :] * cols
    e_t = rel_l2(S_from_T, S_plain)
    net = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=2, batch_capacity=Nx * Ny, seed=4)
    Y = net.forward(X)
    sg.attach_surrogate(m_plain, net, cfg)
    G_surr = m_plain.tendency().astype(np.float64)
    S_surr = (G_surr - G_none).reshape(Nx * Ny, nb, order="F")
    S_pred = Y * cfg.rate_unit * sig[None, :] * cols
    e_s = rel_l2(S_surr, S_pred)
    # the stage reads the rates in FP16 (relative rounding <= 2^-11)
    S_pred16 = Y.astype(np.float16).astype(np.float64) * cfg.rate_unit * sig[None, :] * cols
    e_s16 = rel_l2(S_surr, S_pred16)
    say(f"targets vs bundle source {e_t:.2e}; stage source vs network {e_s:.2e}, vs its FP16 rates {e_s16:.2e}")
    assert e_t < 1e-4
    assert e_s < 1e-3
    assert e_s16 < 1e-4
    net.close()


def test_wind_input_surrogate_keeps_the_plain_dissipation():
    """terms = 1: the stage adds the network's wind input and the plain
    bundle's dissipation; the input-only targets differ from the full ones by
    exactly the dissipation."""
    say("surrogate terms=1 consistency")
    ft = np.float64
    wl = _storm(ft, 45.0, nkappa=8, nphi=8)
    nb = 64
    m_plain = _model(wl, ft)
    ra.time_step_(m_plain, wl.dt, 2)
    N = m_plain.action.astype(np.float64)

    def with_sources(src):
        w2 = _storm(ft, 45.0, nkappa=8, nphi=8)
        if src is not None:
            w2.model_kwargs["sources"] = src
        m = _model(w2, ft, sources=src is not None)
        ra.set_(m, N=N)
        ra._lib.check(m._ctx.lib.rpl_set_clock(m._ctx.ptr, m_plain.clock.time, 2))
        return m

    m_none = with_sources(None)
    m_diss = with_sources(wl.model_kwargs["sources"].dissipation)
    Nx, Ny = N.shape[:2]
    cols = N.reshape(Nx * Ny, nb, order="F")
    G_none = m_none.tendency().astype(np.float64)
    D = (m_diss.tendency().astype(np.float64) - G_none).reshape(Nx * Ny, nb, order="F")
    sigma = (np.sqrt(9.81 * wl.spectral_grid.kappa)[:, None] * np.ones(wl.spectral_grid.phi.size)[None, :]).ravel(
        order="F")
    full = sg.SurrogateConfig(rate_unit=1e-4)
    part = sg.SurrogateConfig(rate_unit=1e-4, terms=1)
    X, T_full = sg.training_pairs(m_plain, full)
    _, T_in = sg.training_pairs(m_plain, part)
    e_t = rel_l2((T_full.astype(np.float64) - T_in) * 1e-4 * sigma[None, :] * cols, D)
    net = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=2, batch_capacity=Nx * Ny, seed=5)
    Y = net.forward(X)
    sg.attach_surrogate(m_plain, net, part)
    S = (m_plain.tendency().astype(np.float64) - G_none).reshape(Nx * Ny, nb, order="F")
    # the stage reads the network's rates in FP16
    e_s = rel_l2(S, Y.astype(np.float16).astype(np.float64) * 1e-4 * sigma[None, :] * cols + D)
    say(f"terms=1: (full - input) targets vs dissipation {e_t:.2e}; stage source vs network + dissipation {e_s:.2e}")
    assert e_t < 1e-5 and e_s < 1e-6
    net.close()


@pytest.mark.parametrize("terms", [0, 1])
def test_trained_surrogate_gains_over_no_sources(terms):
    """Train on plain-bundle pairs from storms of vmax 35, 40 and 50 m/s, run a
    45 m/s storm with the network supplying the sources (terms 0) or the wind
    input (terms 1), and compare the significant wave height and total action
    against the plain-source run and against the run without sources."""
    say(f"surrogate train + run, terms={terms}")
    ft = np.float32
    nk, npp = 12, 12
    nb = nk * npp
    cfg = sg.SurrogateConfig(rate_unit=1.0, scale=0.05, floor=1e-24, terms=terms)
    Xs, Ts = [], []
    bin_hi = np.zeros(nb)
    for vmax in (35.0, 40.0, 50.0):
        m = _model(_storm(ft, vmax, nkappa=nk, nphi=npp), ft)
        X, T, bmax = sg.collect_pairs(m, 60.0, 60, 4, cfg, return_bin_max=True)
        Xs.append(X)
        Ts.append(T)
        bin_hi = np.maximum(bin_hi, bmax)
    bin_hi *= 1.05
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    # The network covers the wind-forced columns (U10 >= 5 m/s); weaker-wind
    # columns stay on the plain bundle through the envelope.
    u10 = np.hypot(X[:, nb], X[:, nb + 1]) / cfg.wind_scale
    keep = u10 >= 5.0
    X, T = X[keep], T[keep]
    unit = float(np.sqrt(np.mean(T[T != 0] ** 2)))
    T = T / unit
    cfg.rate_unit = unit
    sg.envelope_from_pairs(X, cfg)
    cfg.wind_lo = 5.0
    say(f"pairs {X.shape}, rate unit {unit:.3e}, envelope U10 [{cfg.wind_lo:.2f}, {cfg.wind_hi:.2f}], "
        f"action <= {cfg.action_hi:.3e}")
    net = sg.SourceNetwork(nb + 3, nb, hidden=512, hidden_layers=3, batch_capacity=1024, seed=11)
    losses = sg.fit(net, X, T, steps=12000, batch=1024, lr=1e-3, lr_final=1e-5)
    fit_err = rel_l2(net.forward(X[:1024]), T[:1024])
    say(f"standardised loss {losses[0]:.3e} -> {losses[-1]:.3e}; rel-L2 of fitted rates {fit_err:.3e}")
    assert losses[-1] < losses[0] / 10

    wl = _storm(ft, 45.0, nkappa=nk, nphi=npp)
    cfg.audit_stride, cfg.audit_bound = 16, 0.5
    plain, surr, none = _model(wl, ft), _model(wl, ft), _model(wl, ft, sources=False)
    sg.attach_surrogate(surr, net, cfg, bin_action_hi=bin_hi)
    nsteps = 30
    for m in (plain, surr, none):
        ra.time_step_(m, wl.dt, nsteps)
    st = sg.surrogate_stats(surr)
    hs = [ra.significant_wave_height(m) for m in (plain, surr, none)]
    ta = [ra.total_action(m) for m in (plain, surr, none)]
    e_surr = rel_l2(hs[1], hs[0])
    e_none = rel_l2(hs[2], hs[0])
    a_surr = abs(ta[1] - ta[0]) / ta[0]
    a_none = abs(ta[2] - ta[0]) / ta[0]
    say(f"stats {st}")
    say(f"Hs rel-L2 vs plain: surrogate {e_surr:.3e}, no sources {e_none:.3e}; "
        f"total action rel: surrogate {a_surr:.3e}, no sources {a_none:.3e}")
    assert st["plain"] < st["columns"]
    assert np.all(np.isfinite(surr.action))
    assert e_surr < (0.25 if terms == 0 else 0.1) * e_none
    assert a_surr < 0.25 * a_none
    net.close()


def test_surrogate_error_paths():
    say("surrogate error paths")
    wl = _storm(np.float32)
    m = _model(wl, np.float32, sources=False)
    nb = wl.spectral_grid.kappa.size * wl.spectral_grid.phi.size
    net = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=1, batch_capacity=64)
    with pytest.raises(ra.RippleError):
        sg.attach_surrogate(m, net, sg.SurrogateConfig())  # no plain bundle installed
    with pytest.raises(ra.RippleError):
        sg.surrogate_stats(m)
    m2 = _model(wl, np.float32)
    bad = sg.SourceNetwork(nb + 2, nb, hidden=64, hidden_layers=1, batch_capacity=64)
    with pytest.raises(ra.RippleArgumentError):
        sg.attach_surrogate(m2, bad, sg.SurrogateConfig())
    with pytest.raises(ra.RippleArgumentError):
        sg.attach_surrogate(m2, net, sg.SurrogateConfig(rate_unit=0.0))
    with pytest.raises(ValueError):
        net.forward(np.zeros((65, nb + 3), np.float32))
    with pytest.raises(ra.RippleArgumentError):
        sg.SourceNetwork(10, 10, hidden=100)
    net.close()
    bad.close()


@pytest.mark.parametrize("float_type", [np.float32, np.float64])
def test_rates_once_per_step(float_type):
    """refresh = 1: the rates are inferred for each step's first stage only
    (the column counter grows by one plane per step, not three) and held
    through its later stages; the run stays close to per-stage rates but is
    not the same; a tendency evaluation still infers on its own input, and
    an invalid refresh is refused."""
    say(f"surrogate rates once per step {np.dtype(float_type).name}")
    wl = _storm(float_type)
    nb = wl.spectral_grid.kappa.size * wl.spectral_grid.phi.size
    plane = wl.grid.Nx * wl.grid.Ny
    net = sg.SourceNetwork(nb + 3, nb, hidden=64, hidden_layers=2, batch_capacity=256, seed=11)
    every, once = _model(wl, float_type), _model(wl, float_type)
    sg.attach_surrogate(every, net, sg.SurrogateConfig(rate_unit=1e-3))
    sg.attach_surrogate(once, net, sg.SurrogateConfig(rate_unit=1e-3, refresh=1))
    ra.time_step_(every, wl.dt, 3)
    ra.time_step_(once, wl.dt, 3)
    se, so = sg.surrogate_stats(every), sg.surrogate_stats(once)
    assert se["columns"] == 9 * plane and so["columns"] == 3 * plane
    e = rel_l2(once.action, every.action)
    say(f"once-per-step vs every-stage rates after 3 steps: rel-L2 {e:.2e}")
    assert 0 < e < 1e-2
    n_before = sg.surrogate_stats(once)["columns"]
    once.tendency()
    assert sg.surrogate_stats(once)["columns"] == n_before + plane
    plain = _model(wl, float_type)
    with pytest.raises(ra.RippleArgumentError):
        sg.attach_surrogate(plain, net, sg.SurrogateConfig(refresh=2))
    for m in (every, once):
        sg.attach_surrogate(m, None, None)
    net.close()
