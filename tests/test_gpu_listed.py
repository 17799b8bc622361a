"""Device DiscreteInteractionApproximation{ListedQuadruplets}
(Sources/Sources.jl:1268-1385) through the C-ABI against the FP64 oracle:
the reference's known answer (test/sources/source_terms.jl:1098-1187), RK3
steps on a larger spectrum with overlapping interactions, FP32 against FP64,
and the reference's argument checks."""
import ctypes
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import grids, workloads

pytestmark = pytest.mark.gpu
Q = ra.QuadrupletTransferInteraction


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def model(sources, frequency, nphi, float_type=np.float64, nx=4, ny=3, timestepper="RK3"):
    g = grids.RectilinearGrid(float_type=float_type, size=(nx, ny, 1), x=(0, 1), y=(0, 1), z=(0, 1))
    sg = grids.FrequencyDirectionGrid(float_type, frequency=frequency,
                                      phi=np.arange(nphi) * (2 * math.pi / nphi))
    return ra.SpectralWaveModel(g, sg, horizontal_advection=None, sources=sources, timestepper=timestepper)


def weights(sg):
    kf, pf = sg.kappa_faces, sg.phi_faces
    return np.array([[(kf[m + 1] ** 2 - kf[m] ** 2) / 2 * (pf[n + 1] - pf[n]) for n in range(len(pf) - 1)]
                     for m in range(len(kf) - 1)])


def test_reference_known_answer():
    say("listed quadruplets: reference known answer")
    dia = ra.DiscreteInteractionApproximation((Q((1, 1), (1, 3), (1, 2), (1, 4), rate=0.05),), power=1)
    m = model(dia, [0.2], 4, timestepper="ForwardEuler")
    N = np.full(m.size, 0.5, order="F")
    N[:, :, 0, :] = [4.0, 0.25, 3.0, 0.75]
    ra.set_(m, N=N)
    w = weights(m.spectral_grid)
    G = m.tendency()
    flux = 0.05 * 4.0 * 3.0
    want = np.array([-flux / w[0, 0], flux / w[0, 1], -flux / w[0, 2], flux / w[0, 3]])
    np.testing.assert_allclose(G[:, :, 0, :], np.broadcast_to(want, G[:, :, 0, :].shape), rtol=1e-15, atol=0)
    before = (N[0, 0, 0] * w[0]).sum()
    ra.time_step_(m, 0.01)
    A = m.action[0, 0, 0]
    assert abs((A * w[0]).sum() - before) <= 1e-14 * before
    assert A[1] + A[3] > 1.0 and (m.action >= 0).all()


def _resonant_list():
    """Equal-frequency quadruplets whose donor and receiver wavevectors both
    sum to zero (antipodal pairs) on a 24-direction grid, at three
    frequencies, sharing bins so that sums accumulate."""
    out = []
    for m, rate in ((2, 3.0), (5, 1.5), (5, 0.7), (7, 2.2)):
        for n0 in range(0, 12, 5 if rate != 0.7 else 7):
            out.append(Q((m, n0 + 1), (m, n0 + 13), (m, (n0 + 6) % 24 + 1), (m, (n0 + 18) % 24 + 1), rate=rate))
    return tuple(out)


@pytest.mark.parametrize("power", [1.0, 1.5])
def test_fp64_steps_against_oracle(oracle, power):
    say(f"listed quadruplets fp64 power={power}")
    freq = np.geomspace(0.08, 0.3, 8)
    quads = _resonant_list()
    dia = ra.DiscreteInteractionApproximation(quads, power=power)
    m = model(dia, freq, 24, nx=6, ny=5)
    rng = np.random.default_rng(9)
    N0 = np.asfortranarray(rng.uniform(0.0, 1.0, m.size))
    N0[0, 0, 4, 3] = -0.2  # a negative donor switches its interaction off in that column
    ra.set_(m, N=N0)
    sg = m.spectral_grid
    wl = workloads.Workload("listed", m.grid, sg, None, N0, 0.02, 3, dict(timestepper="RK3"))
    om = oracle.OracleModel(wl, transport=oracle.MODE_NONE, timestepper="RK3")
    inter = [tuple((b[0] - 1, b[1] - 1) for b in q.bins) + (q.rate,) for q in quads]
    hook, user, ctx = oracle.listed_hook(m.size, inter, power, weights(sg))
    G = m.tendency()
    Gw = np.zeros(N0.size)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(hook.value)
    Nf = oracle._c(N0)
    fn(Nf.ctypes.data, Gw.ctypes.data, user)
    Gw = Gw.reshape(m.size, order="F")
    err = rel_l2(G, Gw)
    say(f"tendency rel-L2 {err:.3e}, bitwise {np.array_equal(G, Gw)}")
    assert (Gw > 0).any() and (Gw < 0).any()
    if power == 1.0:
        assert np.array_equal(G, Gw)  # same operation order, no contraction
    else:
        assert err < 1e-15  # device pow within an ulp of the host's
    ra.time_step_(m, 0.02, 3)
    om.step(0.02, 3, hook, user)
    err = rel_l2(m.action, om.action)
    say(f"3 RK3 steps rel-L2 {err:.3e}")
    assert err < 1e-14


def test_listed_as_bundle_member_against_oracle(oracle):
    """PrecomputedSources(wind input, local saturation, nonlinear =
    DiscreteInteractionApproximation(listed quadruplets)) (the bundle's
    source_split, precomputed_sources.jl:240-261): the device adds the
    listed source after the bundle's, so the oracle runs its bundle hook and
    then its listed hook.  The reference sums every member's positive and
    damping parts before forming positive - damping N, a rounding-level
    difference (bound below)."""
    from ripple_amd.sources import to_struct
    say("listed quadruplets as a bundle member fp64")
    freq = np.geomspace(0.08, 0.3, 8)
    quads = _resonant_list()
    src = ra.PrecomputedSources(
        wind_input=ra.PressureCorrelationInput(drag=ra.BulkWindDrag("linear"), wind=12.0, direction=0.4),
        dissipation=ra.LocalSaturationDissipation(), nonlinear=ra.DiscreteInteractionApproximation(quads))
    m = model(src, freq, 24, nx=6, ny=5)
    N0 = np.asfortranarray(np.random.default_rng(11).uniform(0.0, 1.0, m.size))
    ra.set_(m, N=N0)
    sg = m.spectral_grid
    wl = workloads.Workload("listed_bundle", m.grid, sg, None, N0, 0.02, 3, dict(timestepper="RK3", sources=src))
    om = oracle.OracleSourceModel(wl, oracle.sources_from(to_struct(src)), transport=oracle.MODE_NONE)
    inter = [tuple((b[0] - 1, b[1] - 1) for b in q.bins) + (q.rate,) for q in quads]
    _, luser, lctx = oracle.listed_hook(m.size, inter, 1.0, weights(sg))
    L = oracle.lib()
    V = ctypes.c_void_p

    def both(N, G, user):
        L.orc_sources_hook_f64(V(N), V(G), ctypes.pointer(om._ctx))
        L.orc_listed_hook_f64(V(N), V(G), luser)

    hook = ctypes.CFUNCTYPE(None, V, V, V)(both)
    G = m.tendency()
    Gw = oracle._c(om.tendency())
    L.orc_listed_hook_f64(V(oracle._c(N0).ctypes.data), V(Gw.ctypes.data), luser)
    Gw = Gw.reshape(m.size, order="F")
    G_alone = model(src.nonlinear, freq, 24, nx=6, ny=5)
    ra.set_(G_alone, N=N0)
    err = rel_l2(G, Gw)
    say(f"tendency rel-L2 {err:.3e}; listed part alone {rel_l2(G_alone.tendency(), Gw):.3e} away")
    assert err < 1e-13 and rel_l2(G_alone.tendency(), Gw) > 1e-6  # the bundle terms are in it
    ra.time_step_(m, wl.dt, 3)
    for _ in range(3):
        om._ctx.time = om.time
        oracle.OracleModel.step(om, wl.dt, 1, ctypes.cast(hook, V), None)
    err = rel_l2(m.action, om.action)
    say(f"3 RK3 steps rel-L2 {err:.3e}")
    assert err < 1e-12
    del lctx


def test_fp32_against_fp64_oracle(oracle):
    say("listed quadruplets fp32")
    freq = np.geomspace(0.08, 0.3, 8)
    quads = _resonant_list()
    m32 = model(ra.DiscreteInteractionApproximation(quads), freq, 24, float_type=np.float32, nx=6, ny=5)
    m64 = model(ra.DiscreteInteractionApproximation(quads), freq, 24, nx=6, ny=5)
    N0 = np.asfortranarray(np.random.default_rng(3).uniform(0.0, 1.0, m64.size))
    for m in (m32, m64):
        ra.set_(m, N=N0)
        ra.time_step_(m, 0.02, 3)
    err = rel_l2(m32.action, m64.action)
    say(f"FP32 vs FP64 3 steps rel-L2 {err:.3e}")
    assert err < 1e-6


def test_argument_checks():
    """source_terms.jl:1163-1187 plus the checks of Sources.jl:1274-1347."""
    say("listed quadruplets argument checks")
    for bad in ((Q((1, 1), (1, 2), (1, 3), (1, 4), rate=0.05),),   # not resonant in wavevector
                (Q((1, 1), (1, 3), (1, 2), (1, 5), rate=0.05),),   # direction index out of range
                (Q((1, 1), (1, 3), (1, 2), (1, 4), rate=-0.05),),  # negative rate
                (Q((1, 1), (1, 3), (1, 3), (1, 1), rate=0.05),)):  # donors are the receivers
        with pytest.raises(ra.RippleArgumentError):
            model(ra.DiscreteInteractionApproximation(bad), [0.2], 4)
    ok = (Q((1, 1), (1, 3), (1, 2), (1, 4), rate=0.05),)
    with pytest.raises(ra.RippleArgumentError):  # frequencies differ: 0.2 + 0.2 against 0.3 + 0.3
        model(ra.DiscreteInteractionApproximation((Q((1, 1), (1, 3), (2, 2), (2, 4), rate=0.05),)), [0.2, 0.3], 4)
    with pytest.raises(ra.RippleArgumentError):
        model(ra.DiscreteInteractionApproximation(ok, power=0.0), [0.2], 4)
    with pytest.raises(ra.RippleArgumentError):
        model(ra.DiscreteInteractionApproximation(ok, resonance_tolerance=-1.0), [0.2], 4)
    g = grids.RectilinearGrid(size=(2, 2, 1), x=(0, 1), y=(0, 1), z=(0, 1))
    with pytest.raises(ValueError):  # source_terms.jl:1184-1187
        ra.SpectralWaveModel(g, grids.PolarWaveVectorGrid(kappa=[1.0], phi=[0.0, 1.6, 3.1, 4.7]),
                             horizontal_advection=None, sources=ra.DiscreteInteractionApproximation(ok))
    # as the bundle's only member it is the listed source itself
    mb = model(ra.PrecomputedSources(nonlinear=ra.DiscreteInteractionApproximation(ok)), [0.2], 4)
    ma = model(ra.DiscreteInteractionApproximation(ok), [0.2], 4)
    N = np.full(ma.size, 0.5, order="F")
    N[:, :, 0, :] = [4.0, 0.25, 3.0, 0.75]
    for mm in (mb, ma):
        ra.set_(mm, N=N)
    np.testing.assert_allclose(mb.tendency(), ma.tendency(), rtol=1e-15, atol=0)
    # An empty list is a no-op source.
    m = model(ra.DiscreteInteractionApproximation(), [0.2], 4)
    ra.set_(m, N=1.0)
    assert not m.tendency().any()


def _field(quads, m, t):
    """(Nx, Ny, count) rates of the listed interactions at time t."""
    from ripple_amd.sources import source_value_field
    from ripple_amd.grids import xnodes, ynodes
    return np.stack([source_value_field(q.rate, xnodes(m.grid), ynodes(m.grid), t) for q in quads], axis=2)


@pytest.mark.parametrize("power", [1.0, 1.5])
def test_rates_per_column_and_in_time_against_oracle(oracle, power):
    """Rates as the reference's source_value takes them (Sources.jl:
    1450-1467): an (Nx, Ny) array, a function of (x, y) and a function of
    (x, y, t), which the host evaluates at the clock before every step."""
    say(f"listed quadruplets fp64 with per-column and time-dependent rates, power={power}")
    freq = np.geomspace(0.08, 0.3, 8)
    base = _resonant_list()
    nx, ny = 6, 5
    rng = np.random.default_rng(11)
    arr = rng.uniform(0.5, 2.5, (nx, ny))
    forms = [arr, lambda x, y: 1.0 + 0.5 * np.sin(7 * x) * np.cos(5 * y),
             lambda x, y, t: 2.0 + np.cos(3 * x + 40 * t) + 0.1 * y]
    quads = tuple(Q(*q.bins, rate=forms[k % 3]) if k % 2 == 0 else q for k, q in enumerate(base))
    dia = ra.DiscreteInteractionApproximation(quads, power=power)
    m = model(dia, freq, 24, nx=nx, ny=ny)
    N0 = np.asfortranarray(rng.uniform(0.0, 1.0, m.size))
    ra.set_(m, N=N0)
    sg = m.spectral_grid
    inter = [tuple((b[0] - 1, b[1] - 1) for b in q.bins) + (0.0,) for q in quads]
    G = m.tendency()
    hook, user, ctx = oracle.listed_hook(m.size, inter, power, weights(sg), rate_field=_field(quads, m, 0.0))
    Gw = np.zeros(N0.size)
    fn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)(hook.value)
    Nf = oracle._c(N0)
    fn(Nf.ctypes.data, Gw.ctypes.data, user)
    Gw = Gw.reshape(m.size, order="F")
    err = rel_l2(G, Gw)
    say(f"tendency rel-L2 {err:.3e}")
    assert err < (1e-300 if power == 1.0 else 1e-15)
    wl = workloads.Workload("listed-field", m.grid, sg, None, N0, 0.02, 3, dict(timestepper="RK3"))
    om = oracle.OracleModel(wl, transport=oracle.MODE_NONE, timestepper="RK3")
    dt = 0.02
    ra.time_step_(m, dt, 3)
    for k in range(3):
        hook, user, ctx = oracle.listed_hook(m.size, inter, power, weights(sg), rate_field=_field(quads, m, k * dt))
        om.step(dt, 1, hook, user)
    err = rel_l2(m.action, om.action)
    say(f"3 RK3 steps rel-L2 {err:.3e}")
    assert err < (1e-300 if power == 1.0 else 1e-14)
    assert m.clock.iteration == 3
    # the time dependence matters: rates frozen at t = 0 give another answer
    frozen = oracle.OracleModel(wl, transport=oracle.MODE_NONE, timestepper="RK3")
    hook, user, ctx = oracle.listed_hook(m.size, inter, power, weights(sg), rate_field=_field(quads, m, 0.0))
    frozen.step(dt, 3, hook, user)
    assert rel_l2(frozen.action, om.action) > 1e-8


def test_rate_field_errors():
    say("listed quadruplet rate-field errors")
    freq = np.geomspace(0.08, 0.3, 8)
    quads = _resonant_list()
    neg = tuple(Q(*q.bins, rate=(lambda x, y: x - 0.5)) if k == 0 else q for k, q in enumerate(quads))
    with pytest.raises(ra.RippleArgumentError):
        model(ra.DiscreteInteractionApproximation(neg), freq, 24, nx=6, ny=5)  # Sources.jl:1347
    m = model(ra.DiscreteInteractionApproximation(quads), freq, 24, nx=6, ny=5)
    lib = m._ctx.lib
    ra._lib.check(lib.rpl_set_quadruplet_rate_field(m._ctx.ptr, None))  # back to the constants
    bad = np.full(6 * 5 * len(quads), np.nan)
    with pytest.raises(ra.RippleArgumentError):
        ra._lib.check(lib.rpl_set_quadruplet_rate_field(m._ctx.ptr, ra._lib.dptr(bad)))
