"""State round trips and failure flags through the C ABI: a run resumed from
a saved (action, clock) continues bit for bit (the reference's state is
model.action plus model.clock, set.jl:20-48), and the per-step non-finite
check flags a NaN forced into the action from the host."""
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def build(wl):
    m = ra.SpectralWaveModel(wl.grid, wl.spectral_grid, velocities=wl.velocities, **wl.model_kwargs)
    ra.set_(m, N=wl.N0)
    return m


@pytest.mark.parametrize("case", ["hurricane_fp32", "vortex_fp64"])
def test_resume_is_bit_identical(case):
    """Six steps straight against three, a save of (interior action, time,
    iteration), a fresh model set from it, and three more.  The hurricane's
    storm moves with the clock, so the resumed clock matters."""
    say(f"resume round trip {case}")
    if case == "hurricane_fp32":
        wl = workloads.translating_hurricane(nx=64, ny=64, nz=4, nkappa=16, nphi=24, dt=600.0)
    else:
        wl = workloads.vortex_refraction(nsteps=6)
    a = build(wl)
    ra.time_step_(a, wl.dt, 6)
    b = build(wl)
    ra.time_step_(b, wl.dt, 3)
    saved, t, it = b.action.copy(), b.clock.time, b.clock.iteration
    del b
    c = build(wl)
    ra.set_(c, N=saved)
    c.set_clock(t, it)
    ra.time_step_(c, wl.dt, 3)
    assert c.clock.iteration == a.clock.iteration == 6
    assert c.clock.time == a.clock.time
    assert np.array_equal(c.action, a.action)
    if case == "hurricane_fp32":
        # without the clock the storm sits elsewhere and the result differs
        d = build(wl)
        ra.set_(d, N=saved)
        ra.time_step_(d, wl.dt, 3)
        assert not np.array_equal(d.action, a.action)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nonfinite_flag(dtype):
    say(f"non-finite check {np.dtype(dtype).name}")
    wl = workloads.vortex_refraction(float_type=dtype, nsteps=4)
    clean = build(wl)
    clean.set_nonfinite_check(1)
    ra.time_step_(clean, wl.dt, 4)
    assert clean.nonfinite() == (-1, 0)
    bad = wl.N0.copy()
    bad[10, 20, 1, 5] = np.nan
    m = build(wl)
    ra.set_(m, N=bad)
    m.set_nonfinite_check(3)
    ra.time_step_(m, wl.dt, 2)
    assert m.nonfinite() == (-1, 0)  # not checked yet
    ra.time_step_(m, wl.dt, 1)
    it, count = m.nonfinite()
    # the count is that of the action the check saw (the NaN has spread
    # through the stencils)
    assert it == 3 and count == int(np.count_nonzero(~np.isfinite(m.action))) > 1
    ra.time_step_(m, wl.dt, 3)
    assert m.nonfinite() == (3, count)  # the first detection is kept
    m.set_nonfinite_check(1)  # clears the record
    assert m.nonfinite() == (-1, 0)
    ra.time_step_(m, wl.dt, 1)
    assert m.nonfinite()[0] == 7
    with pytest.raises(ra.RippleArgumentError):
        m.set_nonfinite_check(-1)
