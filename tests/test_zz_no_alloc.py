"""rpl_weighted_total_noalloc, the allocation-free variant of
rpl_conservation's compensated totals: through the C ABI on device arrays
the test owns (allocated with the HIP runtime the library links), against
rpl_conservation on a context holding the same action, on two calls in a
row (the second smaller), FP64 and FP32."""
import ctypes
import math
import sys

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import grids, workloads
from hip_buffers import HipBuffers

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _model(nx, ny, nk, nphi, float_type, seed):
    rng = np.random.default_rng(seed)
    g = grids.RectilinearGrid(float_type=float_type, size=(nx, ny, 2), x=(0, 0.5 * nx), y=(0, 0.4 * ny), z=(-1, 0))
    sg = grids.PolarWaveVectorGrid(float_type, kappa=np.geomspace(0.2, 1.5, nk),
                                   phi=np.linspace(0, 2 * math.pi, nphi + 1)[:nphi])
    m = ra.SpectralWaveModel(g, sg, timestepper="RK3")
    N = np.asfortranarray(rng.random((nx, ny, nk, nphi)) * 10.0 ** rng.integers(-4, 2, size=(nx, ny, nk, nphi)))
    ra.set_(m, N=N)
    return m, sg


_Hip = HipBuffers


def _noalloc_total(lib, hip, N, measure, area, dtype_code):
    dev_N = hip.upload(N.ravel(order="F"))
    dev_w = hip.upload(np.asarray(measure, dtype=np.float64).ravel(order="F"))
    out = hip.upload(np.full(1, np.nan))
    nx, ny, nk, nphi = N.shape
    ra._lib.check(lib.rpl_weighted_total_noalloc(dev_N, dtype_code, nx, ny, nk, nphi, dev_w, float(area), out,
                                                 None))
    return float(hip.download(out, np.empty(1))[0])


@pytest.mark.parametrize("float_type", [np.float64, np.float32])
def test_noalloc_total_matches_conservation(float_type):
    lib = ra._lib.load()
    hip = _Hip()
    code = ra._lib.RPL_F64 if float_type == np.float64 else ra._lib.RPL_F32
    # the first call spans the whole 1024-workgroup grid; the second is smaller
    for shape, seed in (((700, 500, 5, 7), 1), ((37, 11, 3, 4), 2)):
        say(f"noalloc total {np.dtype(float_type).name} {shape}")
        m, sg = _model(*shape, float_type, seed)
        N = m.action
        dx = float(m.grid.x_faces[1] - m.grid.x_faces[0])
        dy = float(m.grid.y_faces[1] - m.grid.y_faces[0])
        kf, pf = np.asarray(sg.kappa_faces, np.float64), np.asarray(sg.phi_faces, np.float64)
        w = ((kf[1:] * kf[1:] - kf[:-1] * kf[:-1]) / 2)[:, None] * np.diff(pf)[None, :]
        got = _noalloc_total(lib, hip, N, w, dx * dy, code)
        want = ra.total_action(m)
        say(f"noalloc {got!r} vs rpl_conservation {want!r}")
        # same kernel, same grid and the same fold order: equal to rounding of
        # the measures only (the host builds w with the library's formula)
        assert got == pytest.approx(want, rel=1e-14)
    hip.free()


def test_noalloc_rejects_bad_sizes():
    lib = ra._lib.load()
    one = ctypes.c_void_p(1)
    assert lib.rpl_weighted_total_noalloc(one, ra._lib.RPL_F64, 1 << 15, 1 << 15, 1, 1, one, 1.0, one,
                                          None) == -1
    assert b"2^30" in lib.rpl_last_error()
    assert lib.rpl_weighted_total_noalloc(one, 7, 4, 4, 1, 1, one, 1.0, one, None) == -1
    assert lib.rpl_weighted_total_noalloc(None, ra._lib.RPL_F64, 4, 4, 1, 1, one, 1.0, one, None) == -1
