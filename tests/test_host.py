"""CPU-side checks of the host layer and the C ABI boundary (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import ripple_amd as ra
from ripple_amd import _lib, grids, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rpl_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert "rpl_time_step" in names and "rpl_create" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes table and header disagree"


def test_create_rejects_bad_configs_without_touching_the_device():
    lib = _lib.load()
    for cfg, frag in (((2, 8, 4, 8), "nx and ny"), ((8, 8, 4, 2), "nphi"), ((8, 8, 0, 8), "nkappa")):
        c = _lib.rpl_config(*cfg, _lib.RPL_F64, _lib.RPL_RK3, _lib.RPL_TRANSPORT_REFRACTION, 0)
        ptr = ctypes.c_void_p()
        rc = lib.rpl_create(ctypes.byref(ptr), ctypes.byref(c))
        assert rc == -1 and frag in lib.rpl_last_error().decode()
    c = _lib.rpl_config(8, 8, 300, 300, _lib.RPL_F64, _lib.RPL_RK3, _lib.RPL_TRANSPORT_REFRACTION, 0)
    assert lib.rpl_create(ctypes.byref(ctypes.c_void_p()), ctypes.byref(c)) == -4
    assert lib.rpl_time_step(None, 0.1, 1) == -1


def test_grids_follow_reference_constructors():
    g = grids.RectilinearGrid(size=(4, 3, 2), x=(0, 1), y=(0, 3), z=(-1, 0))
    np.testing.assert_allclose(grids.xnodes(g), [0.125, 0.375, 0.625, 0.875])
    assert g.Lz == 1.0
    sg = grids.PolarWaveVectorGrid(kappa=np.linspace(0.25, 1.0, 4), phi=np.linspace(0, 2 * math.pi, 9)[:8])
    assert sg.weights.sum() == pytest.approx(math.pi * (sg.kappa_faces[-1] ** 2 - sg.kappa_faces[0] ** 2))
    with pytest.raises(ValueError):
        grids.PolarWaveVectorGrid(kappa=[0.3, 0.2], phi=[0.0])
    with pytest.raises(ValueError):
        grids.PolarWaveVectorGrid(kappa=[-0.1, 0.2], phi=[0.0])
    fg = grids.FrequencyDirectionGrid(frequency=np.linspace(0.08, 0.2, 4), phi=[0.0, 2 * math.pi / 3, 4 * math.pi / 3])
    np.testing.assert_allclose(fg.kappa, (2 * math.pi * fg.frequency) ** 2 / 9.81)


def test_workloads_are_deterministic():
    a = workloads.sheared_eddies(nx=64, ny=64, nz=8, nkappa=4, nphi=6)
    b = workloads.sheared_eddies(nx=64, ny=64, nz=8, nkappa=4, nphi=6)
    np.testing.assert_array_equal(a.velocities[0], b.velocities[0])
    u = a.velocities[0][:, :, -1]
    v = a.velocities[1][:, :, -1]
    surface_rms = math.sqrt(float(np.mean(u * u + v * v)))
    z_top = grids.znodes(a.grid)[-1]
    assert surface_rms == pytest.approx(0.3 * math.exp(z_top / 0.2), rel=1e-12)
    wl = workloads.vortex_refraction()
    assert wl.N0.shape == (64, 64, 4, 16) and wl.cell_bins == 262144
    assert np.max(wl.N0) <= 1.0 and np.min(wl.N0) > 0


def test_rectilinear_grid_accessors_known_answers():
    """test/product_fields/indexing.jl:81-98: a stretched grid's faces, nodes,
    spacings and topology."""
    g = grids.RectilinearGrid(size=(3, 2, 1), x=[0, 0.25, 1.0, 2.0], y=[-1, 0, 3], z=(0, 1),
                              topology=(grids.BOUNDED, grids.PERIODIC, grids.BOUNDED))
    assert (g.Nx, g.Ny, g.Nz) == (3, 2, 1)
    assert list(grids.xfaces(g)) == [0.0, 0.25, 1.0, 2.0]
    assert list(grids.yfaces(g)) == [-1.0, 0.0, 3.0]
    assert list(grids.zfaces(g)) == [0.0, 1.0]
    assert list(grids.xnodes(g)) == [0.125, 0.625, 1.5]
    assert list(grids.xspacings(g)) == [0.25, 0.75, 1.0]
    assert tuple(g.topology) == (grids.BOUNDED, grids.PERIODIC, grids.BOUNDED)
