"""ctypes front end of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product package never does.  `OracleModel` drives the
oracle from the same Workload inputs the device model takes.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RIPPLE_ORACLE_LIB selects another build of the same sources (the
# address/undefined-behaviour sanitizer build, `make -C oracle sanitize`).
LIB_PATH = os.environ.get("RIPPLE_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_D = ctypes.c_double
_F = ctypes.c_float
_I = ctypes.c_int
_P = ctypes.c_void_p
_lib = None

MODE_NONE, MODE_INTRINSIC, MODE_REFRACTION, MODE_PSEUDOMOMENTUM = 0, 1, 2, 3
STEPPERS = {"RK3": 0, "LSRK3": 1, "LowStorageRK3": 1, "ForwardEuler": 2}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_q_cdf.restype = _D
        _lib.orc_q_cdf.argtypes = [_D, _D, _D, _D]
        _lib.orc_q_cdf_dkappa.restype = _D
        _lib.orc_q_cdf_dkappa.argtypes = [_D, _D, _D, _D]
        for f in ("orc_q_cell_integral", "orc_q_cell_integral_dkappa"):
            getattr(_lib, f).restype = _D
            getattr(_lib, f).argtypes = [_D, _D, _D, _D]
        _lib.orc_total_weighted_f64.restype = _D
        _lib.orc_total_weighted_f32.restype = _D
        _lib.orc_weno5_f64.restype = _D
        _lib.orc_weno5_f64.argtypes = [_D] * 5
        _lib.orc_weno5_f32.restype = _F
        _lib.orc_weno5_f32.argtypes = [_F] * 5
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _c(a, dtype=np.float64):
    """Contiguous i-fastest copy of an array given in Julia index order."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))


def set_threads(n):
    return lib().orc_set_threads(_I(n))


def weno5(f, dtype=np.float64):
    L = lib()
    if dtype == np.float64:
        return L.orc_weno5_f64(*[float(x) for x in f])
    return L.orc_weno5_f32(*[float(x) for x in f])


def centers_to_faces(c, lower_bounded=False, lower=0.0):
    c = _c(c)
    f = np.zeros(len(c) + 1)
    lib().orc_centers_to_faces(_I(len(c)), _p(c), _p(f), _I(int(lower_bounded)), _D(lower))
    return f


def phi_faces(phi):
    phi = _c(phi)
    f = np.zeros(len(phi) + 1)
    lib().orc_phi_faces(_I(len(phi)), _p(phi), _p(f))
    return f


def geometry(kc, kf, pc, pf, finite_depth=False, depth=0.0, gravity=9.81):
    """Tables (Nk, Nphi): w, mu_x, mu_y, trace, cgx, cgy."""
    kc, kf, pc, pf = _c(kc), _c(kf), _c(pc), _c(pf)
    nk, npf = len(kc), len(pc)
    out = {k: np.zeros(nk * npf) for k in ("w", "mu_x", "mu_y", "trace", "cgx", "cgy")}
    L = lib()
    L.orc_cell_measures(_I(nk), _I(npf), _p(kf), _p(pf), _p(out["w"]))
    L.orc_moment_measures(_I(nk), _I(npf), _p(kf), _p(pf), _p(out["mu_x"]), _p(out["mu_y"]), _p(out["trace"]))
    L.orc_group_velocity_tables(_I(nk), _I(npf), _p(kc), _p(kf), _p(pc), _p(pf), _p(out["w"]),
                                _I(int(finite_depth)), _D(depth), _D(gravity), _p(out["cgx"]), _p(out["cgy"]))
    return {k: v.reshape((nk, npf), order="F") for k, v in out.items()}


def energy_measures(kf, pf, gravity=9.81):
    """bulk_statistics.jl:23-26: sqrt(g) * the annular measure of kappa^(1/2),
    (Nk, Nphi)."""
    kf, pf = _c(kf), _c(pf)
    nk, npf = len(kf) - 1, len(pf) - 1
    e = np.zeros(nk * npf)
    lib().orc_energy_measures(_I(nk), _I(npf), _p(kf), _p(pf), _D(gravity), _p(e))
    return e.reshape((nk, npf), order="F")


def q_cdf(kappa, z, depth, small_mu=1e-6):
    return lib().orc_q_cdf(kappa, z, depth, small_mu)


def q_weights(z_faces, kappa, depth):
    zf, kc = _c(z_faces), _c(kappa)
    W = np.zeros((len(zf) - 1) * len(kc))
    lib().orc_q_weights(_I(len(zf) - 1), _p(zf), _I(len(kc)), _p(kc), _D(depth), _p(W))
    return W.reshape((len(zf) - 1, len(kc)), order="F")


def doppler(u, v, W):
    nx, ny, nz = u.shape
    nk = W.shape[1]
    uf, vf, Wf = _c(u), _c(v), _c(W)
    Ux = np.zeros(nx * ny * nk)
    Uy = np.zeros(nx * ny * nk)
    lib().orc_doppler(_I(nx), _I(ny), _I(nz), _p(uf), _p(vf), _I(nk), _p(Wf), _p(Ux), _p(Uy))
    return Ux.reshape((nx, ny, nk), order="F"), Uy.reshape((nx, ny, nk), order="F")


def group_velocity_field(kappa, phi, depth, gravity=9.81):
    """(cgx, cgy) of shape (Nx, Ny, Nk, Nphi): the point-sampled finite-depth
    group velocity of every column at its own depth (transport.jl:111-139)."""
    d = np.asarray(depth, dtype=np.float64)
    nx, ny = d.shape
    kc, pc = _c(kappa), _c(phi)
    nk, npf = len(kc), len(pc)
    cgx, cgy = np.zeros(nx * ny * nk * npf), np.zeros(nx * ny * nk * npf)
    df = _c(d)  # held: a temporary would be freed before the call
    lib().orc_group_velocity_field(_I(nx), _I(ny), _I(nk), _I(npf), _p(kc), _p(pc), _p(df), _D(gravity),
                                   _p(cgx), _p(cgy))
    shape = (nx, ny, nk, npf)
    return cgx.reshape(shape, order="F"), cgy.reshape(shape, order="F")


def q_weights_columns(z_faces, kappa, depth):
    """Q weights of every column at its own depth, shape (Nx, Ny, Nz, Nk)."""
    d = np.asarray(depth, dtype=np.float64)
    nx, ny = d.shape
    zf, kc = _c(z_faces), _c(kappa)
    nz, nk = len(zf) - 1, len(kc)
    W = np.zeros(nx * ny * nz * nk)
    df = _c(d)
    lib().orc_q_weights_columns(_I(nx * ny), _I(nz), _p(zf), _I(nk), _p(kc), _p(df), _p(W))
    return W.reshape((nx, ny, nz, nk), order="F")


def doppler_columns(u, v, Wc):
    """U[i,j,m] = sum_k u[i,j,k] Wc[i,j,k,m], sequential over k."""
    nx, ny, nz = u.shape
    nk = Wc.shape[3]
    Ux, Uy = np.zeros(nx * ny * nk), np.zeros(nx * ny * nk)
    uf, vf, wf = _c(u), _c(v), _c(Wc)
    lib().orc_doppler_columns(_I(nx), _I(ny), _I(nz), _p(uf), _p(vf), _I(nk), _p(wf), _p(Ux), _p(Uy))
    return Ux.reshape((nx, ny, nk), order="F"), Uy.reshape((nx, ny, nk), order="F")


def q_cell_integral(kappa, z1, z2, depth):
    return lib().orc_q_cell_integral(kappa, z1, z2, depth)


def q_cdf_dkappa(kappa, z, depth, small_mu=1e-6):
    return lib().orc_q_cdf_dkappa(kappa, z, depth, small_mu)


def q_cell_integral_dkappa(kappa, z1, z2, depth):
    return lib().orc_q_cell_integral_dkappa(kappa, z1, z2, depth)


def pm_overlap(z_faces, kappa, depth):
    """Pseudomomentum overlap and its kappa derivative, each (Nk, Nk) [target, source]."""
    zf, kc = _c(z_faces), _c(kappa)
    nk = len(kc)
    O, dO = np.zeros(nk * nk), np.zeros(nk * nk)
    lib().orc_pm_overlap(_I(len(zf) - 1), _p(zf), _I(nk), _p(kc), _D(depth), _p(O), _p(dO))
    return O.reshape((nk, nk), order="F"), dO.reshape((nk, nk), order="F")


def pm_overlap_columns(z_faces, kappa, depth):
    """The overlap tables of every column at its depth (an (Nx, Ny) array),
    each (Nx, Ny, Nk, Nk) [i, j, target, source]."""
    d = np.asarray(depth, dtype=np.float64)
    nx, ny = d.shape
    zf, kc, df = _c(z_faces), _c(kappa), _c(d)
    nk = len(kc)
    O, dO = np.zeros(nx * ny * nk * nk), np.zeros(nx * ny * nk * nk)
    lib().orc_pm_overlap_columns(_I(nx * ny), _I(len(zf) - 1), _p(zf), _I(nk), _p(kc), _p(df), _p(O), _p(dO))
    shape = (nx, ny, nk, nk)
    return O.reshape(shape, order="F"), dO.reshape(shape, order="F")


def pm_doppler(N, O, dO, kx, ky, dtype=np.float64):
    """Ux, Uy, dUx/dkappa, dUy/dkappa (each (Nx, Ny, Nk)) of the pseudomomentum
    coupling; O, dO one (Nk, Nk) table or one per column (Nx, Ny, Nk, Nk)."""
    nx, ny, nk, npf = N.shape
    Nf, Of, dOf, kxf, kyf = (_c(a, dtype) for a in (N, O, dO, kx, ky))
    out = [np.zeros(nx * ny * nk, dtype=dtype) for _ in range(4)]
    sfx = "f64" if np.dtype(dtype) == np.float64 else "f32"
    lib().orc_set_pm_columns(_I(int(np.ndim(O) == 4)))
    getattr(lib(), f"orc_pm_doppler_{sfx}")(_I(nx), _I(ny), _I(nk), _I(npf), _p(Nf), _p(Of), _p(dOf), _p(kxf),
                                            _p(kyf), *[_p(a) for a in out])
    return [a.reshape((nx, ny, nk), order="F") for a in out]


def pm_cells(N, W, z_faces, kx, ky, cell_average=True, dtype=np.float64):
    """Pseudomomentum cell averages (or integrals) px, py, each (Nx, Ny, Nz);
    W one (Nz, Nk) table or one per column (Nx, Ny, Nz, Nk)."""
    nx, ny, nk, npf = N.shape
    nz = W.shape[2] if np.ndim(W) == 4 else W.shape[0]
    Nf, Wf, kxf, kyf = (_c(a, dtype) for a in (N, W, kx, ky))
    zf = _c(z_faces)
    px, py = np.zeros(nx * ny * nz, dtype=dtype), np.zeros(nx * ny * nz, dtype=dtype)
    sfx = "f64" if np.dtype(dtype) == np.float64 else "f32"
    lib().orc_set_pm_columns(_I(int(np.ndim(W) == 4)))
    getattr(lib(), f"orc_pm_cells_{sfx}")(_I(nx), _I(ny), _I(nk), _I(npf), _I(nz), _p(Nf), _p(Wf), _p(zf), _p(kxf),
                                          _p(kyf), _I(int(cell_average)), _p(px), _p(py))
    return px.reshape((nx, ny, nz), order="F"), py.reshape((nx, ny, nz), order="F")


def gse_tables(kappa, dtype=np.float64):
    """SpatialAveraging's per-kappa group speed and its half neighbour difference."""
    t = np.dtype(dtype)
    k = _c(kappa, t)
    cg, dcg = np.zeros(len(k), dtype=t), np.zeros(len(k), dtype=t)
    sfx = "f64" if t == np.float64 else "f32"
    getattr(lib(), f"orc_gse_tables_{sfx}")(_I(len(k)), _p(k), _p(cg), _p(dcg))
    return cg, dcg


def spatial_averaging(N, kappa, phi, dphi, dx, dy, dt, alpha_s=0.5, alpha_n=0.5, dtype=np.float64,
                      periodic=(True, True)):
    """One SpatialAveraging pass (Models/propagation_smoothing.jl:160-199) of N (Nx, Ny, Nk, Nphi);
    periodic = (x, y): a bounded axis clamps the corner indices to its edge cells."""
    t = np.dtype(dtype)
    nx, ny, nk, npf = N.shape
    cg, dcg = gse_tables(kappa, t)
    cs = _c([math.cos(p) for p in phi], t)
    sn = _c([math.sin(p) for p in phi], t)
    Nin = _c(N, t)
    out = np.zeros_like(Nin)
    sc = _D if t == np.float64 else _F
    one = t.type(1)
    getattr(lib(), f"orc_spatial_averaging_{'f64' if t == np.float64 else 'f32'}")(
        _I(nx), _I(ny), _I(nk), _I(npf), _p(Nin), _p(out), _p(cg), _p(dcg), _p(cs), _p(sn), sc(dphi),
        sc(alpha_s), sc(alpha_n), sc(dt), sc(one / t.type(dx)), sc(one / t.type(dy)), _I(int(bool(periodic[0]))),
        _I(int(bool(periodic[1]))))
    return out.reshape((nx, ny, nk, npf), order="F")


class OracleModel:
    """The oracle's version of SpectralWaveModel over a Workload's inputs."""

    def __init__(self, workload, dtype=np.float64, depth=None, gravity=9.81, doppler_override=None,
                 transport=None, timestepper=None, pseudomomentum=None, smoothing=None):
        wl = workload
        self.dtype = np.dtype(dtype)
        g, sg = wl.grid, wl.spectral_grid
        self.size = (g.Nx, g.Ny, len(sg.kappa), len(sg.phi))
        kw = wl.model_kwargs
        self.stepper = STEPPERS[timestepper or kw.get("timestepper", "ForwardEuler")]
        if transport is not None:
            self.mode = transport
        elif wl.velocities is not None:
            self.mode = MODE_REFRACTION
        elif kw.get("horizontal_advection", "WENO") is None:
            self.mode = MODE_NONE
        else:
            self.mode = MODE_INTRINSIC
        finite = depth is not None
        # depth: None (deep water), a number, or an (Nx, Ny) array (a
        # spatially varying depth: group velocity and Q depth per column)
        field = depth is not None and np.ndim(depth) == 2
        tabs = geometry(sg.kappa, sg.kappa_faces, sg.phi, sg.phi_faces, finite, 1.0 if field else (depth or 0.0),
                        gravity)
        if field:
            tabs["cgx"], tabs["cgy"] = group_velocity_field(sg.kappa, sg.phi, depth, gravity)
        self._cg_cells = field
        self.tables = tabs
        self.dx = float(g.x_faces[1] - g.x_faces[0])
        self.dy = float(g.y_faces[1] - g.y_faces[0])
        self.dk = float(sg.kappa_faces[1] - sg.kappa_faces[0])
        self.dp = float(sg.phi_faces[1] - sg.phi_faces[0])
        t = self.dtype
        self.kappa = _c(sg.kappa, t)
        self.cgx, self.cgy = _c(tabs["cgx"], t), _c(tabs["cgy"], t)
        # libm cos/sin, as the device library forms its tables on the host.
        self.cos = _c([math.cos(p) for p in sg.phi], t)
        self.sin = _c([math.sin(p) for p in sg.phi], t)
        nx, ny, nk, _ = self.size
        if doppler_override is not None:
            Ux, Uy = doppler_override
        elif wl.velocities is not None:
            q_depth = g.Lz if depth is None else depth
            if field:
                W = q_weights_columns(g.z_faces, sg.kappa, depth)
            else:
                W = q_weights(g.z_faces, sg.kappa, q_depth)
            if t == np.float32:
                W = W.astype(np.float32).astype(np.float64)
                u = wl.velocities[0].astype(np.float32).astype(np.float64)
                v = wl.velocities[1].astype(np.float32).astype(np.float64)
            else:
                u, v = wl.velocities
            Ux, Uy = doppler_columns(u, v, W) if field else doppler(u, v, W)
        else:
            Ux = Uy = np.zeros((nx, ny, nk))
        self.Ux, self.Uy = _c(Ux, t), _c(Uy, t)
        # Self-coupled pseudomomentum current: pseudomomentum=(z_faces, q_depth)
        # (Coupling/current_coupling.jl:93-121); U follows every stage input.
        self.pm = None
        self._pm_cols = False
        if pseudomomentum is not None:
            zf, q_depth = pseudomomentum
            if field:
                # pseudomomentum.jl:72-87: one overlap table per column
                O, _ = pm_overlap_columns(zf, sg.kappa, depth)
                self._pm_cols = True
            else:
                O, _ = pm_overlap(zf, sg.kappa, q_depth)
            self.pm = [_c(O, t), _c(tabs["mu_x"], t), _c(tabs["mu_y"], t)]
            self.mode = MODE_PSEUDOMOMENTUM
            self.Ux, self.Uy = np.zeros(nx * ny * nk, dtype=t), np.zeros(nx * ny * nk, dtype=t)
        # SpatialAveraging after every step: smoothing=(alpha_s, alpha_n),
        # clamped on the grid's bounded axes (propagation_smoothing.jl:175-177).
        self.smoothing = smoothing
        self._sg = sg
        topo = getattr(g, "topology", ("Periodic", "Periodic"))
        self._periodic = (topo[0] == "Periodic", topo[1] == "Periodic")
        self.N = _c(wl.N0, t).copy()  # never alias the workload's array
        self.time, self.iteration = 0.0, 0

    def _inv(self):
        t = np.float64 if self.dtype == np.float64 else np.float32
        one = t(1)
        return [one / t(self.dx), one / t(self.dy), one / t(self.dk), one / t(self.dp)]

    def _scalar(self, x):
        return _D(x) if self.dtype == np.float64 else _F(x)

    def step(self, dt, nsteps=1, hook=None, user=None):
        if self.smoothing is not None:
            # apply_propagation_smoothing! closes every step (time_step.jl:167).
            for _ in range(nsteps):
                self._step(dt, 1, hook, user)
                sg = self._sg
                self.N = _c(spatial_averaging(self.action, sg.kappa, sg.phi, self.dp, self.dx, self.dy, dt,
                                              *self.smoothing, dtype=self.dtype, periodic=self._periodic),
                                    self.dtype)
            return
        self._step(dt, nsteps, hook, user)

    def _step(self, dt, nsteps=1, hook=None, user=None):
        L = lib()
        L.orc_set_cg_cells(_I(int(self._cg_cells)))
        L.orc_set_pm_columns(_I(int(self._pm_cols)))
        nx, ny, nk, npf = self.size
        sfx = "f64" if self.dtype == np.float64 else "f32"
        fn = getattr(L, f"orc_run_{sfx}")
        inv = [self._scalar(v) for v in self._inv()]
        rc = fn(_I(nx), _I(ny), _I(nk), _I(npf), _I(self.mode), _I(self.stepper), _p(self.N),
                _p(self.Ux), _p(self.Uy), _p(self.kappa), _p(self.cgx), _p(self.cgy), _p(self.cos), _p(self.sin),
                *inv, _D(dt), _I(nsteps), hook, user, *self._pm_args())
        if rc != 0:
            raise MemoryError("oracle run failed to allocate")
        self.time += dt * nsteps
        self.iteration += nsteps

    def tendency(self):
        L = lib()
        L.orc_set_cg_cells(_I(int(self._cg_cells)))
        L.orc_set_pm_columns(_I(int(self._pm_cols)))
        nx, ny, nk, npf = self.size
        sfx = "f64" if self.dtype == np.float64 else "f32"
        G = np.zeros_like(self.N)
        inv = [self._scalar(v) for v in self._inv()]
        getattr(L, f"orc_tendency_{sfx}")(_I(nx), _I(ny), _I(nk), _I(npf), _I(self.mode), _p(self.N), _p(G),
                                          _p(self.Ux), _p(self.Uy), _p(self.kappa), _p(self.cgx), _p(self.cgy),
                                          _p(self.cos), _p(self.sin), *inv, *self._pm_args())
        return G.reshape(self.size, order="F")

    def _pm_args(self):
        return [_p(a) for a in self.pm] if self.pm is not None else [None, None, None]

    def doppler_velocity(self):
        """(Ux, Uy): the coupling's velocity (in mode 3, of the last tendency input)."""
        nx, ny, nk, _ = self.size
        return self.Ux.reshape((nx, ny, nk), order="F"), self.Uy.reshape((nx, ny, nk), order="F")

    @property
    def action(self):
        return self.N.reshape(self.size, order="F")

    def diagnostics(self):
        nx, ny, nk, npf = self.size
        t = self.tables
        w, mx, my, tr = (_c(t[k]) for k in ("w", "mu_x", "mu_y", "trace"))
        m0, kr, md = np.zeros(nx * ny), np.zeros(nx * ny), np.zeros(nx * ny)
        sfx = "f64" if self.dtype == np.float64 else "f32"
        getattr(lib(), f"orc_column_sums_{sfx}")(_I(nx), _I(ny), _I(nk), _I(npf), _p(self.N), _p(w), _p(mx), _p(my),
                                                 _p(tr), _p(m0), _p(kr), _p(md))
        shp = (nx, ny)
        return m0.reshape(shp, order="F"), kr.reshape(shp, order="F"), md.reshape(shp, order="F")

    def total(self, measure):
        nx, ny, nk, npf = self.size
        sfx = "f64" if self.dtype == np.float64 else "f32"
        w = _c(measure)
        return getattr(lib(), f"orc_total_weighted_{sfx}")(_I(nx), _I(ny), _I(nk), _I(npf), _p(self.N), _p(w),
                                                           _D(self.dx), _D(self.dy))


# --- source terms (FP64 oracle only) -----------------------------------------

class OrcSources(ctypes.Structure):
    """Mirror of orc_sources (oracle/sources_impl.h); same layout as rpl_sources."""
    _fields_ = [(n, ctypes.c_int) for n in ("wind_input", "drag", "wind_field", "stress_cap", "dissipation",
                                             "rotation")] + \
               [(n, ctypes.c_double) for n in (
                   "wind_speed", "wind_direction",
                   "beta_max", "z_alpha", "p_in", "alpha0", "rho_air", "rho_water", "von_karman", "gravity",
                   "cap_epsilon",
                   "vmax", "rmax", "radius", "shape", "inflow_angle", "background_x", "background_y",
                   "track_t0", "track_t1", "track_x0", "track_y0", "track_x1", "track_y1",
                   "cds", "delta1", "delta2", "p_mean",
                   "ls_cds", "ls_br", "ls_p", "ls_sigma_power",
                   "wc_rate", "wc_threshold", "wc_saturation_power", "wc_wavenumber_power",
                   "wc_reference_wavenumber")] + \
               [("track_n", ctypes.c_int), ("track_t", ctypes.c_void_p), ("track_x", ctypes.c_void_p),
                ("track_y", ctypes.c_void_p),
                ("nonlinear", ctypes.c_int), ("dia_coefficient", ctypes.c_double),
                ("dia_frequency_offset", ctypes.c_double), ("dia_gravity", ctypes.c_double)]


def set_track(src, times, centers):
    """Give OrcSources an N-point LinearStormTrack (keeps the arrays alive on src)."""
    src._track = [_c(times), _c([p[0] for p in centers]), _c([p[1] for p in centers])]
    src.track_n = len(times)
    src.track_t, src.track_x, src.track_y = [a.ctypes.data for a in src._track]
    return src


def track_at(times, centers, t):
    T, X, Y = _c(times), _c([p[0] for p in centers]), _c([p[1] for p in centers])
    cx, cy = ctypes.c_double(), ctypes.c_double()
    lib().orc_track_at(_I(len(T)), _p(T), _p(X), _p(Y), _D(t), ctypes.byref(cx), ctypes.byref(cy))
    return cx.value, cy.value


class OrcSourceCtx(ctypes.Structure):
    _fields_ = [("src", ctypes.POINTER(OrcSources)), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nk", ctypes.c_int), ("nphi", ctypes.c_int), ("x_nodes", ctypes.c_void_p),
                ("y_nodes", ctypes.c_void_p), ("kappa", ctypes.c_void_p), ("phi", ctypes.c_void_p),
                ("w", ctypes.c_void_p), ("time", ctypes.c_double),
                ("kappa_faces", ctypes.c_void_p), ("phi_faces", ctypes.c_void_p)]


def sources_from(struct):
    """Copy any ctypes structure with the rpl_sources field names into OrcSources."""
    s = OrcSources()
    for name, _ in OrcSources._fields_:
        if hasattr(struct, name):
            setattr(s, name, getattr(struct, name))
    return s


def dia_offsets(frequency_offset=0.25):
    dp, dm = ctypes.c_double(), ctypes.c_double()
    lib().orc_dia_offsets(_D(frequency_offset), ctypes.byref(dp), ctypes.byref(dm))
    return dp.value, dm.value


def dia_transfer(N, kappa, phi, coefficient=2.78e7, frequency_offset=0.25, gravity=9.81):
    """SymmetricQuadruplet transfer T (Nx, Ny, Nk, Nphi) of an action field (FP64)."""
    nx, ny, nk, npf = N.shape
    Nf, kc, pc = _c(N), _c(kappa), _c(phi)
    T = np.zeros_like(Nf)
    lib().orc_dia_transfer(_I(nx), _I(ny), _I(nk), _I(npf), _p(Nf), _p(kc), _p(pc), _D(coefficient),
                           _D(frequency_offset), _D(gravity), _p(T))
    return T.reshape((nx, ny, nk, npf), order="F")


class OrcListedCtx(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nk", ctypes.c_int), ("nphi", ctypes.c_int),
                ("count", ctypes.c_int), ("bins", ctypes.c_void_p), ("rates", ctypes.c_void_p),
                ("power", ctypes.c_double), ("w", ctypes.c_void_p), ("rate_field", ctypes.c_void_p)]


def listed_hook(size, interactions, power, w, rate_field=None):
    """(hook, user) for OracleModel.step/tendency: ListedQuadruplets;
    interactions = [((d1m, d1n), (d2m, d2n), (r1m, r1n), (r2m, r2n), rate)], 0-based, with constant rates,
    or per-column rates rate_field of shape (Nx, Ny, count)."""
    nx, ny, nk, npf = size
    bins = np.ascontiguousarray([v for it in interactions for pair in it[:4] for v in pair], dtype=np.int32)
    rates = _c([it[4] for it in interactions])
    wf = _c(w)
    rf = None if rate_field is None else _c(np.asarray(rate_field, dtype=np.float64).reshape(
        (nx, ny, len(interactions)), order="F"))
    ctx = OrcListedCtx(nx, ny, nk, npf, len(interactions), bins.ctypes.data, rates.ctypes.data, float(power),
                       wf.ctypes.data, None if rf is None else rf.ctypes.data)
    ctx._keep = (bins, rates, wf, rf)
    hook = ctypes.cast(lib().orc_listed_hook_f64, ctypes.c_void_p)
    return hook, ctypes.cast(ctypes.pointer(ctx), ctypes.c_void_p), ctx


def wind_columns(src, x_nodes, y_nodes, t):
    xn, yn = _c(x_nodes), _c(y_nodes)
    nx, ny = len(xn), len(yn)
    out = [np.zeros(nx * ny) for _ in range(3)]
    lib().orc_wind_columns(ctypes.byref(src), _I(nx), _I(ny), _p(xn), _p(yn), _D(t), *[_p(a) for a in out])
    return [a.reshape((nx, ny), order="F") for a in out]


class OracleSourceModel(OracleModel):
    """OracleModel plus the FP64 source terms of the PrecomputedSources bundle."""

    def __init__(self, workload, sources, **kw):
        super().__init__(workload, dtype=np.float64, **kw)
        g = workload.grid
        self._src = sources
        self._xn = _c(0.5 * (g.x_faces[1:] + g.x_faces[:-1]))
        self._yn = _c(0.5 * (g.y_faces[1:] + g.y_faces[:-1]))
        sg = workload.spectral_grid
        self._kc, self._pc = _c(sg.kappa), _c(sg.phi)
        self._w = _c(self.tables["w"])
        self._kf, self._pf = _c(sg.kappa_faces), _c(sg.phi_faces)
        nx, ny, nk, npf = self.size
        self._ctx = OrcSourceCtx(ctypes.pointer(self._src), nx, ny, nk, npf, self._xn.ctypes.data,
                                 self._yn.ctypes.data, self._kc.ctypes.data, self._pc.ctypes.data,
                                 self._w.ctypes.data, 0.0, self._kf.ctypes.data, self._pf.ctypes.data)
        self._hook = ctypes.cast(lib().orc_sources_hook_f64, ctypes.c_void_p)

    def step(self, dt, nsteps=1, hook=None, user=None):
        for _ in range(nsteps):
            self._ctx.time = self.time
            super().step(dt, 1, self._hook, ctypes.cast(ctypes.pointer(self._ctx), ctypes.c_void_p))

    def tendency(self):
        G = super().tendency()
        Gf = _c(G)
        self._ctx.time = self.time
        lib().orc_sources_hook_f64(_p(self.N), _p(Gf), ctypes.pointer(self._ctx))
        return Gf.reshape(self.size, order="F")
