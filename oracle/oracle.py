"""ctypes binding of the oracle (oracle.cpp) and of the reference harness
(_ref/libref.so, built from the reference's own CPU sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the test oracle or as the CPU baseline. The
product library never loads it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

# Layers of the reference's default two-tier stack (stack3d.h:153-162):
# (thickness m, k W/mK, rho_c J/m^3K, planes, active)
DEFAULT_STACK = [
    (50e-6, 130.0, 1.66e6, 2, 1),
    (10e-6, 2.0, 1.70e6, 1, 0),
    (50e-6, 130.0, 1.66e6, 2, 1),
    (25e-6, 5.0, 2.00e6, 1, 0),
    (1e-3, 400.0, 3.45e6, 4, 0),
]

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _i(a):
    return a.ctypes.data_as(_ip)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.or_model_new.restype = _vp
        L.or_model_new.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _ip, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_double]
        L.or_model_free.argtypes = [_vp]
        L.or_model_dims.argtypes = [_vp, _ip, _ip, _ip, _dp, _dp]
        L.or_model_add_power_block.argtypes = [_vp, C.c_int] + [C.c_double] * 5
        L.or_model_add_uniform_power.argtypes = [_vp, C.c_int, C.c_double]
        L.or_model_add_hotspot.argtypes = [_vp, C.c_int] + [C.c_double] * 4
        L.or_model_add_hotspot.restype = C.c_int
        L.or_model_add_tsv_region.argtypes = [_vp, C.c_int] + [C.c_double] * 5
        L.or_model_finalize.argtypes = [_vp]
        L.or_model_get_f.argtypes = [_vp, C.c_int, _fp]
        L.or_model_get_d.argtypes = [_vp, C.c_int, _dp]
        L.or_model_set_P.argtypes = [_vp, _dp]
        L.or_model_total_power.argtypes = [_vp]
        L.or_model_total_power.restype = C.c_double
        L.or_model_analytic_1d.argtypes = [_vp, _dp]
        L.or_eff_kz.argtypes = [C.c_double] * 3
        L.or_eff_kz.restype = C.c_double
        L.or_eff_kxy.argtypes = [C.c_double] * 3
        L.or_eff_kxy.restype = C.c_double
        L.or_solver_new.restype = _vp
        L.or_solver_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _fp, _fp, _fp,
                                    _fp, _fp, _dp, C.c_int, C.c_int, C.c_int]
        L.or_solver_free.argtypes = [_vp]
        L.or_solver_set_dt.argtypes = [_vp, C.c_double]
        for f in ("or_solver_set_x_zero", "or_solver_set_rhs_steady", "or_solver_set_rhs_transient"):
            getattr(L, f).argtypes = [_vp]
        L.or_solver_set_x.argtypes = [_vp, _dp]
        L.or_solver_set_power.argtypes = [_vp, _dp]
        L.or_solver_solve.argtypes = [_vp, C.c_double, C.c_int, _ip, _dp]
        L.or_solver_solve.restype = C.c_int
        L.or_solver_get_x.argtypes = [_vp, _dp]
        L.or_solver_precond.argtypes = [_vp, _dp, _dp]
        L.or_solver_apply.argtypes = [_vp, _dp, _dp]
        L.or_solver_nlevels.argtypes = [_vp]
        L.or_solver_nlevels.restype = C.c_int
        L.or_solver_level_dims.argtypes = [_vp, C.c_int, _ip, _ip, _ip]
        L.or_solver_level_get.argtypes = [_vp, C.c_int, C.c_int, _fp]
        L.or_solver_level_factors.argtypes = [_vp, C.c_int, _dp, _dp]
        L.or_solver_set_fast_restrict.argtypes = [_vp, C.c_int]
        L.or_true_relres.argtypes = [C.c_int] * 3 + [_fp] * 5 + [_dp, _dp]
        L.or_true_relres.restype = C.c_double
        L.or_true_relres_flux.argtypes = [C.c_int] * 3 + [_fp] * 5 + [_dp, _dp]
        L.or_true_relres_flux.restype = C.c_double
        L.or_heat_out.argtypes = [C.c_int] * 3 + [_fp, _fp, _dp]
        L.or_heat_out.restype = C.c_double
        L.or_heat_stored.argtypes = [C.c_int] * 3 + [_fp, _dp, _dp, C.c_double]
        L.or_heat_stored.restype = C.c_double
        L.or_run_transient.restype = C.c_int
        L.or_run_transient.argtypes = ([C.c_int] * 3 + [C.c_double] * 2 + [_fp] * 5 + [_dp] +
                                       [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int, C.c_double, C.c_int,
                                                                           _dp, C.c_int, _dp, _ip])
        _lib = L
    return _lib


class Model:
    """Oracle model (stack3d.h:45-148 restated): build, add power/TSVs,
    finalize, then read the arrays. Arrays are in the natural layout
    (x fastest, then y, then z)."""

    def __init__(self, layers, nx, ny, Lx=10e-3, Ly=10e-3, h_top=2e4, h_bot=0.0):
        L = lib()
        self.layers = list(layers)
        nl = len(self.layers)
        th = np.array([l[0] for l in layers], np.float64)
        k = np.array([l[1] for l in layers], np.float64)
        rc = np.array([l[2] for l in layers], np.float64)
        nc = np.array([l[3] for l in layers], np.int32)
        ac = np.array([l[4] for l in layers], np.int32)
        self.h = L.or_model_new(nl, _d(th), _d(k), _d(rc), _i(nc), _i(ac), nx, ny, Lx, Ly, h_top, h_bot)
        if not self.h:
            raise ValueError("too many z-planes")
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        dx, dy = C.c_double(), C.c_double()
        L.or_model_dims(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(dx), C.byref(dy))
        self.nx, self.ny, self.nz, self.dx, self.dy = a.value, b.value, c.value, dx.value, dy.value
        self.Lx, self.Ly, self.h_top, self.h_bot = Lx, Ly, h_top, h_bot

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_model_free(self.h)
            self.h = None

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def add_power_block(self, layer, x0, y0, w, h, watts):
        lib().or_model_add_power_block(self.h, layer, x0, y0, w, h, watts)

    def add_uniform_power(self, layer, watts):
        lib().or_model_add_uniform_power(self.h, layer, watts)

    def add_hotspot(self, layer, xc, yc, size, watts):
        if lib().or_model_add_hotspot(self.h, layer, xc, yc, size, watts) != 0:
            raise ValueError("hotspot covers no cells")

    def add_tsv_region(self, layer, x0, y0, w, h, f):
        lib().or_model_add_tsv_region(self.h, layer, x0, y0, w, h, f)

    def finalize(self):
        lib().or_model_finalize(self.h)
        return self

    def arrays(self):
        """dict of gE, gN, gU, cap (float32, n), gB (float32, nx*ny), P (float64, n)."""
        out = {}
        for w, name in enumerate(("gE", "gN", "gU", "gB", "cap")):
            a = np.empty(self.nx * self.ny if name == "gB" else self.n, np.float32)
            lib().or_model_get_f(self.h, w, _f(a))
            out[name] = a
        for w, name in ((5, "P"), (6, "tsv_frac")):
            a = np.empty(self.n, np.float64)
            lib().or_model_get_d(self.h, w, _d(a))
            out[name] = a
        dz = np.empty(self.nz, np.float64)
        lib().or_model_get_d(self.h, 7, _d(dz))
        out["dz"] = dz
        out.update(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy)
        return out

    def set_P(self, P):
        P = np.ascontiguousarray(P, np.float64)
        lib().or_model_set_P(self.h, _d(P))

    def total_power(self):
        return lib().or_model_total_power(self.h)

    def analytic_1d(self):
        T = np.empty(self.nz, np.float64)
        lib().or_model_analytic_1d(self.h, _d(T))
        return T


def eff_kz(k, kc, f):
    return lib().or_eff_kz(k, kc, f)


def eff_kxy(k, kc, f):
    return lib().or_eff_kxy(k, kc, f)


def default_model(nx, ny=None):
    return Model(DEFAULT_STACK, nx, ny or nx)


def hotspot_model(n, xc=3e-3, yc=3e-3):
    """main3d.cu:60-67 (hotspot at (3,3) mm); config 0 of the benchmark uses (5,5) mm."""
    m = default_model(n)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, xc, yc, 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    return m.finalize()


PRECOND = {"jacobi": 0, "mg": 1, "multigrid": 1}


class Solver:
    """Oracle CpuPcg (pcg3d_cpu.cpp restated). mg_double selects the V-cycle
    precision (False: float, the reference default MgReal)."""

    def __init__(self, arrays, precond="mg", mg_double=False, flexible=False, fast_restrict=False):
        a = arrays
        self.nx, self.ny, self.nz = a["nx"], a["ny"], a["nz"]
        self.n = self.nx * self.ny * self.nz
        self._keep = [np.ascontiguousarray(a[k], np.float32) for k in ("gE", "gN", "gU", "gB", "cap")]
        self._P = np.ascontiguousarray(a["P"], np.float64)
        self.h = lib().or_solver_new(self.nx, self.ny, self.nz, a["dx"], a["dy"],
                                     *[_f(v) for v in self._keep], _d(self._P), PRECOND[precond],
                                     int(mg_double), int(flexible))
        lib().or_solver_set_fast_restrict(self.h, int(fast_restrict))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_solver_free(self.h)
            self.h = None

    def set_dt(self, dt):
        lib().or_solver_set_dt(self.h, dt)

    def set_x_zero(self):
        lib().or_solver_set_x_zero(self.h)

    def set_x(self, x):
        x = np.ascontiguousarray(x, np.float64)
        lib().or_solver_set_x(self.h, _d(x))

    def set_rhs_steady(self):
        lib().or_solver_set_rhs_steady(self.h)

    def set_rhs_transient(self):
        lib().or_solver_set_rhs_transient(self.h)

    def set_power(self, P):
        P = np.ascontiguousarray(P, np.float64)
        lib().or_solver_set_power(self.h, _d(P))

    def solve(self, tol, maxit):
        it, rr = C.c_int(), C.c_double()
        conv = lib().or_solver_solve(self.h, tol, maxit, C.byref(it), C.byref(rr))
        return dict(iters=it.value, relres=rr.value, converged=bool(conv))

    def get_x(self):
        x = np.empty(self.n, np.float64)
        lib().or_solver_get_x(self.h, _d(x))
        return x

    def precond(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n, np.float64)
        lib().or_solver_precond(self.h, _d(r), _d(z))
        return z

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n, np.float64)
        lib().or_solver_apply(self.h, _d(x), _d(y))
        return y

    def nlevels(self):
        return lib().or_solver_nlevels(self.h)

    def level(self, k):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().or_solver_level_dims(self.h, k, C.byref(a), C.byref(b), C.byref(c))
        nx, ny, nz = a.value, b.value, c.value
        out = dict(nx=nx, ny=ny, nz=nz)
        for w, name in enumerate(("gE", "gN", "gU", "gB", "cap")):
            v = np.empty(nx * ny if name == "gB" else nx * ny * nz, np.float32)
            lib().or_solver_level_get(self.h, k, w, _f(v))
            out[name] = v
        fcp = np.empty(nx * ny * nz, np.float64)
        finv = np.empty(nx * ny * nz, np.float64)
        lib().or_solver_level_factors(self.h, k, _d(fcp), _d(finv))
        out["fcp"], out["finv"] = fcp, finv
        return out


def steady(arrays, tol=1e-10, maxit=50000, precond="mg", mg_double=False, flexible=False, fast_restrict=False):
    s = Solver(arrays, precond, mg_double, flexible, fast_restrict)
    s.set_x_zero()
    s.set_rhs_steady()
    st = s.solve(tol, maxit)
    st["T"] = s.get_x()
    return st


def _op_args(a):
    return [a["nx"], a["ny"], a["nz"]]


def true_relres(arrays, T):
    a = arrays
    keep = [np.ascontiguousarray(a[k], np.float32) for k in ("gE", "gN", "gU", "gB", "cap")]
    P = np.ascontiguousarray(a["P"], np.float64)
    T = np.ascontiguousarray(T, np.float64)
    return lib().or_true_relres(*_op_args(a), *[_f(v) for v in keep], _d(P), _d(T))


def true_relres_flux(arrays, T):
    """||P - A T|| / ||P|| with A T in flux form (lower rounding floor)."""
    a = arrays
    keep = [np.ascontiguousarray(a[k], np.float32) for k in ("gE", "gN", "gU", "gB", "cap")]
    P = np.ascontiguousarray(a["P"], np.float64)
    T = np.ascontiguousarray(T, np.float64)
    return lib().or_true_relres_flux(*_op_args(a), *[_f(v) for v in keep], _d(P), _d(T))


def heat_out(arrays, T):
    a = arrays
    gU = np.ascontiguousarray(a["gU"], np.float32)
    gB = np.ascontiguousarray(a["gB"], np.float32)
    T = np.ascontiguousarray(T, np.float64)
    return lib().or_heat_out(*_op_args(a), _f(gU), _f(gB), _d(T))


def heat_stored(arrays, Told, Tnew, dt):
    cap = np.ascontiguousarray(arrays["cap"], np.float32)
    Told = np.ascontiguousarray(Told, np.float64)
    Tnew = np.ascontiguousarray(Tnew, np.float64)
    return lib().or_heat_stored(*_op_args(arrays), _f(cap), _d(Told), _d(Tnew), dt)


def run_transient(arrays, dt0=1e-5, growth=1.1, dt_max=0.05, t_end=1.0, max_steps=0, tol=1e-9,
                  maxit=50000, precond="mg", mg_double=False, flexible=False, rec_cap=100000):
    """transient3d.h:48-90; max_steps > 0 adds a step-count stop."""
    a = arrays
    keep = [np.ascontiguousarray(a[k], np.float32) for k in ("gE", "gN", "gU", "gB", "cap")]
    P = np.ascontiguousarray(a["P"], np.float64)
    rec = np.zeros(6 * rec_cap, np.float64)
    T = np.empty(a["nx"] * a["ny"] * a["nz"], np.float64)
    conv = C.c_int()
    n = lib().or_run_transient(*_op_args(a), a["dx"], a["dy"], *[_f(v) for v in keep], _d(P),
                               PRECOND[precond], int(mg_double), int(flexible), dt0, growth, dt_max,
                               t_end, max_steps, tol, maxit, _d(rec), rec_cap, _d(T), C.byref(conv))
    if n < 0:
        raise RuntimeError("rec_cap too small")
    r = rec[: 6 * n].reshape(n, 6)
    return dict(t=r[:, 0].copy(), dt=r[:, 1].copy(), iters=r[:, 2].astype(int), T_max=r[:, 3].copy(),
                q_out=r[:, 4].copy(), q_stored=r[:, 5].copy(), T=T, converged=bool(conv.value))


# ---------------------------------------------------------------------------
# Reference harness (_ref/libref.so): the reference's own CPU path.
# ---------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.ref_model_default.restype = _vp
        L.ref_model_default.argtypes = [C.c_int, C.c_int]
        L.ref_model_new.restype = _vp
        L.ref_model_new.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _ip, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double]
        L.ref_model_case.restype = _vp
        L.ref_model_case.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.ref_model_example_case.restype = _vp
        L.ref_model_example_case.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ref_model_free.argtypes = [_vp]
        L.ref_model_dims.argtypes = [_vp, _ip, _ip, _ip, _dp, _dp]
        L.ref_add_power_block.argtypes = [_vp, C.c_int] + [C.c_double] * 5
        L.ref_add_uniform_power.argtypes = [_vp, C.c_int, C.c_double]
        L.ref_add_hotspot.argtypes = [_vp, C.c_int] + [C.c_double] * 4
        L.ref_add_tsv_region.argtypes = [_vp, C.c_int] + [C.c_double] * 5
        L.ref_finalize.argtypes = [_vp]
        L.ref_model_get_f.argtypes = [_vp, C.c_int, _fp]
        L.ref_model_get_d.argtypes = [_vp, C.c_int, _dp]
        L.ref_total_power.argtypes = [_vp]
        L.ref_total_power.restype = C.c_double
        L.ref_analytic_1d.argtypes = [_vp, _dp]
        L.ref_true_relres.argtypes = [_vp, _dp]
        L.ref_true_relres.restype = C.c_double
        L.ref_heat_out.argtypes = [_vp, _dp]
        L.ref_heat_out.restype = C.c_double
        L.ref_levels.restype = C.c_int
        L.ref_levels.argtypes = [_vp, C.c_int, _ip, _ip, _fp, _fp, _fp, _fp, _fp]
        L.ref_cpu_pcg.restype = C.c_int
        L.ref_cpu_pcg.argtypes = [_vp, C.c_double, C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp]
        L.ref_precondition.argtypes = [_vp, C.c_double, _dp, _dp]
        L.ref_run_transient.restype = C.c_int
        L.ref_run_transient.argtypes = [_vp, C.c_int] + [C.c_double] * 5 + [C.c_int, _dp, C.c_int, _dp, _dp]
        L.ref_export_3dice.restype = C.c_int
        L.ref_export_3dice.argtypes = [C.c_char_p, C.c_double, C.c_char_p, C.c_char_p, C.c_int]
        _ref = L
    return _ref


def ref_export_3dice(folder, cell_um, outdir):
    """The reference's 3D-ICE exporter (case3d.h:224-383) on a case folder:
    None on success, else the exporter's error message."""
    err = C.create_string_buffer(512)
    rc = ref().ref_export_3dice(str(folder).encode(), cell_um, str(outdir).encode(), err, 512)
    return None if rc == 0 else err.value.decode()


class RefModel:
    """A Model3D built by the reference's own code."""

    def __init__(self, h):
        if not h:
            raise ValueError("reference model construction failed")
        self.h = h
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        dx, dy = C.c_double(), C.c_double()
        ref().ref_model_dims(h, C.byref(a), C.byref(b), C.byref(c), C.byref(dx), C.byref(dy))
        self.nx, self.ny, self.nz, self.dx, self.dy = a.value, b.value, c.value, dx.value, dy.value

    @classmethod
    def default(cls, nx, ny=None):
        return cls(ref().ref_model_default(nx, ny or nx))

    @classmethod
    def from_layers(cls, layers, nx, ny, Lx=10e-3, Ly=10e-3, h_top=2e4, h_bot=0.0):
        th = np.array([l[0] for l in layers], np.float64)
        k = np.array([l[1] for l in layers], np.float64)
        rc = np.array([l[2] for l in layers], np.float64)
        nc = np.array([l[3] for l in layers], np.int32)
        ac = np.array([l[4] for l in layers], np.int32)
        return cls(ref().ref_model_new(len(layers), _d(th), _d(k), _d(rc), _i(nc), _i(ac), nx, ny, Lx, Ly,
                                       h_top, h_bot))

    @classmethod
    def case(cls, folder, nx, ny=None, with_tsv=True):
        err = C.create_string_buffer(512)
        h = ref().ref_model_case(folder.encode(), nx, ny or nx, int(with_tsv), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    @classmethod
    def example_case(cls, nx, ny=None, with_tsv=True):
        return cls(ref().ref_model_example_case(nx, ny or nx, int(with_tsv)))

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_model_free(self.h)
            self.h = None

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def add_power_block(self, *a):
        ref().ref_add_power_block(self.h, *a)

    def add_uniform_power(self, *a):
        ref().ref_add_uniform_power(self.h, *a)

    def add_hotspot(self, *a):
        ref().ref_add_hotspot(self.h, *a)

    def add_tsv_region(self, *a):
        ref().ref_add_tsv_region(self.h, *a)

    def finalize(self):
        ref().ref_finalize(self.h)
        return self

    def arrays(self):
        out = {}
        for w, name in enumerate(("gE", "gN", "gU", "gB", "cap")):
            a = np.empty(self.nx * self.ny if name == "gB" else self.n, np.float32)
            ref().ref_model_get_f(self.h, w, _f(a))
            out[name] = a
        for w, name in ((5, "P"), (6, "tsv_frac")):
            a = np.empty(self.n, np.float64)
            ref().ref_model_get_d(self.h, w, _d(a))
            out[name] = a
        dz = np.empty(self.nz, np.float64)
        ref().ref_model_get_d(self.h, 7, _d(dz))
        out["dz"] = dz
        out.update(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy)
        return out

    def total_power(self):
        return ref().ref_total_power(self.h)

    def analytic_1d(self):
        T = np.empty(self.nz, np.float64)
        ref().ref_analytic_1d(self.h, _d(T))
        return T

    def true_relres(self, T):
        T = np.ascontiguousarray(T, np.float64)
        return ref().ref_true_relres(self.h, _d(T))

    def heat_out(self, T):
        T = np.ascontiguousarray(T, np.float64)
        return ref().ref_heat_out(self.h, _d(T))

    def levels(self):
        out = []
        nl = ref().ref_levels(self.h, -1, None, None, None, None, None, None, None)
        a, b = C.c_int(), C.c_int()
        for k in range(nl):
            ref().ref_levels(self.h, k, C.byref(a), C.byref(b), None, None, None, None, None)
            nx, ny = a.value, b.value
            n = nx * ny * self.nz
            arr = {name: np.empty(nx * ny if name == "gB" else n, np.float32)
                   for name in ("gE", "gN", "gU", "gB", "cap")}
            ref().ref_levels(self.h, k, C.byref(a), C.byref(b), *[_f(arr[x]) for x in ("gE", "gN", "gU", "gB", "cap")])
            arr.update(nx=nx, ny=ny, nz=self.nz)
            out.append(arr)
        return out

    def cpu_pcg(self, tol, maxit=50000, precond="mg"):
        T = np.empty(self.n, np.float64)
        it, rr, su, so = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        conv = ref().ref_cpu_pcg(self.h, tol, maxit, PRECOND[precond], _d(T), C.byref(it), C.byref(rr),
                                 C.byref(su), C.byref(so))
        return dict(T=T, iters=it.value, relres=rr.value, converged=bool(conv), setup_ms=su.value,
                    solve_ms=so.value)

    def precondition(self, r, dt=0.0):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n, np.float64)
        ref().ref_precondition(self.h, dt, _d(r), _d(z))
        return z

    def run_transient(self, precond="mg", dt0=1e-5, growth=1.1, dt_max=0.05, t_end=1.0, tol=1e-9,
                      maxit=50000, rec_cap=100000):
        rec = np.zeros(6 * rec_cap, np.float64)
        T = np.empty(self.n, np.float64)
        ms = C.c_double()
        n = ref().ref_run_transient(self.h, PRECOND[precond], dt0, growth, dt_max, t_end, tol, maxit,
                                    _d(rec), rec_cap, _d(T), C.byref(ms))
        if n < 0:
            raise RuntimeError("rec_cap too small")
        r = rec[: 6 * n].reshape(n, 6)
        return dict(t=r[:, 0].copy(), dt=r[:, 1].copy(), iters=r[:, 2].astype(int), T_max=r[:, 3].copy(),
                    q_out=r[:, 4].copy(), q_stored=r[:, 5].copy(), T=T, solve_ms=ms.value)
