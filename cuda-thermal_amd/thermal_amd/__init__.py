"""Python binding of the MI355X thermal solver's C ABI (include/thermal_hip.h).

Mirrors the reference's interface for the hot path: the model builder of
src/stack3d.h / src/case3d.h, the solver concept of src/pcg3d.h:60-163
(set_dt, set_x_zero, set_rhs_steady, set_rhs_transient, solve, get_x,
set_power), gpu_pcg (steady solve from T = 0) and run_transient
(src/transient3d.h:48-90). Every call goes to lib/libthermal_hip.so; there
is no fallback: if the library is missing, loading fails with an error.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libthermal_hip.so")

TH_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "model not finalized", 4: "case error",
          5: "too many z-planes", 6: "hotspot covers no cells"}

PRECOND = {"jacobi": 0, "mg": 1, "multigrid": 1}
PRECISION = {"mixed": 0, "fp64": 1}
BETA = {"fr": 0, "fletcher-reeves": 0, "flexible": 1}
RESTRICTION = {"lateral": 0, "residual": 1}

# Layers of the reference's default two-tier stack (stack3d.h:153-162):
# (name, thickness m, k W/mK, rho_c J/m^3K, planes, active)
DEFAULT_STACK = [
    ("Si tier 2 (bottom)", 50e-6, 130.0, 1.66e6, 2, 1),
    ("bond + BEOL", 10e-6, 2.0, 1.70e6, 1, 0),
    ("Si tier 1 (top)", 50e-6, 130.0, 1.66e6, 2, 1),
    ("TIM", 25e-6, 5.0, 2.00e6, 1, 0),
    ("Cu spreader", 1e-3, 400.0, 3.45e6, 4, 0),
]


class ThermalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {msg}")
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("gE", C.POINTER(C.c_float)), ("gN", C.POINTER(C.c_float)), ("gU", C.POINTER(C.c_float)),
                ("gB", C.POINTER(C.c_float)), ("cap", C.POINTER(C.c_float)), ("P", C.POINTER(C.c_double))]


class SolverOpts(C.Structure):
    _fields_ = [("precond", C.c_int), ("precision", C.c_int), ("beta", C.c_int), ("use_graphs", C.c_int),
                ("device", C.c_int), ("true_residual", C.c_int), ("restriction", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("relres", C.c_double), ("converged", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("relres", C.c_double), ("converged", C.c_int), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double)]


class SetupTimes(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("alloc_ms", "upload_ms", "levels_ms", "types_ms", "factor_ms", "graphs_ms",
                                          "total_ms")]


class TransientOpts(C.Structure):
    _fields_ = [("dt0", C.c_double), ("growth", C.c_double), ("dt_max", C.c_double), ("t_end", C.c_double),
                ("tol", C.c_double), ("maxit", C.c_int), ("max_steps", C.c_int), ("predictor", C.c_int)]


class StepRecord(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("iters", C.c_int), ("T_max", C.c_double),
                ("q_out", C.c_double), ("q_stored", C.c_double)]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "th_last_error": (C.c_char_p, []),
    "th_version": (C.c_char_p, []),
    "th_model_create": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), _dp, _dp, _dp, _ip, _ip, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(_vp)]),
    "th_model_create_default": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "th_model_load_case": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "th_model_example_case": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "th_model_destroy": (None, [_vp]),
    "th_model_add_power_block": (C.c_int, [_vp, C.c_int] + [C.c_double] * 5),
    "th_model_add_uniform_power": (C.c_int, [_vp, C.c_int, C.c_double]),
    "th_model_add_hotspot": (C.c_int, [_vp, C.c_int] + [C.c_double] * 4),
    "th_model_add_tsv_region": (C.c_int, [_vp, C.c_int] + [C.c_double] * 5),
    "th_model_finalize": (C.c_int, [_vp]),
    "th_model_set_power": (C.c_int, [_vp, _dp]),
    "th_model_dims": (C.c_int, [_vp, _ip, _ip, _ip, _dp, _dp]),
    "th_model_layer_count": (C.c_int, [_vp]),
    "th_model_plane_layer": (C.c_int, [_vp, _ip]),
    "th_model_layer_info": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _ip, _ip, C.POINTER(C.c_char_p)]),
    "th_model_settings": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "th_model_get_f32": (C.c_int, [_vp, C.c_char_p, _fp]),
    "th_model_get_f64": (C.c_int, [_vp, C.c_char_p, _dp]),
    "th_model_total_power": (C.c_double, [_vp]),
    "th_model_analytic_1d": (C.c_int, [_vp, _dp]),
    "th_model_heat_out": (C.c_double, [_vp, _dp]),
    "th_model_describe": (C.c_int, [_vp, C.POINTER(ModelDesc)]),
    "th_model_block_count": (C.c_int, [_vp]),
    "th_model_tsv_count": (C.c_int, [_vp]),
    "th_model_block_info": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_char_p), _ip, _dp, _dp, _dp, _dp, _dp]),
    "th_model_block_temps": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "th_case_export_3dice": (C.c_int, [C.c_char_p, C.c_double, C.c_char_p, _ip, _ip]),
    "th_3dice_read_tmap": (C.c_int, [C.c_char_p, _ip, _ip, _dp, C.c_long]),
    "th_3dice_read_tflp": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, _dp, C.c_int, _ip]),
    "th_solver_default_opts": (None, [C.POINTER(SolverOpts)]),
    "th_solver_create": (C.c_int, [C.POINTER(ModelDesc), C.POINTER(SolverOpts), C.POINTER(_vp)]),
    "th_solver_destroy": (None, [_vp]),
    "th_solver_set_dt": (C.c_int, [_vp, C.c_double]),
    "th_solver_set_x_zero": (C.c_int, [_vp]),
    "th_solver_set_x": (C.c_int, [_vp, _dp]),
    "th_solver_set_rhs_steady": (C.c_int, [_vp]),
    "th_solver_set_rhs_transient": (C.c_int, [_vp]),
    "th_solver_solve": (C.c_int, [_vp, C.c_double, C.c_int, C.POINTER(Stats)]),
    "th_solver_get_x": (C.c_int, [_vp, _dp]),
    "th_solver_set_power": (C.c_int, [_vp, _dp]),
    "th_solver_device_x": (C.c_int, [_vp, C.POINTER(_dp)]),
    "th_solver_true_relres": (C.c_int, [_vp, _dp]),
    "th_solver_heat_out": (C.c_int, [_vp, _dp]),
    "th_solver_nlevels": (C.c_int, [_vp]),
    "th_solver_column_types": (C.c_int, [_vp]),
    "th_solver_level_info": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _ip, _ip]),
    "th_solver_setup_times": (C.c_int, [_vp, C.POINTER(SetupTimes)]),
    "th_solver_set_guide": (C.c_int, [_vp, _dp]),
    "th_solver_set_guide_lifted": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                             C.c_int]),
    "th_solver_guide_stats": (C.c_int, [_vp, _ip, _dp]),
    "th_solver_colorize": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, _vp, _fp, _ip]),
    "th_solver_colorize_device": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, _vp, _fp, _ip]),
    "th_solver_colorize_device_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, _vp, _fp, _ip]),
    "th_solver_precondition": (C.c_int, [_vp, _dp, _dp]),
    "th_solver_apply_operator": (C.c_int, [_vp, _dp, _dp]),
    "th_solver_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "th_steady_solve": (C.c_int, [C.POINTER(ModelDesc), C.POINTER(SolverOpts), C.c_double, C.c_int, _dp,
                                  C.POINTER(SolveResult)]),
    "th_transient_default_opts": (None, [C.POINTER(TransientOpts)]),
    "th_run_transient": (C.c_int, [_vp, C.POINTER(TransientOpts), C.POINTER(StepRecord), C.c_int, _ip, _dp,
                                   _ip, _dp]),
    "th_batch_create": (C.c_int, [C.c_int, C.POINTER(C.POINTER(ModelDesc)), C.POINTER(SolverOpts), C.POINTER(_vp)]),
    "th_batch_destroy": (None, [_vp]),
    "th_batch_count": (C.c_int, [_vp]),
    "th_batch_set_power": (C.c_int, [_vp, C.c_int, _dp]),
    "th_batch_solve": (C.c_int, [_vp, C.c_double, C.c_int, _ip, _dp, _ip, _dp]),
    "th_batch_get_x": (C.c_int, [_vp, C.c_int, _dp]),
    "th_batch_autotune": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _dp]),
    "th_batch_shape": (C.c_int, [_vp, _ip, _ip, _ip]),
    "th_surrogate_param_count": (C.c_long, []),
    "th_surrogate_create": (C.c_int, [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_vp)]),
    "th_surrogate_destroy": (None, [_vp]),
    "th_surrogate_get_params": (C.c_int, [_vp, _fp]),
    "th_surrogate_set_params": (C.c_int, [_vp, _fp]),
    "th_surrogate_predict": (C.c_int, [_vp, _fp, C.c_int, _fp]),
    "th_surrogate_predict_device": (C.c_int, [_vp, _fp, C.POINTER(_fp)]),
    "th_surrogate_predict_device_input": (C.c_int, [_vp, C.c_void_p, C.POINTER(_fp)]),
    "th_surrogate_grad": (C.c_int, [_vp, _fp, _fp, C.c_int, _dp, _fp]),
    "th_surrogate_dataset_reserve": (C.c_int, [_vp, C.c_int]),
    "th_surrogate_dataset_add": (C.c_int, [_vp, _fp, _fp, C.c_int]),
    "th_surrogate_dataset_size": (C.c_int, [_vp]),
    "th_surrogate_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_ulonglong, C.c_int, _fp,
                                     _dp]),
    "th_surrogate_eval": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "th_surrogate_time_forward": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "th_solver_set_x_lifted": (C.c_int, [_vp, _fp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                         C.c_int, _ip, _dp, _dp]),
    "th_solver_set_x_learned": (C.c_int, [_vp, _vp, C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_int]),
    "th_solver_lifted_stats": (C.c_int, [_vp, _ip, _dp, _dp]),
}

_lib = None


def lib():
    """The loaded library. Raises if it has not been built: the product has no
    CPU fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C cuda-thermal_amd`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != TH_OK:
        raise ThermalError(rc, lib().th_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


class DeviceBuffer:
    """A float32 array copied into device memory (hipMalloc + hipMemcpy of
    the HIP runtime the library links): input kept resident between calls,
    e.g. the surrogate's power maps for Surrogate.predict_device_input."""

    def __init__(self, a):
        self._hip = C.CDLL("libamdhip64.so.7")
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        if self._hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) != 0:
            raise ThermalError("hipMalloc failed")
        self.ptr = p.value
        if self._hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) != 0:  # host to device
            raise ThermalError("hipMemcpy failed")
        self.nbytes = a.nbytes

    def __del__(self):
        if getattr(self, "ptr", None):
            self._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None


class Model:
    """A stack model built by the library (stack3d.h:45-148)."""

    def __init__(self, handle):
        self.h = handle
        nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
        dx, dy = C.c_double(), C.c_double()
        _check(lib().th_model_dims(self.h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(dx), C.byref(dy)))
        self.nx, self.ny, self.nz, self.dx, self.dy = nx.value, ny.value, nz.value, dx.value, dy.value
        self._desc = None

    @classmethod
    def from_layers(cls, layers, nx, ny=None, Lx=10e-3, Ly=10e-3, h_top=2e4, h_bot=0.0):
        ny = ny or nx
        n = len(layers)
        names = (C.c_char_p * n)(*[l[0].encode() for l in layers])
        th = np.array([l[1] for l in layers], np.float64)
        k = np.array([l[2] for l in layers], np.float64)
        rc = np.array([l[3] for l in layers], np.float64)
        planes = np.array([l[4] for l in layers], np.int32)
        active = np.array([l[5] for l in layers], np.int32)
        h = _vp()
        _check(lib().th_model_create(n, names, _d(th), _d(k), _d(rc), planes.ctypes.data_as(_ip),
                                     active.ctypes.data_as(_ip), nx, ny, Lx, Ly, h_top, h_bot, C.byref(h)))
        return cls(h)

    @classmethod
    def default(cls, nx, ny=None):
        h = _vp()
        _check(lib().th_model_create_default(nx, ny or nx, C.byref(h)))
        return cls(h)

    @classmethod
    def load_case(cls, folder, nx, ny=None, with_tsv=True):
        h = _vp()
        _check(lib().th_model_load_case(str(folder).encode(), nx, ny or nx, int(with_tsv), C.byref(h)))
        return cls(h)

    @classmethod
    def example_case(cls, nx, ny=None, with_tsv=True):
        h = _vp()
        _check(lib().th_model_example_case(nx, ny or nx, int(with_tsv), C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().th_model_destroy(self.h)
            self.h = None

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def add_power_block(self, layer, x0, y0, w, h, watts):
        _check(lib().th_model_add_power_block(self.h, layer, x0, y0, w, h, watts))

    def add_uniform_power(self, layer, watts):
        _check(lib().th_model_add_uniform_power(self.h, layer, watts))

    def add_hotspot(self, layer, xc, yc, size, watts):
        _check(lib().th_model_add_hotspot(self.h, layer, xc, yc, size, watts))

    def add_tsv_region(self, layer, x0, y0, w, h, f):
        _check(lib().th_model_add_tsv_region(self.h, layer, x0, y0, w, h, f))

    def finalize(self):
        _check(lib().th_model_finalize(self.h))
        self._desc = None
        return self

    def set_power(self, P):
        self._P = np.ascontiguousarray(P, np.float64)
        _check(lib().th_model_set_power(self.h, _d(self._P)))

    def get(self, name):
        if name in ("gE", "gN", "gU", "cap", "gB"):
            a = np.empty(self.nx * self.ny if name == "gB" else self.n, np.float32)
            _check(lib().th_model_get_f32(self.h, name.encode(), _f(a)))
        else:
            a = np.empty(self.nz if name == "dz" else self.n, np.float64)
            _check(lib().th_model_get_f64(self.h, name.encode(), _d(a)))
        return a

    def arrays(self):
        out = {k: self.get(k) for k in ("gE", "gN", "gU", "gB", "cap", "P", "tsv_frac", "dz")}
        out.update(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy)
        return out

    def plane_layer(self):
        a = np.empty(self.nz, np.int32)
        _check(lib().th_model_plane_layer(self.h, a.ctypes.data_as(_ip)))
        return a

    def layers(self):
        out = []
        for L in range(lib().th_model_layer_count(self.h)):
            t, k, rc = C.c_double(), C.c_double(), C.c_double()
            pl, ac = C.c_int(), C.c_int()
            nm = C.c_char_p()
            _check(lib().th_model_layer_info(self.h, L, C.byref(t), C.byref(k), C.byref(rc), C.byref(pl),
                                             C.byref(ac), C.byref(nm)))
            out.append((nm.value.decode(), t.value, k.value, rc.value, pl.value, ac.value))
        return out

    def settings(self):
        v = [C.c_double() for _ in range(5)]
        _check(lib().th_model_settings(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("Lx", "Ly", "h_top", "h_bot", "T_amb"), [x.value for x in v]))

    def total_power(self):
        return lib().th_model_total_power(self.h)

    def analytic_1d(self):
        T = np.empty(self.nz, np.float64)
        _check(lib().th_model_analytic_1d(self.h, _d(T)))
        return T

    def heat_out(self, T):
        T = np.ascontiguousarray(T, np.float64)
        return lib().th_model_heat_out(self.h, _d(T))

    def blocks(self):
        """Power blocks of the case the model was built from (CaseSpec,
        case3d.h:24-33): (name, layer, x0, y0, w, h, watts) each."""
        out = []
        for k in range(lib().th_model_block_count(self.h)):
            nm = C.c_char_p()
            layer = C.c_int()
            v = [C.c_double() for _ in range(5)]
            _check(lib().th_model_block_info(self.h, k, C.byref(nm), C.byref(layer), *[C.byref(x) for x in v]))
            out.append((nm.value.decode(), layer.value) + tuple(x.value for x in v))
        return out

    def block_temps(self, k, T):
        """Max and area-weighted mean rise of T over block k (block_temps,
        main3d.cu:286-301)."""
        T = np.ascontiguousarray(T, np.float64)
        tmax, tavg = C.c_double(), C.c_double()
        _check(lib().th_model_block_temps(self.h, k, _d(T), C.byref(tmax), C.byref(tavg)))
        return tmax.value, tavg.value

    def desc(self):
        d = ModelDesc()
        _check(lib().th_model_describe(self.h, C.byref(d)))
        return d


def export_3dice(folder, cell_um, outdir):
    """Writes case folder `folder` as a 3D-ICE stack (case.stk, floorplans,
    TSV layouts) on cells of cell_um micrometres into outdir (export_3dice,
    case3d.h:224-383); returns the grid (nx, ny)."""
    nx, ny = C.c_int(), C.c_int()
    _check(lib().th_case_export_3dice(str(folder).encode(), cell_um, str(outdir).encode(), C.byref(nx), C.byref(ny)))
    return nx.value, ny.value


def read_3dice_tmap(path):
    """A 3D-ICE Tmap file as a (rows, cols) array of K."""
    rows, cols = C.c_int(), C.c_int()
    _check(lib().th_3dice_read_tmap(str(path).encode(), C.byref(rows), C.byref(cols), None, 0))
    out = np.empty(rows.value * cols.value, np.float64)
    _check(lib().th_3dice_read_tmap(str(path).encode(), C.byref(rows), C.byref(cols), _d(out), out.size))
    return out.reshape(rows.value, cols.value)


def read_3dice_tflp(path):
    """A 3D-ICE Tflp file as {block name: K}, in file order."""
    n = C.c_int()
    _check(lib().th_3dice_read_tflp(str(path).encode(), None, 0, None, 0, C.byref(n)))
    names = C.create_string_buffer(256 * n.value)
    vals = np.empty(n.value, np.float64)
    _check(lib().th_3dice_read_tflp(str(path).encode(), names, 256, _d(vals), n.value, C.byref(n)))
    return {names.raw[256 * k:256 * (k + 1)].split(b"\0", 1)[0].decode(): float(vals[k]) for k in range(n.value)}


def default_model(n, ny=None):
    return Model.default(n, ny)


def hotspot_model(n, xc=3e-3, yc=3e-3):
    """main3d.cu:60-67: 22 W uniform + 3 W in a 1 mm hotspot in the bottom
    tier, 25 W uniform in the top tier."""
    m = Model.default(n)
    m.add_uniform_power(0, 22.0)
    m.add_hotspot(0, xc, yc, 1e-3, 3.0)
    m.add_uniform_power(2, 25.0)
    return m.finalize()


def desc_from_arrays(a):
    """A ModelDesc over numpy arrays (kept alive by the returned holder)."""
    keep = {k: np.ascontiguousarray(a[k], np.float32) for k in ("gE", "gN", "gU", "gB", "cap")}
    keep["P"] = np.ascontiguousarray(a["P"], np.float64)
    d = ModelDesc(a["nx"], a["ny"], a["nz"], a["dx"], a["dy"], _f(keep["gE"]), _f(keep["gN"]), _f(keep["gU"]),
                  _f(keep["gB"]), _f(keep["cap"]), _d(keep["P"]))
    return d, keep


def make_opts(precond="mg", precision="mixed", beta="flexible", graphs=True, device=0, true_residual=False,
              restriction="lateral"):
    return SolverOpts(PRECOND[precond], PRECISION[precision], BETA[beta], int(graphs), device, int(true_residual),
                      RESTRICTION[restriction])


class Solver:
    """GpuPcg's interface (pcg3d.h:115-163) over the C ABI. `model` is a
    Model or a dict of arrays (gE, gN, gU, gB, cap, P, nx, ny, nz, dx, dy)."""

    def __init__(self, model, precond="mg", precision="mixed", beta="flexible", graphs=True, device=0,
                 true_residual=False, restriction="lateral"):
        if isinstance(model, Model):
            d, self._keep = model.desc(), model
            self.nx, self.ny, self.nz = model.nx, model.ny, model.nz
        else:
            d, self._keep = desc_from_arrays(model)
            self.nx, self.ny, self.nz = model["nx"], model["ny"], model["nz"]
        self.n = self.nx * self.ny * self.nz
        self.h = None
        h = _vp()
        opts = make_opts(precond, precision, beta, graphs, device, true_residual, restriction)
        _check(lib().th_solver_create(C.byref(d), C.byref(opts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().th_solver_destroy(self.h)
            self.h = None

    __del__ = close

    def set_dt(self, dt):
        _check(lib().th_solver_set_dt(self.h, dt))

    def set_x_zero(self):
        _check(lib().th_solver_set_x_zero(self.h))

    def set_x(self, x):
        x = np.ascontiguousarray(x, np.float64)
        _check(lib().th_solver_set_x(self.h, _d(x)))

    def set_rhs_steady(self):
        _check(lib().th_solver_set_rhs_steady(self.h))

    def set_rhs_transient(self):
        _check(lib().th_solver_set_rhs_transient(self.h))

    def set_power(self, P):
        P = np.ascontiguousarray(P, np.float64)
        _check(lib().th_solver_set_power(self.h, _d(P)))

    def solve(self, tol, maxit=50000):
        st = Stats()
        _check(lib().th_solver_solve(self.h, tol, maxit, C.byref(st)))
        return dict(iters=st.iters, relres=st.relres, converged=bool(st.converged))

    def get_x(self):
        x = np.empty(self.n, np.float64)
        _check(lib().th_solver_get_x(self.h, _d(x)))
        return x

    def device_x(self):
        p = _dp()
        _check(lib().th_solver_device_x(self.h, C.byref(p)))
        return C.cast(p, C.c_void_p).value

    def true_relres(self):
        v = C.c_double()
        _check(lib().th_solver_true_relres(self.h, C.byref(v)))
        return v.value

    def heat_out(self):
        v = C.c_double()
        _check(lib().th_solver_heat_out(self.h, C.byref(v)))
        return v.value

    def nlevels(self):
        return lib().th_solver_nlevels(self.h)

    def column_types(self):
        """Distinct fine-level column types in use (0: per-cell arrays)."""
        return lib().th_solver_column_types(self.h)

    def setup_times(self):
        """Where the solver's set-up time went, ms (GpuSetupTimes)."""
        t = SetupTimes()
        _check(lib().th_solver_setup_times(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in SetupTimes._fields_}

    def set_guide(self, G):
        """Guide field of the learned step start (natural layout, K)."""
        self._guide = np.ascontiguousarray(G, np.float64)
        _check(lib().th_solver_set_guide(self.h, _d(self._guide)))

    def set_guide_lifted(self, d_field, Hc, Wc, C_, zc_coarse, zc_fine, scale=1.0, passes=2):
        """Guide field from a coarse device field (the surrogate's output),
        lifted as set_x_lifted lifts its guess."""
        zc = np.ascontiguousarray(zc_coarse, np.float64)
        zf = np.ascontiguousarray(zc_fine, np.float64)
        _check(lib().th_solver_set_guide_lifted(self.h, d_field, Hc, Wc, C_, len(zc), _d(zc), _d(zf), scale, passes))

    def guide_stats(self):
        """(steps whose start took the guide direction, their mean w) over
        the last run_transient."""
        n, w = C.c_int(), C.c_double()
        _check(lib().th_solver_guide_stats(self.h, C.byref(n), C.byref(w)))
        return n.value, w.value

    def levels(self):
        """Per multigrid level: dict(nx, ny, nz, types, kind), kind one of
        "column walk", "plane-parallel", "tail"."""
        out = []
        for k in range(self.nlevels()):
            v = [C.c_int() for _ in range(5)]
            _check(lib().th_solver_level_info(self.h, k, *[C.byref(x) for x in v]))
            out.append(dict(nx=v[0].value, ny=v[1].value, nz=v[2].value, types=v[3].value,
                            kind=("column walk", "plane-parallel", "tail")[v[4].value]))
        return out

    def colorize(self, l0, l1, rise_hi, image=True):
        """Die map of the current T (viewer.cu:68-87): the hottest rise of
        planes l0..l1 per column, coloured on [0, rise_hi] as an (ny, nx, 4)
        RGBA uint8 image (None with image=False), and the hottest cell:
        (rise, (i, j))."""
        img = np.empty((self.ny, self.nx, 4), np.uint8) if image else None
        v, cell = C.c_float(), C.c_int()
        ptr = img.ctypes.data_as(_vp) if image else None
        _check(lib().th_solver_colorize(self.h, l0, l1, rise_hi, ptr, C.byref(v), C.byref(cell)))
        return img, v.value, (cell.value % self.nx, cell.value // self.nx)

    def precondition(self, r):
        """z = M^-1 r for the current dt (one V-cycle, or Jacobi)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n, np.float64)
        _check(lib().th_solver_precondition(self.h, _d(r), _d(z)))
        return z

    KERNELS = {"sweep": 0, "apply": 1, "restrict": 2, "copy": 3, "vcycle": 4, "update": 5}

    def time_kernel(self, which, reps=20):
        """(average ms per launch, algorithmic bytes per launch) of one
        hot-path kernel, timed with HIP events on the solver's stream."""
        ms, nb = C.c_double(), C.c_double()
        _check(lib().th_solver_time_kernel(self.h, self.KERNELS[which], reps, C.byref(ms), C.byref(nb)))
        return ms.value, nb.value

    def apply(self, x):
        """y = (A + C/dt) x in FP64."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n, np.float64)
        _check(lib().th_solver_apply_operator(self.h, _d(x), _d(y)))
        return y

    def set_x_lifted(self, d_field, Hc, Wc, C_, zc_coarse, zc_fine, scale=1.0, passes=2):
        """Initial guess lifted from a coarse device field (th_solver_set_x_lifted),
        with the plain warm start as the fallback. Returns (used, relres_plain,
        relres_lift)."""
        zc = np.ascontiguousarray(zc_coarse, np.float64)
        zf = np.ascontiguousarray(zc_fine, np.float64)
        used, rp, rl = C.c_int(), C.c_double(), C.c_double()
        _check(lib().th_solver_set_x_lifted(self.h, C.cast(d_field, _fp), Hc, Wc, C_, len(zc), _d(zc), _d(zf),
                                            scale, passes, C.byref(used), C.byref(rp), C.byref(rl)))
        return bool(used.value), rp.value, rl.value

    def set_x_learned(self, sur, d_in, zc_coarse, zc_fine, scale=1.0, passes=1):
        """The learned start in one asynchronous call (th_solver_set_x_learned):
        the surrogate `sur` predicts from its device input d_in on this
        solver's stream, and the lifted guess is kept or dropped on the
        device; lifted_stats() reads the judgement afterwards."""
        zc = np.ascontiguousarray(zc_coarse, np.float64)
        zf = np.ascontiguousarray(zc_fine, np.float64)
        _check(lib().th_solver_set_x_learned(self.h, sur.h, C.c_void_p(d_in), len(zc), _d(zc), _d(zf), scale,
                                             passes))

    def lifted_stats(self):
        """(used, relres_plain, relres_lift) of the last set_x_learned."""
        used, rp, rl = C.c_int(), C.c_double(), C.c_double()
        _check(lib().th_solver_lifted_stats(self.h, C.byref(used), C.byref(rp), C.byref(rl)))
        return bool(used.value), rp.value, rl.value


def steady_solve(model, tol=1e-10, maxit=50000, **kw):
    """gpu_pcg (pcg3d_gpu.cu:686-719): steady solve from T = 0, timed."""
    if isinstance(model, Model):
        d, keep = model.desc(), model
        n = model.n
    else:
        d, keep = desc_from_arrays(model)
        n = model["nx"] * model["ny"] * model["nz"]
    opts = make_opts(**kw)
    T = np.empty(n, np.float64)
    r = SolveResult()
    _check(lib().th_steady_solve(C.byref(d), C.byref(opts), tol, maxit, _d(T), C.byref(r)))
    return dict(T=T, iters=r.iters, relres=r.relres, converged=bool(r.converged), setup_ms=r.setup_ms,
                solve_ms=r.solve_ms, ms=r.total_ms)


PREDICTORS = {"plain": 0, "extrapolate": 1, "learned": 2, "extrapolate+learned": 3}


def run_transient(solver, dt0=1e-5, growth=1.1, dt_max=0.05, t_end=1.0, tol=1e-9, maxit=50000, max_steps=0,
                  rec_cap=100000, predictor=False):
    """run_transient (transient3d.h:48-90) on a Solver; energy terms on the device.
    predictor: the start of each step - False / "plain" (T_old), True /
    "extrapolate" (linear extrapolation of the last two steps), "learned"
    (T_old moved towards the solver's guide field, see Solver.set_guide),
    "extrapolate+learned"."""
    if isinstance(predictor, str):
        predictor = PREDICTORS[predictor]
    o = TransientOpts(dt0, growth, dt_max, t_end, tol, maxit, max_steps, int(predictor))
    recs = (StepRecord * rec_cap)()
    n, conv = C.c_int(), C.c_int()
    ms = C.c_double()
    T = np.empty(solver.n, np.float64)
    _check(lib().th_run_transient(solver.h, C.byref(o), recs, rec_cap, C.byref(n), _d(T), C.byref(conv),
                                  C.byref(ms)))
    k = n.value
    return dict(t=np.array([recs[i].t for i in range(k)]), dt=np.array([recs[i].dt for i in range(k)]),
                iters=np.array([recs[i].iters for i in range(k)]),
                T_max=np.array([recs[i].T_max for i in range(k)]),
                q_out=np.array([recs[i].q_out for i in range(k)]),
                q_stored=np.array([recs[i].q_stored for i in range(k)]), T=T, converged=bool(conv.value),
                solve_ms=ms.value)


class Surrogate:
    """The learned warm start (th_surrogate_*): a U-Net on an H x W grid from
    3 features per cell to 10 plane temperature rises (units of 50 K)."""

    N_IN, N_OUT, T_SCALE = 3, 10, 50.0

    def __init__(self, H=256, W=256, seed=1):
        self.H, self.W = H, W
        h = _vp()
        _check(lib().th_surrogate_create(H, W, seed, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().th_surrogate_destroy(self.h)
            self.h = None

    __del__ = close

    @staticmethod
    def param_count():
        return lib().th_surrogate_param_count()

    def get_params(self):
        p = np.empty(self.param_count(), np.float32)
        _check(lib().th_surrogate_get_params(self.h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count():
            raise ValueError("wrong parameter count")
        _check(lib().th_surrogate_set_params(self.h, _f(p)))

    def predict(self, x):
        """x [B,H,W,3] float -> [B,H,W,10] float (units of T_SCALE)."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(x.shape[:3] + (self.N_OUT,), np.float32)
        _check(lib().th_surrogate_predict(self.h, _f(x), x.shape[0], _f(out)))
        return out

    def predict_device_input(self, d_in):
        """One image whose input already sits in device memory (d_in: device
        address of float32 [H][W][C], e.g. DeviceBuffer.ptr); device pointer
        to the FP32 [H][W][16] output."""
        p = _fp()
        _check(lib().th_surrogate_predict_device_input(self.h, C.c_void_p(d_in), C.byref(p)))
        return C.cast(p, C.c_void_p).value

    def predict_device(self, x):
        """One image; device pointer to the FP32 [H][W][16] output."""
        x = np.ascontiguousarray(x, np.float32)
        p = _fp()
        _check(lib().th_surrogate_predict_device(self.h, _f(x), C.byref(p)))
        return C.cast(p, C.c_void_p).value

    def grad(self, x, tgt):
        x = np.ascontiguousarray(x, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        loss = C.c_double()
        g = np.empty(self.param_count(), np.float32)
        _check(lib().th_surrogate_grad(self.h, _f(x), _f(tgt), x.shape[0], C.byref(loss), _f(g)))
        return loss.value, g

    def reserve(self, capacity):
        _check(lib().th_surrogate_dataset_reserve(self.h, capacity))

    def add(self, x, tgt):
        x = np.ascontiguousarray(x, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        _check(lib().th_surrogate_dataset_add(self.h, _f(x), _f(tgt), x.shape[0]))

    def size(self):
        return lib().th_surrogate_dataset_size(self.h)

    def train(self, steps, batch=16, lr=2e-3, lr_final=1e-4, seed=7, n_train=None):
        losses = np.empty(steps, np.float32)
        ms = C.c_double()
        _check(lib().th_surrogate_train(self.h, steps, batch, lr, lr_final, seed, n_train or self.size(),
                                        _f(losses), C.byref(ms)))
        return losses, ms.value

    def eval(self, first, count):
        v = C.c_double()
        _check(lib().th_surrogate_eval(self.h, first, count, C.byref(v)))
        return v.value

    def time_forward(self, batch=1, reps=20):
        ms, fl = C.c_double(), C.c_double()
        _check(lib().th_surrogate_time_forward(self.h, batch, reps, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value


class Batch:
    """Batched steady solves (th_batch_*): many stacks, one launch per stage."""

    def __init__(self, models, graphs=True, device=0):
        descs, self._keep = [], []
        for m in models:
            if isinstance(m, Model):
                d, k = m.desc(), m
            else:
                d, k = desc_from_arrays(m)
            descs.append(d)
            self._keep.append((d, k))
        self.sizes = [(d.nx, d.ny, d.nz) for d in descs]
        arr = (C.POINTER(ModelDesc) * len(descs))(*[C.pointer(d) for d in descs])
        opts = make_opts("mg", "mixed", "flexible", graphs, device)
        h = _vp()
        self.h = None
        _check(lib().th_batch_create(len(descs), arr, C.byref(opts), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().th_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def __len__(self):
        return len(self.sizes)

    def set_power(self, i, P):
        P = np.ascontiguousarray(P, np.float64)
        _check(lib().th_batch_set_power(self.h, i, _d(P)))

    def solve(self, tol=1e-10, maxit=500):
        n = len(self)
        it, rr, cv = np.zeros(n, np.int32), np.zeros(n), np.zeros(n, np.int32)
        ms = C.c_double()
        _check(lib().th_batch_solve(self.h, tol, maxit, it.ctypes.data_as(_ip), _d(rr), cv.ctypes.data_as(_ip),
                                    C.byref(ms)))
        return dict(iters=it, relres=rr, converged=cv.astype(bool), ms=ms.value)

    def get_x(self, i):
        nx, ny, nz = self.sizes[i]
        x = np.empty(nx * ny * nz)
        _check(lib().th_batch_get_x(self.h, i, _d(x)))
        return x

    def autotune(self, reps=5):
        r, c, v, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _check(lib().th_batch_autotune(self.h, reps, C.byref(r), C.byref(c), C.byref(v), C.byref(ms)))
        return dict(rows=r.value, cols_per_wave=c.value, vec_threads=v.value, vcycle_ms=ms.value)

    def shape(self):
        r, c, v = C.c_int(), C.c_int(), C.c_int()
        _check(lib().th_batch_shape(self.h, C.byref(r), C.byref(c), C.byref(v)))
        return r.value, c.value, v.value
