"""ctypes binding of the C-ABI in include/mhpc.h (libmhpc.so).

The library is the product: there is no Python or CPU fallback for any compute call.
If libmhpc.so is missing this module raises at import time.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libmhpc.so"

# status codes (mhpc_status)
OK = 0
E_CAP_NODES = 1
E_CAP_NEIGHBOURS = 2
E_DEPTH = 3
E_RHO_NONPOSITIVE = 4
E_STACK = 5
E_ARG = 6
E_HIP = 7
E_STATE = 8
E_SCRATCH = 9

CURVE_LEBESGUE, CURVE_HILBERT = 0, 1
EOS_POLYTROPIC, EOS_ISOTHERMAL, EOS_IDEAL_GAS = 0, 3, 9
EOS_LOCALLY_ISOTHERMAL, EOS_ISOTHERMAL_CONST_CS = 12, 100
MAC_CLASSIC, MAC_SQRT_THETA = 0, 1
BH_BOX, BH_VOTE, BH_PER_PARTICLE = 0, 1, 2
PREC_FP64, PREC_FP32_KAHAN = 0, 1
KERNEL_SPIKY, KERNEL_CUBIC_SPLINE, KERNEL_WENDLAND_C2, KERNEL_WENDLAND_C4, KERNEL_WENDLAND_C6 = 0, 1, 2, 3, 4

RHS_GRAVITY = 1 << 0
RHS_SPH = 1 << 1
RHS_VARIABLE_SML = 1 << 2
RHS_POTENTIAL = 1 << 3
RHS_TIMING = 1 << 4
RHS_LEARNED_H = 1 << 5
RHS_SPECULATIVE_NN = 1 << 6
RHS_FEATURES = 1 << 7

MLP_NPARAM = 4801   # W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3
BH_GROUP = 64       # particles sharing one Barnes-Hut acceptance test and list (MHPC_BH_GROUP)
NOALLOC_SORT_MAX = 2097152   # elements the no-allocation key sort takes (MHPC_NOALLOC_SORT_MAX)
MLP_HIDDEN = 64
MLP_IN = 8          # 7 features + 1 zero pad

FIELDS = [
    "x", "y", "z", "vx", "vy", "vz", "mass", "sml", "e", "rho", "p", "cs",
    "ax", "ay", "az", "g_ax", "g_ay", "g_az", "g_ax_old", "g_ay_old", "g_az_old",
    "drhodt", "dedt", "dsmldt", "u", "muijmax", "curlv", "gradrho", "noi", "key", "sorted", "level",
    "h_iters",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
FIELD_DTYPE = {name: np.float64 for name in FIELDS[: FIELD_ID["noi"]]}
FIELD_DTYPE.update({"noi": np.int32, "key": np.uint64, "sorted": np.int32, "level": np.int32,
                    "h_iters": np.int32})


class Params(C.Structure):
    _fields_ = [
        ("theta", C.c_double),
        ("smoothing", C.c_double),
        ("mac", C.c_int),
        ("quadrupole", C.c_int),
        ("G", C.c_double),
        ("curve", C.c_int),
        ("eos", C.c_int),
        ("gamma", C.c_double),
        ("K", C.c_double),
        ("cs_const", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("n_target", C.c_int),
        ("band_lo", C.c_int),
        ("band_hi", C.c_int),
        ("max_counts", C.c_int),
        ("max_nb", C.c_int),
        ("sml_avg", C.c_int),
        ("precision", C.c_int),
        ("cs_before_p", C.c_int),
        ("sml_max", C.c_double),
        ("bh_accept", C.c_int),
        ("smoothing_kernel", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Diagnostics(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double),
        ("potential", C.c_double),
        ("internal", C.c_double),
        ("total", C.c_double),
        ("momentum", C.c_double * 3),
        ("angular_momentum", C.c_double * 3),
        ("angular_momentum_norm", C.c_double),
        ("mass", C.c_double),
        ("com", C.c_double * 3),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if not isinstance(v, float) else v
        return out


class RhsStats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_float),
        ("ms_bbox", C.c_float),
        ("ms_keys", C.c_float),
        ("ms_sort", C.c_float),
        ("ms_gather", C.c_float),
        ("ms_tree", C.c_float),
        ("ms_com", C.c_float),
        ("ms_gravity", C.c_float),
        ("ms_mlp", C.c_float),
        ("ms_hiter", C.c_float),
        ("ms_nn", C.c_float),
        ("ms_density", C.c_float),
        ("ms_eos", C.c_float),
        ("ms_forces", C.c_float),
        ("ms_scatter", C.c_float),
        ("num_cells", C.c_int64),
        ("n_cell_interactions", C.c_int64),
        ("n_part_interactions", C.c_int64),
        ("n_neighbours", C.c_int64),
        ("n_h_counts", C.c_int64),
        ("max_noi", C.c_int32),
        ("h_passes", C.c_int32),
        ("ms_nn_walk", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built: the HIP library is required (no fallback). "
            "Run `python -c 'import __graft_entry__ as g; g.build()'` first.")
    lib = C.CDLL(str(LIB_PATH))
    vp, i64, i32, u32 = C.c_void_p, C.c_int64, C.c_int, C.c_uint32
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    sig = {
        "mhpc_default_params": (None, [C.POINTER(Params)]),
        "mhpc_create": (i32, [C.POINTER(Params), i64, i64, C.POINTER(vp)]),
        "mhpc_destroy": (i32, [vp]),
        "mhpc_set_params": (i32, [vp, C.POINTER(Params)]),
        "mhpc_get_params": (i32, [vp, C.POINTER(Params)]),
        "mhpc_set_num_particles": (i32, [vp, i64]),
        "mhpc_num_particles": (i64, [vp]),
        "mhpc_last_error": (C.c_char_p, [vp]),
        "mhpc_status_string": (C.c_char_p, [i32]),
        "mhpc_synchronize": (i32, [vp]),
        "mhpc_upload": (i32, [vp, i32, vp, i64]),
        "mhpc_download": (i32, [vp, i32, vp, i64]),
        "mhpc_field_ptr": (i32, [vp, i32, C.POINTER(vp)]),
        "mhpc_upload_device": (i32, [vp, i32, vp, i64]),
        "mhpc_bind_field": (i32, [vp, i32, vp]),
        "mhpc_device_alloc": (i32, [vp, i64, C.POINTER(vp)]),
        "mhpc_device_free": (i32, [vp, vp]),
        "mhpc_device_upload": (i32, [vp, vp, vp, i64]),
        "mhpc_device_download": (i32, [vp, vp, vp, i64]),
        "mhpc_download_neighbours": (i32, [vp, vp, vp, i64, i32]),
        "mhpc_tree_size": (i32, [vp, C.POINTER(i64)]),
        "mhpc_download_tree": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "mhpc_get_box": (i32, [vp, dp]),
        "mhpc_rhs": (i32, [vp, u32, C.POINTER(RhsStats)]),
        "mhpc_bbox": (i32, [vp, dp, fp]),
        "mhpc_keys": (i32, [vp, fp]),
        "mhpc_sort": (i32, [vp, fp]),
        "mhpc_build_tree": (i32, [vp, fp]),
        "mhpc_centres_of_mass": (i32, [vp, fp]),
        "mhpc_gravity": (i32, [vp, i32, fp]),
        "mhpc_h_iterate": (i32, [vp, fp]),
        "mhpc_neighbours": (i32, [vp, fp]),
        "mhpc_density": (i32, [vp, fp]),
        "mhpc_eos": (i32, [vp, fp]),
        "mhpc_internal_forces": (i32, [vp, fp]),
        "mhpc_compute_diagnostics": (i32, [vp, C.POINTER(Diagnostics)]),
        "mhpc_sort_keys_noalloc": (i32, [vp, i64, vp, vp, vp]),
        "mhpc_time_step": (i32, [vp, C.c_double, C.c_double, C.c_double, dp, dp]),
        "mhpc_pc_predictor": (i32, [vp, C.c_double]),
        "mhpc_pc_corrector": (i32, [vp, C.c_double]),
        "mhpc_pc_euler_step": (i32, [vp, u32, C.c_double, C.c_double, C.c_double, C.c_double, dp,
                                     C.POINTER(RhsStats), C.POINTER(RhsStats)]),
        "mhpc_leapfrog_update_x": (i32, [vp, C.c_double]),
        "mhpc_leapfrog_update_v": (i32, [vp, C.c_double]),
        "mhpc_mlp_init": (i32, [vp, C.c_uint64]),
        "mhpc_mlp_set_params": (i32, [vp, vp, i32]),
        "mhpc_mlp_get_params": (i32, [vp, vp, i32]),
        "mhpc_mlp_set_norm": (i32, [vp, vp]),
        "mhpc_mlp_get_norm": (i32, [vp, vp]),
        "mhpc_mlp_features": (i32, [vp, C.c_double]),
        "mhpc_mlp_download_features": (i32, [vp, vp, i64]),
        "mhpc_mlp_predict": (i32, [vp, fp]),
        "mhpc_mlp_train": (i32, [vp, i32, i32, C.c_float, C.c_uint64, fp]),
        "mhpc_mlp_forward": (i32, [vp, vp, i64, vp]),
        "mhpc_mlp_gradient": (i32, [vp, vp, vp, i64, vp, fp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


class MhpcError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = LIB.mhpc_status_string(status).decode()
        super().__init__(f"{where}: status {status} ({msg}) {detail}".strip())


def default_params(**overrides) -> Params:
    p = Params()
    LIB.mhpc_default_params(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Context:
    """One subdomain on the current GPU (mhpc_create / mhpc_destroy)."""

    def __init__(self, cap_particles: int, params: Params | None = None, cap_cells: int = 0, **overrides):
        self.params = params if params is not None else default_params()
        for k, v in overrides.items():
            setattr(self.params, k, v)
        h = C.c_void_p()
        st = LIB.mhpc_create(C.byref(self.params), int(cap_particles), int(cap_cells), C.byref(h))
        if st != OK:
            raise MhpcError(st, "mhpc_create")
        self._h = h
        self.cap = int(cap_particles)
        self.n = 0

    # -- lifetime --------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            LIB.mhpc_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, where: str) -> None:
        if st != OK:
            detail = LIB.mhpc_last_error(self._h).decode(errors="replace") if self._h else ""
            raise MhpcError(st, where, detail)

    # -- parameters / data -----------------------------------------------------------
    def set_params(self, **overrides) -> None:
        for k, v in overrides.items():
            setattr(self.params, k, v)
        self._check(LIB.mhpc_set_params(self._h, C.byref(self.params)), "mhpc_set_params")

    def set_num_particles(self, n: int) -> None:
        self._check(LIB.mhpc_set_num_particles(self._h, int(n)), "mhpc_set_num_particles")
        self.n = int(n)

    def upload(self, field: str, values) -> None:
        dt = FIELD_DTYPE[field]
        a = np.ascontiguousarray(values, dtype=dt)
        if a.shape != (self.n,):
            raise ValueError(f"{field}: expected shape ({self.n},), got {a.shape}")
        self._check(LIB.mhpc_upload(self._h, FIELD_ID[field], _ptr(a), self.n), f"upload {field}")

    def download(self, field: str) -> np.ndarray:
        a = np.empty(self.n, dtype=FIELD_DTYPE[field])
        self._check(LIB.mhpc_download(self._h, FIELD_ID[field], _ptr(a), self.n), f"download {field}")
        return a

    def load(self, **fields) -> None:
        """Set the particle count from the first field and upload every given field."""
        n = len(next(iter(fields.values())))
        self.set_num_particles(n)
        for k, v in fields.items():
            self.upload(k, v)

    def upload_device(self, field: str, device_ptr: int, n: int | None = None) -> None:
        """Device-to-device copy into a field (e.g. from a torch tensor's data_ptr())."""
        n = self.n if n is None else int(n)
        self._check(LIB.mhpc_upload_device(self._h, FIELD_ID[field], C.c_void_p(int(device_ptr)), n),
                    f"upload_device {field}")

    def bind_field(self, field: str, device_ptr: int | None) -> None:
        """Point a field at a caller-owned device array (None: the context's own array)."""
        self._check(LIB.mhpc_bind_field(self._h, FIELD_ID[field], C.c_void_p(device_ptr or None)),
                    f"bind_field {field}")

    def device_alloc(self, nbytes: int) -> int:
        """Zeroed device memory (mhpc_device_alloc); free with device_free."""
        p = C.c_void_p()
        self._check(LIB.mhpc_device_alloc(self._h, int(nbytes), C.byref(p)), "mhpc_device_alloc")
        return int(p.value)

    def device_free(self, device_ptr: int) -> None:
        self._check(LIB.mhpc_device_free(self._h, C.c_void_p(device_ptr)), "mhpc_device_free")

    def device_upload(self, device_ptr: int, host: np.ndarray) -> None:
        a = np.ascontiguousarray(host)
        self._check(LIB.mhpc_device_upload(self._h, C.c_void_p(device_ptr), _ptr(a), a.nbytes),
                    "mhpc_device_upload")

    def device_download(self, device_ptr: int, count: int, dtype=np.float64) -> np.ndarray:
        a = np.empty(int(count), dtype=dtype)
        self._check(LIB.mhpc_device_download(self._h, _ptr(a), C.c_void_p(device_ptr), a.nbytes),
                    "mhpc_device_download")
        return a

    def field_ptr(self, field: str) -> int:
        p = C.c_void_p()
        self._check(LIB.mhpc_field_ptr(self._h, FIELD_ID[field], C.byref(p)), "mhpc_field_ptr")
        return int(p.value or 0)

    # -- RHS and stages --------------------------------------------------------------
    def rhs(self, flags: int) -> dict:
        st = RhsStats()
        self._check(LIB.mhpc_rhs(self._h, C.c_uint32(flags), C.byref(st)), "mhpc_rhs")
        return st.as_dict()

    def _stage(self, name: str, *args) -> float:
        ms = C.c_float(0.0)
        self._check(getattr(LIB, name)(self._h, *args, C.byref(ms)), name)
        return float(ms.value)

    def bbox(self) -> np.ndarray:
        box = np.zeros(6)
        ms = C.c_float(0.0)
        self._check(LIB.mhpc_bbox(self._h, box.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms)), "mhpc_bbox")
        return box

    def keys(self) -> float:
        return self._stage("mhpc_keys")

    def sort(self) -> float:
        return self._stage("mhpc_sort")

    def build_tree(self) -> float:
        return self._stage("mhpc_build_tree")

    def centres_of_mass(self) -> float:
        return self._stage("mhpc_centres_of_mass")

    def gravity(self, potential: bool = False) -> float:
        return self._stage("mhpc_gravity", int(bool(potential)))

    def h_iterate(self) -> float:
        return self._stage("mhpc_h_iterate")

    def neighbours(self) -> float:
        return self._stage("mhpc_neighbours")

    def density(self) -> float:
        return self._stage("mhpc_density")

    def eos(self) -> float:
        return self._stage("mhpc_eos")

    def internal_forces(self) -> float:
        return self._stage("mhpc_internal_forces")

    def time_step(self, courant_fact: float = 0.4, forces_fact: float = 0.2, dt_max: float = -1.0):
        dt = C.c_double(0.0)
        parts = np.zeros(5)
        self._check(LIB.mhpc_time_step(self._h, float(courant_fact), float(forces_fact), float(dt_max),
                                       C.byref(dt), parts.ctypes.data_as(C.POINTER(C.c_double))),
                    "mhpc_time_step")
        return float(dt.value), parts

    def pc_euler_step(self, flags: int, courant_fact: float = 0.4, forces_fact: float = 0.2,
                      dt_max: float = -1.0, t_left: float | None = None,
                      train: dict | None = None) -> tuple[float, dict, dict]:
        """One predictor-corrector Euler step (predictor_corrector_euler.cpp): rhs with
        `flags`, CFL dt, predictor, rhs without self-gravity, corrector.  Returns
        (dt, stats of the first rhs, stats of the predictor rhs).  With RHS_LEARNED_H in
        `flags` and `train` given (mlp_train keyword arguments), the network is trained on
        the pairs each rhs's h-iteration emits, right after that rhs.  Without training and
        learned-h features the whole step is one call (mhpc_pc_euler_step), with no wait between
        its kernels."""
        if train is None and not flags & (RHS_LEARNED_H | RHS_FEATURES):
            dt = C.c_double(0.0)
            a, b = RhsStats(), RhsStats()
            self._check(LIB.mhpc_pc_euler_step(self._h, C.c_uint32(flags), float(courant_fact), float(forces_fact),
                                               float(dt_max), -1.0 if t_left is None else float(t_left),
                                               C.byref(dt), C.byref(a), C.byref(b)), "mhpc_pc_euler_step")
            return float(dt.value), a.as_dict(), b.as_dict()
        s0 = self.rhs(flags)
        if train is not None:
            s0["mlp_loss"] = self.mlp_train(**train)
        dt, _ = self.time_step(courant_fact, forces_fact, dt_max)
        if t_left is not None:
            dt = min(dt, t_left)
        self.pc_predictor(dt)
        s1 = self.rhs(flags & ~RHS_GRAVITY)
        if train is not None:
            s1["mlp_loss"] = self.mlp_train(**train)
        self.pc_corrector(dt)
        return dt, s0, s1

    # ---- learned smoothing length
    def mlp_init(self, seed: int = 1) -> None:
        self._check(LIB.mhpc_mlp_init(self._h, C.c_uint64(seed)), "mhpc_mlp_init")

    def mlp_set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != MLP_NPARAM:
            raise ValueError(f"expected {MLP_NPARAM} parameters, got {p.size}")
        self._check(LIB.mhpc_mlp_set_params(self._h, _ptr(p), MLP_NPARAM), "mhpc_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAM, np.float32)
        self._check(LIB.mhpc_mlp_get_params(self._h, _ptr(p), MLP_NPARAM), "mhpc_mlp_get_params")
        return p

    def mlp_set_norm(self, norm) -> None:
        a = np.ascontiguousarray(norm, dtype=np.float32)
        if a.size != 18:
            raise ValueError("norm has 18 entries")
        self._check(LIB.mhpc_mlp_set_norm(self._h, _ptr(a)), "mhpc_mlp_set_norm")

    def mlp_get_norm(self) -> np.ndarray:
        a = np.zeros(18, np.float32)
        self._check(LIB.mhpc_mlp_get_norm(self._h, _ptr(a)), "mhpc_mlp_get_norm")
        return a

    def mlp_features(self, dt: float) -> None:
        self._check(LIB.mhpc_mlp_features(self._h, float(dt)), "mhpc_mlp_features")

    def mlp_download_features(self) -> np.ndarray:
        X = np.zeros((self.n, MLP_IN), np.float32)
        self._check(LIB.mhpc_mlp_download_features(self._h, _ptr(X), self.n), "mhpc_mlp_download_features")
        return X

    def mlp_predict(self) -> float:
        return self._stage("mhpc_mlp_predict")

    def mlp_train(self, nsamples: int = 65536, iters: int = 50, lr: float = 1e-3, seed: int = 1) -> float:
        loss = C.c_float(0.0)
        self._check(LIB.mhpc_mlp_train(self._h, int(nsamples), int(iters), C.c_float(lr), C.c_uint64(seed),
                                       C.byref(loss)), "mhpc_mlp_train")
        return float(loss.value)

    def mlp_forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        assert X.shape == (n, MLP_IN)
        Y = np.zeros(n, np.float32)
        self._check(LIB.mhpc_mlp_forward(self._h, _ptr(X), n, _ptr(Y)), "mhpc_mlp_forward")
        return Y

    def mlp_gradient(self, X, T) -> tuple[np.ndarray, float]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        n = X.shape[0]
        assert X.shape == (n, MLP_IN) and T.shape == (n,)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float(0.0)
        self._check(LIB.mhpc_mlp_gradient(self._h, _ptr(X), _ptr(T), n, _ptr(g), C.byref(loss)),
                    "mhpc_mlp_gradient")
        return g, float(loss.value)

    def pc_predictor(self, dt: float) -> None:
        self._check(LIB.mhpc_pc_predictor(self._h, float(dt)), "mhpc_pc_predictor")

    def pc_corrector(self, dt: float) -> None:
        self._check(LIB.mhpc_pc_corrector(self._h, float(dt)), "mhpc_pc_corrector")

    def leapfrog_update_x(self, dt: float) -> None:
        self._check(LIB.mhpc_leapfrog_update_x(self._h, float(dt)), "mhpc_leapfrog_update_x")

    def leapfrog_update_v(self, dt: float) -> None:
        self._check(LIB.mhpc_leapfrog_update_v(self._h, float(dt)), "mhpc_leapfrog_update_v")

    def diagnostics(self) -> dict:
        d = Diagnostics()
        self._check(LIB.mhpc_compute_diagnostics(self._h, C.byref(d)), "mhpc_compute_diagnostics")
        return d.as_dict()

    def synchronize(self) -> None:
        self._check(LIB.mhpc_synchronize(self._h), "mhpc_synchronize")

    def get_box(self) -> np.ndarray:
        box = np.zeros(6)
        self._check(LIB.mhpc_get_box(self._h, box.ctypes.data_as(C.POINTER(C.c_double))), "mhpc_get_box")
        return box

    def neighbour_lists(self, max_nb: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        """(nnl[n, max_nb], noi[n]) in the reference layout, original indices, -1 padded."""
        max_nb = int(max_nb or self.params.max_nb)
        nnl = np.empty((self.n, max_nb), dtype=np.int32)
        noi = np.empty(self.n, dtype=np.int32)
        self._check(LIB.mhpc_download_neighbours(self._h, _ptr(nnl), _ptr(noi), self.n, max_nb),
                    "mhpc_download_neighbours")
        return nnl, noi

    def tree(self) -> dict:
        nc = C.c_int64(0)
        self._check(LIB.mhpc_tree_size(self._h, C.byref(nc)), "mhpc_tree_size")
        n = int(nc.value)
        out = {
            "level": np.empty(n, np.int32), "start": np.empty(n, np.int32), "end": np.empty(n, np.int32),
            "child": np.empty((n, 8), np.int32), "mass": np.empty(n), "x": np.empty(n), "y": np.empty(n),
            "z": np.empty(n), "quad": np.empty((6, n)),
        }
        self._check(LIB.mhpc_download_tree(
            self._h, _ptr(out["level"]), _ptr(out["start"]), _ptr(out["end"]), _ptr(out["child"]),
            _ptr(out["mass"]), _ptr(out["x"]), _ptr(out["y"]), _ptr(out["z"]), _ptr(out["quad"])),
            "mhpc_download_tree")
        return out
