"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / host-core baseline, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

OK, E_CAP_NODES, E_CAP_NEIGHBOURS, E_DEPTH, E_RHO, E_STACK, E_ARG = 0, 1, 2, 3, 4, 5, 6


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
    sig = {
        "orc_num_threads": (i32, []),
        "orc_set_num_threads": (None, [i32]),
        "orc_bbox": (None, [i64, vp, vp, vp, vp]),
        "orc_lebesgue_key": (C.c_uint64, [d, d, d, vp]),
        "orc_lebesgue2hilbert": (C.c_uint64, [C.c_uint64]),
        "orc_keys": (None, [i64, vp, vp, vp, vp, i32, vp]),
        "orc_sort": (None, [i64, vp, vp]),
        "orc_tree_build": (i32, [i64, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]),
        "orc_com": (None, [i64, i64, vp, vp, vp, vp, vp, vp]),
        "orc_quadrupole": (None, [i64, i64, vp, vp, vp, vp, vp, vp, vp]),
        "orc_bh_v1": (i32, [i64, i64, vp, vp, vp, vp, vp, d, d, d, vp, vp, vp, vp]),
        "orc_bh_group": (i32, [i64, vp, i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, d, d, i32,
                               vp, vp, vp, vp, vp, vp]),
        "orc_bh_group_subset": (i32, [i64, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, d, d, i32, i64, vp,
                                      vp, vp, vp, vp, vp, vp]),
        "orc_direct": (None, [i64, vp, vp, vp, vp, d, i64, vp, vp, vp, vp, vp]),
        "orc_bench_sph_rhs": (i32, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32,
                                    d, i32, d, d, d, d, i32, vp]),
        "orc_bench_gravity": (i32, [i64, vp, vp, vp, vp, d, d, i32, i32, i32, i32, i64, vp, vp, vp]),
        "orc_nn_brute": (i32, [i64, vp, vp, vp, vp, i32, vp, vp]),
        "orc_nn_grid": (i32, [i64, vp, vp, vp, vp, i32, vp, vp]),
        "orc_nn_refwalk": (i32, [i64, vp, vp, vp, vp, vp, vp, vp, vp, d, i32, vp, vp]),
        "orc_h_factor_table": (None, [i32, i32, vp]),
        "orc_h_iterate": (i32, [i64, vp, vp, vp, vp, i32, i32, i32, i32, i32, vp, i32, d, vp, vp]),
        "orc_cubic_spline": (None, [d, d, d, d, d, vp, vp, vp]),
        "orc_sph_kernel": (i32, [i32, d, d, d, d, d, vp, vp, vp]),
        "orc_density": (i32, [i64, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, vp]),
        "orc_soundspeed": (None, [i64, i32, d, d, vp, vp, vp]),
        "orc_pressure": (None, [i64, i32, d, d, vp, vp, vp, vp]),
        "orc_forces": (None, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32,
                              d, d, vp, vp, vp, vp, vp, vp, vp]),
        "orc_pc_predictor": (None, [i64, d, vp, vp]),
        "orc_pc_corrector": (None, [i64, d, vp, vp]),
        "orc_time_step_terms": (None, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, d, vp]),
        "orc_leapfrog_update_x": (None, [i64, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_leapfrog_update_v": (None, [i64, d, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "orc_sph_features": (None, [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp]),
        "orc_mlp_forward": (None, [i64, i32, i32, vp, vp, vp, vp, vp, vp, d, vp, vp, vp]),
        "orc_mlp_backward": (d, [i64, i32, i32, vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp, vp,
                                 vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def P(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    LIB.orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(LIB.orc_num_threads())


def bbox(x, y, z) -> np.ndarray:
    x, y, z = f64(x), f64(y), f64(z)
    box = np.zeros(6)
    LIB.orc_bbox(len(x), P(x), P(y), P(z), P(box))
    return box


def lebesgue_key(x: float, y: float, z: float, box) -> int:
    box = f64(box)
    return int(LIB.orc_lebesgue_key(float(x), float(y), float(z), P(box)))


def lebesgue2hilbert(k: int) -> int:
    return int(LIB.orc_lebesgue2hilbert(C.c_uint64(int(k))))


def keys(x, y, z, box, curve: int) -> np.ndarray:
    x, y, z, box = f64(x), f64(y), f64(z), f64(box)
    out = np.empty(len(x), np.uint64)
    LIB.orc_keys(len(x), P(x), P(y), P(z), P(box), int(curve), P(out))
    return out


def sort(keys_: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys_, dtype=np.uint64)
    perm = np.empty(len(k), np.int32)
    LIB.orc_sort(len(k), P(k), P(perm))
    return perm


class Tree:
    """Reference insertion octree: nodes [0,n) particles, [n, n+ncells) cells, root n."""

    def __init__(self, x, y, z, box, cap_cells: int | None = None):
        x, y, z, box = f64(x), f64(y), f64(z), f64(box)
        n = len(x)
        cap = cap_cells if cap_cells is not None else 2 * n + 64
        tot = n + cap
        self.n = n
        self.child = np.empty(8 * tot, np.int32)
        self.level = np.empty(tot, np.int32)
        self.nx, self.ny, self.nz = np.empty(tot), np.empty(tot), np.empty(tot)
        nc = C.c_int64(0)
        self.status = LIB.orc_tree_build(n, P(x), P(y), P(z), P(box), cap, P(self.child), P(self.level),
                                         P(self.nx), P(self.ny), P(self.nz), C.byref(nc))
        self.ncells = int(nc.value)
        self.nm = np.zeros(tot)
        self.q = None

    def com(self, m) -> None:
        m = f64(m)
        self.nm[: self.n] = m
        LIB.orc_com(self.n, self.ncells, P(self.child), P(self.level), P(self.nm), P(self.nx),
                    P(self.ny), P(self.nz))

    def quadrupole(self) -> np.ndarray:
        tot = len(self.nm)
        self.q = np.zeros(6 * tot)
        LIB.orc_quadrupole(self.n, self.ncells, P(self.child), P(self.level), P(self.nm), P(self.nx),
                           P(self.ny), P(self.nz), P(self.q))
        return self.q


def bh_group(tree: Tree, perm, box, theta, smoothing, mac=0, group=64, quad=False, potential=False,
             vote=True):
    """Group walk (orc_bh_group).  vote=True: every member tests a cell with its own distance
    and the group accepts only if all do; vote=False: one test with the distance to the
    group's bounding box."""
    n = tree.n
    mac = int(mac) | (2 if vote else 0)
    perm = np.ascontiguousarray(perm, np.int32)
    box = f64(box)
    ax, ay, az = np.empty(n), np.empty(n), np.empty(n)
    phi = np.empty(n) if potential else None
    nci, npi = C.c_int64(0), C.c_int64(0)
    q = tree.q if quad else None
    if quad and q is None:
        q = tree.quadrupole()
    st = LIB.orc_bh_group(n, P(perm), int(group), tree.ncells, P(tree.child), P(tree.level), P(tree.nm),
                          P(tree.nx), P(tree.ny), P(tree.nz), P(q), P(box), float(theta),
                          float(smoothing), int(mac), P(ax), P(ay), P(az), P(phi), C.byref(nci),
                          C.byref(npi))
    return {"status": st, "ax": ax, "ay": ay, "az": az, "phi": phi, "n_cell": int(nci.value),
            "n_part": int(npi.value)}


def bh_v1(tree: Tree, radius, theta, smoothing, potential=False):
    n = tree.n
    ax, ay, az = np.empty(n), np.empty(n), np.empty(n)
    u = np.empty(n) if potential else None
    st = LIB.orc_bh_v1(n, tree.ncells, P(tree.child), P(tree.nm), P(tree.nx), P(tree.ny), P(tree.nz),
                       float(radius), float(theta), float(smoothing), P(ax), P(ay), P(az), P(u))
    return {"status": st, "ax": ax, "ay": ay, "az": az, "u": u}


def direct(x, y, z, m, smoothing, targets=None):
    x, y, z, m = f64(x), f64(y), f64(z), f64(m)
    t = None if targets is None else np.ascontiguousarray(targets, np.int32)
    nt = len(x) if t is None else len(t)
    ax, ay, az, phi = np.empty(nt), np.empty(nt), np.empty(nt), np.empty(nt)
    LIB.orc_direct(len(x), P(x), P(y), P(z), P(m), float(smoothing), nt, P(t), P(ax), P(ay), P(az), P(phi))
    return {"ax": ax, "ay": ay, "az": az, "phi": phi}


def _nn(fn, x, y, z, h, max_nb, lists=True):
    x, y, z, h = f64(x), f64(y), f64(z), f64(h)
    n = len(x)
    nnl = np.full((n, max_nb), -1, np.int32) if lists else None
    noi = np.empty(n, np.int32)
    st = fn(n, P(x), P(y), P(z), P(h), int(max_nb), P(nnl), P(noi))
    return st, nnl, noi


def nn_brute(x, y, z, h, max_nb=180, lists=True):
    return _nn(LIB.orc_nn_brute, x, y, z, h, max_nb, lists)


def nn_grid(x, y, z, h, max_nb=180, lists=True):
    return _nn(LIB.orc_nn_grid, x, y, z, h, max_nb, lists)


def nn_refwalk(x, y, z, h, tree: Tree, diam, max_nb=180):
    x, y, z, h = f64(x), f64(y), f64(z), f64(h)
    n = len(x)
    nnl = np.full((n, max_nb), -1, np.int32)
    noi = np.empty(n, np.int32)
    st = LIB.orc_nn_refwalk(n, P(x), P(y), P(z), P(h), P(tree.child), P(tree.nx), P(tree.ny), P(tree.nz),
                            float(diam), int(max_nb), P(nnl), P(noi))
    return st, nnl, noi


TMAX = 4096


def h_factor_table(n_target: int, tmax: int = TMAX) -> np.ndarray:
    t = np.empty(tmax + 1)
    LIB.orc_h_factor_table(int(n_target), int(tmax), P(t))
    return t


def h_iterate(x, y, z, h, n_target=64, band=(50, 80), max_counts=7, max_nb=180, sml_max=0.0):
    x, y, z = f64(x), f64(y), f64(z)
    h = f64(h).copy()
    n = len(x)
    table = h_factor_table(n_target)
    iters = np.empty(n, np.int32)
    noi = np.empty(n, np.int32)
    st = LIB.orc_h_iterate(n, P(x), P(y), P(z), P(h), int(n_target), int(band[0]), int(band[1]),
                           int(max_counts), int(max_nb), P(table), TMAX, float(sml_max), P(iters), P(noi))
    return st, h, iters, noi


def cubic_spline(r2, dx, dy, dz, h):
    W = np.zeros(1)
    g = np.zeros(3)
    dWdr = np.zeros(1)
    LIB.orc_cubic_spline(float(r2), float(dx), float(dy), float(dz), float(h), P(W), P(g), P(dWdr))
    return float(W[0]), g, float(dWdr[0])


# smoothing kernels (the reference's Smoothing::Kernel ids)
KERNEL_SPIKY, KERNEL_CUBIC_SPLINE, KERNEL_WENDLAND_C2, KERNEL_WENDLAND_C4, KERNEL_WENDLAND_C6 = 0, 1, 2, 3, 4


def sph_kernel(kind, r2, dx, dy, dz, h):
    W = np.zeros(1)
    g = np.zeros(3)
    dWdr = np.zeros(1)
    st = LIB.orc_sph_kernel(int(kind), float(r2), float(dx), float(dy), float(dz), float(h), P(W), P(g), P(dWdr))
    if st != 0:
        raise ValueError(f"unknown smoothing kernel {kind}")
    return float(W[0]), g, float(dWdr[0])


def density(x, y, z, m, h, nnl, noi, sml_avg=1, kernel=KERNEL_CUBIC_SPLINE):
    x, y, z, m, h = f64(x), f64(y), f64(z), f64(m), f64(h)
    nnl = np.ascontiguousarray(nnl, np.int32)
    noi = np.ascontiguousarray(noi, np.int32)
    rho = np.empty(len(x))
    st = LIB.orc_density(len(x), P(x), P(y), P(z), P(m), P(h), P(nnl), P(noi), nnl.shape[1], int(sml_avg),
                         int(kernel), P(rho))
    return st, rho


def soundspeed(eos, gamma, K, rho, p, cs):
    rho, p = f64(rho), f64(p)
    cs = f64(cs).copy()
    LIB.orc_soundspeed(len(rho), int(eos), float(gamma), float(K), P(rho), P(p), P(cs))
    return cs


def pressure(eos, gamma, K, rho, e, cs):
    rho, e, cs = f64(rho), f64(e), f64(cs)
    p = np.empty(len(rho))
    LIB.orc_pressure(len(rho), int(eos), float(gamma), float(K), P(rho), P(e), P(cs), P(p))
    return p


def forces(x, y, z, vx, vy, vz, m, h, rho, p, cs, nnl, noi, alpha=1.0, beta=2.0, sml_avg=1,
           kernel=KERNEL_CUBIC_SPLINE):
    arrs = [f64(a) for a in (x, y, z, vx, vy, vz, m, h, rho, p, cs)]
    nnl = np.ascontiguousarray(nnl, np.int32)
    noi = np.ascontiguousarray(noi, np.int32)
    n = len(arrs[0])
    out = [np.empty(n) for _ in range(7)]
    LIB.orc_forces(n, *[P(a) for a in arrs], P(nnl), P(noi), nnl.shape[1], int(sml_avg), int(kernel),
                   float(alpha), float(beta), *[P(o) for o in out])
    return dict(zip(["ax", "ay", "az", "drhodt", "dedt", "dsmldt", "muijmax"], out))


def sph_features(x, y, z, vx, vy, vz, m, h, rho, nnl, noi, sml_avg=1, kernel=KERNEL_CUBIC_SPLINE):
    """(curl v, grad rho) as (n, 3) arrays over the neighbour lists."""
    arrs = [f64(a) for a in (x, y, z, vx, vy, vz, m, h, rho)]
    nnl = np.ascontiguousarray(nnl, np.int32)
    noi = np.ascontiguousarray(noi, np.int32)
    n = len(arrs[0])
    curl = np.empty((n, 3))
    grad = np.empty((n, 3))
    LIB.orc_sph_features(n, *[P(a) for a in arrs], P(nnl), P(noi), nnl.shape[1], int(sml_avg), int(kernel),
                         P(curl), P(grad))
    return curl, grad


def mlp_forward(X, W1, b1, W2, b2, W3, b3):
    """FP64 forward of the 8 -> H -> H -> 1 tanh MLP (rows of X are samples)."""
    X = f64(X)
    nb, nin = X.shape
    nh = W1.shape[0]
    Y = np.empty(nb)
    A1 = np.empty((nb, nh))
    A2 = np.empty((nb, nh))
    LIB.orc_mlp_forward(nb, nin, nh, P(X), P(f64(W1)), P(f64(b1)), P(f64(W2)), P(f64(b2)), P(f64(W3)),
                        float(b3), P(Y), P(A1), P(A2))
    return Y


def mlp_backward(X, T, W1, b1, W2, b2, W3, b3):
    """(loss, dict of gradients) of mean 0.5 (y - t)^2, FP64."""
    X, T = f64(X), f64(T)
    nb, nin = X.shape
    nh = W1.shape[0]
    g = {"W1": np.zeros((nh, nin)), "b1": np.zeros(nh), "W2": np.zeros((nh, nh)), "b2": np.zeros(nh),
         "W3": np.zeros(nh), "b3": np.zeros(1)}
    L = LIB.orc_mlp_backward(nb, nin, nh, P(X), P(T), P(f64(W1)), P(f64(b1)), P(f64(W2)), P(f64(b2)),
                             P(f64(W3)), float(b3), P(g["W1"]), P(g["b1"]), P(g["W2"]), P(g["b2"]),
                             P(g["W3"]), P(g["b3"]))
    return L, g


PC_STATE = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "g_ax", "g_ay", "g_az", "e", "dedt")


def _ptr_array(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def pc_predictor(dt, s: dict, backup: list) -> None:
    """s: dict of the 14 PC_STATE arrays (float64, modified in place); backup: 11 arrays."""
    LIB.orc_pc_predictor(len(s["x"]), float(dt), C.cast(_ptr_array([s[k] for k in PC_STATE]), C.c_void_p),
                         C.cast(_ptr_array(backup), C.c_void_p))


def pc_corrector(dt, s: dict, backup: list) -> None:
    LIB.orc_pc_corrector(len(s["x"]), float(dt), C.cast(_ptr_array([s[k] for k in PC_STATE]), C.c_void_p),
                         C.cast(_ptr_array(backup), C.c_void_p))


def time_step(s: dict, alpha=1.0, beta=2.0, courant_fact=0.4, forces_fact=0.2, dt_max=-1.0):
    arrs = [f64(s[k]) for k in ("ax", "ay", "az", "g_ax", "g_ay", "g_az", "sml", "cs", "muijmax", "vx", "vy", "vz")]
    t = np.zeros(5)
    LIB.orc_time_step_terms(len(arrs[0]), *[P(a) for a in arrs], float(alpha), float(beta), P(t))
    dt = min(courant_fact * t[1], forces_fact * t[0], courant_fact * t[2])
    if t[3] > 0 and t[4] > 0:
        dt = min(dt, t[4] / t[3])
    if dt_max > 0:
        dt = min(dt, dt_max)
    return dt, t


def leapfrog_update_x(dt, s: dict) -> None:
    LIB.orc_leapfrog_update_x(len(s["x"]), float(dt), *[P(s[k]) for k in (
        "x", "y", "z", "vx", "vy", "vz", "g_ax", "g_ay", "g_az", "g_ax_old", "g_ay_old", "g_az_old")])


def leapfrog_update_v(dt, s: dict) -> None:
    LIB.orc_leapfrog_update_v(len(s["x"]), float(dt), *[P(s[k]) for k in (
        "vx", "vy", "vz", "g_ax", "g_ay", "g_az", "g_ax_old", "g_ay_old", "g_az_old")])


# ---- composite passes used by the tests and the host-core baseline ------------------
def sph_rhs(s: dict, n_target=64, band=(50, 80), max_counts=7, max_nb=180, eos=9, gamma=5.0 / 3.0,
            K=0.0, alpha=1.0, beta=2.0, variable_sml=True, cs_before_p=True, sml_max=0.0,
            kernel=KERNEL_CUBIC_SPLINE):
    """h-iteration -> neighbours -> density -> sound speed -> pressure -> forces, FP64."""
    h = f64(s["sml"])
    out = {}
    if variable_sml:
        st, h, iters, _ = h_iterate(s["x"], s["y"], s["z"], h, n_target, band, max_counts, max_nb, sml_max)
        out["h_status"] = st
        out["h_iters"] = iters
    st, nnl, noi = nn_grid(s["x"], s["y"], s["z"], h, max_nb)
    out["nn_status"] = st
    st, rho = density(s["x"], s["y"], s["z"], s["mass"], h, nnl, noi, kernel=kernel)
    out["rho_status"] = st
    cs = f64(s["cs"])
    p = f64(s["p"])
    if cs_before_p:
        cs = soundspeed(eos, gamma, K, rho, p, cs)
        p = pressure(eos, gamma, K, rho, s["e"], cs)
    else:
        p = pressure(eos, gamma, K, rho, s["e"], cs)
        cs = soundspeed(eos, gamma, K, rho, p, cs)
    f = forces(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["mass"], h, rho, p, cs, nnl, noi,
               alpha, beta, kernel=kernel)
    out.update({"sml": h, "nnl": nnl, "noi": noi, "rho": rho, "p": p, "cs": cs, **f})
    return out


def gravity_rhs(s: dict, theta, smoothing, curve=1, mac=0, group=64, quad=False, potential=False, vote=True):
    """bbox -> keys -> sort -> tree -> COM (+quadrupole) -> group Barnes-Hut, FP64."""
    box = bbox(s["x"], s["y"], s["z"])
    k = keys(s["x"], s["y"], s["z"], box, curve)
    perm = sort(k)
    t = Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    if quad:
        t.quadrupole()
    res = bh_group(t, perm, box, theta, smoothing, mac=mac, group=group, quad=quad, potential=potential,
                   vote=vote)
    res.update({"box": box, "keys": k, "perm": perm, "tree": t})
    return res


# ---- host-core baseline (timed at C level) ----------------------------------------------
def bench_sph_rhs(s: dict, reps: int, n_target=64, band=(50, 80), max_counts=7, max_nb=180, eos=9,
                  gamma=5.0 / 3.0, K=0.0, alpha=1.0, beta=2.0, sml_max=0.0) -> np.ndarray:
    """Wall seconds of `reps` full SPH RHS (orc_bench_sph_rhs) on the state s."""
    a = {k: f64(s[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "mass", "sml", "e", "p", "cs")}
    secs = np.zeros(int(reps))
    st = LIB.orc_bench_sph_rhs(len(a["x"]), *[P(a[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "mass",
                                                                 "sml", "e", "p", "cs")],
                               int(n_target), int(band[0]), int(band[1]), int(max_counts), int(max_nb),
                               float(sml_max), int(eos), float(gamma), float(K), float(alpha), float(beta),
                               int(reps), P(secs))
    if st != OK:
        raise RuntimeError(f"orc_bench_sph_rhs: status {st}")
    return secs


def bench_gravity(s: dict, theta, smoothing, curve=1, mac=0, quad=False, group=64, sample_every=1,
                  vote=True) -> dict:
    """Barnes-Hut RHS timed at C level: tree phases over all particles, walk over every
    sample_every-th group (orc_bench_gravity)."""
    a = {k: f64(s[k]) for k in ("x", "y", "z", "mass")}
    secs = np.zeros(2)
    walked = C.c_int64(0)
    counts = np.zeros(2, np.int64)
    st = LIB.orc_bench_gravity(len(a["x"]), P(a["x"]), P(a["y"]), P(a["z"]), P(a["mass"]), float(theta),
                               float(smoothing), int(curve), int(mac) | (2 if vote else 0), int(bool(quad)),
                               int(group), int(sample_every), P(secs), C.byref(walked), P(counts))
    if st != OK:
        raise RuntimeError(f"orc_bench_gravity: status {st}")
    return {"tree_s": float(secs[0]), "walk_s": float(secs[1]), "groups_walked": int(walked.value),
            "n_cell": int(counts[0]), "n_part": int(counts[1])}
