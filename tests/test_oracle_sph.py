"""Pins the oracle's neighbour search, h-iteration, kernel, density, EOS, forces, leapfrog
and MLP gradients (CPU only), against brute force, closed forms and invariants."""
import math

import numpy as np
import pytest

from miluphpc_amd import ics


def as_sets(nnl, noi):
    return [frozenset(nnl[i, : noi[i]].tolist()) for i in range(len(noi))]


def fma_exact(a, b, c):
    """a * b + c rounded once (exact rationals, then Python's correctly rounded float())."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def numpy_brute_sets(x, y, z, h):
    """Brute-force sets with the neighbour test d^2 = fma(dx, dx, fma(dy, dy, dz^2)) < h^2:
    numpy's unfused d^2 differs from it by a few ulps at most, so only pairs that close to
    either bound are decided again with the exactly rounded fused form."""
    n = len(x)
    out = []
    for i in range(n):
        dx = x[i] - x
        dy = y[i] - y
        dz = z[i] - z
        d = dx * dx + dy * dy + dz * dz
        hi2, hj2 = h[i] * h[i], h * h
        ok = (d < hi2) & (d < hj2)
        near = np.nonzero((np.abs(d - hi2) <= 1e-12 * hi2) | (np.abs(d - hj2) <= 1e-12 * hj2))[0]
        for j in near:
            dj = fma_exact(dx[j], dx[j], fma_exact(dy[j], dy[j], dz[j] * dz[j]))
            ok[j] = (dj < hi2) and (dj < hj2[j])
        ok[i] = False
        out.append(frozenset(np.nonzero(ok)[0].tolist()))
    return out


@pytest.fixture(scope="module")
def blob():
    rng = np.random.default_rng(21)
    n = 3000
    x, y, z = (rng.normal(size=n) for _ in range(3))
    h = 0.25 * np.exp(0.3 * rng.normal(size=n))
    return x, y, z, h


def test_nn_brute_matches_numpy(orc, blob):
    x, y, z, h = blob
    st, nnl, noi = orc.nn_brute(x, y, z, h, max_nb=400)
    assert st == 0
    assert as_sets(nnl, noi) == numpy_brute_sets(x, y, z, h)


def fused_tie_cases(count, seed=5):
    """Offsets (dx, dy, dz) and radii h for which the fused d^2 = fma(dx, dx, fma(dy, dy, dz^2))
    and the unfused dx^2 + dy^2 + dz^2 fall on either side of fl(h * h): found by search, each
    case decided one way by the fused form and the other way by the unfused one."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        dx, dy, dz = (float(v) for v in rng.uniform(0.1, 1.0, 3))
        df = fma_exact(dx, dx, fma_exact(dy, dy, dz * dz))
        du = dx * dx + dy * dy + dz * dz
        if df == du:
            continue
        hi = max(df, du)   # want fl(h*h) == the larger one: then only the smaller one is below it
        h = math.sqrt(hi)
        for _ in range(8):
            if h * h == hi:
                out.append((dx, dy, dz, h, df < du))
                break
            h = np.nextafter(h, np.inf) if h * h < hi else np.nextafter(h, -np.inf)
    return out


def test_nn_pair_test_uses_the_fused_distance(orc):
    """The neighbour test's d^2 is fma(dx, dx, fma(dy, dy, dz^2)): on pairs where that and the
    unfused sum straddle fl(h^2), the oracle's brute force and grid search decide as the fused
    form does."""
    for dx, dy, dz, h, fused_is_pair in fused_tie_cases(12):
        x, y, z = np.array([0.0, dx]), np.array([0.0, dy]), np.array([0.0, dz])
        hh = np.array([h, h])
        for nn in (orc.nn_brute, orc.nn_grid):
            st, _, noi = nn(x, y, z, hh, max_nb=4)
            assert st == 0
            assert noi.tolist() == ([1, 1] if fused_is_pair else [0, 0]), (dx, dy, dz, h)


def test_nn_grid_equals_brute(orc, blob):
    x, y, z, h = blob
    _, nb, kb = orc.nn_brute(x, y, z, h, max_nb=400)
    st, ng, kg = orc.nn_grid(x, y, z, h, max_nb=400)
    assert st == 0
    assert np.array_equal(kb, kg)
    assert np.array_equal(nb, ng)  # both sorted by j


@pytest.mark.parametrize("com", [False, True])
def test_reference_tree_walk_equals_brute(orc, blob, com):
    x, y, z, h = blob
    box = orc.bbox(x, y, z)
    t = orc.Tree(x, y, z, box)
    if com:
        t.com(np.full(len(x), 1.0 / len(x)))
    diam = abs(box[1]) + abs(box[0])
    st, nw, kw = orc.nn_refwalk(x, y, z, h, t, diam, max_nb=400)
    _, nb, kb = orc.nn_brute(x, y, z, h, max_nb=400)
    assert st == 0
    assert as_sets(nw, kw) == as_sets(nb, kb)


def test_nn_capacity_error(orc, blob):
    x, y, z, h = blob
    st, _, _ = orc.nn_grid(x, y, z, h * 3, max_nb=8)
    assert st == orc.E_CAP_NEIGHBOURS


def test_cubic_spline_closed_forms(orc):
    h = 0.37
    W0, _, _ = orc.cubic_spline(0.0, 0, 0, 0, h)
    assert W0 == pytest.approx(8 / (math.pi * h ** 3), rel=1e-15)
    Wh2, _, _ = orc.cubic_spline((h / 2) ** 2, h / 2, 0, 0, h)
    assert Wh2 == pytest.approx(2 / (math.pi * h ** 3), rel=1e-14)
    Wh, _, _ = orc.cubic_spline(h * h, h, 0, 0, h)
    assert Wh == 0.0
    # normalisation: int W dV = 1
    r = np.linspace(0, h, 20001)
    W = np.array([orc.cubic_spline(rr * rr, rr, 0, 0, h)[0] for rr in r[::10]])
    rr = r[::10]
    integral = np.trapezoid(4 * math.pi * rr * rr * W, rr)
    assert integral == pytest.approx(1.0, abs=2e-5)
    # dW/dr by central differences, and grad W = dW/dr * dx / r
    for q in (0.2, 0.45, 0.7, 0.95):
        r0, dr = q * h, 1e-7
        Wp = orc.cubic_spline((r0 + dr) ** 2, r0 + dr, 0, 0, h)[0]
        Wm = orc.cubic_spline((r0 - dr) ** 2, r0 - dr, 0, 0, h)[0]
        W, g, dWdr = orc.cubic_spline(r0 * r0, r0 * 0.6, r0 * 0.8, 0.0, h)
        assert dWdr == pytest.approx((Wp - Wm) / (2 * dr), rel=1e-6)
        assert np.allclose(g, dWdr / r0 * np.array([0.6 * r0, 0.8 * r0, 0.0]), rtol=1e-14)


# W(0) h^3 of each kernel from its published normalisation (3D; Desbrun & Cani 1996 for the
# spiky kernel, Dehnen & Aly 2012 Table 1 for the Wendland functions), and W(q) h^3 in closed
# form -- the shapes the oracle's expressions must reproduce
KERNEL_SHAPES = {
    0: (15 / math.pi, lambda q: (1 - q) ** 3),
    1: (8 / math.pi, lambda q: 1 - 6 * q * q + 6 * q ** 3 if q <= 0.5 else 2 * (1 - q) ** 3),
    2: (21 / (2 * math.pi), lambda q: (1 - q) ** 4 * (1 + 4 * q)),
    3: (495 / (32 * math.pi), lambda q: (1 - q) ** 6 * (1 + 6 * q + 35 / 3 * q * q)),
    4: (1365 / (64 * math.pi), lambda q: (1 - q) ** 8 * (1 + 8 * q + 25 * q * q + 32 * q ** 3)),
}


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_smoothing_kernels_normalised(orc, kind):
    """Every kernel: W(0) = sigma / h^3, the published shape at interior points, W = 0 from
    r = h on, int W dV = 1 by quadrature, and grad W = dW/dr dx / r.  dW/dr equals the
    central difference of W for all but the Wendland C4 kernel, whose gradient expression in
    the reference is not its derivative: there the difference of W matches the analytic
    derivative -(56/3) sigma/h^4 q (1-q)^5 (1+5q) instead, and the restated expression stays
    unpinned."""
    h = 0.37
    sigma, shape = KERNEL_SHAPES[kind]
    W0 = orc.sph_kernel(kind, 0.0, 0, 0, 0, h)[0]
    assert W0 == pytest.approx(sigma / h ** 3, rel=1e-14)
    for q in (0.1, 0.3, 0.5, 0.62, 0.9):
        W = orc.sph_kernel(kind, (q * h) ** 2, q * h, 0, 0, h)[0]
        assert W == pytest.approx(sigma / h ** 3 * shape(q), rel=1e-12)
    assert orc.sph_kernel(kind, h * h, h, 0, 0, h)[0] == 0.0
    assert orc.sph_kernel(kind, (1.01 * h) ** 2, 1.01 * h, 0, 0, h)[0] == 0.0
    rr = np.linspace(0, h, 4001)
    Wr = np.array([orc.sph_kernel(kind, r * r, r, 0, 0, h)[0] for r in rr])
    assert np.trapezoid(4 * math.pi * rr * rr * Wr, rr) == pytest.approx(1.0, abs=2e-5)
    for q in (0.2, 0.45, 0.7, 0.95):
        r0, dr = q * h, 1e-7
        Wp = orc.sph_kernel(kind, (r0 + dr) ** 2, r0 + dr, 0, 0, h)[0]
        Wm = orc.sph_kernel(kind, (r0 - dr) ** 2, r0 - dr, 0, 0, h)[0]
        fd = (Wp - Wm) / (2 * dr)
        W, g, dWdr = orc.sph_kernel(kind, r0 * r0, r0 * 0.6, r0 * 0.8, 0.0, h)
        assert np.allclose(g, dWdr / r0 * np.array([0.6 * r0, 0.8 * r0, 0.0]), rtol=1e-14)
        if kind == 3:
            assert fd == pytest.approx(-(56 / 3) * sigma / h ** 4 * q * (1 - q) ** 5 * (1 + 5 * q), rel=1e-6)
            assert abs(dWdr - fd) > 1e-3 * abs(fd)
        else:
            assert dWdr == pytest.approx(fd, rel=1e-6)
    with pytest.raises(ValueError):
        orc.sph_kernel(5, 0.01, 0.1, 0, 0, h)


def lattice_state(nside, n_nb=64):
    s = ics.sedov_lattice(nside, n_nb=n_nb)
    s["e"] = np.full(len(s["x"]), 1.0)  # uniform pressure
    return s


def interior(s, margin):
    return ((np.abs(s["x"]) < 0.5 - margin) & (np.abs(s["y"]) < 0.5 - margin)
            & (np.abs(s["z"]) < 0.5 - margin))


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4])
def test_lattice_density_and_zero_force(orc, kernel):
    s = lattice_state(24)
    out = orc.sph_rhs(s, variable_sml=False, kernel=kernel)
    inner = interior(s, 2.5 * s["sml"][0])
    assert out["rho_status"] == 0
    # rho = 1 in the interior up to the kernel's discretisation error (the spiky kernel,
    # peaked at r = 0, has the largest)
    assert np.all(np.abs(out["rho"][inner] - 1.0) < (0.12 if kernel == 0 else 0.03))
    # uniform density and pressure: the forces cancel by symmetry
    a = np.sqrt(out["ax"] ** 2 + out["ay"] ** 2 + out["az"] ** 2)
    scale = np.abs(out["p"]).max() / out["rho"].max() / s["sml"][0]
    assert np.max(a[inner]) < 1e-9 * scale


def test_eos_closed_forms(orc):
    rng = np.random.default_rng(22)
    rho = rng.uniform(0.5, 2, 100)
    e = rng.uniform(0.1, 3, 100)
    p_prev = rng.uniform(0.1, 1, 100)
    cs = np.zeros(100)
    g = 5 / 3
    assert np.array_equal(orc.pressure(9, g, 0, rho, e, cs), (g - 1) * rho * e)
    assert np.array_equal(orc.soundspeed(9, g, 0, rho, p_prev, cs), np.sqrt(g * p_prev / rho))
    c = np.full(100, 0.167)
    assert np.array_equal(orc.pressure(100, g, 0, rho, e, c), c * c * rho)
    assert np.array_equal(orc.soundspeed(100, g, 0, rho, p_prev, c), c)
    assert np.array_equal(orc.pressure(3, g, 0, rho, e, c), 41255.407 * rho)
    assert np.allclose(orc.pressure(0, 1.4, 2.0, rho, e, c), 2.0 * rho ** 1.4, rtol=1e-15)


@pytest.fixture(scope="module")
def sedov_small():
    s = ics.sedov_lattice(16)
    rng = np.random.default_rng(23)
    n = len(s["x"])
    # a moving, compressed, unequal-mass state exercises every term
    for k in ("vx", "vy", "vz"):
        s[k] = 0.3 * rng.normal(size=n)
    for k in ("x", "y", "z"):
        s[k] = s[k] + 0.2 / 16 * rng.uniform(-1, 1, n)
    s["mass"] = s["mass"] * rng.uniform(0.8, 1.2, n)
    s["e"] = s["e"] + rng.uniform(0, 0.5, n)
    s["p"] = rng.uniform(0.1, 0.5, n)
    return s


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4])
def test_forces_conserve_momentum_and_energy(orc, sedov_small, kernel):
    s = sedov_small
    out = orc.sph_rhs(s, variable_sml=True, kernel=kernel)
    m = s["mass"]
    for k in ("ax", "ay", "az"):
        tot = (m * out[k]).sum()
        assert abs(tot) < 1e-12 * (m * np.abs(out[k])).sum()
    # d/dt sum m (e + v^2 / 2) = sum m (de/dt + v . a) = 0
    dE = (m * (out["dedt"] + s["vx"] * out["ax"] + s["vy"] * out["ay"] + s["vz"] * out["az"])).sum()
    scale = (m * np.abs(out["dedt"])).sum()
    assert abs(dE) < 1e-12 * scale
    # continuity: sum m drho/dt / rho ... each pair term is antisymmetric in v_ij . grad W
    # when rho_i = rho_j; here check drho/dt against the direct definition on 5 particles
    nnl, noi = out["nnl"], out["noi"]
    for i in (0, 100, 777, 2000, 4000):
        acc = 0.0
        for j in nnl[i, : noi[i]]:
            h = 0.5 * (out["sml"][i] + out["sml"][j])
            dr = np.array([s["x"][i] - s["x"][j], s["y"][i] - s["y"][j], s["z"][i] - s["z"][j]])
            r = np.sqrt(dr @ dr)
            dWdr = orc.sph_kernel(kernel, r * r, dr[0], dr[1], dr[2], h)[2]
            dv = np.array([s["vx"][i] - s["vx"][j], s["vy"][i] - s["vy"][j], s["vz"][i] - s["vz"][j]])
            acc += out["rho"][i] / out["rho"][j] * m[j] * (dv @ (dWdr / r * dr))
        assert acc == pytest.approx(out["drhodt"][i], rel=1e-10, abs=1e-12)


def test_h_iteration_fixed_point(orc, sedov_small):
    s = sedov_small
    st, h, iters, noi = orc.h_iterate(s["x"], s["y"], s["z"], s["sml"] * 1.3)
    assert st == 0
    # converged particles satisfy the band with their final h (their own last count)
    done = iters < 7
    assert np.all((noi[done] >= 50) & (noi[done] <= 80))
    assert np.mean(iters) < 4
    # a second run from the converged h changes nothing for particles whose neighbours
    # also kept their h (the Jacobi fixed point)
    st2, h2, iters2, _ = orc.h_iterate(s["x"], s["y"], s["z"], h)
    assert np.mean(iters2) < 1.5


def test_h_iteration_uniform_lattice_converges_to_one_h(orc):
    s = ics.sedov_lattice(20)
    inner = interior(s, 0.25)
    st, h, iters, noi = orc.h_iterate(s["x"], s["y"], s["z"], s["sml"] * 0.8)
    assert st == 0
    assert np.ptp(h[inner]) == 0.0
    assert np.all(noi[inner] == noi[inner][0])


def test_h_iteration_sml_max_bounds_free_boundary(orc):
    """On a free lattice boundary the corner particles cannot reach the band under
    d < min(h_i, h_j): without a bound their h grows on every pass; with sml_max it stops
    exactly at the bound, and particles that converge below the bound are unchanged."""
    s = ics.sedov_lattice(12)
    h0 = s["sml"].copy()
    st, h_free, it_free, _ = orc.h_iterate(s["x"], s["y"], s["z"], h0)
    hmax = 1.5 * h0.max()
    st, h_cap, it_cap, _ = orc.h_iterate(s["x"], s["y"], s["z"], h0, sml_max=hmax)
    corner = np.argmin(s["x"] + s["y"] + s["z"])
    assert h_free[corner] > 2.0 * hmax            # unbounded growth without the cap
    assert h_cap[corner] == hmax                   # stops exactly at the bound
    assert h_cap.max() == hmax
    below = h_free < hmax
    # particles whose own and neighbours' iterations stayed below the bound: same h
    assert np.mean(h_cap[below] == h_free[below]) > 0.9
    assert np.all(it_cap[below & (it_free == 1)] == 1)


def test_leapfrog_constant_acceleration(orc):
    n = 10
    s = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "g_ax", "g_ay", "g_az",
                                  "g_ax_old", "g_ay_old", "g_az_old")}
    s["vx"][:] = 1.0
    s["g_ax"][:] = -2.0
    t, dt = 0.0, 0.125
    for _ in range(8):
        orc.leapfrog_update_x(dt, s)
        orc.leapfrog_update_v(dt, s)
        t += dt
    assert np.allclose(s["x"], t - t * t, rtol=0, atol=1e-14)
    assert np.allclose(s["vx"], 1.0 - 2.0 * t, rtol=0, atol=1e-14)


def test_mlp_gradient_finite_differences(orc):
    import ctypes as C

    rng = np.random.default_rng(24)
    nb, nin, nh = 17, 7, 8
    X = rng.normal(size=(nb, nin))
    T = rng.normal(size=nb)
    W1, b1 = rng.normal(size=(nh, nin)) * 0.4, rng.normal(size=nh) * 0.1
    W2, b2 = rng.normal(size=(nh, nh)) * 0.4, rng.normal(size=nh) * 0.1
    W3, b3 = rng.normal(size=nh) * 0.4, 0.05
    P = orc.P

    def loss(W1, b1, W2, b2, W3, b3):
        Y = np.empty(nb)
        A1 = np.empty((nb, nh))
        A2 = np.empty((nb, nh))
        orc.LIB.orc_mlp_forward(nb, nin, nh, P(X), P(W1), P(b1), P(W2), P(b2), P(W3), b3, P(Y), P(A1), P(A2))
        return 0.5 * np.mean((Y - T) ** 2)

    g = [np.zeros_like(W1), np.zeros_like(b1), np.zeros_like(W2), np.zeros_like(b2), np.zeros_like(W3)]
    gb3 = np.zeros(1)
    L = orc.LIB.orc_mlp_backward(nb, nin, nh, P(X), P(T), P(W1), P(b1), P(W2), P(b2), P(W3), b3,
                                 P(g[0]), P(g[1]), P(g[2]), P(g[3]), P(g[4]), P(gb3))
    assert L == pytest.approx(loss(W1, b1, W2, b2, W3, b3), rel=1e-14)
    params = [W1, b1, W2, b2, W3]
    eps = 1e-6
    for pi, prm in enumerate(params):
        flat = prm.reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[idx]
            flat[idx] = old + eps
            lp = loss(W1, b1, W2, b2, W3, b3)
            flat[idx] = old - eps
            lm = loss(W1, b1, W2, b2, W3, b3)
            flat[idx] = old
            assert g[pi].reshape(-1)[idx] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5, abs=1e-9)
    lp = loss(W1, b1, W2, b2, W3, b3 + eps)
    lm = loss(W1, b1, W2, b2, W3, b3 - eps)
    assert gb3[0] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5)


def test_pc_euler_constant_acceleration_closed_form(orc):
    """Heun's predictor-corrector is exact for constant acceleration (a + g)."""
    n = 7
    rng = np.random.default_rng(25)
    s = {k: rng.normal(size=n) for k in orc.PC_STATE}
    x0 = {k: s[k].copy() for k in s}
    s["e"] = np.abs(s["e"]) + 1.0
    x0["e"] = s["e"].copy()
    backup = [np.zeros(n) for _ in range(11)]
    dt = 0.01
    orc.pc_predictor(dt, s, backup)
    # the rhs on the predicted state returns the same (constant) derivatives
    for k in ("ax", "ay", "az", "dedt"):
        s[k] = x0[k].copy()
    orc.pc_corrector(dt, s, backup)
    for d in ("x", "y", "z"):
        a = x0["a" + d] + x0["g_a" + d]
        assert np.allclose(s[d], x0[d] + dt * x0["v" + d] + 0.5 * dt * dt * a, rtol=0, atol=1e-15)
        assert np.allclose(s["v" + d], x0["v" + d] + dt * a, rtol=0, atol=1e-15)
        assert np.allclose(s["a" + d], x0["a" + d] + x0["g_a" + d], rtol=0, atol=1e-15)
    assert np.allclose(s["e"], np.maximum(x0["e"] + dt * x0["dedt"], 1e-6), rtol=0, atol=1e-14)


def test_time_step_terms_hand_values(orc):
    s = {"ax": np.array([3.0, 0.0]), "ay": np.array([4.0, 0.0]), "az": np.zeros(2),
         "g_ax": np.zeros(2), "g_ay": np.zeros(2), "g_az": np.array([0.0, 0.25]),
         "sml": np.array([0.5, 0.1]), "cs": np.array([2.0, 0.5]), "muijmax": np.array([0.0, 1.0]),
         "vx": np.array([0.0, 3.0]), "vy": np.array([0.0, 4.0]), "vz": np.zeros(2)}
    dt, t = orc.time_step(s, alpha=1.0, beta=2.0, courant_fact=0.3, forces_fact=0.2)
    # |a| = 5 and 0.25: sqrt(0.5/5) = 0.316.., sqrt(0.1/0.25) = 0.632..
    assert t[0] == pytest.approx(math.sqrt(0.1))
    assert t[1] == pytest.approx(0.2)            # min(0.5/2, 0.1/0.5)
    assert t[2] == pytest.approx(0.1 / (0.5 + 1.2 * (0.5 + 2.0)))   # 0.1/3.5
    assert (t[3], t[4]) == (5.0, 0.5)
    assert dt == pytest.approx(min(0.3 * 0.2, 0.2 * math.sqrt(0.1), 0.3 * 0.1 / 3.5, 0.5 / 5.0))


def _lattice(nside, h_factor=2.5):
    dx = 1.0 / nside
    g = (np.arange(nside) + 0.5) * dx
    x, y, z = [a.ravel().copy() for a in np.meshgrid(g, g, g, indexing="ij")]
    return x, y, z, dx, h_factor * dx


def test_sph_features_linear_fields(orc):
    """Rigid rotation v = Omega x r has curl v = 2 Omega and a linear density rho0 + g.r has
    grad rho = g: the oracle's SPH estimates on a lattice interior reproduce both (to the
    lattice's kernel-moment error), in direction and sign."""
    x, y, z, dx, h0 = _lattice(20)
    n = len(x)
    h = np.full(n, h0)
    om = np.array([0.3, -0.2, 1.0])
    r = np.stack([x, y, z], axis=1) - 0.5
    v = np.cross(om, r)
    gvec = np.array([0.7, -0.4, 0.25])
    rho = 1.0 + r @ gvec
    m = rho * dx ** 3
    st, nnl, noi = orc.nn_grid(x, y, z, h, 180)
    assert st == 0
    curl, grad = orc.sph_features(x, y, z, v[:, 0], v[:, 1], v[:, 2], m, h, rho, nnl, noi)
    inner = np.all(np.abs(r) < 0.5 - 1.5 * h0, axis=1)
    assert inner.sum() > 500
    assert np.allclose(curl[inner], 2.0 * om, rtol=0, atol=0.03 * np.linalg.norm(2 * om))
    assert np.allclose(grad[inner], gvec, rtol=0, atol=0.03 * np.linalg.norm(gvec))
    # uniform state: both vanish by symmetry
    curl0, grad0 = orc.sph_features(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, dx ** 3), h,
                                    np.ones(n), nnl, noi)
    assert np.abs(curl0).max() == 0.0
    assert np.abs(grad0[inner]).max() < 1e-10
