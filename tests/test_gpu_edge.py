"""Edge sizes through the C-ABI: an empty context, Barnes-Hut on a handful of particles and
around the 64-particle group size, and the SPH RHS on small random clouds, against the oracle."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def blob(n, seed):
    rng = np.random.default_rng(seed)
    s = {k: rng.normal(size=n) for k in ("x", "y", "z")}
    s["mass"] = rng.uniform(0.5, 1.5, n) / n
    for k in ("vx", "vy", "vz"):
        s[k] = 0.1 * rng.normal(size=n)
    s["e"] = rng.uniform(0.5, 1.0, n)
    s["p"] = rng.uniform(0.1, 0.2, n)
    s["cs"] = np.zeros(n)
    return s


def make_ctx(lib, s, params):
    ctx = lib.Context(max(len(s["x"]), 1), params)
    ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
    return ctx


def test_empty_context_rhs(gpu_lib):
    """n = 0: every RHS form returns MHPC_OK with zero counts."""
    print_case("empty context rhs")
    with gpu_lib.Context(1024, gpu_lib.default_params()) as ctx:
        ctx.set_num_particles(0)
        for flags in (gpu_lib.RHS_GRAVITY, gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML,
                      gpu_lib.RHS_GRAVITY | gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML | gpu_lib.RHS_SPECULATIVE_NN):
            st = ctx.rhs(flags)
            assert st["n_neighbours"] == 0
            assert st["n_part_interactions"] == 0 and st["n_cell_interactions"] == 0


@pytest.mark.parametrize("n", [2, 3, 17, 63, 64, 65, 129])
def test_barnes_hut_small_n(gpu_lib, orc, n):
    """Tiny and ragged inputs (one partial group, exactly one group, a group plus one): FP64
    accelerations and potentials equal the oracle's walk of the same rule to 1e-12 and, with
    the interaction counts, match; for n <= 3 every particle meets every particle (itself too,
    with zero force), so the result also equals direct summation."""
    print_case(f"barnes-hut fp64 n={n}")
    s = blob(n, n)
    p = gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, mac=0, precision=0)
    with make_ctx(gpu_lib, s, p) as ctx:
        st = ctx.rhs(gpu_lib.RHS_GRAVITY | gpu_lib.RHS_POTENTIAL)
        g = {k: ctx.download(k) for k in ("g_ax", "g_ay", "g_az", "u", "sorted")}
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    o = orc.bh_group(t, g["sorted"], box, 0.5, 1e-6, mac=0, group=gpu_lib.BH_GROUP, potential=True, vote=True)
    assert st["n_cell_interactions"] == o["n_cell"] and st["n_part_interactions"] == o["n_part"]
    mag = np.sqrt(o["ax"] ** 2 + o["ay"] ** 2 + o["az"] ** 2)
    for k, ok in (("g_ax", "ax"), ("g_ay", "ay"), ("g_az", "az")):
        assert np.max(np.abs(g[k] - o[ok]) / mag) < 1e-12, k
    assert np.max(np.abs(g["u"] - 0.5 * s["mass"] * o["phi"]) / np.abs(0.5 * s["mass"] * o["phi"])) < 1e-12
    if n <= 3:
        d = orc.direct(s["x"], s["y"], s["z"], s["mass"], 1e-6)
        assert o["n_part"] == n * n and o["n_cell"] == 0
        assert np.allclose(g["g_ax"], d["ax"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("n", [100, 777])
def test_sph_small_cloud(gpu_lib, orc, n):
    """A small random cloud: fewer particles than some boxes' searches expect, h growing for
    the outer ones until the pass limit.  h, iteration and neighbour counts bit-exact; density,
    pressure and derivatives to the FP64 tolerances."""
    print_case(f"sph small cloud n={n}")
    s = blob(n, 1000 + n)
    r = np.sqrt(s["x"] ** 2 + s["y"] ** 2 + s["z"] ** 2)
    s["sml"] = 0.5 + 0.3 * r
    p = gpu_lib.default_params(curve=1, precision=0)
    flags = gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML | gpu_lib.RHS_SPECULATIVE_NN
    with make_ctx(gpu_lib, s, p) as ctx:
        ctx.rhs(flags)
        g = {k: ctx.download(k) for k in ("sml", "h_iters", "noi", "rho", "p", "ax", "ay", "az", "dedt")}
    o = orc.sph_rhs(s)
    assert o["h_status"] == 0 and o["nn_status"] == 0
    assert np.array_equal(g["sml"], o["sml"])
    assert np.array_equal(g["h_iters"], o["h_iters"])
    assert np.array_equal(g["noi"], o["noi"])
    assert np.max(np.abs(g["rho"] - o["rho"]) / o["rho"]) < 1e-12
    for k in ("ax", "ay", "az", "dedt"):
        scale = np.sqrt(np.mean(o[k] ** 2))
        assert np.max(np.abs(g[k] - o[k]) / np.maximum(np.abs(o[k]), scale)) < 1e-11, k
