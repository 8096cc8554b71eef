"""GPU parity of bbox, keys, sort, octree, centres of mass and Barnes-Hut against the
oracle, through the C-ABI.  Bit-exact for keys, order, tree structure, COM (FP64 in the
same order) and the walk's interaction counts; FP64 forces to 1e-12, FP32+Kahan forces to
the stated per-particle tolerance."""
import numpy as np
import pytest

from miluphpc_amd import ics
from conftest import print_case, vec_rel_err

pytestmark = pytest.mark.gpu


def make_ctx(lib, s, params, cap=None, cap_cells=0):
    n = len(s["x"])
    ctx = lib.Context(cap or n, params, cap_cells=cap_cells)
    ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
    return ctx


def stage_to_tree(ctx):
    box = ctx.bbox()
    ctx.keys()
    ctx.sort()
    ctx.build_tree()
    return box


@pytest.fixture(scope="module")
def plummer64k():
    return ics.plummer(65536, seed=ics.SEED)


@pytest.mark.parametrize("curve", [0, 1])
def test_keys_and_sort_bitexact(gpu_lib, orc, plummer64k, curve):
    print_case(f"keys+sort plummer 65536 curve={curve}")
    s = plummer64k
    p = gpu_lib.default_params(curve=curve)
    with make_ctx(gpu_lib, s, p) as ctx:
        box = ctx.bbox()
        assert np.array_equal(box, orc.bbox(s["x"], s["y"], s["z"]))
        ctx.keys()
        ctx.sort()
        k = ctx.download("key")
        perm = ctx.download("sorted")
    ko = orc.keys(s["x"], s["y"], s["z"], box, curve)
    assert np.array_equal(k, ko)
    assert np.array_equal(perm, orc.sort(ko))


@pytest.mark.parametrize("curve", [0, 1])
def test_sort_depth_hint_and_fallback(gpu_lib, orc, curve):
    """After a tree build the context sorts only the key levels the previous tree needed (plus
    one), checks that no adjacent keys are left equal in them, and falls back to the full sort
    when some are.  The order must be the full stable sort's in all three calls: first (no
    hint), same positions again (partial passes), two particles on one position (fallback)."""
    print_case(f"sort depth hint and fallback, lattice 16^3, curve {curve}")
    s = ics.sedov_lattice(16)
    p = gpu_lib.default_params(curve=curve)
    with make_ctx(gpu_lib, s, p) as ctx:
        for step in range(3):
            if step == 2:
                s = dict(s)
                for c in ("x", "y", "z"):
                    s[c] = s[c].copy()
                    s[c][1] = s[c][0]
                ctx.load(x=s["x"], y=s["y"], z=s["z"])
            box = ctx.bbox()
            ctx.keys()
            ctx.sort()
            k = ctx.download("key")
            perm = ctx.download("sorted")
            if step < 2:
                ctx.build_tree()
                ctx.centres_of_mass()   # the tree depth the next sort uses
            ko = orc.keys(s["x"], s["y"], s["z"], box, curve)
            assert np.array_equal(k, ko), step
            assert np.array_equal(perm, orc.sort(ko)), step
        assert k[0] == k[1]


def test_keys_golden_fixture(gpu_lib):
    """Lebesgue keys of dyadic points against the committed closed-form vectors."""
    from pathlib import Path
    print_case("keys golden fixture (closed form, 4100 dyadic points)")
    g = np.load(Path(__file__).parent / "golden" / "keys_lebesgue_dyadic.npz")
    n = len(g["keys"])
    s = {"x": g["x"].copy(), "y": g["y"].copy(), "z": g["z"].copy(), "mass": np.full(n, 1.0 / n)}
    with make_ctx(gpu_lib, s, gpu_lib.default_params(curve=0)) as ctx:
        box = ctx.bbox()
        assert np.array_equal(box, [-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
        ctx.keys()
        k = ctx.download("key")
    assert np.array_equal(k, g["keys"])


def test_sort_ties_and_ragged_sizes(gpu_lib, orc):
    """Duplicated keys keep index order; sizes that are not multiples of the tile."""
    rng = np.random.default_rng(31)
    for n in (1, 2, 63, 4097, 70001):
        print_case(f"sort ties n={n}")
        # coarse lattice of 8^3 points repeated: many equal keys (the sort alone is fine
        # with duplicates; the tree would reject them)
        g = rng.integers(0, 8, (n, 3)).astype(np.float64) / 8.0 - 0.4
        s = {"x": g[:, 0].copy(), "y": g[:, 1].copy(), "z": g[:, 2].copy(), "mass": np.full(n, 1.0 / n)}
        with make_ctx(gpu_lib, s, gpu_lib.default_params()) as ctx:
            box = ctx.bbox()
            ctx.keys()
            ctx.sort()
            k = ctx.download("key")
            perm = ctx.download("sorted")
        ko = orc.keys(s["x"], s["y"], s["z"], box, 0)
        assert np.array_equal(k, ko)
        assert np.array_equal(perm, np.argsort(ko, kind="stable"))


def test_sort_many_tiles(gpu_lib, orc):
    """Hundreds of 4096-key tiles chained by the look-back: stable order against numpy."""
    rng = np.random.default_rng(32)
    n = 1_500_007
    print_case(f"sort many tiles n={n}")
    g = rng.random((n, 3))
    g[::7] = np.round(g[::7] * 16) / 16      # some duplicated keys across tiles
    s = {"x": g[:, 0].copy(), "y": g[:, 1].copy(), "z": g[:, 2].copy(), "mass": np.full(n, 1.0 / n)}
    with make_ctx(gpu_lib, s, gpu_lib.default_params()) as ctx:
        box = ctx.bbox()
        ctx.keys()
        for _ in range(2):                    # a second sort reuses the look-back words
            ctx.sort()
        k = ctx.download("key")
        perm = ctx.download("sorted")
    ko = orc.keys(s["x"], s["y"], s["z"], box, 0)
    assert np.array_equal(k, ko)
    assert np.array_equal(perm, np.argsort(ko, kind="stable"))


def canonical_gpu(tree, n, perm, lkeys):
    """(level, Lebesgue prefix) -> (cell index) and canonical children of every GPU cell."""
    lev, start = tree["level"], tree["start"]
    slk = lkeys[perm]
    canon = {}
    for c in range(1, len(lev)):
        L = int(lev[c])
        canon[(L, int(slk[start[c]]) >> (3 * (21 - L)))] = c
    return canon


def canonical_oracle(t, lkeys):
    n = t.n
    child = t.child.reshape(-1, 8)
    cells = {}
    stack = [(n, 0, 0)]
    while stack:
        node, L, prefix = stack.pop()
        if L > 0:
            cells[(L, prefix)] = node
        for s_ in range(8):
            c = child[node, s_]
            if c >= n:
                stack.append((c, L + 1, (prefix << 3) | s_))
    return cells


@pytest.mark.parametrize("case", ["plummer", "sedov"])
def test_tree_and_com_bitexact(gpu_lib, orc, plummer64k, case):
    print_case(f"tree+com {case}")
    if case == "plummer":
        check_tree_against_oracle(gpu_lib, orc, plummer64k, 1)
    else:
        check_tree_against_oracle(gpu_lib, orc, ics.sedov_lattice(32), 0)


def cell_at_level(x, y, z, box, levels):
    """The box of the cell containing (x, y, z) after `levels` bisections of the root box (the
    keys' expressions, in FP64)."""
    mn, mx = [box[0], box[2], box[4]], [box[1], box[3], box[5]]
    for _ in range(levels):
        for a, v in enumerate((x, y, z)):
            m = 0.5 * (mn[a] + mx[a])
            if v < m:
                mx[a] = m
            else:
                mn[a] = m
    return mn, mx


def test_tree_pair_sharing_a_key(gpu_lib, orc, plummer64k):
    """Two particles with equal 21-digit keys that separate at the next bisection: the
    reference's insertion puts them below a level-21 cell (leaf level 22); tree, centres of
    mass and the Barnes-Hut walk match the oracle."""
    print_case("tree+com+bh plummer 65536 with a pair sharing a 21-digit key")
    s = {k: v.copy() for k, v in plummer64k.items()}
    r2 = s["x"] ** 2 + s["y"] ** 2 + s["z"] ** 2
    a, b = (int(i) for i in np.argsort(r2)[:2])   # two core particles (neither sets the box)
    box = orc.bbox(s["x"], s["y"], s["z"])
    mn, mx = cell_at_level(s["x"][b], s["y"][b], s["z"][b], box, 21)
    # particle a: the sub-octant of b's level-21 cell opposite to b's on every axis
    for ax, k in enumerate(("x", "y", "z")):
        mid = 0.5 * (mn[ax] + mx[ax])
        q = 0.25 * (mx[ax] - mn[ax])
        s[k][a] = mid + q if s[k][b] < mid else mid - q
    box = orc.bbox(s["x"], s["y"], s["z"])
    lk = orc.keys(s["x"], s["y"], s["z"], box, 0)
    assert lk[a] == lk[b]
    check_tree_against_oracle(gpu_lib, orc, s, 1, expect_level22=(a, b))
    p = gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, mac=0, precision=0)
    g = run_gravity(gpu_lib, s, p, potential=True)
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    o = orc.bh_group(t, g["sorted"], box, 0.5, 1e-6, mac=0, group=gpu_lib.BH_GROUP, potential=True,
                     vote=p.bh_accept == gpu_lib.BH_VOTE)
    assert g["stats"]["n_cell_interactions"] == o["n_cell"]
    assert g["stats"]["n_part_interactions"] == o["n_part"]
    assert np.max(vec_rel_err(g["g_ax"], g["g_ay"], g["g_az"], o["ax"], o["ay"], o["az"])) < 1e-12


def check_tree_against_oracle(gpu_lib, orc, s, curve, expect_level22=()):
    n = len(s["x"])
    p = gpu_lib.default_params(curve=curve, quadrupole=1)
    with make_ctx(gpu_lib, s, p) as ctx:
        box = stage_to_tree(ctx)
        ctx.centres_of_mass()
        tree = ctx.tree()
        perm = ctx.download("sorted")
        plevel = ctx.download("level")
    lk = orc.keys(s["x"], s["y"], s["z"], box, 0)
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    assert t.status == 0
    t.com(s["mass"])
    q = t.quadrupole().reshape(-1, 6)
    og = canonical_oracle(t, lk)
    gg = canonical_gpu(tree, n, perm, lk)
    assert set(og) == set(gg)
    assert len(tree["level"]) == t.ncells
    assert np.array_equal(plevel, t.level[:n])
    for i in expect_level22:
        assert plevel[i] == 22
    # children, canonically: particle children by original index, cells by (L, prefix)
    ochild = t.child.reshape(-1, 8)
    inv = {v: k for k, v in gg.items()}
    for key, oc in og.items():
        gc = gg[key]
        for s_ in range(8):
            a, b = ochild[oc, s_], tree["child"][gc, s_]
            if a == -1:
                assert b == -1
            elif a < n:
                assert b < n and perm[b] == a
            else:
                assert b >= n and inv[b - n] == (key[0] + 1, (key[1] << 3) | s_)
        assert tree["mass"][gc] == t.nm[oc]
        assert tree["x"][gc] == t.nx[oc] and tree["y"][gc] == t.ny[oc] and tree["z"][gc] == t.nz[oc]
        assert np.array_equal(tree["quad"][:, gc], q[oc])
    assert tree["mass"][0] == t.nm[n]


def run_gravity(gpu_lib, s, params, potential=False):
    with make_ctx(gpu_lib, s, params) as ctx:
        st = ctx.rhs(gpu_lib.RHS_GRAVITY | (gpu_lib.RHS_POTENTIAL if potential else 0))
        out = {k: ctx.download(k) for k in ("g_ax", "g_ay", "g_az", "u", "sorted")}
        out["box"] = ctx.get_box()
    out["stats"] = st
    return out


@pytest.mark.parametrize("mac,vote", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_barnes_hut_fp64_parity(gpu_lib, orc, plummer64k, mac, vote):
    print_case(f"bh fp64 plummer 65536 mac={mac} vote={vote}")
    s = plummer64k
    p = gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, mac=mac, precision=0, bh_accept=vote)
    g = run_gravity(gpu_lib, s, p, potential=True)
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    o = orc.bh_group(t, g["sorted"], box, 0.5, 1e-6, mac=mac, group=gpu_lib.BH_GROUP, potential=True,
                     vote=bool(vote))
    assert g["stats"]["n_cell_interactions"] == o["n_cell"]
    assert g["stats"]["n_part_interactions"] == o["n_part"]
    err = vec_rel_err(g["g_ax"], g["g_ay"], g["g_az"], o["ax"], o["ay"], o["az"])
    assert np.max(err) < 1e-12
    u_ref = 0.5 * s["mass"] * o["phi"]
    assert np.max(np.abs(g["u"] - u_ref) / np.abs(u_ref)) < 1e-12


def test_barnes_hut_fp32_quadrupole_parity(gpu_lib, orc, plummer64k):
    """Mixed precision: FP32 pair math + Kahan vs the FP64 oracle of the same walk."""
    print_case("bh fp32+kahan quadrupole plummer 65536")
    s = plummer64k
    p = gpu_lib.default_params(theta=0.6, smoothing=4e-8, curve=1, mac=0, quadrupole=1, precision=1)
    g = run_gravity(gpu_lib, s, p, potential=True)
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    o = orc.bh_group(t, g["sorted"], box, 0.6, 4e-8, mac=0, group=gpu_lib.BH_GROUP, quad=True, potential=True,
                     vote=p.bh_accept == gpu_lib.BH_VOTE)
    assert g["stats"]["n_cell_interactions"] == o["n_cell"]
    assert g["stats"]["n_part_interactions"] == o["n_part"]
    err = vec_rel_err(g["g_ax"], g["g_ay"], g["g_az"], o["ax"], o["ay"], o["az"])
    print(f"fp32 quad: rel err p50 {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} max {err.max():.2e}")
    assert np.percentile(err, 99) < 1e-5
    # the BH approximation error itself is far larger than the FP32 error
    sub = np.random.default_rng(5).choice(len(s["x"]), 512, replace=False).astype(np.int32)
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], 4e-8, targets=sub)
    e_bh = vec_rel_err(o["ax"][sub], o["ay"][sub], o["az"][sub], d["ax"], d["ay"], d["az"])
    assert np.percentile(err, 99) < 0.1 * np.percentile(e_bh, 99)
    u_ref = 0.5 * s["mass"] * o["phi"]
    assert np.percentile(np.abs(g["u"] - u_ref) / np.abs(u_ref), 99) < 1e-5


def test_tree_errors(gpu_lib):
    """Equal positions never separate (the reference asserts at depth MAX_LEVEL + 2,
    tree.cu:949-963); too small a cell capacity."""
    print_case("tree errors: duplicate positions, node capacity")
    s = ics.plummer(5000, seed=9)
    s["x"][10], s["y"][10], s["z"][10] = s["x"][20], s["y"][20], s["z"][20]
    with make_ctx(gpu_lib, s, gpu_lib.default_params()) as ctx:
        with pytest.raises(gpu_lib.MhpcError) as e:
            ctx.rhs(gpu_lib.RHS_GRAVITY)
        assert e.value.status == gpu_lib.E_DEPTH
    s = ics.plummer(5000, seed=9)
    with make_ctx(gpu_lib, s, gpu_lib.default_params(), cap_cells=100) as ctx:
        with pytest.raises(gpu_lib.MhpcError) as e:
            ctx.rhs(gpu_lib.RHS_GRAVITY)
        assert e.value.status == gpu_lib.E_CAP_NODES


def test_leapfrog_matches_oracle(gpu_lib, orc):
    print_case("leapfrog 3 steps plummer 8192")
    s = ics.plummer(8192, seed=10)
    p = gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, precision=0)
    dt = 1e-3
    with make_ctx(gpu_lib, s, p) as ctx:
        ctx.rhs(gpu_lib.RHS_GRAVITY)
        for _ in range(3):
            ctx.leapfrog_update_x(dt)
            ctx.rhs(gpu_lib.RHS_GRAVITY)
            ctx.leapfrog_update_v(dt)
        g = {k: ctx.download(k) for k in ("x", "y", "z", "vx", "vy", "vz", "g_ax", "g_ay", "g_az")}
    st = {k: s[k].copy() for k in ("x", "y", "z", "vx", "vy", "vz")}

    def acc(st):
        r = orc.gravity_rhs({**st, "mass": s["mass"]}, 0.5, 1e-6, curve=1, group=gpu_lib.BH_GROUP)
        return r["ax"], r["ay"], r["az"]

    st["g_ax"], st["g_ay"], st["g_az"] = acc(st)
    for k in ("g_ax_old", "g_ay_old", "g_az_old"):
        st[k] = np.zeros(len(s["x"]))
    for _ in range(3):
        orc.leapfrog_update_x(dt, st)
        st["g_ax"], st["g_ay"], st["g_az"] = acc(st)
        orc.leapfrog_update_v(dt, st)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.max(np.abs(g[k] - st[k])) < 1e-12 * max(1.0, np.abs(st[k]).max())


def test_barnes_hut_per_particle_v1_is_the_reference_walk(gpu_lib, orc):
    """bh_accept = PER_PARTICLE runs computeForces_v1 itself (one depth-first walk per
    particle, classic MAC, the particle itself skipped): accelerations and potential equal the
    oracle's restatement orc_bh_v1, which walks in the same order with the same expressions."""
    print_case("bh per-particle v1, plummer 20000, fp64")
    s = ics.plummer(20000, seed=23)
    p = gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, mac=0, precision=0,
                               bh_accept=gpu_lib.BH_PER_PARTICLE)
    g = run_gravity(gpu_lib, s, p, potential=True)
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    o = orc.bh_v1(t, 0.5 * (box[1] - box[0]), 0.5, 1e-6, potential=True)
    assert o["status"] == 0
    err = vec_rel_err(g["g_ax"], g["g_ay"], g["g_az"], o["ax"], o["ay"], o["az"])
    exact = np.mean((g["g_ax"] == o["ax"]) & (g["g_ay"] == o["ay"]) & (g["g_az"] == o["az"]))
    print(f"v1: max rel err {err.max():.2e}, bit-identical fraction {exact:.4f}")
    assert err.max() < 1e-14
    assert np.max(np.abs(g["u"] - o["u"]) / np.abs(o["u"])) < 1e-14
    # v1 is the G = 1 limit of the group walks: it needs the fewest interactions
    vote = run_gravity(gpu_lib, s, gpu_lib.default_params(theta=0.5, smoothing=1e-6, curve=1, mac=0, precision=0))
    tot = lambda st: st["n_cell_interactions"] + st["n_part_interactions"]
    assert tot(g["stats"]) < tot(vote["stats"])


def test_per_particle_walk_rejects_other_forms(gpu_lib):
    for kw in ({"mac": 1}, {"quadrupole": 1}):
        with pytest.raises(gpu_lib.MhpcError) as e:
            gpu_lib.Context(1000, gpu_lib.default_params(bh_accept=gpu_lib.BH_PER_PARTICLE, **kw))
        assert e.value.status == gpu_lib.E_ARG
