"""Pins the oracle's bounding box, keys, sort and octree (CPU only).

Checks are independent of the oracle's code: closed forms (dyadic coordinates, whose
bisection is exact, give keys from their binary expansion), textbook routines (numpy's
stable argsort), brute-force structure (distinct key prefixes), and the Hilbert curve's
defining properties (bijection, face adjacency of consecutive cells).
"""
import numpy as np
import pytest

from miluphpc_amd import ics


def closed_form_lebesgue(ix, iy, iz, levels=21):
    """Key of dyadic points u = i / 2^21 in a unit box: the lower half sets the bit, so
    digit l = (1 - bit_l(ix)) + 2 (1 - bit_l(iy)) + 4 (1 - bit_l(iz))."""
    key = np.zeros(len(ix), dtype=np.uint64)
    for lv in range(levels):
        sh = levels - 1 - lv
        bx = (ix >> sh) & 1
        by = (iy >> sh) & 1
        bz = (iz >> sh) & 1
        d = (1 - bx) + 2 * (1 - by) + 4 * (1 - bz)
        key |= d.astype(np.uint64) << np.uint64(3 * (levels - 1 - lv))
    return key


def test_bbox_closed_form(orc):
    rng = np.random.default_rng(1)
    for shift in (0.0, 5.0, -5.0):
        x, y, z = (rng.normal(size=1000) + shift for _ in range(3))
        box = orc.bbox(x, y, z)
        lo = min(0.0, x.min(), y.min(), z.min())
        hi = max(0.0, x.max(), y.max(), z.max())
        assert np.array_equal(box, [lo, hi, lo, hi, lo, hi])
    assert np.array_equal(orc.bbox([], [], []), np.zeros(6))


def test_lebesgue_keys_dyadic_golden(orc):
    rng = np.random.default_rng(2)
    n = 20000
    ix, iy, iz = (rng.integers(0, 2 ** 21, n) for _ in range(3))
    # box [-1, 1]: u = (x + 1) / 2 with u = i / 2^21 -> x = 2 i / 2^21 - 1 (exact)
    x = 2.0 * ix / 2 ** 21 - 1.0
    y = 2.0 * iy / 2 ** 21 - 1.0
    z = 2.0 * iz / 2 ** 21 - 1.0
    box = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    k = orc.keys(x, y, z, box, 0)
    assert np.array_equal(k, closed_form_lebesgue(ix, iy, iz))


def test_lebesgue_keys_golden_fixture(orc):
    """Committed closed-form vectors (tests/golden/make_golden.py)."""
    from pathlib import Path
    g = np.load(Path(__file__).parent / "golden" / "keys_lebesgue_dyadic.npz")
    box = orc.bbox(g["x"], g["y"], g["z"])
    assert np.array_equal(box, [-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    assert np.array_equal(orc.keys(g["x"], g["y"], g["z"], box, 0), g["keys"])


def test_lebesgue_hand_points(orc):
    box = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])
    # the lowest corner is in the lower half on every axis at every level: all digits 7
    allsevens = sum(7 << (3 * l) for l in range(21))
    assert orc.lebesgue_key(-1.0, -1.0, -1.0, box) == allsevens
    # the upper corner is never below a midpoint: key 0
    assert orc.lebesgue_key(1.0, 1.0, 1.0, box) == 0
    # a point exactly on the first midpoint (0) goes to the upper half (x < mid is false)
    k = orc.lebesgue_key(0.0, -1.0, -1.0, box)
    assert (k >> 60) & 7 == 6
    # -0.0 compares equal to 0.0: same key
    assert orc.lebesgue_key(-0.0, -1.0, -1.0, box) == k


@pytest.mark.parametrize("L", [1, 2, 3, 4])
def test_hilbert_bijection_and_adjacency(orc, L):
    m = 2 ** L
    ii = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)])
    shift = 21 - L
    leb = closed_form_lebesgue(ii[:, 0] << shift, ii[:, 1] << shift, ii[:, 2] << shift)
    hil = np.array([orc.lebesgue2hilbert(int(v)) >> (3 * (21 - L)) for v in leb])
    assert sorted(hil.tolist()) == list(range(8 ** L))
    order = np.argsort(hil)
    cells = ii[order]
    steps = np.abs(np.diff(cells, axis=0)).sum(1)
    assert np.all(steps == 1), "consecutive Hilbert cells must share a face"


def test_hilbert_keys_consistent(orc):
    rng = np.random.default_rng(3)
    x, y, z = (rng.uniform(-1, 1, 5000) for _ in range(3))
    box = orc.bbox(x, y, z)
    kl = orc.keys(x, y, z, box, 0)
    kh = orc.keys(x, y, z, box, 1)
    assert np.array_equal(kh, np.array([orc.lebesgue2hilbert(int(v)) for v in kl], dtype=np.uint64))


def test_sort_is_stable_argsort(orc):
    rng = np.random.default_rng(4)
    k = rng.integers(0, 2 ** 63, 50000, dtype=np.uint64)
    k[::7] = k[3]  # many ties
    perm = orc.sort(k)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))


def prefix_structure(keys):
    """Brute force: cells = (level, prefix) shared by >= 2 particles; leaf level per particle."""
    cells = {}
    for L in range(1, 22):
        pre = keys >> np.uint64(3 * (21 - L))
        u, cnt = np.unique(pre, return_counts=True)
        for p in u[cnt >= 2]:
            cells[(L, int(p))] = None
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    x = sk[:-1] ^ sk[1:]
    lcp = np.array([(64 - int(v).bit_length() - 1) // 3 for v in x])
    leaf = np.empty(len(keys), np.int64)
    a = np.concatenate([[0], lcp])
    b = np.concatenate([lcp, [0]])
    leaf[order] = np.maximum(a, b) + 1
    return cells, leaf


def canonical_oracle_tree(t, keys):
    """(level, Lebesgue prefix) of every cell of the oracle's pointer tree, and leaf levels."""
    n = t.n
    child = t.child.reshape(-1, 8)
    cells = {}
    leaf = np.zeros(n, np.int64)
    stack = [(n, 0, 0)]
    while stack:
        node, L, prefix = stack.pop()
        if L > 0:
            cells[(L, prefix)] = node
        for s in range(8):
            c = child[node, s]
            if c == -1:
                continue
            if c < n:
                leaf[c] = L + 1
                assert (int(keys[c]) >> (3 * (21 - L - 1))) == (prefix << 3) | s
            else:
                stack.append((c, L + 1, (prefix << 3) | s))
    return cells, leaf


@pytest.mark.parametrize("case", ["plummer", "lattice"])
def test_tree_structure_is_key_prefixes(orc, case):
    if case == "plummer":
        s = ics.plummer(4000, seed=5)
    else:
        s = ics.sedov_lattice(12)
    box = orc.bbox(s["x"], s["y"], s["z"])
    keys = orc.keys(s["x"], s["y"], s["z"], box, 0)
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    assert t.status == 0
    cells, leaf = canonical_oracle_tree(t, keys)
    ref_cells, ref_leaf = prefix_structure(keys)
    assert set(cells) == set(ref_cells)
    assert t.ncells == len(cells) + 1  # + root
    assert np.array_equal(leaf, ref_leaf)
    assert np.array_equal(t.level[: t.n], ref_leaf)


def test_centres_of_mass_closed_form(orc):
    s = ics.plummer(3000, seed=6)
    rng = np.random.default_rng(6)
    m = rng.uniform(0.5, 1.5, len(s["x"]))
    box = orc.bbox(s["x"], s["y"], s["z"])
    keys = orc.keys(s["x"], s["y"], s["z"], box, 0)
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(m)
    cells, _ = canonical_oracle_tree(t, keys)
    for (L, pre), node in list(cells.items())[:400]:
        sel = (keys >> np.uint64(3 * (21 - L))) == pre
        M = m[sel].sum()
        assert abs(t.nm[node] - M) <= 1e-13 * M
        for arr, pos in ((t.nx, s["x"]), (t.ny, s["y"]), (t.nz, s["z"])):
            ref = (m[sel] * pos[sel]).sum() / M
            assert abs(arr[node] - ref) <= 1e-12 * max(1.0, abs(ref))
    root = t.n
    assert abs(t.nm[root] - m.sum()) <= 1e-12 * m.sum()


def test_quadrupole_closed_form(orc):
    s = ics.plummer(2000, seed=7)
    m = s["mass"]
    box = orc.bbox(s["x"], s["y"], s["z"])
    keys = orc.keys(s["x"], s["y"], s["z"], box, 0)
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(m)
    q = t.quadrupole().reshape(-1, 6)
    cells, _ = canonical_oracle_tree(t, keys)
    for (L, pre), node in list(cells.items())[:200]:
        sel = (keys >> np.uint64(3 * (21 - L))) == pre
        d = np.stack([s["x"][sel] - t.nx[node], s["y"][sel] - t.ny[node], s["z"][sel] - t.nz[node]], 1)
        mm = m[sel]
        r2 = (d * d).sum(1)
        ref = [(mm * (3 * d[:, 0] * d[:, 0] - r2)).sum(), (mm * 3 * d[:, 0] * d[:, 1]).sum(),
               (mm * 3 * d[:, 0] * d[:, 2]).sum(), (mm * (3 * d[:, 1] * d[:, 1] - r2)).sum(),
               (mm * 3 * d[:, 1] * d[:, 2]).sum(), (mm * (3 * d[:, 2] * d[:, 2] - r2)).sum()]
        scale = (mm * r2).sum() + 1e-300
        assert np.all(np.abs(q[node] - ref) <= 1e-10 * scale)
        assert abs(q[node][0] + q[node][3] + q[node][5]) <= 1e-10 * scale  # traceless


def test_tree_duplicate_positions_is_depth_error(orc):
    x = np.array([0.1, 0.1, -0.3])
    y = np.array([0.2, 0.2, 0.4])
    z = np.array([0.3, 0.3, 0.1])
    box = orc.bbox(x, y, z)
    t = orc.Tree(x, y, z, box)
    assert t.status == orc.E_DEPTH


def test_tree_capacity_error(orc):
    s = ics.plummer(1000, seed=8)
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box, cap_cells=10)
    assert t.status == orc.E_CAP_NODES
