"""Pins the oracle's Barnes-Hut (CPU only).

* numpy O(N^2) direct summation pins orc_direct;
* theta -> 0 opens every cell, so the group walk must equal direct summation;
* with groups of ONE particle the group acceptance test is the reference's per-particle
  v1 test (computeForces_v1, gravity.cu:252-396), so orc_bh_group(G=1) must equal the
  restatement orc_bh_v1 -- and v1 itself must converge to direct summation;
* the closed-form Plummer field M(<r)/r^2 pins the physics statistically;
* the quadrupole term must reduce the error against direct summation.
"""
import numpy as np
import pytest

from miluphpc_amd import ics
from conftest import vec_rel_err


def numpy_direct(x, y, z, m, eps2):
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dz = z[None, :] - z[:, None]
    r2 = dx * dx + dy * dy + dz * dz + eps2
    np.fill_diagonal(r2, np.inf)
    f = m[None, :] / r2 ** 1.5
    return (f * dx).sum(1), (f * dy).sum(1), (f * dz).sum(1)


def test_direct_matches_numpy(orc):
    s = ics.plummer(800, seed=11)
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], 1e-4)
    ax, ay, az = numpy_direct(s["x"], s["y"], s["z"], s["mass"], 1e-4)
    assert np.max(vec_rel_err(d["ax"], d["ay"], d["az"], ax, ay, az)) < 1e-12
    # Newton's third law: total momentum change vanishes
    tot = np.abs((s["mass"] * d["ax"]).sum()) + np.abs((s["mass"] * d["ay"]).sum())
    assert tot < 1e-12 * np.abs(s["mass"] * d["ax"]).sum()


@pytest.mark.parametrize("group", [1, 64])
def test_theta_to_zero_is_direct(orc, group):
    s = ics.plummer(1500, seed=12)
    eps2 = 1e-6
    g = orc.gravity_rhs(s, theta=1e-9, smoothing=eps2, group=group)
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], eps2)
    assert g["n_cell"] == 0
    assert np.max(vec_rel_err(g["ax"], g["ay"], g["az"], d["ax"], d["ay"], d["az"])) < 1e-11


def test_group_of_one_is_reference_v1(orc):
    s = ics.plummer(3000, seed=13)
    eps2, theta = 1e-6, 0.5
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    perm = orc.sort(orc.keys(s["x"], s["y"], s["z"], box, 1))
    radius = 0.5 * (box[1] - box[0])
    v1 = orc.bh_v1(t, radius, theta, eps2, potential=True)
    g1 = orc.bh_group(t, perm, box, theta, eps2, mac=0, group=1, potential=True)
    assert v1["status"] == 0 and g1["status"] == 0
    err = vec_rel_err(g1["ax"], g1["ay"], g1["az"], v1["ax"], v1["ay"], v1["az"])
    assert np.max(err) < 1e-12
    # v1's potential carries the 1/2 m_i factor (gravity.cu:360-366)
    assert np.allclose(0.5 * s["mass"] * g1["phi"], v1["u"], rtol=1e-12, atol=0)


def test_group_walk_accuracy_and_conservatism(orc):
    s = ics.plummer(6000, seed=14)
    eps2 = 1e-6
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], eps2)
    errs = {}
    for group in (1, 64):
        g = orc.gravity_rhs(s, theta=0.5, smoothing=eps2, group=group)
        errs[group] = vec_rel_err(g["ax"], g["ay"], g["az"], d["ax"], d["ay"], d["az"])
    # a group accepts a cell only if every member would: never coarser than G = 1
    assert np.percentile(errs[64], 99) <= np.percentile(errs[1], 99) * 1.0001
    assert np.percentile(errs[1], 99) < 2e-2


def test_quadrupole_improves_accuracy(orc):
    s = ics.plummer(6000, seed=15)
    eps2 = 4e-8
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], eps2)
    mono = orc.gravity_rhs(s, theta=0.6, smoothing=eps2, quad=False)
    quad = orc.gravity_rhs(s, theta=0.6, smoothing=eps2, quad=True)
    em = np.percentile(vec_rel_err(mono["ax"], mono["ay"], mono["az"], d["ax"], d["ay"], d["az"]), 99)
    eq = np.percentile(vec_rel_err(quad["ax"], quad["ay"], quad["az"], d["ax"], d["ay"], d["az"]), 99)
    assert eq < 0.5 * em
    assert eq < 5e-3


def test_plummer_closed_form_field(orc):
    s = ics.plummer(20000, seed=16)
    g = orc.gravity_rhs(s, theta=0.5, smoothing=1e-6, quad=True)
    r = np.sqrt(s["x"] ** 2 + s["y"] ** 2 + s["z"] ** 2)
    ar = (g["ax"] * s["x"] + g["ay"] * s["y"] + g["az"] * s["z"]) / r
    for lo, hi in ((0.5, 0.8), (0.8, 1.2), (1.2, 2.0), (2.0, 4.0)):
        sel = (r > lo) & (r < hi)
        rm = r[sel]
        model = -(rm ** 3 / (rm ** 2 + 1.0) ** 1.5) / rm ** 2
        assert abs(np.mean(ar[sel]) / np.mean(model) - 1.0) < 0.03


def test_sqrt_theta_mac_is_coarser_than_classic_at_same_theta(orc):
    # the v2 form s^2/theta <= d^2 accepts when s/d <= sqrt(theta), which for theta < 1 is
    # looser than s/d < theta: fewer interactions, larger error; and the v2 form at
    # theta^2 is the classic form at theta (up to the strict/non-strict boundary)
    s = ics.plummer(4000, seed=17)
    a = orc.gravity_rhs(s, theta=0.5, smoothing=1e-6, mac=0)
    b = orc.gravity_rhs(s, theta=0.5, smoothing=1e-6, mac=1)
    c = orc.gravity_rhs(s, theta=0.25, smoothing=1e-6, mac=1)
    assert b["n_part"] + b["n_cell"] < a["n_part"] + a["n_cell"]
    assert (c["n_part"], c["n_cell"]) == (a["n_part"], a["n_cell"])


def _vote_cut_counts(t, perm, box, group, theta, eps2):
    """Independent restatement of the vote form's interaction counts: a recursive descent
    from the root in Python over the oracle's tree arrays; a cell is taken for a group iff
    every member's own classic test (d_i^2 + eps^2) theta^2 > s^2 passes."""
    n = t.n
    R = 0.5 * (box[1] - box[0])
    child = t.child.reshape(-1, 8)
    cells = parts = 0
    for g0 in range(0, n, group):
        mem = perm[g0:g0 + group]
        px, py, pz = t.nx[mem], t.ny[mem], t.nz[mem]
        stack = [n]
        while stack:
            c = stack.pop()
            if c < n:
                parts += len(mem)
                continue
            d2 = (t.nx[c] - px) ** 2 + (t.ny[c] - py) ** 2 + (t.nz[c] - pz) ** 2
            s2 = 4.0 * R * R * 2.0 ** (-2 * int(t.level[c]))
            if c != n and np.all((d2 + eps2) * theta ** 2 > s2):
                cells += len(mem)
            else:
                stack.extend(int(ch) for ch in child[c] if ch != -1)
    return cells, parts


def test_vote_form_counts_match_an_independent_descent(orc):
    s = ics.plummer(3000, seed=21)
    eps2, theta = 1e-6, 0.5
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    perm = orc.sort(orc.keys(s["x"], s["y"], s["z"], box, 1))
    for group in (64, 7):
        g = orc.bh_group(t, perm, box, theta, eps2, mac=0, group=group, vote=True)
        assert (g["n_cell"], g["n_part"]) == _vote_cut_counts(t, perm, box, group, theta, eps2)


def test_vote_form_sits_between_per_particle_and_box_forms(orc):
    """Per member, the vote form's cut refines the member's own (G = 1) cut and is refined by
    the box form's: interaction counts are ordered G=1 <= vote <= box, the vote form with
    groups of one is the v1 walk, and theta -> 0 is direct summation."""
    s = ics.plummer(6000, seed=22)
    eps2 = 1e-6
    box = orc.bbox(s["x"], s["y"], s["z"])
    t = orc.Tree(s["x"], s["y"], s["z"], box)
    t.com(s["mass"])
    perm = orc.sort(orc.keys(s["x"], s["y"], s["z"], box, 1))
    for mac in (0, 1):
        one = orc.bh_group(t, perm, box, 0.5, eps2, mac=mac, group=1, vote=False)
        one_v = orc.bh_group(t, perm, box, 0.5, eps2, mac=mac, group=1, vote=True)
        vote = orc.bh_group(t, perm, box, 0.5, eps2, mac=mac, group=64, vote=True)
        boxf = orc.bh_group(t, perm, box, 0.5, eps2, mac=mac, group=64, vote=False)
        assert (one_v["n_cell"], one_v["n_part"]) == (one["n_cell"], one["n_part"])
        assert np.array_equal(one_v["ax"], one["ax"]) and np.array_equal(one_v["az"], one["az"])
        tot = lambda r: r["n_cell"] + r["n_part"]
        assert tot(one) < tot(vote) < tot(boxf)
        assert vote["n_part"] <= boxf["n_part"]
    d = orc.direct(s["x"], s["y"], s["z"], s["mass"], eps2)
    z = orc.gravity_rhs(s, theta=1e-9, smoothing=eps2, group=64, vote=True)
    assert z["n_cell"] == 0
    assert np.max(vec_rel_err(z["ax"], z["ay"], z["az"], d["ax"], d["ay"], d["az"])) < 1e-11
    e1 = vec_rel_err(one["ax"], one["ay"], one["az"], d["ax"], d["ay"], d["az"])
    ev = vec_rel_err(vote["ax"], vote["ay"], vote["az"], d["ax"], d["ay"], d["az"])
    assert np.percentile(ev, 99) <= np.percentile(e1, 99) * 1.0001
