"""The C++ drop-in facade (Miluphpc / ParticleHandler / Leapfrog, miluphpc_amd/host) and the
global sums, on the GPU.  The driver binary is used the way reference code uses the
classes; its inputs and outputs are checked against the oracle."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from miluphpc_amd import ics
from conftest import print_case, vec_rel_err

pytestmark = pytest.mark.gpu

REPO = Path(__file__).resolve().parents[1]
DRIVER = REPO / "miluphpc_amd" / "bin" / "mhpc_rhs_driver"


def read_fields(path, names):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:8].view(np.int64)[0])
    data = raw[8:].view(np.float64).reshape(len(names), n)
    return {k: data[i].copy() for i, k in enumerate(names)}


def test_facade_leapfrog_matches_oracle(gpu_lib, orc, tmp_path):
    print_case("c++ facade: Leapfrog 2 steps, 6000 particles")
    out = tmp_path / "grav.bin"
    res = subprocess.run([str(DRIVER), "gravity", "6000", "2", str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    names = res.stdout.strip().splitlines()[-1].split(": ")[1].split()
    f = read_fields(out, names)
    s = {"x": f["x0"], "y": f["y0"], "z": f["z0"], "vx": f["vx0"], "vy": f["vy0"], "vz": f["vz0"]}
    m = f["m"]

    def acc(st):
        r = orc.gravity_rhs({**st, "mass": m}, 0.5, 1e-6, curve=1, mac=0, group=gpu_lib.BH_GROUP)
        return r["ax"], r["ay"], r["az"]

    st = {k: v.copy() for k, v in s.items()}
    st["g_ax"], st["g_ay"], st["g_az"] = acc(st)
    for k in ("g_ax_old", "g_ay_old", "g_az_old"):
        st[k] = np.zeros(len(m))
    for _ in range(2):
        orc.leapfrog_update_x(1e-3, st)
        st["g_ax"], st["g_ay"], st["g_az"] = acc(st)
        orc.leapfrog_update_v(1e-3, st)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.max(np.abs(f[k] - st[k])) < 1e-12, k
    assert np.max(vec_rel_err(f["g_ax"], f["g_ay"], f["g_az"], st["g_ax"], st["g_ay"], st["g_az"])) < 1e-12
    # afterIntegrationStep() printed energy and |L|: check |L| against the final state
    line = [ln for ln in res.stdout.splitlines() if ln.startswith("energy")][0].split()
    L = np.array([(m * (f["y"] * f["vz"] - f["z"] * f["vy"])).sum(), (m * (f["z"] * f["vx"] - f["x"] * f["vz"])).sum(),
                  (m * (f["x"] * f["vy"] - f["y"] * f["vx"])).sum()])
    assert float(line[3]) == pytest.approx(np.linalg.norm(L), rel=1e-13)


def test_facade_sph_rhs_matches_oracle(orc, tmp_path):
    print_case("c++ facade: SPH rhs, 20^3 jittered lattice")
    out = tmp_path / "sph.bin"
    res = subprocess.run([str(DRIVER), "sph", "20", str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    names = res.stdout.strip().splitlines()[-1].split(": ")[1].split()
    f = read_fields(out, names)
    s = {"x": f["x"], "y": f["y"], "z": f["z"], "vx": f["vx"], "vy": f["vy"], "vz": f["vz"], "mass": f["m"],
         "sml": f["sml0"], "e": f["e"], "p": f["p0"], "cs": np.zeros(len(f["m"]))}
    o = orc.sph_rhs(s)
    assert np.array_equal(f["sml"], o["sml"])
    assert np.max(np.abs(f["rho"] - o["rho"]) / o["rho"]) < 1e-12
    assert np.max(np.abs(f["p"] - o["p"]) / o["p"]) < 1e-12
    scale = np.sqrt(np.mean(o["ax"] ** 2))
    assert np.max(np.abs(f["ax"] - o["ax"]) / np.maximum(np.abs(o["ax"]), scale)) < 1e-10


def test_diagnostics_are_exactly_rounded_sums(gpu_lib):
    print_case("diagnostics plummer 65536 with potential")
    s = ics.plummer(65536)
    p = ics.params_plummer_small()
    with gpu_lib.Context(len(s["x"]), p) as ctx:
        ctx.load(**s, e=np.linspace(0.1, 2.0, len(s["x"])))
        ctx.rhs(gpu_lib.RHS_GRAVITY | gpu_lib.RHS_POTENTIAL)
        d = ctx.diagnostics()
        u = ctx.download("u")
    m = s["mass"]
    e = np.linspace(0.1, 2.0, len(m))
    v2 = s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2
    ref = {
        "kinetic": math.fsum(0.5 * m * v2),
        "potential": math.fsum(u),
        "internal": math.fsum(m * e),
        "mass": math.fsum(m),
    }
    for k, v in ref.items():
        assert d[k] == pytest.approx(v, rel=1e-14, abs=0), k
    Lz = math.fsum(m * (s["x"] * s["vy"] - s["y"] * s["vx"]))
    assert d["angular_momentum"][2] == pytest.approx(Lz, rel=1e-12, abs=1e-17)
    # Plummer virial equilibrium: 2T + W ~ 0
    assert abs(2 * d["kinetic"] + d["potential"]) < 0.05 * abs(d["potential"])


def _write_ic(d, s, material_id=None, with_sml=True, with_rho=True):
    d.mkdir(parents=True, exist_ok=True)
    n = len(s["x"])
    np.save(d / "m.npy", s["mass"])
    np.save(d / "x.npy", np.stack([s["x"], s["y"], s["z"]], axis=1))
    np.save(d / "v.npy", np.stack([s["vx"], s["vy"], s["vz"]], axis=1))
    np.save(d / "materialId.npy", np.zeros(n, np.int8) if material_id is None else material_id)
    np.save(d / "u.npy", s["e"])
    if with_rho:
        np.save(d / "rho.npy", s.get("rho", np.ones(n)))
    if with_sml:
        np.save(d / "sml.npy", s["sml"])


def _read_snapshot(d):
    return {p.stem: np.load(p, allow_pickle=False) for p in d.glob("*.npy")}


def test_facade_io_round_trip_sph(orc, tmp_path):
    """Initial distribution in the reference's datasets (/m, /x N x 3, /v N x 3, /materialId,
    /u, /rho, /sml as .npy) -> PredictorCorrectorEuler(simulationParameters) -> rhs() ->
    particles2file(0): the snapshot's particle data round-trips exactly and its SPH fields are
    the oracle's for that input; one integrate() later ts000001 holds the advanced state."""
    print_case("c++ facade: .npy initial distribution -> rhs -> ts000000 / ts000001 snapshots")
    s = ics.sedov_lattice(20)
    rng = np.random.default_rng(8)
    n = len(s["x"])
    for k in ("x", "y", "z"):
        s[k] = s[k] + 0.2 / 20 * rng.uniform(-1, 1, n)
    for k in ("vx", "vy", "vz"):
        s[k] = 0.1 * rng.normal(size=n)
    ic, out = tmp_path / "ic", tmp_path / "out"
    _write_ic(ic, s)
    out.mkdir()
    res = subprocess.run([str(DRIVER), "io", "sph", str(ic), str(out), "1"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    t0 = _read_snapshot(out / "ts000000")
    assert set(t0) >= {"time", "x", "v", "key", "m", "proc", "rho", "p", "e", "sml", "noi", "cs"}
    assert t0["x"].shape == (n, 3) and t0["v"].shape == (n, 3) and t0["time"].shape == (1,)
    assert np.array_equal(t0["x"][:, 0], s["x"]) and np.array_equal(t0["x"][:, 2], s["z"])
    assert np.array_equal(t0["v"][:, 1], s["vy"]) and np.array_equal(t0["m"], s["mass"])
    assert t0["time"][0] == 0.0
    o = orc.sph_rhs({**s, "cs": np.zeros(n), "p": np.zeros(n)})
    assert np.array_equal(t0["sml"], o["sml"]) and np.array_equal(t0["noi"], o["noi"])
    assert np.max(np.abs(t0["rho"] - o["rho"]) / o["rho"]) < 1e-12
    assert np.max(np.abs(t0["p"] - o["p"]) / o["p"]) < 1e-12
    # keys of the run (Hilbert, over the run's bounding box)
    box = orc.bbox(s["x"], s["y"], s["z"])
    assert np.array_equal(t0["key"], orc.keys(s["x"], s["y"], s["z"], box, 1))
    t1 = _read_snapshot(out / "ts000001")
    assert t1["time"][0] > 0.0 and not np.array_equal(t1["x"], t0["x"])


def test_facade_io_gravity_and_material_check(tmp_path):
    print_case("c++ facade: gravity run from a .npy distribution; a second material is refused")
    s = ics.plummer(3000, seed=4)
    s["e"] = np.zeros(3000)
    s["sml"] = np.full(3000, 0.1)
    ic, out = tmp_path / "ic", tmp_path / "out"
    _write_ic(ic, s, with_rho=False, with_sml=False)
    out.mkdir()
    res = subprocess.run([str(DRIVER), "io", "gravity", str(ic), str(out), "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    t2 = _read_snapshot(out / "ts000002")
    assert "totalEnergy" in t2 and "totalAngularMomentum" in t2 and "rho" not in t2
    assert t2["time"][0] == pytest.approx(2e-3)
    # the Leapfrog run conserves energy to the integrator's accuracy over two steps
    t0 = _read_snapshot(out / "ts000000")
    assert abs(t2["totalEnergy"][0] / t0["totalEnergy"][0] - 1.0) < 1e-4
    mid = np.zeros(3000, np.int8)
    mid[17] = 1
    bad = tmp_path / "bad"
    _write_ic(bad, {**s}, material_id=mid)
    res = subprocess.run([str(DRIVER), "io", "sph", str(bad), str(out), "0"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 4 and "material id 1" in res.stderr


def test_facade_set_pointer_runs_rhs_on_the_second_set(orc, tmp_path):
    """ParticleHandler::setPointer(IntegratedParticleHandler*) points the context at the second
    particle set: rhs() then computes on it and leaves the particle set's arrays untouched;
    resetPointer() goes back."""
    print_case("c++ facade: setPointer / resetPointer")
    out = tmp_path / "out"
    out.mkdir()
    res = subprocess.run([str(DRIVER), "pointer", "14", str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    a, b = _read_snapshot(out / "ts000002"), _read_snapshot(out / "ts000001")
    n = len(a["m"])
    assert np.array_equal(b["x"][:, 0] - 0.25, a["x"][:, 0]) or np.allclose(b["x"][:, 0] - 0.25, a["x"][:, 0],
                                                                          rtol=0, atol=1e-15)
    assert np.array_equal(b["x"][:, 1], a["x"][:, 1])
    for snap in (a, b):
        st = {"x": snap["x"][:, 0], "y": snap["x"][:, 1], "z": snap["x"][:, 2], "vx": snap["v"][:, 0],
              "vy": snap["v"][:, 1], "vz": snap["v"][:, 2], "mass": snap["m"], "e": snap["e"],
              "sml": np.full(n, 2.5 / 14), "p": np.full(n, 0.5), "cs": np.zeros(n)}
        o = orc.sph_rhs(st)
        assert np.array_equal(snap["sml"], o["sml"])
        assert np.max(np.abs(snap["rho"] - o["rho"]) / o["rho"]) < 1e-12
