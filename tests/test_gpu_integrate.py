"""GPU predictor-corrector Euler and CFL time step against the oracle (Sedov, FP64)."""
import numpy as np
import pytest

from miluphpc_amd import ics
from conftest import print_case

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "sml", "e", "rho", "p", "cs", "ax", "ay", "az", "dedt")


def oracle_pc_step(orc, s, courant=0.3, forces=0.2):
    """One PC-Euler step of the SPH-only RHS with the oracle (FP64)."""
    def rhs(st):
        o = orc.sph_rhs(st)
        for k in ("sml", "rho", "p", "cs", "ax", "ay", "az", "dedt", "drhodt", "dsmldt", "muijmax"):
            st[k] = o[k]
    rhs(s)
    dt, _ = orc.time_step(s, 1.0, 2.0, courant, forces)
    backup = [np.zeros(len(s["x"])) for _ in range(11)]
    orc.pc_predictor(dt, s, backup)
    rhs(s)
    orc.pc_corrector(dt, s, backup)
    return dt


def test_pc_euler_sedov_matches_oracle(gpu_lib, orc):
    print_case("pc-euler 3 steps sedov 16^3 fp64")
    s = ics.sedov_lattice(16)
    n = len(s["x"])
    for k in ("ax", "ay", "az", "dedt", "g_ax", "g_ay", "g_az", "muijmax", "drhodt", "dsmldt"):
        s.setdefault(k, np.zeros(n))
    o = {k: v.copy() for k, v in s.items()}
    p = ics.params_sedov()
    with gpu_lib.Context(n, p) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in gpu_lib.FIELD_ID})
        dts = []
        for _ in range(3):
            dt, _, _ = ctx.pc_euler_step(gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML, courant_fact=0.3)
            dts.append(dt)
        g = {k: ctx.download(k) for k in FIELDS}
    odts = [oracle_pc_step(orc, o) for _ in range(3)]
    assert np.allclose(dts, odts, rtol=1e-12, atol=0)
    assert np.array_equal(g["sml"], o["sml"])
    for k in FIELDS:
        scale = max(np.abs(o[k]).max(), 1e-300)
        assert np.max(np.abs(g[k] - o[k])) <= 1e-10 * scale, k
    # the blast is moving: energy went from the centre into kinetic motion
    assert np.abs(g["vx"]).max() > 0.0
