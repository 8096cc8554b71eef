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


def separate_pc_step(ctx, lib, flags, courant, t_left=None):
    """The five calls mhpc_pc_euler_step stands for, each waiting for its work."""
    s0 = ctx.rhs(flags)
    dt, _ = ctx.time_step(courant, 0.2, -1.0)
    if t_left is not None:
        dt = min(dt, t_left)
    ctx.pc_predictor(dt)
    s1 = ctx.rhs(flags & ~lib.RHS_GRAVITY)
    ctx.pc_corrector(dt)
    return dt, s0, s1


@pytest.mark.parametrize("case", ["sedov", "sedov-speculative", "sedov-ragged", "boss-bodenheimer"])
def test_fused_pc_step_equals_separate_calls(gpu_lib, case):
    """mhpc_pc_euler_step queues rhs, the CFL step (computed on the device), predictor, rhs and
    corrector without waiting in between: every field, every dt and the statistics equal those
    of the five calls made one by one, bit for bit, over four steps (the last one limited by
    the remaining time).  The Boss-Bodenheimer case has self-gravity in the first RHS only and
    FP32 + compensation pair sums; the ragged case (11^3 particles) leaves the last block of every
    512-particle window kernel part empty."""
    print_case(f"fused pc-euler step vs separate calls, {case}")
    if case == "boss-bodenheimer":
        s = ics.boss_bodenheimer(32768)
        params = ics.params_boss_bodenheimer()
        flags = gpu_lib.RHS_GRAVITY | gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML | gpu_lib.RHS_SPECULATIVE_NN
    else:
        s = ics.sedov_lattice(11 if case == "sedov-ragged" else 16)
        params = ics.params_sedov(sml_max=4.0 * float(s["sml"].max()))
        flags = gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML
        if case == "sedov-speculative":
            flags |= gpu_lib.RHS_SPECULATIVE_NN
    n = len(s["x"])
    for k in ("ax", "ay", "az", "dedt", "g_ax", "g_ay", "g_az", "muijmax", "drhodt", "dsmldt"):
        s.setdefault(k, np.zeros(n))
    out = {}
    for fused in (True, False):
        with gpu_lib.Context(n, params) as ctx:
            ctx.load(**{k: v for k, v in s.items() if k in gpu_lib.FIELD_ID})
            res = []
            for i in range(4):
                t_left = 0.25 * res[2][0] if i == 3 else None
                if fused:
                    res.append(ctx.pc_euler_step(flags, courant_fact=0.3, forces_fact=0.2, t_left=t_left))
                else:
                    res.append(separate_pc_step(ctx, gpu_lib, flags, 0.3, t_left))
            out[fused] = (res, {k: ctx.download(k) for k in FIELDS + ("noi", "h_iters", "g_ax", "g_ay", "g_az")})
    (rf, gf), (rs, gs) = out[True], out[False]
    for (dtf, a0, a1), (dts, b0, b1) in zip(rf, rs):
        assert dtf == dts
        for a, b in ((a0, b0), (a1, b1)):
            for k in ("n_neighbours", "n_h_counts", "max_noi", "num_cells", "h_passes", "n_cell_interactions",
                      "n_part_interactions"):
                assert a[k] == b[k], k
    assert rf[3][0] == 0.25 * rf[2][0]
    for k in gf:
        assert np.array_equal(gf[k], gs[k]), k


def test_fused_pc_step_refuses_learned_flags(gpu_lib):
    s = ics.sedov_lattice(8)
    with gpu_lib.Context(len(s["x"]), ics.params_sedov()) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in gpu_lib.FIELD_ID})
        import ctypes as C
        dt = C.c_double(0.0)
        st = gpu_lib.LIB.mhpc_pc_euler_step(ctx._h, C.c_uint32(gpu_lib.RHS_SPH | gpu_lib.RHS_VARIABLE_SML |
                                                             gpu_lib.RHS_FEATURES),
                                            0.3, 0.2, -1.0, -1.0, C.byref(dt), None, None)
        assert st == gpu_lib.E_ARG
