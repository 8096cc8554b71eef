"""RHS throughput on one MI355X.

Workload (the metric's configuration on one GPU): Sedov-Taylor 128^3 = 2,097,152
particles, SPH only, plain h-iteration (band 50-80, target 64), cubic spline, ideal gas,
FP64, advanced by predictor-corrector Euler steps under a CFL 0.3 time step (two
Miluphpc::rhs() calls per step: bbox, keys, sort, tree, h-iteration, neighbours, density,
EOS, forces) with inputs resident in HBM.  One timed step is one such PC-Euler step of the
evolving run; value = RHS particles/s = 2 N / (median step wall time).  The Plummer
16,777,216-particle Barnes-Hut RHS (theta 0.6, quadrupole, FP32 + Kahan) is timed beside it
and reported under "plummer", with its energy drift over five leapfrog steps and the 99th
percentile force error against FP64 direct summation on 4,096 particles.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# kernels the SPH roofline can be quoted on: stats field -> (kernel name as rocprofv3 lists it,
# algorithmic HBM bytes per particle at nnb neighbours).  Each particle's own fields count once,
# the neighbour list (4 B per entry + 4 B count) once; neighbour data is assumed reused through
# LDS / L2.
def sph_kernel_table(nnb: float, precision_fp64: bool = True) -> dict:
    return {
        "ms_nn_walk": ("k_nn_walk<2>", 32 + 4 + 4 * nnb),            # read x,y,z,h; write list + count
        "ms_density": ("k_density<true, true, true>", 40 + 4 + 4 * nnb + 8),   # read x,y,z,h,m + list; write rho
        "ms_forces": ("k_forces<true, false, true>", 88 + 4 + 4 * nnb + 48),   # read 11 fields + list; write 6
    }


def measured_traffic(kernel_name: str, workload: str = "sedov128"):
    """HBM bytes per launch of `kernel_name` from the committed PMC summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a plain copy), or None if there is none."""
    path = REPO / "profiles" / f"traffic_{workload}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split("\t")
        if len(parts) >= 2 and parts[0].strip() == kernel_name:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def measured_busy(kernel_name: str, workload: str = "sedov128"):
    """Busy shares of `kernel_name` from the committed PMC summary (texture addresser, VALU,
    LDS), or None if there is none: what bounds a kernel that is not HBM-bound."""
    path = REPO / "profiles" / f"busy_{workload}.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split("\t")
        if len(parts) >= 6 and parts[0].strip() == kernel_name:
            try:
                return {"valu": float(parts[3]), "texture_addresser": float(parts[2]), "lds": float(parts[5]),
                        "valu_instructions_per_wave": float(parts[4]), "source": f"profiles/busy_{workload}.txt"}
            except ValueError:
                return None
    return None


def sync():
    import torch

    torch.cuda.synchronize()


def median_stats(stats):
    keys = [k for k in stats[0] if k.startswith("ms_")]
    out = {k: float(np.median([s[k] for s in stats])) for k in keys}
    for k in stats[0]:
        if not k.startswith("ms_"):
            out[k] = stats[-1][k]
    return out


def time_steps(step, steps: int, warmup: int, label: str, ctx=None):
    """W untimed steps, then K steps, the timed region bracketed by device synchronisations;
    returns the per-step wall times (s), the total wall time of the K steps and the steps'
    results.  Each step ends with a synchronisation: the context's own (spinning) one when ctx
    is given -- all of a step's work is on the context's stream -- else the device-wide one."""
    for i in range(warmup):
        step()
        log(f"{label}: warmup {i + 1}/{warmup} done")
    sync()
    walls, res = [], []
    t_all = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        res.append(step())
        if ctx is not None:
            ctx.synchronize()
        else:
            sync()
        walls.append(time.perf_counter() - t0)
    sync()
    total = time.perf_counter() - t_all
    log(f"{label}: {steps} timed steps, median {np.median(walls) * 1e3:.3f} ms/step")
    return np.array(walls), total, res


def cpu_baseline_sedov(state: dict, sml_max: float) -> dict:
    """The oracle (FP64, OpenMP, -O3 AVX2) on the same evolved Sedov state: full SPH RHS
    (h-iteration, neighbour lists, density, EOS, forces), timed at C level."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle as orc

    cores = len(os.sched_getaffinity(0))
    orc.set_threads(cores)
    secs = list(orc.bench_sph_rhs(state, 2, sml_max=sml_max))
    while sum(secs) < 10.0 and len(secs) < 12:
        secs += list(orc.bench_sph_rhs(state, 1, sml_max=sml_max))
    dt = float(np.median(secs))
    n = len(state["x"])
    return {"value": n / dt, "unit": "particles/s", "cores": cores, "kind": "port", "impl": "oracle",
            "sample": f"the FP64 oracle (oracle/oracle.c, one CPU restatement of the same algorithm: Jacobi "
                      f"h-iteration, grid neighbour lists, density, EOS, forces; gcc -O3 -march=x86-64-v3 "
                      f"-fopenmp) on the evolved Sedov {round(n ** (1 / 3))}^3 state of the timed run, timed at "
                      f"C level: median of {len(secs)} RHS ({sum(secs):.1f} s of CPU work), {dt:.3f} s per RHS "
                      f"on {cores} host threads"}


def cpu_baseline_plummer(n: int, params, sample_every: int = 16) -> dict:
    """The oracle's Barnes-Hut RHS on a Plummer sphere of n particles (the per-GPU share of the
    16.7M sphere on eight GPUs): tree phases over all particles, the group walk over every
    sample_every-th group, extrapolated to all groups."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle as orc
    from miluphpc_amd import ics

    cores = len(os.sched_getaffinity(0))
    orc.set_threads(cores)
    pl = ics.plummer(n)
    r = orc.bench_gravity(pl, params.theta, params.smoothing, curve=params.curve, mac=params.mac,
                          quad=bool(params.quadrupole), sample_every=sample_every, vote=params.bh_accept == 1)
    groups = (n + 63) // 64
    t = r["tree_s"] + r["walk_s"] * groups / r["groups_walked"]
    return {"value": n / t, "unit": "particles/s", "cores": cores, "kind": "port", "impl": "oracle",
            "sample": f"FP64 oracle Barnes-Hut RHS (bbox, keys, sort, insertion octree, COM + quadrupole, "
                      f"group walk of the same acceptance form) on a {n}-particle Plummer sphere: tree phases "
                      f"{r['tree_s']:.2f} s over all particles, walk {r['walk_s']:.2f} s over "
                      f"{r['groups_walked']} of {groups} groups (every {sample_every}th) extrapolated; "
                      f"{t:.1f} s per RHS on {cores} host threads"}


def measured_copy_ceiling(torch, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Measured device-copy bandwidth (read + write bytes / s) of a 1 GiB elementwise kernel,
    the practical HBM ceiling beside the 8 TB/s peak."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        torch.neg(src, out=dst)
    ev1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del src, dst
    log(f"measured copy ceiling {gbs:.0f} GB/s")
    return gbs


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=18, help="timed PC-Euler steps of the Sedov run")
    ap.add_argument("--warmup", type=int, default=2, help="untimed PC-Euler steps before them")
    ap.add_argument("--nside", type=int, default=128)
    ap.add_argument("--plummer-n", type=int, default=16777216)
    ap.add_argument("--plummer-steps", type=int, default=5)
    ap.add_argument("--no-plummer", action="store_true")
    ap.add_argument("--no-sedov", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-initial", action="store_true", help="skip the initial-state RHS line")
    ap.add_argument("--no-shard", action="store_true", help="skip the N/8 Plummer line")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the Plummer direct-sum check and drift")
    ap.add_argument("--learned-nside", type=int, default=128)
    ap.add_argument("--learned-steps", type=int, default=10)
    ap.add_argument("--learned-train-iters", type=int, default=100)
    ap.add_argument("--learned-workload", choices=("sedov", "bb"), default="sedov")
    ap.add_argument("--learned-bb-n", type=int, default=1 << 20)
    ap.add_argument("--learned-band", type=int, nargs=2, default=None, metavar=("LO", "HI"),
                    help="neighbour-count band of the learned-vs-plain comparison (default: the workload's 50-80)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first (profiler byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one subdomain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU decomposition is out of scope for this build)", file=sys.stderr)
        return 2

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from miluphpc_amd import ics, lib

    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)   # one elementwise kernel: reads and writes nbytes each
        torch.cuda.synchronize()
        log(f"calibration copy (neg kernel): {nbytes} bytes read, {nbytes} bytes written")
        del src, dst

    copy_gbs = measured_copy_ceiling(torch)
    result = {
        "metric": "RHS particles/s: Sedov 2M SPH, Plummer 16M BH; learned-h vs plain, same GPU",
        "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
        "hbm_copy_ceiling_GBs": copy_gbs,
    }
    state = None
    if not args.no_sedov:
        state = run_sedov(args, lib, ics, result)
        if not args.no_initial:
            run_sedov_initial(args, lib, ics, result)
    if not args.no_plummer:
        run_plummer(args, lib, ics, result)
    if not args.no_learned:
        run_learned(args, lib, ics, result)
    if not args.no_cpu and state is not None:
        log("cpu baseline (oracle, Sedov) ...")
        result["cpu_baseline"] = cpu_baseline_sedov(state, result["config"]["sml_max"])
        log(f"cpu baseline {result['cpu_baseline']['value']:.4e} particles/s")
        if "plummer" in result and not args.no_shard:
            log("cpu baseline (oracle, Plummer N/8) ...")
            cb = cpu_baseline_plummer(args.plummer_n // 8, ics.params_plummer_large())
            result["plummer"]["shard_of_8"]["cpu_baseline"] = cb
            log(f"cpu baseline plummer {cb['value']:.4e} particles/s")
    print(json.dumps(result), flush=True)
    return 0


STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "mass", "sml", "e", "rho", "p", "cs", "ax", "ay", "az",
                "g_ax", "g_ay", "g_az", "drhodt", "dedt", "dsmldt", "muijmax", "curlv", "gradrho", "noi")


def run_sedov(args, lib, ics, result):
    """The headline: Sedov 128^3 advanced by PC-Euler steps (CFL 0.3), each step two RHS of the
    plain h-iteration with its neighbour lists written by pass 0 (only particles whose list can
    change are re-listed: the same h and neighbour sets as the reference's count-then-list
    structure)."""
    s = ics.sedov_lattice(args.nside)
    n = len(s["x"])
    log(f"sedov {args.nside}^3 = {n} particles generated")
    for k in ("ax", "ay", "az", "dedt", "g_ax", "g_ay", "g_az", "muijmax", "drhodt", "dsmldt"):
        s.setdefault(k, np.zeros(n))
    # free boundary: h is bounded at 4 h0 so that the corner and edge particles (which can never
    # reach the band under d < min(h_i, h_j)) do not grow h without limit over the run
    sml_max = 4.0 * float(s["sml"].max())
    p = ics.params_sedov(sml_max=sml_max)
    flags = lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_SPECULATIVE_NN
    with lib.Context(n, p) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
        walls, total, res = time_steps(lambda: ctx.pc_euler_step(flags, courant_fact=0.3), args.steps,
                                       args.warmup, "sedov", ctx)
        state = {k: ctx.download(k) for k in STATE_FIELDS}
        # two more steps with every phase timed (per-phase events are left out of the timed steps;
        # those time the total, the pass-0 walk, the density and the forces kernels)
        phase_res = [ctx.pc_euler_step(flags | lib.RHS_TIMING, courant_fact=0.3) for _ in range(2)]
    stats = [st for _, s0, s1 in res for st in (s0, s1)]
    dts = [dt for dt, _, _ in res]
    med = median_stats(stats)
    phase_med = median_stats([st for _, s0, s1 in phase_res for st in (s0, s1)])
    nnb = float(np.mean([st["n_neighbours"] for st in stats])) / n
    table = sph_kernel_table(nnb)
    kernels = {}
    for k, (name, bpp) in table.items():
        gbs = bpp * n / (med[k] * 1e-3) / 1e9
        kernels[k[3:]] = {"kernel": name, "ms": round(med[k], 4), "alg_bytes_per_particle": round(bpp, 1),
                          "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    kernels["nn_fixup"] = {"kernel": "k_nn_mark_dirty + k_nn_walk<0> + k_nn_one<0> (re-lists)",
                           "ms": round(phase_med["ms_nn"], 4), "note": "re-lists only the changed neighbourhoods"}
    kernels["hiter_rest"] = {"kernel": "pass-0 outliers + tail passes (k_nn_one, k_nn_count_direct, k_h_finish)",
                             "ms": round(phase_med["ms_hiter"] - phase_med["ms_nn_walk"], 4)}
    # the roofline is quoted on the kernel with the largest time per RHS
    dom = max(table, key=lambda k: med[k])
    name, bpp = table[dom]
    alg_bytes = bpp * n
    achieved = alg_bytes / (med[dom] * 1e-3) / 1e9
    step_ms = float(np.median(walls)) * 1e3
    result.update({
        "value": 2 * n / (step_ms * 1e-3),
        "ms_per_step": step_ms,
        "ms_per_step_mean": total / args.steps * 1e3,
        "ms_per_rhs": step_ms / 2,
        "dtype": "f64",
        "config": {"workload": f"Sedov-Taylor {args.nside}^3 lattice SPH, PC-Euler steps under CFL 0.3 (2 RHS "
                               f"per step), plain h-iteration (target 64, band 50-80, h bounded at 4 h0), cubic "
                               f"spline, ideal gas, FP64",
                   "n_particles": n, "rhs_per_step": 2, "sml_max": sml_max,
                   "steps_before_timing": args.warmup,
                   "path": "plain h-iteration, lists from pass 0 (speculative)",
                   "sort": "key levels of the previous RHS's tree + 1, checked, full sort when needed"},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name),
                     "algorithmic_bytes_per_launch": alg_bytes, "algorithmic_bytes_per_particle": bpp,
                     "kernel_ms": med[dom], "timing": "hipEvents around the kernel on the context stream, "
                                                      "median over the timed RHS",
                     "busy": measured_busy(name)},
        "sph_kernels": kernels,
        "phases_ms": {k[3:]: round(v, 4) for k, v in phase_med.items() if k.startswith("ms_")},
        "phases_note": "per-phase times from two steps after the timed ones, with every phase timed",
        "mean_neighbours": nnb,
        "mean_h_counts": float(np.mean([st["n_h_counts"] for st in stats])) / n,
        "h_passes_max": int(max(st["h_passes"] for st in stats)),
        "sim_time": float(np.sum(dts)),
    })
    log(f"sedov: {result['value']:.4e} RHS particles/s, {step_ms / 2:.3f} ms/RHS, dominant {name} "
        f"{med[dom]:.3f} ms, {achieved:.1f} GB/s")
    return state


def run_sedov_initial(args, lib, ics, result):
    """The same RHS on the Sedov initial state (h restored on the device before each RHS), and
    the reference's count-then-list structure beside it."""
    import torch

    s = ics.sedov_lattice(args.nside)
    n = len(s["x"])
    p = ics.params_sedov()
    out = {}
    h0 = torch.tensor(s["sml"], dtype=torch.float64, device="cuda")
    with lib.Context(n, p) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
        for name, flags in (("speculative", lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_SPECULATIVE_NN | lib.RHS_TIMING),
                            ("two_walk", lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_TIMING)):
            def one():
                ctx.upload_device("sml", h0.data_ptr())
                return ctx.rhs(flags)

            walls, _, res = time_steps(one, 10, 2, f"sedov-initial-{name}", ctx)
            med = median_stats(res)
            out[name] = {"ms_per_rhs": float(np.median(walls)) * 1e3, "value": n / float(np.median(walls)),
                         "phases_ms": {k[3:]: round(v, 4) for k, v in med.items() if k.startswith("ms_")}}
    result["sedov_initial_state"] = {
        "note": "RHS on the Sedov initial state (h restored before each RHS); two_walk = the reference's "
                "structure, a count-only h-iteration then a full list walk", **out}


def run_learned(args, lib, ics, result):
    """Learned-h vs plain on an evolving Sedov run (same GPU, same bands): 5 plain
    predictor-corrector steps produce the training pairs (the network is trained online
    on the GPU after each RHS), then the same state is advanced K steps three ways --
    plain (reference structure: count passes, then a full list walk), speculative lists
    with the previous h as the guess (no network), and learned (network prediction +
    speculative lists) -- and the RHS time and h-iteration counts are compared."""
    import torch

    nside = args.learned_nside
    band = tuple(args.learned_band) if args.learned_band else (50, 80)
    if args.learned_workload == "bb":
        # Boss-Bodenheimer: self-gravity + SPH, random particles in a sphere (no lattice shells,
        # so narrow bands are reachable), FP32 + Kahan pair sums
        s = ics.boss_bodenheimer(args.learned_bb_n)
        p = ics.params_boss_bodenheimer()
        p.band_lo, p.band_hi = band
        p.sml_max = 4.0 * float(s["sml"].max())
        base = lib.RHS_GRAVITY | lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_TIMING
        wl = f"Boss-Bodenheimer {len(s['x'])} (self-gravity + SPH, FP32 + Kahan)"
    else:
        s = ics.sedov_lattice(nside)
        p = ics.params_sedov(sml_max=4.0 * float(s["sml"].max()), band_lo=band[0], band_hi=band[1])
        base = lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_TIMING
        wl = f"Sedov {nside}^3"
    n = len(s["x"])
    for k in ("ax", "ay", "az", "dedt", "g_ax", "g_ay", "g_az", "muijmax", "drhodt", "dsmldt"):
        s.setdefault(k, np.zeros(n))
    train = dict(nsamples=65536, iters=args.learned_train_iters, lr=2e-3, seed=5)
    out = {}
    with lib.Context(n, p) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
        ctx.mlp_init(20261202)
        t0 = time.perf_counter()
        losses = []
        for step in range(5):
            _, s0, s1 = ctx.pc_euler_step(base | lib.RHS_FEATURES, courant_fact=0.3, train=train)
            losses += [s0["mlp_loss"], s1["mlp_loss"]]
        torch.cuda.synchronize()
        log(f"learned: 5 plain training steps ({n} particles), loss {losses[0]:.3g} -> {losses[-1]:.3g}, "
            f"{time.perf_counter() - t0:.2f} s")
        snap = {k: ctx.download(k) for k in STATE_FIELDS}
        for name, flags in (("plain", base), ("speculative", base | lib.RHS_SPECULATIVE_NN),
                            ("learned", base | lib.RHS_LEARNED_H)):
            ctx.load(**snap)
            rhs_ms, hcounts, hiter_ms, nn_ms, mlp_ms = [], [], [], [], []
            for step in range(args.learned_steps):
                _, s0, s1 = ctx.pc_euler_step(flags, courant_fact=0.3)
                for st in (s0, s1):
                    rhs_ms.append(st["ms_total"])
                    hcounts.append(st["n_h_counts"] / n)
                    hiter_ms.append(st["ms_hiter"])
                    nn_ms.append(st["ms_nn"])
                    mlp_ms.append(st["ms_mlp"])
            out[name] = {"rhs_ms": float(np.median(rhs_ms)), "mean_h_iterations": float(np.mean(hcounts)),
                         "hiter_ms": float(np.median(hiter_ms)), "nn_ms": float(np.median(nn_ms)),
                         "mlp_ms": float(np.median(mlp_ms))}
            log(f"learned-vs-plain: {name}: {out[name]['rhs_ms']:.3f} ms/RHS, "
                f"{out[name]['mean_h_iterations']:.4f} h-iterations")
    result["learned_h"] = {
        "workload": f"{wl} evolving (PC-Euler, CFL 0.3): 5 plain training steps, then "
                    f"{args.learned_steps} steps each way from the same state; band {band[0]}-{band[1]}, FP64",
        "n_particles": n, "train_loss_first_last": [losses[0], losses[-1]],
        **out,
        "speedup_learned_vs_plain": out["plain"]["rhs_ms"] / out["learned"]["rhs_ms"],
        "speedup_speculative_vs_plain": out["plain"]["rhs_ms"] / out["speculative"]["rhs_ms"],
    }


def plummer_accuracy(ctx, lib, pl, params, nsample: int = 4096):
    """99th-percentile relative force error of the RHS just computed against FP64 direct
    summation (the oracle) on a seeded random subsample."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle as orc

    orc.set_threads(len(os.sched_getaffinity(0)))
    sub = np.sort(np.random.default_rng(4096).choice(len(pl["x"]), nsample, replace=False)).astype(np.int32)
    g = {k: ctx.download(k)[sub] for k in ("g_ax", "g_ay", "g_az")}
    t0 = time.perf_counter()
    d = orc.direct(pl["x"], pl["y"], pl["z"], pl["mass"], params.smoothing, targets=sub)
    num = np.sqrt((g["g_ax"] - d["ax"]) ** 2 + (g["g_ay"] - d["ay"]) ** 2 + (g["g_az"] - d["az"]) ** 2)
    err = num / np.sqrt(d["ax"] ** 2 + d["ay"] ** 2 + d["az"] ** 2)
    log(f"plummer: direct-sum check of {nsample} particles {time.perf_counter() - t0:.1f} s, "
        f"p99 {np.percentile(err, 99):.3e}")
    return {"subsample": nsample, "seed": 4096, "p50": float(np.median(err)), "p99": float(np.percentile(err, 99)),
            "max": float(err.max())}


def run_plummer(args, lib, ics, result):
    # ---- Plummer 16M, Barnes-Hut, theta 0.6, quadrupole, FP32 + Kahan ----------------
    pl = ics.plummer(args.plummer_n)
    log(f"plummer {args.plummer_n} generated")
    pp = ics.params_plummer_large()
    with lib.Context(args.plummer_n, pp) as ctx:
        ctx.load(**pl)
        walls, _, pst = time_steps(lambda: ctx.rhs(lib.RHS_GRAVITY | lib.RHS_TIMING), max(1, args.steps // 2), 1,
                                   "plummer", ctx)
        pm = median_stats(pst)
        acc = None if args.no_accuracy else plummer_accuracy(ctx, lib, pl, pp)
        drift = None
        if not args.no_accuracy:
            # energy drift: leapfrog (kick-drift-kick) steps at dt = 5e-4 with the potential
            dt = 5e-4
            flags = lib.RHS_GRAVITY | lib.RHS_POTENTIAL
            ctx.rhs(flags)
            energies = [ctx.diagnostics()["total"]]
            for _ in range(args.plummer_steps):
                ctx.leapfrog_update_x(dt)
                ctx.rhs(flags)
                ctx.leapfrog_update_v(dt)
                energies.append(ctx.diagnostics()["total"])
            e = np.array(energies)
            rel = (e - e[0]) / abs(e[0])
            drift = {"steps": args.plummer_steps, "dt": dt, "E0": float(e[0]),
                     "rel_drift_per_step": float(rel[-1] / args.plummer_steps),
                     "max_abs_rel_drift": float(np.abs(rel).max()), "rel_history": [float(v) for v in rel]}
            log(f"plummer: energy drift per step {drift['rel_drift_per_step']:.3e}")
    flops = 23.0 * pm["n_part_interactions"] + 68.0 * pm["n_cell_interactions"]
    tflops = flops / (pm["ms_gravity"] * 1e-3) / 1e12
    step_ms = float(np.median(walls)) * 1e3
    result["plummer"] = {
        "workload": f"Plummer {args.plummer_n} Barnes-Hut RHS, theta 0.6, eps 2e-4, monopole + quadrupole, "
                    f"{'per-member (vote)' if pp.bh_accept == lib.BH_VOTE else 'group-box'} acceptance, "
                    "FP32 + Kahan far field, FP64 positions and keys",
        "value": args.plummer_n / (step_ms * 1e-3), "unit": "particles/s", "ms_per_step": step_ms,
        "phases_ms": {k[3:]: round(v, 4) for k, v in pm.items() if k.startswith("ms_")},
        "interactions_per_particle": (pm["n_part_interactions"] + pm["n_cell_interactions"]) / args.plummer_n,
        "roofline": {"bound": "fp32", "kernel": "k_bh_walk", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS,
                     "flops_per_interaction": {"particle": 23, "cell_quadrupole": 68}},
        "accuracy_vs_direct": acc,
        "energy_drift": drift,
    }
    log(f"plummer: {args.plummer_n / (step_ms * 1e-3):.4e} particles/s, gravity {pm['ms_gravity']:.2f} ms, "
        f"{tflops:.2f} TFLOP/s")
    del pl
    if not args.no_shard:
        # the per-GPU particle count of the same model on an 8-GPU node: a whole Plummer sphere
        # of N / 8 particles (a proxy for one shard's kernel shape, not a spatial sub-domain)
        n8 = args.plummer_n // 8
        pl = ics.plummer(n8)
        with lib.Context(n8, pp) as ctx:
            ctx.load(**pl)
            w8, _, pst8 = time_steps(lambda: ctx.rhs(lib.RHS_GRAVITY | lib.RHS_TIMING), max(1, args.steps // 2), 1,
                                     "plummer-shard",
                                     ctx)
        pm8 = median_stats(pst8)
        result["plummer"]["shard_of_8"] = {
            "n_particles": n8, "value": n8 / float(np.median(w8)), "ms_per_step": float(np.median(w8)) * 1e3,
            "gravity_ms": pm8["ms_gravity"]}
        del pl


if __name__ == "__main__":
    sys.exit(main())
