"""RHS throughput on one MI355X.

Workload (the metric's configuration on one GPU): Sedov-Taylor 128^3 = 2,097,152
particles, SPH only, plain h-iteration (band 50-80, target 64), cubic spline, ideal gas,
FP64 -- one step = one full Miluphpc::rhs() (bbox, keys, sort, tree, h-iteration,
neighbours, density, EOS, forces) with inputs resident in HBM.  The Plummer
16,777,216-particle Barnes-Hut RHS (theta 0.6, quadrupole, FP32 + Kahan) is timed beside
it and reported under "plummer".

Prints ONE JSON line.  value = RHS particles/s of the Sedov workload.
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


def sph_bytes_per_particle(nnb: float) -> dict:
    """Algorithmic (compulsory) HBM bytes per particle of each SPH kernel: each particle's
    own fields once, the neighbour list (4 B per entry + 4 B count) once, neighbour data
    assumed reused through LDS / L2."""
    return {
        "ms_nn": 32 + 4 + 4 * nnb,           # read x,y,z,h; write list + count
        "ms_density": 40 + 4 + 4 * nnb + 8,  # read x,y,z,h,m + list; write rho (the EOS is fused in)
        "ms_forces": 88 + 4 + 4 * nnb + 48,  # read 11 fields + list; write 6 fields
    }


# the kernel the roofline line is quoted for: the largest total time per RHS by kernel
# name in the rocprofv3 kernel statistics of this workload (one launch per RHS)
ROOFLINE_KERNEL = "ms_forces"
ROOFLINE_KERNEL_NAME = "k_forces<true, false>"


def measured_traffic(kernel_name: str):
    """HBM bytes per launch of `kernel_name` from the committed PMC summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a plain copy), or None if there is none."""
    path = REPO / "profiles" / "traffic_sedov128.txt"
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


def time_rhs(ctx, lib, flags, steps, warmup, label, restore=None):
    """restore: {field: device tensor} copied back (device to device, on the context
    stream) before every RHS, so each step is the same RHS on the same input state."""
    import torch

    def one():
        for field, t in (restore or {}).items():
            ctx.upload_device(field, t.data_ptr())
        return ctx.rhs(flags)

    for i in range(warmup):
        one()
        log(f"{label}: warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(steps):
        stats.append(one())
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    log(f"{label}: {steps} timed steps, {wall * 1e3:.3f} ms/step")
    return wall, stats


def median_stats(stats):
    keys = [k for k in stats[0] if k.startswith("ms_")]
    out = {k: float(np.median([s[k] for s in stats])) for k in keys}
    for k in stats[0]:
        if not k.startswith("ms_"):
            out[k] = stats[-1][k]
    return out


def cpu_baseline_sedov(s) -> dict:
    """The oracle (FP64, OpenMP) on the same Sedov 128^3 input: one full SPH RHS."""
    sys.path.insert(0, str(REPO / "oracle"))
    import pyoracle as orc

    cores = len(os.sched_getaffinity(0))
    orc.set_threads(cores)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < 10.0 and len(times) < 12):
        t0 = time.perf_counter()
        orc.sph_rhs(s)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    n = len(s["x"])
    return {"value": n / dt, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"full SPH RHS (h-iteration, neighbours, density, EOS, forces) of the Sedov "
                      f"{round(n ** (1 / 3))}^3 input, median of {len(times)} runs ({sum(times):.1f} s of CPU "
                      f"work), {dt:.3f} s per RHS on {cores} host threads"}


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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=128)
    ap.add_argument("--plummer-n", type=int, default=16777216)
    ap.add_argument("--no-plummer", action="store_true")
    ap.add_argument("--no-sedov", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-two-walk", action="store_true")
    ap.add_argument("--no-shard", action="store_true", help="skip the N/8 Plummer line")
    ap.add_argument("--learned-nside", type=int, default=128)
    ap.add_argument("--learned-steps", type=int, default=10)
    ap.add_argument("--learned-train-iters", type=int, default=100)
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
    if args.no_sedov:
        s = None
    else:
        s = run_sedov(args, lib, ics, torch, result)
    if not args.no_plummer:
        run_plummer(args, lib, ics, result)
    if not args.no_learned:
        run_learned(args, lib, ics, result)
    if not args.no_cpu and s is not None:
        log("cpu baseline (oracle) ...")
        result["cpu_baseline"] = cpu_baseline_sedov(s)
        log(f"cpu baseline {result['cpu_baseline']['value']:.4e} particles/s")
    print(json.dumps(result), flush=True)
    return 0


def run_sedov(args, lib, ics, torch, result):
    # ---- Sedov 128^3, SPH, plain h-iteration, FP64 ------------------------------------
    s = ics.sedov_lattice(args.nside)
    n = len(s["x"])
    log(f"sedov {args.nside}^3 = {n} particles generated")
    p = ics.params_sedov()
    # the plain h-iteration with its neighbour lists written by pass 0 (only particles whose
    # list can change are re-listed): the same h and neighbour sets as the reference's
    # count-then-list structure, which is timed beside it
    flags = lib.RHS_SPH | lib.RHS_VARIABLE_SML | lib.RHS_SPECULATIVE_NN
    two_walk = lib.RHS_SPH | lib.RHS_VARIABLE_SML
    # the h-iteration rewrites sml; every step restarts from the initial smoothing lengths
    h0 = torch.tensor(s["sml"], dtype=torch.float64, device="cuda")
    with lib.Context(n, p) as ctx:
        ctx.load(**{k: v for k, v in s.items() if k in lib.FIELD_ID})
        wall, stats = time_rhs(ctx, lib, flags, args.steps, args.warmup, "sedov", restore={"sml": h0})
        if not args.no_two_walk:
            wall_s, stats_s = time_rhs(ctx, lib, two_walk, args.steps, 1, "sedov-two-walk", restore={"sml": h0})
    if not args.no_two_walk:
        med_s = median_stats(stats_s)
        result["sedov_two_walk"] = {
            "note": "same RHS with the reference's structure: count-only h-iteration, then a full list walk",
            "value": n / wall_s, "ms_per_step": wall_s * 1e3,
            "phases_ms": {k[3:]: round(v, 4) for k, v in med_s.items() if k.startswith("ms_")}}
    med = median_stats(stats)
    nnb = med["n_neighbours"] / n
    bpp = sph_bytes_per_particle(nnb)
    dom = ROOFLINE_KERNEL
    alg_bytes = bpp[dom] * n
    achieved = alg_bytes / (med[dom] * 1e-3) / 1e9
    kernels = {k[3:]: {"ms": round(med[k], 4), "alg_bytes_per_particle": round(bpp[k], 1),
                       "GB/s": round(bpp[k] * n / (med[k] * 1e-3) / 1e9, 1),
                       "frac": round(bpp[k] * n / (med[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
               for k in bpp}
    result.update({
        "value": n / wall,
        "ms_per_step": wall * 1e3,
        "dtype": "f64",
        "config": {"workload": f"Sedov-Taylor {args.nside}^3 lattice SPH RHS, plain h-iteration "
                               f"(target 64, band 50-80), cubic spline, ideal gas, FP64",
                   "n_particles": n, "path": "plain h-iteration, lists from pass 0 (speculative)",
                   "sort": "key levels of the previous RHS's tree + 1, checked, full sort when needed "
                           "(the timed RHS follow the warm-up)"},
        "roofline": {"bound": "hbm", "kernel": ROOFLINE_KERNEL_NAME, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(ROOFLINE_KERNEL_NAME),
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "algorithmic_bytes_per_particle": bpp[dom], "kernel_ms": med[dom]},
        "sph_kernels": kernels,
        "phases_ms": {k[3:]: round(v, 4) for k, v in med.items() if k.startswith("ms_")},
        "mean_neighbours": nnb,
        "mean_h_counts": med["n_h_counts"] / n,
    })
    log(f"sedov: {n / wall:.4e} particles/s, dominant {dom} {med[dom]:.3f} ms, {achieved:.1f} GB/s")
    return s


STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "mass", "sml", "e", "rho", "p", "cs", "ax", "ay", "az",
                "g_ax", "g_ay", "g_az", "drhodt", "dedt", "dsmldt", "muijmax", "curlv", "gradrho", "noi")


def run_learned(args, lib, ics, result):
    """Learned-h vs plain on an evolving Sedov run (same GPU, same bands): 5 plain
    predictor-corrector steps produce the training pairs (the network is trained online
    on the GPU after each RHS), then the same state is advanced K steps three ways --
    plain (reference structure: count passes, then a full list walk), speculative lists
    with the previous h as the guess (no network), and learned (network prediction +
    speculative lists) -- and the RHS time and h-iteration counts are compared."""
    import torch

    nside = args.learned_nside
    s = ics.sedov_lattice(nside)
    n = len(s["x"])
    for k in ("ax", "ay", "az", "dedt", "g_ax", "g_ay", "g_az", "muijmax", "drhodt", "dsmldt"):
        s.setdefault(k, np.zeros(n))
    # free boundary: bound h at 4 h0 so the corner particles (which can never reach the band)
    # do not grow h without limit over the run
    p = ics.params_sedov(sml_max=4.0 * float(s["sml"].max()))
    base = lib.RHS_SPH | lib.RHS_VARIABLE_SML
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
        "workload": f"Sedov {nside}^3 evolving (PC-Euler, CFL 0.3): 5 plain training steps, then "
                    f"{args.learned_steps} steps each way from the same state; band 50-80, FP64",
        "n_particles": n, "train_loss_first_last": [losses[0], losses[-1]],
        **out,
        "speedup_learned_vs_plain": out["plain"]["rhs_ms"] / out["learned"]["rhs_ms"],
        "speedup_speculative_vs_plain": out["plain"]["rhs_ms"] / out["speculative"]["rhs_ms"],
    }


def run_plummer(args, lib, ics, result):
    # ---- Plummer 16M, Barnes-Hut, theta 0.6, quadrupole, FP32 + Kahan ----------------
    if True:
        pl = ics.plummer(args.plummer_n)
        log(f"plummer {args.plummer_n} generated")
        pp = ics.params_plummer_large()
        with lib.Context(args.plummer_n, pp) as ctx:
            ctx.load(**pl)
            pw, pst = time_rhs(ctx, lib, lib.RHS_GRAVITY, max(1, args.steps // 2), 1, "plummer")
        pm = median_stats(pst)
        flops = 23.0 * pm["n_part_interactions"] + 68.0 * pm["n_cell_interactions"]
        tflops = flops / (pm["ms_gravity"] * 1e-3) / 1e12
        result["plummer"] = {
            "workload": f"Plummer {args.plummer_n} Barnes-Hut RHS, theta 0.6, eps 2e-4, monopole + "
                        "quadrupole, FP32 + Kahan far field, FP64 positions and keys",
            "value": args.plummer_n / pw, "unit": "particles/s", "ms_per_step": pw * 1e3,
            "phases_ms": {k[3:]: round(v, 4) for k, v in pm.items() if k.startswith("ms_")},
            "interactions_per_particle": (pm["n_part_interactions"] + pm["n_cell_interactions"]) / args.plummer_n,
            "roofline": {"bound": "fp32", "kernel": "gravity", "achieved": tflops, "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TFLOPS,
                         "flops_per_interaction": {"particle": 23, "cell_quadrupole": 68}},
        }
        log(f"plummer: {args.plummer_n / pw:.4e} particles/s, gravity {pm['ms_gravity']:.2f} ms, {tflops:.2f} TFLOP/s")
        del pl
        if not args.no_shard:
            # the per-GPU particle count of the same model on an 8-GPU node: a whole Plummer sphere
            # of N / 8 particles (a proxy for one shard's kernel shape, not a spatial sub-domain)
            n8 = args.plummer_n // 8
            pl = ics.plummer(n8)
            with lib.Context(n8, pp) as ctx:
                ctx.load(**pl)
                pw8, pst8 = time_rhs(ctx, lib, lib.RHS_GRAVITY, max(1, args.steps // 2), 1, "plummer-shard")
            pm8 = median_stats(pst8)
            result["plummer"]["shard_of_8"] = {
                "n_particles": n8, "value": n8 / pw8, "ms_per_step": pw8 * 1e3,
                "gravity_ms": pm8["ms_gravity"]}
            del pl


if __name__ == "__main__":
    sys.exit(main())
