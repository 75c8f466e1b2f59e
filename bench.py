"""Benchmark: Huygens superposition on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-extras] [--skip-cpu]

Timed step: the FP64 brute-force chain x -> y -> z at full size
(x: 512^2 sources at z=0, 10 um, amplitude = 64 Gaussian spots, phase U[0,2pi);
 y: 1024^2 projector at z=0.35 m, 7.56 um, sign +1;
 z: 512^2 plane at z=0, 10 um, sign -1; lambda 632.8 nm), 2 x 2.75e11 pairs,
with the time_transform normalisation + polar conversion between the stages.
Inputs are resident in HBM when the timed region starts.

Side measurements printed in the same JSON line: FP32-split vs FP64 on
1024^2 -> 1024^2 (pairs/s and relative L2 / max-abs error), the Monte Carlo
estimator for K in {256, 1024, 4096} (pairs/s and error vs brute force), one
hologram forward + adjoint gradient step 1024^2 -> 512^2 (ms), and the NumPy
restatement of the reference's CPU path timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "phasor pairs/s (FP64 vs FP32-split @rel err≤1e-6) + ms per adjoint grad step"
FLOP_PER_PAIR = 29            # cuda/macros.h:71-80
PEAK_FP64 = 78.6e12           # MI355X FP64 vector (and matrix) peak, FLOP/s
# FP64 FLOPs per pair that k_superpose_f64 executes (SQ_INSTS_VALU_FLOPS_FP64 x 64
# per pair, profiles/k_superpose_f64_counters.md) on the chain's lattice sources
# (one plane, rows of one x): the hand-written sqrt and sin/cos cost more than
# the reference's nominal 29.
COUNTED_FLOP_PER_PAIR = 44.25


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def read_profile(kernel: str, key: str):
    """A figure (`key = value`) from the committed rocprofv3 summary
    profiles/<kernel>_traffic.txt, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            try:
                return float(line.split("=", 1)[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


# The CPU share of a one-GPU job on a shared multi-GPU host (16 of its cores), whatever
# os.cpu_count() reports there; on smaller hosts every available core is used.
CPU_SHARE = 16


def cpu_baseline(stage, n_targets: int, repeats: int):
    """The NumPy restatement of py/util.py:92-147 (map_sum_kernel over a Pool,
    util.py:134-141) on the fixed stride subset of n_targets stage-1 targets:
    one warm-up on 256 of them, then `repeats` timed runs over the whole
    subset (median and std of pairs/s).  Runs before the GPU is initialised,
    so the worker processes fork from a process without it."""
    from oracle import numpy_baseline as nb
    cores = min(CPU_SHARE, nb.cpu_count())
    M = stage.v.shape[0]
    sub = (np.arange(n_targets, dtype=np.int64) * M) // n_targets
    x = np.stack([stage.amp, stage.phase], axis=1)
    y = np.empty((n_targets, 2))
    N = stage.u.shape[0]
    # CUDA sign s corresponds to the Python direction -s (py/util.py:106)
    run = lambda t, out: nb.map_sum_kernel(x, out, stage.u, stage.v[t], direction=-stage.sign,
                                           lam=stage.lam, parallel=cores)
    run(sub[:: max(1, n_targets // 256)], np.empty((len(sub[:: max(1, n_targets // 256)]), 2)))
    rates, walls = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run(sub, y)
        dt = time.perf_counter() - t0
        walls.append(dt)
        rates.append(N * n_targets / dt)
    res = {"value": float(np.median(rates)), "unit": "pairs/s", "cores": cores, "kind": "port",
           "std": float(np.std(rates)), "repeats": repeats,
           "host_cpus_reported": os.cpu_count(),
           "sample": f"x->y stage, {N} sources x {n_targets} stride-sampled targets "
                     f"({N * n_targets:.3g} pairs per run; 1 warm-up + {repeats} runs, "
                     f"median wall {np.median(walls):.2f} s on {cores} processes)"}
    # the C FP64 restatement, single-threaded and with OpenMP, on the same subset
    import oracle
    c_sub = sub[:: max(1, n_targets // 64)]
    t0 = time.perf_counter()
    oracle.transform_full(stage.sign, stage.lam, stage.amp, stage.phase, stage.u, stage.v,
                          targets=c_sub, nthreads=1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.transform_full(stage.sign, stage.lam, stage.amp, stage.phase, stage.u, stage.v,
                          targets=sub, nthreads=cores)
    tm = time.perf_counter() - t0
    res["c_oracle"] = {"single_thread_pairs_per_s": N * c_sub.size / t1,
                       "openmp_pairs_per_s": N * n_targets / tm, "openmp_threads": cores,
                       "sample": f"{c_sub.size} (1 thread) and {n_targets} ({cores} threads) "
                                 "of the same stride-sampled targets, one run each"}
    return res, sub, y[:, 0] * np.exp(1j * y[:, 1])


def time_events(fn, repeats: int, warmup: int = 1):
    """fn() timed with HIP events on torch's current stream (the stream the
    library launches on): `warmup` untimed calls, then `repeats` timed ones.
    Returns median / std / mean / min in ms and the repeat count."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(repeats)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    ms = np.array([a.elapsed_time(b) for a, b in ev])
    return {"median_ms": float(np.median(ms)), "std_ms": float(np.std(ms)),
            "mean_ms": float(np.mean(ms)), "min_ms": float(np.min(ms)), "repeats": repeats}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-targets", type=int, default=4096,
                    help="stride subset of stage-1 targets for the CPU baseline")
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed repeats (after one warm-up) of each side-section measurement")
    ap.add_argument("--sections", default="chain,fp32,mc,adjoint",
                    help="comma list of chain, fp32, mc, adjoint (profiling runs pick one), "
                         "and lattices and distance (not run by default: the FP32-split "
                         "lattice kernel's tile shapes on narrow and square lattices, and its "
                         "speed against FP64 from 0.02 to 0.35 m)")
    ap.add_argument("--hologram-steps", type=int, default=20,
                    help="Adam steps of the hologram section (the full optimisation is 200)")
    ap.add_argument("--lattice-tiles", type=int, choices=(0, 8, 16, 32, 64), default=64,
                    help="tile columns of the FP32-split lattice kernel on uniform lattices "
                         "(64: 16 x 64 tiles where the tile columns form quads, else 16 x 32, "
                         "options grid16, grid32 and grid64; 32: 16 x 32 tiles, grid64 = 0; "
                         "16: 16 x 16 tiles, grid32 = 0; 8: the 8 x 8 kernel, grid16 = 0; "
                         "0: the tiled kernel for arbitrary point sets, grid = 0)")
    ap.add_argument("--source-points", type=int, default=0,
                    help="fp32 and mc sections: this many sources from the reference's plane "
                         "generator (10.24 mm wide; 1105920 = 1024 x 1080, a 1051 x 1052 lattice "
                         "and 268 copies of point 0) instead of the 1024^2 lattice")
    ap.add_argument("--source-jitter", type=float, default=0.0,
                    help="fp32 and mc sections: move every point of the 1024^2 source lattice "
                         "by up to this fraction of the pitch per axis (uniform, seed 5; the "
                         "reference's -r gives 0.4875), which leaves no lattice: the tiled kernel")
    ap.add_argument("--tile", type=int, choices=(8, 16, 32, 64), default=64,
                    help="sources per tile of the tiled FP32-split kernel (the plan's tile)")
    ap.add_argument("--target-z", type=float, default=None,
                    help="distance of the 1024^2 target plane of the fp32 and mc sections "
                         "(default 0.35 m); nearer planes exercise the lattice kernel's "
                         "accuracy bounds")
    ap.add_argument("--mc-passes", type=int, default=0,
                    help="library option mc_passes: source ranges of per-target Monte Carlo, "
                         "one launch each (0 = the library's choice, 1 = one direct launch)")
    ap.add_argument("--no-series", action="store_true",
                    help="library option grid_series = 0: near-field pairs of the lattice "
                         "kernels take the exact path instead of the q-series")
    ap.add_argument("--no-near", action="store_true",
                    help="library option grid_near = 0: waves off the lattice kernels' "
                         "far-field bounds take the q-series and exact paths per pair")
    ap.add_argument("--no-rows", action="store_true",
                    help="library option grid_rows = 0: the lattice kernels' exact pairs gather "
                         "each source's position instead of reading the lattice's row and "
                         "column tables (same bits)")
    ap.add_argument("--target-waves", type=int, default=0,
                    help="library option target_waves (waves a launch aims for when it splits "
                         "sources; 0 = the library default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (HBM counter calibration)")
    a = ap.parse_args()
    sections = set(a.sections.split(","))
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the "
              "superposition would shard by targets, each GPU holding every source, "
              "with no exchange beyond gathering the outputs)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__
    holo = __graft_entry__.load_package()
    holo.set_option("grid16", 1 if a.lattice_tiles in (16, 32, 64) else 0)
    holo.set_option("grid32", 1 if a.lattice_tiles in (32, 64) else 0)
    holo.set_option("grid64", 1 if a.lattice_tiles == 64 else 0)
    holo.set_option("grid", 0 if a.lattice_tiles == 0 else 1)
    if a.target_waves:
        holo.set_option("target_waves", a.target_waves)
    if a.mc_passes:
        holo.set_option("mc_passes", a.mc_passes)
    if a.no_series:
        holo.set_option("grid_series", 0)
    if a.no_near:
        holo.set_option("grid_near", 0)
    if a.no_rows:
        holo.set_option("grid_rows", 0)
    import importlib
    wl = importlib.import_module("holo_amd.workloads")
    w = wl.chain_xyz(0)
    s1, s2 = w.stages

    if a.skip_extras:
        sections &= {"chain"}
    if "chain" not in sections:
        a.skip_cpu = True
    cpu = None
    if not a.skip_cpu and not a.calibrate_copy:
        log(f"[bench] cpu baseline: {a.cpu_targets} targets")
        cpu = cpu_baseline(s1, a.cpu_targets, a.cpu_repeats)
        log(f"[bench] cpu: {cpu[0]}")

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize(dev)
        print(json.dumps({"calibrate_copy_bytes_read": 1 << 30, "bytes_written": 1 << 30,
                          "launches": 3}))
        return
    ctx = holo.Context(0)
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    out = {"metric": METRIC, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic"}

    # ---------------- main workload: FP64 chain x -> y -> z -----------------
    if "chain" in sections:
        out.update(run_chain(a, holo, ctx, s1, s2, T, dev, cpu, w.name))

    # ---------------- FP32-split vs FP64, Monte Carlo, adjoint ---------------
    f64 = None
    R = max(5, a.repeats)
    if sections & {"fp32", "mc"}:
        p3 = (wl.plane_points(0, a.source_points) if a.source_points else wl.plane_1024(0)).stages[0]
        src_desc = (f"{a.source_points} points of the reference's plane generator"
                    if a.source_points else "1024^2")
        if a.target_z is not None:
            p3.v[:, 2] = a.target_z
        if a.source_jitter and not a.source_points:
            gj = np.random.default_rng(5)
            p3.u[:, :2] += (gj.random((len(p3.u), 2)) - 0.5) * (2 * a.source_jitter * 10e-6)
            src_desc = f"1024^2 jittered by up to {a.source_jitter:g} pitch"
        amp3, ph3, u3, v3 = T(p3.amp), T(p3.phase), T(p3.u), T(p3.v)
        plan3 = holo.Plan(ctx, u3, tile=a.tile)
        lattice3 = plan3.lattice() != (0, 0)
        y64 = torch.empty((v3.shape[0], 2), dtype=torch.float64, device=dev)
        y32 = torch.empty_like(y64)
        if "fp32" in sections:
            log(f"[bench] plane {src_desc}: FP64, 1 warm-up + {R} runs")
            t64 = time_events(lambda: plan3.transform(amp3, ph3, v3, p3.sign, p3.lam, holo.FP64,
                                                      out=y64), R)
        else:  # the brute-force field the Monte Carlo estimates are compared with
            plan3.transform(amp3, ph3, v3, p3.sign, p3.lam, holo.FP64, out=y64)
        f64 = torch.view_as_complex(y64).cpu().numpy()
        if "fp32" in sections:
            log(f"[bench] plane {src_desc}: FP32-split, 1 warm-up + {R} runs")
            t32 = time_events(lambda: plan3.transform(amp3, ph3, v3, p3.sign, p3.lam,
                                                      holo.FP32_SPLIT, out=y32), R)
            # the same call with its error report (FP64 at 256 sampled targets)
            t32r = time_events(lambda: plan3.transform(amp3, ph3, v3, p3.sign, p3.lam,
                                                       holo.FP32_SPLIT, out=y32,
                                                       report_error=True), R)
            err_lib = float(ctx.last_rel_l2.item())
            f32 = torch.view_as_complex(y32).cpu().numpy()
            pairs = p3.pairs
            out["fp32_split"] = {
                "workload": f"{src_desc} -> 1024^2 (z 0 -> {p3.v[0, 2]:g} m, 10 / 7.56 um, amp U[0,1), "
                            "phase U[0,2pi)), " + (
                                f"lattice kernel, {min(a.lattice_tiles, 16)}x{a.lattice_tiles} tiles"
                                if a.lattice_tiles and lattice3 else f"tiled kernel, tiles of {a.tile}"),
                "pairs": pairs,
                "fp64_pairs_per_s": pairs / (t64["median_ms"] / 1e3),
                "fp64_ms": t64,
                "fp32_split_pairs_per_s": pairs / (t32["median_ms"] / 1e3),
                "fp32_split_ms": t32,
                "fp32_split_with_error_report_ms": t32r,
                "speedup": t64["median_ms"] / t32["median_ms"],
                "rel_l2_vs_fp64": rel_l2(f32, f64),
                "max_abs_err_rel_to_max": float(np.max(np.abs(f32 - f64)) / np.max(np.abs(f64))),
                "err_rel_l2_reported": err_lib}
            log(f"[bench] fp32-split: {out['fp32_split']}")
        if "mc" in sections:
            mc = {}
            ymc = torch.empty_like(y64)
            for shared in (False, True):
                for K in (256, 1024, 4096):
                    key = ("batched_" if shared else "") + str(K)
                    log(f"[bench] monte carlo {key}: 1 warm-up + {R} runs")

                    def run(report=False):
                        return ctx.transform_mc(amp3, ph3, u3, v3, K, 0, p3.sign, p3.lam,
                                                holo.FP32_SPLIT, shared=shared,
                                                report_error=report)
                    tk = time_events(run, R)
                    ymc = run(True)
                    err_lib = float(ctx.last_rel_l2.item())
                    samples = K * v3.shape[0]
                    mc[key] = {"pairs_per_s": samples / (tk["median_ms"] / 1e3), "ms": tk,
                               "rel_l2_vs_brute_force": rel_l2(ymc.cpu().numpy(), f64),
                               "expected_sqrt_N_over_K": float(np.sqrt(u3.shape[0] / K)),
                               "err_rel_l2_reported": err_lib}
            out["monte_carlo"] = {"workload": f"same {src_desc} -> 1024^2, FP32-split, seed 0; K draws "
                                              "per target, or per group of 1024 targets (batched); "
                                              "rel_l2_vs_brute_force is the estimator's statistical "
                                              "error, expected sqrt(N/K) for this random-phase "
                                              "field; err_rel_l2_reported is the FP32-split "
                                              "arithmetic against the FP64 estimator with the "
                                              "same indices",
                                  "K": mc}
            log(f"[bench] monte carlo: {mc}")

    if "lattices" in sections:
        out["lattices"] = run_lattices(holo, ctx, wl, T, max(5, a.repeats))

    if "distance" in sections:
        out["distance"] = run_distance(holo, ctx, wl, T)

    if "adjoint" in sections:
        hol = importlib.import_module("holo_amd.hologram")
        n = max(20, a.hologram_steps)
        r = hol.optimise(steps=n, warmup=1, log_every=max(1, n // 20) if n > 20 else 0,
                         step_events=True, report_error=True)
        out["adjoint"] = {"workload": "phase-only hologram 1024^2 (7.56 um, z=0) -> 512^2 "
                                      "(10 um, z=0.35 m), FP32-split forward + adjoint, Adam lr 0.05",
                          "ms_per_grad_step": r.step_ms["median_ms"],
                          "step_ms": r.step_ms, "target_ms": 50.0,
                          "pairs_per_step": 2 * 1024 * 1024 * 512 * 512,
                          "steps": n, "losses": r.losses,
                          "forward_err_rel_l2_reported": r.forward_rel_l2}
        log(f"[bench] adjoint: {out['adjoint']}")

    ctx.close()
    print(json.dumps(out), flush=True)


def run_distance(holo, ctx, wl, T, R=3):
    """FP32-split against FP64 on the 1024^2 plane (10 um) -> 1024^2 targets
    (7.56 um) as the target plane moves from 0.02 to 0.35 m, with the lattice
    kernel's near-field GEMM on (default) and off."""
    import torch
    p3 = wl.plane_1024(0).stages[0]
    amp, ph, u = T(p3.amp), T(p3.phase), T(p3.u)
    plan = holo.Plan(ctx, u)
    res = {}
    for z in (0.02, 0.03, 0.04, 0.05, 0.06, 0.075, 0.1, 0.15, 0.2, 0.25, 0.35):
        vz = p3.v.copy()
        vz[:, 2] = z
        v = T(vz)
        y64 = torch.empty((v.shape[0], 2), dtype=torch.float64, device=v.device)
        y32 = torch.empty_like(y64)
        t64 = time_events(lambda: plan.transform(amp, ph, v, p3.sign, p3.lam, holo.FP64,
                                                 out=y64), R)
        f64 = torch.view_as_complex(y64).cpu().numpy()
        row = {"fp64_ms": t64["median_ms"]}
        for near in (1, 0):
            holo.set_option("grid_near", near)
            t32 = time_events(lambda: plan.transform(amp, ph, v, p3.sign, p3.lam,
                                                     holo.FP32_SPLIT, out=y32), R)
            key = "near_gemm" if near else "series_exact"
            row[key] = {"ms": t32["median_ms"], "speedup": t64["median_ms"] / t32["median_ms"],
                        "rel_l2_vs_fp64": rel_l2(torch.view_as_complex(y32).cpu().numpy(), f64)}
        holo.set_option("grid_near", 1)
        res[f"{z:g}"] = row
        log(f"[bench] distance {z:g} m: FP64 {row['fp64_ms']:.1f} ms, FP32-split "
            f"{row['near_gemm']['speedup']:.2f}x (near-field GEMM off: "
            f"{row['series_exact']['speedup']:.2f}x), rel L2 {row['near_gemm']['rel_l2_vs_fp64']:.3g}")
    return res


def run_lattices(holo, ctx, wl, T, R):
    """FP32-split on uniform lattices of 8, 24, 40 and 1024 rows (10 um pitch,
    about 65k or 1M sources) -> 1024^2 targets at 0.35 m, with 16 x 64, 16 x 32,
    16 x 16 and 8 x 8 tiles: narrow lattices pad their rows to 16 (8 rows: half
    the slots are padding)."""
    import torch
    res = {}
    v = T(wl.grid(1024, 1024, 7.56e-6, 0.35))
    for nx, ny in ((8, 8192), (24, 2720), (40, 1632), (1024, 1024)):
        u = wl.grid(nx, ny, 10e-6, 0.0)
        g = np.random.default_rng(0)
        amp, ph, uu = T(g.random(len(u))), T(g.random(len(u)) * 2 * np.pi), T(u)
        plan = holo.Plan(ctx, uu)
        y = torch.empty((v.shape[0], 2), dtype=torch.float64, device=v.device)
        row = {}
        for name, g16, g32, g64 in (("16x64", 1, 1, 1), ("16x32", 1, 1, 0), ("16x16", 1, 0, 0),
                                    ("8x8", 0, 1, 1)):
            holo.set_option("grid16", g16)
            holo.set_option("grid32", g32)
            holo.set_option("grid64", g64)
            t = time_events(lambda: plan.transform(amp, ph, v, +1, 632.8e-9, holo.FP32_SPLIT,
                                                   out=y), R)
            row[name] = {"pairs_per_s": len(u) * v.shape[0] / (t["median_ms"] / 1e3), "ms": t}
        holo.set_option("grid16", 1)
        holo.set_option("grid32", 1)
        holo.set_option("grid64", 1)
        res[f"{nx}x{ny}"] = row
        log(f"[bench] lattice {nx}x{ny}: " + ", ".join(
            f"{k} {r['pairs_per_s']:.3g}" for k, r in row.items()))
    return res


def run_chain(a, holo, ctx, s1, s2, T, dev, cpu, name):
    import torch
    amp_x, ph_x, u_x, v_y, w_z = T(s1.amp), T(s1.phase), T(s1.u), T(s1.v), T(s2.v)
    plan_x = holo.Plan(ctx, u_x)
    plan_y = holo.Plan(ctx, v_y)
    M1, M2 = v_y.shape[0], w_z.shape[0]
    y = torch.empty((M1, 2), dtype=torch.float64, device=dev)
    z = torch.empty((M2, 2), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    launch_ms = []

    def chain_step(record):
        if record:
            ev[0].record(stream)
        plan_x.transform(amp_x, ph_x, v_y, s1.sign, s1.lam, holo.FP64, out=y)
        if record:
            ev[1].record(stream)
        amp_y, ph_y, _ = ctx.normalize_polar(y)
        if record:
            ev[2].record(stream)
        plan_y.transform(amp_y, ph_y, w_z, s2.sign, s2.lam, holo.FP64, out=z)
        if record:
            ev[3].record(stream)

    log(f"[bench] {name}: warmup {a.warmup}")
    for _ in range(a.warmup):
        chain_step(False)
    torch.cuda.synchronize(dev)
    log(f"[bench] timing {a.steps} steps")
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        chain_step(True)
        torch.cuda.synchronize(dev)
        launch_ms += [ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])]
        step_ms.append(ev[0].elapsed_time(ev[3]))
    torch.cuda.synchronize(dev)
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / a.steps
    pairs_step = s1.pairs + s2.pairs
    value = pairs_step / (dt / a.steps)
    log(f"[bench] chain: {ms_per_step:.1f} ms/step, {value:.4g} pairs/s")

    pairs_launch = s1.pairs  # both stages: 512^2 x 1024^2 pairs
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = FLOP_PER_PAIR * pairs_launch / avg_launch_s
    roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                "traffic": read_profile("k_superpose_f64", "traffic_bytes_per_launch"),
                "valu_busy_profiled": read_profile("k_superpose_f64", "valu_busy"),
                "kernel": "k_superpose_f64 (+ partial-sum reduce)",
                "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": pairs_launch,
                "counted_flop_per_pair": COUNTED_FLOP_PER_PAIR,
                "counted_frac": achieved / FLOP_PER_PAIR * COUNTED_FLOP_PER_PAIR / PEAK_FP64,
                "avg_launch_ms": 1e3 * avg_launch_s,
                "median_launch_ms": float(np.median(launch_ms)),
                "std_launch_ms": float(np.std(launch_ms))}
    res = {"value": value, "unit": "pairs/s", "ms_per_step": ms_per_step,
           "config": {"workload": "FP64 brute-force chain x->y->z: 512^2 -> 1024^2 -> 512^2, "
                                  "lambda 632.8 nm, z 0 / 0.35 m, pitch 10 / 7.56 um",
                      "pairs_per_step": pairs_step, "precision": "FP64"},
           "gflops_29": value * FLOP_PER_PAIR / 1e9,
           "step_ms_median": float(np.median(step_ms)), "step_ms_std": float(np.std(step_ms)),
           "roofline": roofline}
    # End to end from host memory (the reference's time_transform includes
    # its host staging): stage 1 through holo_transform_full with host
    # pointers, i.e. H2D of 5 doubles per source and 3 per target, the
    # launch, and D2H of the field.
    log("[bench] stage 1 from host memory")
    ctx.transform_full(s1.amp, s1.phase, s1.u, s1.v, s1.sign, s1.lam,
                       holo.FP64)  # first call sizes the staging buffers
    ths = []
    for _ in range(3):
        t0 = time.perf_counter()
        yh_host = ctx.transform_full(s1.amp, s1.phase, s1.u, s1.v, s1.sign, s1.lam, holo.FP64)
        ths.append(time.perf_counter() - t0)
    th = float(np.median(ths))
    res["host_pointer_stage1"] = {"pairs_per_s": s1.pairs / th, "ms": 1e3 * th,
                                  "ms_std": 1e3 * float(np.std(ths)), "repeats": 3,
                                  "kernel_ms": 1e3 * avg_launch_s,
                                  "rel_l2_vs_device_path": rel_l2(
                                      yh_host, torch.view_as_complex(y).cpu().numpy())}
    if cpu is not None:
        cres, sub, ycpu = cpu
        # y holds the last chain step's stage-1 field (same inputs every step)
        yh = torch.view_as_complex(y).cpu().numpy()
        cres["agreement_rel_l2_vs_gpu"] = rel_l2(ycpu, yh[sub])
        res["cpu_baseline"] = cres
    return res


if __name__ == "__main__":
    main()
