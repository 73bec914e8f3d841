"""Liquid-substep throughput on one MI355X.

Workload: the large dam break -- 256^3 staggered grid (dx = 1/256), liquid
block [0,0.5]x[0,0.6]x[0,1] at 8 particles per cell (~40.3 M particles),
seeded Gaussian velocity (sigma = 1 m/s), dt = 1/1920, MIC(0)-PCG to 1e-6
relative residual.  One step = one full substep (sort, P2G, grid update,
assembly, pressure solve, pressure application, G2P + advection) with every
input resident in HBM.

The same workload runs twice from the same seeds: in the mixed-precision
form (the reported value) and in the plain FP64 form (reported beside it,
with the ratio and the positions' difference after the same steps).  A third
run solves the pressure with FP64 AMG-PCG instead, the reference's as-shipped
solver, for comparison (fp64_amg; --no-amg skips it).

A fourth leg (wet_cloth_128; --no-wet-cloth skips it) times the 128^3
wet-cloth scene in the mixed form: the 128^3 liquid column (0.4 x 0.6 x 1 m,
~4.0 M particles) starting at y = 0.35 over a resting 257 x 257-vertex
porous cloth at y = 0.3, dt = 1/960 -- the liquid path with the mixture
terms (volume fraction, drag, pore pressure).

Prints ONE JSON line (see the keys below).  --gpus N > 1 exits with a message:
the library runs the substep on a single GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "substeps/s at 256³ grid, 40M particles: mixed-prec vs FP64 GPU at 1e-6 rel res"
TIMING_STRIDE = 32


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="grid cells per axis")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--maxit", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-amg", action="store_true",
                    help="skip the FP64 AMG-PCG run of the workload (the reference's "
                         "as-shipped pressure solver, reported under fp64_amg)")
    ap.add_argument("--no-wet-cloth", action="store_true",
                    help="skip the 128^3 wet-cloth leg (reported under wet_cloth_128)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    return ap.parse_args()


HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# algorithmic words per DOF of one launch (8-byte words in FP64, 4-byte in
# the mixed form's FP32 solve): a sweep reads invdiag, three factor entries
# and its input and writes its output; the SpMV reads four coefficients and
# x and writes y
SWEEP_WORDS = 6
SPMV_WORDS = 6


def copy_calibration():
    import torch
    n = 1 << 27  # doubles: 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}))


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of a kernel measured with rocprofv3 counters and
    committed as data next to the summaries (None if absent or measured on
    another workload)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(kernel)
    if not e or e.get("workload") != workload:
        return None
    return e.get("bytes_per_launch")


# particle-state words a transfer reads and writes per particle: P2G reads
# x, v and C (3 + 3 + 9); G2P reads x and writes x, v and C (3 + 3 + 3 + 9).
# Grid words per active node (FP64 grid in both forms): P2G writes mass,
# momentum and volume of three faces (3 x 3), G2P reads three face
# velocities (3).
P2G_WORDS, G2P_WORDS = 15, 18
P2G_GRID_WORDS, G2P_GRID_WORDS = 9, 3


def transfers(stats: dict, word: int, active_nodes: int) -> dict:
    """Achieved HBM rate of the P2G and G2P kernels from the library's
    HIP-event times (every launch of the timed substeps), against the
    particle state they move plus the grid they write or read."""
    out = {}
    for name, words, gwords in (("p2g", P2G_WORDS, P2G_GRID_WORDS),
                                ("g2p", G2P_WORDS, G2P_GRID_WORDS)):
        ms, nl, units = stats[name]
        if nl == 0:
            continue
        pb = units / nl * words * word
        b = pb + active_nodes * gwords * 8
        gbs = b / (ms / nl * 1e-3) / 1e9
        out[name] = {"avg_ms": ms / nl, "launches": nl, "particles": units // nl,
                     "active_nodes": active_nodes, "particle_bytes_per_launch": pb,
                     "bytes_per_launch": b, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def roofline(stats: dict, word: int, workload: str) -> dict:
    """HBM roofline of the dominant kernel (the slower MIC(0) sweep) from the
    HIP-event times the library recorded on its own stream; word is the
    solve's storage word (8 FP64, 4 FP32)."""
    per = {}
    for name, words in (("sweep_fwd", SWEEP_WORDS), ("sweep_bwd", SWEEP_WORDS),
                        ("spmv", SPMV_WORDS)):
        ms, nl, units = stats[name]
        if nl == 0:
            continue
        avg_s = ms / nl * 1e-3
        bytes_per_launch = units / nl * words * word
        per[name] = {"avg_ms": ms / nl, "launches": nl, "dofs": units // nl,
                     "bytes_per_launch": bytes_per_launch,
                     "GBs": bytes_per_launch / avg_s / 1e9,
                     "frac": bytes_per_launch / avg_s / 1e9 / HBM_PEAK_GBS}
    if not per:
        return None
    dom = max(("sweep_fwd", "sweep_bwd"), key=lambda k: per.get(k, {}).get("avg_ms", -1))
    d = per[dom]
    # the kernels' names as rocprofv3 reports them
    vt = "double" if word == 8 else "float"
    kname = {"sweep_fwd": f"k_sweep_sub<0, {vt}>", "sweep_bwd": f"k_sweep_sub<1, {vt}>"}[dom]
    return {
        "bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": d["GBs"] / HBM_PEAK_GBS, "traffic": measured_traffic(kname, workload),
        "kernel": kname, "avg_ms": d["avg_ms"], "algorithmic_bytes": d["bytes_per_launch"],
        "kernels": per,
    }


def cpu_baseline(n_particles_gpu: int, steps: int) -> dict:
    """The FP64 oracle (a plain-C port of the reference path) on one host core,
    on a bounded sample of the workload: the same block, velocity field and
    tolerance at 128^3 (dt 1/960), whose substep takes ~18 s on one core."""
    from oracle import oracle as O
    O.build()
    n = 128
    g, x, v = O.dam_break(n, block_hi=(0.5, 0.6, 1.0), seed=4, sigma=1.0)
    s = O.OracleScene((n, n, n), 1.0 / n, x, v)
    dt = 1.0 / 960
    t0 = time.perf_counter()
    iters = []
    for _ in range(steps):
        s.substep(dt, 1e-6, 2000)
        iters.append(s.last_solve[1])
    t = (time.perf_counter() - t0) / steps
    np_sample = x.shape[1]
    scaled = (1.0 / t) * np_sample / max(n_particles_gpu, 1)
    return {
        "value": scaled,
        "unit": "substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{steps} substep(s) of the workload's dam break at 128^3 ({np_sample} "
                   f"particles, dt 1/960, {t:.1f} s each, PCG iterations {iters}, "
                   f"{np_sample / t:.3e} particle-updates/s) on 1 core; scaled to the "
                   f"{n_particles_gpu}-particle workload by particle count (the 256^3 solve "
                   "takes more iterations, so this overstates the CPU rate)"),
    }


def run(W, precision, n, dt, a, solver=0):
    import numpy as np
    """Seed the workload in one precision, step it warmup + steps times and
    time the last steps (the library's stream is synchronized by every step:
    the solver reads its residual back each iteration).  solver 1 solves the
    pressure with AMG-PCG instead of MIC(0)-PCG."""
    scene = W.Scene((n, n, n), 1.0 / n, precision=precision)
    scene.set_option("solver", solver)
    npart = scene.seed_block((0.0, 0.0, 0.0), (0.5, 0.6, 1.0), seed=4, sigma=1.0)
    iters, passes = [], []
    for _ in range(a.warmup):
        ok, it, res = scene.step(dt, a.tol, a.maxit)
    phase_tot = dict.fromkeys(W.PHASES, 0.0)
    # HIP events around the sweep and SpMV launches of every 32nd PCG
    # iteration (each event is a marker packet that idles the GPU ~5 us;
    # timing every launch costs ~3 % of the substep); also clears the counters
    scene.set_option("kernel_timing", TIMING_STRIDE)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ok, it, res = scene.step(dt, a.tol, a.maxit)
        if not ok:
            raise RuntimeError(f"pressure solve did not converge ({it} iterations)")
        iters.append(it)
        passes.append(scene.refinement_passes())
        for k, v in scene.phase_times().items():
            phase_tot[k] += v
    wall = time.perf_counter() - t0
    active_nodes = int(scene.field("act").astype(np.int64).sum()) * 64
    return scene, {
        "active_nodes": active_nodes,
        "particles": npart,
        "value": a.steps / wall,
        "ms_per_step": 1e3 * wall / a.steps,
        "pcg_iterations": iters,
        "refinement_passes": passes,
        "phase_ms_per_step": {k: v / a.steps for k, v in phase_tot.items()},
        "kernel_stats": scene.kernel_stats(),
    }


def wet_cloth(W, a, n=128, dt=1.0 / 960):
    """The 128^3 wet-cloth scene in the mixed form: warmup + steps substeps,
    the last steps timed."""
    from libwetcloth_amd import scenes
    scene = W.Scene((n, n, n), 1.0 / n, precision=W.WC_MIXED)
    npart = scene.seed_block((0.0, 0.35, 0.0), (0.4, 0.95, 1.0), seed=2)
    cloth = scenes.config2_cloth()
    scene.set_elasto(*scenes.elasto_args(cloth), prm=W.mix_params())
    for _ in range(a.warmup):
        scene.substep(dt, a.tol, a.maxit)
    iters = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ok, it, _ = scene.substep(dt, a.tol, a.maxit)
        if not ok:
            raise RuntimeError(f"wet cloth: pressure solve did not converge ({it} iterations)")
        iters.append(it)
    wall = time.perf_counter() - t0
    scene.close()
    return {"workload": (f"wet cloth {n}^3: block [0,0.4]x[0.35,0.95]x[0,1] at 8 ppc over a "
                         "resting 257x257-vertex cloth at y = 0.3 (1 m square, 2 mm, "
                         f"0.2 kg/m^2, porosity U[0.3,0.7]), dt {dt:.6g}, mixed, "
                         f"MIC(0)-PCG tol {a.tol:g}"),
            "particles": npart, "cloth_vertices": int(cloth["x"].shape[1]),
            "value": a.steps / wall, "unit": "substeps/s", "ms_per_step": 1e3 * wall / a.steps,
            "pcg_iterations": iters}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "the substep runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import libwetcloth_amd as W

    n = a.n
    dt = 1.0 / 1920 if n >= 256 else (1.0 / 960 if n >= 128 else 1.0 / 480)
    workload = (f"dam break {n}^3, block [0,0.5]x[0,0.6]x[0,1], 8 ppc, "
                f"sigma 1 m/s, dt {dt:.6g}, MIC(0)-PCG tol {a.tol:g}")
    mixed, m = run(W, W.WC_MIXED, n, dt, a)
    fp64, f = run(W, W.WC_FP64, n, dt, a)
    # both forms took warmup + steps substeps from the same seeds
    xm, _, _ = mixed.particles()
    xf, _, _ = fp64.particles()
    pos_dx = float(np.abs(xm - xf).max() * n)
    del xm, xf
    mixed.close()
    fp64.close()
    npart = m["particles"]
    sps = m["value"]
    out = {
        "metric": METRIC,
        "value": sps,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload, "grid": n, "particles": npart,
                   "precision": "mixed (FP32 particles/transfers/PCG vectors; FP64 grid, "
                                "factor, inner products, refinement residual)"},
        "particle_updates_per_s": npart * sps,
        "pcg_iterations": m["pcg_iterations"],
        "refinement_passes": m["refinement_passes"],
        "phase_ms_per_step": m["phase_ms_per_step"],
        "fp64": {k: f[k] for k in ("value", "ms_per_step", "pcg_iterations",
                                   "phase_ms_per_step")},
        "mixed_vs_fp64": sps / f["value"],
        "mixed_vs_fp64_position_diff_dx": pos_dx,
    }
    out["roofline"] = roofline(m["kernel_stats"], 4, workload)
    out["fp64"]["roofline"] = roofline(f["kernel_stats"], 8, workload)
    out["transfers"] = transfers(m["kernel_stats"], 4, m["active_nodes"])
    out["fp64"]["transfers"] = transfers(f["kernel_stats"], 8, f["active_nodes"])
    if not a.no_amg:
        amg, g = run(W, W.WC_FP64, n, dt, a, solver=1)
        amg.close()
        out["fp64_amg"] = {k: g[k] for k in ("value", "ms_per_step", "pcg_iterations",
                                             "phase_ms_per_step")}
    if not a.no_wet_cloth:
        out["wet_cloth_128"] = wet_cloth(W, a)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(npart, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
