#!/usr/bin/env python3
"""Benchmark of the NUFFT hot path on one MI355X.

Workload (one "step"): a 2D type-1 NUFFT in single precision (complex64
strengths, float32 points), M = 10,000,000 points uniform on [-pi, pi)^2,
1024 x 1024 output modes, eps = 1e-6, iflag = +1: setpts (fold, bin, sort)
followed by execute (spread, FFT, deconvolve) on a warm plan.  Inputs are
synthetic, generated on the device before timing.

Prints ONE JSON line:
  value         NU points per second, end to end (M / time per step)
  roofline      the spread kernel (the dominant one): algorithmic bytes per
                launch (2*4*M coordinates + 8*M strengths + 8*F fine grid)
                over its average launch time from HIP events recorded on the
                plan's stream inside the timed region, against 8 TB/s
  cpu_baseline  the FINUFFT-method CPU restatement (oracle/, C + OpenMP +
                scipy.fft) on a bounded sample of the same workload
Also: per-stage times, and the same step on clustered points (Gaussian,
sigma = 0.1 about the origin).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jax-finufft_amd"))
sys.path.insert(0, ROOT)

METRIC = "NU pts/s & HBM-roofline % (2D t1 fp32 eps1e-6); bwd/fwd time ratio vs plain grad"
HBM_PEAK_GBS = 8000.0
M_POINTS = 10_000_000
MODES = (1024, 1024)
EPS = 1e-6
# plan stages timed with HIP events on the plan's stream (units / spread / merge
# are the strip spreader's piece table, spreading kernel and tile merge)
STAGES = ("setpts", "zero", "units", "spread", "merge", "fft", "deconv")


def make_points(M, dist, seed, device):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if dist == "uniform":
        pts = [(torch.rand(M, generator=g, device=device, dtype=torch.float32) * 2 - 1) * math.pi
               for _ in range(2)]
    else:
        pts = []
        for _ in range(2):
            v = torch.randn(M, generator=g, device=device, dtype=torch.float32) * 0.1
            pts.append(torch.remainder(v + math.pi, 2 * math.pi) - math.pi)
    c = torch.complex(torch.randn(M, generator=g, device=device),
                      torch.randn(M, generator=g, device=device))
    return pts, c


def run_steps(plan, pts, c, fk, steps, warmup):
    import torch

    for _ in range(warmup):
        plan.setpts(*pts)
        plan.execute(c, fk)
    for st in STAGES:
        plan.timing(st, reset=True)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.setpts(*pts)
        plan.execute(c, fk)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    stages = {}
    for st in STAGES:
        tot, n = plan.timing(st)
        stages[st] = tot / n if n else None
    return ms, stages


def sampled_accuracy(pts, c, fk, n_sample=48, seed=11):
    """Relative l2 error of the timed output on a random sample of modes,
    against the exact sum f[k1, k2] = sum_j c_j exp(+i (k1 x_j + k2 y_j)) over
    all M points, evaluated in float64 on the device."""
    import torch

    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    n1, n2 = fk.shape
    i1 = torch.randint(0, n1, (n_sample,), generator=g)
    i2 = torch.randint(0, n2, (n_sample,), generator=g)
    x = pts[0].double()
    y = pts[1].double()
    cd = c.to(torch.complex128)
    exact = []
    for a, b in zip(i1.tolist(), i2.tolist()):
        k1, k2 = a - n1 // 2, b - n2 // 2
        ph = k1 * x + k2 * y
        exact.append(torch.sum(cd * torch.polar(torch.ones_like(ph), ph)))
    exact = torch.stack(exact)
    got = fk[i1.to(fk.device), i2.to(fk.device)].to(torch.complex128)
    return float(torch.linalg.norm(got - exact) / torch.linalg.norm(exact))


def traffic_from_profiles():
    """HBM bytes per spread launch from committed rocprofv3 --pmc summaries."""
    path = os.path.join(ROOT, "profiles", "spread_2d_t1_f32_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(sample_M, ns, reps=3):
    """FINUFFT-method CPU restatement on a sample of the workload (same grid,
    same kernel width); NU points per second."""
    import numpy as np

    from oracle import finufft_cpu

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(2)
    pts = [rng.uniform(-np.pi, np.pi, sample_M) for _ in range(2)]
    c = rng.normal(size=sample_M) + 1j * rng.normal(size=sample_M)
    plan = finufft_cpu.Plan(MODES, EPS, 1, workers=threads, ns=ns)
    plan.type1(pts, c)  # warm-up (FFT plan, page faults)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.type1(pts, c)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return {
        "value": sample_M / t,
        "unit": "NU pts/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"M={sample_M:.0e} uniform points (the full workload when 1e7), full 1024x1024 modes (2048^2 fine "
                   f"grid), width {ns}; FINUFFT-method restatement (C/OpenMP fp64 spread + "
                   f"scipy.fft + deconvolve), median of {reps} after 1 warm-up"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=10_000_000)
    ap.add_argument("--no-clustered", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this backend has no multi-GPU path, so --gpus must "
              "be 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from jax_finufft_amd import Plan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    pts, c = make_points(M_POINTS, "uniform", 2, dev)
    fk = torch.empty(MODES, dtype=torch.complex64, device=dev)
    plan = Plan(1, MODES, iflag=1, ntransf=1, eps=EPS, dtype=torch.complex64, device=dev,
                debug=2)
    info = plan.info()
    ms, stages = run_steps(plan, pts, c, fk, args.steps, args.warmup)
    value = M_POINTS / (ms * 1e-3)
    accuracy = sampled_accuracy(pts, c, fk)

    F = info["nf"][0] * info["nf"][1]
    spread_bytes = 2 * 4 * M_POINTS + 8 * M_POINTS + 8 * F
    spread_ms = stages["spread"]
    achieved = spread_bytes / (spread_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "kernel": ("spread_strip8_kernel<float,%d>" if info["ns"] <= 8
                   else "spread_strip_kernel<float,%d>") % info["ns"],
        "algorithmic_bytes": spread_bytes,
        "avg_launch_ms": round(spread_ms, 4),
    }

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "NU pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2D type-1 NUFFT, complex64/float32, M=1e7 uniform points on "
                        "[-pi,pi)^2, 1024x1024 modes, eps=1e-6, iflag=+1; step = setpts + "
                        "execute on a warm plan",
            "M": M_POINTS,
            "modes": list(MODES),
            "eps": EPS,
            "kernel_width": info["ns"],
            "fine_grid": list(info["nf"]),
        },
        "roofline": roofline,
        "stages_ms": {k: (round(v, 4) if v is not None else None) for k, v in stages.items()},
        "accuracy": {"rel_l2_sampled_modes": accuracy, "modes_sampled": 48,
                     "against": "float64 direct sum over all M points, on the device"},
    }
    del plan

    if not args.no_clustered:
        cpts, cc = make_points(M_POINTS, "clustered", 3, dev)
        cplan = Plan(1, MODES, iflag=1, ntransf=1, eps=EPS, dtype=torch.complex64, device=dev,
                     debug=2)
        cms, cst = run_steps(cplan, cpts, cc, fk, max(2, args.steps // 2), 1)
        out["clustered"] = {
            "value": round(M_POINTS / (cms * 1e-3), 1),
            "ms_per_step": round(cms, 4),
            "stages_ms": {k: (round(v, 4) if v is not None else None) for k, v in cst.items()},
            "points": "Gaussian sigma=0.1 about the origin, wrapped",
        }
        del cplan

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, info["ns"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
