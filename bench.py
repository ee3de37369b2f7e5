#!/usr/bin/env python3
"""Energy+gradient throughput of the NUwU structure-factor path on one MI355X.

Workload (the reference's 3D NUwU configuration on one GPU): N = 2^20 points
uniform in the unit box (= [0, 2pi)^3 NUFFT coordinates), 129^3 integer mode
grid (K' = 64), stealthy constraint S0(k) = 0 with weight 1 on 0 < |k| <= 24,
radii 1.  A step is one energy+gradient evaluation of all N points with the
coordinates resident in HBM.

  value         E+grad evaluations / s of the FP32 + kernel-derivative-adjoint
                path (eps 1e-6, sigma 2)
  fp64_plain    the same for the FP64 path with d separate type-2 transforms
                (eps 1e-9), and the relative energy / gradient / S(k) error
                of FP32 against it (S(k) on the constrained modes)
  fp64_adjoint  the FP64 path on the FP32 path's band grid with one inverse
                transform (eps 1e-9): splits the speed-up into grid and
                transform count on one side, precision on the other
  roofline      dominant kernel over its average launch duration (HIP events
                on the evaluation stream): the matrix-core spreader against
                the FP32 MFMA peak (algorithmic FLOPs: one multiply-add per
                support cell and point), other kernels against 8 TB/s
                (algorithmic bytes); both views are on the line
  cpu_baseline  CPU FP64 restatement of the reference path (oracle/nufft_cpu,
                eps 1e-9, all host threads), a bounded sample of evaluations

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3


def stealthy_weights(kp, kmax):
    kv = np.arange(-kp, kp + 1)
    kz, ky, kx = np.meshgrid(kv, kv, kv, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2).ravel()
    return ((kk > 0) & (kk <= kmax)).astype(np.float64)


PROF_PERIOD = 8  # headline run: the dominant kernel is timed on one evaluation in 8


def timed_evals(pot, x, g, steps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = None
    for _ in range(steps):
        e = pot.energy_gradient_device(x, g)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (the HBM byte-counter calibration kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the N = 1 throughput); "
              "multi-GPU sharding of the path is design-only.", file=sys.stderr)
        sys.exit(2)

    import torch

    from fresco_amd import _lib
    from fresco_amd.potentials import NUwU

    N, d, kp, kmax = args.n, 3, 64, 24
    V = stealthy_weights(kp, kmax)
    Sk = np.zeros_like(V)
    radii = np.ones(N)
    L = np.ones(d)
    rng = np.random.default_rng(0)
    p = rng.uniform(0.0, 1.0, N * d)
    x = torch.tensor(p, dtype=torch.float64, device="cuda")
    g32 = torch.empty_like(x)
    g64 = torch.empty_like(x)

    pot32 = NUwU(radii, kp, Sk, V, L, 1e-6, precision="fp32")
    pot64 = NUwU(radii, kp, Sk, V, L, 1e-9, precision="fp64")
    info32 = pot32.plan_info()
    info64 = pot64.plan_info()

    for _ in range(args.warmup):
        pot32.energy_gradient_device(x, g32)
        pot64.energy_gradient_device(x, g64)

    # stage breakdown: every stage between HIP events (not the headline run:
    # ~16 events per evaluation cost ~10% of its time)
    _lib.prof_reset()
    _lib.prof_enable(True)
    timed_evals(pot32, x, g32, args.steps, torch)
    _lib.prof_enable(False)
    stages32 = _lib.prof_read()
    per_stage = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages32.items()}
    nf = info32["nf"]
    G = nf[0] * nf[1] * nf[2]
    alg_bytes = {
        # sorted records (3 offsets + weight, fp32) + packed cell in; real fp32 grid out
        "spread": N * (16 + 4) + G * 4,
        # sorted records + cell + original index in, real fp32 grid in, fp64 gradient out
        "interp": N * (16 + 4 + 4) + G * 4 + N * d * 8,
    }
    dom = max(alg_bytes, key=lambda k: per_stage.get(k, 0.0))

    # headline: FP32 + adjoint; the only events on the evaluation stream are
    # the two around the dominant kernel (its live launch duration), on one
    # evaluation in PROF_PERIOD (each event pair costs ~6 us of the
    # evaluation's 0.3 ms)
    _lib.prof_reset()
    _lib.prof_sample_period(PROF_PERIOD)
    _lib.prof_enable_stages([dom])
    dt32, e32 = timed_evals(pot32, x, g32, args.steps, torch)
    _lib.prof_enable(False)
    _lib.prof_sample_period(1)
    live = _lib.prof_read()[dom]
    ev32 = args.steps / dt32

    k64 = args.fp64_steps or max(3, args.steps // 5)
    dt64, e64 = timed_evals(pot64, x, g64, k64, torch)
    ev64 = k64 / dt64
    _lib.prof_reset()
    _lib.prof_enable(True)
    timed_evals(pot64, x, g64, max(2, k64 // 2), torch)
    _lib.prof_enable(False)
    stages64 = _lib.prof_read()

    # FP64 with the one-transform adjoint on the same band grid as the FP32
    # path: separates the grid / transform-count savings from precision
    pota = NUwU(radii, kp, Sk, V, L, 1e-9, precision="fp64-adjoint")
    ga = torch.empty_like(x)
    for _ in range(2):
        pota.energy_gradient_device(x, ga)
    dta, _ = timed_evals(pota, x, ga, k64, torch)
    eva = k64 / dta
    infoa = pota.plan_info()

    e32 = pot32.energy_gradient_device(x, g32)
    e64 = pot64.energy_gradient_device(x, g64)
    ea = pota.energy_gradient_device(x, ga)
    info32 = pot32.plan_info()  # after the first evaluation: the autotuned spreader
    rel_e = abs(e32 - e64) / abs(e64)
    rel_g = float(torch.linalg.norm(g32 - g64) / torch.linalg.norm(g64))
    # S(k) = |rho_k|^2 / N on the constrained (active) modes
    act = V > 0
    s64 = (np.abs(pot64.get_rho()) ** 2 / N)[act]
    s32 = (np.abs(pot32.get_rho()) ** 2 / N)[act]
    sa = (np.abs(pota.get_rho()) ** 2 / N)[act]
    rel_s = float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))
    rel_a = {"energy": abs(ea - e64) / abs(e64),
             "gradient": float(torch.linalg.norm(ga - g64) / torch.linalg.norm(g64)),
             "Sk": float(np.linalg.norm(sa - s64) / np.linalg.norm(s64))}
    del pota

    if args.calib_copy:
        # known traffic: 256 MiB read + 256 MiB written per launch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # dominant single kernel: algorithmic bytes (and, for the spreader,
    # FLOPs) per launch over its live average launch duration in the
    # headline run
    dom_ms = live[0] / live[1]
    achieved = alg_bytes[dom] / (dom_ms * 1e-3) / 1e9
    kernels = {"halom": "k_spread3_halo_mfma2", "halor": "k_spread3_halo_ring", "halo4": "k_spread3_halo_mw",
               "halo": "k_spread3_halo", "tiles": "k_spread3_reg"}
    kname = kernels.get(info32["spreader"], "k_spread") if dom == "spread" else "k_interp3_g8"
    w = info32["w"]
    alg_flops = 2.0 * N * w ** d  # one multiply-add per support cell and point
    hbm_view = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": alg_bytes[dom]}
    if dom == "spread" and info32["spreader"] in ("halom", "halor"):
        tflops = alg_flops / (dom_ms * 1e-3) / 1e12
        roofline = {
            "kernel": kname,
            "bound": "mfma",
            "achieved": round(tflops, 3),
            "peak": MFMA_FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(tflops / MFMA_FP32_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(dom, info32),
            "avg_launch_ms": round(dom_ms, 4),
            "alg_flops_per_launch": alg_flops,
            "hbm_view": hbm_view,
        }
    else:
        roofline = {
            "kernel": kname,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom, info32),
            "avg_launch_ms": round(dom_ms, 4),
            "alg_bytes_per_launch": alg_bytes[dom],
        }
    # whole evaluation: real-grid algorithmic bytes 3 N d s_r + 6 G s_g + 2 M_active s_c
    m_active = int(np.count_nonzero(V))
    b_eval = 3 * N * d * 8 + 6 * G * 4 + 2 * m_active * 8
    eval_roofline = {
        "alg_bytes_per_eval": b_eval,
        "achieved_GBs": round(b_eval * ev32 / 1e9, 1),
        "frac": round(b_eval * ev32 / 1e9 / HBM_PEAK_GBS, 4),
        "note": "3*N*d*8 (fp64 coords/grad) + 6*G*4 (real fp32 fine grid) + 2*M_active*8",
    }

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(radii, kp, Sk, V, L, p, N)

    line = {
        "metric": "E+grad evals/s, N=2^20 3D NUwU: FP32+adjoint vs FP64 plain @ rel err 1e-5",
        "value": round(ev32, 2),
        "unit": "E+grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt32 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform random points, seed 0)",
        "config": {
            "workload": "NUwU 3D energy+gradient, N=2^20 points, 129^3 modes (K'=64), "
                        "stealthy S0=0 V=1 on 0<|k|<=24, radii 1, FP32 adjoint eps=1e-6 sigma=2",
            "N": N,
            "modes": (2 * kp + 1) ** 3,
            "active_modes": m_active,
            "nf": nf,
            "w": info32["w"],
            "tile": info32["tile"],
            "spreader": spreader_label(info32),
        },
        "points_per_s": round(ev32 * N, 1),
        "fp64_plain": {
            "evals_per_s": round(ev64, 2),
            "ms_per_step": round(1e3 * dt64 / k64, 4),
            "steps": k64,
            "eps": 1e-9,
            "nf": info64["nf"],
            "w": info64["w"],
        },
        "fp64_adjoint": {
            "evals_per_s": round(eva, 2),
            "ms_per_step": round(1e3 * dta / k64, 4),
            "steps": k64,
            "eps": 1e-9,
            "nf": infoa["nf"],
            "w": infoa["w"],
            "rel_err_vs_fp64_plain": rel_a,
        },
        "speedup_fp32_adjoint_vs_fp64_plain": round(ev32 / ev64, 3),
        # the speed-up in two factors: band grid + one inverse transform (FP64
        # plain -> FP64 adjoint), then FP32 and its kernels (FP64 adjoint -> FP32)
        "speedup_split": {"grid_and_transforms": round(eva / ev64, 3), "precision": round(ev32 / eva, 3)},
        "rel_err_vs_fp64": {"energy": rel_e, "gradient": rel_g, "Sk": rel_s},
        "stages_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "stages_fp64_ms": {k: round(v[0] / v[1], 4) if v[1] else 0.0 for k, v in stages64.items()},
        "roofline": roofline,
        "eval_roofline": eval_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def spreader_label(info):
    """The spreader in use, its halo tile parts per bin and where the pick came from."""
    name, parts = info["spreader"], info["spread_parts"]
    if name == "halom":
        split = "%d slice(s) x 2 x windows, one halo tile per slice" % (parts // 2)
    elif name == "halor":
        split = "%d slice(s) x 2 x windows, one tile per window" % (parts // 2)
    else:
        split = "%d slice(s)" % parts
    return "%s:%d, %s (%s)" % (name, parts, split, info["spread_source"])


def measured_traffic(stage, info):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)
    for the spreader configuration in use, or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "nuwu3d_n2p20_stealthy24_hbm_traffic.md")
    cfg = "%s:%d" % (info.get("spreader"), info.get("spread_parts"))
    key = {"spread": "spread_traffic_bytes_per_launch[%s]:" % cfg, "interp": "interp_traffic_bytes_per_launch:"}[stage]
    try:
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith(key):
                    return int(float(line[len(key):].strip()))
    except OSError:
        pass
    return None


def cpu_baseline(radii, kp, Sk, V, L, p, N):
    """CPU FP64 restatement of the reference path on the host cores."""
    from oracle import nufft_cpu

    pot = nufft_cpu.NUwU(radii, kp, Sk, V, L, eps=1e-9)
    times = []
    budget = 25.0
    t_start = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_start) < budget:
        t0 = time.perf_counter()
        pot.get_energy_gradient(p)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {
        "value": round(1.0 / med, 4),
        "unit": "E+grad evals/s",
        "cores": nufft_cpu.threads(),
        "kind": "port",
        "sample": "%d full E+grad evaluations of the same N=2^20 workload (median), FP64, eps=1e-9, "
                  "sigma=2, OpenMP spread/interp + scipy.fft; CPU %s" % (len(times), cpu_model()),
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
