"""Benchmark: bead-force evaluations per second on one MI355X.

Workload (the headline configuration): 4096 q-TIP4P/F molecules (16384
sites incl. the massless M site), cubic L = 49.66 A, P = 32 beads at 300 K,
r_c = 9 A, Ewald alpha = 0.35/A, PME 64^3 of order 6, FP64.  One step is one
energy_control_pimd call over all 32 beads (32 bead-force evaluations:
forces with ghost forces redistributed, energies and the 3x3 virial); the
Verlet list is checked every step and rebuilt by the skin criterion, as the
reference does.  Positions are resident on the device when timing starts.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bead-force evals/s, 4096 q-TIP4P/F P=32: mixed vs FP64 GPU; force RMS err"
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)
FLOP_PER_PAIR = 100.0       # nominal per-pair operation count of the pair kernel
DT_FS = 0.25


def cpu_baseline(cfg, seconds=12.0):
    """The oracle (FP64 C restatement, one core) on a bounded sample: bead 1
    of the same box evaluated as a classical configuration, repeated."""
    from oracle.oracle import Oracle
    from piny_amd import water as w
    s = cfg.system
    s1 = w.make_system(s.model, s.nmol, s.box, 1, s.settings)
    n = s.natm
    x, y, z = cfg.x[:n].copy(), cfg.y[:n].copy(), cfg.z[:n].copy()
    o = Oracle(s1, x, y, z)
    o.energy_control(x, y, z)                 # builds the Verlet list
    t0 = time.perf_counter()
    k = 0
    while True:
        o.energy_control(x, y, z)
        k += 1
        el = time.perf_counter() - t0
        if el > seconds or k >= 50:
            break
    return {"value": k / el, "unit": "bead-force evals/s", "cores": 1, "kind": "port",
            "sample": f"{k} force calls on bead 1 of the workload box (16384 sites), list reused, "
                      f"{el:.1f} s on one host core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--beads", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(beads and boxes would shard across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from piny_amd import abi, water as w
    from piny_amd.engine import Engine

    cfg = w.config3(P=args.beads)
    s = cfg.system
    P = s.pi_beads
    eng = Engine(s, cfg.x, cfg.y, cfg.z)
    eng.set_positions(cfg.x, cfg.y, cfg.z)
    flags = abi.PMX_EVAL_PIMD
    for _ in range(args.warmup):
        e = eng.energy_control(flags)
    stage = {"list": 0.0, "real": 0.0, "recip": 0.0, "intra": 0.0, "total": 0.0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = eng.energy_control(flags)
        tm = eng.timings()
        for k in stage:
            stage[k] += tm[k]
    t1 = time.perf_counter()
    el = t1 - t0
    for k in stage:
        stage[k] /= args.steps
    npairs = eng.pair_count()
    value = args.steps * P / el
    ms_step = el / args.steps * 1e3
    pair_ms = stage["real"]
    achieved = FLOP_PER_PAIR * npairs / (pair_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bead-force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "q-TIP4P/F 4096 molecules, 16384 sites, L=49.66 A, P=32, 300 K, "
                               "r_c=9 A, skin 1 A, PME 64^3 order 6, FP64",
                   "molecules": s.nmol, "sites": s.natm, "beads": P, "precision": "fp64"},
        "ns_per_day": (args.steps / el) * DT_FS * 86400 * 1e-6,
        "stage_ms": stage,
        "pairs_in_cutoff_all_beads": npairs,
        "verlet_pairs": e["nver"],
        "energies_per_bead_hartree": {"vintert": e["vintert"], "vintrat": e["vintrat"]},
        "mixed": None,
        "roofline": {"kernel": "k_pair_f64 (real-space LJ + erfc pair forces, all beads)",
                     "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "note": "FP64 vector-ALU bound; MI355X's FP64 vector peak equals its FP64 matrix "
                             "peak (78.6 TFLOP/s); 100 flop/pair nominal"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
