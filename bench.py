"""Benchmark: bead-force evaluations per second on one MI355X.

Workload (the headline configuration): 4096 q-TIP4P/F molecules (16384
sites incl. the massless M site), cubic L = 49.66 A, P = 32 beads at 300 K,
r_c = 9 A, skin 1 A, Ewald alpha = 0.35/A, PME 64^3 of order 6.  One step is
one complete energy_control_pimd call over all 32 beads (32 bead-force
evaluations: forces with ghost forces redistributed, energies and the 3x3
virial, then the forces transformed to normal-mode coordinates with the
ring-polymer spring forces added, energy_control_pimd.c:198-214); the Verlet
list is checked every step and rebuilt by the skin criterion, as the
reference does.  Positions are resident on the device when timing starts.
`rebuild_every_20` times the same call with the list rebuilt on one step
in 20 (a Verlet list lives 10-50 steps in water MD).

Three paths are timed on the same box:
  fp64        PINY's arithmetic: FP64 with cubic spline-table interpolation
              (the drop-in path, parity with the reference to ~1e-12);
  fp64_direct FP64 with the tabulated pair functions evaluated directly;
  mixed       FP32 pair terms and PME grid/FFT with FP32/FP64 split sums
              (separations' image part, energies, virial, bonded terms,
              springs in FP64).
`value` is the mixed path's throughput, reported with the fp64 path's and
the mixed path's force RMS error relative to fp64.  `fp64_per_bead` times
the plain FP64 form beside them: each bead as its own classical force call,
one after another.  `respa` times the force calls of one outer RESPA step
with the step counts of the reference's q-TIP4P/F example (intramolecular,
short-range and full calls) against as many plain calls.

Prints one JSON line.
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
PEAK = {"f64": 78.6, "f32": 157.3}   # MI355X vector TFLOP/s (FP64 vector = FP64 matrix peak)
# HBM bytes per launch of the pair kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (separate runs) scaled by a 512 MiB copy probe (FETCH x 2.00,
# WRITE x 1.00); profiles/kernels_q-tip4p-f_4096_p32_{mixed,fp64}.md
PAIR_TRAFFIC = {"mixed": (58888 * 2.0 + 21779) * 1024, "fp64": (129379 * 2.0 + 22829) * 1024}
FLOP_PER_PAIR = 100.0                 # nominal algorithmic flops per pair term inside the cutoff
DT_FS = 0.25


def cpu_baseline(cfg, seconds=12.0):
    """The reference on one host core: oracle/_ref/piny_ref (the reference's
    own C and Fortran, built from its sources by oracle/ref.mk) runs its
    energy_control on bead 1 of the workload box as a classical
    configuration, once to build the Verlet list and then `repeat` timed
    calls.  Without that build, the oracle (the FP64 C restatement) is timed
    the same way."""
    from piny_amd import water as w
    s = cfg.system
    s1 = w.make_system(s.model, s.nmol, s.box, 1, s.settings)
    n = s.natm
    x, y, z = cfg.x[:n].copy(), cfg.y[:n].copy(), cfg.z[:n].copy()
    exe = os.path.join(ROOT, "oracle", "_ref", "piny_ref")
    if os.path.exists(exe):
        import subprocess
        import tempfile
        from tests.golden.make_ref_vectors import write_inputs
        repeat = 40
        with tempfile.TemporaryDirectory() as d:
            write_inputs(d, w.Config(system=s1, x=x, y=y, z=z))
            r = subprocess.run([exe, "water.input", "pos.bin", "out.bin", str(repeat)], cwd=d,
                               capture_output=True, text=True, timeout=600)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("REPEAT")]
        if r.returncode == 0 and line:
            k, el = int(line[0].split()[1]), float(line[0].split()[2])
            return {"value": k / el, "unit": "bead-force evals/s", "cores": 1, "kind": "reference",
                    "sample": f"{k} energy_control calls of the reference (built from its sources) on bead 1 of "
                              f"the workload box (16384 sites) as a classical configuration, list reused, "
                              f"{el:.1f} s on one host core"}
    from oracle.oracle import Oracle
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
            "sample": f"{k} force calls of the oracle on bead 1 of the workload box (16384 sites), list reused, "
                      f"{el:.1f} s on one host core"}


def cpu_baseline_all_cores(cfg, procs=None, repeat=10):
    """The reference on all host cores the way PINYEmpirical runs replicas
    (one energy_control per replica per host thread, piny.pyx:426-436): one
    oracle/_ref/piny_ref process per core, each on its own bead of the
    workload box as a classical configuration, all at once; the throughput
    is the sum over processes.  None without the reference build."""
    import subprocess
    import tempfile
    from piny_amd import water as w
    exe = os.path.join(ROOT, "oracle", "_ref", "piny_ref")
    if not os.path.exists(exe):
        return None
    from tests.golden.make_ref_vectors import write_inputs
    s = cfg.system
    n = s.natm
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    procs = procs or max(1, min(16, ncpu, s.pi_beads))
    s1 = w.make_system(s.model, s.nmol, s.box, 1, s.settings)
    with tempfile.TemporaryDirectory() as root:
        runs = []
        for p in range(procs):
            d = os.path.join(root, str(p))
            os.makedirs(d)
            x, y, z = (v[p * n:(p + 1) * n].copy() for v in (cfg.x, cfg.y, cfg.z))
            write_inputs(d, w.Config(system=s1, x=x, y=y, z=z))
            runs.append(subprocess.Popen([exe, "water.input", "pos.bin", "out.bin", str(repeat)], cwd=d,
                                         stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True))
        total, secs = 0.0, []
        for r in runs:
            out, _ = r.communicate(timeout=900)
            line = [ln for ln in out.splitlines() if ln.startswith("REPEAT")]
            if r.returncode != 0 or not line:
                return None
            k, el = int(line[0].split()[1]), float(line[0].split()[2])
            total += k / el
            secs.append(el)
    return {"value": total, "unit": "bead-force evals/s", "cores": procs, "kind": "reference",
            "sample": f"{procs} concurrent processes of the reference (built from its sources), each {repeat} "
                      f"energy_control calls on its own bead of the workload box (16384 sites) as a classical "
                      f"configuration, list reused; {min(secs):.1f}-{max(secs):.1f} s each"}


FLAGS = None  # energy_control_pimd with mode forces and springs (set in main)


def run_path(cfg, steps, warmup, rebuild_every=20, tune=True, **engine_kw):
    from piny_amd import abi
    from piny_amd.engine import Engine
    s = cfg.system
    eng = Engine(s, cfg.x, cfg.y, cfg.z, **engine_kw)
    eng.set_positions(cfg.x, cfg.y, cfg.z)
    flags = FLAGS
    tuned = None
    if tune:  # the project's autotuner picks the launch shapes first (untimed)
        t0 = time.perf_counter()
        tuned = eng.autotune(reps=10, flags=flags)
        tuned["seconds"] = time.perf_counter() - t0
    for _ in range(warmup):
        e = eng.energy_control(flags)
    stage = {"list": 0.0, "real": 0.0, "recip": 0.0, "intra": 0.0, "total": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        e = eng.energy_control(flags)
        tm = eng.timings()
        for k in stage:
            stage[k] += tm[k]
    el = time.perf_counter() - t0
    for k in stage:
        stage[k] /= steps
    out = {"value": steps * s.pi_beads / el, "ms_per_step": el / steps * 1e3, "stage_ms": stage,
           "pairs": eng.pair_count(), "energies": e, "forces": eng.forces(), "autotune": tuned}
    # the list rebuilt on one call in rebuild_every (PMX_EVAL_FORCE_LIST),
    # and the list stage of a rebuilding call
    n = max(rebuild_every, steps)
    build_ms = []
    t0 = time.perf_counter()
    for k in range(n):
        rb = k % rebuild_every == 0
        eng.energy_control(flags | (abi.PMX_EVAL_FORCE_LIST if rb else 0))
        if rb:
            build_ms.append(eng.timings()["list"])
    el = time.perf_counter() - t0
    out["rebuild_every_20"] = {"value": n * s.pi_beads / el, "ms_per_step": el / n * 1e3,
                               "list_stage_ms_when_rebuilt": float(np.mean(build_ms))}
    eng.close()
    return out


def run_respa(cfg, reps, **engine_kw):
    """The force calls of one outer step of int_NVT_pimd under the RESPA
    steps of the reference's q-TIP4P/F example (respa_steps_lrf 4,
    respa_steps_intra 2, respa_steps_torsion 1, respa_steps_pimd 5,
    examples/q-TIP4P-F/build.py:54,66-68) on the benchmark box, the inter
    split at res_dist: int_NVT_pimd.c:143-215 calls energy_control_pimd once
    per intramolecular step (ir_pimd = nres_pimd), 4 x 2 = 8 calls per outer
    step: 4 intramolecular only, 3 short-range, 1 full (which runs an
    unweighted pass first for the virial estimator).  Reported: ms per call
    of each kind, ms per outer step, and the same 8 calls as full plain
    calls for comparison."""
    import dataclasses
    from piny_amd import abi
    from piny_amd.engine import Engine
    st = dataclasses.replace(cfg.system.settings, respa_lrf=4, respa_intra=2, respa_torsion=1, respa_pimd=5)
    s = dataclasses.replace(cfg.system, settings=st)
    eng = Engine(s, cfg.x, cfg.y, cfg.z, **engine_kw)
    eng.set_positions(cfg.x, cfg.y, cfg.z)
    base = abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES
    kinds = {"intra_only": base | abi.PMX_EVAL_NO_FULL_INTER | abi.PMX_EVAL_NO_FULL_INTRA,
             "short_range": base | abi.PMX_EVAL_NO_FULL_INTER | abi.PMX_EVAL_RES_INTER,
             "full": base | abi.PMX_EVAL_RES_INTER}
    seq = []
    for ir_ter in range(1, 5):
        for ir_tra in range(1, 3):
            seq.append("full" if (ir_ter == 4 and ir_tra == 2) else ("short_range" if ir_tra == 2 else "intra_only"))

    def timed(fn, n):
        fn()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        return (time.perf_counter() - t0) / n * 1e3

    ms_plain = timed(lambda: eng.energy_control(base), reps)
    eng.set_respa(abi.respa_params(s))
    ms = {k: timed(lambda f=f: eng.energy_control(f), reps) for k, f in kinds.items()}
    ms_outer = timed(lambda: [eng.energy_control(kinds[k]) for k in seq], reps)
    eng.close()
    return {"steps": {"respa_steps_lrf": 4, "respa_steps_intra": 2, "respa_steps_torsion": 1,
                      "respa_steps_pimd": 5},
            "res_dist_A": st.res_dist, "calls_per_outer_step": seq, "ms_per_call": ms,
            "ms_per_outer_step": ms_outer, "ms_plain_full_call": ms_plain,
            "speedup_vs_plain_calls": len(seq) * ms_plain / ms_outer}


def run_batch(steps, warmup, nbox=64):
    """The uneven-box workload: nbox SPC/Fw boxes (molecule counts
    log-uniform in [64, 4096], 1 to 64 beads).  One step evaluates every box:
    sequentially, one force call per box (FP64), or as one batched graph
    launch (FP64 and mixed)."""
    from piny_amd import abi, water as w
    from piny_amd.engine import Batch, Engine
    cfgs = w.config5(nbox=nbox)
    evals = sum(c.system.pi_beads for c in cfgs)
    res = {}
    tuned = {}
    for name, prec, batched, tune in (("sequential_fp64", abi.PMX_PREC_FP64, False, False),
                                      ("batched_fp64", abi.PMX_PREC_FP64, True, False),
                                      ("batched_mixed", abi.PMX_PREC_MIXED, True, False),
                                      ("batched_mixed_tuned", abi.PMX_PREC_MIXED, True, True)):
        engs = [Engine(c.system, c.x, c.y, c.z, precision=prec) for c in cfgs]
        for e, c in zip(engs, cfgs):
            e.set_positions(c.x, c.y, c.z)
        b = Batch(engs) if batched else None
        if tune:  # the batch autotuner picks the launch shapes first
            t0 = time.perf_counter()
            tuned = b.autotune(reps=5)
            tuned["seconds"] = time.perf_counter() - t0

        def step():
            if b is not None:
                b.energy_control()
            else:
                for e in engs:
                    e.energy_control()
        for _ in range(warmup):
            step()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        el = time.perf_counter() - t0
        res[name] = {"value": steps * evals / el, "ms_per_step": el / steps * 1e3}
        if b is not None:
            b.close()
        for e in engs:
            e.close()
    head = max(("batched_mixed", "batched_mixed_tuned"), key=lambda k: res[k]["value"])
    return {"metric": "bead-force evals/s, 64 uneven SPC/Fw boxes (64-4096 molecules, 1-64 beads), one batched "
                      "graph launch per step", "value": res[head]["value"], "unit": "bead-force evals/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": res[head]["ms_per_step"],
            "autotune": tuned,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"{nbox} SPC/Fw boxes at 0.997 g/cm^3, seeded", "boxes": nbox,
                       "bead_evals_per_step": evals, "sites_x_beads": int(sum(c.x.size for c in cfgs))},
            "paths": res, "path": head,
            "batched_vs_sequential_fp64": res["batched_fp64"]["value"] / res["sequential_fp64"]["value"]}


def run_per_bead(cfg, steps, warmup):
    """The plain FP64 form: every bead of the workload box evaluated by its own
    classical force call, one after another (one context per bead, each with
    its own list), as PINYEmpirical evaluates replicas
    (python/piny/piny.pyx:426-436).  One step = one call per bead."""
    from piny_amd import water as w
    from piny_amd.engine import Engine
    s = cfg.system
    P, n = s.pi_beads, s.natm
    s1 = w.make_system(s.model, s.nmol, s.box, 1, s.settings)
    engs = []
    for b in range(P):
        x, y, z = (v[b * n:(b + 1) * n].copy() for v in (cfg.x, cfg.y, cfg.z))
        e = Engine(s1, x, y, z)
        e.set_positions(x, y, z)
        engs.append(e)
    for _ in range(warmup):
        for e in engs:
            e.energy_control(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in engs:
            e.energy_control(0)
    el = time.perf_counter() - t0
    for e in engs:
        e.close()
    return {"value": steps * P / el, "ms_per_step": el / steps * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--beads", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shapes)")
    ap.add_argument("--paths", default="fp64,fp64_direct,mixed")
    ap.add_argument("--no-per-bead", action="store_true",
                    help="skip the one-bead-per-call FP64 baseline (kernel profiles of the batched paths)")
    ap.add_argument("--no-respa", action="store_true", help="skip the RESPA call timings")
    ap.add_argument("--batch", action="store_true",
                    help="the uneven-box workload (64 boxes, batched graph launch vs sequential calls)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a device copy of 512 MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(beads and boxes would shard across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copies": 3}))
        return

    if args.batch:
        print(json.dumps(run_batch(args.steps, args.warmup)))
        return

    from piny_amd import abi, water as w

    global FLAGS
    FLAGS = abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES
    cfg = w.config3(P=args.beads)
    s = cfg.system
    P = s.pi_beads
    kw = {"fp64": {}, "fp64_direct": {"pair_eval": abi.PMX_PAIR_DIRECT},
          "mixed": {"precision": abi.PMX_PREC_MIXED}}
    paths = [p for p in args.paths.split(",") if p]
    res = {p: run_path(cfg, args.steps, args.warmup, tune=not args.no_tune, **kw[p]) for p in paths}
    head_kw = kw["mixed" if "mixed" in res else paths[-1]]
    respa = run_respa(cfg, max(5, args.steps // 4), **head_kw) if not args.no_respa else None
    per_bead = (run_per_bead(cfg, max(2, args.steps // 4), 1)
                if "fp64" in paths and not args.no_per_bead else None)
    head = res["mixed"] if "mixed" in res else res[paths[-1]]
    head_name = "mixed" if "mixed" in res else paths[-1]
    ref = res.get("fp64")
    fo = ref["forces"] if ref else None

    def rms(f):
        return float(np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())) if fo is not None else None

    # the roofline kernel alone: the head path again on a context whose k space
    # runs on the same stream (PMX_OVERLAP=0), so the pair kernel's HIP-event
    # duration is not stretched by PME kernels beside it
    os.environ["PMX_OVERLAP"] = "0"
    try:
        serial = run_path(cfg, max(10, args.steps // 2), 2, tune=False, **kw[head_name])
    finally:
        del os.environ["PMX_OVERLAP"]
    pair_ms = serial["stage_ms"]["real"]
    npairs = head["pairs"]
    prec = "f32" if head_name == "mixed" else "f64"
    achieved = FLOP_PER_PAIR * npairs / (pair_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "bead-force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_name == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "q-TIP4P/F 4096 molecules, 16384 sites, L=49.66 A, P=32, 300 K, "
                               "r_c=9 A, skin 1 A, PME 64^3 order 6",
                   "molecules": s.nmol, "sites": s.natm, "beads": P, "path": head_name},
        "ns_per_day": (1e3 / head["ms_per_step"]) * DT_FS * 86400 * 1e-6,
        "paths": {p: {"value": r["value"], "ms_per_step": r["ms_per_step"], "stage_ms": r["stage_ms"],
                      "rebuild_every_20": r["rebuild_every_20"], "autotune": r["autotune"],
                      "force_rms_err_vs_fp64": rms(r["forces"]),
                      "vintert_rel_err_vs_fp64": (abs(r["energies"]["vintert"] - ref["energies"]["vintert"])
                                                  / abs(ref["energies"]["vintert"])) if ref else None}
                  for p, r in res.items()},
        "speedup_vs_fp64": head["value"] / ref["value"] if ref else None,
        "fp64_per_bead": per_bead,
        "respa": respa,
        "speedup_vs_fp64_per_bead": head["value"] / per_bead["value"] if per_bead else None,
        "force_rms_err": rms(head["forces"]),
        "pairs_in_cutoff_all_beads": npairs,
        "verlet_pairs": head["energies"]["nver"],
        "energies_per_bead_hartree": {"vintert": head["energies"]["vintert"],
                                      "vintrat": head["energies"]["vintrat"]},
        "roofline": {"kernel": f"pair kernel of the {head_name} path (real-space LJ + erfc, all beads)",
                     "bound": "valu", "achieved": achieved, "peak": PEAK[prec], "unit": "TFLOP/s",
                     "frac": achieved / PEAK[prec], "traffic": PAIR_TRAFFIC.get(head_name),
                     "note": "vector-ALU bound (no MFMA work: nothing on this path is a dense contraction); "
                             f"peak is MI355X's published vector peak for {prec}; achieved counts the nominal "
                             f"{FLOP_PER_PAIR:.0f} algorithmic flop per pair term inside the cutoff (minimum image "
                             "15, r^2 5, LJ 11, erfc 30, switch 10, force 9, energy and virial 14, rounded), "
                             "not issued instructions; duration from HIP events around the launch on the "
                             "context's stream, on a context with the k space on that stream too "
                             "(PMX_OVERLAP=0); the code is built without packed FP32 instructions, which "
                             "halves the FP32 issue rate the published peak assumes",
                     "kernel_ms": pair_ms},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg)
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
