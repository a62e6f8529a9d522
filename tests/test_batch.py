"""Batched force calls (pmx_batch_energy_control) on the GPU: one graph
launch evaluates several uneven boxes, each with its own beads.  Every
replay of the graph is checked against the same boxes evaluated one call at
a time (same kernels; only the order of the bonded kernels' atomic force
additions differs), with positions moved between calls so that lists are
rebuilt and the graph is replayed on new data; one box is also checked
against the oracle."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd import water as w
from piny_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu


def _boxes():
    return w.config5(nbox=6, max_mol=400)


def _set(e, c, dx=None):
    x, y, z = c.x, c.y, c.z
    if dx is not None:
        x, y, z = x + dx[0], y + dx[1], z + dx[2]
    e.set_positions(x, y, z)
    return x, y, z


@pytest.mark.parametrize("prec", [abi.PMX_PREC_FP64, abi.PMX_PREC_MIXED])
def test_batch_matches_single_calls(prec):
    cfgs = _boxes()
    print("[batch] boxes", [(c.system.nmol, c.system.pi_beads) for c in cfgs], "prec", prec, flush=True)
    eb = [Engine(c.system, c.x, c.y, c.z, precision=prec) for c in cfgs]
    es = [Engine(c.system, c.x, c.y, c.z, precision=prec) for c in cfgs]
    b = Batch(eb)
    rng = np.random.default_rng(5)
    # FP64: only the order of atomic additions differs between two contexts.
    # Mixed: the PME spread sums in LDS with atomics in FP64 and rounds the
    # box to FP32, so an FP32 grid point can differ by one unit in the last
    # place between two runs.
    tol = 1e-12 if prec == abi.PMX_PREC_FP64 else 1e-6
    for call in range(4):
        # call 2 moves every site by up to 0.6 Bohr (past half the skin:
        # the lists are rebuilt), the others by a little
        amp = 0.6 if call == 2 else (0.0 if call == 0 else 0.02)
        for c, e1, e2 in zip(cfgs, eb, es):
            n = c.x.size
            d = rng.uniform(-amp, amp, (3, n)) if amp else None
            _set(e1, c, d)
            _set(e2, c, d)
        out = b.energy_control()
        for k, (c, e1, e2) in enumerate(zip(cfgs, eb, es)):
            ref = e2.energy_control()
            got = out[k]
            assert got["list_rebuilt"] == ref["list_rebuilt"]
            for key in abi.ENERGY_KEYS:
                assert abs(got[key] - ref[key]) <= tol * max(abs(ref[key]), 1e-3), (call, k, key)
            dpv = float(np.max(np.abs(got["pvten"] - ref["pvten"])) / np.max(np.abs(ref["pvten"])))
            assert dpv <= tol, (call, k, dpv)
            f1, f2 = e1.forces(), e2.forces()
            df = float(np.max(np.abs(f1 - f2)) / np.max(np.abs(f2)))
            assert df <= tol, (call, k, df)
        print(f"  call {call}: rebuilt {[o['list_rebuilt'] for o in out]}", flush=True)


def test_batch_box_matches_oracle():
    cfgs = _boxes()
    eb = [Engine(c.system, c.x, c.y, c.z) for c in cfgs]
    for e, c in zip(eb, cfgs):
        e.set_positions(c.x, c.y, c.z)
    out = Batch(eb).energy_control()
    for k in (0, 1):
        c = cfgs[k]
        s = c.system
        print(f"[batch box {k}] N={s.natm} P={s.pi_beads}", flush=True)
        flags = abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0
        eo, fo, _ = Oracle(s, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, flags)
        scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
        for key in abi.ENERGY_KEYS:
            assert abs(out[k][key] - eo[key]) <= 1e-9 * max(abs(eo[key]), scale), (k, key)
        fg = eb[k].forces()
        assert np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum()) < 1e-9


def test_autotune_keeps_results():
    # the tuned launch shapes change only the order of sums
    cfgs = w.config5(nbox=3, max_mol=600)
    for c in cfgs:
        s = c.system
        print(f"[autotune] N={s.natm} P={s.pi_beads}", flush=True)
        a = Engine(s, c.x, c.y, c.z)
        b = Engine(s, c.x, c.y, c.z)
        for e in (a, b):
            e.set_positions(c.x, c.y, c.z)
        ref = b.energy_control()
        ch = a.autotune(reps=1)
        print("  choice", ch, flush=True)
        assert ch["pair_waves"] in (1024, 2048, 4096, 8192, 16384)
        got = a.energy_control()
        for key in abi.ENERGY_KEYS:
            assert abs(got[key] - ref[key]) <= 1e-10 * max(abs(ref[key]), 1e-3), key
        df = float(np.max(np.abs(a.forces() - b.forces())) / np.max(np.abs(b.forces())))
        assert df < 1e-10


def test_batch_autotune_keeps_results():
    cfgs = _boxes()
    eb = [Engine(c.system, c.x, c.y, c.z) for c in cfgs]
    es = [Engine(c.system, c.x, c.y, c.z) for c in cfgs]
    for e1, e2, c in zip(eb, es, cfgs):
        e1.set_positions(c.x, c.y, c.z)
        e2.set_positions(c.x, c.y, c.z)
    b = Batch(eb)
    ch = b.autotune(reps=1)
    print("[batch autotune] choice", ch, flush=True)
    out = b.energy_control()
    for k, e2 in enumerate(es):
        ref = e2.energy_control()
        for key in abi.ENERGY_KEYS:
            assert abs(out[k][key] - ref[key]) <= 1e-10 * max(abs(ref[key]), 1e-3), (k, key)
