"""The device PIMD NVT integrator (pmx_pimd_nvt) against the reference's own
int_NVT_pimd trajectories (tests/golden/ref_pimd_*.npz, see test_pimd_nvt)
and against the oracle, and the mixed precision's conserved-quantity drift
against the FP64 path's with beads.

Bars: FP64 trajectories (bead positions, mode velocities, every chain's
state) 1e-9 relative to the largest value after 20 steps; the mixed path's
extended-energy fluctuation over 50 fs within 2x of the FP64 path's (the
north-star bound)."""
import os
import tempfile

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi, pinyio
from piny_amd.engine import Engine
from tests import cases
from tests.test_pimd_nvt import PIMD, compare_trajectory, load_pimd

pytestmark = pytest.mark.gpu


def _device_run(name, nsteps, v0, precision=abi.PMX_PREC_FP64, params=None, g=None):
    c = cases.get(name)
    s = c.system
    g = g or Engine(s, c.x, c.y, c.z, precision=precision)
    g.set_positions(c.x, c.y, c.z)
    v0 = np.asarray(v0).reshape(s.pi_beads * s.natm, 3)
    g.set_velocities(*(np.ascontiguousarray(v0[:, d]) for d in range(3)))
    e, ec = g.pimd_nvt(nsteps, params)
    return g, e, ec


@pytest.mark.parametrize("name", PIMD)
def test_device_pimd_nvt_matches_reference(name):
    d = load_pimd(name)
    print(f"[pimd nvt {name}] {int(d['nsteps'])} steps", flush=True)
    g, e, ec = _device_run(name, int(d["nsteps"]), d["v0"])
    pos, vel, st = g.positions(), g.velocities(), g.nhc_state(4)
    print(f"  pos {np.abs(pos - d['pos']).max():.2e} vel {np.abs(vel - d['vel']).max():.2e}", flush=True)
    compare_trajectory(name, d, pos, vel, st, e, 1e-9)
    # the device's final state as a restart file (pinyio), against the one
    # the reference's own writer made of its final state
    s = cases.get(name).system
    cls, bead = pinyio.nhc_state_to_restart(s, st, 4)
    mine = pinyio.restart_for(s, pos, vel, nhc_class_v=cls, nhc_bead_v=bead, dt=float(d["dt"]))
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "device.restart")
        pinyio.write_restart(p, mine)
        back = pinyio.read_restart(p)
        q = os.path.join(t, "reference.restart")
        open(q, "w").write(str(d["restart"]))
        ref = pinyio.read_restart(q)
    assert back.labels == ref.labels and back.nfree == ref.nfree and back.ensemble == ref.ensemble
    for a, b in ((back.pos, ref.pos), (back.vel, ref.vel), (back.nhc_class_v, ref.nhc_class_v),
                 (back.nhc_bead_v, ref.nhc_bead_v)):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))


def test_device_pimd_nvt_matches_oracle_over_calls():
    # two calls of 10 steps (state kept in the context between them) against
    # one oracle run of 20 steps, staging
    name = "tip64_p4_stag"
    d = load_pimd(name)
    c = cases.get(name)
    g, _, _ = _device_run(name, 10, d["v0"])
    e, ec = g.pimd_nvt(10)
    o = Oracle(c.system, c.x, c.y, c.z)
    pos, vel, st, eo, eco = o.pimd_nvt(20, c.x, c.y, c.z, d["v0"])
    assert np.max(np.abs(g.positions() - pos)) < 1e-9 * np.max(np.abs(pos))
    assert np.max(np.abs(g.velocities() - vel)) < 1e-9 * np.max(np.abs(vel))
    assert np.allclose(ec, eco, rtol=1e-9, atol=1e-12), (ec, eco)


def test_mixed_pimd_drift_within_2x_of_fp64():
    # the extended-system energy over 200 steps of 0.25 fs at P = 4 (the
    # configuration-2 box), every 20 steps, in FP64 and mixed precision
    d = load_pimd("cfg2_p4")
    spread = {}
    for mode, prec in (("fp64", abi.PMX_PREC_FP64), ("mixed", abi.PMX_PREC_MIXED)):
        g, e, ec = _device_run("cfg2_p4", 20, d["v0"], precision=prec)
        tot = [ec[3]]
        for _ in range(9):
            e, ec = g.pimd_nvt(20)
            tot.append(ec[3])
        tot = np.array(tot)
        spread[mode] = float(np.max(np.abs(tot - tot[0])))
        print(f"  {mode}: max |E(t) - E(0)| = {spread[mode]:.3e} Ha over 200 PIMD steps", flush=True)
    assert spread["mixed"] <= 2.0 * spread["fp64"] + 1e-6
