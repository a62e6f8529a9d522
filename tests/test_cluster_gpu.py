"""The molecule-cluster form of the mixed-precision pair kernel (switched on
by the tuning override PMX_CLUSTER=1 at context creation; the per-site form
is the default) against the per-site form and against the oracle: water boxes, classical
and with beads, cubic and triclinic, that qualify (one molecule template,
one cutoff, cutoff plus skin and spread well inside half the cell), and one
small box that does not (the per-site kernel then runs in both).  Both forms evaluate the same pairs in FP32 with FP64
sums; they differ in the order of sums and in the tables (whole switched
functions against erfc table plus direct LJ and switch), so the bar between
them is 2e-6 relative (energies) and 2e-6 (force RMS); against the oracle the
mixed bars of the other tests hold."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd.engine import Engine
from tests import cases

pytestmark = pytest.mark.gpu

MODES = abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES


def _engine(c, cluster):
    old = os.environ.get("PMX_CLUSTER")
    os.environ["PMX_CLUSTER"] = "1" if cluster else "0"
    try:
        g = Engine(c.system, c.x, c.y, c.z, precision=abi.PMX_PREC_MIXED)
    finally:
        if old is None:
            os.environ.pop("PMX_CLUSTER", None)
        else:
            os.environ["PMX_CLUSTER"] = old
    g.set_positions(c.x, c.y, c.z)
    return g


@pytest.mark.parametrize("name", ["spc512", "tip512_p4", "tip512_p4_tric", "tip64_p4"])
def test_cluster_matches_site_kernel_and_oracle(name):
    c = cases.get(name)
    s = c.system
    flags = MODES if s.pi_beads > 1 else 0
    print(f"[cluster {name}] N={s.natm} P={s.pi_beads}", flush=True)
    gc, gs = _engine(c, True), _engine(c, False)
    ec, es = gc.energy_control(flags), gs.energy_control(flags)
    fc, fs = gc.forces(), gs.forces()
    eo, fo, _ = Oracle(s, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, flags)
    scale = max(abs(eo["vreal"]), abs(eo["vintert"]), 1e-3)
    for k in ("vreal", "vvdw", "vcoul", "vintert"):
        assert abs(ec[k] - es[k]) <= 2e-6 * scale, (k, ec[k], es[k])
        assert abs(ec[k] - eo[k]) <= 1e-5 * scale, (k, ec[k], eo[k])
    assert np.allclose(ec["pvten"], es["pvten"], atol=2e-6 * np.max(np.abs(eo["pvten"])))
    rcs = float(np.sqrt(((fc - fs) ** 2).sum() / (fs ** 2).sum()))
    rco = float(np.sqrt(((fc - fo) ** 2).sum() / (fo ** 2).sum()))
    print(f"  cluster vs site force rms {rcs:.2e}, vs oracle {rco:.2e}", flush=True)
    assert rcs < 2e-6 and rco < 1e-4
    # the two forms differ in their rounding, so results that agree to the
    # last bit mean the per-site kernel ran in both (the small box)
    assert (ec["vcoul"] != es["vcoul"]) == (name != "tip64_p4"), (ec["vcoul"], es["vcoul"])
    # the pair count of the counting pass against the per-site kernel's count
    nc, ns_ = gc.pair_count(), gs.pair_count()
    assert abs(nc - ns_) <= 1e-4 * ns_ + 2, (nc, ns_)
    gc.close()
    gs.close()


def test_cluster_follows_list_rebuilds():
    # positions moved past the skin: the rows are rebuilt with the list
    c = cases.get("tip512_p4")
    s = c.system
    gc, gs = _engine(c, True), _engine(c, False)
    o = Oracle(s, c.x, c.y, c.z)
    rng = np.random.default_rng(5)
    x, y, z = c.x.copy(), c.y.copy(), c.z.copy()
    for step in range(3):
        x += rng.normal(0, 0.3, x.shape)
        for g in (gc, gs):
            g.set_positions(x, y, z)
        ec, es = gc.energy_control(MODES), gs.energy_control(MODES)
        eo, fo, _ = o.energy_control(x.copy(), y.copy(), z.copy(), MODES)
        assert ec["list_rebuilt"] == es["list_rebuilt"] == eo["list_rebuilt"]
        assert abs(ec["vcoul"] - es["vcoul"]) <= 2e-6 * abs(es["vcoul"])
        fc, fs = gc.forces(), gs.forces()
        assert np.sqrt(((fc - fs) ** 2).sum() / (fs ** 2).sum()) < 2e-6
    gc.close()
    gs.close()
