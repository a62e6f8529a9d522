"""pmx_energy_control_noalloc (the force call with its partial sums in the
library's static device scratch and no allocation) against
pmx_energy_control, through the C ABI: two calls in a row, the second on a
smaller system, in FP64 and in the mixed precision, each compared with the
allocating entry point on its own context with the same input.  Both run
the same kernels in the same order; only the order of the PME spread's
atomic additions may differ: in FP64 that moves the last bits (bar 1e-12
relative); the mixed precision rounds the accumulated grid to FP32, where
a different order can flip a rounding (bar 1e-6, inside the path's 1e-4)."""
import numpy as np
import pytest

from piny_amd import abi
from piny_amd.engine import Engine
from tests import cases

pytestmark = pytest.mark.gpu

TOLS = {abi.PMX_PREC_FP64: 1e-12, abi.PMX_PREC_MIXED: 1e-6}
MODES = abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES


def _run(name, precision, noalloc):
    c = cases.get(name)
    s = c.system
    g = Engine(s, c.x, c.y, c.z, precision=precision)
    g.set_positions(c.x, c.y, c.z)
    flags = MODES if s.pi_beads > 1 else 0
    e = g.energy_control_noalloc(flags) if noalloc else g.energy_control(flags)
    f = g.forces()
    g.close()
    return e, f


def _same(e, f, er, fr, label, precision=abi.PMX_PREC_FP64):
    TOL = TOLS[precision]
    for k in abi.ENERGY_KEYS:
        assert abs(e[k] - er[k]) <= TOL * max(abs(er[k]), 1e-6), (label, k, e[k], er[k])
    assert np.allclose(e["pvten"], er["pvten"], rtol=TOL, atol=TOL * np.max(np.abs(er["pvten"])))
    assert e["nver"] == er["nver"]
    assert abs(e["kin_harm"] - er["kin_harm"]) <= TOL * max(abs(er["kin_harm"]), 1e-12)
    assert np.max(np.abs(f - fr)) <= TOL * np.max(np.abs(fr)), label


@pytest.mark.parametrize("precision", [abi.PMX_PREC_FP64, abi.PMX_PREC_MIXED])
def test_noalloc_matches_energy_control(precision):
    # the larger system first, then the smaller one: the scratch keeps the
    # first call's contents when the second runs
    for name in ("tip256_p32", "spc64"):
        print(f"[noalloc {name} precision {precision}]", flush=True)
        e, f = _run(name, precision, True)
        er, fr = _run(name, precision, False)
        _same(e, f, er, fr, name, precision)


def test_noalloc_twice_on_one_context():
    # two calls of the no-allocation entry on one context (the list kept),
    # then the allocating one: all three agree
    c = cases.get("tip64_p4")
    s = c.system
    g = Engine(s, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    e1 = g.energy_control_noalloc(MODES)
    f1 = g.forces()
    e2 = g.energy_control_noalloc(MODES)
    f2 = g.forces()
    er = g.energy_control(MODES)
    fr = g.forces()
    g.close()
    _same(e1, f1, er, fr, "first")
    _same(e2, f2, er, fr, "second")
