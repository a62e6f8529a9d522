"""The workload configurations at full size on the GPU, against the oracle.

* configuration 3 (the benchmark's box): 4096 q-TIP4P/F, 16384 sites,
  L = 49.66 A, P = 32, PME 64^3 of order 6 (its FP32 plan is two radix-8
  passes), the whole energy_control_pimd step (mode forces and springs
  included), FP64 and mixed;
* configuration 2: 256 SPC/Fw at P = 32;
* configuration 5: the 64 uneven SPC/Fw boxes in one batched call, a
  stratified sample of boxes (the most molecules, a 64-bead box, the fewest
  molecules, a classical box, a 32-bead box) against the oracle.

Bars as in test_gpu_parity: FP64 energies, virial and force RMS 1e-9
relative; mixed energies 1e-5 of the largest term, virial 1e-4, force RMS
1e-4 (the north-star bound for the mixed path)."""
import functools

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd import water as w
from piny_amd.engine import Batch, Engine

pytestmark = pytest.mark.gpu

ETOL, FTOL = 1e-9, 1e-9
MIXED_ETOL, MIXED_VTOL, MIXED_FTOL = 1e-5, 1e-4, 1e-4
PIMD_MODES = abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES


@functools.lru_cache(maxsize=4)
def _oracle(name):
    c = {"config3": w.config3, "config2": w.config2}[name]()
    s = c.system
    print(f"[oracle {name}] N={s.natm} P={s.pi_beads}", flush=True)
    e, f, _ = Oracle(s, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, PIMD_MODES)
    return c, e, f


def _check(e, f, eo, fo, mode, label):
    etol, vtol, ftol = (ETOL, ETOL, FTOL) if mode == "fp64" else (MIXED_ETOL, MIXED_VTOL, MIXED_FTOL)
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), abs(eo["vreal"]), abs(eo["vrecip"]), 1e-3)
    if mode == "fp64":
        scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(e[k] - eo[k]) <= etol * max(abs(eo[k]), scale), (label, k, e[k], eo[k])
    rms = float(np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum()))
    dpv = float(np.max(np.abs(e["pvten"] - eo["pvten"])) / np.max(np.abs(eo["pvten"])))
    print(f"  {label} {mode}: force rms rel err {rms:.3e}, vintert rel err "
          f"{abs(e['vintert'] - eo['vintert']) / abs(eo['vintert']):.3e}, pvten {dpv:.3e}", flush=True)
    assert rms < ftol, (label, rms)
    assert dpv <= vtol, (label, dpv)
    return rms


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["config3", "config2"])
def test_full_size_config_against_oracle(name, mode):
    c, eo, fo = _oracle(name)
    s = c.system
    kw = {} if mode == "fp64" else {"precision": abi.PMX_PREC_MIXED}
    g = Engine(s, c.x, c.y, c.z, **kw)
    g.set_positions(c.x, c.y, c.z)
    e = g.energy_control(PIMD_MODES)
    f = g.forces()
    assert e["nver"] == eo["nver"]
    _check(e, f, eo, fo, mode, name)
    assert abs(e["kin_harm"] - eo["kin_harm"]) < 1e-10 * abs(eo["kin_harm"])
    # a second call on the same positions (list kept) gives the same result
    e2 = g.energy_control(PIMD_MODES)
    assert e2["list_rebuilt"] == 0
    tol = 1e-12 if mode == "fp64" else 1e-6
    assert abs(e2["vintert"] - e["vintert"]) <= tol * abs(e["vintert"])
    g.close()


# the most molecules, a 64-bead box, the fewest molecules, a classical box
# and a 32-bead box of the 64 (indices into water.config5())
def _sample(cfgs):
    n = [c.system.nmol for c in cfgs]
    p = [c.system.pi_beads for c in cfgs]
    cost = [c.system.natm * c.system.pi_beads for c in cfgs]
    pick = [int(np.argmax(n)), int(np.argmin(n))]
    for beads in (64, 1, 32):
        cand = [k for k in range(len(cfgs)) if p[k] == beads and k not in pick]
        cand.sort(key=lambda k: cost[k])
        pick.append(cand[len(cand) // 2])  # the median-cost box of that bead count
    return pick


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_config5_batched_against_oracle(mode):
    cfgs = w.config5()
    assert len(cfgs) == 64
    prec = abi.PMX_PREC_FP64 if mode == "fp64" else abi.PMX_PREC_MIXED
    engs = [Engine(c.system, c.x, c.y, c.z, precision=prec) for c in cfgs]
    for e, c in zip(engs, cfgs):
        e.set_positions(c.x, c.y, c.z)
    b = Batch(engs)
    print(f"[config5 {mode}] 64 boxes, {sum(c.system.pi_beads for c in cfgs)} bead evaluations", flush=True)
    out = b.energy_control(abi.PMX_EVAL_MODE_FORCES)
    for k in _sample(cfgs):
        c = cfgs[k]
        s = c.system
        flags = PIMD_MODES if s.pi_beads > 1 else 0
        print(f"  box {k}: {s.nmol} molecules, P={s.pi_beads}", flush=True)
        eo, fo, _ = Oracle(s, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, flags)
        assert out[k]["nver"] == eo["nver"]
        _check(out[k], engs[k].forces(), eo, fo, mode, f"box {k}")
    b.close()
    for e in engs:
        e.close()
