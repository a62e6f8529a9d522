"""The oracle against the reference itself.

tests/golden/ref_<case>.npz hold the outputs of the reference's own
energy_control (P = 1) and energy_control_pimd (P > 1), built from the
reference's sources and driven through its own parse() on the case's input
files (tests/golden/make_ref_vectors.py, oracle/ref.mk, oracle/ref_harness.c).
For P > 1 the reference returns normal-mode forces including the springs,
which is this library's PMX_EVAL_MODE_FORCES form.

Right after one force call the reference's stat_avg.vrecip holds the
k-space sum plus the self, background and excluded-pair terms, and the
real-space energy is vintert - vrecip - vlong; the other per-term fields are
filled by its output routines and are not compared."""
import glob
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(n for n in (os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))
                  if not n.startswith(("md_", "pimd_")))
TOL = 1e-12


def load(name):
    d = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
    ref = {str(k): float(v) for k, v in zip(d["energy_keys"], d["energies"])}
    ref["vreal_implied"] = ref["vintert"] - ref["vrecip"] - ref["vlong"]
    return d, ref


def compare(name, e, f, d, ref, etol, ftol, vtol=None, term_scale=False):
    """Energies within etol of max(|value|, scale): scale is the larger of
    vintert and vintrat, or with term_scale the largest of the terms whose sum
    vintert is (real space, k space incl. self/background/ecor), the bar of a
    reduced-precision path whose errors scale with the terms, not with the
    cancelled total."""
    vtol = etol if vtol is None else vtol
    scale = max(abs(ref["vintert"]), abs(ref["vintrat"]), 1e-3)
    if term_scale:
        scale = max(scale, abs(ref["vreal_implied"]), abs(ref["vrecip"]))
    got = dict(e)
    got["vrecip_total"] = e["vrecip"] + e["vself"] + e["vbgr"] + e["vecor"]
    got["vreal_implied"] = e["vreal"]
    pairs = [("vintert", "vintert"), ("vintrat", "vintrat"), ("vlong", "vlong"), ("vbond", "vbond"),
             ("vbend", "vbend"), ("vrecip", "vrecip_total"), ("vreal_implied", "vreal_implied"),
             ("pi_ke_vir", "pi_ke_vir")]
    if int(d["forces"].shape[0]) > 1:
        pairs += [("pi_ke_prim", "pi_ke_prim"), ("kin_harm", "kin_harm")]
    for rk, gk in pairs:
        assert abs(got[gk] - ref[rk]) <= etol * max(abs(ref[rk]), scale), (name, rk, got[gk], ref[rk])
    assert e["nver"] == int(d["nver"]), (e["nver"], int(d["nver"]))
    fr = d["forces"]
    rms = np.sqrt(((f - fr) ** 2).sum() / (fr ** 2).sum())
    assert rms < ftol, (name, rms)
    for k in ("pvten", "pvten_tot"):
        assert np.max(np.abs(e[k] - d[k])) <= vtol * np.max(np.abs(d[k])), (name, k)
    return rms


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference(name):
    d, ref = load(name)
    c = cases.get(name)
    assert np.array_equal(d["x"], c.x) and np.array_equal(d["y"], c.y) and np.array_equal(d["z"], c.z)
    s = c.system
    flags = (abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES) if s.pi_beads > 1 else 0
    e, f, _ = Oracle(s, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, flags)
    compare(name, e, f, d, ref, TOL, TOL)


def test_fixtures_present():
    assert len(FIXTURES) >= 8
