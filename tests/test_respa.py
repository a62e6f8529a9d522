"""RESPA force calls: the oracle against the reference itself.

tests/golden/respa_<case>.npz hold the reference's own force call
(oracle/ref_harness.c respa mode, tests/golden/make_ref_vectors.py) on an
input with respa_steps_lrf/intra/torsion (and respa_steps_pimd, pressure,
pten_inter_respa for the bead case), for each combination of
iget_full_inter, iget_res_inter and iget_full_intra the fixture lists.  The
reference weights the force terms by the step counts
(energy_control_initial.c:212-243), switches the short-range part of the
pair forces off below inter_cutoff_res (force_npol.c:807-816,1033-1060),
adds RESPA terms to the pressure tensor (energy_control_inter_real.c:
206-318) and keeps energies only in full calls.

The plain-call limit pins the weights independently: with every flag of a
full call and every step count 1 the result is the plain fixture's."""
import dataclasses
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
RESPA_CASES = ["spc64_respa", "spc64_ewald_respa", "tip64_p4_respa"]
TOL = 1e-12


def load(name):
    return np.load(os.path.join(HERE, "golden", f"respa_{name}.npz"))


def flags_of(full, res, full_intra, P):
    f = (abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES) if P > 1 else 0
    if not full:
        f |= abi.PMX_EVAL_NO_FULL_INTER
    if res:
        f |= abi.PMX_EVAL_RES_INTER
    if not full_intra:
        f |= abi.PMX_EVAL_NO_FULL_INTRA
    return f


def compare_respa(name, k, e, f, d, etol, ftol, vtol=None, no_fxt=False, term_scale=False):
    """One flag combination k of fixture d: the energies the reference keeps
    (vintert, vintrat, vbond, vbend, its vrecip = k-space + self +
    background + ecor, the implied real-space energy in full calls, the
    estimators), forces, pvten and pvten_tot.  no_fxt: the result comes from
    a path that does not keep the unweighted forces, so in a full call the
    virial estimator and pvten_tot must be NaN (the device under RESPA
    weights)."""
    vtol = etol if vtol is None else vtol
    ref = {str(a): float(b) for a, b in zip(d["energy_keys"], d["energies"][k])}
    full = int(d["flags"][k][0])
    skip_fxt = no_fxt and bool(full)
    got = dict(e)
    got["vrecip_total"] = e["vrecip"] + e["vself"] + e["vbgr"] + e["vecor"]
    scale = max(abs(ref["vintert"]), abs(ref["vintrat"]), abs(ref["vrecip"]), 1e-3)
    if term_scale and full:
        scale = max(scale, abs(ref["vintert"] - ref["vrecip"] - ref["vlong"]))
    pairs = [("vintert", "vintert"), ("vintrat", "vintrat"), ("vlong", "vlong"), ("vbond", "vbond"),
             ("vbend", "vbend"), ("vrecip", "vrecip_total")]
    if full:
        ref["vreal_implied"] = ref["vintert"] - ref["vrecip"] - ref["vlong"]
        got["vreal_implied"] = e["vreal"]
        pairs.append(("vreal_implied", "vreal_implied"))
    if int(d["forces"].shape[1]) > 1:
        pairs += [("pi_ke_prim", "pi_ke_prim"), ("kin_harm", "kin_harm")]
        if skip_fxt:
            assert np.isnan(e["pi_ke_vir"]), (name, k)
        else:
            pairs.append(("pi_ke_vir", "pi_ke_vir"))
    for rk, gk in pairs:
        assert abs(got[gk] - ref[rk]) <= etol * max(abs(ref[rk]), scale), (name, k, rk, got[gk], ref[rk])
    fr = d["forces"][k]
    den = max((fr ** 2).sum(), 1e-30)
    rms = np.sqrt(((f - fr) ** 2).sum() / den)
    if (fr ** 2).sum() < 1e-24:  # no forces in this call (molecules at their equilibrium geometry)
        assert np.max(np.abs(f)) < 1e-12, (name, k)
    else:
        assert rms < ftol, (name, k, rms)
    # tensors relative to their size in the case's calls (a call whose only
    # terms are bonds at their equilibrium length has a tensor of rounding
    # noise)
    pv_scale = float(np.max(np.abs(d["pvten"])))
    for key in ("pvten", "pvten_tot"):
        ref_v = d[key][k]
        if key == "pvten_tot" and skip_fxt:
            assert np.all(np.isnan(e[key])), (name, k)
            continue
        assert np.max(np.abs(e[key] - ref_v)) <= vtol * max(np.max(np.abs(ref_v)), pv_scale), (name, k, key)
    return rms


@pytest.mark.parametrize("name", RESPA_CASES)
def test_oracle_matches_reference_respa(name):
    d = load(name)
    c = cases.get(name)
    assert np.array_equal(d["x"], c.x)
    s = c.system
    o = Oracle(s, c.x, c.y, c.z)
    resp = abi.respa_params(s)
    assert resp is not None and resp.nfree == int(d["nfree"])  # the reference's own count
    o.set_respa(resp)
    for k, (full, res, fintra) in enumerate(d["flags"]):
        e, f, _ = o.energy_control(c.x, c.y, c.z, flags_of(full, res, fintra, s.pi_beads))
        assert e["nver"] == int(d["nver"])
        compare_respa(name, k, e, f, d, TOL, TOL)


def test_respa_weights_reduce_to_the_plain_call():
    """Step counts all 1 with the inter split on: a full call with the
    short-range part too weighs every term by 1, so the blended pair weight
    sw + (1 - sw) is 1 and the result is the plain call's (the RESPA
    pressure terms vanish with the weights, the long-range one becomes the
    plain correction)."""
    c = cases.get("spc64_respa")
    s = c.system
    o = Oracle(s, c.x, c.y, c.z)
    e0, f0, _ = o.energy_control(c.x, c.y, c.z, 0)
    st1 = dataclasses.replace(s.settings, respa_lrf=1, respa_intra=1, respa_torsion=1)
    s1 = dataclasses.replace(s, settings=st1)
    o.set_respa(abi.respa_params(s1))
    e1, f1, _ = o.energy_control(c.x, c.y, c.z, abi.PMX_EVAL_RES_INTER)
    assert np.max(np.abs(f1 - f0)) <= 1e-13 * np.max(np.abs(f0))
    for k in ("vintert", "vreal", "vrecip", "vecor"):
        assert abs(e1[k] - e0[k]) <= 1e-13 * max(abs(e0[k]), 1e-3), k
    # the short-range part alone: forces of the size of the full ones, no
    # intermolecular energy
    e_res, f_res, _ = o.energy_control(c.x, c.y, c.z, abi.PMX_EVAL_RES_INTER | abi.PMX_EVAL_NO_FULL_INTER)
    assert np.max(np.abs(f_res)) > 0.1 * np.max(np.abs(f0))
    assert np.max(np.abs(f_res)) < np.max(np.abs(f0)) * 10
    assert e_res["vintert"] == 0.0 and e_res["vreal"] == 0.0


def test_respa_argument_checks():
    c = cases.get("spc64_respa")
    o = Oracle(c.system, c.x, c.y, c.z)
    bad = abi.PmxRespa(respa_steps_lrf=2, respa_steps_torsion=1, respa_steps_intra=0)
    with pytest.raises(RuntimeError):
        o.set_respa(bad)
