"""RESPA force calls on the device, through the C ABI: against the
reference's own RESPA calls (tests/golden/respa_*.npz, every flag
combination) in FP64 and mixed precision, against the oracle with the half
list, in a batch, and across calls that change the flags on one context.

The reference builds the virial estimator and pvten_tot of a full call
from the unweighted forces fxt; under RESPA weights the device's single
force call forms them in an unweighted first pass, a batched call returns
them as NaN."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd.engine import Batch, Engine
from tests import cases
from tests.test_respa import RESPA_CASES, compare_respa, flags_of, load

pytestmark = pytest.mark.gpu

ETOL = FTOL = 1e-9
MIXED_ETOL, MIXED_VTOL, MIXED_FTOL = 1e-5, 1e-4, 1e-4


@pytest.mark.parametrize("mode", ["fp64", "direct", "mixed"])
@pytest.mark.parametrize("name", RESPA_CASES)
def test_device_respa_matches_reference(name, mode):
    d = load(name)
    c = cases.get(name)
    s = c.system
    print(f"[respa fixture {name}] N={s.natm} P={s.pi_beads} {mode}", flush=True)
    kw = {"fp64": {}, "direct": {"pair_eval": abi.PMX_PAIR_DIRECT},
          "mixed": {"precision": abi.PMX_PREC_MIXED}}[mode]
    g = Engine(s, c.x, c.y, c.z, **kw)
    g.set_respa(abi.respa_params(s, nfree=int(d["nfree"])))
    g.set_positions(c.x, c.y, c.z)
    etol, ftol, vtol = (ETOL, FTOL, ETOL) if mode != "mixed" else (MIXED_ETOL, MIXED_FTOL, MIXED_VTOL)
    if mode == "direct":  # the tabulated functions evaluated directly (their interpolation error)
        etol, ftol, vtol = 1e-7, 2e-6, 1e-6
    for k, (full, res, fintra) in enumerate(d["flags"]):
        e = g.energy_control(flags_of(full, res, fintra, s.pi_beads))
        f = g.forces()
        assert e["nver"] == int(d["nver"])
        rms = compare_respa(name, k, e, f, d, etol, ftol, vtol, term_scale=(mode == "mixed"))
        if s.pi_beads > 1 and full:  # fxt: the unweighted forces, as the reference keeps them
            fxt = g.fxt()
            assert np.isfinite(fxt).all()
        print(f"  flags {tuple(int(v) for v in d['flags'][k])}: force rms rel err {rms:.2e}", flush=True)
    g.close()


def test_respa_half_list_matches_oracle(monkeypatch):
    monkeypatch.setenv("PMX_PAIR_LIST", "half")
    c = cases.get("tip64_p4_respa")
    s = c.system
    print("[respa half list tip64_p4_respa]", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    o.set_respa(abi.respa_params(s))
    g = Engine(s, c.x, c.y, c.z)
    g.set_respa(abi.respa_params(s))
    g.set_positions(c.x, c.y, c.z)
    for full, res in ((1, 1), (0, 1), (1, 0)):
        fl = flags_of(full, res, 1, s.pi_beads)
        eo, fo, _ = o.energy_control(c.x, c.y, c.z, fl)
        eg = g.energy_control(fl)
        fg = g.forces()
        assert np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum()) < 1e-9
        assert np.max(np.abs(eg["pvten"] - eo["pvten"])) <= 1e-9 * np.max(np.abs(eo["pvten"]))
        for k in ("vintert", "vintrat"):
            assert abs(eg[k] - eo[k]) <= 1e-9 * max(abs(eo[k]), 1e-3), k
    g.close()


def test_respa_flags_change_between_calls_and_in_a_batch():
    """One context through a sequence of RESPA calls (full, short-range,
    intramolecular only, full again) against fresh single calls, and two
    boxes with RESPA in one batched launch against their own calls."""
    names = ["tip64_p4_respa", "spc64_respa"]
    cfgs = [cases.get(n) for n in names]
    seq = [(1, 1, 1), (0, 1, 1), (0, 0, 0), (1, 1, 1)]
    print("[respa call sequence and batch]", flush=True)
    engs = []
    for c in cfgs:
        g = Engine(c.system, c.x, c.y, c.z)
        g.set_respa(abi.respa_params(c.system))
        g.set_positions(c.x, c.y, c.z)
        engs.append(g)
    ref = {}
    for c, g, n in zip(cfgs, engs, names):
        for fl in seq:
            h = Engine(c.system, c.x, c.y, c.z)
            h.set_respa(abi.respa_params(c.system))
            h.set_positions(c.x, c.y, c.z)
            e1 = h.energy_control(flags_of(*fl, c.system.pi_beads))
            ref[(n, fl)] = (e1, h.forces())
            h.close()
        for fl in seq:
            e = g.energy_control(flags_of(*fl, c.system.pi_beads))
            e1, f1 = ref[(n, fl)]
            assert np.max(np.abs(g.forces() - f1)) <= 1e-12 * max(np.max(np.abs(f1)), 1e-30), (n, fl)
            assert abs(e["vintrat"] - e1["vintrat"]) <= 1e-12 * max(abs(e1["vintrat"]), 1e-3)
    b = Batch(engs)
    for fl in seq:
        # PMX_EVAL_PIMD is implied for the beads' box, mode forces ignored for the classical one
        out = b.energy_control(flags_of(*fl, 1) | abi.PMX_EVAL_MODE_FORCES)
        for c, g, n, e in zip(cfgs, engs, names, out):
            e1, f1 = ref[(n, fl)]
            assert np.max(np.abs(g.forces() - f1)) <= 1e-12 * max(np.max(np.abs(f1)), 1e-30), (n, fl)
            assert abs(e["vintert"] - e1["vintert"]) <= 1e-12 * max(abs(e1["vintert"]), 1e-3), (n, fl)
    b.close()
    for g in engs:
        g.close()


def test_respa_fxt_and_integrators():
    """fxt of a weighted full call equals the forces of the same call without
    RESPA (the bead forces, before the mode transform); a short-range call
    has none; the RESPA-off integrators refuse a RESPA context; a batch
    returns the estimator as NaN."""
    c = cases.get("tip64_p4_respa")
    s = c.system
    print("[respa fxt tip64_p4_respa]", flush=True)
    g = Engine(s, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    e0 = g.energy_control(abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES)
    fxt0 = g.fxt()
    g.set_respa(abi.respa_params(s))
    e = g.energy_control(flags_of(1, 1, 1, s.pi_beads))
    assert np.max(np.abs(g.fxt() - fxt0)) <= 1e-12 * np.max(np.abs(fxt0))
    assert abs(e["pi_ke_vir"] - e0["pi_ke_vir"]) <= 1e-12 * abs(e0["pi_ke_vir"])
    assert np.max(np.abs(e["pvten_tot"] - e0["pvten"])) <= 1e-12 * np.max(np.abs(e0["pvten"]))
    from piny_amd.engine import PinyAmdError
    g.energy_control(flags_of(0, 1, 1, s.pi_beads))
    with pytest.raises(PinyAmdError):
        g.fxt()
    with pytest.raises(PinyAmdError):
        g.pimd_nvt(1)
    b = Batch([g])
    eb = b.energy_control(flags_of(1, 1, 1, s.pi_beads))[0]
    assert np.isnan(eb["pi_ke_vir"]) and np.all(np.isnan(eb["pvten_tot"]))
    assert abs(eb["vintert"] - e["vintert"]) <= 1e-12 * abs(e["vintert"])
    b.close()
    # off again: the plain call, the estimator and fxt are back
    g.set_respa(None)
    e = g.energy_control(abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES)
    assert np.isfinite(e["pi_ke_vir"])
    g.fxt()
    g.close()


def test_respa_mixed_cutoff_per_type_pair():
    """Mixed precision under RESPA with type pairs of different cutoffs (the
    kernel switches the erfc term itself, as without one erfc function for
    every type pair): full, short-range and both, against the oracle."""
    import dataclasses
    c = cases.get("tip64_p4_respa")
    cut = np.array(c.system.inter_cutoff, dtype=float).copy()
    cut[0] *= 0.92
    s = dataclasses.replace(c.system, inter_cutoff=cut)
    print(f"[respa mixed, cutoffs {cut}]", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    o.set_respa(abi.respa_params(s))
    g = Engine(s, c.x, c.y, c.z, precision=abi.PMX_PREC_MIXED)
    g.set_respa(abi.respa_params(s))
    g.set_positions(c.x, c.y, c.z)
    for full, res in ((1, 1), (0, 1), (1, 0)):
        fl = flags_of(full, res, 1, s.pi_beads)
        eo, fo, _ = o.energy_control(c.x, c.y, c.z, fl)
        eg = g.energy_control(fl)
        fg = g.forces()
        rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
        print(f"  flags ({full}, {res}): force rms rel err {rms:.2e}", flush=True)
        assert rms < MIXED_FTOL
        assert np.max(np.abs(eg["pvten"] - eo["pvten"])) <= MIXED_VTOL * np.max(np.abs(eo["pvten"]))
        scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
        for k in ("vintert", "vintrat"):
            assert abs(eg[k] - eo[k]) <= MIXED_ETOL * max(abs(eo[k]), scale), k
    g.close()
