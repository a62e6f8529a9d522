"""PINY's on-disk formats (piny_amd/pinyio.py), on the CPU.

The pin: the input directories that the reference's own parse() read to
write the golden fixtures (tests/golden/make_ref_vectors.py writes them,
oracle/ref_harness.c parses them) are read back by the product reader; the
system it builds must hold the case's arrays, and the oracle run on it must
reproduce the reference's forces and energies from those fixtures.  Then
round trips of the writers (input directory, .initial, .confp) and the
header layout of a configuration file (write_gen_header.c:78-150)."""
import os
import tempfile

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi, pinyio
from piny_amd import water as w
from tests import cases
from tests.golden.make_ref_vectors import write_inputs
from tests.test_ref_golden import TOL, compare, load

ARRAYS = ("atm_type", "q", "mass", "inter_label", "inter_sig", "inter_eps", "inter_cutti", "inter_cutoff",
          "inter_cutoff_res", "bond_j1", "bond_j2", "bond_type", "bond_c", "bond_eq", "bend_j1", "bend_j2",
          "bend_j3", "bend_type", "bend_c", "bend_s", "ghost_atom", "ghost_ncomp", "ghost_comp", "ghost_coef")

# cases whose inputs the reference parsed: classical and beads, SPC/Fw and
# q-TIP4P/F (a ghost site, Morse bonds), a triclinic cell, staging, a box
# without charges
CASES = ["spc64", "tip64_p4", "cfg1", "tip64_p4_tric", "tip64_p4_stag", "spc64_shift", "spc64_none"]
SETTINGS = ("skin", "rheal", "kmax_ewd", "kmax_pme", "n_interp", "nsplin", "ishift", "pme_on", "intra_perds",
            "temperature", "pimd_typ", "respa_lrf", "respa_torsion", "respa_intra", "respa_pimd", "pressure",
            "pten_inter_respa")


@pytest.mark.parametrize("name", CASES)
def test_reader_rebuilds_what_the_reference_parsed(name):
    c = cases.get(name)
    s = c.system
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        r = pinyio.read_input_dir(d, "water.input")
    t = r.system
    assert t.natm == s.natm and t.ntype == s.ntype and t.pi_beads == s.pi_beads
    for k in ARRAYS:
        a, b = np.asarray(getattr(t, k)), np.asarray(getattr(s, k))
        assert a.shape == b.shape and np.array_equal(a, b), (k, a, b)
    assert np.allclose(t.hmat, s.hmat, rtol=1e-14, atol=0)
    assert abs(t.ewald_alpha - s.ewald_alpha) <= 1e-15 * s.ewald_alpha
    for k in SETTINGS:
        assert getattr(t.settings, k) == getattr(s.settings, k), k
    # the .initial positions (15 decimals in Angstrom) to rounding
    for u, v in ((r.x, c.x), (r.y, c.y), (r.z, c.z)):
        assert np.max(np.abs(u - v)) < 1e-13 * np.max(np.abs(v))
    # the oracle on the system the reader built, at the positions the
    # reference was given, against the reference's own results
    d, ref = load(name)
    flags = (abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES) if t.pi_beads > 1 else 0
    e, f, _ = Oracle(t, c.x, c.y, c.z).energy_control(c.x, c.y, c.z, flags)
    compare(name, e, f, d, ref, TOL, TOL)


def test_input_dir_round_trip():
    c = cases.get("tip64_p4_tric")
    with tempfile.TemporaryDirectory() as d:
        pinyio.write_input_dir(d, c, name="box")
        r = pinyio.read_input_dir(d)
    for k in ARRAYS:
        assert np.array_equal(np.asarray(getattr(r.system, k)), np.asarray(getattr(c.system, k))), k
    assert np.allclose(r.system.hmat, c.system.hmat, rtol=1e-14, atol=0)
    assert np.max(np.abs(r.x - c.x)) < 1e-13 * np.max(np.abs(c.x))


def test_dictionary_grammar():
    text = "~sim_gen_def[\\temperature{300} \\restart_type{initial}]\n\n" \
           "~atom_def[\n    \\atom_typ{M}\n    \\def_ghost1{1, 0.73612}]\n~atom_def[\\atom_typ{H}]"
    secs = pinyio.parse_dict_text(text)
    assert secs == [("sim_gen_def", {"temperature": "300", "restart_type": "initial"}),
                    ("atom_def", {"atom_typ": "M", "def_ghost1": "1, 0.73612"}), ("atom_def", {"atom_typ": "H"})]
    assert pinyio.parse_dict_text(pinyio.dict_text(secs)) == secs


def test_conf_file_round_trip_and_header():
    c = cases.get("tip64_p4")
    s = c.system
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "water.confp")
        h = pinyio.conf_header_for(s, dt=0.25 / abi.TIME_CONV, write_freq=10)
        cw = pinyio.ConfWriter(p, h)
        cw.write(c.x, c.y, c.z, s.hmat)
        cw.write(c.x + 0.5, c.y, c.z, s.hmat)
        lines = open(p).read().splitlines()
        hr, frames = pinyio.read_conf(p)
    head = lines[0].split()
    # ibinary mach_typ file_typ ntime dt freq P nmol_typ nres_typ natm_typ natm ...
    assert head[0] == "0" and head[2] == "pos_file" and int(head[6]) == s.pi_beads and int(head[10]) == s.natm
    assert lines[1:4] == ["water", "water", "O"] and lines[5] == "M"
    assert len(lines) == 1 + 1 + 1 + 3 + s.natm + 2 * (s.pi_beads * s.natm + 3)
    assert hr.natm == s.natm and hr.pi_beads == s.pi_beads and hr.atm_types == ("O", "H", "M")
    assert len(frames) == 2
    for (x, y, z, hm), dx in zip(frames, (0.0, 0.5)):
        assert np.max(np.abs(x - (c.x + dx))) < 1e-10 * np.max(np.abs(c.x))
        assert np.allclose(hm, s.hmat, rtol=1e-12)


def test_respa_input_read_back():
    """The RESPA keywords of an input the reference parsed for the RESPA
    fixtures come back as the settings, and the pmx_respa they give matches
    the reference's counted degrees of freedom."""
    from tests.test_respa import load as load_respa
    c = cases.get("tip64_p4_respa")
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        r = pinyio.read_input_dir(d, "water.input")
    for k in SETTINGS:
        assert getattr(r.system.settings, k) == getattr(c.system.settings, k), k
    assert np.array_equal(r.system.inter_cutoff_res, c.system.inter_cutoff_res)
    resp = abi.respa_params(r.system)
    assert resp.respa_steps_lrf == 2 and resp.respa_steps_pimd == 2
    assert resp.nfree == int(load_respa("tip64_p4_respa")["nfree"])
    with tempfile.TemporaryDirectory() as d:
        pinyio.write_input_dir(d, c, name="box")
        r2 = pinyio.read_input_dir(d)
    for k in SETTINGS:
        assert getattr(r2.system.settings, k) == getattr(c.system.settings, k), k


def _golden(name):
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))


def test_restart_text_is_the_reference_writers():
    """The reference's own restart writers (write_dump_file_pimd after 20
    PIMD NVT steps, write_dump_file_md after 100 NVE steps; oracle/ref_harness.c)
    wrote the fixtures' restart files from the states the fixtures hold;
    restart_text writes the same states byte for byte, and read_restart
    reads the reference's files back to those states."""
    # PIMD: bead positions, mode velocities, the global class chain and the
    # per-degree-of-freedom bead chains
    d = _golden("ref_pimd_tip64_p4.npz")
    c = cases.get("tip64_p4")
    s = c.system
    L = int(d["len_nhc"])
    state = np.concatenate([d["nhc_class"], d["nhc_bead"].reshape(-1)])
    cls, bead = pinyio.nhc_state_to_restart(s, state, L)
    r = pinyio.restart_for(s, d["pos"], d["vel"], nhc_class_v=cls, nhc_bead_v=bead, dt=float(d["dt"]))
    ref_text = str(d["restart"])
    assert pinyio.restart_text(r) == ref_text
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "water.restart")
        open(p, "w").write(ref_text)
        back = pinyio.read_restart(p)
    assert back.natm == s.natm and back.pi_beads == s.pi_beads
    assert np.allclose(back.pos, d["pos"], rtol=1e-12, atol=0) and np.allclose(back.vel, d["vel"], rtol=1e-12, atol=0)
    assert np.allclose(back.nhc_class_v, cls, rtol=1e-12, atol=0)
    assert np.allclose(back.nhc_bead_v, bead, rtol=1e-9, atol=1e-300)  # the bead chains are written with 10 digits
    assert back.labels[3] == ("M", "water", "water", 1) and back.nfree == 576 and back.ensemble == (0, 1, 0, 0, 0)
    assert pinyio.restart_text(back) == ref_text
    # classical NVE: Cartesian velocities, no chains
    d = _golden("ref_md_cfg1.npz")
    c = cases.get("cfg1")
    s = c.system
    r = pinyio.restart_for(s, d["x"][None], d["v"][None], dt=float(d["dt"]))
    assert pinyio.restart_text(r) == str(d["restart"])
