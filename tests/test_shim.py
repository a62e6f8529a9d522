"""The C side of the drop-in (piny_amd/shim/pmx_piny.c) inside a PINY
program: oracle/_ref/piny_shim links the reference's own objects (built from
its sources by oracle/ref.mk), the shim and the library.  It reads a PINY
input directory with the reference's parse(), fills pmx_system / pmx_params
from the structures parse() leaves (pmx_piny_fill) and, on the GPU, makes
the reference's force call and the library's (pmx_piny_energy_control) on
the same PINY structures.

CPU: every filled array equals the arrays the case was written from (the
parse-time LJ parameters, k-space keyword and shift option, which PINY does
not keep, are handed over as pmx_piny_parse).  GPU: forces, energies and
pressure tensors of the two force calls agree to 1e-9."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from piny_amd import abi
from tests import cases
from tests.golden.make_ref_vectors import write_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "piny_shim")
needs_shim = pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/piny_shim is built only where the "
                                                                  "reference's sources are (oracle/ref.mk)")


def run_shim(name, run=False):
    c = cases.get(name)
    s = c.system
    npair = s.ntype * (s.ntype + 1) // 2
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        parse = np.concatenate([[npair], np.asarray(s.inter_label, float), s.inter_sig, s.inter_eps,
                                [s.settings.kmax_ewd, s.settings.ishift]]).astype(np.float64)
        parse.tofile(os.path.join(d, "parse.bin"))
        cmd = [EXE, "water.input", "pos.bin", "parse.bin", "out.bin"] + (["run"] if run else [])
        with open(os.path.join(d, "stdout.txt"), "w") as log:
            r = subprocess.run(cmd, cwd=d, stdout=log, stderr=subprocess.STDOUT, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(open(os.path.join(d, "stdout.txt")).read()[-3000:])
        raw = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
    return c, raw


def unpack(raw):
    N, P, nt, npair, nb, nbt, na, nat, ng = (int(v) for v in raw[:9])
    o = 9
    out = {}

    def take(key, n, as_int=False):
        nonlocal o
        v = raw[o:o + n]
        o += n
        out[key] = v.astype(np.int64) if as_int else v.copy()

    take("atm_type", N, True)
    take("q", N)
    take("mass", N)
    take("inter_label", npair, True)
    for k in ("inter_sig", "inter_eps", "inter_cutti", "inter_cutoff", "inter_cutoff_res"):
        take(k, npair)
    for k in ("bond_j1", "bond_j2", "bond_type"):
        take(k, nb, True)
    take("bond_c", 7 * nbt)
    take("bond_eq", nbt)
    for k in ("bend_j1", "bend_j2", "bend_j3", "bend_type"):
        take(k, na, True)
    take("bend_c", 7 * nat)
    take("bend_s", 7 * nat)
    take("ghost_atom", ng, True)
    take("ghost_ncomp", ng, True)
    take("ghost_comp", 4 * ng, True)
    take("ghost_coef", 4 * ng)
    take("hmat", 9)
    take("params", 13)
    results = []
    while o < len(raw):
        blk = {}
        blk["e"] = raw[o:o + 12]
        blk["pvten"] = raw[o + 12:o + 21]
        blk["pvten_tot"] = raw[o + 21:o + 30]
        o += 30
        blk["forces"] = raw[o:o + 3 * N * P].reshape(P, N, 3)
        o += 3 * N * P
        results.append(blk)
    return (N, P), out, results


FILL_CASES = ["spc64", "tip64_p4", "tip64_p4_tric", "tip64_p4_stag", "spc64_shift", "spc64_none"]


@needs_shim
@pytest.mark.parametrize("name", FILL_CASES)
def test_shim_fills_the_arrays_the_case_was_written_from(name):
    c, raw = run_shim(name)
    s = c.system
    (N, P), got, _ = unpack(raw)
    assert N == s.natm and P == s.pi_beads
    for k in ("atm_type", "inter_label", "bond_j1", "bond_j2", "bond_type", "bend_j1", "bend_j2", "bend_j3",
              "bend_type", "ghost_atom", "ghost_ncomp"):
        assert np.array_equal(got[k], np.asarray(getattr(s, k))), k
    for k in ("q", "mass", "inter_sig", "inter_eps", "inter_cutti", "inter_cutoff", "inter_cutoff_res", "bond_c",
              "bond_eq", "bend_c", "bend_s"):
        a, b = got[k], np.asarray(getattr(s, k), dtype=np.float64).reshape(-1)
        assert np.allclose(a, b, rtol=1e-14, atol=0), (k, np.max(np.abs(a - b)))
    comp = np.asarray(s.ghost_comp).reshape(-1, 4)
    coef = np.asarray(s.ghost_coef).reshape(-1, 4)
    for g in range(len(s.ghost_atom)):
        n = int(s.ghost_ncomp[g])
        assert np.array_equal(got["ghost_comp"].reshape(-1, 4)[g, :n], comp[g, :n])
        assert np.allclose(got["ghost_coef"].reshape(-1, 4)[g, :n], coef[g, :n], rtol=1e-15, atol=0)
    st = s.settings
    assert np.allclose(got["hmat"], np.asarray(s.hmat).reshape(-1), rtol=1e-13, atol=1e-13)
    pr = got["params"]
    assert int(pr[0]) == P and pr[1] == st.temperature and int(pr[2]) == st.pimd_typ
    assert int(pr[3]) == st.nsplin and int(pr[4]) == st.ishift
    assert abs(pr[5] - st.rheal / 0.529177) <= 1e-14 * pr[5] and abs(pr[6] - st.skin / 0.529177) <= 1e-14 * pr[6]
    assert abs(pr[7] - s.ewald_alpha) <= 1e-13 * s.ewald_alpha
    assert (int(pr[8]), int(pr[9]), int(pr[10]), int(pr[11]), int(pr[12])) == (
        st.kmax_ewd, st.pme_on, st.kmax_pme, st.n_interp, st.intra_perds)


@needs_shim
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["spc64", "tip64_p4", "tip64_p4_stag"])
def test_shim_force_call_matches_the_reference_in_one_program(name):
    print(f"[shim {name}]", flush=True)
    c, raw = run_shim(name, run=True)
    (N, P), _, res = unpack(raw)
    assert len(res) == 2
    ref, mine = res
    fr = ref["forces"]
    rms = np.sqrt(((mine["forces"] - fr) ** 2).sum() / (fr ** 2).sum())
    print(f"  force rms rel err {rms:.2e}", flush=True)
    assert rms < 1e-9
    scale = max(abs(ref["e"][0]), abs(ref["e"][1]), 1e-3)
    for k in range(6 if P == 1 else 9):  # vintert vintrat vrecip vlong vbondt vbendt (pi_ke_prim pi_ke_vir kin_harm)
        assert abs(mine["e"][k] - ref["e"][k]) <= 1e-9 * max(abs(ref["e"][k]), scale), (k, mine["e"][k], ref["e"][k])
    for key in ("pvten", "pvten_tot"):
        assert np.max(np.abs(mine[key] - ref[key])) <= 1e-9 * np.max(np.abs(ref[key])), key
