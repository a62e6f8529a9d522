"""Setup products and Verlet lists against the reference itself, bit for bit.

tests/golden/setup_<case>.npz hold what the reference's own build
(oracle/ref.mk, oracle/ref_harness.c in dump mode) leaves after parse() and
one force call on the case: the Verlet half list as verlist_create.c:373-503
stores it (nter, jver_off, jter), the spline tables of every type pair
(cv0, cdv0, cv0_c, cdv0_c through inter_map_index, control_inter_params.c:
643-772), the half-space k-vectors with their B-spline weights
(set_cp_ewald.c:428-573, 1288-1370) and the excluded-pair Ewald-correction
table (splin_ecor, set_exclude.c:548-779).  The oracle must reproduce every
one of them exactly; the device is held to the same in test_gpu_parity."""
import glob
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
SETUP = sorted(os.path.basename(p)[6:-4] for p in glob.glob(os.path.join(HERE, "golden", "setup_*.npz")))


def load_setup(name):
    return np.load(os.path.join(HERE, "golden", f"setup_{name}.npz"))


def compare_setup(d, verlet_list, spline, kvectors, npair):
    nter, off, jter = verlet_list
    assert np.array_equal(nter, d["nter"])
    assert np.array_equal(off, d["jver_off"])
    assert np.array_equal(jter, d["jter"])
    for which in range(4):
        for ip in range(npair):
            assert np.array_equal(spline(which, ip), d["tables"][which, ip]), (which, ip)
    assert np.array_equal(spline(4, 0), d["ecor"][0])
    assert np.array_equal(spline(5, 0), d["ecor"][1])
    _, ka, kb, kc, bw = kvectors
    for a, k in ((ka, "ka"), (kb, "kb"), (kc, "kc"), (bw, "bweight")):
        assert np.array_equal(a, d[k]), k


@pytest.mark.parametrize("name", SETUP)
def test_oracle_setup_matches_reference(name):
    d = load_setup(name)
    c = cases.get(name)
    s = c.system
    o = Oracle(s, c.x, c.y, c.z)
    o.energy_control(c.x, c.y, c.z, abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0)
    compare_setup(d, o.verlet_list(), o.spline, o.kvectors(), s.ntype * (s.ntype + 1) // 2)


def test_setup_fixtures_present():
    assert len(SETUP) >= 4
