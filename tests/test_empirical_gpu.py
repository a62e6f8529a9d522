"""PINYEmpirical (piny_amd/empirical.py), the drop-in for the reference's
python/piny/piny.pyx:76-436, on the device: P replicas read from PINY input
directories (written here with piny_amd.pinyio), their energies and
gradients against the oracle per replica (FP64, 1e-9), the array layouts
(n_atoms, 3, P), set_x recomputing the ghost sites, and get_V / get_dV_dx
keeping the last update()."""
import os
import tempfile

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import pinyio
from piny_amd import water as w
from piny_amd.empirical import PINYEmpirical

pytestmark = pytest.mark.gpu


def test_replicas_against_oracle():
    st = w.Settings(cutoff=5.5, alpha=0.62, kmax_ewd=6, kmax_pme=6)
    cfgs = [w.build_config(w.Q_TIP4P_F, 64, w.density_box(64), 1, 40 + i, st) for i in range(3)]
    with tempfile.TemporaryDirectory() as d:
        for i, c in enumerate(cfgs):
            pinyio.write_input_dir(os.path.join(d, "rep%d" % i), c)
        print("[PINYEmpirical 3 replicas of 64 q-TIP4P/F]", flush=True)
        m = PINYEmpirical(os.path.join(d, "rep%d"), "water.input", fn_out="parse.out", P=3)
        assert os.path.exists(os.path.join(d, "rep0", "parse.out"))
    assert m.get_P() == 3 and m.get_n_atoms() == cfgs[0].system.natm
    x = m.get_x()
    assert x.shape == (m.get_n_atoms(), 3, 3)
    m.update()
    V, dV = m.get_V(), m.get_dV_dx()
    for i, c in enumerate(cfgs):
        xi, yi, zi = x[:, 0, i].copy(), x[:, 1, i].copy(), x[:, 2, i].copy()
        eo, fo, _ = Oracle(c.system, xi, yi, zi).energy_control(xi, yi, zi, 0)
        vo = eo["vintert"] + eo["vintrat"]
        assert abs(V[i] - vo) <= 1e-9 * abs(vo), (i, V[i], vo)
        fo = fo.reshape(-1, 3)
        assert np.sqrt(((dV[:, :, i] + fo) ** 2).sum() / (fo ** 2).sum()) < 1e-9
    # set_x moves one replica; the ghost (M) sites follow their components
    x2 = x.copy()
    x2[:, 0, 1] += 0.01
    x2[3::4, :, 1] = 0.0  # ghost positions given wrong on purpose: set_x recomputes them
    m.set_x(x2)
    got = m.get_x()
    assert np.allclose(got[3::4, :, 1], x2[0::4, :, 1] * 0.73612 + (x2[1::4, :, 1] + x2[2::4, :, 1]) * 0.13194,
                       rtol=0, atol=1e-12)
    # energies stay those of the last update until the next one
    assert np.array_equal(m.get_V(), V)
    m.update()
    assert m.get_V()[1] != V[1] and m.get_V()[0] == V[0]
    m.close()
