"""Velocity Verlet (int_NVE, md/int_NVE.c:29-125) against the reference's
own 100-step trajectory of configuration 1 (256 SPC/Fw, 0.25 fs steps,
tests/golden/ref_md_cfg1.npz, written by tests/golden/make_ref_vectors.py
from the reference build).  On the CPU the oracle's forces drive the same
steps in the reference's order of operations."""
import os

import numpy as np

from oracle.oracle import Oracle
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return np.load(os.path.join(HERE, "golden", "ref_md_cfg1.npz"))


def oracle_nve(c, v0, nsteps, dt):
    s = c.system
    m = np.asarray(s.mass, dtype=np.float64)[:, None]
    x = np.stack([c.x, c.y, c.z], 1).copy()
    v = v0.copy()
    o = Oracle(s, c.x, c.y, c.z)
    e, f, _ = o.energy_control(x[:, 0], x[:, 1], x[:, 2])
    f = f[0]
    dt2 = dt / 2.0
    for _ in range(nsteps):
        v += f * dt2 / m
        x += v * dt
        e, f, xyz = o.energy_control(x[:, 0], x[:, 1], x[:, 2])
        x = np.stack(xyz, 1)[: s.natm].copy()   # ghost positions as the force call set them
        f = f[0]
        v += f * dt2 / m
    ke = 0.5 * (m[:, 0] * (v * v).sum(1)).sum()
    return e, ke, x, v, f


def test_oracle_nve_matches_reference():
    d = load()
    c = cases.get(str(d["case"]))
    e, ke, x, v, f = oracle_nve(c, d["v0"], int(d["nsteps"]), float(d["dt"]))
    # observed: positions 3e-14 Bohr, velocities 5e-14, forces 7e-14 RMS,
    # vintert 4e-15, kinetic energy 5e-16 (relative)
    assert np.max(np.abs(x - d["x"])) < 1e-12
    assert np.max(np.abs(v - d["v"])) < 1e-12 * np.max(np.abs(d["v"]))
    assert np.sqrt(((f - d["f"]) ** 2).sum() / (d["f"] ** 2).sum()) < 1e-11
    assert abs(e["vintert"] - d["vintert"]) < 1e-12 * abs(d["vintert"])
    assert abs(ke - d["kinet"]) < 1e-12 * d["kinet"]
