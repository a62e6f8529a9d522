"""Gradients of the force call (the backward kernels and the autograd
Function) on the GPU.

Checks: central finite differences of the oracle's energies (h = 1e-5
relative) for every physical parameter of the water models and for single
site charges, on small boxes; on the 1024 q-TIP4P/F box at P = 8 (five
seeded configurations) central differences of the device's own FP64
energies, and -dE/dx against the oracle's forces; the mixed precision against
FP64.  The reference computes no parameter derivatives, so these checks pin
them by the mathematics (the FD limit) on energies that are themselves
pinned to the reference."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd import autograd as A
from piny_amd import water as w
from piny_amd.engine import Engine
from tests import cases

pytestmark = pytest.mark.gpu

H_REL = 1e-5
# FD truncation (h^2) and rounding are far below this; the bar is set by the
# tables: the analytic LJ derivatives use the LJ form itself, the energies
# PINY's cubic interpolation of it (about 1e-7 relative next to the
# switching onset)
GRAD_TOL = 2e-6


def _pos(c, P, N):
    return torch.tensor(np.stack([c.x, c.y, c.z], 1).reshape(P, N, 3))


def _oracle_energy(system, c):
    o = Oracle(system, c.x, c.y, c.z)
    flags = abi.PMX_EVAL_PIMD if system.pi_beads > 1 else 0
    e, _, _ = o.energy_control(c.x, c.y, c.z, flags)
    return e["vintert"] + e["vintrat"]


def _analytic(eng, c, params):
    s = c.system
    p = {k: v.clone().requires_grad_(True) for k, v in params.items()}
    pos = _pos(c, s.pi_beads, s.natm).requires_grad_(True)
    E = A.water_energy(eng, pos, p)
    E.mean().backward()
    return E.detach(), pos.grad, {k: v.grad.item() for k, v in p.items()}


def _check(name, g, fd, scale):
    err = abs(g - fd) / max(abs(fd), scale)
    print(f"  {name}: analytic {g:.12e} fd {fd:.12e} rel {err:.2e}", flush=True)
    assert err < GRAD_TOL, (name, g, fd)


@pytest.mark.parametrize("case", ["spc64", "spc64_shift", "spc64_none", "tip64_p4"])
def test_param_grad_vs_oracle_fd(case):
    c = cases.get(case)
    s = c.system
    print(f"[case {case}] N={s.natm} P={s.pi_beads}", flush=True)
    params = A.water_params(s.model)
    eng = Engine(s, c.x, c.y, c.z)
    E, gpos, g = _analytic(eng, c, params)
    e_ref = _oracle_energy(s, c)
    assert abs(E.mean().item() - e_ref) < 1e-9 * abs(e_ref)
    scale = max(abs(v) for v in g.values()) * 1e-3
    for k, v in params.items():
        h = H_REL * abs(v.item())
        ep = _oracle_energy(A.system_with_params(s, {**params, k: v + h}), c)
        em = _oracle_energy(A.system_with_params(s, {**params, k: v - h}), c)
        _check(k, g[k], (ep - em) / (2 * h), scale)
    # -d(mean E)/dx_b is the force call's bead force (f/P with PIMD)
    o = Oracle(s, c.x, c.y, c.z)
    _, fo, _ = o.energy_control(c.x, c.y, c.z, abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0)
    rms = np.sqrt(((-gpos.numpy() - fo) ** 2).sum() / (fo ** 2).sum())
    assert rms < 1e-9


@pytest.mark.parametrize("case,sites", [("spc64", (0, 1, 95)), ("tip64_p4", (1, 3, 250))])
def test_site_charge_grad_vs_oracle_fd(case, sites):
    c = cases.get(case)
    s = c.system
    print(f"[case {case}] sites {sites}", flush=True)
    eng = Engine(s, c.x, c.y, c.z)
    eng.set_positions(c.x, c.y, c.z)
    eng.energy_control(abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0)
    dq = eng.param_grad()["dq"].mean(0)
    assert np.all(dq[s.q == 0.0] == 0.0)
    for i in sites:
        assert s.q[i] != 0.0
        h = H_REL * abs(s.q[i])
        e = []
        for sign in (1.0, -1.0):
            q = s.q.copy()
            q[i] += sign * h
            e.append(_oracle_energy(dataclasses.replace(s, q=q), c))
        _check(f"q[{i}]", dq[i], (e[0] - e[1]) / (2 * h), 1e-6)


@pytest.mark.parametrize("sample", [0, 1, 2, 3, 4])
def test_autograd_config4_fd(sample):
    # 1024 q-TIP4P/F, L = 31.28 A, P = 8, FP64: parameter gradients against
    # central differences of the device's energies
    c = w.config4(sample=sample)
    s = c.system
    print(f"[config4 sample {sample}] N={s.natm} P={s.pi_beads}", flush=True)
    params = A.water_params(s.model)
    eng = Engine(s, c.x, c.y, c.z)
    E, gpos, g = _analytic(eng, c, params)
    pos = _pos(c, s.pi_beads, s.natm)
    scale = max(abs(v) for v in g.values()) * 1e-3
    for k, v in params.items():
        h = H_REL * abs(v.item())
        with torch.no_grad():
            ep = A.water_energy(eng, pos, {**params, k: v + h}).mean().item()
            em = A.water_energy(eng, pos, {**params, k: v - h}).mean().item()
        _check(k, g[k], (ep - em) / (2 * h), scale)
    if sample == 0:
        o = Oracle(s, c.x, c.y, c.z)
        _, fo, _ = o.energy_control(c.x, c.y, c.z, abi.PMX_EVAL_PIMD)
        rms = np.sqrt(((-gpos.numpy() - fo) ** 2).sum() / (fo ** 2).sum())
        print(f"  -dE/dx vs oracle forces: rms rel {rms:.2e}", flush=True)
        assert rms < 1e-9


def test_autograd_mixed_vs_fp64():
    c = w.config4(sample=0)
    s = c.system
    params = A.water_params(s.model)
    e64 = Engine(s, c.x, c.y, c.z)
    emx = Engine(s, c.x, c.y, c.z, precision=abi.PMX_PREC_MIXED)
    E1, gp1, g1 = _analytic(e64, c, params)
    E2, gp2, g2 = _analytic(emx, c, params)
    rms = torch.sqrt(((gp2 - gp1) ** 2).sum() / (gp1 ** 2).sum()).item()
    print(f"  mixed: force rms rel {rms:.2e}", flush=True)
    assert rms < 1e-4
    for k in g1:
        err = abs(g2[k] - g1[k]) / abs(g1[k])
        print(f"  mixed d/d{k}: rel {err:.2e}", flush=True)
        assert err < 1e-4, k


def test_set_force_field_matches_new_context():
    c = cases.get("tip64_p4")
    s = c.system
    p = A.water_params(s.model)
    p2 = {**p, "sigma": p["sigma"] * 1.01, "q_h": p["q_h"] * 0.98, "morse_a": p["morse_a"] * 1.02,
          "theta0": p["theta0"] + 1.0}
    s2 = A.system_with_params(s, p2)
    a = Engine(s, c.x, c.y, c.z)
    a.set_force_field(s2)
    b = Engine(s2, c.x, c.y, c.z)
    for g in (a, b):
        g.set_positions(c.x, c.y, c.z)
    ea, eb = a.energy_control(), b.energy_control()
    for k in abi.ENERGY_KEYS:
        assert abs(ea[k] - eb[k]) <= 1e-13 * max(abs(eb[k]), 1e-3), k
    # the bonded kernels add forces with atomics, so two contexts agree to
    # rounding, not bit for bit
    fa, fb = a.forces(), b.forces()
    assert np.max(np.abs(fa - fb)) <= 1e-13 * np.max(np.abs(fb))
