"""The torch parameter mapping of the autograd Function (host only): the
coefficient formulas against the C ABI's coefficient helpers and the numpy
model module, and their torch derivatives against central differences."""
import ctypes as C

import numpy as np
import torch

from piny_amd import autograd as A
from piny_amd import water as w
from piny_amd.engine import lib


def test_coefficients_match_c_helpers():
    L = lib()
    for model in (w.SPC_FW, w.Q_TIP4P_F):
        p = A.water_params(model)
        bc, beq = A.bond_coef(p)
        c_np, eq_np = w.bond_coef(model)
        assert np.allclose(bc.numpy(), c_np, rtol=1e-15, atol=0) and abs(beq.item() - eq_np) < 1e-15
        ac, as_ = A.bend_coef(p)
        c7, s7 = (C.c_double * 7)(), (C.c_double * 7)()
        fk = (model.bend[1] / w.KB_KCAL) / w.BOLTZ
        L.pmx_bend_coef_harm(fk, model.bend[2] * np.pi / 180.0, c7, s7)
        assert np.allclose(ac.numpy(), np.array(c7[:]), rtol=1e-14, atol=1e-30)
        assert np.allclose(as_.numpy(), np.array(s7[:]), rtol=1e-14, atol=1e-30)
        if model.bond[0] == "morse":
            m7 = (C.c_double * 7)()
            L.pmx_bond_coef_morse((model.bond[1] / w.KB_KCAL) / w.BOLTZ, model.bond[2] * w.BOHR, m7)
            assert np.allclose(bc.numpy(), np.array(m7[:]), rtol=1e-14, atol=0)


def test_mapping_derivatives_vs_fd():
    s = w.make_system(w.Q_TIP4P_F, 4, 8.0)
    p = A.water_params(s.model)
    for k in p:
        v = p[k].clone().requires_grad_(True)
        a = A.piny_arrays(s, {**p, k: v})
        f = sum(t.sum() for t in a.values())
        f.backward()
        h = 1e-6 * abs(p[k].item())
        fp = sum(t.sum() for t in A.piny_arrays(s, {**p, k: p[k] + h}).values()).item()
        fm = sum(t.sum() for t in A.piny_arrays(s, {**p, k: p[k] - h}).values()).item()
        fd = (fp - fm) / (2 * h)
        assert abs(v.grad.item() - fd) <= 1e-6 * max(abs(fd), 1e-12), k


def test_charges_keep_molecules_neutral():
    for model in (w.SPC_FW, w.Q_TIP4P_F):
        s = w.make_system(model, 8, 10.0)
        p = A.water_params(model)
        q = A.piny_arrays(s, {**p, "q_h": p["q_h"] * 1.1})["q"]
        assert abs(q.sum().item()) < 1e-12
        assert np.allclose(A.piny_arrays(s, p)["q"].numpy(), s.q, rtol=1e-15, atol=1e-16)
