"""The force call as a differentiable torch function.

``PinyEnergy`` is a ``torch.autograd.Function`` whose forward is the
per-bead potential energy of the force call (``vintert + vintrat`` of each
bead as a classical configuration, the quantity ``PINYEmpirical.get_V``
returns per replica, python/piny/piny.pyx:361-384) and whose backward uses
the device's own gradients: the forces for the positions
(``get_dV_dx`` = -f, piny.pyx:386-424) and the parameter-gradient kernels
(``pmx_get_param_grad``) for the charges, LJ sigma/epsilon and the bond and
bend coefficients.  ``water_energy`` maps the physical parameters of a water
model (q_H, sigma, epsilon, Morse D/a/r0 or the harmonic bond constant, the
bend constant and angle) onto PINY's arrays with torch operations
(restating intra_coefs.c:62-122, 236-277 and the unit conversions of the
parameter files), so autograd carries the chain rule back to them.

The energies and gradients come from the HIP library; torch is only the
container.  Positions are in Bohr, [P, natm, 3], ghost sites included (their
coordinates are recomputed from the other sites, so their gradient is 0).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import abi
from . import water as w

BOHR, BOLTZ, KB_KCAL = w.BOHR, w.BOLTZ, w.KB_KCAL


def water_params(model: w.WaterModel, dtype=torch.float64) -> dict:
    """The physical parameters of a water model as scalar tensors: q_h (e),
    sigma (A), epsilon (kcal/mol), the bond (morse_d kcal/mol, morse_a 1/A
    and r0 A, or bond_k kcal/mol/A^2 and r0), k_theta (kcal/mol/rad^2) and
    theta0 (degrees)."""
    t = lambda v: torch.tensor(float(v), dtype=dtype)
    h = model.sites.index("H")
    p = {"q_h": t(model.charges[h]), "sigma": t(model.lj_sig), "epsilon": t(model.lj_eps_kcal)}
    if model.bond[0] == "morse":
        p.update(morse_d=t(model.bond[1]), morse_a=t(model.bond[2]), r0=t(model.bond[3]))
    else:
        p.update(bond_k=t(model.bond[1]), r0=t(model.bond[2]))
    p.update(k_theta=t(model.bend[1]), theta0=t(model.bend[2]))
    return p


def bond_coef(params: dict):
    """c_0..c_6 (Hartree/Bohr^m) and r_eq (Bohr) of the bond, intra_coefs.c:62-122
    (PINY replaces the Morse potential by its degree-6 Taylor polynomial)."""
    z = torch.zeros((), dtype=params["r0"].dtype)
    if "morse_d" in params:
        al = params["morse_a"] * BOHR
        d = (params["morse_d"] / KB_KCAL) / BOLTZ
        c = torch.stack([z, z, al ** 2 * d, -(al ** 3) * d, 7.0 * al ** 4 * d / 12.0, -(al ** 5) * d / 4.0,
                         62.0 * al ** 6 * d / 720.0])
    else:
        fk = params["bond_k"] / KB_KCAL * (BOHR * BOHR / BOLTZ)
        c = torch.stack([z, z, 0.5 * fk, z, z, z, z])
    return c, params["r0"] / BOHR


def bend_coef(params: dict):
    """Cosine and sine series of the harmonic bend, intra_coefs.c:236-277."""
    fk = (params["k_theta"] / KB_KCAL) / BOLTZ
    th0 = params["theta0"] * (math.pi / 180.0)
    cn1, cn2, cn3, cn4 = 49.0 / 72.0, 3.0 / 4.0, 3.0 / 40.0, 1.0 / 180.0
    cn5, cn6, cn7, cn8 = 3.0 / 2.0, 3.0 / 5.0, 1.0 / 10.0, 4.0 / 15.0
    cn9, cn10, cn11 = 8.0 / 45.0, 3.0 / 10.0, 1.0 / 30.0
    c2, c4, c6 = torch.cos(2.0 * th0), torch.cos(4.0 * th0), torch.cos(6.0 * th0)
    s2, s4, s6 = torch.sin(2.0 * th0), torch.sin(4.0 * th0), torch.sin(6.0 * th0)
    z = torch.zeros((), dtype=fk.dtype)
    h = 0.5 * fk
    c = torch.stack([h * (cn1 + cn2 * c2 + cn3 * c4 + cn4 * c6), z, h * (-cn5 * c2 - cn6 * c4 - cn7 * c6), z,
                     h * (cn6 * c4 + cn8 * c6), z, h * (-cn9 * c6)])
    s = torch.stack([z, z, h * (-cn5 * s2 + cn10 * s4 - cn11 * s6), z, h * (-cn6 * s4 + cn9 * s6), z,
                     h * (-cn9 * s6)])
    return c, s


def piny_arrays(system: w.System, params: dict) -> dict:
    """PINY's force-field arrays for `system` from the physical parameters:
    per-site charges (every charge scales with q_h, keeping the model's
    ratios, so molecules stay neutral), LJ sigma/eps of the O-O pair in Bohr
    and Hartree, bond and bend coefficients."""
    m = system.model
    h = m.sites.index("H")
    ratio = torch.tensor(np.array(m.charges) / m.charges[h], dtype=params["q_h"].dtype)
    q = (params["q_h"] * ratio).repeat(system.nmol)
    npair = system.ntype * (system.ntype + 1) // 2
    lj = torch.tensor(system.inter_label == abi.PMX_INTER_LJ)
    sig = torch.where(lj, params["sigma"] / BOHR, torch.zeros(npair, dtype=q.dtype))
    eps = torch.where(lj, (params["epsilon"] / KB_KCAL) / BOLTZ, torch.zeros(npair, dtype=q.dtype))
    bc, beq = bond_coef(params)
    ac, as_ = bend_coef(params)
    return {"q": q, "inter_sig": sig, "inter_eps": eps, "bond_c": bc[None, :], "bond_eq": beq[None],
            "bend_c": ac[None, :], "bend_s": as_[None, :]}


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


class PinyEnergy(torch.autograd.Function):
    """E [P] = per-bead potential energy of engine's box at positions pos
    [P, natm, 3] (Bohr) and force-field arrays q [natm], inter_sig and
    inter_eps [npair], bond_c [nbt, 7], bond_eq [nbt], bend_c and bend_s
    [nat, 7] (PINY units).  With P > 1 the force call is energy_control_pimd's
    (shared list, xmod image convention); the mean of E is then its
    vintert + vintrat."""

    @staticmethod
    def forward(ctx, engine, pos, q, sig, eps, bond_c, bond_eq, bend_c, bend_s):
        sysm = engine.system
        new = {"q": _np(q), "inter_sig": _np(sig), "inter_eps": _np(eps), "bond_c": _np(bond_c).reshape(-1),
               "bond_eq": _np(bond_eq).reshape(-1), "bend_c": _np(bend_c).reshape(-1),
               "bend_s": _np(bend_s).reshape(-1)}
        if any(not np.array_equal(np.asarray(getattr(sysm, k)).reshape(-1), v.reshape(-1)) for k, v in new.items()):
            engine.set_force_field(dataclasses.replace(sysm, **new))
        p = _np(pos)
        P, N = engine.P, engine.N
        if p.shape != (P, N, 3):
            raise ValueError(f"positions must be [{P}, {N}, 3], got {p.shape}")
        engine.set_positions(p[:, :, 0].reshape(-1), p[:, :, 1].reshape(-1), p[:, :, 2].reshape(-1))
        engine.energy_control(abi.PMX_EVAL_PIMD if P > 1 else 0)
        # PIMD forces are f/P: dE_b/dx_b = -P f
        f = engine.forces() * (P if P > 1 else 1)
        ctx.f = torch.as_tensor(f, dtype=pos.dtype, device=pos.device)
        ctx.dev = [t.device for t in (q, sig, eps, bond_c, bond_eq, bend_c, bend_s)]
        if any(ctx.needs_input_grad[2:]):
            g = engine.param_grad()
            ctx.g = {k: torch.as_tensor(v) for k, v in g.items()}
            energy = g["energy"]
        else:
            ctx.g = None
            energy = engine.bead_potential()
        return torch.as_tensor(energy, dtype=pos.dtype, device=pos.device)

    @staticmethod
    def backward(ctx, gE):
        grad_pos = -(gE[:, None, None].to(ctx.f.dtype) * ctx.f) if ctx.needs_input_grad[1] else None
        out = [None, grad_pos]
        g = ctx.g
        gc = gE.detach().to("cpu", torch.float64)
        for k, (key, idx) in enumerate((("dq", 2), ("dsig", 3), ("deps", 4), ("dbond_c", 5), ("dbond_eq", 6),
                                         ("dbend_c", 7), ("dbend_s", 8))):
            if ctx.needs_input_grad[idx] and g is not None:
                out.append(torch.tensordot(gc, g[key], dims=([0], [0])).to(ctx.dev[k]))
            else:
                out.append(None)
        return tuple(out)


def water_energy(engine, pos: torch.Tensor, params: dict) -> torch.Tensor:
    """Per-bead energies [P] of engine's water box as a function of the
    positions and the physical parameters of water_params()."""
    a = piny_arrays(engine.system, params)
    return PinyEnergy.apply(engine, pos, a["q"], a["inter_sig"], a["inter_eps"], a["bond_c"], a["bond_eq"],
                            a["bend_c"], a["bend_s"])


def system_with_params(system: w.System, params: dict) -> w.System:
    """A copy of `system` whose force-field arrays follow `params` (for a
    new context, or for finite differences)."""
    a = {k: v.detach().cpu().numpy().astype(np.float64) for k, v in piny_arrays(system, params).items()}
    return dataclasses.replace(system, q=a["q"], inter_sig=a["inter_sig"], inter_eps=a["inter_eps"],
                               bond_c=a["bond_c"].reshape(-1), bond_eq=a["bond_eq"].reshape(-1),
                               bend_c=a["bend_c"].reshape(-1), bend_s=a["bend_s"].reshape(-1))
