"""Parity of the HIP path with the oracle, through the C ABI, on the GPU.

Bars: Verlet lists, k-vectors, grid sizes and spline tables bit-exact;
energies, forces and virials in FP64 within 1e-9 relative (the north-star
bound for the FP64 path is a force RMS error below 1e-5 relative)."""
import dataclasses
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle
from piny_amd import abi
from piny_amd import water as w
from piny_amd.engine import Engine
from tests import cases

pytestmark = pytest.mark.gpu

ETOL = 1e-9
FTOL = 1e-9


def _run(case, flags_extra=0, **engine_kw):
    c = cases.get(case)
    s = c.system
    flags = (abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0) | flags_extra
    print(f"[case {case}] N={s.natm} P={s.pi_beads} {engine_kw}", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    eo, fo, _ = o.energy_control(c.x, c.y, c.z, flags)
    g = Engine(s, c.x, c.y, c.z, **engine_kw)
    g.set_positions(c.x, c.y, c.z)
    eg = g.energy_control(flags)
    fg = g.forces()
    return c, o, g, eo, fo, eg, fg


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("case", ["spc27", "spc64", "spc64_shift", "spc64_none", "tip64_p4", "cfg1", "cfg2_p8",
                                  "tip256_p32", "spc64_k20", "tip64_p4_k24", "cfg1_k40", "tip128_p4_k48",
                                  "spc64_ortho", "tip64_p4_tric", "tip64_p6", "spc27_p64", "spc64_noq",
                                  "spc64_ewald", "tip64_p4_ewald", "spc64_k80", "tip64_p4_k96", "tip64_p4_stag",
                                  "tip64_p4_k64"])
def test_energy_control_parity(case):
    c, o, g, eo, fo, eg, fg = _run(case)
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= ETOL * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    assert eg["nver"] == eo["nver"]
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    print(f"  force rms rel err {rms:.3e}", flush=True)
    assert rms < FTOL
    pv_scale = np.max(np.abs(eo["pvten"]))
    assert np.max(np.abs(eg["pvten"] - eo["pvten"])) <= ETOL * pv_scale
    if c.system.pi_beads > 1:
        assert _rel(eg["pi_ke_vir"], eo["pi_ke_vir"]) < ETOL
    assert np.allclose(g.bead_energies(), o.bead_energies(), rtol=1e-9, atol=1e-12)


# Direct evaluation of the tabulated functions differs from PINY's cubic
# interpolation by the tables' interpolation error: about 1e-10 relative for
# the energy table, but up to 1e-7 relative for the force table next to the
# switching onset r_c - r_heal, where the switch's second derivative jumps
# and the 4-point cubic across that kink is first-order accurate (the table
# equals the direct function at its nodes to 1e-15).  Bars: energies 1e-7,
# virial 1e-6 of its largest element, force RMS 2e-6 (north-star FP64 bound:
# 1e-5).  The mixed path (FP32 pair terms, FP64 sums) is held to the
# north-star bound of 1e-4 relative force RMS error and virial, and its
# energies to 1e-5 of the largest term.
DIRECT_ETOL, DIRECT_VTOL, DIRECT_FTOL = 1e-7, 1e-6, 2e-6
MIXED_ETOL, MIXED_VTOL, MIXED_FTOL = 1e-5, 1e-4, 1e-4


@pytest.mark.parametrize("mode", ["direct", "mixed"])
@pytest.mark.parametrize("case", ["spc64", "spc64_shift", "spc64_none", "tip64_p4", "cfg2_p8", "tip256_p32",
                                  "tip64_p4_tric", "tip64_p6", "spc27_p64", "spc64_noq", "tip64_p4_ewald",
                                  "spc64_k80", "tip64_p4_k96", "tip64_p4_k64", "tip64_p4_stag"])
def test_pair_direct_and_mixed(case, mode):
    kw = {"pair_eval": abi.PMX_PAIR_DIRECT} if mode == "direct" else {"precision": abi.PMX_PREC_MIXED}
    c, o, g, eo, fo, eg, fg = _run(case, **kw)
    etol, vtol, ftol = ((DIRECT_ETOL, DIRECT_VTOL, DIRECT_FTOL) if mode == "direct"
                        else (MIXED_ETOL, MIXED_VTOL, MIXED_FTOL))
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= etol * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    de = abs(eg["vintert"] - eo["vintert"]) / abs(eo["vintert"])
    print(f"  {mode}: force rms rel err {rms:.3e}, vintert rel err {de:.3e}", flush=True)
    assert rms < ftol
    pv_scale = np.max(np.abs(eo["pvten"]))
    assert np.max(np.abs(eg["pvten"] - eo["pvten"])) <= vtol * pv_scale


@pytest.mark.parametrize("case", ["spc64", "tip64_p4"])
def test_mixed_cutoff_per_type_pair(case):
    """Type pairs with cutoffs of their own: no one switched erfc table
    serves them all, so the mixed pair kernel switches the erfc term itself
    (with one cutoff it reads the switched term from its table, as every
    other mixed test does).  Against the oracle on the same system."""
    c = cases.get(case)
    cut = np.array(c.system.inter_cutoff, dtype=float).copy()
    cut[0] *= 0.92
    s = dataclasses.replace(c.system, inter_cutoff=cut)
    flags = abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0
    print(f"[case {case} with cutoffs {cut}]", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    eo, fo, _ = o.energy_control(c.x, c.y, c.z, flags)
    g = Engine(s, c.x, c.y, c.z, precision=abi.PMX_PREC_MIXED)
    g.set_positions(c.x, c.y, c.z)
    eg = g.energy_control(flags)
    fg = g.forces()
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= MIXED_ETOL * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    print(f"  force rms rel err {rms:.3e}", flush=True)
    assert rms < MIXED_FTOL
    pv_scale = np.max(np.abs(eo["pvten"]))
    assert np.max(np.abs(eg["pvten"] - eo["pvten"])) <= MIXED_VTOL * pv_scale


@pytest.mark.parametrize("mode", ["table", "mixed"])
@pytest.mark.parametrize("case", ["spc64", "tip64_p4", "cfg2_p8"])
def test_half_list_with_atomics(case, mode, monkeypatch):
    # the half-list force kernel (j forces scattered with atomics), selected
    # by the tuning switch, against the oracle
    monkeypatch.setenv("PMX_PAIR_LIST", "half")
    kw = {} if mode == "table" else {"precision": abi.PMX_PREC_MIXED}
    c, o, g, eo, fo, eg, fg = _run(case, **kw)
    etol, ftol = (ETOL, FTOL) if mode == "table" else (MIXED_ETOL, MIXED_FTOL)
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= etol * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    print(f"  half list {mode}: force rms rel err {rms:.3e}", flush=True)
    assert rms < ftol
    for a, b in zip(g.verlet_list(), o.verlet_list()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["spc512", "tip512_p4"])
def test_link_cell_list(case, monkeypatch):
    # the link-cell build (forced on; these boxes have 3 cells per axis)
    # gives the reference's rows in the reference's order
    monkeypatch.setenv("PMX_LIST", "cells")
    c, o, g, eo, fo, eg, fg = _run(case)
    for a, b in zip(g.verlet_list(), o.verlet_list()):
        assert np.array_equal(a, b)
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= ETOL * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    assert rms < FTOL


@pytest.mark.parametrize("case", ["spc64", "tip64_p4", "cfg1", "cfg2_p8", "spc512", "spc64_ortho", "tip64_p4_tric"])
def test_verlet_list_bit_exact(case):
    c, o, g, eo, fo, eg, fg = _run(case)
    for a, b in zip(g.verlet_list(), o.verlet_list()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["spc64", "cfg1", "tip256_p32"])
def test_setup_bit_exact(case):
    c = cases.get(case)
    s = c.system
    print(f"[case {case}] setup", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    g = Engine(s, c.x, c.y, c.z)
    assert np.array_equal(g.scalars()[:19], o.scalars()[:19])
    for a, b in zip(g.kvectors(), o.kvectors()):
        assert np.array_equal(a, b)
    npair = s.ntype * (s.ntype + 1) // 2
    for which in range(4):
        for ip in range(npair):
            assert np.array_equal(g.spline(which, ip), o.spline(which, ip)), (which, ip)
    for which in (4, 5):
        assert np.array_equal(g.spline(which), o.spline(which)), which


@pytest.mark.parametrize("case", ["tip64_p4", "cfg2_p8", "tip64_p6", "tip256_p32", "spc27_p64", "tip64_p4_stag",
                                  "cfg2_p4_stag"])
def test_mode_forces_and_springs(case):
    flags = abi.PMX_EVAL_MODE_FORCES
    c, o, g, eo, fo, eg, fg = _run(case, flags)
    assert _rel(eg["kin_harm"], eo["kin_harm"]) < 1e-10
    assert _rel(eg["pi_ke_prim"], eo["pi_ke_prim"]) < 1e-10
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    assert rms < FTOL


def test_list_rebuild_after_motion():
    c = cases.get("spc64")
    s = c.system
    print("[case spc64 rebuild]", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    g = Engine(s, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    g.energy_control()
    o.energy_control(c.x, c.y, c.z)
    rng = np.random.default_rng(5)
    x, y, z = (v + rng.normal(scale=0.05, size=v.shape) for v in (c.x, c.y, c.z))   # small: no rebuild
    g.set_positions(x, y, z)
    e1 = g.energy_control()
    eo1, fo1, _ = o.energy_control(x, y, z)
    assert e1["list_rebuilt"] == 0 and eo1["list_rebuilt"] == 0
    x2 = x.copy()
    x2[5] += 1.5  # more than half the skin: rebuild
    g.set_positions(x2, y, z)
    e2 = g.energy_control()
    eo2, fo2, _ = o.energy_control(x2, y, z)
    assert e2["list_rebuilt"] == 1 and eo2["list_rebuilt"] == 1
    assert _rel(e2["vintert"], eo2["vintert"]) < ETOL
    for a, b in zip(g.verlet_list(), o.verlet_list()):
        assert np.array_equal(a, b)


def test_list_rebuild_decision_with_beads():
    # the skin criterion of check_ver_list with beads: bead 1's displacement
    # since the last build against skin/2 + max(0, spread_now - spread)
    # (nbr_list_control.c:182-239), where the spread is the largest distance
    # of a bead from bead 1 (get_pimd_spread), checked call by call against
    # the oracle
    c = cases.get("tip64_p4")
    s = c.system
    N, P = s.natm, s.pi_beads
    skin = s.settings.skin / w.BOHR
    print("[case tip64_p4 rebuild with beads]", flush=True)
    o = Oracle(s, c.x, c.y, c.z)
    g = Engine(s, c.x, c.y, c.z)
    flags = abi.PMX_EVAL_PIMD
    pos = [c.x.copy(), c.y.copy(), c.z.copy()]

    def move(site, d, beads):
        for b in beads:
            pos[0][b * N + site] += d

    def call(expect):
        g.set_positions(*pos)
        eg = g.energy_control(flags)
        eo, fo, _ = o.energy_control(*[v.copy() for v in pos], flags)
        print(f"  rebuilt {eg['list_rebuilt']} (oracle {eo['list_rebuilt']}), spread {eg['spread']:.4f}", flush=True)
        assert eg["list_rebuilt"] == eo["list_rebuilt"] == expect
        assert eg["nver"] == eo["nver"]
        assert abs(eg["spread"] - eo["spread"]) <= 1e-12 * eo["spread"]
        assert _rel(eg["vintert"], eo["vintert"]) < ETOL
        fg = g.forces()
        assert np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum()) < FTOL
        return eg

    call(1)                                   # first build
    move(5, 0.7 * skin, range(P))             # the whole ring past skin/2: rebuild
    spread = call(1)["spread"]                # the spread of this build (over all sites and beads)
    # beads 2..P of site 9 at spread + skin/2 from its bead 1: the spread grows
    # by half the skin, so 0.7 skin of bead-1 motion is inside skin/2 + dspread
    for b in range(1, P):
        pos[0][b * N + 9] = pos[0][9] + spread + 0.5 * skin
        pos[1][b * N + 9] = pos[1][9]
        pos[2][b * N + 9] = pos[2][9]
    move(12, 0.7 * skin, range(P))
    call(0)
    move(12, 0.5 * skin, range(P))            # 1.2 skin is not
    call(1)
    for a, b in zip(g.verlet_list(), o.verlet_list()):
        assert np.array_equal(a, b)


def test_repeat_runs_reproduce():
    c = cases.get("cfg1")
    print("[case cfg1 repeat]", flush=True)
    g = Engine(c.system, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    e1 = g.energy_control()
    f1 = g.forces()
    e2 = g.energy_control()
    f2 = g.forces()
    assert np.max(np.abs(f1 - f2)) <= 1e-12 * np.max(np.abs(f1))
    assert abs(e1["vintert"] - e2["vintert"]) <= 1e-12 * abs(e1["vintert"])


# --- the device against the reference itself (tests/golden/ref_*.npz) -------

from tests.test_ref_setup import SETUP, compare_setup, load_setup  # noqa: E402


@pytest.mark.parametrize("name", SETUP)
def test_device_setup_matches_reference(name):
    # the reference's own Verlet list, spline tables, k-vectors, B-spline
    # weights and excluded-pair table, bit for bit
    d = load_setup(name)
    c = cases.get(name)
    s = c.system
    print(f"[reference setup {name}] N={s.natm} P={s.pi_beads}", flush=True)
    g = Engine(s, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    g.energy_control(abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0)
    compare_setup(d, g.verlet_list(), g.spline, g.kvectors(), s.ntype * (s.ntype + 1) // 2)

from tests.test_ref_golden import FIXTURES, compare, load  # noqa: E402


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("name", FIXTURES)
def test_device_matches_reference(name, mode):
    d, ref = load(name)
    c = cases.get(name)
    s = c.system
    print(f"[reference fixture {name}] N={s.natm} P={s.pi_beads} {mode}", flush=True)
    kw = {} if mode == "fp64" else {"precision": abi.PMX_PREC_MIXED}
    g = Engine(s, c.x, c.y, c.z, **kw)
    g.set_positions(c.x, c.y, c.z)
    flags = (abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES) if s.pi_beads > 1 else 0
    e = g.energy_control(flags)
    f = g.forces()
    etol, ftol, vtol = (ETOL, FTOL, ETOL) if mode == "fp64" else (MIXED_ETOL, MIXED_FTOL, MIXED_VTOL)
    rms = compare(name, e, f, d, ref, etol, ftol, vtol, term_scale=(mode == "mixed"))
    de = abs(e["vintert"] - ref["vintert"]) / abs(ref["vintert"])
    print(f"  force rms rel err vs the reference {rms:.3e}, vintert rel err {de:.3e}", flush=True)


# --- velocity Verlet on the device -------------------------------------------

from tests.test_md_nve import load as load_md  # noqa: E402


def test_device_nve_matches_reference():
    d = load_md()
    c = cases.get(str(d["case"]))
    s = c.system
    print(f"[md {d['case']}] {int(d['nsteps'])} NVE steps", flush=True)
    g = Engine(s, c.x, c.y, c.z)
    g.set_positions(c.x, c.y, c.z)
    v0 = d["v0"]
    g.set_velocities(v0[:, 0].copy(), v0[:, 1].copy(), v0[:, 2].copy())
    e, ke = g.md_nve(int(d["nsteps"]), float(d["dt"]))
    x, v, f = g.positions()[0], g.velocities()[0], g.forces()[0]
    print(f"  x {np.abs(x - d['x']).max():.2e}  v {np.abs(v - d['v']).max():.2e}  ke {ke} vs {float(d['kinet'])}",
          flush=True)
    assert np.max(np.abs(x - d["x"])) < 1e-9
    assert np.max(np.abs(v - d["v"])) < 1e-9 * np.max(np.abs(d["v"]))
    assert abs(e["vintert"] - float(d["vintert"])) < 1e-9 * abs(float(d["vintert"]))
    assert abs(ke - float(d["kinet"])) < 1e-9 * float(d["kinet"])


def test_mixed_energy_conservation_within_2x_of_fp64():
    # total energy vintert + vintrat + kinetic along 1000 steps of 0.25 fs
    # (configuration 1), in FP64 and in the mixed precision
    d = load_md()
    c = cases.get(str(d["case"]))
    v0 = d["v0"]
    spread = {}
    for mode, kw in (("fp64", {}), ("mixed", {"precision": abi.PMX_PREC_MIXED})):
        g = Engine(c.system, c.x, c.y, c.z, **kw)
        g.set_positions(c.x, c.y, c.z)
        g.set_velocities(v0[:, 0].copy(), v0[:, 1].copy(), v0[:, 2].copy())
        tot = []
        for _ in range(20):
            e, ke = g.md_nve(50, float(d["dt"]))
            tot.append(e["vintert"] + e["vintrat"] + ke)
        tot = np.array(tot)
        spread[mode] = np.max(np.abs(tot - tot[0]))
        print(f"  {mode}: max |E(t) - E(0)| = {spread[mode]:.3e} Ha over 1000 steps", flush=True)
    assert spread["mixed"] <= 2.0 * spread["fp64"] + 1e-6


def test_ewald_matches_pme_limit():
    # standard Ewald against PME on a fine grid of high order: the k-space
    # totals agree to PME's interpolation error
    import dataclasses
    c = cases.get("spc64_ewald")
    s = c.system
    print("[ewald vs pme]", flush=True)
    ge = Engine(s, c.x, c.y, c.z)
    ge.set_positions(c.x, c.y, c.z)
    ee = ge.energy_control()
    st = dataclasses.replace(s.settings, pme_on=1, kmax_pme=14, n_interp=10)
    sp = dataclasses.replace(s, settings=st)
    gp = Engine(sp, c.x, c.y, c.z)
    gp.set_positions(c.x, c.y, c.z)
    ep = gp.energy_control()
    assert abs(ep["vrecip"] - ee["vrecip"]) < 1e-6 * abs(ee["vrecip"]), (ep["vrecip"], ee["vrecip"])
    fe, fp = ge.forces(), gp.forces()
    assert np.sqrt(((fe - fp) ** 2).sum() / (fe ** 2).sum()) < 1e-4


@pytest.mark.parametrize("tiles", ["32,16,16", "4,4,2", "8,16,4"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_pme_tilings(mode, tiles, monkeypatch):
    # tiles whose halo box needs more than 64 KB of LDS (32 x 16 x 16 on a
    # 32^3 grid: 37 x 21 x 21 points, 130 KB in FP64), and tiles smaller
    # than the order-6 support, whose points overlap up to 3 or 4 halo boxes
    # per axis (4 x 4 x 2, 8 x 16 x 4)
    monkeypatch.setenv("PMX_PME_TILE", tiles)
    print(f"[pme tiles {tiles} {mode}]", flush=True)
    kw = {} if mode == "fp64" else {"precision": abi.PMX_PREC_MIXED}
    c, o, g, eo, fo, eg, fg = _run("cfg1", **kw)
    etol, ftol = (ETOL, FTOL) if mode == "fp64" else (MIXED_ETOL, MIXED_FTOL)
    scale = max(abs(eo["vintert"]), abs(eo["vintrat"]), 1e-3)
    for k in abi.ENERGY_KEYS:
        assert abs(eg[k] - eo[k]) <= etol * max(abs(eo[k]), scale), (k, eg[k], eo[k])
    rms = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    assert rms < ftol


@pytest.mark.parametrize("case", ["cfg1", "tip64_p4", "tip256_p32"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_pme_beside_pair_kernel(case, mode, monkeypatch):
    # PME on the context's second stream, concurrent with the pair kernel:
    # the same results as the serial sequence (forces up to the order of
    # the bonded kernels' atomic additions), checked over two calls with
    # the positions moved in between
    kw = {} if mode == "fp64" else {"precision": abi.PMX_PREC_MIXED}
    c = cases.get(case)
    s = c.system
    flags = abi.PMX_EVAL_PIMD if s.pi_beads > 1 else 0
    print(f"[pme stream {case} {mode}]", flush=True)
    res = []
    for ov in ("0", "1"):
        monkeypatch.setenv("PMX_OVERLAP", ov)
        g = Engine(s, c.x, c.y, c.z, **kw)
        out = []
        for call in range(2):
            d = 0.01 * call
            g.set_positions(c.x + d, c.y - d, c.z + 0.5 * d)
            e = g.energy_control(flags)
            out.append((e, g.forces()))
        res.append(out)
    tol = 1e-12 if mode == "fp64" else 1e-6
    for (e0, f0), (e1, f1) in zip(*res):
        for k in abi.ENERGY_KEYS:
            assert abs(e1[k] - e0[k]) <= tol * max(abs(e0[k]), 1e-3), (k, e1[k], e0[k])
        assert np.max(np.abs(f1 - f0)) <= tol * np.max(np.abs(f0))


@pytest.mark.parametrize("case", ["spc64", "tip64_p4", "tip256_p32", "spc64_ewald", "spc64_noq"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_molecule_pass_matches_term_kernels(case, mode, monkeypatch):
    # the one-pass molecule kernel (bonded terms, excluded pairs, PIMD
    # finish and ghosts per (molecule, bead), no atomics) against the
    # per-term kernels with global atomics, over two calls with the
    # positions moved in between; both are also checked against the oracle
    # by the other tests (the molecule pass is the default for water)
    kw = {} if mode == "fp64" else {"precision": abi.PMX_PREC_MIXED}
    c = cases.get(case)
    s = c.system
    flags = (abi.PMX_EVAL_PIMD | abi.PMX_EVAL_MODE_FORCES) if s.pi_beads > 1 else 0
    print(f"[molecule pass {case} {mode}]", flush=True)
    res = []
    for how in ("terms", "molecules"):
        monkeypatch.setenv("PMX_INTRA", how)
        g = Engine(s, c.x, c.y, c.z, **kw)
        out = []
        for call in range(2):
            d = 0.01 * call
            g.set_positions(c.x + d, c.y - d, c.z + 0.5 * d)
            e = g.energy_control(flags)
            out.append((e, g.forces()))
        res.append(out)
    tol = 1e-12 if mode == "fp64" else 1e-6
    for (e0, f0), (e1, f1) in zip(*res):
        for k in abi.ENERGY_KEYS:
            assert abs(e1[k] - e0[k]) <= tol * max(abs(e0[k]), 1e-3), (k, e1[k], e0[k])
        assert np.max(np.abs(f1 - f0)) <= tol * np.max(np.abs(f0))
