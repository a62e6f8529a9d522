"""PINY's on-disk formats for the force path: the input directory a PINY run
starts from, and the coordinate files it reads and writes.

* Dictionary files: sections ``~name[\\key{value} ...]``, the format of every
  PINY input file (the simulation input, the molecule set file, molecule
  parameter files and the bond, bend and inter databases); restated from
  the writer python/piny/tools.py:137-182 and the reader's grammar in
  src/interface/search_base and friends.
* ``read_input_dir`` builds, from such a directory, the arrays the library
  takes (a ``water.System``: types, charges, masses, the LJ parameters per
  type pair, bonds, bends, ghosts, the cell, the settings) and the initial
  bead positions, with PINY's parse-time unit conversions
  (control_inter_params.c:438-470, intra_coefs.c:30-375,
  close_intra_params.c:215, read_coord.c:37-400).  Molecules of any
  template are read; the molecule set file lists them in order.
* ``write_input_dir`` writes the same directory for a ``water.Config``.
* ``.initial`` files (restart_type initial, tools.py:42-67,
  read_coord.c:383-400, read_hmat.c:255-301): a header ``natm 1 P``, the
  P*natm positions in Angstrom, then the cell's three rows.
* Configuration files (``atm_pos_file`` .confp, velocities, forces, the
  centroid file; write_gen_header.c:30-200, output_md.c:1046-1090): a
  formatted header line, the molecule, residue and atom type names, one
  line per atom (mass, charge, molecule type and number, residue type and
  number, atom type), then per frame P*natm lines of x y z (Bohr) and the
  cell's three rows.

* Restart files (restart_type restart_all; write_dump_file_md,
  output_md.c:757-880, and write_dump_file_pimd, output_pimd.c:786-930):
  the bead positions with their type names, the cell, the velocities (mode
  velocities with beads), the thermostat chains' velocities (per bead
  chain with beads), the barostat's (zero in NVE and NVT), and the time
  step, degrees of freedom, ensemble and state point.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re

import numpy as np

from piny_amd import abi
from piny_amd import water as w

# ---------------------------------------------------------------------------
# dictionary files

_SECTION = re.compile(r"~([A-Za-z_0-9]+)\s*\[(.*?)\]", re.S)
_ITEM = re.compile(r"\\([A-Za-z_0-9]+)\s*\{([^}]*)\}")


def parse_dict_text(text: str) -> list:
    """The sections of a PINY dictionary file in file order, as
    ``[(name, {key: value})]`` with values as stripped strings."""
    out = []
    for m in _SECTION.finditer(text):
        items = {k: v.strip() for k, v in _ITEM.findall(m.group(2))}
        out.append((m.group(1), items))
    return out


def read_dict_file(path: str) -> list:
    with open(path) as f:
        return parse_dict_text(f.read())


def dict_text(sections) -> str:
    """Sections ``[(name, [(key, value), ...] or {key: value})]`` as PINY
    dictionary text, one key a line."""
    parts = []
    for name, items in sections:
        pairs = list(items.items()) if isinstance(items, dict) else list(items)
        if not pairs:
            continue
        body = "\n".join("    \\%s{%s}" % (k, v) for k, v in pairs)
        parts.append("~%s[\n%s]" % (name, body))
    return "\n\n".join(parts) + "\n"


def write_dict_file(path: str, sections) -> None:
    with open(path, "w") as f:
        f.write(dict_text(sections))


def _first(sections, name):
    for n, items in sections:
        if n == name:
            return items
    return {}


def _all(sections, name):
    return [items for n, items in sections if n == name]


def _on(v: str) -> bool:
    return v.strip().lower() in ("on", "yes", "1", "true")


# ---------------------------------------------------------------------------
# coordinate files

def read_initial(path: str):
    """Positions of a ``restart_type initial`` file: (x, y, z) in Bohr, each
    [P*natm] bead-major, the cell in Bohr as PINY's hmat[1..9]
    (column-major), and (natm, P)."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    head = lines[0]
    natm, P = int(head[0]), int(head[2])
    pos = np.array([[float(v) for v in ln[:3]] for ln in lines[1:1 + natm * P]], dtype=np.float64)
    if pos.shape[0] != natm * P:
        raise ValueError(f"{path}: {pos.shape[0]} positions, the header announces {natm * P}")
    rows = np.array([[float(v) for v in ln[:3]] for ln in lines[1 + natm * P:4 + natm * P]], dtype=np.float64)
    # row i holds hmat[1+i], hmat[4+i], hmat[7+i] (read_hmat.c:255-301)
    h = np.zeros(9)
    for i in range(3):
        h[i], h[3 + i], h[6 + i] = rows[i]
    pos = pos / w.BOHR
    return pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), h / w.BOHR, (natm, P)


def initial_text(x, y, z, hmat_bohr, natm: int, P: int) -> str:
    """A ``restart_type initial`` file for bead-major positions in Bohr."""
    lines = ["%d 1 %d" % (natm, P)]
    for i in range(natm * P):
        lines.append("%.15f %.15f %.15f" % (x[i] * w.BOHR, y[i] * w.BOHR, z[i] * w.BOHR))
    h = np.asarray(hmat_bohr, dtype=np.float64) * w.BOHR
    for r in range(3):
        lines.append("%.15f %.15f %.15f" % (h[r], h[3 + r], h[6 + r]))
    return "\n".join(lines) + "\n"


@dataclasses.dataclass
class ConfHeader:
    """The header of a PINY configuration file (write_gen_header.c:78-150)."""
    file_typ: str = "pos_file"     # pos_file, vel_file, force_file, cen_file
    ntime: int = 0
    dt: float = 0.0                # atomic time units
    write_freq: int = 1
    pi_beads: int = 1
    natm: int = 0
    nfree: int = 0
    ensemble: tuple = (1, 0, 0, 0, 0)  # nve nvt npt_i npt_f nst
    t_ext: float = 300.0
    pext: float = 0.0
    stens_ext: float = 0.0
    iperd: int = 3
    mol_types: tuple = ("water",)
    res_types: tuple = ("water",)
    atm_types: tuple = ()
    # per atom: mass, charge, mol type, mol number, res type, res number, atm type (1-based)
    atoms: list = dataclasses.field(default_factory=list)


def conf_header_for(system, file_typ="pos_file", dt=0.0, write_freq=1, ntime=0, mol_name="water",
                    type_names=None) -> ConfHeader:
    """A configuration-file header describing ``system``, whose sites come
    in molecules of one template (water.System)."""
    ns = len(system.model.sites) if system.model is not None else 1
    names = list(type_names) if type_names else []
    if not names:
        for n in (system.model.sites if system.model is not None else []):
            if n not in names:
                names.append(n)
    atoms = []
    for i in range(system.natm):
        atoms.append((float(system.mass[i]), float(system.q[i]), 1, i // ns + 1, 1, 1, int(system.atm_type[i]) + 1))
    return ConfHeader(file_typ=file_typ, ntime=ntime, dt=dt, write_freq=write_freq, pi_beads=system.pi_beads,
                      natm=system.natm, nfree=3 * (system.natm - len(system.ghost_atom)),
                      t_ext=system.settings.temperature, mol_types=(mol_name,), res_types=(mol_name,),
                      atm_types=tuple(names), atoms=atoms)


def conf_header_text(h: ConfHeader) -> str:
    e = h.ensemble
    line = "%d %s %s %d %.10g %d %d %d %d %d %d %d %d %d %d %d %d %.10g %.10g %.10g %d" % (
        0, "linux", h.file_typ, h.ntime, h.dt, h.write_freq, h.pi_beads, len(h.mol_types), len(h.res_types),
        len(h.atm_types), h.natm, h.nfree, e[0], e[1], e[2], e[3], e[4], h.t_ext, h.pext, h.stens_ext, h.iperd)
    out = [line]
    out += list(h.mol_types) + list(h.res_types) + list(h.atm_types)
    for a in h.atoms:
        out.append("%g %g %d %d %d %d %d" % a)
    return "\n".join(out) + "\n"


def conf_frame_text(x, y, z, hmat_bohr) -> str:
    """One frame: P*natm lines of x y z (Bohr, as PINY keeps them) and the
    cell's rows (output_md.c:1058-1066)."""
    lines = ["%.12g  %.12g  %.12g" % (x[i], y[i], z[i]) for i in range(len(x))]
    h = hmat_bohr
    for i in range(3):
        lines.append("%.13g %.13g %.13g" % (h[i], h[3 + i], h[6 + i]))
    return "\n".join(lines) + "\n"


class ConfWriter:
    """Appends frames to a formatted configuration file, header first."""

    def __init__(self, path: str, header: ConfHeader):
        self.path = path
        self.header = header
        with open(path, "w") as f:
            f.write(conf_header_text(header))

    def write(self, x, y, z, hmat_bohr):
        with open(self.path, "a") as f:
            f.write(conf_frame_text(x, y, z, hmat_bohr))


def read_conf(path: str):
    """A formatted configuration file: its ConfHeader and the frames as a
    list of (x, y, z, hmat) with bead-major [P*natm] arrays (one bead for a
    centroid file), Bohr."""
    with open(path) as f:
        lines = f.read().splitlines()
    it = iter(lines)
    head = next(it).split()
    h = ConfHeader(file_typ=head[2], ntime=int(head[3]), dt=float(head[4]), write_freq=int(head[5]),
                   pi_beads=int(head[6]), natm=int(head[10]), nfree=int(head[11]),
                   ensemble=tuple(int(v) for v in head[12:17]), t_ext=float(head[17]), pext=float(head[18]),
                   stens_ext=float(head[19]), iperd=int(head[20]))
    nmol, nres, ntyp = int(head[7]), int(head[8]), int(head[9])
    h.mol_types = tuple(next(it).strip() for _ in range(nmol))
    h.res_types = tuple(next(it).strip() for _ in range(nres))
    h.atm_types = tuple(next(it).strip() for _ in range(ntyp))
    for _ in range(h.natm):
        v = next(it).split()
        h.atoms.append((float(v[0]), float(v[1])) + tuple(int(u) for u in v[2:7]))
    P = 1 if h.file_typ == "cen_file" else h.pi_beads
    frames = []
    rest = [ln for ln in it if ln.strip()]
    per = P * h.natm + 3
    for k in range(0, len(rest) - per + 1, per):
        blk = np.array([[float(u) for u in ln.split()[:3]] for ln in rest[k:k + per]])
        hm = np.zeros(9)
        for i in range(3):
            hm[i], hm[3 + i], hm[6 + i] = blk[P * h.natm + i]
        frames.append((blk[:P * h.natm, 0].copy(), blk[:P * h.natm, 1].copy(), blk[:P * h.natm, 2].copy(), hm))
    return h, frames


# ---------------------------------------------------------------------------
# restart files

@dataclasses.dataclass
class Restart:
    """A restart_all file's contents (PINY's 1-based indices made 0-based;
    arrays bead-major [P][natm] like the library's)."""
    natm: int
    pi_beads: int
    itime: int = 0
    pos: np.ndarray = None          # [P, natm, 3] Bohr
    labels: list = None             # per atom: (atm_typ, res_typ, mol_typ, mol_num[, res_num])
    hmat: np.ndarray = None         # [9] column-major (PINY hmat[1..9])
    hmat_ewd: np.ndarray = None
    x_lnv: float = 0.0
    vel: np.ndarray = None          # [P, natm, 3] (mode velocities with beads)
    nhc_class_v: np.ndarray = None  # [len, num] therm_class.v_nhc
    nhc_bead_v: np.ndarray = None   # [P-1, len, num] therm_bead[2..P].v_nhc (beads only)
    vgmat: np.ndarray = None        # [9]
    v_lnv: float = 0.0
    v_vol_nhc: np.ndarray = None    # [len]
    dt: float = 0.0
    nfree: int = 0
    ensemble: tuple = (0, 1, 0, 0, 0)  # nve nvt npt_i npt_f nst
    t_ext: float = 300.0
    pext: float = 0.0
    stens_ext: float = 0.0
    nfree_terms: tuple = (0, 0, 0)  # nbond_free nbend_free ntors_free


def restart_labels(system, mol_name: str = "water", type_names=None):
    """The per-atom names of a restart file for a system of molecules of one
    template: atom type, residue type and molecule type (the molecule's name
    for both, as PINY names the single residue of a molecule without
    residues), the 1-based molecule number and residue number."""
    ns = len(system.model.sites)
    names = list(type_names) if type_names else []
    if not names:
        for n in system.model.sites:
            if n not in names:
                names.append(n)
    return [(names[int(system.atm_type[i])], mol_name, mol_name, i // ns + 1, 1) for i in range(system.natm)]


def restart_text(r: Restart) -> str:
    """The text write_dump_file_pimd (pi_beads > 1) or write_dump_file_md
    writes for r."""
    g = lambda v: "%.13g" % v  # noqa: E731
    P, N = r.pi_beads, r.natm
    out = ["natm_tot restart_typ itime pi_beads", "%d restart_all %d %d" % (N, r.itime, P),
           "atm pos, atm_typ, mol_typ mol_num"]
    for ip in range(P):
        for i in range(N):
            lab = r.labels[i]
            tail = "%s %s %s %d" % lab[:4] if P > 1 else "%s %s %s %d %d" % lab[:5]
            out.append("%s %s %s %s" % (g(r.pos[ip, i, 0]), g(r.pos[ip, i, 1]), g(r.pos[ip, i, 2]), tail))
    h, he = r.hmat, r.hmat_ewd
    out.append("h matrix")
    out += ["%s %s %s" % (g(h[k]), g(h[3 + k]), g(h[6 + k])) for k in range(3)]
    out.append("h matrix for Ewald setup")
    out += ["%s %s %s" % (g(he[k]), g(he[3 + k]), g(he[6 + k])) for k in range(3)]
    out += ["1/3 log(Vol)", g(r.x_lnv), "atm vel"]
    for ip in range(P):
        for i in range(N):
            out.append("%s %s %s" % (g(r.vel[ip, i, 0]), g(r.vel[ip, i, 1]), g(r.vel[ip, i, 2])))
    cl = np.atleast_2d(np.asarray(r.nhc_class_v, dtype=np.float64))
    lenc, numc = cl.shape if cl.size else (0, 0)
    out += ["number of atm nhc, length of nhc", "%d %d" % (numc, lenc), "atm nhc velocities"]
    out += [g(cl[j, k]) for j in range(lenc) for k in range(numc)]
    if P > 1:
        bd = np.asarray(r.nhc_bead_v)
        lenb, numb = bd.shape[1], bd.shape[2]
        out += ["number of bead nhc, length of nhc", "%d %d" % (numb, lenb), "bead nhc velocities"]
        out += ["%.10g" % bd[ip, j, k] for ip in range(P - 1) for j in range(lenb) for k in range(numb)]
    vg = r.vgmat
    out.append("vol velocities")
    out += ["%s %s %s" % (g(vg[k]), g(vg[3 + k]), g(vg[6 + k])) for k in range(3)]
    out += ["log(vol) velocity", g(r.v_lnv), "vol nhc velocities"]
    out += [g(v) for v in r.v_vol_nhc]
    e = r.ensemble
    out += ["dt=%s" % g(r.dt), "nfree=%d" % r.nfree,
            "nve=%d nvt=%d npt_i=%d npt_f=%d nst=%d" % tuple(e),
            "nbond_free=%d nbend_free=%d ntors_free=%d" % tuple(r.nfree_terms),
            "t_ext=%s,pext=%s,stens_ext=%s" % (g(r.t_ext), g(r.pext), g(r.stens_ext))]
    return "\n".join(out) + "\n"


def restart_for(system, pos, vel, itime=0, nhc_class_v=None, nhc_bead_v=None, len_nhc=None, dt=0.0,
                nfree=None, ensemble=None, mol_name="water") -> Restart:
    """A Restart of a system's state: positions and velocities [P, natm, 3]
    (mode velocities with beads), the chains' velocities (class [len, num],
    beads [P-1, len, num]; without them the NVE layout of no chains), in an
    NVT (beads) or NVE (classical) ensemble unless given."""
    P, N = system.pi_beads, system.natm
    h = np.asarray(system.hmat, dtype=np.float64).reshape(9)
    hm = h.reshape(3, 3).T  # columns are the box vectors: hmat[3 c + r] is row r, column c
    if ensemble is None:
        ensemble = (0, 1, 0, 0, 0) if nhc_class_v is not None else (1, 0, 0, 0, 0)
    if len_nhc is None:
        len_nhc = np.shape(nhc_class_v)[0] if nhc_class_v is not None else 0
    st = system.settings
    return Restart(natm=N, pi_beads=P, itime=itime, pos=np.asarray(pos, dtype=np.float64).reshape(P, N, 3),
                   labels=restart_labels(system, mol_name), hmat=h.copy(), hmat_ewd=h.copy(),
                   x_lnv=math.log(abs(np.linalg.det(hm))) / 3.0,
                   vel=np.asarray(vel, dtype=np.float64).reshape(P, N, 3),
                   nhc_class_v=(np.zeros((0, 0)) if nhc_class_v is None else np.asarray(nhc_class_v)),
                   nhc_bead_v=nhc_bead_v, vgmat=np.zeros(9), v_lnv=0.0, v_vol_nhc=np.zeros(len_nhc), dt=dt,
                   nfree=3 * (N - len(system.ghost_atom)) if nfree is None else nfree, ensemble=tuple(ensemble),
                   t_ext=st.temperature, pext=st.pressure * abi.PCONV)


def write_restart(path: str, r: Restart) -> None:
    with open(path, "w") as f:
        f.write(restart_text(r))


def read_restart(path: str) -> Restart:
    """A restart_all file written by PINY (or by write_restart)."""
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0

    def nxt():
        nonlocal k
        k += 1
        return lines[k - 1]

    nxt()
    head = nxt().split()
    N, itime, P = int(head[0]), int(head[2]), int(head[3])
    if head[1] != "restart_all":
        raise ValueError(f"{path}: restart type {head[1]}, not restart_all")
    nxt()
    pos = np.zeros((P, N, 3))
    labels = []
    for ip in range(P):
        for i in range(N):
            v = nxt().split()
            pos[ip, i] = [float(u) for u in v[:3]]
            if ip == 0:
                labels.append(tuple(v[3:6]) + tuple(int(u) for u in v[6:]))

    def mat():
        m = np.zeros(9)
        for r_ in range(3):
            a, b, c = (float(u) for u in nxt().split())
            m[r_], m[3 + r_], m[6 + r_] = a, b, c
        return m

    nxt()
    hmat = mat()
    nxt()
    hmat_ewd = mat()
    nxt()
    x_lnv = float(nxt())
    nxt()
    vel = np.zeros((P, N, 3))
    for ip in range(P):
        for i in range(N):
            vel[ip, i] = [float(u) for u in nxt().split()]
    nxt()
    numc, lenc = (int(u) for u in nxt().split())
    nxt()
    cl = np.array([float(nxt()) for _ in range(lenc * numc)]).reshape(lenc, numc)
    bead = None
    if P > 1:
        nxt()
        numb, lenb = (int(u) for u in nxt().split())
        nxt()
        bead = np.array([float(nxt()) for _ in range((P - 1) * lenb * numb)]).reshape(P - 1, lenb, numb)
    nxt()
    vg = mat()
    nxt()
    v_lnv = float(nxt())
    nxt()
    vvol = []
    while not lines[k].startswith("dt="):
        vvol.append(float(nxt()))
    dt = float(nxt().split("=")[1])
    nfree = int(nxt().split("=")[1])
    ens = tuple(int(t.split("=")[1]) for t in nxt().split())
    nft = tuple(int(t.split("=")[1]) for t in nxt().split())
    sp = dict(t.split("=") for t in nxt().split(","))
    return Restart(natm=N, pi_beads=P, itime=itime, pos=pos, labels=labels, hmat=hmat, hmat_ewd=hmat_ewd,
                   x_lnv=x_lnv, vel=vel, nhc_class_v=cl, nhc_bead_v=bead, vgmat=vg, v_lnv=v_lnv,
                   v_vol_nhc=np.array(vvol), dt=dt, nfree=nfree, ensemble=ens, t_ext=float(sp["t_ext"]),
                   pext=float(sp["pext"]), stens_ext=float(sp["stens_ext"]), nfree_terms=nft)


def nhc_state_to_restart(system, state, len_nhc):
    """The chain velocities of a library chain state (pmx_get_nhc_state's
    layout: the centroid chain, then per mode 2..P, site and direction) in a
    restart file's order: the one global class chain [len, 1], and per bead
    2..P the chains numbered 3 (real site) + direction, ghosts skipped
    (set_atm_NHC.c:648-700), [P-1, len, 3 real sites]."""
    P, N, L = system.pi_beads, system.natm, len_nhc
    st = np.asarray(state, dtype=np.float64)
    cls = st[:L].reshape(L, 1)
    ghost = np.zeros(N, bool)
    ghost[np.asarray(system.ghost_atom, dtype=int)] = True
    real = np.nonzero(~ghost)[0]
    bead = np.zeros((P - 1, L, 3 * len(real)))
    blocks = st[2 * L:].reshape(P - 1, N, 3, 2, L)
    for r_, i in enumerate(real):
        for d in range(3):
            bead[:, :, 3 * r_ + d] = blocks[:, i, d, 0, :]
    return cls, bead


# ---------------------------------------------------------------------------
# input directories

@dataclasses.dataclass
class MoleculeType:
    """One molecule template of a parameter file (molecule_name_def,
    atom_def, bond_def, bend_def)."""
    name: str
    sites: tuple          # atom type name per site
    charges: tuple
    masses_amu: tuple
    bonds: tuple          # ((i, j), ...) 0-based
    bends: tuple          # ((i, j, k), ...) 0-based, j the apex
    ghosts: tuple         # ((site, ((comp, coef), ...)), ...)


def read_parm(path: str) -> MoleculeType:
    secs = read_dict_file(path)
    name = _first(secs, "molecule_name_def").get("molecule_name", "molecule")
    atoms = sorted(_all(secs, "atom_def"), key=lambda a: int(a["atom_ind"]))
    sites = tuple(a["atom_typ"] for a in atoms)
    charges = tuple(float(a.get("charge", "0")) for a in atoms)
    masses = tuple(float(a["mass"]) for a in atoms)
    ghosts = []
    for k, a in enumerate(atoms):
        comps = []
        for g in range(1, 9):
            v = a.get("def_ghost%d" % g)
            if v is None:
                continue
            ci, cf = v.replace(",", " ").split()
            comps.append((int(ci) - 1, float(cf)))
        if comps:
            ghosts.append((k, tuple(comps)))
    bonds = tuple((int(b["atom1"]) - 1, int(b["atom2"]) - 1) for b in _all(secs, "bond_def"))
    bends = tuple((int(b["atom1"]) - 1, int(b["atom2"]) - 1, int(b["atom3"]) - 1) for b in _all(secs, "bend_def"))
    return MoleculeType(name, sites, charges, masses, bonds, bends, tuple(ghosts))


def bond_coef_input(kind: str, eq_A: float, fk_K=None, d0_K=None, alpha_invA=None):
    """c_0..c_6 (Hartree / Bohr^m) and the equilibrium length (Bohr) of a
    bond_parm entry in input units (intra_coefs.c:62-122): harmonic
    V = fk/2 (r - eq)^2 with fk in K/A^2, or PINY's degree-6 Taylor form of
    the Morse potential with d0 in K and alpha in 1/A."""
    c = np.zeros(7)
    if kind == "harm":
        c[2] = 0.5 * (fk_K * (w.BOHR * w.BOHR / w.BOLTZ))
    elif kind == "morse":
        al = alpha_invA * w.BOHR
        d = d0_K / w.BOLTZ
        c[2] = al * al * d
        c[3] = -(al * al * al) * d
        c[4] = 7.0 * (al * al * al * al) * d / 12.0
        c[5] = -(al * al * al * al * al) * d / 4.0
        c[6] = 62.0 * (al * al * al * al * al * al) * d / 720.0
    else:
        raise ValueError(f"bond pot_type {kind!r} is not supported (harm, morse)")
    return c, eq_A / w.BOHR


def bend_coef_input(fk_K: float, eq_deg: float):
    """The cosine/sine series of a harmonic bend_parm entry (fk_bend in
    K/rad^2, eq_bend in degrees), intra_coefs.c:236-277."""
    return w.bend_series(fk_K / w.BOLTZ, eq_deg * (math.pi / 180.0))


def _match(entries, names, keys):
    """The database entry whose atom names match `names` in order or reversed."""
    names = tuple(names)
    for e in entries:
        got = tuple(e[k] for k in keys)
        if got == names or got == names[::-1]:
            return e
    raise KeyError(f"no database entry for {names}")


def read_input_dir(directory: str, input_file: str = None, positions: bool = True):
    """The system of a PINY input directory as a ``water.Config``: the
    arrays and settings the library takes, and (with positions) the initial
    bead positions of the ``in_restart_file`` (restart_type initial).  The
    Config's system has ``model`` set to the directory's single molecule
    template when there is one (None otherwise)."""
    if input_file is None:
        cands = [f for f in os.listdir(directory) if f.endswith(".input")]
        if len(cands) != 1:
            raise ValueError(f"{directory}: name the input file ({cands})")
        input_file = cands[0]
    inp = read_dict_file(os.path.join(directory, input_file))
    gen, pimd = _first(inp, "sim_gen_def"), _first(inp, "sim_pimd_def")
    lst, run = _first(inp, "sim_list_def"), _first(inp, "sim_run_def")
    vol, pe = _first(inp, "sim_vol_def"), _first(inp, "sim_class_PE_def")
    wr = _first(inp, "sim_write_def")
    shift = {"swit": 2, "on": 1, "off": 0}  # shift_inter_pe (set_sim_params.c:1187-1197)
    st = w.Settings(
        skin=float(lst.get("verlist_skin", "1.0")),
        rheal=float(run.get("respa_rheal", "1.0")),
        kmax_ewd=int(pe.get("ewald_kmax", "14")),
        kmax_pme=int(pe.get("ewald_kmax_pme", "14")),
        n_interp=int(pe.get("ewald_interp_pme", "6")),
        nsplin=int(pe.get("inter_spline_pts", "2000")),
        ishift=shift[pe.get("shift_inter_pe", "swit").lower()],
        pme_on=1 if _on(pe.get("ewald_pme_opt", "off")) else 0,
        intra_perds=1 if _on(vol.get("intra_perds", "off")) else 0,
        temperature=float(gen.get("temperature", "300")),
        pimd_typ={"staging": 1, "centroid": 2, "cent": 2}[pimd.get("path_int_md_typ", "centroid").lower()],
        # RESPA (set_sim_params.c:1585-1660, set_sim_dict.c:749-756,921-940,1494-1498)
        respa_lrf=int(float(run.get("respa_steps_lrf", "0"))),
        respa_torsion=int(float(run.get("respa_steps_torsion", "0"))),
        respa_intra=int(float(run.get("respa_steps_intra", "0"))),
        respa_pimd=int(float(pimd.get("respa_steps_pimd", "0"))),
        pressure=float(gen.get("pressure", "0")),
        pten_inter_respa=float(pe.get("pten_inter_respa", "0")))
    P = int(pimd.get("path_int_beads", "1"))
    # molecules, in the set file's order of mol_index
    sets = read_dict_file(os.path.join(directory, wr["mol_set_file"]))
    db = _first(sets, "data_base_def")
    mols = sorted(_all(sets, "molecule_def"), key=lambda m: int(m.get("mol_index", "1")))
    mtypes = [(read_parm(os.path.join(directory, m["mol_parm_file"])), int(m["num_mol"])) for m in mols]
    bond_db = _all(read_dict_file(os.path.join(directory, db["bond_file"])), "bond_parm") if "bond_file" in db else []
    bend_db = _all(read_dict_file(os.path.join(directory, db["bend_file"])), "bend_parm") if "bend_file" in db else []
    inter_db = _all(read_dict_file(os.path.join(directory, db["inter_file"])), "inter_parm")
    # atom types in order of first appearance
    names = []
    for mt, _ in mtypes:
        for n in mt.sites:
            if n not in names:
                names.append(n)
    ntype = len(names)
    npair = ntype * (ntype + 1) // 2
    lab = np.full(npair, abi.PMX_INTER_NULL, dtype=np.int32)
    sig, eps = np.zeros(npair), np.zeros(npair)
    cutti, cutoff, cutres = np.zeros(npair), np.zeros(npair), np.zeros(npair)
    for a in range(ntype):
        for b in range(a, ntype):
            t = a * ntype - a * (a - 1) // 2 + (b - a)
            e = _match(inter_db, (names[a], names[b]), ("atom1", "atom2"))
            cutti[t] = float(e["min_dist"]) / w.BOHR
            cutoff[t] = float(e["max_dist"]) / w.BOHR
            cutres[t] = float(e.get("res_dist", e["max_dist"])) / w.BOHR
            if e["pot_type"].lower() == "lennard-jones":
                lab[t] = abi.PMX_INTER_LJ
                sig[t] = float(e["sig"]) / w.BOHR
                eps[t] = float(e["eps"]) / w.BOLTZ
            elif e["pot_type"].lower() != "null":
                raise ValueError(f"inter pot_type {e['pot_type']!r} is not supported (lennard-jones, null)")
    # sites, bonds, bends, ghosts of every molecule
    atm_type, q, mass = [], [], []
    bj1, bj2, btyp, aj1, aj2, aj3, atyp = [], [], [], [], [], [], []
    gatom, gncomp, gcomp, gcoef = [], [], [], []
    bond_types, bend_types = [], []
    base = 0
    for mt, nm in mtypes:
        ns = len(mt.sites)
        tb = []
        for (i, j) in mt.bonds:
            e = _match(bond_db, (mt.sites[i], mt.sites[j]), ("atom1", "atom2"))
            key = tuple(sorted(e.items()))
            if key not in bond_types:
                bond_types.append(key)
            tb.append(bond_types.index(key))
        ta = []
        for (i, j, k) in mt.bends:
            e = _match(bend_db, (mt.sites[i], mt.sites[j], mt.sites[k]), ("atom1", "atom2", "atom3"))
            key = tuple(sorted(e.items()))
            if key not in bend_types:
                bend_types.append(key)
            ta.append(bend_types.index(key))
        for m in range(nm):
            o = base + m * ns
            atm_type += [names.index(n) for n in mt.sites]
            q += list(mt.charges)
            mass += [v * w.PROT_MASS for v in mt.masses_amu]
            for (i, j), t in zip(mt.bonds, tb):
                bj1.append(o + i)
                bj2.append(o + j)
                btyp.append(t)
            for (i, j, k), t in zip(mt.bends, ta):
                aj1.append(o + i)
                aj2.append(o + j)
                aj3.append(o + k)
                atyp.append(t)
            for g, comps in mt.ghosts:
                gatom.append(o + g)
                gncomp.append(len(comps))
                cc, cf = [0] * 4, [0.0] * 4
                for k, (ci, coef) in enumerate(comps):
                    cc[k], cf[k] = o + ci, coef
                gcomp += cc
                gcoef += cf
        base += nm * ns
    bond_c, bond_eq = [], []
    for key in bond_types:
        e = dict(key)
        kind = e["pot_type"].lower()
        c, eq = bond_coef_input(kind, float(e["eq"]), fk_K=float(e["fk"]) if "fk" in e else None,
                                d0_K=float(e["d0"]) if "d0" in e else None,
                                alpha_invA=float(e["alpha"]) if "alpha" in e else None)
        bond_c.append(c)
        bond_eq.append(eq)
    bend_c, bend_s = [], []
    for key in bend_types:
        e = dict(key)
        if e.get("pot_type_bend", "harm").lower() != "harm":
            raise ValueError(f"bend pot_type {e['pot_type_bend']!r} is not supported (harm)")
        c, s = bend_coef_input(float(e["fk_bend"]), float(e["eq_bend"]))
        bend_c.append(c)
        bend_s.append(s)
    model = None
    if len(mtypes) == 1:
        mt = mtypes[0][0]
        model = w.WaterModel(name=mt.name, sites=mt.sites, charges=mt.charges, masses_amu=mt.masses_amu,
                             lj_sig=0.0, lj_eps_kcal=0.0, bond=("harm", 0.0, 0.0), bend=("harm", 0.0, 0.0),
                             ghost=mt.ghosts[0] if len(mt.ghosts) == 1 else None, r_oh=0.0, theta=0.0)
    nmol = sum(nm for _, nm in mtypes)
    s = w.System(model=model, nmol=nmol, box=0.0, pi_beads=P, settings=st)
    s.natm = len(atm_type)
    s.ntype = ntype
    s.atm_type = np.array(atm_type, dtype=np.int32)
    s.q = np.array(q, dtype=np.float64)
    s.mass = np.array(mass, dtype=np.float64)
    s.inter_label, s.inter_sig, s.inter_eps = lab, sig, eps
    s.inter_cutti, s.inter_cutoff, s.inter_cutoff_res = cutti, cutoff, cutres
    s.bond_j1, s.bond_j2 = np.array(bj1, dtype=np.int32), np.array(bj2, dtype=np.int32)
    s.bond_type = np.array(btyp, dtype=np.int32)
    s.bond_c = np.concatenate(bond_c) if bond_c else np.zeros(7)
    s.bond_eq = np.array(bond_eq) if bond_eq else np.zeros(1)
    s.bend_j1, s.bend_j2, s.bend_j3 = (np.array(v, dtype=np.int32) for v in (aj1, aj2, aj3))
    s.bend_type = np.array(atyp, dtype=np.int32)
    s.bend_c = np.concatenate(bend_c) if bend_c else np.zeros(7)
    s.bend_s = np.concatenate(bend_s) if bend_s else np.zeros(7)
    s.ghost_atom = np.array(gatom, dtype=np.int32)
    s.ghost_ncomp = np.array(gncomp, dtype=np.int32)
    s.ghost_comp = np.array(gcomp, dtype=np.int32)
    s.ghost_coef = np.array(gcoef, dtype=np.float64)
    x = y = z = None
    init = wr.get("in_restart_file")
    if positions and init and os.path.exists(os.path.join(directory, init)):
        x, y, z, h, (natm, Pi) = read_initial(os.path.join(directory, init))
        if natm != s.natm or Pi != P:
            raise ValueError(f"{init}: {natm} atoms and {Pi} beads, the input has {s.natm} and {P}")
        s.cell = h * w.BOHR
        if np.count_nonzero(h - np.diag(np.diag(h.reshape(3, 3))).reshape(9)) == 0 and \
                h[0] == h[4] == h[8]:
            s.box, s.cell = float(h[0] * w.BOHR), None
    # ewald_alpha is alpha times the cube root of the volume
    s.ewald_alpha_in = float(pe.get("ewald_alpha", "7.0"))
    s.settings = dataclasses.replace(st, alpha=s.ewald_alpha_in / s.side if s.side else st.alpha,
                                     cutoff=float(np.max(cutoff)) * w.BOHR, min_dist=float(np.min(cutti)) * w.BOHR,
                                     res_dist=float(np.max(cutres)) * w.BOHR)
    return w.Config(system=s, x=x, y=y, z=z)


def write_input_dir(directory: str, cfg, name: str = "water", nstep: int = 0, dt_fs: float = 0.25) -> None:
    """A PINY input directory for ``cfg`` (a water.Config whose system has one
    molecule template): <name>.input, .set, .parm, the bond, bend and inter
    databases and the .initial positions, in the units PINY reads
    (Angstrom, Kelvin, amu, degrees)."""
    os.makedirs(directory, exist_ok=True)
    s = cfg.system
    st = s.settings
    m = s.model
    P = s.pi_beads
    ns = len(m.sites)
    kb = w.KB_KCAL
    pimd_name = {1: "staging", 2: "centroid"}
    shift_name = {0: "off", 1: "on", 2: "swit"}
    gen = [("simulation_typ", "pimd" if P > 1 else "md"), ("ensemble_typ", "nvt" if P > 1 else "nve"),
           ("num_time_step", nstep), ("restart_type", "initial"), ("time_step", dt_fs),
           ("temperature", st.temperature), ("generic_fft_opt", "on"), ("num_proc_tot", 1), ("num_proc_beads", 1),
           ("num_proc_class_forc", 1)]
    if st.pressure:
        gen.append(("pressure", repr(st.pressure)))
    inp = [("sim_gen_def", gen)]
    if P > 1:
        pim = [("path_int_beads", P), ("path_int_md_typ", pimd_name[st.pimd_typ]), ("initial_spread_opt", "off")]
        if st.respa_pimd:
            pim.append(("respa_steps_pimd", st.respa_pimd))
        inp.append(("sim_pimd_def", pim))
        inp.append(("sim_nhc_def", [("atm_nhc_tau_def", 20.0), ("atm_nhc_len", 4), ("respa_steps_nhc", 4),
                                    ("yosh_steps_nhc", 3), ("respa_xi_opt", 1)]))
    inp += [("sim_list_def", [("neighbor_list", "ver_list"), ("verlist_skin", repr(st.skin)),
                              ("update_type", "no_list")]),
            ("sim_run_def", [("respa_rheal", repr(st.rheal))] +
             [(k, v) for k, v in (("respa_steps_lrf", st.respa_lrf), ("respa_steps_torsion", st.respa_torsion),
                                  ("respa_steps_intra", st.respa_intra)) if v]),
            ("sim_vol_def", [("periodicity", 3), ("intra_perds", "on" if st.intra_perds else "off")]),
            ("sim_class_PE_def", [("shift_inter_pe", shift_name[st.ishift]), ("inter_spline_pts", st.nsplin),
                                  ("ewald_alpha", repr(s.ewald_alpha)), ("ewald_kmax", st.kmax_ewd),
                                  ("ewald_pme_opt", "on" if st.pme_on else "off"), ("ewald_kmax_pme", st.kmax_pme),
                                  ("ewald_interp_pme", st.n_interp), ("ewald_respa_pme_opt", "off"),
                                  ("inter_PE_calc_freq", 1)] +
             ([("pten_inter_respa", repr(st.pten_inter_respa))] if st.pten_inter_respa else [])),
            ("sim_write_def", [("sim_name", name + ".input-all"), ("write_screen_freq", 1000000),
                               ("instant_file", name + ".iavg"), ("write_inst_freq", 1000000),
                               ("atm_pos_file", name + ".confp"), ("write_pos_freq", 1000000),
                               ("path_cent_file", name + "-centroid.confp"), ("path_cent_freq", 1000000),
                               ("atm_vel_file", name + ".confv"), ("write_vel_freq", 1000000),
                               ("atm_force_file", name + ".conff"), ("write_force_freq", 1000000),
                               ("out_restart_file", name + ".restart"), ("write_dump_freq", 1000000),
                               ("in_restart_file", name + ".initial"), ("mol_set_file", name + ".set"),
                               ("conf_file_format", "formatted")])]
    write_dict_file(os.path.join(directory, name + ".input"), inp)
    qs = np.asarray(s.q).reshape(s.nmol, ns)
    if not np.all(qs == qs[0]):
        raise ValueError("write_input_dir: every molecule must carry the same charges")
    ghost = {m.ghost[0]: m.ghost[1]} if m.ghost is not None else {}
    parm = [("molecule_name_def", [("molecule_name", name), ("natom", ns)])]
    for k, (site, qk, mk) in enumerate(zip(m.sites, qs[0], m.masses_amu)):
        items = [("atom_typ", site), ("atom_ind", k + 1), ("mass", repr(float(mk))), ("charge", repr(float(qk)))]
        for g, (comp, coef) in enumerate(ghost.get(k, ())):
            items.append(("def_ghost%d" % (g + 1), "%d, %r" % (comp + 1, coef)))
        parm.append(("atom_def", items))
    parm += [("bond_def", [("atom1", 1), ("atom2", 2)]), ("bond_def", [("atom1", 1), ("atom2", 3)]),
             ("bend_def", [("atom1", 2), ("atom2", 1), ("atom3", 3)])]
    write_dict_file(os.path.join(directory, name + ".parm"), parm)
    if m.bond[0] == "harm":
        bond = [("atom1", m.sites[0]), ("atom2", m.sites[1]), ("pot_type", "harm"), ("eq", repr(m.bond[2])),
                ("fk", repr(m.bond[1] / kb))]
    else:
        bond = [("atom1", m.sites[0]), ("atom2", m.sites[1]), ("pot_type", "morse"), ("eq", repr(m.bond[3])),
                ("alpha", repr(m.bond[2])), ("d0", repr(m.bond[1] / kb))]
    write_dict_file(os.path.join(directory, name + ".bond"), [("bond_parm", bond)])
    write_dict_file(os.path.join(directory, name + ".bend"), [("bend_parm", [
        ("atom1", m.sites[1]), ("atom2", m.sites[0]), ("atom3", m.sites[2]), ("pot_type_bend", "harm"),
        ("eq_bend", repr(m.bend[2])), ("fk_bend", repr(m.bend[1] / kb))])])
    types = []
    for n in m.sites:
        if n not in types:
            types.append(n)
    inter = []
    for a in range(len(types)):
        for b in range(a, len(types)):
            t = a * len(types) - a * (a - 1) // 2 + (b - a)
            items = [("atom1", types[a]), ("atom2", types[b])]
            if s.inter_label[t] == 2:
                items += [("pot_type", "lennard-jones"), ("eps", repr(float(s.inter_eps[t]) * w.BOLTZ)),
                          ("sig", repr(float(s.inter_sig[t]) * w.BOHR))]
            else:
                items += [("pot_type", "null")]
            items += [("min_dist", repr(float(s.inter_cutti[t]) * w.BOHR)),
                      ("max_dist", repr(float(s.inter_cutoff[t]) * w.BOHR)),
                      ("res_dist", repr(float(s.inter_cutoff_res[t]) * w.BOHR))]
            inter.append(("inter_parm", items))
    write_dict_file(os.path.join(directory, name + ".inter"), inter)
    write_dict_file(os.path.join(directory, name + ".set"), [
        ("molecule_def", [("mol_parm_file", name + ".parm"), ("mol_opt_nhc", "global"), ("num_mol", s.nmol),
                          ("mol_index", 1), ("mol_name", name)]),
        ("data_base_def", [("bond_file", name + ".bond"), ("bend_file", name + ".bend"),
                           ("inter_file", name + ".inter")])])
    with open(os.path.join(directory, name + ".initial"), "w") as f:
        f.write(initial_text(cfg.x, cfg.y, cfg.z, s.hmat, s.natm, P))
