"""Water models and seeded synthetic boxes in PINY's internal units.

PINY converts its input (Angstrom, Kelvin, amu) at parse time; the same
conversions are applied here so that the arrays handed to the library are
the values PINY itself would hold after ``parse()``:

* lengths  / BOHR                      (defines.h:78, e.g. control_inter_params.c:438-440)
* energies in K / BOLTZ                 (defines.h:80, control_inter_params.c:466)
* masses   * PROT_MASS                  (close_intra_params.c:215)
* kcal/mol -> K with kB = 0.001987204118 kcal/mol/K
  (examples/q-TIP4P-F/water-q-TIP4P-F.py:1)
* Morse alpha * BOHR, D / BOLTZ         (intra_coefs.c:92-99)
* harmonic bond fk * BOHR^2 / BOLTZ     (intra_coefs.c:64-66)
* harmonic bend fk / BOLTZ, theta0 in radians (intra_coefs.c:236-241)
* ewald_alpha is dimensionless alpha*L  (control_set_cp_ewald.c:162-163)

Topology follows the reference's water definitions: SPC/Fw sites O,H,H
(examples/q-SPC-Fw) and q-TIP4P/F sites O,H,H,M with the massless M site
M = 0.73612 O + 0.13194 H1 + 0.13194 H2 (examples/q-TIP4P-F/water-q-TIP4P-F.py:28-35).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BOHR = 0.529177
BOLTZ = 315777.0
PROT_MASS = 1822.8885
KB_KCAL = 0.001987204118
HBAR = 1.0  # atomic units

SEED_BASE = 20261202


@dataclasses.dataclass
class WaterModel:
    name: str
    sites: tuple            # site names per molecule
    charges: tuple
    masses_amu: tuple
    lj_sig: float           # O-O, Angstrom
    lj_eps_kcal: float      # O-O, kcal/mol
    bond: tuple             # ("harm", k_kcal_A2, r0) or ("morse", D_kcal, a_invA, r0)
    bend: tuple             # ("harm", k_kcal_rad2, theta0_deg)
    ghost: tuple | None     # (site index, ((comp, coef), ...))
    r_oh: float             # geometry used to build molecules, Angstrom
    theta: float            # degrees


SPC_FW = WaterModel(
    name="SPC/Fw", sites=("O", "H", "H"), charges=(-0.82, 0.41, 0.41),
    masses_amu=(15.9994, 1.008, 1.008), lj_sig=3.165492, lj_eps_kcal=0.1554253,
    bond=("harm", 1059.162, 1.012), bend=("harm", 75.90, 113.24), ghost=None,
    r_oh=1.012, theta=113.24)

Q_TIP4P_F = WaterModel(
    name="q-TIP4P/F", sites=("O", "H", "H", "M"), charges=(0.0, 0.5564, 0.5564, -1.1128),
    masses_amu=(15.9994, 1.008, 1.008, 1.0), lj_sig=3.1589, lj_eps_kcal=0.1852,
    bond=("morse", 116.09, 2.287, 0.9419), bend=("harm", 87.85, 107.4),
    ghost=(3, ((0, 0.73612), (1, 0.13194), (2, 0.13194))), r_oh=0.9419, theta=107.4)


@dataclasses.dataclass
class Settings:
    """PINY input keywords (sim_class_PE_def / sim_list_def / sim_pimd_def)."""
    cutoff: float = 9.0          # max_dist, Angstrom
    min_dist: float = 0.1        # min_dist (spline start), Angstrom
    res_dist: float = 5.0        # res_dist, Angstrom
    skin: float = 1.0            # verlist_skin, Angstrom
    rheal: float = 1.0           # respa_rheal, Angstrom
    alpha: float = 0.35          # Ewald alpha, 1/Angstrom
    kmax_ewd: int = 14
    kmax_pme: int = 14
    n_interp: int = 6
    nsplin: int = 5000           # inter_spline_pts (the examples use 5000)
    ishift: int = 2              # shift_inter_pe = swit
    pme_on: int = 1
    intra_perds: int = 0         # default off (set_sim_dict.c:1231-1234)
    temperature: float = 300.0
    pimd_typ: int = 2            # centroid (normal modes)
    # RESPA (multiple time steps): 0 leaves a keyword out, which turns that
    # split off (set_sim_params.c:1585-1640)
    respa_lrf: int = 0           # respa_steps_lrf: the inter split (short range inside res_dist)
    respa_torsion: int = 0       # respa_steps_torsion
    respa_intra: int = 0         # respa_steps_intra
    respa_pimd: int = 0          # respa_steps_pimd
    pressure: float = 0.0        # pressure, atm (enters the RESPA pressure-tensor terms)
    pten_inter_respa: float = 0.0  # pten_inter_respa, atm


def bond_coef(model: WaterModel):
    """Polynomial coefficients c_0..c_6 (Hartree/Bohr^m) and r_eq (Bohr),
    restating intra_coefs.c:62-122."""
    kind = model.bond[0]
    c = np.zeros(7)
    if kind == "harm":
        fk = model.bond[1] / KB_KCAL * (BOHR * BOHR / BOLTZ)
        c[2] = 0.5 * fk
        eq = model.bond[2] / BOHR
    elif kind == "morse":
        al = model.bond[2] * BOHR
        d = (model.bond[1] / KB_KCAL) / BOLTZ
        a2, a3, a4 = al * al, al * al * al, al * al * al * al
        a5, a6 = al * al * al * al * al, al * al * al * al * al * al
        c[2] = a2 * d
        c[3] = -a3 * d
        c[4] = 7.0 * a4 * d / 12.0
        c[5] = -a5 * d / 4.0
        c[6] = 62.0 * a6 * d / 720.0
        eq = model.bond[3] / BOHR
    else:
        raise ValueError(kind)
    return c, eq


def bend_coef(model: WaterModel):
    """Cosine/sine series of the harmonic bend, intra_coefs.c:236-277."""
    return bend_series((model.bend[1] / KB_KCAL) / BOLTZ, model.bend[2] * (math.pi / 180.0))


def bend_series(fk: float, th0: float):
    """c_m, s_m of the harmonic bend 1/2 fk (theta - th0)^2 (fk in Hartree,
    th0 in radians) as PINY expands it in cos and sin (intra_coefs.c:236-277)."""
    cn1, cn2, cn3, cn4 = 49.0 / 72.0, 3.0 / 4.0, 3.0 / 40.0, 1.0 / 180.0
    cn5, cn6, cn7, cn8 = 3.0 / 2.0, 3.0 / 5.0, 1.0 / 10.0, 4.0 / 15.0
    cn9, cn10, cn11 = 8.0 / 45.0, 3.0 / 10.0, 1.0 / 30.0
    c2, c4, c6 = math.cos(2.0 * th0), math.cos(4.0 * th0), math.cos(6.0 * th0)
    s2, s4, s6 = math.sin(2.0 * th0), math.sin(4.0 * th0), math.sin(6.0 * th0)
    c = np.zeros(7)
    s = np.zeros(7)
    c[0] = 0.5 * fk * (cn1 + cn2 * c2 + cn3 * c4 + cn4 * c6)
    c[2] = 0.5 * fk * (-cn5 * c2 - cn6 * c4 - cn7 * c6)
    c[4] = 0.5 * fk * (cn6 * c4 + cn8 * c6)
    c[6] = 0.5 * fk * (-cn9 * c6)
    s[2] = 0.5 * fk * (-cn5 * s2 + cn10 * s4 - cn11 * s6)
    s[4] = 0.5 * fk * (-cn6 * s4 + cn9 * s6)
    s[6] = 0.5 * fk * (-cn9 * s6)
    return c, s


@dataclasses.dataclass
class System:
    """Arrays of a pmx_system plus the settings of a pmx_params."""
    model: WaterModel
    nmol: int
    box: float                 # Angstrom (cubic)
    pi_beads: int
    settings: Settings
    natm: int = 0
    ntype: int = 0
    atm_type: np.ndarray = None
    q: np.ndarray = None
    mass: np.ndarray = None
    inter_label: np.ndarray = None
    inter_sig: np.ndarray = None
    inter_eps: np.ndarray = None
    inter_cutti: np.ndarray = None
    inter_cutoff: np.ndarray = None
    inter_cutoff_res: np.ndarray = None
    bond_j1: np.ndarray = None
    bond_j2: np.ndarray = None
    bond_type: np.ndarray = None
    bond_c: np.ndarray = None
    bond_eq: np.ndarray = None
    bend_j1: np.ndarray = None
    bend_j2: np.ndarray = None
    bend_j3: np.ndarray = None
    bend_type: np.ndarray = None
    bend_c: np.ndarray = None
    bend_s: np.ndarray = None
    ghost_atom: np.ndarray = None
    ghost_ncomp: np.ndarray = None
    ghost_comp: np.ndarray = None
    ghost_coef: np.ndarray = None
    # a general cell (column-major like PINY's hmat[1..9], cell vectors as
    # columns, Angstrom) instead of the cube of edge `box`
    cell: np.ndarray = None
    # the ewald_alpha keyword as read from an input file (alpha times the
    # cube root of the volume), kept to the bit; None: settings.alpha * side
    ewald_alpha_in: float = None

    @property
    def hmat(self):
        """The cell in Bohr, column-major (PINY's hmat[1..9])."""
        if self.cell is not None:
            return np.asarray(self.cell, dtype=np.float64) / BOHR
        h = np.zeros(9)
        h[0] = h[4] = h[8] = self.box / BOHR
        return h

    @property
    def side(self):
        """Cube root of the cell volume, Angstrom."""
        if self.cell is None:
            return self.box
        return float(abs(np.linalg.det(np.asarray(self.cell, dtype=np.float64).reshape(3, 3).T))) ** (1.0 / 3.0)

    @property
    def ewald_alpha(self):
        if self.ewald_alpha_in is not None:
            return self.ewald_alpha_in
        return self.settings.alpha * self.side


def make_system(model: WaterModel, nmol: int, box: float, pi_beads: int = 1,
                settings: Settings | None = None) -> System:
    """Topology, types and force-field arrays for nmol molecules."""
    st = settings or Settings()
    ns = len(model.sites)
    type_names = []
    for s in model.sites:
        if s not in type_names:
            type_names.append(s)
    ntype = len(type_names)
    tmol = np.array([type_names.index(s) for s in model.sites], dtype=np.int32)
    natm = ns * nmol
    atm_type = np.tile(tmol, nmol).astype(np.int32)
    q = np.tile(np.array(model.charges), nmol)
    mass = np.tile(np.array(model.masses_amu) * PROT_MASS, nmol)
    npair = ntype * (ntype + 1) // 2
    lab = np.full(npair, 4, dtype=np.int32)
    sig = np.zeros(npair)
    eps = np.zeros(npair)
    o = type_names.index("O")
    ipo = o * ntype - o * (o - 1) // 2
    lab[ipo] = 2
    sig[ipo] = model.lj_sig / BOHR
    eps[ipo] = (model.lj_eps_kcal / KB_KCAL) / BOLTZ
    cutti = np.full(npair, st.min_dist / BOHR)
    cutoff = np.full(npair, st.cutoff / BOHR)
    cutres = np.full(npair, st.res_dist / BOHR)
    base = np.arange(nmol, dtype=np.int32) * ns
    # bonds listed per molecule as the parm files do: (O,H1), (O,H2)
    bond_j1 = np.stack([base, base], 1).reshape(-1).astype(np.int32)
    bond_j2 = np.stack([base + 1, base + 2], 1).reshape(-1).astype(np.int32)
    bc, beq = bond_coef(model)
    ac, as_ = bend_coef(model)
    sys_ = System(model=model, nmol=nmol, box=box, pi_beads=pi_beads, settings=st)
    sys_.natm = natm
    sys_.ntype = ntype
    sys_.atm_type = atm_type
    sys_.q = q.astype(np.float64)
    sys_.mass = mass.astype(np.float64)
    sys_.inter_label = lab
    sys_.inter_sig = sig
    sys_.inter_eps = eps
    sys_.inter_cutti = cutti
    sys_.inter_cutoff = cutoff
    sys_.inter_cutoff_res = cutres
    sys_.bond_j1 = bond_j1
    sys_.bond_j2 = bond_j2
    sys_.bond_type = np.zeros(2 * nmol, dtype=np.int32)
    sys_.bond_c = bc.copy()
    sys_.bond_eq = np.array([beq])
    sys_.bend_j1 = (base + 1).astype(np.int32)
    sys_.bend_j2 = (base + 0).astype(np.int32)
    sys_.bend_j3 = (base + 2).astype(np.int32)
    sys_.bend_type = np.zeros(nmol, dtype=np.int32)
    sys_.bend_c = ac.copy()
    sys_.bend_s = as_.copy()
    if model.ghost is not None:
        gsite, comps = model.ghost
        sys_.ghost_atom = (base + gsite).astype(np.int32)
        sys_.ghost_ncomp = np.full(nmol, len(comps), dtype=np.int32)
        gc = np.zeros((nmol, 4), dtype=np.int32)
        gf = np.zeros((nmol, 4))
        for k, (cs, cf) in enumerate(comps):
            gc[:, k] = base + cs
            gf[:, k] = cf
        sys_.ghost_comp = gc.reshape(-1)
        sys_.ghost_coef = gf.reshape(-1)
    else:
        sys_.ghost_atom = np.zeros(0, dtype=np.int32)
        sys_.ghost_ncomp = np.zeros(0, dtype=np.int32)
        sys_.ghost_comp = np.zeros(0, dtype=np.int32)
        sys_.ghost_coef = np.zeros(0)
    return sys_


def _random_rotations(rng, n):
    """Haar-uniform rotation matrices from unit quaternions."""
    qv = rng.normal(size=(n, 4))
    qv /= np.linalg.norm(qv, axis=1, keepdims=True)
    w, x, y, z = qv.T
    r = np.empty((n, 3, 3))
    r[:, 0, 0] = 1 - 2 * (y * y + z * z)
    r[:, 0, 1] = 2 * (x * y - z * w)
    r[:, 0, 2] = 2 * (x * z + y * w)
    r[:, 1, 0] = 2 * (x * y + z * w)
    r[:, 1, 1] = 1 - 2 * (x * x + z * z)
    r[:, 1, 2] = 2 * (y * z - x * w)
    r[:, 2, 0] = 2 * (x * z - y * w)
    r[:, 2, 1] = 2 * (y * z + x * w)
    r[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return r


def lattice_molecules(model: WaterModel, nmol: int, box: float, rng, jitter=0.3):
    """O on a jittered cubic lattice (first nmol of n^3 sites in lexicographic
    order, spacing box/n, uniform +-jitter), random orientations; returns
    site positions in Angstrom, shape (nmol*nsite, 3), ghosts filled."""
    n = int(math.ceil(nmol ** (1.0 / 3.0) - 1e-12))
    while n ** 3 < nmol:
        n += 1
    a = box / n
    idx = np.arange(nmol)
    ijk = np.stack([idx // (n * n), (idx // n) % n, idx % n], 1).astype(np.float64)
    o = (ijk + 0.5) * a + rng.uniform(-jitter, jitter, size=(nmol, 3))
    half = math.radians(model.theta) / 2.0
    body = np.array([[0.0, 0.0, 0.0],
                     [model.r_oh * math.sin(half), 0.0, model.r_oh * math.cos(half)],
                     [-model.r_oh * math.sin(half), 0.0, model.r_oh * math.cos(half)]])
    rot = _random_rotations(rng, nmol)
    h = np.einsum("mij,kj->mki", rot, body[1:]) + o[:, None, :]
    ns = len(model.sites)
    pos = np.zeros((nmol, ns, 3))
    pos[:, 0] = o
    pos[:, 1:3] = h
    if model.ghost is not None:
        g, comps = model.ghost
        pos[:, g] = sum(cf * pos[:, cs] for cs, cf in comps)
    return pos.reshape(-1, 3)


def ring_polymer_beads(centroid_A: np.ndarray, masses_amu: np.ndarray, P: int,
                       temperature: float, rng) -> np.ndarray:
    """Bead positions (P, natm, 3) in Angstrom: centroid plus free
    ring-polymer normal-mode displacements.

    Mode k = 1..P-1 has frequency w_k = 2 P kT/hbar sin(k pi/P).  Written as
    x_j = c + sum_k C_jk q_k with C the orthonormal real Fourier basis, the
    free-particle path-integral density exp(-(m P/(2 beta hbar^2))
    sum_j (x_j - x_{j+1})^2) makes q_k Gaussian with variance
    P kT/(m w_k^2) = beta hbar^2 / (4 m P sin^2(k pi/P)); equivalently
    kT/(m w_k^2) per mode in the unit-amplitude (sqrt(P)-scaled) basis."""
    natm = centroid_A.shape[0]
    if P == 1:
        return centroid_A[None].copy()
    beta = BOLTZ / temperature                    # 1/Hartree
    m = masses_amu * PROT_MASS                    # electron masses
    j = np.arange(P)
    disp = np.zeros((P, natm, 3))
    for k in range(1, P):
        lam = 4.0 * math.sin(math.pi * k / P) ** 2
        # orthonormal-mode variance of the discretised free-particle path
        var = beta * HBAR ** 2 / (m * P * lam)    # Bohr^2, per atom
        if 2 * k < P:
            basis = math.sqrt(2.0 / P) * np.cos(2 * math.pi * k * j / P)
        elif 2 * k == P:
            basis = math.sqrt(1.0 / P) * np.cos(math.pi * j)
        else:
            basis = math.sqrt(2.0 / P) * np.sin(2 * math.pi * (P - k) * j / P)
        qk = rng.normal(size=(natm, 3)) * np.sqrt(var)[:, None]
        disp += basis[:, None, None] * qk[None]
    return centroid_A[None] + disp * BOHR


@dataclasses.dataclass
class Config:
    system: System
    x: np.ndarray   # [P*N] Bohr, bead-major
    y: np.ndarray
    z: np.ndarray


def build_config(model: WaterModel, nmol: int, box: float, P: int, seed: int,
                 settings: Settings | None = None, jitter: float = 0.3) -> Config:
    rng = np.random.Generator(np.random.PCG64(seed))
    st = settings or Settings()
    sysd = make_system(model, nmol, box, P, st)
    cent = lattice_molecules(model, nmol, box, rng, jitter)
    masses = np.tile(np.array(model.masses_amu), nmol)
    beads = ring_polymer_beads(cent, masses, P, st.temperature, rng)
    if model.ghost is not None:
        g, comps = model.ghost
        ns = len(model.sites)
        b = beads.reshape(P, nmol, ns, 3)
        b[:, :, g] = sum(cf * b[:, :, cs] for cs, cf in comps)
        beads = b.reshape(P, -1, 3)
    pos = beads / BOHR
    return Config(sysd, np.ascontiguousarray(pos[..., 0].reshape(-1)),
                  np.ascontiguousarray(pos[..., 1].reshape(-1)),
                  np.ascontiguousarray(pos[..., 2].reshape(-1)))


def density_box(nmol: int, rho_gcc: float = 0.997) -> float:
    """Cubic box edge (Angstrom) of nmol waters at rho (g/cm^3)."""
    mw = 18.01528 / 6.02214076e23  # g per molecule
    vol_cm3 = nmol * mw / rho_gcc
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def config1(settings: Settings | None = None) -> Config:
    """256 SPC/Fw, L = 19.73 A, P = 1, PME 32^3 order 6, kmax 14."""
    st = settings or Settings(kmax_ewd=14, kmax_pme=14)
    return build_config(SPC_FW, 256, 19.73, 1, SEED_BASE + 1, st)


def config2(settings: Settings | None = None, P: int = 32) -> Config:
    st = settings or Settings(kmax_ewd=14, kmax_pme=14)
    return build_config(SPC_FW, 256, 19.73, P, SEED_BASE + 2, st)


def config3(settings: Settings | None = None, P: int = 32) -> Config:
    """4096 q-TIP4P/F (16384 sites), L = 49.66 A, P = 32, PME 64^3 order 6."""
    st = settings or Settings(kmax_ewd=30, kmax_pme=30)
    return build_config(Q_TIP4P_F, 4096, 49.66, P, SEED_BASE + 3, st)


def config4(settings: Settings | None = None, P: int = 8, sample: int = 0) -> Config:
    st = settings or Settings(kmax_ewd=20, kmax_pme=20)
    return build_config(Q_TIP4P_F, 1024, 31.28, P, SEED_BASE + 4 + 100 * sample, st)


# radixme's table of FFT-friendly sizes (set_cp_ewald.c:898-947)
RADIX_SIZES = (4, 8, 12, 16, 20, 24, 32, 36, 40, 48, 60, 64, 72, 80, 96, 100, 108, 120, 128, 144, 160, 180,
               192, 200, 216, 240, 256)


def config5(nbox: int = 64, seed: int = SEED_BASE + 5, max_mol: int = 4096) -> list:
    """64 SPC/Fw boxes at 0.997 g/cm^3: molecule counts log-uniform in
    [64, max_mol], bead counts from {1, 4, 8, 16, 32, 64}, per-box cutoff
    min(9 A, L/2 - 0.05 A), PME grid about L / 1 A snapped to the radix
    table (kmax_pme chosen so that PINY's sizing gives that grid), order 6,
    alpha = 3.5 / r_c.  Seeded; box k uses seed + 1000 k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nmol = np.exp(rng.uniform(np.log(64), np.log(max_mol), nbox)).astype(int)
    beads = rng.choice(np.array([1, 4, 8, 16, 32, 64]), nbox)
    out = []
    for k in range(nbox):
        L = density_box(int(nmol[k]))
        rc = min(9.0, 0.5 * L - 0.05)
        K = min(RADIX_SIZES, key=lambda r: abs(r - L))
        kp = max(2, 2 * (K // 4 - 1))
        st = Settings(cutoff=rc, alpha=3.5 / rc, kmax_ewd=min(kp, K // 2 - 2), kmax_pme=kp)
        out.append(build_config(SPC_FW, int(nmol[k]), L, int(beads[k]), seed + 1000 * k, st))
    return out


def with_cell(cfg: Config, cell_rows) -> Config:
    """The configuration in another cell: cell_rows are the rows of the
    cell matrix in Angstrom (cell vectors as columns, the layout of PINY's
    initial file), positions unchanged."""
    H = np.asarray(cell_rows, dtype=np.float64).reshape(3, 3)
    sysd = dataclasses.replace(cfg.system, cell=H.T.reshape(-1).copy())
    return Config(sysd, cfg.x, cfg.y, cfg.z)


def without_charges(cfg: Config) -> Config:
    """The configuration with every charge zero (LJ and bonded terms only)."""
    sysd = dataclasses.replace(cfg.system, q=np.zeros_like(cfg.system.q))
    return Config(sysd, cfg.x, cfg.y, cfg.z)


def small_config(model: WaterModel = SPC_FW, nmol: int = 27, P: int = 1, seed: int = 7,
                 box: float | None = None, cutoff: float | None = None,
                 kmax: int | None = None) -> Config:
    """A small box for fast oracle-vs-device checks."""
    L = box or density_box(nmol)
    rc = cutoff if cutoff is not None else min(9.0, 0.5 * L - 0.05)
    k = kmax if kmax is not None else max(4, int(2 * 0.35 * L * 1.0) // 2)
    st = Settings(cutoff=rc, kmax_ewd=k, kmax_pme=k, alpha=3.5 / rc)
    return build_config(model, nmol, L, P, seed, st)
