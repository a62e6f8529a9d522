"""ctypes mirror of include/piny_amd.h (structs, enums) and helpers that turn a
``water.System`` into the plain-pointer structs the C ABI takes."""
from __future__ import annotations

import ctypes as C

import numpy as np

PMX_OK = 0
PMX_EVAL_PIMD = 1 << 0
PMX_EVAL_FORCE_LIST = 1 << 1
PMX_EVAL_NO_RECIP = 1 << 2
PMX_EVAL_MODE_FORCES = 1 << 3
PMX_EVAL_RES_INTER = 1 << 4
PMX_EVAL_NO_FULL_INTER = 1 << 5
PMX_EVAL_NO_FULL_INTRA = 1 << 6
PMX_PREC_FP64 = 0
PMX_PREC_MIXED = 1
PMX_PAIR_TABLE = 0
PMX_PAIR_DIRECT = 1
PMX_INTER_LJ = 2
PMX_INTER_NULL = 4

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class PmxSystem(C.Structure):
    _fields_ = [
        ("natm", C.c_int), ("ntype", C.c_int),
        ("atm_type", _ip), ("q", _dp), ("mass", _dp),
        ("inter_label", _ip), ("inter_sig", _dp), ("inter_eps", _dp),
        ("inter_cutti", _dp), ("inter_cutoff", _dp), ("inter_cutoff_res", _dp),
        ("nbond", C.c_int), ("bond_j1", _ip), ("bond_j2", _ip), ("bond_type", _ip),
        ("nbond_type", C.c_int), ("bond_c", _dp), ("bond_eq", _dp),
        ("nbend", C.c_int), ("bend_j1", _ip), ("bend_j2", _ip), ("bend_j3", _ip),
        ("bend_type", _ip), ("nbend_type", C.c_int), ("bend_c", _dp), ("bend_s", _dp),
        ("nghost", C.c_int), ("ghost_atom", _ip), ("ghost_ncomp", _ip),
        ("ghost_comp", _ip), ("ghost_coef", _dp),
    ]


class PmxParams(C.Structure):
    _fields_ = [
        ("hmat", C.c_double * 9), ("pi_beads", C.c_int), ("temperature", C.c_double),
        ("pimd_typ", C.c_int), ("nsplin", C.c_int), ("ishift", C.c_int),
        ("rheal", C.c_double), ("skin", C.c_double), ("ewald_alpha", C.c_double),
        ("kmax_ewd", C.c_int), ("pme_on", C.c_int), ("kmax_pme", C.c_int),
        ("n_interp", C.c_int), ("intra_perds", C.c_int), ("precision", C.c_int),
        ("pair_eval", C.c_int),
    ]


class PmxNhcParams(C.Structure):
    """pmx_nhc_params: the PIMD NVT integrator's settings (atomic units)."""
    _fields_ = [
        ("dt", C.c_double), ("tau_nhc", C.c_double), ("len_nhc", C.c_int), ("nres_nhc", C.c_int),
        ("nyosh_nhc", C.c_int), ("gamma_adb", C.c_double),
    ]


TIME_CONV = 0.0241888  # fs per atomic time unit (defines.h:85)
PCONV = 3.3989242e-09  # atomic pressure units per atm (defines.h:83)


class PmxRespa(C.Structure):
    """pmx_respa: RESPA step counts (0: keyword left out) and the inputs of
    the RESPA pressure-tensor terms (atomic units)."""
    _fields_ = [
        ("respa_steps_lrf", C.c_int), ("respa_steps_torsion", C.c_int), ("respa_steps_intra", C.c_int),
        ("respa_steps_pimd", C.c_int), ("nfree", C.c_int), ("pext", C.c_double),
        ("pten_inter_guess", C.c_double),
    ]


def respa_params(system, nfree=None):
    """The pmx_respa of a system's settings (respa_steps_*, pressure and
    pten_inter_respa in atm), or None when every RESPA keyword is left out.
    nfree defaults to three per real (non-ghost) site, the count
    close_intra_params.c:222-247 makes for molecules without constraints or
    frozen atoms."""
    st = system.settings
    if not (st.respa_lrf or st.respa_torsion or st.respa_intra or st.respa_pimd):
        return None
    if nfree is None:
        nfree = 3 * (system.natm - len(system.ghost_atom))
    return PmxRespa(respa_steps_lrf=st.respa_lrf, respa_steps_torsion=st.respa_torsion,
                    respa_steps_intra=st.respa_intra, respa_steps_pimd=st.respa_pimd, nfree=int(nfree),
                    pext=st.pressure * PCONV, pten_inter_guess=st.pten_inter_respa * PCONV)


def nhc_params(dt_fs=0.25, tau_fs=20.0, len_nhc=4, nres_nhc=4, nyosh_nhc=3, gamma_adb=1.0):
    """The settings of PINY's PIMD inputs (time_step, atm_nhc_tau_def in fs,
    atm_nhc_len, respa_steps_nhc, yosh_steps_nhc, path_int_gamma_adb)."""
    return PmxNhcParams(dt=dt_fs / TIME_CONV, tau_nhc=tau_fs / TIME_CONV, len_nhc=len_nhc, nres_nhc=nres_nhc,
                        nyosh_nhc=nyosh_nhc, gamma_adb=gamma_adb)


def nhc_state_size(N, P, len_nhc):
    """Doubles of the chain state: the centroid chain, then one chain per
    degree of freedom of modes 2..P (v[len], x[len] each)."""
    return 2 * len_nhc * (1 + 3 * N * (P - 1))


class PmxEnergies(C.Structure):
    _fields_ = [
        ("vintert", C.c_double), ("vintrat", C.c_double),
        ("vreal", C.c_double), ("vvdw", C.c_double), ("vcoul", C.c_double),
        ("vlong", C.c_double), ("vrecip", C.c_double), ("vself", C.c_double),
        ("vbgr", C.c_double), ("vecor", C.c_double), ("vbond", C.c_double),
        ("vbend", C.c_double), ("kin_harm", C.c_double), ("pi_ke_prim", C.c_double),
        ("pi_ke_vir", C.c_double), ("pvten", C.c_double * 9),
        ("pvten_tot", C.c_double * 9), ("spread", C.c_double), ("nver", C.c_int),
        ("list_rebuilt", C.c_int),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pvten", "pvten_tot")}
        d["pvten"] = np.array(self.pvten[:])
        d["pvten_tot"] = np.array(self.pvten_tot[:])
        return d


class PmxParamGrad(C.Structure):
    _fields_ = [(k, _dp) for k in ("energy", "dq", "deps", "dsig", "dbond_c", "dbond_eq", "dbend_c", "dbend_s")]


ENERGY_KEYS = ("vintert", "vintrat", "vreal", "vvdw", "vcoul", "vlong", "vrecip",
               "vself", "vbgr", "vecor", "vbond", "vbend")


def _iarr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _darr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class SystemStruct:
    """Holds a PmxSystem / PmxParams pair and the numpy arrays they point to."""

    def __init__(self, s, precision: int = PMX_PREC_FP64, pair_eval: int = PMX_PAIR_TABLE):
        self.keep = []

        def ia(x):
            arr, p = _iarr(x)
            self.keep.append(arr)
            return p

        def da(x):
            arr, p = _darr(x)
            self.keep.append(arr)
            return p

        st = s.settings
        self.sys = PmxSystem(
            natm=s.natm, ntype=s.ntype, atm_type=ia(s.atm_type), q=da(s.q), mass=da(s.mass),
            inter_label=ia(s.inter_label), inter_sig=da(s.inter_sig), inter_eps=da(s.inter_eps),
            inter_cutti=da(s.inter_cutti), inter_cutoff=da(s.inter_cutoff),
            inter_cutoff_res=da(s.inter_cutoff_res),
            nbond=len(s.bond_j1), bond_j1=ia(s.bond_j1), bond_j2=ia(s.bond_j2),
            bond_type=ia(s.bond_type), nbond_type=len(s.bond_eq),
            bond_c=da(np.asarray(s.bond_c).reshape(-1)), bond_eq=da(s.bond_eq),
            nbend=len(s.bend_j1), bend_j1=ia(s.bend_j1), bend_j2=ia(s.bend_j2),
            bend_j3=ia(s.bend_j3), bend_type=ia(s.bend_type),
            nbend_type=max(1, np.asarray(s.bend_c).size // 7),
            bend_c=da(np.asarray(s.bend_c).reshape(-1)), bend_s=da(np.asarray(s.bend_s).reshape(-1)),
            nghost=len(s.ghost_atom), ghost_atom=ia(s.ghost_atom), ghost_ncomp=ia(s.ghost_ncomp),
            ghost_comp=ia(s.ghost_comp), ghost_coef=da(s.ghost_coef))
        bohr = 0.529177
        self.par = PmxParams(
            pi_beads=s.pi_beads, temperature=st.temperature, pimd_typ=st.pimd_typ,
            nsplin=st.nsplin, ishift=st.ishift, rheal=st.rheal / bohr, skin=st.skin / bohr,
            ewald_alpha=s.ewald_alpha, kmax_ewd=st.kmax_ewd, pme_on=st.pme_on,
            kmax_pme=st.kmax_pme, n_interp=st.n_interp, intra_perds=st.intra_perds,
            precision=precision, pair_eval=pair_eval)
        for k, v in enumerate(s.hmat):
            self.par.hmat[k] = float(v)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)
