"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the C oracle (liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from piny_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(abi.PmxSystem), C.POINTER(abi.PmxParams), dp, dp, dp,
                                C.c_char_p, C.c_size_t]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_energy_control.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp, C.c_uint, dp, dp, dp,
                                        C.POINTER(abi.PmxEnergies)]
        L.or_get_verlet_list.argtypes = [C.c_void_p, ip, ip, ip, C.c_int, ip]
        L.or_get_nktot.argtypes = [C.c_void_p, ip, ip]
        L.or_get_kvectors.argtypes = [C.c_void_p, ip, ip, ip, dp, C.c_int]
        L.or_get_spline.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_int]
        L.or_get_setup_scalars.argtypes = [C.c_void_p, dp]
        L.or_get_bead_energies.argtypes = [C.c_void_p, dp]
        L.or_recip_energy.restype = C.c_double
        L.or_recip_energy.argtypes = [C.c_void_p, dp, dp, dp, C.c_int, dp, dp, dp, dp]
        for nm in ("or_nm_pos_to_mode", "or_nm_mode_to_pos", "or_nm_force_to_mode", "or_stag_pos_to_mode",
                   "or_stag_mode_to_pos", "or_stag_force_to_mode"):
            getattr(L, nm).argtypes = [C.c_int, dp, dp]
        L.or_gauss_legendre.argtypes = [C.c_int, dp, dp]
        L.or_set_respa.argtypes = [C.c_void_p, C.POINTER(abi.PmxRespa)]
        L.or_pimd_nvt.argtypes = [C.c_void_p, C.POINTER(abi.PmxNhcParams), C.c_int, dp, dp, dp, dp, dp, dp, dp,
                                  C.POINTER(abi.PmxEnergies), dp]
        L.or_bond_coef_morse.argtypes = [C.c_double, C.c_double, dp]
        L.or_bond_coef_harm.argtypes = [C.c_double, dp]
        L.or_bend_coef_harm.argtypes = [C.c_double, C.c_double, dp, dp]
        _lib = L
    return _lib


class Oracle:
    """CPU FP64 restatement of energy_control / energy_control_pimd."""

    def __init__(self, system, x=None, y=None, z=None):
        self.system = system
        self.st = abi.SystemStruct(system)
        L = lib()
        err = C.create_string_buffer(512)
        args = [None, None, None]
        if x is not None:
            self._init = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
            args = [abi.dptr(v) for v in self._init]
        self.h = L.or_create(C.byref(self.st.sys), C.byref(self.st.par), *args, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.N = system.natm
        self.P = system.pi_beads

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def set_respa(self, respa):
        """RESPA settings (an abi.PmxRespa, or None for off)."""
        self._respa = respa
        rc = lib().or_set_respa(self.h, C.byref(respa) if respa is not None else None)
        if rc != 0:
            raise RuntimeError(f"or_set_respa failed: {rc}")

    def energy_control(self, x, y, z, flags=None, xmod=None):
        P, N = self.P, self.N
        if flags is None:
            flags = abi.PMX_EVAL_PIMD if P > 1 else 0
        x = np.array(x, dtype=np.float64)
        y = np.array(y, dtype=np.float64)
        z = np.array(z, dtype=np.float64)
        fx = np.zeros(P * N)
        fy = np.zeros(P * N)
        fz = np.zeros(P * N)
        e = abi.PmxEnergies()
        xm = [None, None, None]
        if xmod is not None:
            self._xm = [np.ascontiguousarray(v, dtype=np.float64) for v in xmod]
            xm = [abi.dptr(v) for v in self._xm]
        rc = lib().or_energy_control(self.h, abi.dptr(x), abi.dptr(y), abi.dptr(z), *xm, flags,
                                     abi.dptr(fx), abi.dptr(fy), abi.dptr(fz), C.byref(e))
        if rc != 0:
            raise RuntimeError(f"or_energy_control failed: {rc}")
        return e.as_dict(), np.stack([fx, fy, fz], 1).reshape(P, N, 3), (x, y, z)

    def pimd_nvt(self, nsteps, x, y, z, v, nhc=None, params=None):
        """nsteps of PIMD NVT (or_pimd_nvt) from bead positions x, y, z [P*N]
        and mode velocities v [P, N, 3]; returns (positions [P, N, 3], mode
        velocities [P, N, 3], chain state, energies of the last force call,
        econs = [kinetic, chain energy, potential with springs, total])."""
        P, N = self.P, self.N
        params = params or abi.nhc_params()
        x, y, z = (np.array(a, dtype=np.float64) for a in (x, y, z))
        v = np.asarray(v, dtype=np.float64).reshape(P, N, 3)
        vx, vy, vz = (np.ascontiguousarray(v[:, :, d].reshape(-1)) for d in range(3))
        st = (np.zeros(abi.nhc_state_size(N, P, params.len_nhc)) if nhc is None
              else np.array(nhc, dtype=np.float64))
        e = abi.PmxEnergies()
        ec = np.zeros(4)
        rc = lib().or_pimd_nvt(self.h, C.byref(params), int(nsteps), abi.dptr(x), abi.dptr(y), abi.dptr(z),
                               abi.dptr(vx), abi.dptr(vy), abi.dptr(vz), abi.dptr(st), C.byref(e), abi.dptr(ec))
        if rc != 0:
            raise RuntimeError(f"or_pimd_nvt failed: {rc}")
        pos = np.stack([x, y, z], 1).reshape(P, N, 3)
        vel = np.stack([vx, vy, vz], 1).reshape(P, N, 3)
        return pos, vel, st, e.as_dict(), ec

    def verlet_list(self):
        N = self.N
        nter = np.zeros(N, np.int32)
        off = np.zeros(N, np.int32)
        nver = C.c_int(0)
        cap = 0
        dummy = np.zeros(1, np.int32)
        lib().or_get_verlet_list(self.h, abi.iptr(nter), abi.iptr(off), abi.iptr(dummy), 0,
                                 C.byref(nver))
        jter = np.zeros(max(1, nver.value), np.int32)
        rc = lib().or_get_verlet_list(self.h, abi.iptr(nter), abi.iptr(off), abi.iptr(jter),
                                      len(jter), C.byref(nver))
        assert rc == 0, rc
        return nter, off, jter[:nver.value]

    def kvectors(self):
        nk = C.c_int(0)
        ng = np.zeros(3, np.int32)
        lib().or_get_nktot(self.h, C.byref(nk), abi.iptr(ng))
        n = nk.value
        ka, kb, kc = (np.zeros(n, np.int32) for _ in range(3))
        bw = np.zeros(n)
        lib().or_get_kvectors(self.h, abi.iptr(ka), abi.iptr(kb), abi.iptr(kc), abi.dptr(bw), n)
        return ng, ka, kb, kc, bw

    def spline(self, which, ipair=0):
        out = np.zeros(self.system.settings.nsplin)
        rc = lib().or_get_spline(self.h, which, ipair, abi.dptr(out), len(out))
        assert rc == 0
        return out

    def scalars(self):
        out = np.zeros(32)
        lib().or_get_setup_scalars(self.h, abi.dptr(out))
        return out

    def bead_energies(self):
        out = np.zeros(4 * self.P)
        lib().or_get_bead_energies(self.h, abi.dptr(out))
        return out.reshape(self.P, 4)

    def recip_energy(self, x, y, z, use_pme=True):
        x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
        n = len(x)
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        pv = np.zeros(9)
        v = lib().or_recip_energy(self.h, abi.dptr(x), abi.dptr(y), abi.dptr(z), int(use_pme),
                                  abi.dptr(fx), abi.dptr(fy), abi.dptr(fz), abi.dptr(pv))
        return v, np.stack([fx, fy, fz], 1), pv


def nm_pos_to_mode(xb):
    P = len(xb)
    u = np.zeros(P)
    lib().or_nm_pos_to_mode(P, abi.dptr(np.ascontiguousarray(xb, dtype=np.float64)), abi.dptr(u))
    return u


def nm_mode_to_pos(u):
    P = len(u)
    x = np.zeros(P)
    lib().or_nm_mode_to_pos(P, abi.dptr(np.ascontiguousarray(u, dtype=np.float64)), abi.dptr(x))
    return x


def nm_force_to_mode(f):
    P = len(f)
    fm = np.zeros(P)
    lib().or_nm_force_to_mode(P, abi.dptr(np.ascontiguousarray(f, dtype=np.float64)), abi.dptr(fm))
    return fm


def _stag(fn, v):
    P = len(v)
    out = np.zeros(P)
    getattr(lib(), fn)(P, abi.dptr(np.ascontiguousarray(v, dtype=np.float64)), abi.dptr(out))
    return out


def stag_pos_to_mode(xb):
    return _stag("or_stag_pos_to_mode", xb)


def stag_mode_to_pos(u):
    return _stag("or_stag_mode_to_pos", u)


def stag_force_to_mode(f):
    return _stag("or_stag_force_to_mode", f)


def gauss_legendre(n):
    a, w = np.zeros(n), np.zeros(n)
    lib().or_gauss_legendre(n, abi.dptr(a), abi.dptr(w))
    return a, w
