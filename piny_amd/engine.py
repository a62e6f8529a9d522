"""ctypes binding of libpiny_amd.so (include/piny_amd.h).

This is the host-side mirror of the reference's force call: ``Engine`` owns
one device context and exposes ``energy_control`` with PINY's semantics.
There is no CPU fallback: if the HIP library is missing or no device is
visible, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpiny_amd.so")

_lib = None


class PinyAmdError(RuntimeError):
    pass


def lib():
    """Loads libpiny_amd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PinyAmdError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    cdp = C.POINTER(C.POINTER(C.c_double))
    L.pmx_abi_version.restype = C.c_int
    L.pmx_device_count.restype = C.c_int
    L.pmx_create.restype = C.c_void_p
    L.pmx_create.argtypes = [C.POINTER(abi.PmxSystem), C.POINTER(abi.PmxParams), dp, dp, dp, C.c_int,
                             C.c_char_p, C.c_size_t]
    L.pmx_destroy.argtypes = [C.c_void_p]
    L.pmx_last_error.restype = C.c_char_p
    L.pmx_last_error.argtypes = [C.c_void_p]
    L.pmx_set_positions.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_set_positions_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.pmx_set_xmod.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_energy_control.argtypes = [C.c_void_p, C.c_uint, C.POINTER(abi.PmxEnergies)]
    L.pmx_energy_control_noalloc.argtypes = [C.c_void_p, C.c_uint, C.POINTER(abi.PmxEnergies)]
    L.pmx_get_forces.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_get_forces_device.argtypes = [C.c_void_p, cdp, cdp, cdp]
    L.pmx_get_fxt.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_get_bead_energies.argtypes = [C.c_void_p, dp]
    L.pmx_get_verlet_list.argtypes = [C.c_void_p, ip, ip, ip, C.c_int, ip]
    L.pmx_get_nktot.argtypes = [C.c_void_p, ip, ip]
    L.pmx_get_kvectors.argtypes = [C.c_void_p, ip, ip, ip, dp, C.c_int]
    L.pmx_get_spline.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_int]
    L.pmx_get_setup_scalars.argtypes = [C.c_void_p, dp]
    L.pmx_get_timings.argtypes = [C.c_void_p, dp]
    L.pmx_get_pair_count.argtypes = [C.c_void_p, dp]
    L.pmx_set_velocities.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_get_velocities.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_get_positions.argtypes = [C.c_void_p, dp, dp, dp]
    L.pmx_md_nve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(abi.PmxEnergies), dp]
    L.pmx_pimd_nvt.argtypes = [C.c_void_p, C.POINTER(abi.PmxNhcParams), C.c_int, C.POINTER(abi.PmxEnergies), dp]
    L.pmx_set_nhc_state.argtypes = [C.c_void_p, C.c_int, dp]
    L.pmx_get_nhc_state.argtypes = [C.c_void_p, C.c_int, dp]
    L.pmx_bond_coef_harm.argtypes = [C.c_double, dp]
    L.pmx_bond_coef_morse.argtypes = [C.c_double, C.c_double, dp]
    L.pmx_bend_coef_harm.argtypes = [C.c_double, C.c_double, dp, dp]
    L.pmx_set_force_field.argtypes = [C.c_void_p, C.POINTER(abi.PmxSystem)]
    L.pmx_get_param_grad.argtypes = [C.c_void_p, C.POINTER(abi.PmxParamGrad)]
    L.pmx_batch_create.restype = C.c_void_p
    L.pmx_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_char_p, C.c_size_t]
    L.pmx_batch_energy_control.argtypes = [C.c_void_p, C.c_uint, C.POINTER(abi.PmxEnergies)]
    L.pmx_batch_last_error.restype = C.c_char_p
    L.pmx_batch_last_error.argtypes = [C.c_void_p]
    L.pmx_batch_destroy.argtypes = [C.c_void_p]
    L.pmx_autotune.argtypes = [C.c_void_p, C.c_uint, C.c_int, C.POINTER(C.c_int)]
    L.pmx_batch_autotune.argtypes = [C.c_void_p, C.c_uint, C.c_int, C.POINTER(C.c_int)]
    L.pmx_set_respa.argtypes = [C.c_void_p, C.POINTER(abi.PmxRespa)]
    if L.pmx_abi_version() != 8:
        raise PinyAmdError("libpiny_amd.so ABI version mismatch")
    _lib = L
    return L


# every symbol include/piny_amd.h declares
EXPORTED = ("pmx_abi_version", "pmx_device_count", "pmx_create", "pmx_destroy", "pmx_last_error",
            "pmx_set_positions", "pmx_set_positions_device", "pmx_set_xmod", "pmx_energy_control",
            "pmx_energy_control_noalloc",
            "pmx_get_forces", "pmx_get_forces_device", "pmx_get_fxt", "pmx_get_bead_energies",
            "pmx_get_verlet_list", "pmx_get_nktot", "pmx_get_kvectors", "pmx_get_spline",
            "pmx_get_setup_scalars", "pmx_get_timings", "pmx_get_pair_count", "pmx_set_velocities",
            "pmx_get_velocities", "pmx_get_positions", "pmx_md_nve", "pmx_bond_coef_harm", "pmx_bond_coef_morse",
            "pmx_bend_coef_harm", "pmx_set_force_field", "pmx_get_param_grad", "pmx_batch_create",
            "pmx_batch_energy_control", "pmx_batch_last_error", "pmx_batch_destroy", "pmx_autotune",
            "pmx_batch_autotune", "pmx_pimd_nvt", "pmx_set_nhc_state", "pmx_get_nhc_state", "pmx_set_respa")


class Engine:
    """One device context: setup at construction, then set_positions /
    energy_control / forces, like PINY's energy_control on its structs."""

    def __init__(self, system, x=None, y=None, z=None, precision=abi.PMX_PREC_FP64, device=0,
                 pair_eval=abi.PMX_PAIR_TABLE):
        L = lib()
        self.system = system
        self.st = abi.SystemStruct(system, precision, pair_eval)
        self.N = system.natm
        self.P = system.pi_beads
        args = [None, None, None]
        if x is not None:
            self._init = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
            args = [abi.dptr(v) for v in self._init]
        err = C.create_string_buffer(1024)
        self.h = L.pmx_create(C.byref(self.st.sys), C.byref(self.st.par), *args, device, err, 1024)
        if not self.h:
            raise PinyAmdError(err.value.decode())

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().pmx_last_error(self.h)
            raise PinyAmdError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_positions(self, x, y, z):
        self._pos = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        self._chk(lib().pmx_set_positions(self.h, *[abi.dptr(v) for v in self._pos]), "pmx_set_positions")

    def set_positions_device(self, px: int, py: int, pz: int):
        self._chk(lib().pmx_set_positions_device(self.h, C.c_void_p(px), C.c_void_p(py), C.c_void_p(pz)),
                  "pmx_set_positions_device")

    def set_xmod(self, xm=None):
        if xm is None:
            self._chk(lib().pmx_set_xmod(self.h, None, None, None), "pmx_set_xmod")
            return
        self._xm = [np.ascontiguousarray(v, dtype=np.float64) for v in xm]
        self._chk(lib().pmx_set_xmod(self.h, *[abi.dptr(v) for v in self._xm]), "pmx_set_xmod")

    def set_respa(self, respa):
        """RESPA settings for the later force calls (an abi.PmxRespa, or None
        for off)."""
        self._respa = respa
        self._chk(lib().pmx_set_respa(self.h, C.byref(respa) if respa is not None else None), "pmx_set_respa")

    def energy_control(self, flags=None):
        if flags is None:
            flags = abi.PMX_EVAL_PIMD if self.P > 1 else 0
        e = abi.PmxEnergies()
        self._chk(lib().pmx_energy_control(self.h, flags, C.byref(e)), "pmx_energy_control")
        return e.as_dict()

    def energy_control_noalloc(self, flags=None):
        """The force call through pmx_energy_control_noalloc (partial sums in
        the library's static scratch, no allocation)."""
        if flags is None:
            flags = abi.PMX_EVAL_PIMD if self.P > 1 else 0
        e = abi.PmxEnergies()
        self._chk(lib().pmx_energy_control_noalloc(self.h, flags, C.byref(e)), "pmx_energy_control_noalloc")
        return e.as_dict()

    def forces(self):
        n = self.N * self.P
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        self._chk(lib().pmx_get_forces(self.h, abi.dptr(fx), abi.dptr(fy), abi.dptr(fz)), "pmx_get_forces")
        return np.stack([fx, fy, fz], 1).reshape(self.P, self.N, 3)

    def forces_device(self):
        px, py, pz = (C.POINTER(C.c_double)() for _ in range(3))
        self._chk(lib().pmx_get_forces_device(self.h, C.byref(px), C.byref(py), C.byref(pz)),
                  "pmx_get_forces_device")
        return [C.cast(p, C.c_void_p).value for p in (px, py, pz)]

    def _bead_major(self, fn, what):
        n = self.N * self.P
        a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
        self._chk(fn(self.h, abi.dptr(a), abi.dptr(b), abi.dptr(c)), what)
        return np.stack([a, b, c], 1).reshape(self.P, self.N, 3)

    def fxt(self):
        return self._bead_major(lib().pmx_get_fxt, "pmx_get_fxt")

    def positions(self):
        return self._bead_major(lib().pmx_get_positions, "pmx_get_positions")

    def velocities(self):
        return self._bead_major(lib().pmx_get_velocities, "pmx_get_velocities")

    def set_velocities(self, vx, vy, vz):
        self._vel = [np.ascontiguousarray(v, dtype=np.float64) for v in (vx, vy, vz)]
        self._chk(lib().pmx_set_velocities(self.h, *[abi.dptr(v) for v in self._vel]), "pmx_set_velocities")

    def md_nve(self, nsteps, dt):
        """nsteps of velocity Verlet (int_NVE); dt in atomic time units.
        Returns (energies of the last force call, kinetic energy)."""
        e = abi.PmxEnergies()
        ke = C.c_double(0.0)
        self._chk(lib().pmx_md_nve(self.h, int(nsteps), float(dt), C.byref(e), C.byref(ke)), "pmx_md_nve")
        return e.as_dict(), ke.value

    def pimd_nvt(self, nsteps, params=None):
        """nsteps of PIMD NVT on the device (pmx_pimd_nvt); the velocities are
        mode velocities.  Returns (energies of the last force call, econs =
        [kinetic, chain energy, potential with springs, total])."""
        params = params or abi.nhc_params()
        self._nhc_len = params.len_nhc
        e = abi.PmxEnergies()
        ec = np.zeros(4)
        self._chk(lib().pmx_pimd_nvt(self.h, C.byref(params), int(nsteps), C.byref(e), abi.dptr(ec)),
                  "pmx_pimd_nvt")
        return e.as_dict(), ec

    def nhc_state(self, len_nhc=None):
        L = len_nhc or getattr(self, "_nhc_len", 4)
        st = np.zeros(abi.nhc_state_size(self.N, self.P, L))
        self._chk(lib().pmx_get_nhc_state(self.h, L, abi.dptr(st)), "pmx_get_nhc_state")
        return st

    def set_nhc_state(self, st, len_nhc):
        self._nhc_st = np.ascontiguousarray(st, dtype=np.float64)
        self._nhc_len = len_nhc
        self._chk(lib().pmx_set_nhc_state(self.h, int(len_nhc), abi.dptr(self._nhc_st)), "pmx_set_nhc_state")

    def set_force_field(self, system):
        """New charges, LJ sigma/eps and bond/bend coefficients (same topology)."""
        st = abi.SystemStruct(system, self.st.par.precision, self.st.par.pair_eval)
        self._chk(lib().pmx_set_force_field(self.h, C.byref(st.sys)), "pmx_set_force_field")
        self.system = system
        self.st = st

    def param_grad(self):
        """dE_b/d(parameters) per bead at the last force call (pmx_get_param_grad):
        dict of arrays energy [P], dq [P, N], deps/dsig [P, npair],
        dbond_c [P, nbt, 7], dbond_eq [P, nbt], dbend_c/dbend_s [P, nat, 7]."""
        P, sy = self.P, self.st.sys
        nbt, nat, npair = sy.nbond_type, sy.nbend_type, self.system.ntype * (self.system.ntype + 1) // 2
        out = {"energy": np.zeros(P), "dq": np.zeros((P, self.N)), "deps": np.zeros((P, npair)),
               "dsig": np.zeros((P, npair)), "dbond_c": np.zeros((P, nbt, 7)), "dbond_eq": np.zeros((P, nbt)),
               "dbend_c": np.zeros((P, nat, 7)), "dbend_s": np.zeros((P, nat, 7))}
        g = abi.PmxParamGrad(**{k: abi.dptr(v) for k, v in out.items()})
        self._chk(lib().pmx_get_param_grad(self.h, C.byref(g)), "pmx_get_param_grad")
        return out

    def autotune(self, reps=3, flags=None):
        """Picks launch shapes by timing (pmx_autotune); returns them."""
        if flags is None:
            flags = abi.PMX_EVAL_PIMD if self.P > 1 else 0
        ch = (C.c_int * 5)()
        self._chk(lib().pmx_autotune(self.h, flags, int(reps), ch), "pmx_autotune")
        return {"pme_tile": tuple(ch[:3]), "pair_waves": ch[3], "list": ("plain", "cells")[ch[4] - 1]}

    def bead_potential(self):
        """E_b [P] of the last force call (pmx_get_param_grad, energies only)."""
        e = np.zeros(self.P)
        g = abi.PmxParamGrad(energy=abi.dptr(e))
        self._chk(lib().pmx_get_param_grad(self.h, C.byref(g)), "pmx_get_param_grad")
        return e

    def bead_energies(self):
        out = np.zeros(4 * self.P)
        self._chk(lib().pmx_get_bead_energies(self.h, abi.dptr(out)), "pmx_get_bead_energies")
        return out.reshape(self.P, 4)

    def verlet_list(self):
        N = self.N
        nter = np.zeros(N, np.int32)
        off = np.zeros(N, np.int32)
        nver = C.c_int(0)
        dummy = np.zeros(1, np.int32)
        lib().pmx_get_verlet_list(self.h, abi.iptr(nter), abi.iptr(off), abi.iptr(dummy), 0, C.byref(nver))
        jter = np.zeros(max(1, nver.value), np.int32)
        self._chk(lib().pmx_get_verlet_list(self.h, abi.iptr(nter), abi.iptr(off), abi.iptr(jter), len(jter),
                                            C.byref(nver)), "pmx_get_verlet_list")
        return nter, off, jter[:nver.value]

    def kvectors(self):
        nk = C.c_int(0)
        ng = np.zeros(3, np.int32)
        self._chk(lib().pmx_get_nktot(self.h, C.byref(nk), abi.iptr(ng)), "pmx_get_nktot")
        n = nk.value
        ka, kb, kc = (np.zeros(n, np.int32) for _ in range(3))
        bw = np.zeros(n)
        self._chk(lib().pmx_get_kvectors(self.h, abi.iptr(ka), abi.iptr(kb), abi.iptr(kc), abi.dptr(bw), n),
                  "pmx_get_kvectors")
        return ng, ka, kb, kc, bw

    def spline(self, which, ipair=0):
        out = np.zeros(self.system.settings.nsplin)
        self._chk(lib().pmx_get_spline(self.h, which, ipair, abi.dptr(out), len(out)), "pmx_get_spline")
        return out

    def scalars(self):
        out = np.zeros(32)
        self._chk(lib().pmx_get_setup_scalars(self.h, abi.dptr(out)), "pmx_get_setup_scalars")
        return out

    def pair_count(self):
        v = C.c_double(0.0)
        self._chk(lib().pmx_get_pair_count(self.h, C.byref(v)), "pmx_get_pair_count")
        return v.value

    def timings(self):
        out = np.zeros(5)
        self._chk(lib().pmx_get_timings(self.h, abi.dptr(out)), "pmx_get_timings")
        return dict(zip(("list", "real", "recip", "intra", "total"), out))


class Batch:
    """Several engines (boxes) on one device evaluated by one graph launch per
    call (pmx_batch_energy_control)."""

    def __init__(self, engines):
        L = lib()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        err = C.create_string_buffer(1024)
        self.h = L.pmx_batch_create(arr, len(self.engines), err, 1024)
        if not self.h:
            raise PinyAmdError(err.value.decode())

    def energy_control(self, flags=0):
        out = (abi.PmxEnergies * len(self.engines))()
        rc = lib().pmx_batch_energy_control(self.h, flags, out)
        if rc != 0:
            msg = lib().pmx_batch_last_error(self.h)
            raise PinyAmdError(f"pmx_batch_energy_control failed ({rc}): {msg.decode() if msg else ''}")
        return [e.as_dict() for e in out]

    def autotune(self, reps=2, flags=0):
        """Launch shapes for the whole batch, by timing batched calls."""
        ch = (C.c_int * 4)()
        rc = lib().pmx_batch_autotune(self.h, flags, int(reps), ch)
        if rc != 0:
            msg = lib().pmx_batch_last_error(self.h)
            raise PinyAmdError(f"pmx_batch_autotune failed ({rc}): {msg.decode() if msg else ''}")
        return {"pair_waves": ch[0], "pme_tile_cap": tuple(ch[1:4])}

    def close(self):
        if getattr(self, "h", None):
            lib().pmx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
