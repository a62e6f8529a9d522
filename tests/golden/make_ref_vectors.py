"""Writes golden vectors from the reference itself: oracle/_ref/piny_ref is
the reference's own C and Fortran compiled from /root/reference by
oracle/ref.mk and linked with oracle/ref_harness.c, which runs the
reference's parse() on the input files written here and then one
energy_control() (P = 1) or energy_control_pimd() (P > 1) on the case's
exact positions.

Run in a container where /root/reference is mounted and the reference has
been built (make -C oracle -f ref.mk); the .npz files it writes next to this
script are the committed fixtures, holding the inputs (case name and
positions) and the reference's outputs.

The input files follow the formats of python/piny/tools.py:137-182 and the
example inputs examples/q-SPC-Fw/classical/build.py and
examples/q-TIP4P-F/build.py; the force-field values are the ones
piny_amd/water.py converts (Angstrom, Kelvin via kB = 0.001987204118
kcal/mol/K, amu, degrees).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from piny_amd import water as w  # noqa: E402
from tests import cases  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "piny_ref")
CASES = ["spc64", "spc64_shift", "spc64_none", "tip64", "cfg1", "spc64_k20", "tip64_p4", "cfg2_p4", "spc64_ortho",
         "tip64_p4_tric", "tip64_p6", "spc27_p64", "spc64_noq", "spc64_ewald",
         "tip64_p4_ewald", "spc64_k80", "w512", "w512_ewald", "tip64_p4_stag", "cfg2_p4_stag"]
SHIFT = {0: "off", 1: "on", 2: "swit"}  # shift_inter_pe values (set_sim_dict.c:736)
PIMD_TYP = {1: "staging", 2: "centroid"}  # path_int_md_typ values (transform_cnt.c:104,190)


def section(name, items):
    body = "\n".join("    \\%s{%s}" % (k, v) for k, v in items)
    return "~%s[\n%s]" % (name, body)


def write_inputs(d, c):
    s = c.system
    st = s.settings
    m = s.model
    P = s.pi_beads
    # PIMD needs a thermostatted ensemble in the reference's input checks; the
    # thermostat does not enter the force call
    gen = [("simulation_typ", "pimd" if P > 1 else "md"), ("ensemble_typ", "nvt" if P > 1 else "nve"),
           ("num_time_step", 0),
           ("restart_type", "initial"), ("time_step", 0.25), ("temperature", st.temperature),
           ("generic_fft_opt", "on"), ("num_proc_tot", 1), ("num_proc_beads", 1), ("num_proc_class_forc", 1)]
    if st.pressure:
        gen.append(("pressure", repr(st.pressure)))
    parts = [section("sim_gen_def", gen)]
    if P > 1:
        pimd = [("path_int_beads", P), ("path_int_md_typ", PIMD_TYP[st.pimd_typ]), ("initial_spread_opt", "off")]
        if st.respa_pimd:
            pimd.append(("respa_steps_pimd", st.respa_pimd))
        parts.append(section("sim_pimd_def", pimd))
    if P > 1:
        parts.append(section("sim_nhc_def", [("atm_nhc_tau_def", 20.0), ("atm_nhc_len", 4), ("respa_steps_nhc", 4),
                                             ("yosh_steps_nhc", 3), ("respa_xi_opt", 1)]))
    parts.append(section("sim_list_def", [("neighbor_list", "ver_list"), ("verlist_skin", repr(st.skin)),
                                          ("update_type", "no_list")]))
    run = [("respa_rheal", repr(st.rheal))]
    for key, v in (("respa_steps_lrf", st.respa_lrf), ("respa_steps_torsion", st.respa_torsion),
                   ("respa_steps_intra", st.respa_intra)):
        if v:
            run.append((key, v))
    parts.append(section("sim_run_def", run))
    parts.append(section("sim_vol_def", [("periodicity", 3), ("intra_perds", "on" if st.intra_perds else "off")]))
    parts.append(section("sim_class_PE_def", [
        ("shift_inter_pe", SHIFT[st.ishift]), ("inter_spline_pts", st.nsplin),
        ("ewald_alpha", repr(s.ewald_alpha)), ("ewald_kmax", st.kmax_ewd),
        ("ewald_pme_opt", "on" if st.pme_on else "off"), ("ewald_kmax_pme", st.kmax_pme),
        ("ewald_interp_pme", st.n_interp), ("ewald_respa_pme_opt", "off"), ("inter_PE_calc_freq", 1)]
        + ([("pten_inter_respa", repr(st.pten_inter_respa))] if st.pten_inter_respa else [])))
    parts.append(section("sim_write_def", [
        ("sim_name", "water.input-all"), ("write_screen_freq", 1000000), ("instant_file", "water.iavg"),
        ("write_inst_freq", 1000000), ("atm_pos_file", "water.confp"), ("write_pos_freq", 1000000),
        ("path_cent_file", "water-centroid.confp"), ("path_cent_freq", 1000000),
        ("atm_vel_file", "water.confv"), ("write_vel_freq", 1000000), ("atm_force_file", "water.conff"),
        ("write_force_freq", 1000000), ("out_restart_file", "water.restart"), ("write_dump_freq", 1000000),
        ("in_restart_file", "water.initial"), ("mol_set_file", "water.set"),
        ("conf_file_format", "formatted")]))
    open(os.path.join(d, "water.input"), "w").write("\n\n".join(parts) + "\n")

    kb = w.KB_KCAL
    parm = [section("molecule_name_def", [("molecule_name", "water"), ("natom", len(m.sites))])]
    ghost = dict()
    if m.ghost is not None:
        ghost[m.ghost[0]] = m.ghost[1]
    # the system's charges (the same in every molecule)
    qs = np.asarray(s.q).reshape(s.nmol, len(m.sites))
    assert np.all(qs == qs[0])
    for k, (name, q, mass) in enumerate(zip(m.sites, qs[0], m.masses_amu)):
        items = [("atom_typ", name), ("atom_ind", k + 1), ("mass", repr(float(mass))), ("charge", repr(float(q)))]
        for g, (comp, coef) in enumerate(ghost.get(k, ())):
            items.append(("def_ghost%d" % (g + 1), "%d, %r" % (comp + 1, coef)))
        parm.append(section("atom_def", items))
    parm.append(section("bond_def", [("atom1", 1), ("atom2", 2)]))
    parm.append(section("bond_def", [("atom1", 1), ("atom2", 3)]))
    parm.append(section("bend_def", [("atom1", 2), ("atom2", 1), ("atom3", 3)]))
    open(os.path.join(d, "water.parm"), "w").write("\n\n".join(parm) + "\n")

    if m.bond[0] == "harm":
        bond = [("atom1", "O"), ("atom2", "H"), ("pot_type", "harm"), ("eq", repr(m.bond[2])),
                ("fk", repr(m.bond[1] / kb))]
    else:
        bond = [("atom1", "O"), ("atom2", "H"), ("pot_type", "morse"), ("eq", repr(m.bond[3])),
                ("alpha", repr(m.bond[2])), ("d0", repr(m.bond[1] / kb))]
    open(os.path.join(d, "water.bond"), "w").write(section("bond_parm", bond) + "\n")
    bend = [("atom1", "H"), ("atom2", "O"), ("atom3", "H"), ("pot_type_bend", "harm"),
            ("eq_bend", repr(m.bend[2])), ("fk_bend", repr(m.bend[1] / kb))]
    open(os.path.join(d, "water.bend"), "w").write(section("bend_parm", bend) + "\n")

    types = []
    for n in m.sites:
        if n not in types:
            types.append(n)
    inter = []
    for a in range(len(types)):
        for b in range(a, len(types)):
            items = [("atom1", types[a]), ("atom2", types[b])]
            if types[a] == "O" and types[b] == "O":
                items += [("pot_type", "lennard-jones"), ("eps", repr(m.lj_eps_kcal / kb)), ("sig", repr(m.lj_sig))]
            else:
                items += [("pot_type", "null")]
            items += [("min_dist", repr(st.min_dist)), ("max_dist", repr(st.cutoff)), ("res_dist", repr(st.res_dist))]
            inter.append(section("inter_parm", items))
    open(os.path.join(d, "water.inter"), "w").write("\n\n".join(inter) + "\n")

    open(os.path.join(d, "water.set"), "w").write("\n\n".join([
        section("molecule_def", [("mol_parm_file", "water.parm"), ("mol_opt_nhc", "global"),
                                 ("num_mol", s.nmol), ("mol_index", 1), ("mol_name", "water")]),
        section("data_base_def", [("bond_file", "water.bond"), ("bend_file", "water.bend"),
                                  ("inter_file", "water.inter")])]) + "\n")

    # initial file (Angstrom): P blocks of N sites (read_coord.c:383-400); the
    # harness then sets every bead to the exact Bohr values
    N = s.natm
    lines = ["%d 1 %d" % (N, P)]
    for i in range(N * P):
        lines.append("%.15f %.15f %.15f" % (c.x[i] * w.BOHR, c.y[i] * w.BOHR, c.z[i] * w.BOHR))
    # cell rows (read_hmat.c:255-301 reads line i as hmat[1+i], hmat[4+i], hmat[7+i])
    h = s.hmat * w.BOHR
    for r in range(3):
        lines.append("%.15f %.15f %.15f" % (h[r], h[3 + r], h[6 + r]))
    open(os.path.join(d, "water.initial"), "w").write("\n".join(lines) + "\n")

    pos = np.stack([c.x, c.y, c.z], 1).reshape(P, N, 3)
    pos.astype(np.float64).tofile(os.path.join(d, "pos.bin"))


def run_case(name):
    c = cases.get(name)
    s = c.system
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        with open(os.path.join(d, "stdout.txt"), "w") as log:
            r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin"], cwd=d, stdout=log,
                               stderr=subprocess.STDOUT, timeout=600)
        if r.returncode != 0:
            print(open(os.path.join(d, "stdout.txt")).read()[-3000:])
            raise RuntimeError(f"{name}: reference run failed ({r.returncode})")
        raw = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
    N, P, nver = int(raw[0]), int(raw[1]), int(raw[2])
    assert N == s.natm and P == s.pi_beads, (N, P)
    e = raw[3:27]
    pvten_tot = raw[27:36]
    f = raw[36:36 + 3 * N * P].reshape(P, N, 3)
    fxt = raw[36 + 3 * N * P:36 + 6 * N * P].reshape(P, N, 3) if P > 1 else np.zeros((0, N, 3))
    keys = ["vintert", "vintrat", "vreal", "vrecip", "vvdw", "vcoul", "vlong", "vbond", "vbend",
            "pi_ke_prim", "pi_ke_vir", "kin_harm"]
    out = os.path.join(HERE, f"ref_{name}.npz")
    np.savez_compressed(out, case=name, x=c.x, y=c.y, z=c.z, nver=nver,
                        energy_keys=np.array(keys), energies=e[:12], pvten=e[12:21], pvten_tot=pvten_tot,
                        forces=f, fxt=fxt)
    print(f"{name}: N={N} P={P} nver={nver} vintert={e[0]:.12g} vintrat={e[1]:.12g} -> {out}")


# RESPA: the force call with combinations of (iget_full_inter,
# iget_res_inter, iget_full_intra), iget_res_intra on (oracle/ref_harness.c
# respa mode)
RESPA_CASES = ["spc64_respa", "spc64_ewald_respa", "tip64_p4_respa"]
RESPA_FLAGS = [(1, 1, 1), (0, 1, 1), (0, 0, 1), (1, 0, 1), (0, 1, 0), (0, 0, 0)]


def run_respa_case(name):
    c = cases.get(name)
    s = c.system
    res = {k: [] for k in ("energies", "pvten", "pvten_tot", "forces", "fxt")}
    nver = nfree = None
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        for full, rs, fintra in RESPA_FLAGS:
            # parse() refuses to overwrite the summary file of a previous run
            if os.path.exists(os.path.join(d, "water.input-all")):
                os.remove(os.path.join(d, "water.input-all"))
            with open(os.path.join(d, "stdout.txt"), "w") as log:
                r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin", "respa", str(full), str(rs), str(fintra)], cwd=d,
                                   stdout=log, stderr=subprocess.STDOUT, timeout=600)
            text = open(os.path.join(d, "stdout.txt")).read()
            if r.returncode != 0:
                print(text[-3000:])
                raise RuntimeError(f"{name}: reference run failed ({r.returncode})")
            # the degrees of freedom the reference counted (close_intra_params.c:247,258)
            nfree = int(text.split("number of atomic degrees of freedom:")[1].split()[0])
            raw = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
            N, P = int(raw[0]), int(raw[1])
            nver = int(raw[2])
            e = raw[3:27]
            res["energies"].append(e[:12])
            res["pvten"].append(e[12:21])
            res["pvten_tot"].append(raw[27:36])
            res["forces"].append(raw[36:36 + 3 * N * P].reshape(P, N, 3))
            res["fxt"].append(raw[36 + 3 * N * P:36 + 6 * N * P].reshape(P, N, 3) if P > 1 else np.zeros((0, N, 3)))
    keys = ["vintert", "vintrat", "vreal", "vrecip", "vvdw", "vcoul", "vlong", "vbond", "vbend",
            "pi_ke_prim", "pi_ke_vir", "kin_harm"]
    out = os.path.join(HERE, f"respa_{name}.npz")
    np.savez_compressed(out, case=name, x=c.x, y=c.y, z=c.z, nver=nver, nfree=nfree, flags=np.array(RESPA_FLAGS),
                        energy_keys=np.array(keys), **{k: np.array(v) for k, v in res.items()})
    print(f"{name}: N={s.natm} P={s.pi_beads} nver={nver} nfree={nfree} vintert(full+res)="
          f"{res['energies'][0][0]:.12g} -> {out}")


# the reference's setup products and Verlet list (oracle/ref_harness.c dump
# mode): tables with 500 spline points keep the fixtures small
SETUP_CASES = ["spc64_ns500", "tip64_p4_ns500", "tip64_p4_tric_ns500", "cfg1_ns500"]


def run_setup_case(name):
    """The reference's Verlet list (nter, jver_off, jter), spline tables
    (cv0, cdv0, cv0_c, cdv0_c per type pair), k-vectors with their B-spline
    weights and the excluded-pair table, 0-based where they are indices."""
    c = cases.get(name)
    s = c.system
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        with open(os.path.join(d, "stdout.txt"), "w") as log:
            r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin", "dump", "setup.bin"], cwd=d, stdout=log,
                               stderr=subprocess.STDOUT, timeout=600)
        if r.returncode != 0:
            print(open(os.path.join(d, "stdout.txt")).read()[-3000:])
            raise RuntimeError(f"{name}: reference run failed ({r.returncode})")
        raw = open(os.path.join(d, "setup.bin"), "rb").read()
    hdr = np.frombuffer(raw, np.int32, 8)
    N, nver, ninter, nsp, nk, nse = (int(v) for v in hdr[:6])
    assert N == s.natm
    o = 32

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(raw, dtype, n, o)
        o += a.nbytes
        return a.copy()
    nter = take(np.int32, N)
    jver_off = take(np.int32, N)
    jter = take(np.int32, nver)
    tabs = take(np.float64, 4 * ninter * nsp).reshape(4, ninter, nsp)
    ka, kb, kc = take(np.int32, nk), take(np.int32, nk), take(np.int32, nk)
    bweight = take(np.float64, nk)
    ecor = take(np.float64, 2 * nse).reshape(2, nse)
    assert o == len(raw)
    out = os.path.join(HERE, f"setup_{name}.npz")
    np.savez_compressed(out, case=name, nter=nter, jver_off=jver_off, jter=jter - 1, tables=tabs, ka=ka, kb=kb,
                        kc=kc, bweight=bweight, ecor=ecor)
    print(f"setup {name}: N={N} nver={nver} ninter={ninter} nsplin={nsp} nktot={nk} -> {out}")


def run_w512_debug():
    """The reference's own debug check of the reciprocal-space sum
    (test_force with inter = 3, test_energy.c:163-208, as the interactive
    debug driver that wrote examples/q-TIP4P-F/Ewald-log.txt runs it) on the
    shipped W512 coordinates, for the five settings of the log: standard
    Ewald with kmax 13 and 26, PME of order 8 with kmax_pme 17, 22, 25."""
    import dataclasses
    kat = np.load(os.path.join(HERE, "w512_ewald_kat.npz"))
    vals = []
    for pme, kmax, kpme, order in zip(kat["use_pme"], kat["kmax_ewd"], kat["kmax_pme"], kat["order"]):
        c = cases.get("w512" if pme else "w512_ewald")
        st = dataclasses.replace(c.system.settings, kmax_ewd=int(kmax), kmax_pme=int(kpme) if pme else 25,
                                 n_interp=int(order) if pme else 8)
        s = w.make_system(w.Q_TIP4P_F, 512, float(kat["box_A"]), 1, st)
        with tempfile.TemporaryDirectory() as d:
            write_inputs(d, w.Config(s, c.x, c.y, c.z))
            r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin", "recip"], cwd=d, capture_output=True,
                               text=True, timeout=600)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("RECIP")]
        if r.returncode != 0 or not line:
            raise RuntimeError("w512 debug run failed")
        vals.append(float(line[0].split()[1]))
    out = os.path.join(HERE, "w512_debug_recip.npz")
    np.savez(out, labels=kat["labels"], energy=np.array(vals))
    print("w512 debug recip", vals, "->", out)


# PIMD NVT trajectories of the reference's own int_NVT_pimd (20 steps of
# 0.25 fs, chains of length 4 with 4 RESPA steps and 3 Yoshida weights,
# tau 20 fs, from zeroed chains), normal modes and staging
PIMD_CASES = {"tip64_p4": 20, "tip64_p4_stag": 20, "cfg2_p4": 20}


def pimd_velocities(c, seed):
    """Seeded mode velocities at the case temperature: Gaussian with the mode
    masses m veig (transform_cnt.c:175-233, gamma_adb = 1), zero on ghost
    sites, [P, N, 3] in Bohr per atomic time unit."""
    s = c.system
    P, N = s.pi_beads, s.natm
    rng = np.random.Generator(np.random.PCG64(seed))
    kt = s.settings.temperature / w.BOLTZ
    m = np.asarray(s.mass, dtype=np.float64)
    veig = np.ones(P)
    if s.settings.pimd_typ == 1:
        veig[1:] = [ip / (ip - 1.0) for ip in range(2, P + 1)]
    else:
        veig[P - 1] = 4.0 * P
        for ip in range(2, P // 2 + 1):
            veig[2 * ip - 3] = veig[2 * ip - 2] = (1.0 - np.cos(2 * np.pi / P * (ip - 1))) * 4.0 * P
    v = rng.normal(size=(P, N, 3)) * np.sqrt(kt / (m[None, :] * veig[:, None]))[:, :, None]
    v[:, np.asarray(s.ghost_atom, dtype=int)] = 0.0
    return v


def run_pimd_case(name, nsteps):
    c = cases.get(name)
    s = c.system
    P, N = s.pi_beads, s.natm
    v0 = pimd_velocities(c, w.SEED_BASE + 200)
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        v0.astype(np.float64).tofile(os.path.join(d, "vel.bin"))
        with open(os.path.join(d, "stdout.txt"), "w") as log:
            r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin", "pimd", str(nsteps), "vel.bin"], cwd=d,
                               stdout=log, stderr=subprocess.STDOUT, timeout=1200)
        if r.returncode != 0:
            print(open(os.path.join(d, "stdout.txt")).read()[-3000:])
            raise RuntimeError(f"pimd {name}: reference run failed ({r.returncode})")
        raw = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
        # the reference's restart file of the final state (write_dump_file_pimd)
        restart = open(os.path.join(d, "water.restart")).read()
    hdr = raw[:9]
    assert int(hdr[0]) == N and int(hdr[1]) == P
    lc, nc, lb, nb = (int(v) for v in hdr[5:9])
    o = 9
    e = raw[o:o + 5]
    o += 5
    pos = raw[o:o + 3 * N * P].reshape(P, N, 3)
    o += 3 * N * P
    vel = raw[o:o + 3 * N * P].reshape(P, N, 3)
    o += 3 * N * P
    cls = raw[o:o + 2 * lc]
    o += 2 * lc
    bead = raw[o:o + (P - 1) * N * 3 * 2 * lb].reshape(P - 1, N, 3, 2, lb)
    o += (P - 1) * N * 3 * 2 * lb
    assert o == len(raw), (o, len(raw))
    out = os.path.join(HERE, f"ref_pimd_{name}.npz")
    np.savez_compressed(out, case=name, nsteps=nsteps, dt=hdr[4], len_nhc=lc, nchain_class=nc, nchain_bead=nb,
                        v0=v0, energies=e, energy_keys=np.array(["vintert", "vintrat", "kin_harm", "pi_ke_prim",
                                                                 "pi_ke_vir"]),
                        pos=pos, vel=vel, nhc_class=cls, nhc_bead=bead, restart=np.array(restart))
    print(f"pimd {name}: {nsteps} steps dt={hdr[4]:.6g} au, chains {nc} + {nb}, vintert={e[0]:.12g} -> {out}")


MD_CASES = {"cfg1": 100}   # configuration 1: 100 velocity-Verlet steps at 0.25 fs


def md_velocities(c, seed):
    """Seeded Maxwell-Boltzmann velocities at the case temperature (Bohr per
    atomic time unit), no centre-of-mass motion, zero on ghost sites."""
    s = c.system
    rng = np.random.Generator(np.random.PCG64(seed))
    kt = s.settings.temperature / w.BOLTZ
    m = np.asarray(s.mass, dtype=np.float64)
    v = rng.normal(size=(s.natm, 3)) * np.sqrt(kt / m)[:, None]
    ghost = np.zeros(s.natm, bool)
    ghost[np.asarray(s.ghost_atom, dtype=int)] = True
    v[ghost] = 0.0
    real = ~ghost
    v[real] -= (m[real, None] * v[real]).sum(0) / m[real].sum()
    return v


def run_md_case(name, nsteps):
    c = cases.get(name)
    s = c.system
    assert s.pi_beads == 1
    v0 = md_velocities(c, w.SEED_BASE + 100)
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, c)
        v0.astype(np.float64).tofile(os.path.join(d, "vel.bin"))
        with open(os.path.join(d, "stdout.txt"), "w") as log:
            r = subprocess.run([EXE, "water.input", "pos.bin", "out.bin", "md", str(nsteps), "vel.bin"], cwd=d,
                               stdout=log, stderr=subprocess.STDOUT, timeout=1200)
        if r.returncode != 0:
            print(open(os.path.join(d, "stdout.txt")).read()[-3000:])
            raise RuntimeError(f"md {name}: reference run failed ({r.returncode})")
        raw = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
        restart = open(os.path.join(d, "water.restart")).read()  # write_dump_file_md
    N = int(raw[0])
    vintert, vintrat, kinet, dt = raw[3:7]
    xvf = raw[7:7 + 9 * N].reshape(N, 9)
    out = os.path.join(HERE, f"ref_md_{name}.npz")
    np.savez_compressed(out, case=name, nsteps=nsteps, dt=dt, v0=v0, vintert=vintert, vintrat=vintrat,
                        kinet=kinet, x=xvf[:, 0:3], v=xvf[:, 3:6], f=xvf[:, 6:9], restart=np.array(restart))
    print(f"md {name}: {nsteps} steps dt={dt:.6g} au, vintert={vintert:.12g} kinet={kinet:.12g} -> {out}")


def main():
    names = sys.argv[1:] or CASES + ["md", "setup", "w512_debug", "pimd", "respa"]
    for n in names:
        if n == "md":
            for m, k in MD_CASES.items():
                run_md_case(m, k)
        elif n == "setup":
            for m in SETUP_CASES:
                run_setup_case(m)
        elif n == "w512_debug":
            run_w512_debug()
        elif n == "pimd":
            for m, k in PIMD_CASES.items():
                run_pimd_case(m, k)
        elif n == "respa":
            for m in RESPA_CASES:
                run_respa_case(m)
        else:
            run_case(n)


if __name__ == "__main__":
    main()
