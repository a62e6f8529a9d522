/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A driver for the reference build (oracle/ref.mk): it links the
 * reference's own objects, reads a PINY input directory with the
 * reference's parse(), overwrites the bead positions with exact values, and
 * runs the reference force call once, as python/piny/piny.pyx:170-188,357-359
 * and 426-436 do.  Results are written as raw doubles for the golden-vector
 * script under tests/golden/.  Nothing in the product links it.
 *
 * usage: piny_ref <input file> <positions.bin> <out.bin> [repeat]
 *        piny_ref <input file> <positions.bin> <out.bin> md <nsteps> <velocities.bin>
 *        piny_ref <input file> <positions.bin> <out.bin> dump <setup.bin>
 *        piny_ref <input file> <positions.bin> <out.bin> recip
 *        piny_ref <input file> <positions.bin> <out.bin> pimd <nsteps> <velocities.bin>
 *   positions.bin: P*N*3 doubles, Bohr, [bead][site][x,y,z] (ghost sites
 *                  are recomputed by get_ghost_pos)
 *   out.bin:       header {N, P, nver} as doubles, then 24 energy/virial
 *                  doubles (see the list below), pvten_tot, then forces
 *                  [bead][site][3] (for P > 1: normal-mode forces including
 *                  the springs, as energy_control_pimd leaves them) and,
 *                  for P > 1, the PIMD virial forces fxt in the same layout
 *   repeat:        optional; after the first call (which builds the Verlet
 *                  list) the force call is made this many more times on the
 *                  same positions and the wall time of those calls
 *                  (CLOCK_MONOTONIC, as src/output/timer.c:55-64) is printed
 *                  as "REPEAT <calls> <seconds>"
 *   dump:          after the force call, the setup products and the Verlet
 *                  list of that call are written to setup.bin: int32
 *                  {N, nver, ninter, nsplin, nktot, ecor_nsplin, 0, 0}, then
 *                  int32 nter[N], jver_off[N], jter[nver] as the reference
 *                  stores them (jter 1-based, jver_off the 0-based exclusive prefix), float64 cv0, cdv0, cv0_c, cdv0_c
 *                  [ninter][nsplin] (each type pair's table through
 *                  inter_map_index), int32 kastr, kbstr, kcstr [nktot],
 *                  float64 bweight_tot [nktot] (PME only, else zeros),
 *                  float64 ecor cv0, cdv0 [ecor_nsplin]
 *   recip:         also the reference's debug driver's reciprocal-space
 *                  check (test_force with inter = 3, test_energy.c:163-208,
 *                  489-600) on bead 1, printed as "RECIP <energy>"
 *   pimd:          P > 1; after the first force call the mode velocities
 *                  (P*N*3 doubles, [bead][site][3], Bohr per atomic time
 *                  unit) are set, every Nose-Hoover chain position and
 *                  velocity is zeroed, and int_NVT_pimd runs nsteps times;
 *                  out.bin then holds {N, P, nver, nsteps, dt, len and
 *                  number of chains of the centroid and of the bead
 *                  thermostats}, {vintert, vintrat, kin_harm, pi_ke_prim,
 *                  pi_ke_vir}, the bead positions and the mode velocities
 *                  [bead][site][3], the centroid chain's v[len], x[len], and
 *                  for beads 2..P, sites and x, y, z the chain's v[len],
 *                  x[len] (zeros for ghost sites); the reference's own
 *                  restart writer (write_dump_file_pimd) then writes the
 *                  final state to the input's out_restart_file
 *   respa:         the force call with RESPA flags: argv[5], argv[6] and
 *                  argv[7] are iget_full_inter, iget_res_inter and
 *                  iget_full_intra; iget_res_intra is 1, as
 *                  int_NVT_pimd.c:174-187 and the RESPA integrators of md/
 *                  set it at the end of the innermost intramolecular step.  The weights follow from
 *                  the input's respa_steps_* keywords
 *                  (energy_control_initial.c:212-243); the real-space energy
 *                  and pressure flags are left to the force call itself
 *                  (force_npol.c:146-158).  out.bin as for a plain call.
 *   md:            P = 1; after the first force call the velocities
 *                  (N*3 doubles, Bohr per atomic time unit) are set and
 *                  int_NVE (md/int_NVE.c) runs nsteps times with the
 *                  input's time step; out.bin then holds the header, the
 *                  energies of the last force call, kinet, and the final
 *                  positions, velocities and forces, [site][3] each; the
 *                  reference's restart writer (write_dump_file_md) then
 *                  writes the final state to the input's out_restart_file
 */
#include <time.h>
#include "standard_include.h"
#include "../typ_defs/typedefs_class.h"
#include "../typ_defs/typedefs_bnd.h"
#include "../typ_defs/typedefs_gen.h"
#include "../typ_defs/typedefs_cp.h"
#include "../typ_defs/typedefs_par.h"
#include "../typ_defs/typedefs_stat.h"
#include "../proto_defs/proto_parse_entry.h"
#include "../proto_defs/proto_communicate_wrappers.h"
#include "../proto_defs/proto_energy_ctrl_entry.h"
#include "../proto_defs/proto_intra_con_entry.h"
#include "../proto_defs/proto_pimd_entry.h"
#include "../proto_defs/proto_integrate_md_entry.h"

/* the debug driver's per-term energy (src/energy/control/test_energy.c:489) */
void test_force(CLASS *, BONDED *, GENERAL_DATA *, int, double *, double *, double *);
/* the restart writers (output/md/output_md.c:757, output/pimd/output_pimd.c:786) */
void write_dump_file_md(CLASS *, BONDED *, GENERAL_DATA *);
void write_dump_file_pimd(CLASS *, BONDED *, GENERAL_DATA *);

static CLASS cls;
static BONDED bonded;
static GENERAL_DATA gd;
static CP cp;
static ANALYSIS analysis;

int main(int argc, char *argv[]) {
  const int md = argc == 7 && strcmp(argv[4], "md") == 0;
  const int dump = argc == 6 && strcmp(argv[4], "dump") == 0;
  const int recip = argc == 5 && strcmp(argv[4], "recip") == 0;
  const int pimd = argc == 7 && strcmp(argv[4], "pimd") == 0;
  const int respa = argc == 8 && strcmp(argv[4], "respa") == 0;
  if (argc != 4 && argc != 5 && !md && !dump && !pimd && !respa) {
    fprintf(stderr, "usage: %s <input file> <positions.bin> <out.bin> [repeat | md nsteps velocities.bin]\n",
            argv[0]);
    return 2;
  }
  const int repeat = (argc == 5 && !recip) ? atoi(argv[4]) : 0;
  const char *dump_file = dump ? argv[5] : NULL;
  int zero = 0;
  char **noargs = NULL;
  Init(&zero, &noargs, &(cls.communicate.world));
  Comm_size(cls.communicate.world, &(cls.communicate.np));
  Comm_rank(cls.communicate.world, &(cls.communicate.myid));
  gd.error_check_on = 1;
  parse(&cls, &bonded, &gd, &cp, &analysis, argv[1]);

  cls.energy_ctrl.iget_full_inter = 1;
  cls.energy_ctrl.iget_res_inter = 0;
  cls.energy_ctrl.iget_full_intra = 1;
  cls.energy_ctrl.iget_res_intra = 0;
  /* real-space energy and pressure every call (test_energy.c:506-507) */
  cls.energy_ctrl.iget_pe_real_inter = 1;
  cls.energy_ctrl.iget_pv_real_inter = 1;
  if (respa) {
    cls.energy_ctrl.iget_full_inter = atoi(argv[5]);
    cls.energy_ctrl.iget_res_inter = atoi(argv[6]);
    cls.energy_ctrl.iget_full_intra = atoi(argv[7]);
    cls.energy_ctrl.iget_res_intra = 1;
    cls.energy_ctrl.itime = 1;
  }

  const int N = cls.clatoms_info.natm_tot, P = cls.clatoms_info.pi_beads;
  FILE *f = fopen(argv[2], "rb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)N * P);
  if (fread(pos, sizeof(double), 3 * (size_t)N * P, f) != 3 * (size_t)N * P) {
    fprintf(stderr, "short positions file\n");
    return 1;
  }
  fclose(f);
  for (int ip = 1; ip <= P; ip++) {
    for (int i = 1; i <= N; i++) {
      const double *p = pos + 3 * ((size_t)(ip - 1) * N + (i - 1));
      cls.clatoms_pos[ip].x[i] = p[0];
      cls.clatoms_pos[ip].y[i] = p[1];
      cls.clatoms_pos[ip].z[i] = p[2];
    }
    get_ghost_pos(&(cls.clatoms_info), &(cls.clatoms_pos[ip]), &(cls.ghost_atoms));
  }
  if (P > 1) {
    /* the integrator's order (int_0_to_dt2_pimd.c:156-175): beads to
     * modes, spring forces fxm and kin_harm in mode coordinates, modes back
     * to beads (which sets xmod from mode 1), then the force call, which
     * returns mode forces plus fxm (energy_control_pimd.c:198-214) */
    cls.clatoms_info.wght_pimd = 1.0; /* as control_pimd.c:311 */
    control_pimd_trans_mode(&cls, &gd);
    mode_energy_control(&cls, &gd);
    control_pimd_trans_pos(&cls, &gd);
    for (int ip = 1; ip <= P; ip++) get_ghost_pos(&(cls.clatoms_info), &(cls.clatoms_pos[ip]), &(cls.ghost_atoms));
    energy_control_pimd(&cls, &bonded, &gd);
    if (pimd) {
      const int nsteps = atoi(argv[5]);
      FILE *fv = fopen(argv[6], "rb");
      if (!fv) return 1;
      double *vel = (double *)malloc(sizeof(double) * 3 * (size_t)N * P);
      if (fread(vel, sizeof(double), 3 * (size_t)N * P, fv) != 3 * (size_t)N * P) return 1;
      fclose(fv);
      for (int ip = 1; ip <= P; ip++)
        for (int i = 1; i <= N; i++) {
          const double *v = vel + 3 * ((size_t)(ip - 1) * N + (i - 1));
          cls.clatoms_pos[ip].vx[i] = v[0];
          cls.clatoms_pos[ip].vy[i] = v[1];
          cls.clatoms_pos[ip].vz[i] = v[2];
        }
      free(vel);
      THERM_INFO *tc = &cls.therm_info_class, *tb = &cls.therm_info_bead;
      for (int j = 1; j <= tc->len_nhc; j++)
        for (int k = 1; k <= tc->num_nhc; k++) cls.therm_class.v_nhc[j][k] = cls.therm_class.x_nhc[j][k] = 0.0;
      for (int ip = 2; ip <= P; ip++)
        for (int j = 1; j <= tb->len_nhc; j++)
          for (int k = 1; k <= tb->num_nhc; k++) cls.therm_bead[ip].v_nhc[j][k] = cls.therm_bead[ip].x_nhc[j][k] = 0.0;
      for (int st = 0; st < nsteps; st++) int_NVT_pimd(&cls, &bonded, &gd, &analysis);
      write_dump_file_pimd(&cls, &bonded, &gd);
      STAT_AVG *sa = &gd.stat_avg;
      double hdr[9] = {(double)N, (double)P, (double)cls.nbr_list.verlist.nver_lst_now, (double)nsteps,
                       gd.timeinfo.dt, (double)tc->len_nhc, (double)tc->num_nhc, (double)tb->len_nhc,
                       (double)tb->num_nhc};
      double e[5] = {sa->vintert, sa->vintrat, sa->kin_harm, sa->pi_ke_prim, sa->pi_ke_vir};
      FILE *fo = fopen(argv[3], "wb");
      if (!fo) return 1;
      fwrite(hdr, sizeof(double), 9, fo);
      fwrite(e, sizeof(double), 5, fo);
      for (int ip = 1; ip <= P; ip++)
        for (int i = 1; i <= N; i++) {
          double v[3] = {cls.clatoms_pos[ip].x[i], cls.clatoms_pos[ip].y[i], cls.clatoms_pos[ip].z[i]};
          fwrite(v, sizeof(double), 3, fo);
        }
      for (int ip = 1; ip <= P; ip++)
        for (int i = 1; i <= N; i++) {
          double v[3] = {cls.clatoms_pos[ip].vx[i], cls.clatoms_pos[ip].vy[i], cls.clatoms_pos[ip].vz[i]};
          fwrite(v, sizeof(double), 3, fo);
        }
      for (int j = 1; j <= tc->len_nhc; j++) fwrite(&cls.therm_class.v_nhc[j][1], sizeof(double), 1, fo);
      for (int j = 1; j <= tc->len_nhc; j++) fwrite(&cls.therm_class.x_nhc[j][1], sizeof(double), 1, fo);
      for (int ip = 2; ip <= P; ip++)
        for (int i = 1; i <= N; i++) {
          const int *maps[3] = {tb->inhc_x, tb->inhc_y, tb->inhc_z};
          for (int d = 0; d < 3; d++) {
            const int k = maps[d][i];
            for (int j = 1; j <= tb->len_nhc; j++) {
              double v = (k >= 1 && k <= tb->num_nhc) ? cls.therm_bead[ip].v_nhc[j][k] : 0.0;
              fwrite(&v, sizeof(double), 1, fo);
            }
            for (int j = 1; j <= tb->len_nhc; j++) {
              double v = (k >= 1 && k <= tb->num_nhc) ? cls.therm_bead[ip].x_nhc[j][k] : 0.0;
              fwrite(&v, sizeof(double), 1, fo);
            }
          }
        }
      fclose(fo);
      free(pos);
      return 0;
    }
  } else {
    energy_control(&cls, &bonded, &gd);
    if (md) {
      const int nsteps = atoi(argv[5]);
      FILE *fv = fopen(argv[6], "rb");
      if (!fv) return 1;
      double *vel = (double *)malloc(sizeof(double) * 3 * (size_t)N);
      if (fread(vel, sizeof(double), 3 * (size_t)N, fv) != 3 * (size_t)N) return 1;
      fclose(fv);
      for (int i = 1; i <= N; i++) {
        cls.clatoms_pos[1].vx[i] = vel[3 * (i - 1)];
        cls.clatoms_pos[1].vy[i] = vel[3 * (i - 1) + 1];
        cls.clatoms_pos[1].vz[i] = vel[3 * (i - 1) + 2];
      }
      free(vel);
      for (int st = 0; st < nsteps; st++) int_NVE(&cls, &bonded, &gd);
      write_dump_file_md(&cls, &bonded, &gd);
      STAT_AVG *s = &gd.stat_avg;
      double hdr[3] = {(double)N, (double)P, (double)cls.nbr_list.verlist.nver_lst_now};
      double e[4] = {s->vintert, s->vintrat, s->kinet, gd.timeinfo.dt};
      FILE *fo = fopen(argv[3], "wb");
      if (!fo) return 1;
      fwrite(hdr, sizeof(double), 3, fo);
      fwrite(e, sizeof(double), 4, fo);
      for (int i = 1; i <= N; i++) {
        double v[9] = {cls.clatoms_pos[1].x[i],  cls.clatoms_pos[1].y[i],  cls.clatoms_pos[1].z[i],
                       cls.clatoms_pos[1].vx[i], cls.clatoms_pos[1].vy[i], cls.clatoms_pos[1].vz[i],
                       cls.clatoms_pos[1].fx[i], cls.clatoms_pos[1].fy[i], cls.clatoms_pos[1].fz[i]};
        fwrite(v, sizeof(double), 9, fo);
      }
      fclose(fo);
      free(pos);
      return 0;
    }
    if (repeat > 0) {
      struct timespec t0, t1;
      clock_gettime(CLOCK_MONOTONIC, &t0);
      for (int r = 0; r < repeat; r++) energy_control(&cls, &bonded, &gd);
      clock_gettime(CLOCK_MONOTONIC, &t1);
      printf("REPEAT %d %.6f\n", repeat, (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));
      fflush(stdout);
    }
  }

  if (recip) {
    double vg = 0.0, vg2 = 0.0, vg3 = 0.0;
    get_ghost_pos(&(cls.clatoms_info), &(cls.clatoms_pos[1]), &(cls.ghost_atoms));
    test_force(&cls, &bonded, &gd, 3, &vg, &vg2, &vg3);
    printf("RECIP %.13g\n", vg);
    fflush(stdout);
  }
  if (dump_file) {
    VERLIST *vl = &cls.nbr_list.verlist;
    INTERACT *it = &cls.interact;
    EWALD *ew = &gd.ewald;
    ECOR *ec = &bonded.ecor;
    const int nver = vl->nver_lst_now, ninter = (cls.atommaps.natm_typ * (cls.atommaps.natm_typ + 1)) / 2;
    const int nsp = it->nsplin, nk = ew->nktot, nse = ec->nsplin;
    int hdr8[8] = {N, nver, ninter, nsp, nk, nse, 0, 0};
    FILE *fd = fopen(dump_file, "wb");
    if (!fd) return 1;
    fwrite(hdr8, sizeof(int), 8, fd);
    fwrite(&vl->nter[1], sizeof(int), N, fd);
    fwrite(&vl->jver_off[1], sizeof(int), N, fd);
    fwrite(&vl->jter[1], sizeof(int), nver, fd);
    double *tabs[4] = {it->cv0, it->cdv0, it->cv0_c, it->cdv0_c};
    for (int w = 0; w < 4; w++)
      for (int t = 1; t <= ninter; t++) fwrite(&tabs[w][(it->inter_map_index[t] - 1) * nsp + 1], sizeof(double), nsp, fd);
    fwrite(&ew->kastr[1], sizeof(int), nk, fd);
    fwrite(&ew->kbstr[1], sizeof(int), nk, fd);
    fwrite(&ew->kcstr[1], sizeof(int), nk, fd);
    for (int k = 1; k <= nk; k++) {
      double b = cls.part_mesh.bweight_tot ? cls.part_mesh.bweight_tot[k] : 0.0;
      fwrite(&b, sizeof(double), 1, fd);
    }
    fwrite(&ec->cv0[1], sizeof(double), nse, fd);
    fwrite(&ec->cdv0[1], sizeof(double), nse, fd);
    fclose(fd);
  }

  STAT_AVG *s = &gd.stat_avg;
  double hdr[3] = {(double)N, (double)P, (double)cls.nbr_list.verlist.nver_lst_now};
  double e[24] = {s->vintert, s->vintrat, s->vreal, s->vrecip, s->vvdw, s->vcoul, s->vlong,
                  s->vbondt,  s->vbendt,  s->pi_ke_prim, s->pi_ke_vir, s->kin_harm};
  /* e[12..20]: pvten[1..9]; e[21..23] unused */
  for (int k = 0; k < 9; k++) e[12 + k] = gd.ptens.pvten[k + 1];
  f = fopen(argv[3], "wb");
  if (!f) return 1;
  fwrite(hdr, sizeof(double), 3, f);
  fwrite(e, sizeof(double), 24, f);
  for (int k = 0; k < 9; k++) fwrite(&gd.ptens.pvten_tot[k + 1], sizeof(double), 1, f);
  for (int ip = 1; ip <= P; ip++)
    for (int i = 1; i <= N; i++) {
      double v[3] = {cls.clatoms_pos[ip].fx[i], cls.clatoms_pos[ip].fy[i], cls.clatoms_pos[ip].fz[i]};
      fwrite(v, sizeof(double), 3, f);
    }
  if (P > 1)
    for (int ip = 1; ip <= P; ip++)
      for (int i = 1; i <= N; i++) {
        double v[3] = {cls.clatoms_pos[ip].fxt[i], cls.clatoms_pos[ip].fyt[i], cls.clatoms_pos[ip].fzt[i]};
        fwrite(v, sizeof(double), 3, f);
      }
  fclose(f);
  free(pos);
  return 0;
}
