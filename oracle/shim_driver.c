/*
 * shim_driver.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A PINY program around the drop-in (piny_amd/shim/pmx_piny.c): links the
 * reference's own objects (oracle/ref.mk), reads a PINY input directory
 * with the reference's parse(), sets exact bead positions as
 * oracle/ref_harness.c does, and fills pmx_system / pmx_params with
 * pmx_piny_fill.  With "run" it then makes the reference's force call
 * (energy_control, or for P > 1 the integrator's order: modes, springs,
 * back transform, energy_control_pimd) and the library's through
 * pmx_piny_energy_control on device 0, on the same PINY structures.
 *
 * usage: piny_shim <input file> <positions.bin> <parse.bin> <out.bin> [run]
 *   parse.bin: the values parse() does not keep (pmx_piny_parse), doubles:
 *              npair, then npair labels, npair sigma (Bohr), npair epsilon
 *              (Hartree), kmax_ewd, ishift
 *   out.bin:   doubles: N, P, ntype, npair, nbond, nbond_type, nbend,
 *              nbend_type, nghost; atm_type[N], q[N], mass[N]; per pair
 *              label, sig, eps, cutti, cutoff, cutoff_res; bonds j1, j2,
 *              type; bond c[7 ntype], eq; bends j1, j2, j3, type; bend c,
 *              s; ghosts atom, ncomp, comp[4], coef[4]; params hmat[9],
 *              pi_beads, temperature, pimd_typ, nsplin, ishift, rheal, skin,
 *              ewald_alpha, kmax_ewd, pme_on, kmax_pme, n_interp,
 *              intra_perds; with run: for the reference and then the
 *              library, 12 doubles (vintert, vintrat, vrecip, vlong, vbondt,
 *              vbendt, pi_ke_prim, pi_ke_vir, kin_harm, and 3 zeros), pvten[9],
 *              pvten_tot[9] and the forces [P][N][3]
 */
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
#include "pmx_piny.h"

static CLASS cls;
static BONDED bonded;
static GENERAL_DATA gd;
static CP cp;
static ANALYSIS analysis;
static FILE *fo;

static void put(double v) { fwrite(&v, sizeof v, 1, fo); }
static void puti(const int *a, int n) {
  for (int k = 0; k < n; k++) put((double)a[k]);
}
static void putd(const double *a, int n) { fwrite(a, sizeof(double), (size_t)n, fo); }

static void put_results(int N, int P) {
  STAT_AVG *s = &gd.stat_avg;
  double e[12] = {s->vintert, s->vintrat, s->vrecip, s->vlong, s->vbondt, s->vbendt,
                  s->pi_ke_prim, s->pi_ke_vir, s->kin_harm, 0, 0, 0};
  putd(e, 12);
  putd(&gd.ptens.pvten[1], 9);
  putd(&gd.ptens.pvten_tot[1], 9);
  for (int ip = 1; ip <= P; ip++)
    for (int i = 1; i <= N; i++) {
      put(cls.clatoms_pos[ip].fx[i]);
      put(cls.clatoms_pos[ip].fy[i]);
      put(cls.clatoms_pos[ip].fz[i]);
    }
}

int main(int argc, char *argv[]) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "usage: %s <input file> <positions.bin> <parse.bin> <out.bin> [run]\n", argv[0]);
    return 2;
  }
  const int run = argc == 6 && strcmp(argv[5], "run") == 0;
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
  cls.energy_ctrl.iget_pe_real_inter = 1;
  cls.energy_ctrl.iget_pv_real_inter = 1;
  const int N = cls.clatoms_info.natm_tot, P = cls.clatoms_info.pi_beads;
  FILE *f = fopen(argv[2], "rb");
  if (!f) return 1;
  double *pos = (double *)malloc(sizeof(double) * 3 * (size_t)N * P);
  if (fread(pos, sizeof(double), 3 * (size_t)N * P, f) != 3 * (size_t)N * P) return 1;
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
  /* the parse-time values */
  f = fopen(argv[3], "rb");
  if (!f) return 1;
  double np_d;
  if (fread(&np_d, sizeof np_d, 1, f) != 1) return 1;
  const int npair = (int)np_d;
  double *pv = (double *)malloc(sizeof(double) * (3 * (size_t)npair + 2));
  if (fread(pv, sizeof(double), 3 * (size_t)npair + 2, f) != 3 * (size_t)npair + 2) return 1;
  fclose(f);
  int *lab = (int *)malloc(sizeof(int) * npair);
  for (int t = 0; t < npair; t++) lab[t] = (int)pv[t];
  pmx_piny_parse pp = {lab, pv + npair, pv + 2 * npair, (int)pv[3 * npair], (int)pv[3 * npair + 1]};
  pmx_piny_sys ps;
  if (pmx_piny_fill(&cls, &bonded, &gd, &pp, &ps) != PMX_OK) {
    fprintf(stderr, "pmx_piny_fill failed\n");
    return 1;
  }
  const pmx_system *y = &ps.sys;
  const pmx_params *r = &ps.par;
  fo = fopen(argv[4], "wb");
  if (!fo) return 1;
  const int np = y->ntype * (y->ntype + 1) / 2;
  double hdr[9] = {y->natm, P, y->ntype, np, y->nbond, y->nbond_type, y->nbend, y->nbend_type, y->nghost};
  putd(hdr, 9);
  puti(y->atm_type, N);
  putd(y->q, N);
  putd(y->mass, N);
  puti(y->inter_label, np);
  putd(y->inter_sig, np);
  putd(y->inter_eps, np);
  putd(y->inter_cutti, np);
  putd(y->inter_cutoff, np);
  putd(y->inter_cutoff_res, np);
  puti(y->bond_j1, y->nbond);
  puti(y->bond_j2, y->nbond);
  puti(y->bond_type, y->nbond);
  putd(y->bond_c, 7 * y->nbond_type);
  putd(y->bond_eq, y->nbond_type);
  puti(y->bend_j1, y->nbend);
  puti(y->bend_j2, y->nbend);
  puti(y->bend_j3, y->nbend);
  puti(y->bend_type, y->nbend);
  putd(y->bend_c, 7 * y->nbend_type);
  putd(y->bend_s, 7 * y->nbend_type);
  puti(y->ghost_atom, y->nghost);
  puti(y->ghost_ncomp, y->nghost);
  puti(y->ghost_comp, 4 * y->nghost);
  putd(y->ghost_coef, 4 * y->nghost);
  putd(r->hmat, 9);
  double pr[14] = {r->pi_beads, r->temperature, r->pimd_typ, r->nsplin, r->ishift, r->rheal, r->skin,
                   r->ewald_alpha, r->kmax_ewd, r->pme_on, r->kmax_pme, r->n_interp, r->intra_perds, 0};
  putd(pr, 13);
  if (run) {
    /* the reference's force call */
    if (P > 1) {
      cls.clatoms_info.wght_pimd = 1.0; /* as control_pimd.c:311 */
      control_pimd_trans_mode(&cls, &gd);
      mode_energy_control(&cls, &gd);
      control_pimd_trans_pos(&cls, &gd);
      for (int ip = 1; ip <= P; ip++) get_ghost_pos(&(cls.clatoms_info), &(cls.clatoms_pos[ip]), &(cls.ghost_atoms));
      energy_control_pimd(&cls, &bonded, &gd);
    } else {
      energy_control(&cls, &bonded, &gd);
    }
    put_results(N, P);
    /* the library's, on the same structures */
    char err[512];
    pmx_ctx *ctx = pmx_create(&ps.sys, &ps.par, NULL, NULL, NULL, 0, err, sizeof err);
    if (!ctx) {
      fprintf(stderr, "pmx_create: %s\n", err);
      return 1;
    }
    pmx_energies e;
    const unsigned flags = P > 1 ? (PMX_EVAL_PIMD | PMX_EVAL_MODE_FORCES) : 0;
    if (pmx_piny_energy_control(ctx, &ps, &cls, &gd, flags, &e) != PMX_OK) {
      fprintf(stderr, "pmx_piny_energy_control: %s\n", pmx_last_error(ctx));
      return 1;
    }
    put_results(N, P);
    pmx_destroy(ctx);
  }
  fclose(fo);
  pmx_piny_free(&ps);
  free(pos);
  free(pv);
  free(lab);
  return 0;
}
