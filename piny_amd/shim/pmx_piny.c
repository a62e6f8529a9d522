/*
 * pmx_piny.c -- the PINY side of the drop-in (see pmx_piny.h): PINY's
 * structures after parse() to the library's flat arrays, and the force
 * call on PINY's position and force arrays.  PINY indexes atoms, types and
 * terms from 1; the library from 0.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "standard_include.h"
#include "../typ_defs/typedefs_class.h"
#include "../typ_defs/typedefs_bnd.h"
#include "../typ_defs/typedefs_gen.h"
#include "pmx_piny.h"

typedef struct {
  int *atm_type, *bond_j1, *bond_j2, *bond_type, *bend_j1, *bend_j2, *bend_j3, *bend_type;
  int *ghost_atom, *ghost_ncomp, *ghost_comp, *inter_label;
  double *q, *mass, *sig, *eps, *cutti, *cutoff, *cutoff_res, *bond_c, *bond_eq, *bend_c, *bend_s, *ghost_coef;
  double *px, *py, *pz, *fx, *fy, *fz, *xm, *ym, *zm;  /* [P*N] scratch of the force call */
} store_t;

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

int pmx_piny_fill(CLASS *class, BONDED *bonded, GENERAL_DATA *gd, const pmx_piny_parse *pp, pmx_piny_sys *out) {
  memset(out, 0, sizeof(*out));
  store_t *s = (store_t *)zalloc(sizeof(store_t));
  out->store = s;
  const int N = class->clatoms_info.natm_tot, P = class->clatoms_info.pi_beads;
  const int nt = class->atommaps.natm_typ, npair = nt * (nt + 1) / 2;
  pmx_system *y = &out->sys;
  y->natm = N;
  y->ntype = nt;
  s->atm_type = (int *)zalloc(sizeof(int) * N);
  s->q = (double *)zalloc(sizeof(double) * N);
  s->mass = (double *)zalloc(sizeof(double) * N);
  for (int i = 0; i < N; i++) {
    s->atm_type[i] = class->atommaps.iatm_atm_typ[i + 1] - 1;
    s->q[i] = class->clatoms_info.q[i + 1];
    s->mass[i] = class->clatoms_info.mass[i + 1]; /* amu * PROT_MASS (close_intra_params.c:215) */
  }
  /* type pairs: cutoffs per pair (interact.cutoff, cutoff_res, cutti are
   * indexed by pair, force_npol.c:760-763); the LJ parameters from pp */
  s->inter_label = (int *)zalloc(sizeof(int) * npair);
  s->sig = (double *)zalloc(sizeof(double) * npair);
  s->eps = (double *)zalloc(sizeof(double) * npair);
  s->cutti = (double *)zalloc(sizeof(double) * npair);
  s->cutoff = (double *)zalloc(sizeof(double) * npair);
  s->cutoff_res = (double *)zalloc(sizeof(double) * npair);
  for (int t = 0; t < npair; t++) {
    s->inter_label[t] = pp->inter_label[t];
    s->sig[t] = pp->inter_sig[t];
    s->eps[t] = pp->inter_eps[t];
    s->cutti[t] = class->interact.cutti[t + 1];
    s->cutoff[t] = class->interact.cutoff[t + 1];
    s->cutoff_res[t] = class->interact.cutoff_res[t + 1];
  }
  y->inter_label = s->inter_label;
  y->inter_sig = s->sig;
  y->inter_eps = s->eps;
  y->inter_cutti = s->cutti;
  y->inter_cutoff = s->cutoff;
  y->inter_cutoff_res = s->cutoff_res;
  /* bonds and bends: the power-series terms (intra_coefs.c:30-375) */
  const BOND *b = &bonded->bond;
  y->nbond = b->npow;
  y->nbond_type = b->ntyp_pow;
  s->bond_j1 = (int *)zalloc(sizeof(int) * b->npow);
  s->bond_j2 = (int *)zalloc(sizeof(int) * b->npow);
  s->bond_type = (int *)zalloc(sizeof(int) * b->npow);
  for (int k = 0; k < b->npow; k++) {
    s->bond_j1[k] = b->j1_pow[k + 1] - 1;
    s->bond_j2[k] = b->j2_pow[k + 1] - 1;
    s->bond_type[k] = b->jtyp_pow[k + 1] - 1;
  }
  s->bond_c = (double *)zalloc(sizeof(double) * 7 * b->ntyp_pow);
  s->bond_eq = (double *)zalloc(sizeof(double) * b->ntyp_pow);
  for (int t = 0; t < b->ntyp_pow; t++) {
    const double c[7] = {b->c_0[t + 1], b->c_1[t + 1], b->c_2[t + 1], b->c_3[t + 1],
                         b->c_4[t + 1], b->c_5[t + 1], b->c_6[t + 1]};
    memcpy(s->bond_c + 7 * t, c, sizeof c);
    s->bond_eq[t] = b->eq_pow[t + 1];
  }
  y->bond_j1 = s->bond_j1;
  y->bond_j2 = s->bond_j2;
  y->bond_type = s->bond_type;
  y->bond_c = s->bond_c;
  y->bond_eq = s->bond_eq;
  const BEND *a = &bonded->bend;
  y->nbend = a->npow;
  y->nbend_type = a->ntyp_pow;
  s->bend_j1 = (int *)zalloc(sizeof(int) * a->npow);
  s->bend_j2 = (int *)zalloc(sizeof(int) * a->npow);
  s->bend_j3 = (int *)zalloc(sizeof(int) * a->npow);
  s->bend_type = (int *)zalloc(sizeof(int) * a->npow);
  for (int k = 0; k < a->npow; k++) {
    s->bend_j1[k] = a->j1_pow[k + 1] - 1;
    s->bend_j2[k] = a->j2_pow[k + 1] - 1;
    s->bend_j3[k] = a->j3_pow[k + 1] - 1;
    s->bend_type[k] = a->jtyp_pow[k + 1] - 1;
  }
  s->bend_c = (double *)zalloc(sizeof(double) * 7 * a->ntyp_pow);
  s->bend_s = (double *)zalloc(sizeof(double) * 7 * a->ntyp_pow);
  for (int t = 0; t < a->ntyp_pow; t++) {
    const double c[7] = {a->c_0[t + 1], a->c_1[t + 1], a->c_2[t + 1], a->c_3[t + 1],
                         a->c_4[t + 1], a->c_5[t + 1], a->c_6[t + 1]};
    const double sn[7] = {a->s_0[t + 1], a->s_1[t + 1], a->s_2[t + 1], a->s_3[t + 1],
                          a->s_4[t + 1], a->s_5[t + 1], a->s_6[t + 1]};
    memcpy(s->bend_c + 7 * t, c, sizeof c);
    memcpy(s->bend_s + 7 * t, sn, sizeof sn);
  }
  y->bend_j1 = s->bend_j1;
  y->bend_j2 = s->bend_j2;
  y->bend_j3 = s->bend_j3;
  y->bend_type = s->bend_type;
  y->bend_c = s->bend_c;
  y->bend_s = s->bend_s;
  /* ghost sites (ghost_control.c:28-64) */
  const GHOST_ATOMS *g = &class->ghost_atoms;
  y->nghost = g->nghost_tot;
  s->ghost_atom = (int *)zalloc(sizeof(int) * g->nghost_tot);
  s->ghost_ncomp = (int *)zalloc(sizeof(int) * g->nghost_tot);
  s->ghost_comp = (int *)zalloc(sizeof(int) * 4 * g->nghost_tot);
  s->ghost_coef = (double *)zalloc(sizeof(double) * 4 * g->nghost_tot);
  for (int k = 0; k < g->nghost_tot; k++) {
    const int nc = g->natm_comp[k + 1];
    if (nc > 4) return PMX_ERR_ARG;
    s->ghost_atom[k] = g->ighost_map[k + 1] - 1;
    s->ghost_ncomp[k] = nc;
    for (int m = 0; m < nc; m++) {
      s->ghost_comp[4 * k + m] = g->iatm_comp[m + 1][k + 1] - 1;
      s->ghost_coef[4 * k + m] = g->coef[m + 1][k + 1];
    }
  }
  y->ghost_atom = s->ghost_atom;
  y->ghost_ncomp = s->ghost_ncomp;
  y->ghost_comp = s->ghost_comp;
  y->ghost_coef = s->ghost_coef;
  y->atm_type = s->atm_type;
  y->q = s->q;
  y->mass = s->mass;
  /* settings */
  pmx_params *r = &out->par;
  for (int k = 0; k < 9; k++) r->hmat[k] = gd->cell.hmat[k + 1];
  r->pi_beads = P;
  r->temperature = gd->statepoint.t_ext;
  r->pimd_typ = gd->simopts.pi_md_typ == 1 ? PMX_PIMD_STAGING : PMX_PIMD_CENTROID;
  r->nsplin = class->interact.nsplin;
  r->ishift = class->interact.iswit_vdw ? PMX_SHIFT_SWIT : pp->ishift;
  r->rheal = class->interact.rheal_res;
  r->skin = class->interact.skin;
  r->ewald_alpha = gd->ewald.alp_ewd * cbrt(gd->cell.vol);  /* the keyword is alpha V^(1/3) */
  r->kmax_ewd = pp->kmax_ewd;
  r->pme_on = class->part_mesh.pme_on;
  r->kmax_pme = class->part_mesh.kmax_pme;
  r->n_interp = class->part_mesh.n_interp;
  r->intra_perds = gd->cell.intra_perds;
  r->precision = PMX_PREC_FP64;
  r->pair_eval = PMX_PAIR_TABLE;
  const size_t NP = (size_t)N * P;
  double **bufs[9] = {&s->px, &s->py, &s->pz, &s->fx, &s->fy, &s->fz, &s->xm, &s->ym, &s->zm};
  for (int k = 0; k < 9; k++) *bufs[k] = (double *)zalloc(sizeof(double) * NP);
  return PMX_OK;
}

void pmx_piny_free(pmx_piny_sys *o) {
  store_t *s = (store_t *)o->store;
  if (!s) return;
  void *p[] = {s->atm_type, s->bond_j1, s->bond_j2, s->bond_type, s->bend_j1, s->bend_j2, s->bend_j3, s->bend_type,
               s->ghost_atom, s->ghost_ncomp, s->ghost_comp, s->inter_label, s->q, s->mass, s->sig, s->eps,
               s->cutti, s->cutoff, s->cutoff_res, s->bond_c, s->bond_eq, s->bend_c, s->bend_s, s->ghost_coef,
               s->px, s->py, s->pz, s->fx, s->fy, s->fz, s->xm, s->ym, s->zm};
  for (size_t k = 0; k < sizeof p / sizeof p[0]; k++) free(p[k]);
  free(s);
  o->store = NULL;
}

int pmx_piny_energy_control(pmx_ctx *ctx, pmx_piny_sys *o, CLASS *class, GENERAL_DATA *gd, unsigned flags,
                            pmx_energies *e) {
  store_t *s = (store_t *)o->store;
  const int N = class->clatoms_info.natm_tot, P = class->clatoms_info.pi_beads;
  for (int b = 0; b < P; b++) {
    const CLATOMS_POS *c = &class->clatoms_pos[b + 1];
    for (int i = 0; i < N; i++) {
      s->px[(size_t)b * N + i] = c->x[i + 1];
      s->py[(size_t)b * N + i] = c->y[i + 1];
      s->pz[(size_t)b * N + i] = c->z[i + 1];
    }
  }
  int rc = pmx_set_positions(ctx, s->px, s->py, s->pz);
  if (rc) return rc;
  if (P > 1) {  /* mode-1 coordinates (control_pimd_trans.c:175-181) */
    for (int i = 0; i < N; i++) {
      s->xm[i] = class->clatoms_info.xmod[i + 1];
      s->ym[i] = class->clatoms_info.ymod[i + 1];
      s->zm[i] = class->clatoms_info.zmod[i + 1];
    }
    if ((rc = pmx_set_xmod(ctx, s->xm, s->ym, s->zm))) return rc;
  }
  if ((rc = pmx_energy_control(ctx, flags, e))) return rc;
  if ((rc = pmx_get_forces(ctx, s->fx, s->fy, s->fz))) return rc;
  for (int b = 0; b < P; b++) {
    CLATOMS_POS *c = &class->clatoms_pos[b + 1];
    for (int i = 0; i < N; i++) {
      c->fx[i + 1] = s->fx[(size_t)b * N + i];
      c->fy[i + 1] = s->fy[(size_t)b * N + i];
      c->fz[i + 1] = s->fz[(size_t)b * N + i];
    }
  }
  if (P > 1 && pmx_get_fxt(ctx, s->fx, s->fy, s->fz) == PMX_OK)
    for (int b = 0; b < P; b++) {
      CLATOMS_POS *c = &class->clatoms_pos[b + 1];
      for (int i = 0; i < N; i++) {
        c->fxt[i + 1] = s->fx[(size_t)b * N + i];
        c->fyt[i + 1] = s->fy[(size_t)b * N + i];
        c->fzt[i + 1] = s->fz[(size_t)b * N + i];
      }
    }
  /* the stat_avg fields as energy_control_inter_real.c:186-197,
   * _inter_recip.c:339-347 and _intra.c:284-300 fill them */
  STAT_AVG *a = &gd->stat_avg;
  const double vrecip = e->vrecip + e->vself + e->vbgr + e->vecor;
  a->vintert = e->vintert;
  a->vintrat = e->vintrat;
  a->vrecip = vrecip;
  a->vlong = e->vlong;
  a->vvdw = e->vvdw + e->vlong;
  a->vcoul = e->vcoul + vrecip;
  a->vbondt = e->vbond;
  a->vbendt = e->vbend;
  if (P > 1) {
    a->pi_ke_vir = e->pi_ke_vir;
    if (flags & PMX_EVAL_MODE_FORCES) {
      a->kin_harm = e->kin_harm;
      a->pi_ke_prim = e->pi_ke_prim;
    }
  }
  for (int k = 0; k < 9; k++) {
    gd->ptens.pvten[k + 1] = e->pvten[k];
    gd->ptens.pvten_tot[k + 1] = e->pvten_tot[k];
  }
  return PMX_OK;
}
