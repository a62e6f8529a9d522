/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * FP64, single-threaded C restatement of PINY_MD's force evaluation for
 * path-integral water.  Each function names the reference file:lines it
 * follows (paths relative to the PINY source tree).  Operation order is
 * kept where it decides bits: spline tables, k-vector and grid sizing,
 * minimum image, the Verlet r^2 test.  Summation order of the large sums is
 * not kept (the reference unrolls by 5); that changes results in the last
 * bits only.  The FFT of the reference (vendored FFTPACK) is replaced by a
 * direct DFT with the same sign conventions; the 128-point Gauss-Legendre
 * table (weights_nodes_128.h) is recomputed by Newton iteration.
 *
 * Built with -O2 -ffp-contract=off so that no fused multiply-add changes
 * the bits of the r^2 test or of the tables.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_BOHR 0.529177      /* defines.h:78 */
#define OR_BOLTZ 315777.0     /* defines.h:80 */
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* NINT with SIMP_NINT (defines.h:96-97; standard_include.h defines it) */
static int nint_piny(double x) { return (int)((x) >= 0.0 ? ((x) + 0.5) : ((x)-0.5)); }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* erfc polynomial of spline_fit.c:327-339 and set_exclude.c:566-570 */
static const double ERF_P = 0.3614;
static const double ERF_E[10] = {0.0,
                                 0.2041422096422003,
                                 0.1997535956961481,
                                 0.2213176596405576,
                                 0.03360430734640255,
                                 0.4732592578721755,
                                 -0.509078520069735,
                                 0.6772631491947646,
                                 -0.369912979092217,
                                 0.06965131976970335};

struct or_ctx {
  int N, P, ntype, npair;
  pmx_params par;
  /* copies of the system description */
  int *type;
  double *q, *mass;
  int *ilab;
  double *sig, *eps, *cutti, *cutoff, *cutoff_res;
  int nbond, nbond_type, *b1, *b2, *btyp;
  double *bc, *bdc, *beq;
  int nbend, nbend_type, *a1, *a2, *a3, *atyp;
  double *ac, *as, *adc, *ads;
  int nghost, *gatom, *gncomp, *gcomp;
  double *gcoef;
  /* cell (1-based like PINY) */
  double hmat[10], hmati[10], vol;
  int cubic;
  /* spline tables: [npair][nsplin] */
  int nsplin;
  double *cv0, *cdv0, *cv0_c, *cdv0_c;
  double *dr_spl, *dri_spl, *rmin_spl;
  double cutoff_max;
  double *cutskin; /* squared, per pair */
  double clong, clong_res;
  int iswit;
  /* exclusions, 1-based atoms; excl_j sorted per atom */
  int *ex_num, *ex_off, *ex_j, nexcl;
  /* ecor */
  int necor, *e1, *e2;
  double *ecv0, *ecdv0, e_rmin, e_rmax, e_dr, e_dri;
  /* ewald */
  double alp_ewd, ecut, self_erf;
  int kmaxv[4], nktot, *kka, *kkb, *kkc;
  double *bweight;
  int ngrid[3];
  int nchrg, *ichrg;
  /* verlet list (1-based atoms) */
  int *nter, *jver_off, *jter, nver, nver_cap;
  double *x0, *y0, *z0;
  int have_list;
  double spread;
  /* pimd */
  double *prekf, *veig;
  double *bead_e; /* [P*4] */
  /* RESPA settings (zeros: off) */
  pmx_respa respa;
};

/* ------------------------------------------------------------------------
 * cell: gethinv, mathlib.c:353-384 (iperd = 3); cubic flag read_hmat.c:498
 */
static void or_setup_cell(or_ctx *c) {
  double *h = c->hmat, *hi = c->hmati, vol;
  for (int i = 1; i <= 9; i++) h[i] = c->par.hmat[i - 1];
  vol = (h[1] * (h[5] * h[9] - h[8] * h[6]) + h[4] * (h[8] * h[3] - h[2] * h[9]) +
         h[7] * (h[2] * h[6] - h[5] * h[3]));
  c->vol = vol;
  hi[1] = (h[5] * h[9] - h[8] * h[6]) / vol;
  hi[5] = (h[1] * h[9] - h[7] * h[3]) / vol;
  hi[9] = (h[1] * h[5] - h[4] * h[2]) / vol;
  hi[4] = (h[7] * h[6] - h[4] * h[9]) / vol;
  hi[2] = (h[3] * h[8] - h[2] * h[9]) / vol;
  hi[7] = (h[4] * h[8] - h[7] * h[5]) / vol;
  hi[3] = (h[2] * h[6] - h[3] * h[5]) / vol;
  hi[8] = (h[7] * h[2] - h[8] * h[1]) / vol;
  hi[6] = (h[3] * h[4] - h[6] * h[1]) / vol;
  c->cubic = (h[4] == 0.0 && h[7] == 0.0 && h[2] == 0.0 && h[8] == 0.0 &&
              h[3] == 0.0 && h[6] == 0.0);
}

/* minimum image of one separation: period.c:27-111 (iperd = 3) */
static void or_period1(const or_ctx *c, double *dx, double *dy, double *dz) {
  const double *h = c->hmat, *hi = c->hmati;
  double sx, sy, sz, asx, asy, asz;
  if (c->cubic) {
    asx = (*dx) * hi[1];
    asy = (*dy) * hi[5];
    asz = (*dz) * hi[9];
    sx = asx - nint_piny(asx);
    sy = asy - nint_piny(asy);
    sz = asz - nint_piny(asz);
    *dx = sx * h[1];
    *dy = sy * h[5];
    *dz = sz * h[9];
  } else {
    double dxt = *dx, dyt = *dy, dzt = *dz;
    asx = dxt * hi[1] + dyt * hi[4] + dzt * hi[7];
    asy = dxt * hi[2] + dyt * hi[5] + dzt * hi[8];
    asz = dxt * hi[3] + dyt * hi[6] + dzt * hi[9];
    sx = asx - nint_piny(asx);
    sy = asy - nint_piny(asy);
    sz = asz - nint_piny(asz);
    *dx = sx * h[1] + sy * h[4] + sz * h[7];
    *dy = sx * h[2] + sy * h[5] + sz * h[8];
    *dz = sx * h[3] + sy * h[6] + sz * h[9];
  }
}

/* bead separation imaged with the xmod separation: period.c:121-206 */
static void or_period_pimd1(const or_ctx *c, double *dx, double *dy, double *dz,
                            double dxs, double dys, double dzs) {
  const double *h = c->hmat, *hi = c->hmati;
  double sx, sy, sz, asx, asy, asz;
  if (c->cubic) {
    asx = dxs * hi[1];
    asy = dys * hi[5];
    asz = dzs * hi[9];
    sx = -nint_piny(asx);
    sy = -nint_piny(asy);
    sz = -nint_piny(asz);
    *dx += (sx * h[1]);
    *dy += (sy * h[5]);
    *dz += (sz * h[9]);
  } else {
    asx = dxs * hi[1] + dys * hi[4] + dzs * hi[7];
    asy = dxs * hi[2] + dys * hi[5] + dzs * hi[8];
    asz = dxs * hi[3] + dys * hi[6] + dzs * hi[9];
    sx = -nint_piny(asx);
    sy = -nint_piny(asy);
    sz = -nint_piny(asz);
    *dx += (sx * h[1] + sy * h[4] + sz * h[7]);
    *dy += (sx * h[2] + sy * h[5] + sz * h[8]);
    *dz += (sx * h[3] + sy * h[6] + sz * h[9]);
  }
}

/* 0-based type pair index of PINY's inter list (force_npol.c:577-580) */
static int or_pair(const or_ctx *c, int ta, int tb) {
  int a = ta < tb ? ta : tb, b = ta + tb - a;
  return a * c->ntype - (a * (a - 1)) / 2 + (b - a);
}

/* ------------------------------------------------------------------------
 * spline tables: spline_vdv (spline_fit.c:28-113), vlj_bin (:149-177),
 * vnull_bin (:120-140), vcoul_bin real-space Ewald branch (:314-398),
 * driven as set_inter_splin (control_inter_params.c:643-772).
 * Table index i (1..nsplin) is stored at [i-1].
 */
static void or_spline_one(const or_ctx *c, int itype, double rmin, double rmax,
                          double *v, double *dv, double *dr_out, double *dri_out,
                          double sig, double eps, double qij) {
  int n = c->nsplin, ishift = c->par.ishift;
  double rheal = c->par.rheal, dr, dri;
  double *r = malloc(sizeof(double) * (n + 1));
  double *sw = malloc(sizeof(double) * (n + 1));
  double *dsw = malloc(sizeof(double) * (n + 1));
  dr = (rmax - rmin) / (double)(n - 5);
  dri = 1.0 / dr;
  *dr_out = dr;
  *dri_out = dri;
  for (int i = 1; i <= n; i++) {
    r[i] = dr * ((double)(i - 3)) + rmin;
    sw[i] = 1.0;
    dsw[i] = 0.0;
  }
  if (ishift == 2) {
    for (int i = 1; i <= n; i++) {
      double rp = (r[i] - rmax + rheal) / rheal;
      rp = dmin(rp, 1.0);
      rp = dmax(rp, 0.0);
      sw[i] = 1.0 + rp * rp * (2.0 * rp - 3.0);
      dsw[i] = -6.0 * (rp / rheal) * (rp - 1.0) / r[i];
    }
  }
  if (itype == 2) { /* Lennard-Jones */
    double sig6 = sig * sig * sig * sig * sig * sig;
    for (int i = 1; i <= n; i++) {
      double r2 = r[i] * r[i];
      double sig6r = sig6 / (r2 * r2 * r2);
      double vnow = (eps * 4. * sig6r * (sig6r - 1.));
      double dvnow = (eps * 24. * sig6r * (sig6r * 2. - 1.) / r2);
      v[i - 1] = vnow * sw[i];
      dv[i - 1] = dvnow * sw[i] + vnow * dsw[i];
    }
  } else if (itype == 4) { /* null */
    for (int i = 1; i <= n; i++) {
      v[i - 1] = 0.0;
      dv[i - 1] = 0.0;
    }
  } else if (itype == 1) { /* erfc Coulomb, iperd > 0 */
    const double p = ERF_P;
    const double *e = ERF_E;
    double de[10];
    for (int k = 1; k <= 9; k++) de[k] = ((double)k) * e[k];
    double alp = c->alp_ewd;
    double talp2 = 2.0 * alp * alp, palp = p * alp;
    for (int i = 1; i <= n; i++) {
      double r2 = r[i] * r[i];
      double ralp = r[i] * alp;
      double eee = exp(-ralp * ralp);
      double tt = 1.0 / (1.0 + p * ralp);
      double gerfc = ((((((((e[9] * tt + e[8]) * tt + e[7]) * tt + e[6]) * tt + e[5]) * tt + e[4]) * tt +
                        e[3]) * tt + e[2]) * tt + e[1]) * tt * eee;
      double dgerfc = ((((((((de[9] * tt + de[8]) * tt + de[7]) * tt + de[6]) * tt + de[5]) * tt + de[4]) * tt +
                         de[3]) * tt + de[2]) * tt + de[1]) * tt * tt * eee * palp +
                      talp2 * gerfc * r[i];
      double vnow = qij * gerfc / r[i];
      double dvnow = (gerfc / r2 + dgerfc / r[i]) * qij / r[i];
      v[i - 1] = vnow * sw[i];
      dv[i - 1] = dvnow * sw[i] + vnow * dsw[i];
    }
  }
  if (ishift == 1) {
    double shift = v[(n - 2) - 1];
    for (int i = 1; i <= n; i++) v[i - 1] -= shift;
  }
  free(r);
  free(sw);
  free(dsw);
}

static void or_setup_splines(or_ctx *c) {
  int n = c->nsplin, np = c->npair;
  c->cv0 = calloc((size_t)n * np, sizeof(double));
  c->cdv0 = calloc((size_t)n * np, sizeof(double));
  c->cv0_c = calloc((size_t)n * np, sizeof(double));
  c->cdv0_c = calloc((size_t)n * np, sizeof(double));
  c->dr_spl = calloc(np, sizeof(double));
  c->dri_spl = calloc(np, sizeof(double));
  c->rmin_spl = calloc(np, sizeof(double));
  c->cutoff_max = 0.0;
  for (int t = 0; t < np; t++) {
    c->rmin_spl[t] = c->cutti[t];
    c->cutoff_max = dmax(c->cutoff[t], c->cutoff_max);
    or_spline_one(c, c->ilab[t], c->cutti[t], c->cutoff[t], c->cv0 + (size_t)t * n,
                  c->cdv0 + (size_t)t * n, &c->dr_spl[t], &c->dri_spl[t], c->sig[t],
                  c->eps[t], 0.0);
  }
  if (c->nchrg > 0) {
    for (int t = 0; t < np; t++) {
      double drt, drit;
      or_spline_one(c, 1, c->cutti[t], c->cutoff[t], c->cv0_c + (size_t)t * n,
                    c->cdv0_c + (size_t)t * n, &drt, &drit, 1.0, 0.0, 1.0);
    }
  }
}

/* get_clong.c:27-114 with the type-pair counts of :45-60 */
static void or_setup_clong(or_ctx *c) {
  int nt = c->ntype;
  long *num = calloc(nt, sizeof(long));
  double *clongm = calloc(c->npair, sizeof(double));
  for (int i = 0; i < c->N; i++) num[c->type[i]]++;
  for (int i = 0; i < nt; i++) {
    for (int j = 0; j < i; j++) clongm[or_pair(c, j, i)] = (double)(num[i] * num[j]);
    clongm[or_pair(c, i, i)] = (double)((num[i] * (num[i] - 1)) / 2);
  }
  double fpi3 = 4.0 * M_PI / 3.0, fpi = 4.0 * M_PI, rheal = c->par.rheal;
  c->clong = 0.0;
  c->clong_res = 0.0;
  for (int t = 0; t < c->npair; t++) {
    double c6m = 0.0;
    if (c->ilab[t] == 2) c6m = (4.0 * c->eps[t]) * (c->sig[t] * c->sig[t] * c->sig[t] * c->sig[t] * c->sig[t] * c->sig[t]);
    double rc = c->cutoff[t], rr = c->cutoff_res[t];
    double rmin = (rc < rr ? rc : rr);
    if (c->iswit == 0) {
      c->clong -= fpi3 * c6m * clongm[t] / pow(rc, 3.0);
      c->clong_res -= fpi * c6m * clongm[t] *
                      ((1.0 / (rheal * rheal)) * (1.0 / rmin + 1.0 / (rmin - rheal)) +
                       (2.0 / (rheal * rheal * rheal)) * log((rmin - rheal) / rmin));
    } else if ((c6m > 0) && (rc - rheal > 0)) {
      double fact = (1.0 / (rheal * rheal)) * (1.0 / rc + 1.0 / (rc - rheal)) +
                    (2.0 / (rheal * rheal * rheal)) * log((rc - rheal) / rc);
      c->clong -= (fpi * c6m * clongm[t] * fact);
      double fact_res = (1.0 / (rheal * rheal)) * (1.0 / rmin + 1.0 / (rmin - rheal)) +
                        (2.0 / (rheal * rheal * rheal)) * log((rmin - rheal) / rmin);
      c->clong_res -= fpi * c6m * clongm[t] * fact_res;
    }
  }
  free(num);
  free(clongm);
}

/* ------------------------------------------------------------------------
 * Ewald / PME setup: alpha conversion (control_set_cp_ewald.c:157-163),
 * calc_cutoff (set_cp_ewald.c:125-196, cp off), setkvec3d (:428-573),
 * set_pme_grid (:204-293), radixme (:815-1008), get_bspline_wght1d
 * (:1288-1370), set_pme_wght (:1136-1275).
 */
static void or_radixme(int *kmax1, int *kmax2, int *kmax3) {
  static const int krad[50] = {4,   8,   12,  16,  20,  24,  32,  36,  40,  48,  60,  64,  72,
                               80,  96,  100, 108, 120, 128, 144, 160, 180, 192, 200, 216, 240,
                               256, 288, 300, 320, 324, 360, 384, 400, 432, 480, 500, 512, 540,
                               576, 600, 640, 648, 720, 768, 800, 864, 900, 960, 972};
  int k[3] = {(*kmax1 + 1) << 2, (*kmax2 + 1) << 2, (*kmax3 + 1) << 2};
  for (int i = 1; i <= 50; i++) {
    for (int d = 0; d < 3; d++) {
      if (krad[i - 1] > k[d]) {
        int lo = imax(i - 1, 1);
        int hi_v = krad[i - 1], lo_v = krad[lo - 1];
        if (k[d] > lo_v && k[d] < hi_v) {
          int dhi = abs(k[d] - krad[i - 1]);
          int dlo = abs(k[d] - krad[lo - 1]);
          k[d] = (dhi < dlo) ? krad[i - 1] : krad[lo - 1];
        }
      }
    }
  }
  *kmax1 = k[0] / 4 - 1;
  *kmax2 = k[1] / 4 - 1;
  *kmax3 = k[2] / 4 - 1;
}

static double or_row_norm(const double *hi, int a, int b, int cc) {
  double d1 = hi[a], d2 = hi[b], d3 = hi[cc];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

static void or_calc_cutoff(or_ctx *c) {
  double rtwoth = -(2. / 3.);
  double rvol23 = pow(c->vol, rtwoth);
  int kmax = c->par.kmax_ewd;
  c->ecut = M_PI * .5 * M_PI * (double)(kmax * kmax) * rvol23;
  double try1 = or_row_norm(c->hmati, 1, 4, 7);
  double try2 = or_row_norm(c->hmati, 2, 5, 8);
  double try3 = or_row_norm(c->hmati, 3, 6, 9);
  double t1 = sqrt(c->ecut * .5) / (M_PI * try1);
  double t2 = sqrt(c->ecut * .5) / (M_PI * try2);
  double t3 = sqrt(c->ecut * .5) / (M_PI * try3);
  c->kmaxv[1] = (int)(2.0 * t1);
  c->kmaxv[2] = (int)(2.0 * t2);
  c->kmaxv[3] = (int)(2.0 * t3);
}

/* k-vector try value of setkvec3d */
static double or_ktry(const double *hm, double aka, double akb, double akc) {
  double tpi = 2.0 * M_PI;
  double xk = (aka * hm[1] + akb * hm[2] + akc * hm[3]) * tpi;
  double yk = (aka * hm[4] + akb * hm[5] + akc * hm[6]) * tpi;
  double zk = (aka * hm[7] + akb * hm[8] + akc * hm[9]) * tpi;
  return (xk * xk + yk * yk + zk * zk) * .5;
}

static int or_setkvec3d(or_ctx *c, int fill) {
  const double *hm = c->hmati;
  double ec4 = c->ecut * 4.;
  int icount = 0, i;
  for (int ka = 0; ka <= c->kmaxv[1]; ++ka) {
    double aka = (double)ka;
    int kbmin = (ka == 0) ? 0 : -c->kmaxv[2];
    for (i = kbmin; i <= 0; ++i) {
      if (or_ktry(hm, aka, (double)i, 0.0) <= ec4) break;
    }
    kbmin = i;
    for (i = 1; i <= c->kmaxv[2]; ++i) {
      if (or_ktry(hm, aka, (double)i, 0.0) > ec4) break;
    }
    int kbmax = i - 1;
    for (int kb = kbmin; kb <= kbmax; ++kb) {
      double akb = (double)kb;
      int kcmin = (ka == 0 && kb == 0) ? 1 : -c->kmaxv[3];
      for (i = kcmin; i <= 0; ++i) {
        if (or_ktry(hm, aka, akb, (double)i) <= ec4) break;
      }
      kcmin = i;
      for (i = 1; i <= c->kmaxv[3]; ++i) {
        if (or_ktry(hm, aka, akb, (double)i) > ec4) break;
      }
      int kcmax = i - 1;
      for (int kc = kcmin; kc <= kcmax; ++kc) {
        if (fill) {
          c->kka[icount] = ka;
          c->kkb[icount] = kb;
          c->kkc[icount] = kc;
        }
        ++icount;
      }
    }
  }
  return icount;
}

static void or_set_pme_grid(or_ctx *c) {
  double rtwoth = -(2. / 3.);
  double rvol23 = pow(c->vol, rtwoth);
  int kp = c->par.kmax_pme;
  double ecut_pme = M_PI * .5 * M_PI * (double)(kp * kp) * rvol23;
  double ecut = dmax(ecut_pme, c->ecut);
  double try1 = or_row_norm(c->hmati, 1, 4, 7);
  double try2 = or_row_norm(c->hmati, 2, 5, 8);
  double try3 = or_row_norm(c->hmati, 3, 6, 9);
  double t1 = sqrt(ecut * .5) / (M_PI * try1);
  double t2 = sqrt(ecut * .5) / (M_PI * try2);
  double t3 = sqrt(ecut * .5) / (M_PI * try3);
  int k1 = nint_piny(t1), k2 = nint_piny(t2), k3 = nint_piny(t3);
  or_radixme(&k1, &k2, &k3);
  int *ng = c->ngrid;
  ng[0] = 4 * (k1 + 1);
  ng[1] = 4 * (k2 + 1);
  ng[2] = 4 * (k3 + 1);
  int *kv = c->kmaxv;
  while (ng[0] < 2 * (kv[1] + 2) || ng[1] < 2 * (kv[2] + 2) || ng[2] < 2 * (kv[3] + 2)) {
    if (ng[0] < 2 * (kv[1] + 2)) t1 += 1.0;
    if (ng[1] < 2 * (kv[2] + 2)) t2 += 1.0;
    if (ng[2] < 2 * (kv[3] + 2)) t3 += 1.0;
    k1 = nint_piny(t1);
    k2 = nint_piny(t2);
    k3 = nint_piny(t3);
    or_radixme(&k1, &k2, &k3);
    ng[0] = 4 * (k1 + 1);
    ng[1] = 4 * (k2 + 1);
    ng[2] = 4 * (k3 + 1);
  }
}

static void or_bspline_wght1d(int n_interp, int ngrid, double *bweight) {
  double aj[32], rn[32], rn1[32], mn_k[32], uk[32];
  double grid = (double)ngrid;
  for (int j = 1; j <= n_interp; j++) {
    aj[j] = (double)(j - 1);
    rn[j] = (double)(j);
    if (j > 1) rn1[j] = 1.0 / ((double)(j - 1));
  }
  rn1[1] = 0.0;
  mn_k[1] = 1.0;
  uk[1] = 1.0;
  for (int k = 2; k <= n_interp; k++) {
    uk[k] = (double)(k);
    mn_k[k] = 0.0;
  }
  for (int n = 3; n <= n_interp; n++) {
    for (int k = n; k >= 2; k--) {
      double t = (uk[k] * mn_k[k] + (rn[n] - uk[k]) * mn_k[k - 1]) * rn1[n];
      mn_k[k] = t;
    }
    mn_k[1] = uk[1] * mn_k[1] * rn1[n];
  }
  double *bre = malloc(sizeof(double) * (ngrid + 1));
  double *bim = malloc(sizeof(double) * (ngrid + 1));
  double tpi_n = 2.0 * M_PI / grid;
  for (int m = 1; m <= ngrid; m++) {
    bre[m] = 0.0;
    bim[m] = 0.0;
    for (int k = 1; k <= n_interp - 1; k++) {
      double arg = tpi_n * ((double)((k - 1) * (m - 1)));
      bre[m] += cos(arg) * mn_k[k];
      bim[m] += sin(arg) * mn_k[k];
    }
  }
  tpi_n = 2.0 * M_PI * ((double)(n_interp - 1)) / grid;
  for (int m = 1; m <= ngrid; m++) {
    double arg = tpi_n * ((double)(m - 1));
    double bnr = cos(arg), bni = sin(arg);
    double bdr = bre[m], bdi = bim[m];
    double denom = bdr * bdr + bdi * bdi;
    double tr = (bnr * bdr + bni * bdi) / denom;
    double ti = (bni * bdr - bnr * bdi) / denom;
    bweight[m] = tr * tr + ti * ti;
  }
  free(bre);
  free(bim);
  (void)aj;
}

static void or_setup_ewald(or_ctx *c) {
  double side = pow(c->vol, (1.0 / 3.0));
  c->alp_ewd = c->par.ewald_alpha / side;
  or_calc_cutoff(c);
  c->nktot = or_setkvec3d(c, 0);
  c->kka = malloc(sizeof(int) * (c->nktot + 1));
  c->kkb = malloc(sizeof(int) * (c->nktot + 1));
  c->kkc = malloc(sizeof(int) * (c->nktot + 1));
  or_setkvec3d(c, 1);
  c->bweight = calloc(c->nktot + 1, sizeof(double));
  if (c->par.pme_on) {
    or_set_pme_grid(c);
    double *bw[3];
    for (int d = 0; d < 3; d++) {
      bw[d] = calloc(c->ngrid[d] + 1, sizeof(double));
      or_bspline_wght1d(c->par.n_interp, c->ngrid[d], bw[d]);
    }
    for (int i = 0; i < c->nktot; i++) {
      int ka = c->kka[i], kb = c->kkb[i], kc = c->kkc[i];
      int kap = ka < 0 ? ka + c->ngrid[0] + 1 : ka + 1;
      int kbp = kb < 0 ? kb + c->ngrid[1] + 1 : kb + 1;
      int kcp = kc < 0 ? kc + c->ngrid[2] + 1 : kc + 1;
      c->bweight[i] = bw[0][kap] * bw[1][kbp] * bw[2][kcp];
    }
    for (int d = 0; d < 3; d++) free(bw[d]);
  }
}

/* ------------------------------------------------------------------------
 * exclusions (set_exclude.c:30-545): bonds j1-j2, bends j1-j3, ghost-
 * constituent pairs; sorted and unique per atom.  Ewald corrections for
 * excluded pairs whose charges are both nonzero (:425-453).
 */
static int cmp_pair(const void *a, const void *b) {
  const int *x = a, *y = b;
  if (x[0] != y[0]) return x[0] - y[0];
  return x[1] - y[1];
}

static void or_setup_exclusions(or_ctx *c) {
  int cap = c->nbond + c->nbend;
  for (int g = 0; g < c->nghost; g++) cap += c->gncomp[g];
  int *pr = malloc(sizeof(int) * 2 * (cap + 1));
  int n = 0;
  for (int b = 0; b < c->nbond; b++) {
    int i = c->b1[b] + 1, j = c->b2[b] + 1;
    pr[2 * n] = imax(i, j);
    pr[2 * n + 1] = imin(i, j);
    n++;
  }
  for (int b = 0; b < c->nbend; b++) {
    int i = c->a1[b] + 1, j = c->a3[b] + 1;
    pr[2 * n] = imax(i, j);
    pr[2 * n + 1] = imin(i, j);
    n++;
  }
  for (int g = 0; g < c->nghost; g++) {
    for (int k = 0; k < c->gncomp[g]; k++) {
      int i = c->gatom[g] + 1, j = c->gcomp[4 * g + k] + 1;
      pr[2 * n] = imax(i, j);
      pr[2 * n + 1] = imin(i, j);
      n++;
    }
  }
  qsort(pr, n, 2 * sizeof(int), cmp_pair);
  int m = 0;
  for (int k = 0; k < n; k++) {
    if (m > 0 && pr[2 * (m - 1)] == pr[2 * k] && pr[2 * (m - 1) + 1] == pr[2 * k + 1]) continue;
    pr[2 * m] = pr[2 * k];
    pr[2 * m + 1] = pr[2 * k + 1];
    m++;
  }
  c->nexcl = m;
  c->ex_num = calloc(c->N + 2, sizeof(int));
  c->ex_off = calloc(c->N + 2, sizeof(int));
  c->ex_j = malloc(sizeof(int) * (m + 1));
  for (int k = 0; k < m; k++) {
    c->ex_j[k + 1] = pr[2 * k + 1];
    c->ex_num[pr[2 * k]]++;
  }
  c->ex_off[1] = 0;
  for (int i = 2; i <= c->N; i++) c->ex_off[i] = c->ex_off[i - 1] + c->ex_num[i - 1];
  /* Ewald corrections */
  c->necor = 0;
  c->e1 = malloc(sizeof(int) * (m + 1));
  c->e2 = malloc(sizeof(int) * (m + 1));
  for (int k = 0; k < m; k++) {
    int i = pr[2 * k], j = pr[2 * k + 1];
    if (c->q[i - 1] != 0.0 && c->q[j - 1] != 0.0) {
      c->e1[c->necor] = i;
      c->e2[c->necor] = j;
      c->necor++;
    }
  }
  free(pr);
}

/* Gauss-Legendre on [-1,1] by Newton iteration on P_n */
void or_gauss_legendre(int n, double *node, double *weight) {
  /* Newton's iteration on P_n in quadruple precision (__float128), each
   * node and weight rounded to double once: the correctly rounded values,
   * which are also what 33-digit decimal tables of them parse to (the
   * reference's weights_nodes_128.h) */
  for (int i = 0; i < n; i++) {
    __float128 x = cos(M_PI * (i + 0.75) / (n + 0.5)), p0 = 1, p1 = x, dp = 0;
    for (int it = 0; it < 8; it++) {  /* quadratic convergence from a double start */
      p0 = 1;
      p1 = x;
      for (int k = 2; k <= n; k++) {
        __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1);
      x -= p1 / dp;
    }
    p0 = 1;
    p1 = x;
    for (int k = 2; k <= n; k++) {
      __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1);
    node[n - 1 - i] = (double)x;
    weight[n - 1 - i] = (double)(2 / ((1 - x * x) * dp * dp));
  }
}

/* splin_ecor (set_exclude.c:548-779), RESPA k-vectors off */
static void or_setup_ecor(or_ctx *c, const double *x, const double *y, const double *z) {
  int n = c->nsplin, N = c->N;
  double ten = 10.0 / OR_BOHR;
  double rmax = 0.0;
  if (x != NULL) {
    for (int ip = 0; ip < c->P; ip++) {
      for (int k = 0; k < c->necor; k++) {
        size_t i1 = (size_t)ip * N + c->e1[k] - 1, i2 = (size_t)ip * N + c->e2[k] - 1;
        double dx = x[i1] - x[i2], dy = y[i1] - y[i2], dz = z[i1] - z[i2];
        double r2 = dx * dx + dy * dy + dz * dz;
        rmax = dmax(r2, rmax);
      }
    }
  }
  rmax += 2.0 / OR_BOHR;
  rmax = dmax(ten, rmax);
  double rmin = 0.1;
  double dr = (rmax - rmin) / (double)(n - 5), dri = 1.0 / dr;
  c->e_rmin = rmin;
  c->e_rmax = rmax;
  c->e_dr = dr;
  c->e_dri = dri;
  c->ecv0 = calloc(n, sizeof(double));
  c->ecdv0 = calloc(n, sizeof(double));
  const double p = ERF_P, *e = ERF_E;
  double de[10];
  for (int k = 1; k <= 9; k++) de[k] = ((double)k) * e[k];
  double alp = c->alp_ewd, alp2 = alp * alp, talp2 = 2.0 * alp * alp, palp = p * alp;
  for (int i = 1; i <= n; i++) {
    double r12 = dr * (double)(i - 3) + rmin;
    double r12i = 1.0 / r12;
    double ralp = r12 * alp;
    double eee = exp(-ralp * ralp);
    double tt = 1.0 / (1.0 + p * ralp);
    double temp = ((((((((e[9] * tt + e[8]) * tt + e[7]) * tt + e[6]) * tt + e[5]) * tt + e[4]) * tt + e[3]) *
                        tt + e[2]) * tt + e[1]) * tt * eee;
    double gerf = 1.0 - temp;
    double dgerf = ((((((((de[9] * tt + de[8]) * tt + de[7]) * tt + de[6]) * tt + de[5]) * tt + de[4]) * tt +
                       de[3]) * tt + de[2]) * tt + de[1]) * tt * tt * eee * palp +
                   talp2 * temp * r12;
    c->ecv0[i - 1] = -gerf * r12i;
    double dvecor = -(dgerf * r12i - gerf / (r12 * r12));
    c->ecdv0[i - 1] = -dvecor * r12i;
  }
  /* self term correction and finite-k correction */
  double ecut4 = 4.0 * c->ecut;
  double kcut = sqrt(2.0 * ecut4);
  double arg = kcut / (2.0 * alp);
  double eee = exp(-arg * arg);
  double tt = 1.0 / (1.0 + p * arg);
  double self_erfc = ((((((((e[9] * tt + e[8]) * tt + e[7]) * tt + e[6]) * tt + e[5]) * tt + e[4]) * tt + e[3]) *
                           tt + e[2]) * tt + e[1]) * tt * eee;
  c->self_erf = 1.0 - self_erfc;
  double kmax = 15.0 * alp;
  if (kmax > kcut) {
    int nk = 128;
    double anode[128], weight[128], know[128], fknow[128];
    or_gauss_legendre(nk, anode, weight);
    double dk = kmax - kcut, pk = kmax + kcut;
    for (int k = 0; k < nk; k++) {
      know[k] = 0.5 * (anode[k] * dk + pk);
      fknow[k] = (dk / M_PI) * exp(-0.25 * know[k] * know[k] / alp2) * weight[k];
    }
    for (int i = 1; i <= n; i++) {
      double r12 = dr * (double)(i - 3) + rmin;
      double chi = 0.0, dchi = 0.0;
      for (int k = 0; k < nk; k++) {
        double a = know[k] * r12;
        double j0 = (sin(a) / a);
        double j1 = (sin(a) / a - cos(a)) / a;
        chi += (fknow[k] * j0);
        dchi += (know[k] * fknow[k] * j1 / r12);
      }
      c->ecv0[i - 1] += chi;
      c->ecdv0[i - 1] += dchi;
    }
  }
}

/* path_integral_init (transform_cnt.c:140-201): spring prefactors and
 * centroid eigenvalues (pi_md_typ = 2) or staging eigenvalues (= 1). */
static void or_setup_pimd(or_ctx *c) {
  int P = c->P;
  c->prekf = calloc(c->N, sizeof(double));
  c->veig = calloc(P + 1, sizeof(double));
  double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
  double dP = (double)P;
  for (int i = 0; i < c->N; i++) {
    double beta = OR_BOLTZ / T;
    double tau = OR_BOLTZ / (T * dP);
    c->prekf[i] = c->mass[i] / (2.0 * tau * beta);
  }
  for (int g = 0; g < c->nghost; g++) c->prekf[c->gatom[g]] = 0.0;
  double *v = c->veig;
  if (c->par.pimd_typ == PMX_PIMD_STAGING) {
    v[1] = 1.0;
    for (int ip = 2; ip <= P; ip++) v[ip] = ((double)(ip)) / ((double)(ip - 1));
    v[1] = 0.0;
  } else {
    double tpip = 2.0 * M_PI / dP, pre = 4.0 * dP;
    v[1] = 1.0;
    if (P > 1) v[P] = pre;
    for (int ip = 2; ip <= P / 2; ip++) {
      double arg = tpip * (double)(ip - 1);
      double temp = (1.0 - cos(arg)) * pre;
      v[2 * ip - 2] = temp;
      v[2 * ip - 1] = temp;
    }
    v[1] = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
static void *dup_arr(const void *src, size_t nbytes) {
  if (nbytes == 0) return calloc(1, 1);
  void *p = malloc(nbytes);
  memcpy(p, src, nbytes);
  return p;
}

or_ctx *or_create(const pmx_system *s, const pmx_params *par, const double *ix,
                  const double *iy, const double *iz, char *err, size_t errlen) {
  if (!s || !par || s->natm <= 1 || s->ntype <= 0 || par->pi_beads <= 0 || par->nsplin < 10) {
    if (err) snprintf(err, errlen, "or_create: invalid arguments");
    return NULL;
  }
  or_ctx *c = calloc(1, sizeof(or_ctx));
  c->N = s->natm;
  c->P = par->pi_beads;
  c->ntype = s->ntype;
  c->npair = s->ntype * (s->ntype + 1) / 2;
  c->par = *par;
  c->nsplin = par->nsplin;
  c->iswit = (par->ishift == 2);
  int N = c->N, np = c->npair;
  c->type = dup_arr(s->atm_type, sizeof(int) * N);
  c->q = dup_arr(s->q, sizeof(double) * N);
  c->mass = dup_arr(s->mass, sizeof(double) * N);
  c->ilab = dup_arr(s->inter_label, sizeof(int) * np);
  c->sig = dup_arr(s->inter_sig, sizeof(double) * np);
  c->eps = dup_arr(s->inter_eps, sizeof(double) * np);
  c->cutti = dup_arr(s->inter_cutti, sizeof(double) * np);
  c->cutoff = dup_arr(s->inter_cutoff, sizeof(double) * np);
  c->cutoff_res = dup_arr(s->inter_cutoff_res, sizeof(double) * np);
  c->nbond = s->nbond;
  c->nbond_type = s->nbond_type;
  c->b1 = dup_arr(s->bond_j1, sizeof(int) * s->nbond);
  c->b2 = dup_arr(s->bond_j2, sizeof(int) * s->nbond);
  c->btyp = dup_arr(s->bond_type, sizeof(int) * s->nbond);
  c->bc = dup_arr(s->bond_c, sizeof(double) * 7 * s->nbond_type);
  c->beq = dup_arr(s->bond_eq, sizeof(double) * s->nbond_type);
  c->bdc = calloc(7 * (s->nbond_type + 1), sizeof(double));
  for (int t = 0; t < s->nbond_type; t++) {  /* intra_coefs.c:181-188 */
    double *cc = c->bc + 7 * t, *d = c->bdc + 7 * t;
    d[0] = 0.0;
    d[1] = cc[1];
    d[2] = 2.0 * cc[2];
    d[3] = 3.0 * cc[3];
    d[4] = 4.0 * cc[4];
    d[5] = 5.0 * cc[5];
    d[6] = 6.0 * cc[6];
  }
  c->nbend = s->nbend;
  c->nbend_type = s->nbend_type;
  c->a1 = dup_arr(s->bend_j1, sizeof(int) * s->nbend);
  c->a2 = dup_arr(s->bend_j2, sizeof(int) * s->nbend);
  c->a3 = dup_arr(s->bend_j3, sizeof(int) * s->nbend);
  c->atyp = dup_arr(s->bend_type, sizeof(int) * s->nbend);
  c->ac = dup_arr(s->bend_c, sizeof(double) * 7 * s->nbend_type);
  c->as = dup_arr(s->bend_s, sizeof(double) * 7 * s->nbend_type);
  c->adc = calloc(7 * (s->nbend_type + 1), sizeof(double));
  c->ads = calloc(7 * (s->nbend_type + 1), sizeof(double));
  for (int t = 0; t < s->nbend_type; t++) {  /* intra_coefs.c:355-369 */
    double *cc = c->ac + 7 * t, *ss = c->as + 7 * t, *d = c->adc + 7 * t, *e = c->ads + 7 * t;
    d[0] = 0.0;
    d[1] = cc[1];
    d[2] = 2.0 * cc[2];
    d[3] = 3.0 * cc[3];
    d[4] = 4.0 * cc[4];
    d[5] = 5.0 * cc[5];
    d[6] = 6.0 * cc[6];
    e[0] = 0.0;
    e[1] = ss[1];
    e[2] = ss[2];
    e[3] = 2.0 * ss[3];
    e[4] = 3.0 * ss[4];
    e[5] = 4.0 * ss[5];
    e[6] = 5.0 * ss[6];
  }
  c->nghost = s->nghost;
  c->gatom = dup_arr(s->ghost_atom, sizeof(int) * s->nghost);
  c->gncomp = dup_arr(s->ghost_ncomp, sizeof(int) * s->nghost);
  c->gcomp = dup_arr(s->ghost_comp, sizeof(int) * 4 * s->nghost);
  c->gcoef = dup_arr(s->ghost_coef, sizeof(double) * 4 * s->nghost);
  /* charged list (close_intra_params.c:211-219) */
  c->ichrg = malloc(sizeof(int) * N);
  c->nchrg = 0;
  for (int i = 0; i < N; i++)
    if (c->q[i] != 0.0) c->ichrg[c->nchrg++] = i;

  or_setup_cell(c);
  or_setup_ewald(c);
  or_setup_splines(c);
  or_setup_clong(c);
  or_setup_exclusions(c);
  or_setup_ecor(c, ix, iy, iz);
  or_setup_pimd(c);
  c->cutskin = calloc(np, sizeof(double));
  c->nter = calloc(N + 2, sizeof(int));
  c->jver_off = calloc(N + 2, sizeof(int));
  c->x0 = calloc(N + 1, sizeof(double));
  c->y0 = calloc(N + 1, sizeof(double));
  c->z0 = calloc(N + 1, sizeof(double));
  c->bead_e = calloc(4 * c->P, sizeof(double));
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->type, c->q,    c->mass, c->ilab, c->sig,  c->eps,  c->cutti, c->cutoff,
                  c->cutoff_res, c->b1, c->b2, c->btyp, c->bc, c->bdc, c->beq, c->a1, c->a2,
                  c->a3, c->atyp, c->ac, c->as, c->adc, c->ads, c->gatom, c->gncomp, c->gcomp,
                  c->gcoef, c->cv0, c->cdv0, c->cv0_c, c->cdv0_c, c->dr_spl, c->dri_spl,
                  c->rmin_spl, c->cutskin, c->ex_num, c->ex_off, c->ex_j, c->e1, c->e2, c->ecv0,
                  c->ecdv0, c->kka, c->kkb, c->kkc, c->bweight, c->ichrg, c->nter, c->jver_off,
                  c->jter, c->x0, c->y0, c->z0, c->prekf, c->veig, c->bead_e};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  free(c);
}

/* ------------------------------------------------------------------------
 * ghost sites: get_ghost_pos (ghost_control.c:28-64),
 * distrib_ghost_force (:71-125)
 */
static void or_ghost_pos(const or_ctx *c, double *x, double *y, double *z) {
  for (int g = 0; g < c->nghost; g++) {
    int ig = c->gatom[g];
    x[ig] = 0.0;
    y[ig] = 0.0;
    z[ig] = 0.0;
    for (int k = 0; k < c->gncomp[g]; k++) {
      int ia = c->gcomp[4 * g + k];
      double cf = c->gcoef[4 * g + k];
      x[ig] += cf * x[ia];
      y[ig] += cf * y[ia];
      z[ig] += cf * z[ia];
    }
  }
}

static void or_ghost_force(const or_ctx *c, double *fx, double *fy, double *fz) {
  for (int g = 0; g < c->nghost; g++) {
    int ig = c->gatom[g];
    for (int k = 0; k < c->gncomp[g]; k++) {
      int ia = c->gcomp[4 * g + k];
      double cf = c->gcoef[4 * g + k];
      fx[ia] += cf * fx[ig];
      fy[ia] += cf * fy[ig];
      fz[ia] += cf * fz[ig];
    }
    fx[ig] = 0.0;
    fy[ig] = 0.0;
    fz[ig] = 0.0;
  }
}

/* ------------------------------------------------------------------------
 * neighbour list: get_pimd_spread (pimd_estimators.c:121-222),
 * get_cut_skin (nbr_list_control.c:249-283), check_ver_list (:182-239),
 * nolst_ver_gen + verlist_create count/fill (verlist_control.c:134-441,
 * verlist_create.c:223-503).  Positions are those of bead 1 (0-based
 * arrays); xmod enters through period_pimd when P > 1.
 */
static double or_spread(const or_ctx *c, const double *x, const double *y, const double *z) {
  double s = 0.0;
  int N = c->N;
  for (int ip = 1; ip < c->P; ip++) {
    for (int i = 0; i < N; i++) {
      double dx = x[(size_t)ip * N + i] - x[i];
      double dy = y[(size_t)ip * N + i] - y[i];
      double dz = z[(size_t)ip * N + i] - z[i];
      double r2 = dx * dx + dy * dy + dz * dz;
      s = dmax(s, r2);
    }
  }
  return sqrt(s);
}

static void or_cut_skin(or_ctx *c, double spread) {
  for (int t = 0; t < c->npair; t++) {
    double cs = c->cutoff[t] + c->par.skin + spread;
    c->cutskin[t] = cs * cs;
  }
}

static int or_check_list(or_ctx *c, const double *x, const double *y, const double *z,
                         double spread_now) {
  if (!c->have_list) return 1;
  double rmax = 0.0;
  for (int i = 0; i < c->N; i++) {
    double dx = x[i] - c->x0[i + 1], dy = y[i] - c->y0[i + 1], dz = z[i] - c->z0[i + 1];
    double r2 = dx * dx + dy * dy + dz * dz;
    rmax = dmax(rmax, r2);
  }
  rmax = sqrt(rmax);
  double dspread = 0.0;
  if (c->P > 1) dspread = dmax(0.0, spread_now - c->spread);
  double skin2 = (c->par.skin) / 2.0 + dspread;
  if (rmax > skin2) {
    if (c->P > 1) c->spread = spread_now;
    return 1;
  }
  return 0;
}

static void or_build_list(or_ctx *c, const double *x, const double *y, const double *z,
                          const double *xm, const double *ym, const double *zm) {
  int N = c->N, cap = c->nver_cap, nver = 0;
  if (c->jter == NULL) {
    cap = 1024;
    c->jter = malloc(sizeof(int) * cap);
  }
  for (int i = 1; i <= N; i++) {
    c->x0[i] = x[i - 1];
    c->y0[i] = y[i - 1];
    c->z0[i] = z[i - 1];
    c->nter[i] = 0;
  }
  for (int ip = 2; ip <= N; ip++) {
    int ex_lo = c->ex_off[ip] + 1, ex_hi = c->ex_off[ip] + c->ex_num[ip], ek = ex_lo;
    int ti = c->type[ip - 1];
    double xi = x[ip - 1], yi = y[ip - 1], zi = z[ip - 1];
    for (int jp = 1; jp < ip; jp++) {
      if (ek <= ex_hi && c->ex_j[ek] == jp) {
        ek++;
        continue;
      }
      double dx = xi - x[jp - 1], dy = yi - y[jp - 1], dz = zi - z[jp - 1];
      if (c->P == 1) {
        or_period1(c, &dx, &dy, &dz);
      } else {
        or_period_pimd1(c, &dx, &dy, &dz, xm[ip - 1] - xm[jp - 1], ym[ip - 1] - ym[jp - 1],
                        zm[ip - 1] - zm[jp - 1]);
      }
      double r2 = dx * dx + dy * dy + dz * dz;
      int t = or_pair(c, ti, c->type[jp - 1]);
      if (r2 < c->cutskin[t]) {
        if (nver == cap) {
          cap *= 2;
          c->jter = realloc(c->jter, sizeof(int) * cap);
        }
        c->jter[nver++] = jp;
        c->nter[ip]++;
      }
    }
  }
  c->nver_cap = cap;
  c->nver = nver;
  c->jver_off[1] = 0;
  for (int i = 2; i <= N; i++) c->jver_off[i] = c->jver_off[i - 1] + c->nter[i - 1];
  c->have_list = 1;
}

/* ------------------------------------------------------------------------
 * real-space pair forces of one bead over the Verlet list: force_verlst
 * (force_verlst.c:24-176) + force_npol (force_npol.c:28-328) with
 * shave_skin_opt on, RESPA off (all weights 1).  Also the virial estimator
 * forces of npol_vir_reduc.
 */
static void or_spline_eval4(const double *tab, int k, double del, double swit, double *out) {
  /* npol_vspl_fetch, force_npol.c:339-376; k is PINY's 1-based index */
  static const double oneth = (1.0 / 3.0), onesi = (1.0 / 6.0);
  double c0 = tab[k - 1], fp1 = tab[k], fm1 = tab[k - 2], fm2 = tab[k - 3];
  double c1 = oneth * fp1 + 0.5 * c0 - fm1 + onesi * fm2;
  double c2 = 0.5 * fp1 - c0 + 0.5 * fm1;
  double c3 = onesi * fp1 - 0.5 * c0 + 0.5 * fm1 - onesi * fm2;
  *out = swit * (c0 + del * (c1 + del * (c2 + del * c3)));
}

/* One bead's real-space pair terms (force_npol.c:28-328).  full / res are
 * iget_full_inter / iget_res_inter: the hard cutoff is the full one or,
 * for the short-range part alone, inter_cutoff_res (npol_shave_skin,
 * :1291-1395); the force of each term is weighted by wght_ter (full),
 * by the RESPA switch times wght_ter_res (res), or by their blend (both)
 * (npol_spldata_fetch :807-816, npol_force_calc :1029-1060); energies only
 * with full (:243-256).  pvt gets the weighted virial (pten_sum_roll), pvtt
 * (full only, may be NULL) the unweighted one, fxt (may be NULL) the
 * unweighted forces (npol_vir_reduc, :1222-1287). */
static void or_real_bead(const or_ctx *c, const double *x, const double *y, const double *z,
                         const double *xm, const double *ym, const double *zm, double *fx,
                         double *fy, double *fz, double *vvdw, double *vcoul, double *pvt, int full,
                         int res, double wght_full, double wght_res, double *pvtt, double *fxt,
                         double *fyt, double *fzt) {
  int N = c->N, n = c->nsplin;
  double vv = 0.0, vt = 0.0;
  int icoul = (c->nchrg > 0);
  const double rheal_res = c->par.rheal, rheal_resi = 1.0 / c->par.rheal;
  for (int ip = 2; ip <= N; ip++) {
    int ti = c->type[ip - 1];
    double xi = x[ip - 1], yi = y[ip - 1], zi = z[ip - 1], qi = c->q[ip - 1];
    double fxi = 0.0, fyi = 0.0, fzi = 0.0;
    for (int k = 0; k < c->nter[ip]; k++) {
      int jp = c->jter[c->jver_off[ip] + k];
      double dx = xi - x[jp - 1], dy = yi - y[jp - 1], dz = zi - z[jp - 1];
      if (c->P == 1) {
        or_period1(c, &dx, &dy, &dz);
      } else {
        or_period_pimd1(c, &dx, &dy, &dz, xm[ip - 1] - xm[jp - 1], ym[ip - 1] - ym[jp - 1],
                        zm[ip - 1] - zm[jp - 1]);
      }
      double r2 = (dx * dx + dy * dy + dz * dz);
      double r = sqrt(r2);
      int t = or_pair(c, ti, c->type[jp - 1]);
      const double cut = full ? c->cutoff[t] : c->cutoff_res[t];
      if (!(r < cut)) continue; /* npol_shave_skin */
      double rmin = c->rmin_spl[t], dr = c->dr_spl[t], dri = c->dri_spl[t];
      int kk = (int)(((r - rmin) * dri) + 0.5) + 3;
      kk = imin(kk, n - 2);
      kk = imax(kk, 3);
      double rsp = ((double)(kk - 3)) * dr + rmin;
      double del = (r - rsp) * dri;
      double swit_hard = ((cut - r) >= 0.0 ? 1.0 : 0.0);
      const double *tv = c->cv0 + (size_t)t * n, *tdv = c->cdv0 + (size_t)t * n;
      const double *tvc = c->cv0_c + (size_t)t * n, *tdvc = c->cdv0_c + (size_t)t * n;
      double q2 = qi * c->q[jp - 1];
      double vlj, vc = 0.0, dvlj, dvc = 0.0;
      or_spline_eval4(tv, kk, del, swit_hard, &vlj);
      or_spline_eval4(tdv, kk, del, swit_hard, &dvlj);
      double dvtemp;
      if (icoul) {
        or_spline_eval4(tvc, kk, del, swit_hard, &vc);
        or_spline_eval4(tdvc, kk, del, swit_hard, &dvc);
        dvtemp = (q2 * dvc + dvlj);
      } else {
        dvtemp = dvlj;
      }
      if (full) {
        vv += vlj;
        vt += vlj + q2 * vc;
      }
      double fx1 = dx * dvtemp, fy1 = dy * dvtemp, fz1 = dz * dvtemp;
      double fmult = wght_full;
      if (res) {
        double br = (r - c->cutoff_res[t] + rheal_res) * rheal_resi;
        br = dmax(br, 0.0);
        br = dmin(br, 1.0);
        const double sw = br * br * (br * 2.0 - 3.0) + 1.0;
        fmult = full ? (sw * wght_res + (1.0 - sw) * wght_full) : sw * wght_res;
      }
      double fx2 = fx1 * fmult, fy2 = fy1 * fmult, fz2 = fz1 * fmult;
      fxi += fx2;
      fyi += fy2;
      fzi += fz2;
      fx[jp - 1] -= fx2;
      fy[jp - 1] -= fy2;
      fz[jp - 1] -= fz2;
      pvt[1] += dx * fx2;
      pvt[5] += dy * fy2;
      pvt[9] += dz * fz2;
      pvt[2] += dx * fy2;
      pvt[3] += dx * fz2;
      pvt[6] += dy * fz2;
      if (pvtt) {
        pvtt[1] += dx * fx1;
        pvtt[5] += dy * fy1;
        pvtt[9] += dz * fz1;
        pvtt[2] += dx * fy1;
        pvtt[3] += dx * fz1;
        pvtt[6] += dy * fz1;
      }
      if (fxt) {
        fxt[ip - 1] += fx1;
        fyt[ip - 1] += fy1;
        fzt[ip - 1] += fz1;
        fxt[jp - 1] -= fx1;
        fyt[jp - 1] -= fy1;
        fzt[jp - 1] -= fz1;
      }
    }
    fx[ip - 1] += fxi;
    fy[ip - 1] += fyi;
    fz[ip - 1] += fzi;
  }
  *vvdw += vv;
  *vcoul += vt - vv;
}

/* ------------------------------------------------------------------------
 * k-space.  A direct DFT along each axis with the FFTPACK conventions of
 * fft_gen1d (fft_package.c:1359-1421): sign = +1 is DCFFTB (exp(+i)),
 * sign = -1 is DCFFTF (exp(-i)); unnormalised.  Grid layout as
 * ewald3d_pme.c:282-298 (a fastest), complex interleaved.
 */
/* in-place radix-2 transform of one gathered line (K a power of two) */
static void or_fft_line(double *re, double *im, int K, int sign) {
  for (int i = 1, j = 0; i < K; i++) {
    int bit = K >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= K; len <<= 1) {
    int half = len >> 1;
    for (int st = 0; st < K; st += len) {
      for (int p = 0; p < half; p++) {
        double a = 2.0 * M_PI * (double)p / (double)len;
        double wr = cos(a), wi = sign * sin(a);
        int e0 = st + p, e1 = e0 + half;
        double vr = re[e1] * wr - im[e1] * wi, vi = re[e1] * wi + im[e1] * wr;
        re[e1] = re[e0] - vr;
        im[e1] = im[e0] - vi;
        re[e0] += vr;
        im[e0] += vi;
      }
    }
  }
}

static void or_dft_axis(double *g, const int *ng, int axis, int sign) {
  int K = ng[axis];
  if ((K & (K - 1)) == 0) {
    size_t stride = 1;
    for (int d = 0; d < axis; d++) stride *= ng[d];
    size_t total = (size_t)ng[0] * ng[1] * ng[2];
    double *re = malloc(sizeof(double) * K), *im = malloc(sizeof(double) * K);
    for (size_t base = 0; base < total; base++) {
      if ((base / stride) % K != 0) continue;
      for (int j = 0; j < K; j++) {
        re[j] = g[2 * (base + j * stride)];
        im[j] = g[2 * (base + j * stride) + 1];
      }
      or_fft_line(re, im, K, sign);
      for (int j = 0; j < K; j++) {
        g[2 * (base + j * stride)] = re[j];
        g[2 * (base + j * stride) + 1] = im[j];
      }
    }
    free(re);
    free(im);
    return;
  }
  double *tw = malloc(sizeof(double) * 2 * K);
  for (int m = 0; m < K; m++) {
    double a = 2.0 * M_PI * (double)m / (double)K;
    tw[2 * m] = cos(a);
    tw[2 * m + 1] = sign * sin(a);
  }
  size_t stride = 1;
  for (int d = 0; d < axis; d++) stride *= ng[d];
  size_t total = (size_t)ng[0] * ng[1] * ng[2];
  double *in = malloc(sizeof(double) * 2 * K), *out = malloc(sizeof(double) * 2 * K);
  for (size_t base = 0; base < total; base++) {
    if ((base / stride) % K != 0) continue;
    for (int j = 0; j < K; j++) {
      in[2 * j] = g[2 * (base + j * stride)];
      in[2 * j + 1] = g[2 * (base + j * stride) + 1];
    }
    for (int k = 0; k < K; k++) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < K; j++) {
        int m = (int)(((long)k * j) % K);
        double cr = tw[2 * m], ci = tw[2 * m + 1];
        sr += in[2 * j] * cr - in[2 * j + 1] * ci;
        si += in[2 * j] * ci + in[2 * j + 1] * cr;
      }
      out[2 * k] = sr;
      out[2 * k + 1] = si;
    }
    for (int j = 0; j < K; j++) {
      g[2 * (base + j * stride)] = out[2 * j];
      g[2 * (base + j * stride) + 1] = out[2 * j + 1];
    }
  }
  free(tw);
  free(in);
  free(out);
}

/* one charged site's B-spline data: ewald3d_pme.c:241-337 and :516-530 */
typedef struct {
  int ig[3][32];
  double mn[3][32], dmn[3][32];
} or_bsp;

static void or_bspline_site(const or_ctx *c, double x, double y, double z, or_bsp *b) {
  const double *hi = c->hmati;
  int n_interp = c->par.n_interp;
  double s[3];
  s[0] = x * hi[1] + y * hi[4] + z * hi[7];
  s[1] = x * hi[2] + y * hi[5] + z * hi[8];
  s[2] = x * hi[3] + y * hi[6] + z * hi[9];
  double aj[32], rn[32], rn1[32];
  for (int j = 1; j <= n_interp; j++) {
    aj[j] = (double)(j - 1);
    rn[j] = (double)(j);
    if (j > 1) rn1[j] = 1.0 / ((double)(j - 1));
  }
  rn1[1] = 0.0;
  for (int d = 0; d < 3; d++) {
    int K = c->ngrid[d];
    double t = s[d];
    t = t - nint_piny((t - 0.5));
    double sc = t * (double)K;
    int it = (int)(sc);
    double frac = sc - (double)(it);
    double u[32], *mn = b->mn[d], *dmn = b->dmn[d];
    for (int j = 1; j <= n_interp; j++) {
      u[j] = frac + aj[j];
      int ia = it - (j - 2);
      ia = (ia > 0 ? ia : K + ia);
      if (ia > K) ia -= K; /* only for a coordinate exactly on the cell face */
      b->ig[d][j] = ia - 1;
    }
    mn[1] = 1.0 - fabs(u[1] - 1.0);
    mn[2] = 1.0 - fabs(u[2] - 1.0);
    for (int j = 3; j <= n_interp; j++) mn[j] = 0.0;
    for (int n = 3; n <= n_interp; n++) {
      for (int j = n; j >= 2; j--) {
        double tmp = (u[j] * mn[j] + (rn[n] - u[j]) * mn[j - 1]) * rn1[n];
        mn[j] = tmp;
      }
      mn[1] = u[1] * mn[1] * rn1[n];
      if (n == (n_interp - 1)) {
        dmn[1] = mn[1];
        for (int j = 2; j <= n_interp; j++) dmn[j] = mn[j] - mn[j - 1];
      }
    }
  }
}

/* ewald3d_recip_pme_none (ewald3d_pme.c:95-591) for one configuration */
static double or_pme(const or_ctx *c, const double *x, const double *y, const double *z, double *fx,
                     double *fy, double *fz, double *pvt) {
  const int *ng = c->ngrid;
  int n_interp = c->par.n_interp;
  size_t ntot = (size_t)ng[0] * ng[1] * ng[2];
  double *grid = calloc(2 * ntot, sizeof(double));
  or_bsp *bs = malloc(sizeof(or_bsp) * (c->nchrg > 0 ? c->nchrg : 1));
  size_t sb = ng[0], sc = (size_t)ng[0] * ng[1];
  for (int a = 0; a < c->nchrg; a++) {
    int i = c->ichrg[a];
    or_bspline_site(c, x[i], y[i], z[i], &bs[a]);
    double q = c->q[i];
    for (int jc = 1; jc <= n_interp; jc++)
      for (int jb = 1; jb <= n_interp; jb++)
        for (int ja = 1; ja <= n_interp; ja++) {
          size_t idx = bs[a].ig[0][ja] + sb * bs[a].ig[1][jb] + sc * bs[a].ig[2][jc];
          grid[2 * idx] += bs[a].mn[0][ja] * bs[a].mn[1][jb] * bs[a].mn[2][jc] * q;
        }
  }
  for (int d = 0; d < 3; d++) or_dft_axis(grid, ng, d, +1);
  /* influence function over the half-space sphere list (:367-430) */
  const double *hi = c->hmati;
  double falp2 = 4.0 * c->alp_ewd * c->alp_ewd, tpi = 2.0 * M_PI, pivol = c->vol / (4.0 * M_PI);
  double *conv = calloc(2 * ntot, sizeof(double));
  double vnow = 0.0, pt[10] = {0};
  for (int k = 0; k < c->nktot; k++) {
    double aka = tpi * ((double)c->kka[k]), akb = tpi * ((double)c->kkb[k]), akc = tpi * ((double)c->kkc[k]);
    double xk = (aka * hi[1] + akb * hi[2] + akc * hi[3]);
    double yk = (aka * hi[4] + akb * hi[5] + akc * hi[6]);
    double zk = (aka * hi[7] + akb * hi[8] + akc * hi[9]);
    double g2 = xk * xk + yk * yk + zk * zk;
    double preg = exp((-g2 / falp2)) / (g2 * pivol);
    double prep = -2.0 * preg * ((g2 / falp2) + 1.0) / g2;
    int ia = (c->kka[k] % ng[0] + ng[0]) % ng[0], ib = (c->kkb[k] % ng[1] + ng[1]) % ng[1],
        ic = (c->kkc[k] % ng[2] + ng[2]) % ng[2];
    size_t idx = ia + sb * ib + sc * ic;
    double sr = grid[2 * idx], si = grid[2 * idx + 1];
    double smag = (sr * sr + si * si) * c->bweight[k];
    vnow += (smag * preg);
    prep *= smag;
    pt[1] += prep * xk * xk;
    pt[5] += prep * yk * yk;
    pt[9] += prep * zk * zk;
    pt[2] += prep * xk * yk;
    pt[3] += prep * xk * zk;
    pt[6] += prep * yk * zk;
    double w = preg * c->bweight[k];
    conv[2 * idx] = sr * w;
    conv[2 * idx + 1] = si * w;
    int ja = ((-c->kka[k]) % ng[0] + ng[0]) % ng[0], jb = ((-c->kkb[k]) % ng[1] + ng[1]) % ng[1],
        jc = ((-c->kkc[k]) % ng[2] + ng[2]) % ng[2];
    size_t jdx = ja + sb * jb + sc * jc;
    conv[2 * jdx] = sr * w;
    conv[2 * jdx + 1] = -si * w;
  }
  pt[4] = pt[2];
  pt[7] = pt[3];
  pt[8] = pt[6];
  pt[1] += vnow;
  pt[5] += vnow;
  pt[9] += vnow;
  for (int i = 1; i <= 9; i++) pvt[i] += pt[i];
  for (int d = 0; d < 3; d++) or_dft_axis(conv, ng, d, -1);
  /* gather (:534-557) */
  double ga = (double)ng[0], gb = (double)ng[1], gc = (double)ng[2];
  for (int a = 0; a < c->nchrg; a++) {
    int i = c->ichrg[a];
    double q = c->q[i], sfx = 0.0, sfy = 0.0, sfz = 0.0;
    for (int jc = 1; jc <= n_interp; jc++)
      for (int jb = 1; jb <= n_interp; jb++)
        for (int ja = 1; ja <= n_interp; ja++) {
          size_t idx = bs[a].ig[0][ja] + sb * bs[a].ig[1][jb] + sc * bs[a].ig[2][jc];
          double at = bs[a].dmn[0][ja] * bs[a].mn[1][jb] * bs[a].mn[2][jc] * ga;
          double bt = bs[a].mn[0][ja] * bs[a].dmn[1][jb] * bs[a].mn[2][jc] * gb;
          double ct = bs[a].mn[0][ja] * bs[a].mn[1][jb] * bs[a].dmn[2][jc] * gc;
          double qdx = at * hi[1] + bt * hi[2] + ct * hi[3];
          double qdy = at * hi[4] + bt * hi[5] + ct * hi[6];
          double qdz = at * hi[7] + bt * hi[8] + ct * hi[9];
          double qn = conv[2 * idx] * q;
          sfx -= (qdx * qn);
          sfy -= (qdy * qn);
          sfz -= (qdz * qn);
        }
    fx[i] += sfx;
    fy[i] += sfy;
    fz[i] += sfz;
  }
  free(grid);
  free(conv);
  free(bs);
  return vnow;
}

/* standard Ewald, ewald3d_recip (ewald3d.c:27-373), direct phases */
static double or_ewald(const or_ctx *c, const double *x, const double *y, const double *z, double *fx,
                       double *fy, double *fz, double *pvt) {
  const double *hi = c->hmati;
  int nq = c->nchrg;
  double *sa = malloc(sizeof(double) * (nq + 1) * 3), *ff = calloc((nq + 1) * 3, sizeof(double));
  double *cs = malloc(sizeof(double) * (nq + 1) * 2);
  for (int a = 0; a < nq; a++) {
    int i = c->ichrg[a];
    sa[3 * a] = x[i] * hi[1] + y[i] * hi[4] + z[i] * hi[7];
    sa[3 * a + 1] = x[i] * hi[2] + y[i] * hi[5] + z[i] * hi[8];
    sa[3 * a + 2] = x[i] * hi[3] + y[i] * hi[6] + z[i] * hi[9];
  }
  double falp2 = 4.0 * c->alp_ewd * c->alp_ewd, tpi = 2.0 * M_PI, pivol = c->vol / (4.0 * M_PI);
  double vnow = 0.0, pt[10] = {0};
  for (int k = 0; k < c->nktot; k++) {
    double aka = tpi * ((double)c->kka[k]), akb = tpi * ((double)c->kkb[k]), akc = tpi * ((double)c->kkc[k]);
    double xk = (aka * hi[1] + akb * hi[2] + akc * hi[3]);
    double yk = (aka * hi[4] + akb * hi[5] + akc * hi[6]);
    double zk = (aka * hi[7] + akb * hi[8] + akc * hi[9]);
    double g2 = xk * xk + yk * yk + zk * zk;
    double preg = exp((-g2 / falp2)) / (g2 * pivol);
    double prep = -2.0 * preg * ((g2 / falp2) + 1.0) / g2;
    double sumr = 0.0, sumi = 0.0;
    for (int a = 0; a < nq; a++) {
      double arg = aka * sa[3 * a] + akb * sa[3 * a + 1] + akc * sa[3 * a + 2];
      cs[2 * a] = cos(arg);
      cs[2 * a + 1] = sin(arg);
      double qi = c->q[c->ichrg[a]];
      sumr += cs[2 * a] * qi;
      sumi += cs[2 * a + 1] * qi;
    }
    double smag = sumr * sumr + sumi * sumi;
    vnow += smag * preg;
    prep *= smag;
    pt[1] += prep * xk * xk;
    pt[5] += prep * yk * yk;
    pt[9] += prep * zk * zk;
    pt[2] += prep * xk * yk;
    pt[3] += prep * xk * zk;
    pt[6] += prep * yk * zk;
    sumr *= (preg * 2.0);
    sumi *= (preg * 2.0);
    for (int a = 0; a < nq; a++) {
      double fdiff = cs[2 * a + 1] * sumr - cs[2 * a] * sumi;
      ff[3 * a] += xk * fdiff;
      ff[3 * a + 1] += yk * fdiff;
      ff[3 * a + 2] += zk * fdiff;
    }
  }
  for (int a = 0; a < nq; a++) {
    int i = c->ichrg[a];
    double qi = c->q[i];
    fx[i] += ff[3 * a] * qi;
    fy[i] += ff[3 * a + 1] * qi;
    fz[i] += ff[3 * a + 2] * qi;
  }
  pt[4] = pt[2];
  pt[7] = pt[3];
  pt[8] = pt[6];
  pt[1] += vnow;
  pt[5] += vnow;
  pt[9] += vnow;
  for (int i = 1; i <= 9; i++) pvt[i] += pt[i];
  free(sa);
  free(ff);
  free(cs);
  return vnow;
}

/* ------------------------------------------------------------------------
 * intramolecular terms of one bead: ecor (ecorr.c:28-265), bond
 * (bond.c:29-312; bond_both is identical with eq_res = eq and all RESPA
 * weights 1), bend (bend.c:28-434).  intra_perds selects the image.
 */
static void or_intra_sep(const or_ctx *c, int i, int j, const double *x, const double *y,
                         const double *z, const double *xm, const double *ym, const double *zm,
                         double *dx, double *dy, double *dz) {
  *dx = x[i] - x[j];
  *dy = y[i] - y[j];
  *dz = z[i] - z[j];
  if (c->par.intra_perds == 1) {
    if (c->P == 1)
      or_period1(c, dx, dy, dz);
    else
      or_period_pimd1(c, dx, dy, dz, xm[i] - xm[j], ym[i] - ym[j], zm[i] - zm[j]);
  }
}

static double or_ecor(const or_ctx *c, const double *x, const double *y, const double *z,
                      const double *xm, const double *ym, const double *zm, double *fx, double *fy,
                      double *fz, double *pvt) {
  double v = 0.0;
  int n = c->nsplin;
  for (int k = 0; k < c->necor; k++) {
    int i = c->e1[k] - 1, j = c->e2[k] - 1;
    double dx, dy, dz;
    or_intra_sep(c, i, j, x, y, z, xm, ym, zm, &dx, &dy, &dz);
    double r122 = dx * dx + dy * dy + dz * dz;
    double r12 = sqrt(r122);
    int kk = (int)(((r12 - c->e_rmin) * c->e_dri) + 0.5) + 3;
    kk = imin(kk, n - 2);
    kk = imax(kk, 3);
    double rsp = ((double)(kk - 3)) * c->e_dr + c->e_rmin;
    double del = (r12 - rsp) * c->e_dri;
    double q12 = c->q[i] * c->q[j];
    double sv, sdv;
    or_spline_eval4(c->ecv0, kk, del, 1.0, &sv);
    or_spline_eval4(c->ecdv0, kk, del, 1.0, &sdv);
    v += sv * q12;
    double dvt = sdv * q12;
    double f1x = dx * dvt, f1y = dy * dvt, f1z = dz * dvt;
    pvt[1] += dx * f1x;
    pvt[5] += dy * f1y;
    pvt[9] += dz * f1z;
    pvt[2] += dx * f1y;
    pvt[3] += dx * f1z;
    pvt[6] += dy * f1z;
    fx[i] += f1x;
    fy[i] += f1y;
    fz[i] += f1z;
    fx[j] -= f1x;
    fy[j] -= f1y;
    fz[j] -= f1z;
  }
  return v;
}

static double or_bond(const or_ctx *c, const double *x, const double *y, const double *z,
                      const double *xm, const double *ym, const double *zm, double *fx, double *fy,
                      double *fz, double *pvt) {
  double v = 0.0;
  for (int b = 0; b < c->nbond; b++) {
    int i = c->b1[b], j = c->b2[b], t = c->btyp[b];
    const double *cc = c->bc + 7 * t, *dc = c->bdc + 7 * t;
    double dx, dy, dz;
    or_intra_sep(c, i, j, x, y, z, xm, ym, zm, &dx, &dy, &dz);
    double r122 = dx * dx + dy * dy + dz * dz;
    double r12 = sqrt(r122);
    double rr0 = r12 - c->beq[t];
    double vb = (((((cc[6] * rr0 + cc[5]) * rr0 + cc[4]) * rr0 + cc[3]) * rr0 + cc[2]) * rr0 + cc[1]) * rr0 + cc[0];
    double dvb = ((((dc[6] * rr0 + dc[5]) * rr0 + dc[4]) * rr0 + dc[3]) * rr0 + dc[2]) * rr0 + dc[1];
    double pre = -dvb / r12;
    double f1x = dx * pre, f1y = dy * pre, f1z = dz * pre;
    v += vb;
    pvt[1] += dx * f1x;
    pvt[5] += dy * f1y;
    pvt[9] += dz * f1z;
    pvt[2] += dx * f1y;
    pvt[3] += dx * f1z;
    pvt[6] += dy * f1z;
    fx[i] += f1x;
    fy[i] += f1y;
    fz[i] += f1z;
    fx[j] += -f1x;
    fy[j] += -f1y;
    fz[j] += -f1z;
  }
  return v;
}

static double or_bend(const or_ctx *c, const double *x, const double *y, const double *z,
                      const double *xm, const double *ym, const double *zm, double *fx, double *fy,
                      double *fz, double *pvt) {
  double v = 0.0, seps = 1.0e-8;
  for (int b = 0; b < c->nbend; b++) {
    int i1 = c->a1[b], i2 = c->a2[b], i3 = c->a3[b], t = c->atyp[b];
    const double *cc = c->ac + 7 * t, *ss = c->as + 7 * t, *dc = c->adc + 7 * t, *ds = c->ads + 7 * t;
    double dx12, dy12, dz12, dx23, dy23, dz23;
    or_intra_sep(c, i1, i2, x, y, z, xm, ym, zm, &dx12, &dy12, &dz12);
    or_intra_sep(c, i3, i2, x, y, z, xm, ym, zm, &dx23, &dy23, &dz23);
    double r122i = 1.0 / (dx12 * dx12 + dy12 * dy12 + dz12 * dz12);
    double r12i = sqrt(r122i);
    double r322i = 1.0 / (dx23 * dx23 + dy23 * dy23 + dz23 * dz23);
    double r32i = sqrt(r322i);
    double rpmag = r12i * r32i;
    double cost = (dx12 * dx23 + dy12 * dy23 + dz12 * dz23) * rpmag;
    cost = (cost < 1.0 ? cost : 1.0);
    cost = (cost > -1.0 ? cost : -1.0);
    double sint = sqrt(1.0 - cost * cost);
    sint = (sint > seps ? sint : seps);
    double cos122 = cost * r122i, cos322 = cost * r322i;
    double vbc = ((((((cc[6] * cost + cc[5]) * cost + cc[4]) * cost + cc[3]) * cost + cc[2]) * cost + cc[1]) * cost + cc[0]);
    double sisum = (((((ss[6] * sint + ss[5]) * sint + ss[4]) * sint + ss[3]) * sint + ss[2]) * sint);
    double vbs = sisum * cost + ss[1] * sint;
    v += vbc + vbs;
    double dvbc = (((((dc[6] * cost + dc[5]) * cost + dc[4]) * cost + dc[3]) * cost + dc[2]) * cost + dc[1]);
    double dvbs = ((((ds[6] * sint + ds[5]) * sint + ds[4]) * sint + ds[3]) * sint + ds[2]);
    dvbs = dvbs * cost + ds[1];
    double pre = -(dvbc + sisum - dvbs * cost / sint);
    double f1x = (dx23 * rpmag - dx12 * cos122) * pre;
    double f1y = (dy23 * rpmag - dy12 * cos122) * pre;
    double f1z = (dz23 * rpmag - dz12 * cos122) * pre;
    double f3x = (dx12 * rpmag - dx23 * cos322) * pre;
    double f3y = (dy12 * rpmag - dy23 * cos322) * pre;
    double f3z = (dz12 * rpmag - dz23 * cos322) * pre;
    double f2x = -(f1x + f3x), f2y = -(f1y + f3y), f2z = -(f1z + f3z);
    pvt[1] += dx12 * f1x + dx23 * f3x;
    pvt[5] += dy12 * f1y + dy23 * f3y;
    pvt[9] += dz12 * f1z + dz23 * f3z;
    pvt[2] += dx12 * f1y + dx23 * f3y;
    pvt[3] += dx12 * f1z + dx23 * f3z;
    pvt[6] += dy12 * f1z + dy23 * f3z;
    fx[i1] += f1x;
    fy[i1] += f1y;
    fz[i1] += f1z;
    fx[i2] += f2x;
    fy[i2] += f2y;
    fz[i2] += f2z;
    fx[i3] += f3x;
    fy[i3] += f3y;
    fz[i3] += f3z;
  }
  return v;
}

/* ------------------------------------------------------------------------
 * normal-mode transforms of transform_cnt.c:240-489, naive DFT over beads.
 * Mode storage order: u[0] centroid, u[P-1] the alternating mode, and for
 * k = 1..P/2-1 the real part at u[2k-1] and the imaginary part at u[2k].
 */
static void or_dft_beads(int P, const double *re, const double *im, int sign, double *ore,
                         double *oim) {
  for (int k = 0; k < P; k++) {
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < P; j++) {
      double a = 2.0 * M_PI * (double)(((long)k * j) % P) / (double)P;
      double cr = cos(a), ci = sign * sin(a);
      sr += re[j] * cr - im[j] * ci;
      si += re[j] * ci + im[j] * cr;
    }
    ore[k] = sr;
    oim[k] = si;
  }
}

void or_nm_pos_to_mode(int P, const double *x, double *u) {
  double *zr = calloc(P, sizeof(double)), *zi = calloc(P, sizeof(double));
  double *orr = calloc(P, sizeof(double)), *oi = calloc(P, sizeof(double));
  for (int j = 0; j < P; j++) zr[j] = x[j];
  or_dft_beads(P, zr, zi, +1, orr, oi); /* DCFFTB, then /P */
  for (int k = 0; k < P; k++) {
    orr[k] /= P;
    oi[k] /= P;
  }
  u[0] = orr[0];
  u[P - 1] = orr[P / 2];
  for (int k = 1; k < P / 2; k++) {
    u[2 * k - 1] = orr[k];
    u[2 * k] = oi[k];
  }
  free(zr);
  free(zi);
  free(orr);
  free(oi);
}

void or_nm_mode_to_pos(int P, const double *u, double *x) {
  double *zr = calloc(P, sizeof(double)), *zi = calloc(P, sizeof(double));
  double *orr = calloc(P, sizeof(double)), *oi = calloc(P, sizeof(double));
  zr[0] = u[0];
  zr[P / 2] = u[P - 1];
  for (int k = 1; k < P / 2; k++) {
    zr[k] = u[2 * k - 1];
    zi[k] = u[2 * k];
    zr[P - k] = u[2 * k - 1];
    zi[P - k] = -u[2 * k];
  }
  or_dft_beads(P, zr, zi, -1, orr, oi); /* DCFFTF */
  for (int j = 0; j < P; j++) x[j] = orr[j];
  free(zr);
  free(zi);
  free(orr);
  free(oi);
}

void or_nm_force_to_mode(int P, const double *f, double *fm) {
  double *zr = calloc(P, sizeof(double)), *zi = calloc(P, sizeof(double));
  double *orr = calloc(P, sizeof(double)), *oi = calloc(P, sizeof(double));
  for (int j = 0; j < P; j++) zr[j] = f[j];
  or_dft_beads(P, zr, zi, +1, orr, oi);
  fm[0] = orr[0];
  fm[P - 1] = orr[P / 2];
  for (int k = 1; k < P / 2; k++) {
    fm[2 * k - 1] = orr[k] * 2.0;
    fm[2 * k] = oi[k] * 2.0;
  }
  free(zr);
  free(zi);
  free(orr);
  free(oi);
}

/* ------------------------------------------------------------------------
 * staging transforms of transform_stg.c for one site (path_int_md_typ =
 * staging), with rat1[ip] = (ip-1)/ip and rat2[ip] = 1/ip for 1-based beads
 * ip = 2..P (transform_cnt.c:206-218).  Mode 1 is bead 1 itself.
 */
static double or_rat1(int ip) { return ((double)(ip - 1)) / ((double)(ip)); }
static double or_rat2(int ip) { return 1.0 / ((double)(ip)); }

/* convert_pimd_mode_stag, transform_stg.c:94-160: for ip = 2..P-1,
 * u_ip = x_ip - (rat1 x_{ip+1} + rat2 x_1) with the untransformed x_{ip+1};
 * u_P = x_P - (rat1 x_1 + rat2 x_1). */
void or_stag_pos_to_mode(int P, const double *x, double *u) {
  u[0] = x[0];
  for (int ip = 2; ip <= P - 1; ip++) {
    double xstar = or_rat1(ip) * x[ip] + or_rat2(ip) * x[0];
    u[ip - 1] = x[ip - 1] - xstar;
  }
  if (P > 1) {
    double xstar = or_rat1(P) * x[0] + or_rat2(P) * x[0];
    u[P - 1] = x[P - 1] - xstar;
  }
}

/* convert_pimd_pos_stag, transform_stg.c:168-226: x_P = u_P + x_1, then for
 * ip = P-1 down to 2, x_ip = u_ip + (rat1 x_{ip+1} + x_1 rat2). */
void or_stag_mode_to_pos(int P, const double *u, double *x) {
  for (int k = 0; k < P; k++) x[k] = u[k];
  if (P == 1) return;
  x[P - 1] += x[0];
  for (int ip = P - 1; ip >= 2; ip--) {
    double xadd = or_rat1(ip) * x[ip] + x[0] * or_rat2(ip);
    x[ip - 1] += xadd;
  }
}

/* convert_pimd_force_stag, transform_stg.c:31-88: f_1 += f_ip for
 * ip = 2..P in order, then f_{ip+1} += rat1[ip] f_ip for ip = 2..P-1 (the
 * updated f_ip: a recurrence). */
void or_stag_force_to_mode(int P, const double *f, double *fm) {
  for (int k = 0; k < P; k++) fm[k] = f[k];
  for (int ip = 2; ip <= P; ip++) fm[0] += fm[ip - 1];
  for (int ip = 2; ip <= P - 1; ip++) fm[ip] += (or_rat1(ip) * fm[ip - 1]);
}

/* ------------------------------------------------------------------------
 * energy_control (energy_control.c:33-77) / energy_control_pimd
 * (energy_control_pimd.c:37-278) with RESPA off.
 */
/* RESPA on: any of the step keywords given (set_sim_params.c:1585-1660) */
static int or_respa_on(const or_ctx *c) {
  const pmx_respa *r = &c->respa;
  return r->respa_steps_lrf > 0 || r->respa_steps_torsion > 0 || r->respa_steps_intra > 0 ||
         r->respa_steps_pimd > 0;
}

/* energy_control_initial.c:212-243: wght_ter, wght_ter_res, wght_tra_res
 * from the step counts (a keyword left out counts 1) */
static void or_respa_weights(const or_ctx *c, int pimd_on, double *wt, double *wtr, double *wtrr) {
  const pmx_respa *r = &c->respa;
  const int nter = imax(r->respa_steps_lrf, 1), ntor = imax(r->respa_steps_torsion, 1);
  const int ntra = imax(r->respa_steps_intra, 1), npimd = imax(r->respa_steps_pimd, 1);
  if (pimd_on) {
    *wt = (double)(npimd * ntra * ntor * nter);
    *wtr = (double)(npimd * ntra * ntor);
    *wtrr = (double)(npimd);
  } else {
    *wt = (double)(ntra * ntor * nter);
    *wtr = (double)(ntra * ntor);
    *wtrr = 1.0;
  }
}

/* Adds one term's forces t and virial tp with weight w, as the reference
 * scales each term's forces before adding them (ewald3d_pme.c:575-577,
 * ecorr.c:229-231, bond.c:275-277, bend.c:386-394): f += w t, pvt += w tp;
 * unweighted into pvtt (NULL outside full calls) and ft (NULL without the
 * virial estimator). */
static void or_add_term(int N, double w, const double *tx, const double *ty, const double *tz, const double *tp,
                        double *fx, double *fy, double *fz, double *pvt, double *pvtt, double *ftx, double *fty,
                        double *ftz) {
  for (int i = 0; i < N; i++) {
    fx[i] += tx[i] * w;
    fy[i] += ty[i] * w;
    fz[i] += tz[i] * w;
  }
  for (int k = 1; k <= 9; k++) pvt[k] += tp[k] * w;
  if (pvtt)
    for (int k = 1; k <= 9; k++) pvtt[k] += tp[k];
  if (ftx)
    for (int i = 0; i < N; i++) {
      ftx[i] += tx[i];
      fty[i] += ty[i];
      ftz[i] += tz[i];
    }
}

int or_set_respa(or_ctx *c, const pmx_respa *r) {
  if (!c) return PMX_ERR_ARG;
  if (!r) {
    memset(&c->respa, 0, sizeof(c->respa));
    return PMX_OK;
  }
  if (r->respa_steps_lrf < 0 || r->respa_steps_torsion < 0 || r->respa_steps_intra < 0 ||
      r->respa_steps_pimd < 0 || r->nfree < 0)
    return PMX_ERR_ARG;
  /* set_sim_params.c:3346-3370 */
  if (r->respa_steps_lrf > 0 && (r->respa_steps_intra == 0 || r->respa_steps_torsion == 0)) return PMX_ERR_ARG;
  if (r->respa_steps_torsion > 0 && r->respa_steps_intra == 0) return PMX_ERR_ARG;
  c->respa = *r;
  return PMX_OK;
}

int or_energy_control(or_ctx *c, double *x, double *y, double *z, const double *xmi,
                      const double *ymi, const double *zmi, unsigned flags, double *fx,
                      double *fy, double *fz, pmx_energies *out) {
  int N = c->N, P = c->P;
  if (P > 1 && !(flags & PMX_EVAL_PIMD)) return PMX_ERR_ARG;
  memset(out, 0, sizeof(*out));
  /* ghosts of every bead */
  for (int ip = 0; ip < P; ip++) or_ghost_pos(c, x + (size_t)ip * N, y + (size_t)ip * N, z + (size_t)ip * N);
  /* xmod: given, else mode 1 of the bead transform (control_pimd_trans.c:
   * 175-181 copies it from mode 1 before the back transform): bead 1 with
   * staging, the bead centroid with normal modes */
  double *xm = calloc(N, sizeof(double)), *ym = calloc(N, sizeof(double)), *zm = calloc(N, sizeof(double));
  if (xmi) {
    memcpy(xm, xmi, sizeof(double) * N);
    memcpy(ym, ymi, sizeof(double) * N);
    memcpy(zm, zmi, sizeof(double) * N);
  } else if (c->par.pimd_typ == PMX_PIMD_STAGING) {
    memcpy(xm, x, sizeof(double) * N);
    memcpy(ym, y, sizeof(double) * N);
    memcpy(zm, z, sizeof(double) * N);
  } else {
    for (int i = 0; i < N; i++) {
      double sx = 0, sy = 0, sz = 0;
      for (int ip = 0; ip < P; ip++) {
        sx += x[(size_t)ip * N + i];
        sy += y[(size_t)ip * N + i];
        sz += z[(size_t)ip * N + i];
      }
      xm[i] = sx / P;
      ym[i] = sy / P;
      zm[i] = sz / P;
    }
  }
  memset(fx, 0, sizeof(double) * (size_t)P * N);
  memset(fy, 0, sizeof(double) * (size_t)P * N);
  memset(fz, 0, sizeof(double) * (size_t)P * N);
  double pvt[10] = {0}, dP = (double)P;
  /* neighbour list (nbr_list_control.c:32-118) */
  double spread_now = 0.0;
  if (P > 1) {
    spread_now = or_spread(c, x, y, z);
    or_cut_skin(c, spread_now);
  } else if (!c->have_list) {
    or_cut_skin(c, 0.0);
  }
  int rebuild = or_check_list(c, x, y, z, spread_now) || (flags & PMX_EVAL_FORCE_LIST);
  if (rebuild) {
    if (P > 1) c->spread = spread_now;
    or_build_list(c, x, y, z, xm, ym, zm);
  }
  out->list_rebuilt = rebuild;
  out->nver = c->nver;
  out->spread = c->spread;
  /* RESPA (energy_control_initial.c:212-243): the force weights from the
   * step counts; the RESPA integrators make every force call with
   * iget_res_intra = 1 (int_NVT_pimd.c:174-187, int_NVE_res.c:106-116),
   * the others with 0 */
  const int full = !(flags & PMX_EVAL_NO_FULL_INTER), res = (flags & PMX_EVAL_RES_INTER) != 0;
  const int full_intra = !(flags & PMX_EVAL_NO_FULL_INTRA);
  const int res_intra = or_respa_on(c);
  const int pimd_on = (flags & PMX_EVAL_PIMD) != 0;
  double wt, wtr, wtrr;
  or_respa_weights(c, pimd_on, &wt, &wtr, &wtrr);
  /* iver_get (energy_control_intra.c:111): the unweighted forces fxt of the
   * virial estimator */
  const int iver = pimd_on && full && P > 1;
  double *fxt = NULL, *fyt = NULL, *fzt = NULL;
  if (iver) {
    fxt = calloc((size_t)P * N, sizeof(double));
    fyt = calloc((size_t)P * N, sizeof(double));
    fzt = calloc((size_t)P * N, sizeof(double));
  }
  double pvtt[10] = {0};
  double *tx = calloc(N, sizeof(double)), *ty = calloc(N, sizeof(double)), *tz = calloc(N, sizeof(double));
  /* real space per bead (energy_control_inter_real.c:112-146) */
  double vvdw = 0.0, vcoul = 0.0;
  for (int ip = 0; ip < P; ip++) {
    size_t o = (size_t)ip * N;
    double a = vvdw, b = vcoul;
    if (full || res)
      or_real_bead(c, x + o, y + o, z + o, xm, ym, zm, fx + o, fy + o, fz + o, &vvdw, &vcoul, pvt, full, res, wt,
                   wtr, full ? pvtt : NULL, iver ? fxt + o : NULL, iver ? fyt + o : NULL, iver ? fzt + o : NULL);
    c->bead_e[4 * ip + 0] = (vvdw - a) + (vcoul - b);
  }
  out->vvdw = vvdw / dP;
  out->vcoul = vcoul / dP;
  out->vreal = out->vvdw + out->vcoul;
  /* long range correction with its RESPA pressure terms
   * (energy_control_inter_real.c:148-157,206-318) */
  double vol = c->vol;
  out->vlong = 0.0;
  {
    const double amult = c->iswit ? 1.0 : 2.0, amult_bead = amult * dP;
    const double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
    double pd_bead = 0.0, pd_tra, pd_ter, d;
    if (P > 1 && wtrr != 1.0) {
      const double pkin = (double)(c->respa.nfree / 3) * T / (OR_BOLTZ * vol);
      pd_bead = (c->respa.pext - pkin) * vol * dP;
      pvt[1] += pd_bead;
      pvt[5] += pd_bead;
      pvt[9] += pd_bead;
    }
    if (res_intra && wt != 1.0) {
      pd_tra = c->respa.pten_inter_guess * vol * dP;
      d = (pd_tra - pd_bead) * wtrr;
      pvt[1] += d;
      pvt[5] += d;
      pvt[9] += d;
    } else {
      pd_tra = pd_bead;
    }
    if (res) {
      pd_ter = (c->clong_res / vol) * amult_bead;
      d = (pd_ter - pd_tra) * wtr;
      pvt[1] += d;
      pvt[5] += d;
      pvt[9] += d;
    } else {
      pd_ter = pd_tra;
    }
    if (full) {
      const double pd_full = (c->clong / vol) * amult_bead;
      d = (pd_full - pd_ter) * wt;
      pvt[1] += d;
      pvt[5] += d;
      pvt[9] += d;
      out->vlong = c->clong / vol;
      pvtt[1] += pd_full;
      pvtt[5] += pd_full;
      pvtt[9] += pd_full;
    }
  }
  /* reciprocal space (energy_control_inter_recip.c:119-355) with no RESPA
   * k vectors (nktot_res = 0): the k-space sum and the excluded-pair
   * correction in full calls, weighted by wght_ter */
  double vrecip = 0.0, vecor = 0.0;
  if (c->nchrg > 0) {
    if (full_intra || !res_intra) {
      double qq = 0.0, qs = 0.0;
      for (int i = 0; i < N; i++) {
        qq += c->q[i] * c->q[i];
        qs += c->q[i];
      }
      out->vself = -qq * (c->alp_ewd) * (c->self_erf) / sqrt(M_PI);
      double bgr = -0.5 * qs * qs * M_PI / (c->alp_ewd * c->alp_ewd * vol);
      out->vbgr = bgr;
      pvt[1] += bgr * wtrr * dP;
      pvt[5] += bgr * wtrr * dP;
      pvt[9] += bgr * wtrr * dP;
      if (full) {
        pvtt[1] += bgr * dP;
        pvtt[5] += bgr * dP;
        pvtt[9] += bgr * dP;
      }
    }
    if (full && !(flags & PMX_EVAL_NO_RECIP)) {
      for (int ip = 0; ip < P; ip++) {
        size_t o = (size_t)ip * N;
        double tp[10] = {0};
        memset(tx, 0, sizeof(double) * N);
        memset(ty, 0, sizeof(double) * N);
        memset(tz, 0, sizeof(double) * N);
        double v = c->par.pme_on ? or_pme(c, x + o, y + o, z + o, tx, ty, tz, tp)
                                 : or_ewald(c, x + o, y + o, z + o, tx, ty, tz, tp);
        or_add_term(N, wt, tx, ty, tz, tp, fx + o, fy + o, fz + o, pvt, pvtt, iver ? fxt + o : NULL,
                    iver ? fyt + o : NULL, iver ? fzt + o : NULL);
        vrecip += v;
        c->bead_e[4 * ip + 1] = v;
      }
    }
    if (full) {
      for (int ip = 0; ip < P; ip++) {
        size_t o = (size_t)ip * N;
        double tp[10] = {0};
        memset(tx, 0, sizeof(double) * N);
        memset(ty, 0, sizeof(double) * N);
        memset(tz, 0, sizeof(double) * N);
        double v = or_ecor(c, x + o, y + o, z + o, xm, ym, zm, tx, ty, tz, tp);
        or_add_term(N, wt, tx, ty, tz, tp, fx + o, fy + o, fz + o, pvt, pvtt, iver ? fxt + o : NULL,
                    iver ? fyt + o : NULL, iver ? fzt + o : NULL);
        vecor += v;
        c->bead_e[4 * ip + 2] = v;
      }
    }
  }
  out->vrecip = vrecip / dP;
  out->vecor = vecor / dP;
  /* intramolecular (energy_control_intra.c:33-331): bonds and bends
   * weighted by wght_tra_res (bond.c:118, bend.c:144; bond_both.c splits
   * a bond at eq_res, which equals eq here) */
  double vbond = 0.0, vbend = 0.0;
  if (full_intra || res_intra) {
    for (int ip = 0; ip < P; ip++) {
      size_t o = (size_t)ip * N;
      double tp[10] = {0};
      memset(tx, 0, sizeof(double) * N);
      memset(ty, 0, sizeof(double) * N);
      memset(tz, 0, sizeof(double) * N);
      double vb = or_bond(c, x + o, y + o, z + o, xm, ym, zm, tx, ty, tz, tp);
      double va = or_bend(c, x + o, y + o, z + o, xm, ym, zm, tx, ty, tz, tp);
      or_add_term(N, wtrr, tx, ty, tz, tp, fx + o, fy + o, fz + o, pvt, full ? pvtt : NULL,
                  iver ? fxt + o : NULL, iver ? fyt + o : NULL, iver ? fzt + o : NULL);
      vbond += vb;
      vbend += va;
      c->bead_e[4 * ip + 3] = vb + va;
    }
  }
  const int keep_e = full_intra || full; /* energies stored (energy_control_intra.c:284) */
  out->vbond = keep_e ? vbond / dP : 0.0;
  out->vbend = keep_e ? vbend / dP : 0.0;
  pvt[4] = pvt[2];
  pvt[7] = pvt[3];
  pvt[8] = pvt[6];
  pvtt[4] = pvtt[2];
  pvtt[7] = pvtt[3];
  pvtt[8] = pvtt[6];
  if (!keep_e) out->vself = out->vbgr = 0.0;
  out->vintert = full ? out->vreal + out->vlong + out->vrecip + out->vself + out->vbgr + out->vecor : 0.0;
  out->vintrat = out->vbond + out->vbend;
  /* PIMD: f/P, pvten/P, virial correction (energy_control_pimd.c:116-144,
   * pimd_estimators.c:223-323): pvten with the (weighted) forces, the
   * virial estimator and pvten_tot with the unweighted fxt, in full calls */
  if (flags & PMX_EVAL_PIMD) {
    for (size_t k = 0; k < (size_t)P * N; k++) {
      fx[k] /= dP;
      fy[k] /= dP;
      fz[k] /= dP;
    }
    if (iver)
      for (size_t k = 0; k < (size_t)P * N; k++) {
        fxt[k] /= dP;
        fyt[k] /= dP;
        fzt[k] /= dP;
      }
    for (int i = 1; i <= 9; i++) {
      pvt[i] /= dP;
      pvtt[i] /= dP;
    }
    if (P > 1) {
      int nreal = N - c->nghost;
      double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
      double beta = OR_BOLTZ / T;
      const int vir = full_intra && full;
      double akvirx = vir ? 3.0 * ((double)nreal) / beta : 0.0;
      for (int ip = 0; ip < P; ip++) {
        size_t o = (size_t)ip * N;
        for (int i = 0; i < N; i++) {
          double dx = x[o + i] - xm[i], dy = y[o + i] - ym[i], dz = z[o + i] - zm[i];
          double gx = fx[o + i], gy = fy[o + i], gz = fz[o + i];
          pvt[1] -= dx * gx;
          pvt[5] -= dy * gy;
          pvt[9] -= dz * gz;
          pvt[2] -= dx * gy;
          pvt[4] -= dy * gx;
          pvt[3] -= dx * gz;
          pvt[7] -= dz * gx;
          pvt[6] -= dy * gz;
          pvt[8] -= dz * gy;
          if (vir) {
            double hx = fxt[o + i], hy = fyt[o + i], hz = fzt[o + i];
            akvirx -= (dx * hx + dy * hy + dz * hz);
            pvtt[1] -= dx * hx;
            pvtt[5] -= dy * hy;
            pvtt[9] -= dz * hz;
            pvtt[2] -= dx * hy;
            pvtt[4] -= dy * hx;
            pvtt[3] -= dx * hz;
            pvtt[7] -= dz * hx;
            pvtt[6] -= dy * hz;
            pvtt[8] -= dz * hy;
          }
        }
      }
      out->pi_ke_vir = 0.5 * akvirx;
    }
  }
  for (int i = 1; i <= 9; i++) {
    out->pvten[i - 1] = pvt[i];
    out->pvten_tot[i - 1] = pvtt[i];
  }
  free(tx);
  free(ty);
  free(tz);
  free(fxt);
  free(fyt);
  free(fzt);
  /* ghost forces (energy_control_final.c:174-181) */
  for (int ip = 0; ip < P; ip++)
    or_ghost_force(c, fx + (size_t)ip * N, fy + (size_t)ip * N, fz + (size_t)ip * N);
  /* PIMD mode forces and springs (energy_control_pimd.c:200-218,
   * pimd_estimators.c:31-113, transform_cnt.c:415-489 for normal modes,
   * transform_stg.c:31-160 for staging) */
  if ((flags & PMX_EVAL_PIMD) && (flags & PMX_EVAL_MODE_FORCES) && P > 1) {
    const int stag = c->par.pimd_typ == PMX_PIMD_STAGING;
    double *a = malloc(sizeof(double) * P), *b = malloc(sizeof(double) * P);
    double ake = 0.0;
    for (int d = 0; d < 3; d++) {
      double *pos = d == 0 ? x : (d == 1 ? y : z);
      double *frc = d == 0 ? fx : (d == 1 ? fy : fz);
      for (int i = 0; i < N; i++) {
        for (int ip = 0; ip < P; ip++) a[ip] = pos[(size_t)ip * N + i];
        if (stag)
          or_stag_pos_to_mode(P, a, b); /* b = modes */
        else
          or_nm_pos_to_mode(P, a, b);
        for (int ip = 0; ip < P; ip++) a[ip] = frc[(size_t)ip * N + i];
        double *fm = malloc(sizeof(double) * P);
        if (stag)
          or_stag_force_to_mode(P, a, fm);
        else
          or_nm_force_to_mode(P, a, fm);
        for (int ip = 0; ip < P; ip++) {
          double fk = c->prekf[i] * c->veig[ip + 1];
          ake += fk * b[ip] * b[ip];
          frc[(size_t)ip * N + i] = fm[ip] + (-2.0 * fk * b[ip]);
        }
        free(fm);
      }
    }
    out->kin_harm = ake;
    double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
    double tau = (OR_BOLTZ / T) / P;
    out->pi_ke_prim = 1.5 * ((double)(N - c->nghost)) / tau - ake;
    free(a);
    free(b);
  }
  free(xm);
  free(ym);
  free(zm);
  return PMX_OK;
}

/* ------------------------------------------------------------------------ */
double or_recip_energy(or_ctx *c, const double *x, const double *y, const double *z, int use_pme,
                       double *fx, double *fy, double *fz, double *pvten9) {
  int N = c->N;
  double *gx = calloc(N, sizeof(double)), *gy = calloc(N, sizeof(double)), *gz = calloc(N, sizeof(double));
  double pvt[10] = {0};
  double v = use_pme ? or_pme(c, x, y, z, gx, gy, gz, pvt) : or_ewald(c, x, y, z, gx, gy, gz, pvt);
  if (fx) {
    memcpy(fx, gx, sizeof(double) * N);
    memcpy(fy, gy, sizeof(double) * N);
    memcpy(fz, gz, sizeof(double) * N);
  }
  if (pvten9)
    for (int i = 0; i < 9; i++) pvten9[i] = pvt[i + 1];
  free(gx);
  free(gy);
  free(gz);
  return v;
}

int or_get_verlet_list(or_ctx *c, int *nter, int *jver_off, int *jter, int cap, int *nver) {
  if (!c->have_list) return PMX_ERR_STATE;
  *nver = c->nver;
  if (cap < c->nver) return PMX_ERR_RANGE;
  for (int i = 0; i < c->N; i++) {
    nter[i] = c->nter[i + 1];
    jver_off[i] = c->jver_off[i + 1];
  }
  for (int k = 0; k < c->nver; k++) jter[k] = c->jter[k] - 1;
  return PMX_OK;
}

int or_get_nktot(or_ctx *c, int *nktot, int *ngrid) {
  *nktot = c->nktot;
  for (int d = 0; d < 3; d++) ngrid[d] = c->ngrid[d];
  return PMX_OK;
}

int or_get_kvectors(or_ctx *c, int *ka, int *kb, int *kc, double *bw, int cap) {
  if (cap < c->nktot) return PMX_ERR_RANGE;
  for (int k = 0; k < c->nktot; k++) {
    ka[k] = c->kka[k];
    kb[k] = c->kkb[k];
    kc[k] = c->kkc[k];
    bw[k] = c->bweight[k];
  }
  return PMX_OK;
}

int or_get_spline(or_ctx *c, int which, int ipair, double *out, int cap) {
  int n = c->nsplin;
  if (cap < n) return PMX_ERR_RANGE;
  const double *src = NULL;
  if (which >= 0 && which <= 3 && (ipair < 0 || ipair >= c->npair)) return PMX_ERR_ARG;
  switch (which) {
    case 0: src = c->cv0 + (size_t)ipair * n; break;
    case 1: src = c->cdv0 + (size_t)ipair * n; break;
    case 2: src = c->cv0_c + (size_t)ipair * n; break;
    case 3: src = c->cdv0_c + (size_t)ipair * n; break;
    case 4: src = c->ecv0; break;
    case 5: src = c->ecdv0; break;
    default: return PMX_ERR_ARG;
  }
  memcpy(out, src, sizeof(double) * n);
  return PMX_OK;
}

/* scalars: [0] alp_ewd [1] ecut [2] vol [3] clong [4] clong_res
 * [5] self_erf [6] ecor rmin [7] ecor rmax [8] ecor dr [9] nktot
 * [10..12] ngrid [13..15] kmaxv [16] necor [17] nexcl [18] cutoff_max */
int or_get_setup_scalars(or_ctx *c, double *o) {
  o[0] = c->alp_ewd;
  o[1] = c->ecut;
  o[2] = c->vol;
  o[3] = c->clong;
  o[4] = c->clong_res;
  o[5] = c->self_erf;
  o[6] = c->e_rmin;
  o[7] = c->e_rmax;
  o[8] = c->e_dr;
  o[9] = c->nktot;
  for (int d = 0; d < 3; d++) {
    o[10 + d] = c->ngrid[d];
    o[13 + d] = c->kmaxv[d + 1];
  }
  o[16] = c->necor;
  o[17] = c->nexcl;
  o[18] = c->cutoff_max;
  return PMX_OK;
}

int or_get_bead_energies(or_ctx *c, double *out) {
  memcpy(out, c->bead_e, sizeof(double) * 4 * c->P);
  return PMX_OK;
}

/* intra_coefs.c:30-198 (harm :62-83, Morse :90-122) and :204-375 (harm
 * bend :242-277); inputs already in Hartree / Bohr / radians. */
void or_bond_coef_harm(double fk, double *c7) {
  for (int m = 0; m < 7; m++) c7[m] = 0.0;
  c7[2] = 0.5 * fk;
}

void or_bond_coef_morse(double d0, double al, double *c7) {
  double a2 = al * al, a3 = al * al * al, a4 = al * al * al * al, a5 = al * al * al * al * al,
         a6 = al * al * al * al * al * al;
  c7[0] = 0.0;
  c7[1] = 0.0;
  c7[2] = a2 * d0;
  c7[3] = -a3 * d0;
  c7[4] = 7.0 * a4 * d0 / 12.0;
  c7[5] = -a5 * d0 / 4.0;
  c7[6] = 62.0 * a6 * d0 / 720.0;
}

void or_bend_coef_harm(double fk, double th0, double *c7, double *s7) {
  static const double cn1 = 49.0 / 72.0, cn2 = 3.0 / 4.0, cn3 = 3.0 / 40.0, cn4 = 1.0 / 180.0;
  static const double cn5 = 3.0 / 2.0, cn6 = 3.0 / 5.0, cn7 = 1.0 / 10.0, cn8 = 4.0 / 15.0;
  static const double cn9 = 8.0 / 45.0, cn10 = 3.0 / 10.0, cn11 = 1.0 / 30.0;
  double c2 = cos(2.0 * th0), c4 = cos(4.0 * th0), c6 = cos(6.0 * th0);
  double s2 = sin(2.0 * th0), s4 = sin(4.0 * th0), s6 = sin(6.0 * th0);
  c7[0] = 0.5 * fk * (cn1 + cn2 * c2 + cn3 * c4 + cn4 * c6);
  c7[1] = 0.0;
  c7[2] = 0.5 * fk * (-cn5 * c2 - cn6 * c4 - cn7 * c6);
  c7[3] = 0.0;
  c7[4] = 0.5 * fk * (cn6 * c4 + cn8 * c6);
  c7[5] = 0.0;
  c7[6] = 0.5 * fk * (-cn9 * c6);
  s7[0] = 0.0;
  s7[1] = 0.0;
  s7[2] = 0.5 * fk * (-cn5 * s2 + cn10 * s4 - cn11 * s6);
  s7[3] = 0.0;
  s7[4] = 0.5 * fk * (-cn6 * s4 + cn9 * s6);
  s7[5] = 0.0;
  s7[6] = 0.5 * fk * (-cn9 * s6);
}

/* ------------------------------------------------------------------------
 * PIMD in the NVT ensemble: int_NVT_pimd (pimd/int_NVT_pimd.c:36-380) with
 * RESPA off, i.e. one pass of int_0_to_dt2_nvt_pimd (int_0_to_dt2_pimd.c:
 * 34-196), energy_control_pimd and int_dt2_to_dt_nvt_pimd
 * (int_dt2_to_dt_pimd.c:31-172) per step.  Velocities are mode velocities
 * (bead-major [P*N]); positions stay bead Cartesian between steps.
 * Thermostats as set_atm_NHC (set_atm_NHC.c:330-790) sets them for
 * mol_opt_nhc = global and path-integral beads: mode 1 (the centroid, or
 * bead 1 with staging) on one chain for every real site with the physical
 * masses, gkt = 3 N_real (no constraints) times kT; modes 2..P each degree
 * of freedom of each real site on its own chain with the mode masses.
 * Ghost sites sit on the null thermostat.
 *
 * NHC state layout: class chain v[len], x[len]; then for mode ip = 1..P-1
 * (0-based) and chain k = 3 i + d: v[len], x[len] at
 * 2 len + ((ip - 1) 3N + k) 2 len.
 */
typedef struct {
  int len, nres, nyosh;
  double wdti2[10], wdti4[10], wdti8[10];
  double cmass[8], cgkt[8];  /* class chain */
  double bmass[8], bgkt[8];  /* bead chains (the same for every chain) */
} or_nhc;

/* set_yosh, int_utils.c:244-318 */
static int or_set_yosh(int nyosh, double dti, or_nhc *h) {
  double w[10] = {0};
  if (nyosh == 1) {
    w[1] = 1.0;
  } else if (nyosh == 3) {
    double onethird = 1.0 / 3.0, temp = pow(2.0, onethird);
    w[1] = 1.0 / (2.0 - temp);
    w[2] = -temp / (2.0 - temp);
    w[3] = 1.0 / (2.0 - temp);
  } else if (nyosh == 5) {
    double onethird = 1.0 / 3.0, temp = pow(4.0, onethird), p2 = 1.0 / (4.0 - temp);
    w[1] = p2;
    w[2] = p2;
    w[3] = 1.0 - 4.0 * p2;
    w[4] = p2;
    w[5] = p2;
  } else {
    return PMX_ERR_NOTIMPL;
  }
  for (int i = 1; i <= nyosh; i++) {
    double wd = w[i] * dti;
    h->wdti2[i] = wd / 2.0;
    h->wdti4[i] = wd / 4.0;
    h->wdti8[i] = wd / 8.0;
  }
  return PMX_OK;
}

/* one thermostat chain of apply_NHC_par (md/int_NVT.c:144-330): kin is the
 * sum of m v^2 over the chain's degrees of freedom; returns the factor the
 * velocities are scaled by */
static double or_nhc_chain(const or_nhc *h, const double *mass, const double *gkt, double *v, double *x,
                           double kin) {
  int L = h->len;
  double f[8];
  double sc = 1.0;
  f[1] = (kin - gkt[1]) / mass[1];
  for (int ic = 1; ic <= L - 1; ic++) f[ic + 1] = (mass[ic] * v[ic] * v[ic] - gkt[ic + 1]) / mass[ic + 1];
  for (int ir = 1; ir <= h->nres; ir++) {
    for (int iy = 1; iy <= h->nyosh; iy++) {
      v[L] += f[L] * h->wdti4[iy];
      for (int ic = 1; ic <= L - 1; ic++) {
        double aa = exp(-h->wdti8[iy] * v[L + 1 - ic]);
        v[L - ic] = v[L - ic] * aa * aa + h->wdti4[iy] * f[L - ic] * aa;
      }
      double aa = exp(-h->wdti2[iy] * v[1]);
      sc *= aa;
      kin *= aa * aa;
      for (int ic = 1; ic <= L; ic++) x[ic] += v[ic] * h->wdti2[iy];
      f[1] = (kin - gkt[1]) / mass[1];
      for (int ic = 1; ic <= L - 1; ic++) {
        double a2 = exp(-h->wdti8[iy] * v[ic + 1]);
        v[ic] = v[ic] * a2 * a2 + h->wdti4[iy] * f[ic] * a2;
        f[ic + 1] = (mass[ic] * v[ic] * v[ic] - gkt[ic + 1]) / mass[ic + 1];
      }
      v[L] += f[L] * h->wdti4[iy];
    }
  }
  return sc;
}

/* veig_ip of the mode masses of clatoms_pos[ip].mass (transform_cnt.c:
 * 175-233), with veig_1 = 1 */
static double or_mode_veig(const or_ctx *c, int ip) {
  int P = c->P;
  double dP = (double)P;
  double v;
  if (ip == 1) return 1.0;
  if (c->par.pimd_typ == PMX_PIMD_STAGING) {
    v = ((double)(ip)) / ((double)(ip - 1));
  } else if (ip == P) {
    v = 4.0 * dP;
  } else {
    int k = ip / 2 + 1; /* veig[2k-2] = veig[2k-1] for k = 2..P/2 */
    v = (1.0 - cos(2.0 * M_PI / dP * (double)(k - 1))) * 4.0 * dP;
  }
  return v;
}

static void or_apply_nhc(const or_ctx *c, const or_nhc *h, const double *mm, double *vx, double *vy, double *vz,
                         double *st, const int *ghost) {
  int N = c->N, P = c->P, L = h->len;
  /* mode 1: one chain over the real sites (x, then y, then z, in site order) */
  double kin = 0.0;
  double *vv[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < N; i++)
      if (!ghost[i]) kin += c->mass[i] * vv[d][i] * vv[d][i];
  double sc = or_nhc_chain(h, h->cmass, h->cgkt, st - 1, st + L - 1, kin);
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < N; i++)
      if (!ghost[i]) vv[d][i] *= sc;
  /* modes 2..P: one chain per degree of freedom */
  for (int ip = 1; ip < P; ip++)
    for (int d = 0; d < 3; d++)
      for (int i = 0; i < N; i++) {
        if (ghost[i]) continue;
        size_t t = (size_t)ip * N + i;
        double m = mm[t];
        double *s = st + 2 * L + ((size_t)(ip - 1) * 3 * N + 3 * i + d) * 2 * L;
        double sk = or_nhc_chain(h, h->bmass, h->bgkt, s - 1, s + L - 1, m * vv[d][t] * vv[d][t]);
        vv[d][t] *= sk;
      }
}

int or_pimd_nvt(or_ctx *c, const pmx_nhc_params *np, int nsteps, double *x, double *y, double *z, double *vx,
                double *vy, double *vz, double *nhc, pmx_energies *out, double *econs) {
  int N = c->N, P = c->P;
  if (P < 2 || !np || np->len_nhc < 1 || np->len_nhc > 7 || np->nres_nhc < 1) return PMX_ERR_ARG;
  double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
  or_nhc h;
  memset(&h, 0, sizeof h);
  h.len = np->len_nhc;
  h.nres = np->nres_nhc;
  h.nyosh = np->nyosh_nhc;
  if (or_set_yosh(h.nyosh, np->dt / (double)h.nres, &h)) return PMX_ERR_NOTIMPL;
  int *ghost = calloc(N, sizeof(int));
  int nreal = N;
  for (int g = 0; g < c->nghost; g++) {
    ghost[c->gatom[g]] = 1;
    nreal--;
  }
  /* set_atm_NHC.c:355-370 (global), 510-547 (NHC masses), 715-755 (beads) */
  double text = T, tau = np->tau_nhc;
  double m1 = tau * tau * text, g1 = 3.0 * (double)nreal;
  for (int j = 2; j <= h.len; j++) {
    h.cgkt[j] = text / OR_BOLTZ;
    h.cmass[j] = m1 / OR_BOLTZ;
  }
  h.cmass[1] = m1 * g1 / OR_BOLTZ;
  h.cgkt[1] = g1 * text / OR_BOLTZ;
  double pre = (OR_BOLTZ * OR_BOLTZ * np->gamma_adb * np->gamma_adb) / ((double)P);
  double bm1 = pre / (text * text);
  for (int j = 2; j <= h.len; j++) {
    h.bgkt[j] = text / OR_BOLTZ;
    h.bmass[j] = bm1 * text / OR_BOLTZ;
  }
  h.bmass[1] = bm1 * text / OR_BOLTZ;
  h.bgkt[1] = text / OR_BOLTZ;
  /* mode masses */
  size_t NP = (size_t)N * P;
  double *mm = malloc(sizeof(double) * NP);
  for (int ip = 0; ip < P; ip++)
    for (int i = 0; i < N; i++) {
      /* m veig for mode 1, m veig gamma^2 for the others */
      double m = c->mass[i] * or_mode_veig(c, ip + 1);
      mm[(size_t)ip * N + i] = ip == 0 ? m : m * np->gamma_adb * np->gamma_adb;
    }
  double *fx = calloc(NP, sizeof(double)), *fy = calloc(NP, sizeof(double)), *fz = calloc(NP, sizeof(double));
  double *xm = malloc(sizeof(double) * N), *ym = malloc(sizeof(double) * N), *zm = malloc(sizeof(double) * N);
  double *a = malloc(sizeof(double) * P), *b = malloc(sizeof(double) * P);
  const int stag = c->par.pimd_typ == PMX_PIMD_STAGING;
  unsigned flags = PMX_EVAL_PIMD | PMX_EVAL_MODE_FORCES;
  int rc = or_energy_control(c, x, y, z, NULL, NULL, NULL, flags, fx, fy, fz, out);
  double dt = np->dt, dt2 = dt / 2.0;
  for (int step = 0; step < nsteps && !rc; step++) {
    or_apply_nhc(c, &h, mm, vx, vy, vz, nhc, ghost);
    for (size_t t = 0; t < NP; t++) {
      vx[t] += fx[t] * dt2 / mm[t];
      vy[t] += fy[t] * dt2 / mm[t];
      vz[t] += fz[t] * dt2 / mm[t];
    }
    /* positions to modes, drift, back (mode 1 becomes xmod) */
    double *pos[3] = {x, y, z}, *vel[3] = {vx, vy, vz}, *md[3] = {xm, ym, zm};
    for (int d = 0; d < 3; d++)
      for (int i = 0; i < N; i++) {
        for (int ip = 0; ip < P; ip++) a[ip] = pos[d][(size_t)ip * N + i];
        if (stag)
          or_stag_pos_to_mode(P, a, b);
        else
          or_nm_pos_to_mode(P, a, b);
        for (int ip = 0; ip < P; ip++) b[ip] += vel[d][(size_t)ip * N + i] * dt;
        md[d][i] = b[0];
        if (stag)
          or_stag_mode_to_pos(P, b, a);
        else
          or_nm_mode_to_pos(P, b, a);
        for (int ip = 0; ip < P; ip++) pos[d][(size_t)ip * N + i] = a[ip];
      }
    rc = or_energy_control(c, x, y, z, xm, ym, zm, flags, fx, fy, fz, out);
    for (size_t t = 0; t < NP; t++) {
      vx[t] += fx[t] * dt2 / mm[t];
      vy[t] += fy[t] * dt2 / mm[t];
      vz[t] += fz[t] * dt2 / mm[t];
    }
    or_apply_nhc(c, &h, mm, vx, vy, vz, nhc, ghost);
  }
  if (econs && !rc) {
    /* get_tvten_pimd kinetic energy and the chains' energies */
    double ke = 0.0, en = 0.0;
    for (size_t t = 0; t < NP; t++) ke += 0.5 * mm[t] * (vx[t] * vx[t] + vy[t] * vy[t] + vz[t] * vz[t]);
    int L = h.len;
    for (int j = 1; j <= L; j++) en += 0.5 * h.cmass[j] * nhc[j - 1] * nhc[j - 1] + h.cgkt[j] * nhc[L + j - 1];
    for (int ip = 1; ip < P; ip++)
      for (int i = 0; i < N; i++) {
        if (ghost[i]) continue;
        for (int d = 0; d < 3; d++) {
          const double *s = nhc + 2 * L + ((size_t)(ip - 1) * 3 * N + 3 * i + d) * 2 * L;
          for (int j = 1; j <= L; j++) en += 0.5 * h.bmass[j] * s[j - 1] * s[j - 1] + h.bgkt[j] * s[L + j - 1];
        }
      }
    econs[0] = ke;
    econs[1] = en;
    econs[2] = out->vintert + out->vintrat + out->kin_harm;
    econs[3] = econs[0] + econs[1] + econs[2];
  }
  free(ghost);
  free(mm);
  free(fx);
  free(fy);
  free(fz);
  free(xm);
  free(ym);
  free(zm);
  free(a);
  free(b);
  return rc;
}
