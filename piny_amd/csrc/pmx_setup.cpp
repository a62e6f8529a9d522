// Parse-time setup of the force library, run once per context on the host.
//
// Restates the parts of PINY's parse() that the force call depends on:
//   * spline tables      control_inter_params.c:643-772, spline_fit.c:28-398
//   * LJ tail constant   get_clong.c:27-114
//   * alpha, cutoffs,
//     k-vectors, grid    control_set_cp_ewald.c:157-163, set_cp_ewald.c:125-1370
//   * exclusions, ecor   set_exclude.c:30-779
//   * spring constants   transform_cnt.c:140-201
// Compiled with -ffp-contract=off so the tables carry the reference's bits.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "pmx_internal.h"

namespace pmx {
namespace {

int nint_piny(double x) { return (int)((x) >= 0.0 ? ((x) + 0.5) : ((x)-0.5)); }



double erfc_poly(double tt) {
  const double *e = kErfE;
  return ((((((((e[9] * tt + e[8]) * tt + e[7]) * tt + e[6]) * tt + e[5]) * tt + e[4]) * tt + e[3]) * tt +
           e[2]) * tt + e[1]) * tt;
}

double derfc_poly(double tt) {
  const double *e = kErfE;
  double d[10];
  for (int k = 1; k <= 9; k++) d[k] = ((double)k) * e[k];
  return ((((((((d[9] * tt + d[8]) * tt + d[7]) * tt + d[6]) * tt + d[5]) * tt + d[4]) * tt + d[3]) * tt +
           d[2]) * tt + d[1]) * tt * tt;
}

// One spline table pair (V, -(1/r) dV/dr) on PINY's grid r_i = dr (i-3) + rmin.
// itype: 1 erfc-Coulomb with unit charge product, 2 LJ, 4 null.
void spline_table(const Setup &s, const pmx_params &par, int itype, double rmin, double rmax,
                  double sig, double eps, std::vector<double> &v, std::vector<double> &dv,
                  double &dr_out, double &dri_out, double *shift_out = nullptr) {
  const int n = s.nsplin;
  const double rheal = par.rheal;
  v.assign(n, 0.0);
  dv.assign(n, 0.0);
  const double dr = (rmax - rmin) / (double)(n - 5);
  const double dri = 1.0 / dr;
  dr_out = dr;
  dri_out = dri;
  std::vector<double> r(n), sw(n, 1.0), dsw(n, 0.0);
  for (int i = 1; i <= n; i++) r[i - 1] = dr * ((double)(i - 3)) + rmin;
  if (par.ishift == PMX_SHIFT_SWIT) {
    for (int i = 0; i < n; i++) {
      double rp = (r[i] - rmax + rheal) / rheal;
      rp = std::min(rp, 1.0);
      rp = std::max(rp, 0.0);
      sw[i] = 1.0 + rp * rp * (2.0 * rp - 3.0);
      dsw[i] = -6.0 * (rp / rheal) * (rp - 1.0) / r[i];
    }
  }
  if (itype == PMX_INTER_LJ) {
    const double sig6 = sig * sig * sig * sig * sig * sig;
    for (int i = 0; i < n; i++) {
      const double r2 = r[i] * r[i];
      const double sig6r = sig6 / (r2 * r2 * r2);
      const double e_pair = (eps * 4. * sig6r * (sig6r - 1.));
      const double de_pair = (eps * 24. * sig6r * (sig6r * 2. - 1.) / r2);
      v[i] = e_pair * sw[i];
      dv[i] = de_pair * sw[i] + e_pair * dsw[i];
    }
  } else if (itype == 1) {
    const double alp = s.alp_ewd, qij = 1.0;
    const double two_alpha2 = 2.0 * alp * alp, p_alpha = kErfP * alp;
    for (int i = 0; i < n; i++) {
      const double r2 = r[i] * r[i];
      const double ralp = r[i] * alp;
      const double eee = exp(-ralp * ralp);
      const double tt = 1.0 / (1.0 + kErfP * ralp);
      const double screen = erfc_poly(tt) * eee;
      const double dscreen = derfc_poly(tt) * eee * p_alpha + two_alpha2 * screen * r[i];
      const double e_pair = qij * screen / r[i];
      const double de_pair = (screen / r2 + dscreen / r[i]) * qij / r[i];
      v[i] = e_pair * sw[i];
      dv[i] = de_pair * sw[i] + e_pair * dsw[i];
    }
  }
  if (shift_out) *shift_out = 0.0;
  if (par.ishift == PMX_SHIFT_SHIFT) {
    const double shift = v[(n - 2) - 1];
    for (int i = 0; i < n; i++) v[i] -= shift;
    if (shift_out) *shift_out = shift;
  }
}

void radixme(int *k1p, int *k2p, int *k3p) {
  static const int krad[50] = {4,   8,   12,  16,  20,  24,  32,  36,  40,  48,  60,  64,  72,
                               80,  96,  100, 108, 120, 128, 144, 160, 180, 192, 200, 216, 240,
                               256, 288, 300, 320, 324, 360, 384, 400, 432, 480, 500, 512, 540,
                               576, 600, 640, 648, 720, 768, 800, 864, 900, 960, 972};
  int *kp[3] = {k1p, k2p, k3p};
  for (int d = 0; d < 3; d++) {
    int k = (*kp[d] + 1) << 2;
    for (int i = 1; i <= 50; i++) {
      if (krad[i - 1] <= k) continue;
      const int lo = std::max(i - 1, 1);
      if (k > krad[lo - 1] && k < krad[i - 1]) {
        const int dhi = std::abs(k - krad[i - 1]);
        const int dlo = std::abs(k - krad[lo - 1]);
        k = (dhi < dlo) ? krad[i - 1] : krad[lo - 1];
      }
    }
    *kp[d] = k / 4 - 1;
  }
}

double row_norm(const double *hi, int a, int b, int c) {
  const double d1 = hi[a], d2 = hi[b], d3 = hi[c];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

double ktry(const double *hm, double ga, double gb, double gc) {
  const double tpi = 2.0 * M_PI;
  const double xk = (ga * hm[1] + gb * hm[2] + gc * hm[3]) * tpi;
  const double yk = (ga * hm[4] + gb * hm[5] + gc * hm[6]) * tpi;
  const double zk = (ga * hm[7] + gb * hm[8] + gc * hm[9]) * tpi;
  return (xk * xk + yk * yk + zk * zk) * .5;
}

// setkvec3d, set_cp_ewald.c:428-573: spherically truncated half space.
void kvectors(Setup &s) {
  const double *hm = s.hmati;
  const double ec4 = s.ecut * 4.;
  s.ka.clear();
  s.kb.clear();
  s.kc.clear();
  int i;
  for (int ka = 0; ka <= s.kmaxv[1]; ++ka) {
    const double ga = (double)ka;
    int kbmin = (ka == 0) ? 0 : -s.kmaxv[2];
    for (i = kbmin; i <= 0; ++i)
      if (ktry(hm, ga, (double)i, 0.0) <= ec4) break;
    kbmin = i;
    for (i = 1; i <= s.kmaxv[2]; ++i)
      if (ktry(hm, ga, (double)i, 0.0) > ec4) break;
    const int kbmax = i - 1;
    for (int kb = kbmin; kb <= kbmax; ++kb) {
      const double gb = (double)kb;
      int kcmin = (ka == 0 && kb == 0) ? 1 : -s.kmaxv[3];
      for (i = kcmin; i <= 0; ++i)
        if (ktry(hm, ga, gb, (double)i) <= ec4) break;
      kcmin = i;
      for (i = 1; i <= s.kmaxv[3]; ++i)
        if (ktry(hm, ga, gb, (double)i) > ec4) break;
      const int kcmax = i - 1;
      for (int kc = kcmin; kc <= kcmax; ++kc) {
        s.ka.push_back(ka);
        s.kb.push_back(kb);
        s.kc.push_back(kc);
      }
    }
  }
  s.nktot = (int)s.ka.size();
}

// get_bspline_wght1d, set_cp_ewald.c:1288-1370
void bspline_weight1d(int n_interp, int ngrid, std::vector<double> &bw) {
  double rn[MAX_INTERP + 2], rn1[MAX_INTERP + 2], mn[MAX_INTERP + 2], uk[MAX_INTERP + 2];
  const double grid = (double)ngrid;
  for (int j = 1; j <= n_interp; j++) {
    rn[j] = (double)(j);
    if (j > 1) rn1[j] = 1.0 / ((double)(j - 1));
  }
  rn1[1] = 0.0;
  mn[1] = 1.0;
  uk[1] = 1.0;
  for (int k = 2; k <= n_interp; k++) {
    uk[k] = (double)(k);
    mn[k] = 0.0;
  }
  for (int n = 3; n <= n_interp; n++) {
    for (int k = n; k >= 2; k--) {
      const double t = (uk[k] * mn[k] + (rn[n] - uk[k]) * mn[k - 1]) * rn1[n];
      mn[k] = t;
    }
    mn[1] = uk[1] * mn[1] * rn1[n];
  }
  std::vector<double> bre(ngrid + 1, 0.0), bim(ngrid + 1, 0.0);
  double phase_k = 2.0 * M_PI / grid;
  for (int m = 1; m <= ngrid; m++) {
    for (int k = 1; k <= n_interp - 1; k++) {
      const double arg = phase_k * ((double)((k - 1) * (m - 1)));
      bre[m] += cos(arg) * mn[k];
      bim[m] += sin(arg) * mn[k];
    }
  }
  bw.assign(ngrid + 1, 0.0);
  phase_k = 2.0 * M_PI * ((double)(n_interp - 1)) / grid;
  for (int m = 1; m <= ngrid; m++) {
    const double arg = phase_k * ((double)(m - 1));
    const double bnr = cos(arg), bni = sin(arg);
    const double bdr = bre[m], bdi = bim[m];
    const double denom = bdr * bdr + bdi * bdi;
    const double tr = (bnr * bdr + bni * bdi) / denom;
    const double ti = (bni * bdr - bnr * bdi) / denom;
    bw[m] = tr * tr + ti * ti;
  }
}

void gauss_legendre(int n, double *node, double *weight) {
  // Newton's iteration on P_n in quadruple precision (__float128), each
  // node and weight rounded to double once: the correctly rounded values,
  // which are also what 33-digit decimal tables of them parse to
  for (int i = 0; i < n; i++) {
    __float128 x = cos(M_PI * (i + 0.75) / (n + 0.5)), p0 = 1, p1 = x, dp = 0;
    for (int it = 0; it < 8; it++) {  // quadratic convergence from a double start
      p0 = 1;
      p1 = x;
      for (int k = 2; k <= n; k++) {
        const __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1);
      x -= p1 / dp;
    }
    p0 = 1;
    p1 = x;
    for (int k = 2; k <= n; k++) {
      const __float128 p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1);
    node[n - 1 - i] = (double)x;
    weight[n - 1 - i] = (double)(2 / ((1 - x * x) * dp * dp));
  }
}

}  // namespace

void bond_coef_harm(double fk, double *c7) {
  for (int m = 0; m < 7; m++) c7[m] = 0.0;
  c7[2] = 0.5 * fk;
}

void bond_coef_morse(double d0, double al, double *c7) {
  const double a2 = al * al, a3 = al * al * al, a4 = al * al * al * al;
  const double a5 = al * al * al * al * al, a6 = al * al * al * al * al * al;
  c7[0] = 0.0;
  c7[1] = 0.0;
  c7[2] = a2 * d0;
  c7[3] = -a3 * d0;
  c7[4] = 7.0 * a4 * d0 / 12.0;
  c7[5] = -a5 * d0 / 4.0;
  c7[6] = 62.0 * a6 * d0 / 720.0;
}

void bend_coef_harm(double fk, double th0, double *c7, double *s7) {
  const double cn1 = 49.0 / 72.0, cn2 = 3.0 / 4.0, cn3 = 3.0 / 40.0, cn4 = 1.0 / 180.0;
  const double cn5 = 3.0 / 2.0, cn6 = 3.0 / 5.0, cn7 = 1.0 / 10.0, cn8 = 4.0 / 15.0;
  const double cn9 = 8.0 / 45.0, cn10 = 3.0 / 10.0, cn11 = 1.0 / 30.0;
  const double c2 = cos(2.0 * th0), c4 = cos(4.0 * th0), c6 = cos(6.0 * th0);
  const double s2 = sin(2.0 * th0), s4 = sin(4.0 * th0), s6 = sin(6.0 * th0);
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

std::string build_setup(const pmx_system *sys, const pmx_params *par, const double *ix,
                        const double *iy, const double *iz, Setup &s) {
  char msg[256];
  s.N = sys->natm;
  s.P = par->pi_beads;
  s.ntype = sys->ntype;
  s.npair = s.ntype * (s.ntype + 1) / 2;
  s.nsplin = par->nsplin;
  if (s.ntype > MAX_TYPES) return "too many atom types (max 8)";
  if (par->n_interp < 3 || par->n_interp > MAX_INTERP) return "ewald_interp_pme out of range";
  // --- cell: gethinv (mathlib.c:353-384), cubic flag read_hmat.c:498-505
  double *h = s.hmat, *hi = s.hmati;
  for (int i = 1; i <= 9; i++) h[i] = par->hmat[i - 1];
  const double vol = (h[1] * (h[5] * h[9] - h[8] * h[6]) + h[4] * (h[8] * h[3] - h[2] * h[9]) +
                      h[7] * (h[2] * h[6] - h[5] * h[3]));
  if (!(vol > 0.0)) return "cell matrix has no positive volume";
  s.vol = vol;
  hi[1] = (h[5] * h[9] - h[8] * h[6]) / vol;
  hi[5] = (h[1] * h[9] - h[7] * h[3]) / vol;
  hi[9] = (h[1] * h[5] - h[4] * h[2]) / vol;
  hi[4] = (h[7] * h[6] - h[4] * h[9]) / vol;
  hi[2] = (h[3] * h[8] - h[2] * h[9]) / vol;
  hi[7] = (h[4] * h[8] - h[7] * h[5]) / vol;
  hi[3] = (h[2] * h[6] - h[3] * h[5]) / vol;
  hi[8] = (h[7] * h[2] - h[8] * h[1]) / vol;
  hi[6] = (h[3] * h[4] - h[6] * h[1]) / vol;
  s.cubic = (h[4] == 0.0 && h[7] == 0.0 && h[2] == 0.0 && h[8] == 0.0 && h[3] == 0.0 && h[6] == 0.0);
  // --- topology indices must address existing sites / types
  {
    auto bad = [&](int v, int n) { return v < 0 || v >= n; };
    for (int i = 0; i < s.N; i++)
      if (bad(sys->atm_type[i], s.ntype)) return "atom type index out of range";
    for (int b = 0; b < sys->nbond; b++)
      if (bad(sys->bond_j1[b], s.N) || bad(sys->bond_j2[b], s.N) || bad(sys->bond_type[b], sys->nbond_type))
        return "bond index out of range";
    for (int b = 0; b < sys->nbend; b++)
      if (bad(sys->bend_j1[b], s.N) || bad(sys->bend_j2[b], s.N) || bad(sys->bend_j3[b], s.N) ||
          bad(sys->bend_type[b], sys->nbend_type))
        return "bend index out of range";
    for (int g = 0; g < sys->nghost; g++) {
      if (bad(sys->ghost_atom[g], s.N) || sys->ghost_ncomp[g] < 1 || sys->ghost_ncomp[g] > 4)
        return "ghost definition out of range";
      for (int k = 0; k < sys->ghost_ncomp[g]; k++)
        if (bad(sys->ghost_comp[4 * g + k], s.N)) return "ghost constituent out of range";
    }
  }
  // --- charged sites
  s.ichrg.clear();
  s.qsq = 0.0;
  s.qsum = 0.0;
  for (int i = 0; i < s.N; i++) {
    if (sys->q[i] != 0.0) s.ichrg.push_back(i);
    s.qsq += sys->q[i] * sys->q[i];
    s.qsum += sys->q[i];
  }
  // --- alpha and Ewald cutoff (control_set_cp_ewald.c:159-163, calc_cutoff)
  const double side = pow(s.vol, (1.0 / 3.0));
  s.alp_ewd = par->ewald_alpha / side;
  {
    const double rvol23 = pow(s.vol, -(2. / 3.));
    const int kmax = par->kmax_ewd;
    s.ecut = M_PI * .5 * M_PI * (double)(kmax * kmax) * rvol23;
    const double t1 = sqrt(s.ecut * .5) / (M_PI * row_norm(hi, 1, 4, 7));
    const double t2 = sqrt(s.ecut * .5) / (M_PI * row_norm(hi, 2, 5, 8));
    const double t3 = sqrt(s.ecut * .5) / (M_PI * row_norm(hi, 3, 6, 9));
    s.kmaxv[1] = (int)(2.0 * t1);
    s.kmaxv[2] = (int)(2.0 * t2);
    s.kmaxv[3] = (int)(2.0 * t3);
  }
  kvectors(s);
  // --- PME grid (set_pme_grid) and B-spline weights (set_pme_wght)
  s.bweight.assign(s.nktot, 1.0);
  if (par->pme_on) {
    const double rvol23 = pow(s.vol, -(2. / 3.));
    const int kp = par->kmax_pme;
    const double ecut_pme = M_PI * .5 * M_PI * (double)(kp * kp) * rvol23;
    const double ecut = std::max(ecut_pme, s.ecut);
    double t[3];
    t[0] = sqrt(ecut * .5) / (M_PI * row_norm(hi, 1, 4, 7));
    t[1] = sqrt(ecut * .5) / (M_PI * row_norm(hi, 2, 5, 8));
    t[2] = sqrt(ecut * .5) / (M_PI * row_norm(hi, 3, 6, 9));
    int k[3];
    for (;;) {
      for (int d = 0; d < 3; d++) k[d] = nint_piny(t[d]);
      radixme(&k[0], &k[1], &k[2]);
      for (int d = 0; d < 3; d++) s.ngrid[d] = 4 * (k[d] + 1);
      bool go = false;
      for (int d = 0; d < 3; d++) {
        if (s.ngrid[d] < 2 * (s.kmaxv[d + 1] + 2)) {
          t[d] += 1.0;
          go = true;
        }
      }
      if (!go) break;
    }
    for (int d = 0; d < 3; d++)
      if (par->n_interp > s.ngrid[d]) return "PME n_interp larger than the grid";
    std::vector<double> bw[3];
    for (int d = 0; d < 3; d++) bspline_weight1d(par->n_interp, s.ngrid[d], bw[d]);
    for (int i = 0; i < s.nktot; i++) {
      const int kap = s.ka[i] < 0 ? s.ka[i] + s.ngrid[0] + 1 : s.ka[i] + 1;
      const int kbp = s.kb[i] < 0 ? s.kb[i] + s.ngrid[1] + 1 : s.kb[i] + 1;
      const int kcp = s.kc[i] < 0 ? s.kc[i] + s.ngrid[2] + 1 : s.kc[i] + 1;
      s.bweight[i] = bw[0][kap] * bw[1][kbp] * bw[2][kcp];
    }
  }
  // --- spline tables per type pair (set_inter_splin)
  s.label.assign(s.npair, PMX_INTER_NULL);
  s.rmin.assign(s.npair, 0);
  s.dr.assign(s.npair, 0);
  s.dri.assign(s.npair, 0);
  s.cutoff.assign(s.npair, 0);
  s.cutoff_res.assign(s.npair, 0);
  s.cutskin2.assign(s.npair, 0);
  s.tab.assign((size_t)s.npair * s.nsplin * 4, 0.0);
  s.lj_eps.assign(s.npair, 0.0);
  s.lj_sig6.assign(s.npair, 0.0);
  s.lj_sig.assign(s.npair, 0.0);
  s.vshift_lj.assign(s.npair, 0.0);
  s.vshift_c.assign(s.npair, 0.0);
  s.cutoff_max = 0.0;
  for (int t = 0; t < s.npair; t++) {
    const int lab = sys->inter_label[t];
    if (lab != PMX_INTER_LJ && lab != PMX_INTER_NULL) {
      snprintf(msg, sizeof msg, "type pair %d: only Lennard-Jones (2) and null (4) are supported", t);
      return msg;
    }
    s.label[t] = lab;
    s.rmin[t] = sys->inter_cutti[t];
    s.cutoff[t] = sys->inter_cutoff[t];
    s.cutoff_res[t] = sys->inter_cutoff_res[t];
    s.cutoff_max = std::max(s.cutoff[t], s.cutoff_max);
    std::vector<double> v, dv, vc, dvc;
    spline_table(s, *par, lab, s.rmin[t], s.cutoff[t], sys->inter_sig[t], sys->inter_eps[t], v, dv,
                 s.dr[t], s.dri[t], &s.vshift_lj[t]);
    double drt, drit;
    if (!s.ichrg.empty())
      spline_table(s, *par, 1, s.rmin[t], s.cutoff[t], 1.0, 0.0, vc, dvc, drt, drit, &s.vshift_c[t]);
    else
      vc.assign(s.nsplin, 0.0), dvc.assign(s.nsplin, 0.0);
    const double sg = sys->inter_sig[t];
    s.lj_eps[t] = sys->inter_eps[t];
    s.lj_sig6[t] = sg * sg * sg * sg * sg * sg;
    s.lj_sig[t] = sg;
    for (int i = 0; i < s.nsplin; i++) {
      double *e = &s.tab[((size_t)t * s.nsplin + i) * 4];
      e[0] = v[i];
      e[1] = dv[i];
      e[2] = vc[i];
      e[3] = dvc[i];
    }
  }
  // --- LJ long-range constant (get_clong.c:27-114)
  {
    std::vector<long> num(s.ntype, 0);
    for (int i = 0; i < s.N; i++) num[sys->atm_type[i]]++;
    std::vector<double> clongm(s.npair, 0.0);
    for (int i = 0; i < s.ntype; i++) {
      for (int j = 0; j < i; j++) clongm[pair_index(s.ntype, j, i)] = (double)(num[i] * num[j]);
      clongm[pair_index(s.ntype, i, i)] = (double)((num[i] * (num[i] - 1)) / 2);
    }
    const double fpi3 = 4.0 * M_PI / 3.0, fpi = 4.0 * M_PI, rheal = par->rheal;
    const int iswit = (par->ishift == PMX_SHIFT_SWIT);
    s.clong = 0.0;
    s.clong_res = 0.0;
    s.clong_u.assign(s.npair, 0.0);
    for (int t = 0; t < s.npair; t++) {
      double c6m = 0.0, c6u = 0.0;
      if (s.label[t] == PMX_INTER_LJ) {
        const double sg = sys->inter_sig[t];
        c6m = (4.0 * sys->inter_eps[t]) * (sg * sg * sg * sg * sg * sg);
        c6u = 4.0 * (sg * sg * sg * sg * sg * sg);
      }
      const double rc = s.cutoff[t], rr = s.cutoff_res[t];
      const double rmin = (rc < rr ? rc : rr);
      if (iswit == 0) {
        s.clong_u[t] = -fpi3 * c6u * clongm[t] / pow(rc, 3.0);
        s.clong -= fpi3 * c6m * clongm[t] / pow(rc, 3.0);
        s.clong_res -= fpi * c6m * clongm[t] *
                       ((1.0 / (rheal * rheal)) * (1.0 / rmin + 1.0 / (rmin - rheal)) +
                        (2.0 / (rheal * rheal * rheal)) * log((rmin - rheal) / rmin));
      } else if ((c6m > 0) && (rc - rheal > 0)) {
        const double fact = (1.0 / (rheal * rheal)) * (1.0 / rc + 1.0 / (rc - rheal)) +
                            (2.0 / (rheal * rheal * rheal)) * log((rc - rheal) / rc);
        s.clong_u[t] = -(fpi * c6u * clongm[t] * fact);
        s.clong -= (fpi * c6m * clongm[t] * fact);
        const double fres = (1.0 / (rheal * rheal)) * (1.0 / rmin + 1.0 / (rmin - rheal)) +
                            (2.0 / (rheal * rheal * rheal)) * log((rmin - rheal) / rmin);
        s.clong_res -= fpi * c6m * clongm[t] * fres;
      }
    }
  }
  // --- exclusions and Ewald corrections (set_exclude.c:30-545)
  {
    std::vector<std::pair<int, int>> pr;
    for (int b = 0; b < sys->nbond; b++) {
      const int i = sys->bond_j1[b], j = sys->bond_j2[b];
      pr.emplace_back(std::max(i, j), std::min(i, j));
    }
    for (int b = 0; b < sys->nbend; b++) {
      const int i = sys->bend_j1[b], j = sys->bend_j3[b];
      pr.emplace_back(std::max(i, j), std::min(i, j));
    }
    for (int g = 0; g < sys->nghost; g++) {
      if (sys->ghost_ncomp[g] < 1 || sys->ghost_ncomp[g] > 4) return "ghost with more than 4 constituents";
      for (int k = 0; k < sys->ghost_ncomp[g]; k++) {
        const int i = sys->ghost_atom[g], j = sys->ghost_comp[4 * g + k];
        pr.emplace_back(std::max(i, j), std::min(i, j));
      }
    }
    std::sort(pr.begin(), pr.end());
    pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
    for (auto &p : pr)
      if (p.first == p.second || p.first < 0 || p.first >= s.N) return "invalid bonded topology index";
    s.nexcl = (int)pr.size();
    s.ex_num.assign(s.N, 0);
    s.ex_off.assign(s.N, 0);
    s.ex_j.resize(pr.size());
    for (size_t k = 0; k < pr.size(); k++) {
      s.ex_j[k] = pr[k].second;
      s.ex_num[pr[k].first]++;
    }
    for (int i = 1; i < s.N; i++) s.ex_off[i] = s.ex_off[i - 1] + s.ex_num[i - 1];
    // both directions, for the full list of the force kernels
    {
      std::vector<std::pair<int, int>> sym;
      sym.reserve(2 * pr.size());
      for (auto &p : pr) {
        sym.emplace_back(p.first, p.second);
        sym.emplace_back(p.second, p.first);
      }
      std::sort(sym.begin(), sym.end());
      s.exs_num.assign(s.N, 0);
      s.exs_off.assign(s.N, 0);
      s.exs_j.resize(sym.size());
      for (size_t k = 0; k < sym.size(); k++) {
        s.exs_j[k] = sym[k].second;
        s.exs_num[sym[k].first]++;
      }
      for (int i = 1; i < s.N; i++) s.exs_off[i] = s.exs_off[i - 1] + s.exs_num[i - 1];
    }
    s.e1.clear();
    s.e2.clear();
    for (auto &p : pr) {
      if (sys->q[p.first] != 0.0 && sys->q[p.second] != 0.0) {
        s.e1.push_back(p.first);
        s.e2.push_back(p.second);
      }
    }
  }
  // --- excluded-pair table (splin_ecor, set_exclude.c:548-779)
  {
    const int n = s.nsplin;
    const double ten = 10.0 / BOHR;
    double rmax = 0.0;
    if (ix != nullptr) {
      for (int ip = 0; ip < s.P; ip++) {
        for (size_t k = 0; k < s.e1.size(); k++) {
          const size_t i1 = (size_t)ip * s.N + s.e1[k], i2 = (size_t)ip * s.N + s.e2[k];
          const double dx = ix[i1] - ix[i2], dy = iy[i1] - iy[i2], dz = iz[i1] - iz[i2];
          rmax = std::max(dx * dx + dy * dy + dz * dz, rmax);
        }
      }
    }
    rmax += 2.0 / BOHR;
    rmax = std::max(ten, rmax);
    const double rmin = 0.1, dr = (rmax - rmin) / (double)(n - 5), dri = 1.0 / dr;
    s.e_rmin = rmin;
    s.e_rmax = rmax;
    s.e_dr = dr;
    s.e_dri = dri;
    s.ecor_tab.assign((size_t)n * 2, 0.0);
    const double alp = s.alp_ewd, alp2 = alp * alp, two_alpha2 = 2.0 * alp * alp, p_alpha = kErfP * alp;
    for (int i = 1; i <= n; i++) {
      const double r12 = dr * (double)(i - 3) + rmin;
      const double r12i = 1.0 / r12;
      const double ralp = r12 * alp;
      const double eee = exp(-ralp * ralp);
      const double tt = 1.0 / (1.0 + kErfP * ralp);
      const double temp = erfc_poly(tt) * eee;
      const double gerf = 1.0 - temp;
      const double dgerf = derfc_poly(tt) * eee * p_alpha + two_alpha2 * temp * r12;
      s.ecor_tab[2 * (i - 1)] = -gerf * r12i;
      const double dvecor = -(dgerf * r12i - gerf / (r12 * r12));
      s.ecor_tab[2 * (i - 1) + 1] = -dvecor * r12i;
    }
    const double kcut = sqrt(2.0 * (4.0 * s.ecut));
    const double arg = kcut / (2.0 * alp);
    const double self_erfc = erfc_poly(1.0 / (1.0 + kErfP * arg)) * exp(-arg * arg);
    s.self_erf = 1.0 - self_erfc;
    const double kmax = 15.0 * alp;
    if (kmax > kcut) {
      const int nk = 128;
      double anode[128], weight[128], know[128], fknow[128];
      gauss_legendre(nk, anode, weight);
      const double dk = kmax - kcut, pk = kmax + kcut;
      for (int k = 0; k < nk; k++) {
        know[k] = 0.5 * (anode[k] * dk + pk);
        fknow[k] = (dk / M_PI) * exp(-0.25 * know[k] * know[k] / alp2) * weight[k];
      }
      for (int i = 1; i <= n; i++) {
        const double r12 = dr * (double)(i - 3) + rmin;
        double chi = 0.0, dchi = 0.0;
        for (int k = 0; k < nk; k++) {
          const double a = know[k] * r12;
          chi += (fknow[k] * (sin(a) / a));
          dchi += (know[k] * fknow[k] * ((sin(a) / a - cos(a)) / a) / r12);
        }
        s.ecor_tab[2 * (i - 1)] += chi;
        s.ecor_tab[2 * (i - 1) + 1] += dchi;
      }
    }
  }
  // --- springs (path_integral_init, transform_cnt.c:140-201)
  {
    const double T = par->temperature > 0 ? par->temperature : 300.0;
    const double dP = (double)s.P;
    s.prekf.assign(s.N, 0.0);
    for (int i = 0; i < s.N; i++) {
      const double beta = BOLTZ / T, tau = BOLTZ / (T * dP);
      s.prekf[i] = sys->mass[i] / (2.0 * tau * beta);
    }
    for (int g = 0; g < sys->nghost; g++) s.prekf[sys->ghost_atom[g]] = 0.0;
    s.veig.assign(s.P + 1, 0.0);
    if (par->pimd_typ == PMX_PIMD_STAGING) {
      for (int ip = 2; ip <= s.P; ip++) s.veig[ip] = ((double)(ip)) / ((double)(ip - 1));
    } else {
      const double tpip = 2.0 * M_PI / dP, pre = 4.0 * dP;
      if (s.P > 1) s.veig[s.P] = pre;
      for (int ip = 2; ip <= s.P / 2; ip++) {
        const double temp = (1.0 - cos(tpip * (double)(ip - 1))) * pre;
        s.veig[2 * ip - 2] = temp;
        s.veig[2 * ip - 1] = temp;
      }
    }
    s.veig[1] = 0.0;
  }
  // A cutoff beyond half the box only draws a warning in the reference
  // (energy_control_inter_real.c:323-348); the minimum image is used as is.
  return "";
}

}  // namespace pmx
