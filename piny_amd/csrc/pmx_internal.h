// Internal declarations of the MI355X force library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/piny_amd.h"

namespace pmx {

constexpr double BOHR = 0.529177;     // defines.h:78
constexpr double BOLTZ = 315777.0;    // defines.h:80
constexpr int MAX_TYPES = 8;
constexpr int MAX_PAIRS = MAX_TYPES * (MAX_TYPES + 1) / 2;
constexpr int MAX_INTERP = 12;
// PINY's erfc(x) ~ exp(-x^2) sum_k e_k t^k, t = 1/(1 + p x) (spline_fit.c:327-339)
constexpr double kErfP = 0.3614;
constexpr double kErfE[10] = {0.0,
                              0.2041422096422003,
                              0.1997535956961481,
                              0.2213176596405576,
                              0.03360430734640255,
                              0.4732592578721755,
                              -0.509078520069735,
                              0.6772631491947646,
                              -0.369912979092217,
                              0.06965131976970335};

// ---------------------------------------------------------------------------
// Host-side setup products (restating PINY's parse-time setup).
struct Setup {
  int N = 0, P = 1, ntype = 0, npair = 0, nsplin = 0;
  // cell, 1-based copies as PINY stores them (index 0 unused)
  double hmat[10] = {0}, hmati[10] = {0}, vol = 0.0;
  int cubic = 1;
  // per type pair
  std::vector<int> label;
  std::vector<double> rmin, dr, dri, cutoff, cutoff_res, cutskin2;
  // tables [npair][nsplin][4] = {cv0, cdv0, cv0_c, cdv0_c}
  std::vector<double> tab;
  // the tabulated functions' parameters, for direct evaluation: LJ eps and
  // sigma^6, and the constant removed from V by shift_inter_pe = shift
  std::vector<double> lj_eps, lj_sig, lj_sig6, vshift_lj, vshift_c;
  double cutoff_max = 0.0, clong = 0.0, clong_res = 0.0;
  // each type pair's share of clong per unit LJ eps (clong = sum eps*clong_u)
  std::vector<double> clong_u;
  // charges and the sums of ewald3d_selfbgr (ewald3d_self_bgr.c:55-61)
  std::vector<int> ichrg;
  double qsq = 0.0, qsum = 0.0;
  // exclusions: per atom (0-based) sorted list of j < i
  std::vector<int> ex_off, ex_num, ex_j;
  int nexcl = 0;
  // the same exclusions listed under both sites (j < i and j > i)
  std::vector<int> exs_off, exs_num, exs_j;
  // Ewald-corrected excluded pairs (0-based, e1 > e2)
  std::vector<int> e1, e2;
  std::vector<double> ecor_tab;  // [nsplin][2] = {cv0, cdv0}
  double e_rmin = 0, e_rmax = 0, e_dr = 0, e_dri = 0;
  // Ewald / PME
  double alp_ewd = 0, ecut = 0, self_erf = 1.0;
  int kmaxv[4] = {0};
  int nktot = 0, ngrid[3] = {0, 0, 0};
  std::vector<int> ka, kb, kc;
  std::vector<double> bweight;
  // springs
  std::vector<double> prekf, veig;
};

// Builds Setup from the ABI description; returns an empty string or an
// error message.
std::string build_setup(const pmx_system *sys, const pmx_params *par, const double *ix,
                        const double *iy, const double *iz, Setup &s);

void bond_coef_harm(double fk, double *c7);
void bond_coef_morse(double d0, double alpha, double *c7);
void bend_coef_harm(double fk, double th0, double *c7, double *s7);

// ---------------------------------------------------------------------------
// Small POD blocks passed to kernels by value.
struct Cell {
  double h[9], hi[9];  // 0-based copies of hmat[1..9], hmati[1..9]
  int cubic;
};

struct PairTypes {
  int ntype;
  int lj[MAX_PAIRS];          // 1 if the pair has a nonzero LJ table
  double rmin[MAX_PAIRS], dr[MAX_PAIRS], dri[MAX_PAIRS], cut[MAX_PAIRS];
  double cutskin2[MAX_PAIRS];
};

__host__ __device__ inline int pair_index(int ntype, int ta, int tb) {
  int a = ta < tb ? ta : tb, b = ta + tb - a;
  return a * ntype - (a * (a - 1)) / 2 + (b - a);
}

// Number of per-bead accumulators written by every stage into the
// partial-sum buffer (see Ctx::acc).
enum Acc {
  ACC_VVDW = 0,   // real-space LJ energy
  ACC_VCOUL,      // real-space erfc Coulomb energy
  ACC_VRECIP,     // k-space sum
  ACC_VECOR,      // excluded-pair correction
  ACC_VBOND,
  ACC_VBEND,
  ACC_P11, ACC_P22, ACC_P33, ACC_P12, ACC_P13, ACC_P23,  // symmetric virial
  // PIMD virial correction -sum (x - xmod) (x) f/P, all nine components
  ACC_C11, ACC_C22, ACC_C33, ACC_C12, ACC_C21, ACC_C13, ACC_C31, ACC_C23, ACC_C32,
  ACC_AKVIR,      // -sum (x - xmod) . f/P
  ACC_NPAIR,      // pairs inside the cutoff (work counter)
  ACC_NUM
};

}  // namespace pmx
