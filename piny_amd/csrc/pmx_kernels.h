// Kernel argument blocks and launch declarations (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "pmx_internal.h"

namespace pmx {

struct ListArgs {
  int N, cap, pimd;
  int full;  // 1: jforce holds every interacting j != i (both directions)
  int ref_only;  // k_list_build: write the reference's half list only (jfull, nfull)
  // link cells (when use_cells): nc[3] cells per axis of fractional space,
  // sites of cell k at cell_sites[cell_start[k] .. + cell_cnt[k]]
  int use_cells, nc[3];
  int *site_cell, *site_pos, *cell_cnt, *cell_start, *cell_sites;
  // k_list_cells: the sites in cell order, (bead-1 x, y, z, q) and
  // (xmod x, y, z, type as a double), and the site index
  double4 *cell_p, *cell_m;
  int *cell_i;
  const double *bx, *by, *bz;  // bead-1 positions [N]
  const double *xm, *ym, *zm;  // xmod [N]
  const int *type;
  const double *q;
  const int *ex_off, *ex_num, *ex_j;  // exclusions (both directions when full)
  PairTypes pt;
  Cell cell;
  int *jfull, *jforce;  // ELL rows of width cap
  int *nfull, *nforce;
  int *overflow;
};

// Per type pair constants of the force kernels, PP_NUM doubles each: cutoff,
// table origin and spacing, and the tabulated functions' parameters for
// direct evaluation (spline_fit.c:28-398): LJ 4 eps (s^6/r^6)((s^6/r^6) - 1)
// and the PINY erfc polynomial, both times the switch of width rheal below
// the cutoff when swit is set, minus vshift (shift_inter_pe = shift).
// For the parameter gradients: the LJ shift per unit eps (VSLJ1) and
// sigma d/dsigma of it (DSLJ1), and the type pair's LJ gradient slot (GSLOT,
// -1 when the pair has no LJ term).  C12 = 4 eps s^12 and C6 = 4 eps s^6
// give the FP32 kernels the LJ term as r^-6 (C12 r^-6 - C6).  CULL and CUT2
// are the cutoff of the call, where the kernels cull, and its square (the
// RESPA cutoff in short-range RESPA calls, else CUT); CUT stays the
// functions' cutoff, where their switch ends; CUTRES is the type pair's
// RESPA cutoff (inter_cutoff_res).
enum {
  PP_CUT = 0, PP_RMIN, PP_DRI, PP_DR, PP_EPS, PP_SIG6, PP_VSLJ, PP_VSC, PP_LJ,
  PP_VSLJ1, PP_DSLJ1, PP_GSLOT, PP_C12, PP_C6, PP_CUT2, PP_CUTRES, PP_CULL, PP_NUM
};
// RESPA weighting of the pair forces (force_npol.c:807-816,1033-1060):
// off; the short-range part alone (switch times w_res); full and short-range
// (switch times w_res plus its complement times w_full); full alone, times
// w_full.
enum { RESPA_OFF = 0, RESPA_RES = 1, RESPA_BOTH = 2, RESPA_FULL = 3 };
// LJ type pairs whose eps/sigma gradients one context forms
constexpr int MAX_LJ_GRAD = 4;
constexpr int CT_N = 1024;  // intervals of the mixed kernel's tables (16 KB of LDS each)
constexpr int MAX_FTAB = 4;  // switched pair-function tables the mixed kernel holds in LDS

// List entries of the force kernels carry the type pair in the top byte.
constexpr int ENTRY_SHIFT = 24;
constexpr int ENTRY_MASK = (1 << ENTRY_SHIFT) - 1;

struct PairArgs {
  int N, P, cap, nsplin, pimd, npair;
  int gx;  // workgroups along x (pair_grid_x)
  double alp, p_alpha, two_alpha2, rheal, irheal;  // erfc and switch constants
  double spalp;                           // 2 alp / sqrt(pi)
  // mixed precision: erfc(alp r)/r (PINY's polynomial) as cubic pieces in r,
  // CT_N intervals from ct_r0 with 1/width ct_idr: (c0, c1, c2, c3) of
  // c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1] across the interval
  const float4 *ctab;
  // the same erfc term switched and shifted as the type pairs' tables hold
  // it (build_pair_tables), when every type pair shares one; else null
  const float4 *cswtab;
  float ct_r0, ct_idr;
  // the switched, shifted pair functions whole (erfc term per unit charge
  // product, LJ term) on the same grid: nftab tables of CT_N pieces, and per
  // type pair the offsets (erfc table, LJ table or -1) in pieces; ftab is
  // null when the functions need more than MAX_FTAB tables
  const float4 *ftab;
  int nftab;
  int2 ftoff[MAX_PAIRS];
  int swit;
  // RESPA: one of RESPA_*, the weights of the full and the short-range
  // forces (the switch runs over rheal below PP_CUTRES)
  int respa;
  double w_full, w_res;
  const double *pp;            // [npair][PP_NUM]
  const double *x, *y, *z;     // [N*P]
  const double4 *site;         // [N] (xmod, ymod, zmod, q)
  const float4 *u;             // mixed precision, PIMD: (x - xmod) per (site, bead) [N*P]
  const double *umax;          // mixed precision, PIMD: the largest |u| of every site over its beads [N]
  const int *type;
  const int *jlist, *cnt;      // entries j | t << ENTRY_SHIFT
  const double2 *tab_lj, *tab_c;  // [npair*nsplin] (V, -V'/r)
  Cell cell;
  double *fx, *fy, *fz;
  double *part;  // per-block partial sums [gridDim.x][P][PAIR_NV]
  double *phi;   // gradient kernel: dE_b/dq_i, real-space part [N*P]
};

// PME on a grid cut into tiles of T[0] x T[1] x T[2] points.  Charge (c, b)
// (charge c, bead b) is binned by the tile that holds the first point of its
// interpolation support; its support then lies in that tile's halo box of
// H = T + n - 1 points per axis (held in LDS by the gather), and reaches
// nbr[d] tiles along axis d (the spread of a tile reads nbr[0]*nbr[1]*nbr[2]
// bins).
struct PmeArgs {
  int N, P, nchrg, n;  // n: interpolation order
  int f32;             // grids in FP32 (mixed precision) instead of FP64
  int K[3], T[3], nt[3], H[3], nbr[3];
  int ntile;           // tiles per bead
  size_t gstride;      // doubles per bead grid
  size_t hvol;         // H[0]*H[1]*H[2]
  // k / d as __umulhi(k, m) with m = 2^32 / d + 1 (exact for k * d < 2^32):
  // for H[0], H[1], T[0], T[1]
  unsigned mH0, mH1, mT0, mT1;
  const double *x, *y, *z;
  const double *q;
  const int *ichrg;
  Cell cell;
  int *item_s;         // [P*nchrg][3] first support point per axis (0-based)
  double *item_f;      // [P*nchrg][3] fractional grid offset per axis
  int *item_tile, *item_pos;
  int *bin_cnt, *bin_off;  // [P*ntile]
  int *sorted;         // [P*nchrg] item ids in bin order
  void *hbuf;          // [P*ntile*hvol] halo boxes of spread charge (grid type)
  void *grid;          // [P*gstride] (grid type)
  double *rx, *ry, *rz;  // gather: the k-space force of every (charge, bead), bead-major [P][N]
  double *phi;         // potential gather: dE_b/dq_i += k-space part [N*P]
};

// Standard Ewald (ewald_pme_opt off): the sum over the half-space k list
// with direct phases.
struct EwaldArgs {
  int N, P, nchrg, nk;
  const double *x, *y, *z, *q;
  const int *ichrg;
  Cell cell;
  const double4 *kv;  // [nk] (2 pi ka, 2 pi kb, 2 pi kc, preg)
  const double4 *kg;  // [nk] (kx, ky, kz, prep / |S|^2 factor)
  double2 *sk;        // [P][nk] 2 preg (Re S, Im S)
  double *rx, *ry, *rz;  // the force of every (charge, bead), bead-major [P][N]
  double *part;       // partial rows, FFT_NV values
};
constexpr int EWALD_BLOCK = 256;

constexpr int FFT_MAX_PASSES = 12;
constexpr int FFT_MAX_PLANE = 4096;       // complex values per plane of the smaller class
constexpr size_t FFT_LDS_MAX = 163840;    // bytes of LDS one workgroup may hold on MI355X
// dynamic LDS the FFT kernels may ask for: the workgroup's LDS less the
// static 256 doubles of block_partials (the convolution kernel's sums)
constexpr size_t FFT_LDS_DYN = FFT_LDS_MAX - 256 * sizeof(double);

struct FftArgs {
  int P, f32;
  int K[3];
  int npass[3], radix[3][FFT_MAX_PASSES];  // Stockham passes per axis (radices 2..5)
  size_t gstride, sstride;
  void *grid;             // [P][Kc][Kb][Ka] real (grid type)
  void *spec;             // [P][Kc][Kb][Ka/2+1] complex (grid type)
  const double2 *tw[3];   // exp(-2 pi i m/K) per axis
  const float2 *twf[3];   // the same in FP32
  const double *infl;    // [Kc][Kb][Ka/2+1] preg*bweight on the sphere list
  double falp2;
  Cell cell;
  double *part;  // [Kb][P][FFT_NV]
  // The forward transform reads the spread charge from the halo boxes of the
  // PME tiling (PmeArgs) and sums their overlaps as it loads each plane.
  const void *hbuf;  // [P*ntile*hvol] (grid type)
  int T[3], nt[3], H[3], nbr[3], ntile;
  size_t hvol;
  unsigned mKa, mT0, mT1;  // magic divisors of K[0], T[0], T[1]
};

struct IntraArgs {
  int N, P, pimd, intra_perds, nsplin;
  const double *x, *y, *z;
  const double *xm, *ym, *zm;
  const double *q;
  Cell cell;
  double *fx, *fy, *fz;
  double *part;  // [gridDim.x][P][INTRA_NV], set per kernel
  // bonds
  int nbond;
  const int *b1, *b2, *btyp;
  const double *bc, *beq;
  // bends
  int nbend;
  const int *a1, *a2, *a3, *atyp;
  const double *ac, *as;
  // Ewald corrections
  int necor;
  const int *e1, *e2;
  const double2 *ecor_tab;
  double e_rmin, e_dr, e_dri;
  // gradients: ecor partners of every charged site (both directions),
  // d(vself + vbgr)/dq_i = cself*q_i + cbgr, outputs
  int nchrg;
  const int *ichrg, *ecs_off, *ecs_num, *ecs_j;
  double cself, cbgr;
  double *phi;   // [N*P] += excluded-pair and self/background parts
  int nbond_type, nbend_type;
  double *gbond;  // [P][nbond_type][7] sum over bonds of (r - eq)^m
  double *gbend;  // [P][nbend_type][BEND_GRAD_NV]
  // k_intra_mol: molecules of consecutive sites (first site, template),
  // their templates, the k-space forces to add (bead-major [3][P][N], or
  // null) and the partial rows it fills
  int nmol, mol_max, ntpl;  // molecules, sites of the largest, templates
  const int2 *mol;
  const struct MolTpl *tpl;
  const double *rec;
  double *part_ecor, *part_bond, *part_bend, *part_fin;
  // RESPA force weights (energy_control_initial.c:212-243) of the excluded-
  // pair terms and the k-space forces (wght_ter) and of the bonds and bends
  // (wght_tra_res); 1 without RESPA
  double w_ecor, w_rec, w_intra;
};
constexpr int MOL_MAX = 8;   // sites of a molecule of k_intra_mol
constexpr int TPL_MAXT = 16;  // terms of one kind in a molecule template
constexpr int TPL_MAX = 4;   // templates of one context

// The intramolecular terms of one kind of molecule, by local site index:
// excluded pairs (sites, q_i q_j), bonds (sites, type), bends (sites with
// the apex second, type), ghosts (site, components, their sites and
// coefficients).
struct MolTpl {
  int n, ne, nb, na, ng, pad;
  int e[TPL_MAXT][2];
  int bd[TPL_MAXT][3];
  int an[TPL_MAXT][4];
  int gh[MOL_MAX][6];
  double eq[TPL_MAXT];
  double gc[MOL_MAX][4];
};

// Bend moments per (bead, bend type): cos^m (m = 0..6), sin, cos sin^(m-1)
// (m = 2..6): the derivatives of V with respect to c_m and s_m.
constexpr int BEND_GRAD_NV = 13;

struct FinishArgs {
  int N, P;
  const double *x, *y, *z, *xm, *ym, *zm;
  double *fx, *fy, *fz;
  const double *rec;  // PME forces to add, bead-major [3][P][N], or null
  double w_rec;       // their weight (1 without RESPA)
  double *part;  // [gridDim.x][P][FIN_NV]
};

struct ModeArgs {
  int N, P;
  int stag;  // staging transform (path_int_md_typ staging) instead of normal modes
  const double *x, *y, *z;
  double *fx, *fy, *fz;
  double *fxt, *fyt, *fzt;  // the bead forces before the transform (may be null)
  const double *prekf, *veig;
  const double *cs;  // [2P] cos, sin of 2 pi m/P
  double *part;      // spring energy per workgroup
};
// sites of one k_mode_forces workgroup (four wavefronts of 64 / P sites)
inline int mode_blocks(int N, int P) {
  const int per = 4 * (64 / P);
  return (N + per - 1) / per;
}

// PIMD NVT integrator (int_NVT_pimd with RESPA off).  Velocities are mode
// velocities, atom-major like the positions: v[i*P + m] for mode m.  Chain
// state on the device, chain-major: the chains of modes 2..P for every
// (site, mode) with m >= 1 as vnhc[(j*3 + d)*N*P + i*P + m] and xnhc
// likewise (j < len, d the direction; entries of mode 0 and of ghost sites
// unused); the centroid chain as cls[j] (v) and cls[len + j] (x).
struct NhcArgs {
  int N, P, len, nres, nyosh;
  double wdti2[10], wdti4[10], wdti8[10];  // set_yosh weights times the step, 1-based
  double cmass[8], cgkt[8];                // the centroid chain (1-based)
  double bmass[8], bgkt[8];                // every bead chain (1-based)
  double dt, dt2;
  const double *mass;   // [N] physical masses
  const double *mfac;   // [P] mode mass factor: veig (1 for mode 1) times gamma^2 for modes 2..P
  const int *ghost;     // [N] 1 for ghost sites (the null thermostat)
  const double *fx, *fy, *fz;
  double *vx, *vy, *vz;
  double *vnhc, *xnhc;  // [len][3][N*P]
  double *cls;          // [2 len] centroid chain; cls[2 len] holds the velocity scale of the last chain step
  double *part;         // per-workgroup partial sums
  int stag;
  const double *cs;     // [2P] cos, sin of 2 pi m/P
  double *x, *y, *z, *xm, *ym, *zm;
};
__global__ void k_nhc_ckin(NhcArgs a);
__global__ void k_nhc_class(NhcArgs a, int G);
__global__ void k_nhc_kick(NhcArgs a, int phase);
__global__ void k_nhc_scale0(NhcArgs a);
__global__ void k_mode_drift(NhcArgs a);
__global__ void k_pimd_energy(NhcArgs a);
__global__ void k_pimd_energy_sum(NhcArgs a, int G, double *out);
constexpr int NHC_BLOCKS = 256;  // workgroups of the grid-stride kinetic and energy sums

// Values per (block, bead) partial-sum row of each stage.
constexpr int PAIR_NV = 9;   // vvdw, vcoul, 6 virial, pair count
constexpr int FFT_NV = 7;    // vrecip, 6 virial
constexpr int INTRA_NV = 7;  // energy, 6 virial
constexpr int FIN_NV = 10;   // 9 virial corrections, akvir
constexpr int PAIRG_NV = 2 * MAX_LJ_GRAD;  // per LJ slot: energy per unit eps, sigma dV/dsigma per unit eps

// One stage of partial sums: part[off + (g*P + b)*nv + v] for g < G holds
// the contribution of block g to accumulator slot[v] of bead b.
struct RedStage {
  size_t off;
  int G, nv;
  int slot[12];
  double wv;  // weight of values 1 .. nv-1 (the virial of a RESPA-weighted stage), 1 otherwise
};
constexpr int MAX_RED_STAGES = 8;
struct RedArgs {
  int P, nstage;
  int nslot;  // accumulator slots per bead (ACC_NUM for the energies)
  RedStage st[MAX_RED_STAGES];
  const double *part;
  double *acc;  // [P][nslot], every entry written
  // one scalar besides: *sacc = the sum of spart[0 .. sG) (0 when sG = 0),
  // by the workgroup after the per-bead ones
  const double *spart;
  int sG;
  double *sacc;
};

__global__ void k_bead_to_atom(int N, int P, const double *src, double *dst);
__global__ void k_atom_to_bead(int N, int P, const double *src, double *dst);
__global__ void k_ghost_pos(int ng, int P, const int *gatom, const int *gncomp, const int *gcomp,
                            const double *gcoef, double *x, double *y, double *z);
__global__ void k_pack_sites(int N, int pimd, const double *xm, const double *ym, const double *zm,
                             const double *q, double4 *site);
__global__ void k_pack_disp(int N, int P, const double *x, const double *y, const double *z, const double *xm,
                            const double *ym, const double *zm, float4 *u);
__global__ void k_site_umax(int N, int P, const float4 *u, double *umax);
__global__ void k_pack_disp_umax(int N, int P, const double *x, const double *y, const double *z, const double *xm,
                                 const double *ym, const double *zm, float4 *u, double *umax);
__global__ void k_site_umax64(int N, int P, const double *x, const double *y, const double *z, const double *xm,
                              const double *ym, const double *zm, double *umax);
// Pack: what k_pack_sites and k_pack_disp_umax write, formed in the same
// pass (disp non-null: mixed precision with P a power of two up to 64).
struct CheckPack {
  float4 *disp;
  double *umax;
  double4 *site;
  const double *q;
};
__global__ void k_list_check(int N, int P, const double *x, const double *y, const double *z, const double *x0,
                             const double *y0, const double *z0, double *b0x, double *b0y, double *b0z,
                             double *maxes, int centroid, double *xm, double *ym, double *zm, CheckPack pk);
constexpr int LIST_CHECK_BLOCKS = 512;  // workgroups of k_list_check, one pair of maxima each
__global__ void k_list_build(ListArgs a);
__global__ void k_cell_count(ListArgs a);
__global__ void k_cell_scan(int ncell, const int *cnt, int *start);
__global__ void k_cell_scatter(ListArgs a);
constexpr int CELL_SORT_MAX = 1024;  // sites of one cell k_cell_order sorts in LDS
constexpr int CELL_ROWS_MAX = 512;   // sites of one cell whose rows one k_list_cells workgroup keeps
constexpr int CELL_TILE = 320;       // candidate sites one k_list_cells workgroup stages in LDS at a time
constexpr int SPLIT_CHUNKS = 32;     // 64-entry chunks of a row whose inside flags k_split_rows keeps
__global__ void k_cell_order(ListArgs a);
hipError_t launch_list_cells(const ListArgs &a, int halfwidth, int ncell, int nw, hipStream_t st);
__global__ void k_sort_rows(int cap, int n2, const int *nfull, const int *nforce, int *jfull, int *jforce);
__global__ void k_pme_bin(PmeArgs a);
__global__ void k_pme_scan(PmeArgs a);
__global__ void k_pme_scatter(PmeArgs a);
__global__ void k_fill_int(int *p, int v, size_t n);
__global__ void k_fill_double(double *p, double v, size_t n);
__global__ void k_reduce_partials(RedArgs a);
__global__ void k_split_rows(ListArgs a, int *out);
__global__ void k_ecor_phi(IntraArgs a);
__global__ void k_bond_moments(IntraArgs a);
__global__ void k_bend_moments(IntraArgs a);
__global__ void k_ecor(IntraArgs a);
__global__ void k_bond(IntraArgs a);
__global__ void k_bend(IntraArgs a);
__global__ void k_intra_mol(IntraArgs a);
__global__ void k_pimd_finish(FinishArgs a);
__global__ void k_add_rec(int N, int P, const double *rec, double w, double *fx, double *fy, double *fz);
__global__ void k_ghost_force(int ng, int P, const int *gatom, const int *gncomp, const int *gcomp,
                              const double *gcoef, double *fx, double *fy, double *fz);
hipError_t launch_mode_forces(const ModeArgs &a, hipStream_t st);
__global__ void k_vv_kick(int n, double dt2, const double *mass, const double *fx, const double *fy,
                          const double *fz, double *vx, double *vy, double *vz);
__global__ void k_vv_drift(int n, double dt, const double *vx, const double *vy, const double *vz, double *x,
                           double *y, double *z);
__global__ void k_kinetic(int n, const double *mass, const double *vx, const double *vy, const double *vz,
                          double *ke);

// mode: 0 FP64 tables, 1 FP64 direct, 2 FP32 direct with FP64 accumulation;
// full: the list holds both directions and each wave stores its sites'
// forces (no atomics); otherwise j forces are scattered with atomics
hipError_t launch_pair(const PairArgs &a, int mode, bool full, hipStream_t st);
// The mixed-precision pair kernel in molecule-cluster form, for boxes whose
// sites are nmol molecules of one template of ns <= CLU_MAX_NS consecutive
// sites.  Row I lists the molecules J with at least one pair (I a, J b) in
// the force list, and the site pairs the force list holds as bits
// a*ns + b (the list's own pairs, exclusions already removed).  The site
// pair separation is D_00 + v_Ia - v_Jb: D_00 the imaged separation of the
// molecules' site-0 reference positions (FP64, rounded), v the FP32
// position of a bead relative to its molecule's site-0 reference.  It equals
// the reference's separation whenever the site pair's image is the site-0
// pair's, which the staging checks for every listed pair (flag[1] counts
// the pairs where it is not; the host then repeats the call with the
// per-site kernel).
constexpr int CLU_MAX_NS = 4;
constexpr int CLU_HASH = 2048;  // LDS hash slots of the row build (unique partners of one molecule)

struct CluArgs {
  int nmol, ns, N, P, pimd;
  int mcap;                     // row width
  const int2 *mrow;             // [nmol][mcap] (J, site-pair bits)
  const int *mcnt;              // [nmol]
  // row build: the force list's site rows, the rows written, and
  // flag[0] the longest row, flag[2] the largest distance of a site from its
  // molecule's site 0 (float bits); flag[1] image mismatches (pair kernel)
  const int *jforce, *nforce;
  int cap;
  int2 *mrow_w;
  int *mcnt_w;
  int *flag;
  const double *xr, *yr, *zr;   // reference positions: xmod (PIMD) or bead positions (P = 1) [N]
  const double *x, *y, *z;      // bead positions [N*P] (packing)
  float4 *vw;                   // packing output
  const float4 *v;              // x - xr of the molecule's site 0, per (site, bead) [N*P]
  double *umax_w;               // packing output (PIMD)
  const double *umax;           // PIMD: the largest |x - xmod| of every site over its beads [N]
  Cell cell;
  const float4 *ftab;           // the whole pair functions, CT_N + 1 pieces each (the last zero)
  int npieces;                  // pieces held in LDS
  float r0i, idr;               // piece coordinate u = r idr - r0i
  int cmask, lmask;             // site pairs (bit a*ns + b) with an erfc term, with an LJ term
  int tcoul;                    // the erfc table (all site pairs share it), offset in pieces
  int toff_l[16];               // per site pair: LJ table offset in pieces
  float qq[16];                 // per site pair: q_a q_b
  double cut;                   // the cutoff of every site pair (Bohr)
  double *fx, *fy, *fz;
  double *part;                 // per-block partial sums [gridDim.x][P][PAIR_NV]
  unsigned long long *count;    // set: count the site pairs inside the cutoff instead (nothing else written)
};

hipError_t launch_clu_pack(const CluArgs &a, hipStream_t st);
hipError_t launch_clu_rows(const CluArgs &a, hipStream_t st);
hipError_t launch_pair_clu(const CluArgs &a, int gx, hipStream_t st);

hipError_t launch_pair_grad(const PairArgs &a, int mode, hipStream_t st);
int pair_grid_x(int N, int max_waves);
constexpr int PAIR_MAX_WAVES = 16384;  // the largest pair grid the autotuner tries
// LDS bytes of the spread and gather workgroups for a given tiling
size_t pme_lds_bytes(const PmeArgs &a, bool gather);
hipError_t launch_pme_spread(const PmeArgs &a, hipStream_t st);
hipError_t launch_pme_gather(const PmeArgs &a, hipStream_t st);
hipError_t launch_pme_phi(const PmeArgs &a, hipStream_t st);
hipError_t launch_ewald(const EwaldArgs &a, hipStream_t st);
hipError_t launch_pme_fft(const FftArgs &f, hipStream_t st);
// block size for grid-stride kernels whose threads keep one bead: the
// largest multiple of P not above 256
inline int bead_block(int P) { return 256 - (256 % P); }

}  // namespace pmx
