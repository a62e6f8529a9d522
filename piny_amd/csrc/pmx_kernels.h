// Kernel argument blocks and launch declarations (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "pmx_internal.h"

namespace pmx {

struct ListArgs {
  int N, cap, pimd;
  const double *bx, *by, *bz;  // bead-1 positions [N]
  const double *xm, *ym, *zm;  // xmod [N]
  const int *type;
  const double *q;
  const int *ex_off, *ex_num, *ex_j;
  PairTypes pt;
  Cell cell;
  int *jfull, *jforce;  // ELL rows of width cap
  int *nfull, *nforce;
  int *overflow;
};

struct PairArgs {
  int N, P, cap, nsplin, pimd;
  const double *x, *y, *z;     // [N*P]
  const double *xm, *ym, *zm;  // [N]
  const int *type;
  const double *q;
  const int *jlist, *cnt;
  const double2 *tab_lj, *tab_c;  // [npair*nsplin] (V, -V'/r)
  PairTypes pt;
  Cell cell;
  double *fx, *fy, *fz;
  double *acc;  // [P][ACC_NUM]
};

struct PmeArgs {
  int N, P, nchrg, n_interp;
  int K[3];
  size_t gstride;  // doubles per bead grid
  const double *x, *y, *z;
  const double *q;
  const int *ichrg;
  Cell cell;
  double *grid;
  double *fx, *fy, *fz;
};

struct FftArgs {
  int P;
  int K[3], logK[3];
  size_t gstride, sstride;
  double *grid;
  double2 *spec;
  const double2 *tw[3];  // exp(-2 pi i m/K) per axis
  const double *infl;    // [Kc][Kb][Ka/2+1] preg*bweight on the sphere list
  double falp2;
  Cell cell;
  double *acc;
};

struct IntraArgs {
  int N, P, pimd, intra_perds, nsplin;
  const double *x, *y, *z;
  const double *xm, *ym, *zm;
  const double *q;
  Cell cell;
  double *fx, *fy, *fz;
  double *acc;
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
};

struct FinishArgs {
  int N, P;
  const double *x, *y, *z, *xm, *ym, *zm;
  double *fx, *fy, *fz;
  double *acc;
};

struct ModeArgs {
  int N, P;
  const double *x, *y, *z;
  double *fx, *fy, *fz;
  const double *prekf, *veig;
  const double *cs;  // [2P] cos, sin of 2 pi m/P
  double *ake;
};

__global__ void k_bead_to_atom(int N, int P, const double *src, double *dst);
__global__ void k_atom_to_bead(int N, int P, const double *src, double *dst);
__global__ void k_ghost_pos(int ng, int P, const int *gatom, const int *gncomp, const int *gcomp,
                            const double *gcoef, double *x, double *y, double *z);
__global__ void k_centroid(int N, int P, const double *x, const double *y, const double *z, double *xm,
                           double *ym, double *zm);
__global__ void k_list_check(int N, int P, const double *x, const double *y, const double *z,
                             const double *x0, const double *y0, const double *z0, double *b0x, double *b0y,
                             double *b0z, unsigned long long *maxes);
__global__ void k_list_build(ListArgs a);
__global__ void k_pme_spread(PmeArgs a);
__global__ void k_pme_gather(PmeArgs a);
__global__ void k_fft_fwd_ab(FftArgs a);
__global__ void k_fft_c_conv(FftArgs a);
__global__ void k_fft_inv_ba(FftArgs a);
__global__ void k_ecor(IntraArgs a);
__global__ void k_bond(IntraArgs a);
__global__ void k_bend(IntraArgs a);
__global__ void k_pimd_finish(FinishArgs a);
__global__ void k_ghost_force(int ng, int P, const int *gatom, const int *gncomp, const int *gcomp,
                              const double *gcoef, double *fx, double *fy, double *fz);
__global__ void k_mode_forces(ModeArgs a);

hipError_t launch_pair(const PairArgs &a, hipStream_t st);

}  // namespace pmx
