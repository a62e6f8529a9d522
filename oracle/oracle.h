/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU, FP64, single-threaded restatement of PINY_MD's classical force
 * evaluation for path-integral water.  It is the checker the HIP library is
 * compared against; nothing in the product links or calls it.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark
 * load it.
 *
 * It consumes the same pmx_system / pmx_params descriptions as the product
 * (include/piny_amd.h) so that both see identical inputs.
 */
#ifndef PINY_ORACLE_H
#define PINY_ORACLE_H

#include "../include/piny_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx *or_create(const pmx_system *sys, const pmx_params *par,
                  const double *init_x, const double *init_y,
                  const double *init_z, char *err, size_t errlen);
void or_destroy(or_ctx *c);

/* One force call on positions x,y,z [P*N] (ghost positions recomputed
 * in place).  Forces [P*N] are overwritten.  flags as pmx_energy_control.
 * xmod may be NULL (bead centroid). */
int or_energy_control(or_ctx *c, double *x, double *y, double *z,
                      const double *xmod, const double *ymod,
                      const double *zmod, unsigned flags, double *fx,
                      double *fy, double *fz, pmx_energies *out);

/* RESPA settings (NULL: off), as pmx_set_respa. */
int or_set_respa(or_ctx *c, const pmx_respa *r);

int or_get_verlet_list(or_ctx *c, int *nter, int *jver_off, int *jter,
                       int cap, int *nver);
int or_get_nktot(or_ctx *c, int *nktot, int *ngrid);
int or_get_kvectors(or_ctx *c, int *ka, int *kb, int *kc, double *bweight,
                    int cap);
int or_get_spline(or_ctx *c, int which, int ipair, double *out, int cap);
int or_get_setup_scalars(or_ctx *c, double *out);
int or_get_bead_energies(or_ctx *c, double *out);

/* Reciprocal-space energy of one configuration (N sites, the context's
 * charges and cell).  use_pme = 0: standard Ewald (ewald3d.c); 1: PME.
 * Forces (may be NULL) are overwritten with the k-space force. */
double or_recip_energy(or_ctx *c, const double *x, const double *y,
                       const double *z, int use_pme, double *fx, double *fy,
                       double *fz, double *pvten9);

/* intra_coefs.c restatements */
void or_bond_coef_harm(double fk, double *c7);
void or_bond_coef_morse(double d0, double alpha, double *c7);
void or_bend_coef_harm(double fk, double theta0, double *c7, double *s7);

/* Normal-mode transforms of transform_cnt.c for one site: x[P] (bead
 * Cartesian) -> u[P] (modes, PINY's storage order), and the force
 * transform f[P] -> fm[P].  Naive DFT. */
void or_nm_pos_to_mode(int P, const double *x, double *u);
void or_nm_mode_to_pos(int P, const double *u, double *x);
void or_nm_force_to_mode(int P, const double *f, double *fm);

/* Staging transforms of transform_stg.c for one site (path_int_md_typ
 * staging): positions to modes, modes to positions, forces to modes. */
void or_stag_pos_to_mode(int P, const double *x, double *u);
void or_stag_mode_to_pos(int P, const double *u, double *x);
void or_stag_force_to_mode(int P, const double *f, double *fm);

/* nsteps of PIMD in the NVT ensemble (int_NVT_pimd with RESPA off and
 * global / per-degree-of-freedom Nose-Hoover chains), starting with a force
 * call at the given positions.  x, y, z: bead Cartesian [P*N], in/out;
 * vx, vy, vz: mode velocities [P*N], in/out; nhc: chain state, in/out
 * (2 len for the centroid chain, then 2 len per mode 2..P and degree of
 * freedom); econs[4] (may be NULL): kinetic, chain energy, potential with
 * springs, their sum. */
int or_pimd_nvt(or_ctx *c, const pmx_nhc_params *np, int nsteps, double *x, double *y, double *z, double *vx,
                double *vy, double *vz, double *nhc, pmx_energies *out, double *econs);

/* Gauss-Legendre nodes and weights on [-1,1] (n points). */
void or_gauss_legendre(int n, double *node, double *weight);

#ifdef __cplusplus
}
#endif
#endif
