/*
 * piny_amd.h -- C ABI of the MI355X force evaluation for path-integral water.
 *
 * This library replaces PINY_MD's classical force call
 *     void energy_control(CLASS*, BONDED*, GENERAL_DATA*)
 *         src/energy/control/energy_control.c:33-77,
 *         prototype src/energy/control/proto_energy_ctrl_entry.h:4
 *     void energy_control_pimd(CLASS*, BONDED*, GENERAL_DATA*)
 *         src/energy/control/energy_control_pimd.c:37-278,
 *         prototype src/energy/control/proto_energy_ctrl_entry.h:17
 * for the water force fields (LJ + erfc-Ewald real space, PME, bonds,
 * bends, excluded-pair Ewald corrections, massless ghost sites, ring-polymer
 * springs).  The reference keeps everything inside its CLASS / BONDED /
 * GENERAL_DATA structs; here the same data crosses the boundary as plain
 * pointers and sizes, in PINY's internal units (Bohr, Hartree, elementary
 * charge, electron masses, Kelvin) and 0-based indices.
 *
 * Arrays that PINY keeps per bead (clatoms_pos[ip].x[1..N],
 * typedefs_class.h:209-220) are passed bead-major: x[ip*N + i] for bead
 * ip = 0..P-1 and site i = 0..N-1.
 *
 * Every entry point returns PMX_OK (0) or an error code; the text of the
 * last error of a context is available from pmx_last_error().  Nothing in
 * this library calls exit() (the reference does, e.g. force_npol.c:174-182).
 * A context is not re-entrant; distinct contexts may be used concurrently
 * from different host threads (as PINYEmpirical does with replicas,
 * python/piny/piny.pyx:426-436).
 */
#ifndef PINY_AMD_H
#define PINY_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_ABI_VERSION 8

/* return codes */
enum {
  PMX_OK = 0,
  PMX_ERR_ARG = 1,      /* invalid argument or inconsistent input        */
  PMX_ERR_DEVICE = 2,   /* HIP runtime error                             */
  PMX_ERR_STATE = 3,    /* call out of order (e.g. no positions yet)     */
  PMX_ERR_RANGE = 4,    /* a size exceeds what was allocated             */
  PMX_ERR_NOMEM = 5,    /* allocation failure                            */
  PMX_ERR_NOTIMPL = 6   /* option not supported by this build            */
};

/* intermolecular potential labels, as PINY's inter_label
 * (control_inter_params.c:447-470) */
enum { PMX_INTER_LJ = 2, PMX_INTER_NULL = 4 };

/* shift_inter_pe (set_sim_params.c:1185-1197): 0 none, 1 shift, 2 swit */
enum { PMX_SHIFT_NONE = 0, PMX_SHIFT_SHIFT = 1, PMX_SHIFT_SWIT = 2 };

/* arithmetic of the pair, spline and FFT work */
enum { PMX_PREC_FP64 = 0, PMX_PREC_MIXED = 1 };

/* how FP64 pair terms are evaluated: PINY's 4-point cubic interpolation of
 * the spline tables (force_npol.c:339-376), or the tabulated function itself
 * (LJ and PINY's erfc polynomial with the same switch and shift,
 * spline_fit.c:28-398) computed per pair; the two agree to the tables'
 * interpolation error.  The mixed precision always evaluates directly. */
enum { PMX_PAIR_TABLE = 0, PMX_PAIR_DIRECT = 1 };

/* path_int_md_typ (transform_cnt.c:104,190): 1 staging, 2 centroid */
enum { PMX_PIMD_STAGING = 1, PMX_PIMD_CENTROID = 2 };

/* flags of pmx_energy_control */
enum {
  PMX_EVAL_PIMD = 1u << 0,        /* energy_control_pimd semantics: shared
                                     list from bead 1 + spread, xmod image
                                     convention, f/P, E/P, pvten/P       */
  PMX_EVAL_FORCE_LIST = 1u << 1,  /* rebuild the Verlet list regardless of
                                     the skin criterion                   */
  PMX_EVAL_NO_RECIP = 1u << 2,    /* skip the k-space sum (real, intra
                                     terms only)                          */
  PMX_EVAL_MODE_FORCES = 1u << 3, /* PIMD: also transform the forces to
                                     mode coordinates and add the spring
                                     forces (energy_control_pimd.c:200-218) */
  /* RESPA force calls (energy_ctrl.iget_res_inter / iget_full_inter, as
   * int_NVT_pimd.c:174-187 and the RESPA integrators set them; the
   * intramolecular flags are on).  Without flags a call is a full one. */
  PMX_EVAL_RES_INTER = 1u << 4,   /* iget_res_inter = 1: the short-range
                                     part of the pair forces, switched off
                                     over rheal below each type pair's
                                     inter_cutoff_res (force_npol.c:807-816,
                                     1033-1060); needs respa_steps_lrf > 0 */
  PMX_EVAL_NO_FULL_INTER = 1u << 5,/* iget_full_inter = 0: no full
                                     intermolecular forces, no k space, no
                                     excluded-pair correction and no
                                     intermolecular energies */
  PMX_EVAL_NO_FULL_INTRA = 1u << 6 /* iget_full_intra = 0 (an inner
                                     intramolecular RESPA step): no self and
                                     background terms, no energies kept
                                     unless the inter forces are full */
};

/*
 * Topology and force field after PINY's parse(): the values PINY holds in
 * CLATOMS_INFO (typedefs_class.h:124-177), INTERACT (:610-690), BOND, BEND,
 * GHOST_ATOMS (:225-243).  Type pairs are numbered like PINY's inter list
 * (control_inter_params.c:118-126): for 0-based types a <= b,
 *     pair(a,b) = a*ntype - a*(a-1)/2 + (b - a).
 */
typedef struct pmx_system {
  int natm;                     /* sites, ghosts included                  */
  int ntype;                    /* atom types                              */
  const int *atm_type;          /* [natm] 0..ntype-1                       */
  const double *q;              /* [natm] charge                           */
  const double *mass;           /* [natm] electron masses (amu*PROT_MASS)  */

  /* per type pair, ntype*(ntype+1)/2 entries */
  const int *inter_label;       /* PMX_INTER_LJ or PMX_INTER_NULL          */
  const double *inter_sig;      /* LJ sigma, Bohr                          */
  const double *inter_eps;      /* LJ epsilon, Hartree                     */
  const double *inter_cutti;    /* spline start (min_dist), Bohr           */
  const double *inter_cutoff;   /* cutoff (max_dist), Bohr                 */
  const double *inter_cutoff_res; /* RESPA cutoff (res_dist), Bohr         */

  /* bonds: V = sum_m c_m (r - eq)^m, m = 0..6 (bond.c:150-166) */
  int nbond;
  const int *bond_j1, *bond_j2, *bond_type;  /* [nbond]                   */
  int nbond_type;
  const double *bond_c;         /* [nbond_type*7]                          */
  const double *bond_eq;        /* [nbond_type] Bohr                       */

  /* bends j1-j2-j3 (j2 the apex):
   * V = sum_m c_m cos^m + cos * sum_{m>=2} s_m sin^(m-1) + s_1 sin
   * (bend.c:253-270) */
  int nbend;
  const int *bend_j1, *bend_j2, *bend_j3, *bend_type;
  int nbend_type;
  const double *bend_c;         /* [nbend_type*7]                          */
  const double *bend_s;         /* [nbend_type*7]                          */

  /* massless ghost sites r_g = sum_k coef_k r_comp_k (ghost_control.c) */
  int nghost;
  const int *ghost_atom;        /* [nghost]                                */
  const int *ghost_ncomp;       /* [nghost] <= 4                           */
  const int *ghost_comp;        /* [nghost*4]                              */
  const double *ghost_coef;     /* [nghost*4]                              */
} pmx_system;

/* Simulation settings (PINY input keywords in brackets). */
typedef struct pmx_params {
  double hmat[9];       /* cell, column-major like PINY hmat[1..9], Bohr   */
  int pi_beads;         /* [path_int_beads] P                              */
  double temperature;   /* [temperature] K, for spring constants           */
  int pimd_typ;         /* [path_int_md_typ] PMX_PIMD_*                    */
  int nsplin;           /* [inter_spline_pts]                              */
  int ishift;           /* [shift_inter_pe] PMX_SHIFT_*                    */
  double rheal;         /* [respa_rheal] switching length, Bohr            */
  double skin;          /* [verlist_skin] Bohr                             */
  double ewald_alpha;   /* [ewald_alpha] dimensionless alpha * V^(1/3)     */
  int kmax_ewd;         /* [ewald_kmax]                                    */
  int pme_on;           /* [ewald_pme_opt] 1 on (0: standard Ewald)        */
  int kmax_pme;         /* [ewald_kmax_pme]                                */
  int n_interp;         /* [ewald_interp_pme] B-spline order               */
  int intra_perds;      /* [intra_perds] 1: minimum image on bonded terms  */
  int precision;        /* PMX_PREC_*                                      */
  int pair_eval;        /* PMX_PAIR_* (FP64 only)                          */
} pmx_params;

/* Settings of the PIMD NVT integrator (int_NVT_pimd with RESPA off; PINY
 * input keywords in brackets).  The centroid mode (bead 1 with staging) is
 * coupled to one Nose-Hoover chain over every real site (mol_opt_nhc
 * global); every other mode's every degree of freedom to a chain of its own
 * (set_atm_NHC.c:330-790). */
typedef struct pmx_nhc_params {
  double dt;            /* [time_step] in atomic time units (fs / TIME_CONV) */
  double tau_nhc;       /* [atm_nhc_tau_def] in atomic time units           */
  int len_nhc;          /* [atm_nhc_len] chain length, 1..7                 */
  int nres_nhc;         /* [respa_steps_nhc]                                */
  int nyosh_nhc;        /* [yosh_steps_nhc] 1, 3 or 5                       */
  double gamma_adb;     /* [path_int_gamma_adb] adiabaticity (mass factor) */
} pmx_nhc_params;

/* RESPA settings of the force call (PINY input keywords in brackets;
 * set_sim_params.c:1585-1660).  A steps keyword left out of the input is 0
 * here, which turns that split off (its step count is then 1).  The steps
 * fix the force weights of energy_control_initial.c:212-243: with
 * PMX_EVAL_PIMD (pimd_on) and n_x = max(respa_steps_x, 1)
 *     wght_ter     = n_pimd n_intra n_torsion n_lrf   (full inter forces)
 *     wght_ter_res = n_pimd n_intra n_torsion         (short-range part)
 *     wght_tra_res = n_pimd                           (bonds, bends)
 * and without it the same with n_pimd = 1 and wght_tra_res = 1.  pext,
 * pten_inter_guess and nfree enter the RESPA terms of the pressure tensor
 * (energy_control_inter_real.c:206-318). */
typedef struct pmx_respa {
  int respa_steps_lrf;     /* [respa_steps_lrf] > 0: the inter split on     */
  int respa_steps_torsion; /* [respa_steps_torsion]                         */
  int respa_steps_intra;   /* [respa_steps_intra]                           */
  int respa_steps_pimd;    /* [respa_steps_pimd]                            */
  int nfree;               /* classical degrees of freedom, clatoms_info.nfree
                              (close_intra_params.c:222-247)                */
  double pext;             /* [pressure] in Hartree/Bohr^3 (atm * PCONV)    */
  double pten_inter_guess; /* [pten_inter_respa] in Hartree/Bohr^3         */
} pmx_respa;

/* Energies (Hartree) and pressure tensors (Hartree), stat_avg / ptens. */
typedef struct pmx_energies {
  double vintert;       /* inter: vreal + vlong + vrecip+vself+vbgr+vecor  */
  double vintrat;       /* intra: vbond + vbend                            */
  double vreal, vvdw, vcoul, vlong;
  double vrecip;        /* k-space sum only                                */
  double vself, vbgr, vecor;
  double vbond, vbend;
  double kin_harm;      /* spring energy (mode_energy_control)             */
  double pi_ke_prim;    /* primitive kinetic estimator (mode forces only)  */
  double pi_ke_vir;     /* virial kinetic estimator (PIMD)                 */
  double pvten[9];      /* weighted pressure*volume tensor                 */
  double pvten_tot[9];
  double spread;        /* PIMD bead spread used for the list skin         */
  int nver;             /* Verlet-list pairs                               */
  int list_rebuilt;     /* 1 if this call rebuilt the list                 */
} pmx_energies;

typedef struct pmx_ctx pmx_ctx;

/* ABI version of the loaded library (== PMX_ABI_VERSION). */
int pmx_abi_version(void);

/* Number of HIP devices visible (0 when none). */
int pmx_device_count(void);

/*
 * Builds a context: PINY's setup (control_inter_params spline tables,
 * control_set_cp_ewald k-vectors / PME grid / B-spline weights,
 * set_exclude exclusions and Ewald-correction list, splin_ecor table,
 * path_integral_init spring constants), then device allocation.
 * init_x/y/z ([P*N], may be NULL) are the initial positions used only to
 * size the excluded-pair table as splin_ecor does (set_exclude.c:600-622).
 * Returns NULL on failure and writes a message into err (if not NULL).
 */
pmx_ctx *pmx_create(const pmx_system *sys, const pmx_params *par,
                    const double *init_x, const double *init_y,
                    const double *init_z, int device, char *err,
                    size_t errlen);

void pmx_destroy(pmx_ctx *ctx);

const char *pmx_last_error(const pmx_ctx *ctx);

/* Positions of all beads, bead-major [P*N], from host memory.  Ghost
 * positions are recomputed on the device (get_ghost_pos). */
int pmx_set_positions(pmx_ctx *ctx, const double *x, const double *y,
                      const double *z);

/* Same, from device memory (three [P*N] arrays on the context's device). */
int pmx_set_positions_device(pmx_ctx *ctx, const double *x, const double *y,
                             const double *z);

/* Mode-1 coordinates xmod ([N]) used by the PIMD minimum-image convention
 * (period.c:121-206).  If never set, the bead centroid is used. */
int pmx_set_xmod(pmx_ctx *ctx, const double *x, const double *y,
                 const double *z);

/* RESPA settings for the context's later force calls (NULL: off, all
 * weights 1).  As parse() insists (set_sim_params.c:3346-3370), the inter
 * split needs the intramolecular and torsion splits: respa_steps_lrf > 0
 * with respa_steps_intra or respa_steps_torsion 0 returns PMX_ERR_ARG.
 * The reference forms the virial estimator pi_ke_vir and pvten_tot from the
 * unweighted forces fxt of a full call: under weights other than 1,
 * pmx_energy_control makes such a call twice, first without weights (for
 * fxt, pi_ke_vir and pvten_tot), then with them; a batched call does not,
 * and returns both as NaN.  After a short-range or intramolecular call
 * pmx_get_fxt returns PMX_ERR_STATE.  The cluster pair kernel is not used
 * under RESPA. */
int pmx_set_respa(pmx_ctx *ctx, const pmx_respa *respa);

/* The force call. */
int pmx_energy_control(pmx_ctx *ctx, unsigned flags, pmx_energies *out);

/* The force call without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * inputs, flags and results as pmx_energy_control, with two differences:
 *  - the per-workgroup partial sums of the call (energies, virial, k-space,
 *    bonded and spring sums) live in the library's static device array of
 *    PMX_NOALLOC_SCRATCH_BYTES per device instead of the context's own
 *    buffer; a context whose partial sums need more returns PMX_ERR_RANGE;
 *  - a Verlet list that outgrows its rows returns PMX_ERR_RANGE instead of
 *    growing them (one pmx_energy_control call grows them).
 * Every call on a device shares that device's array: calls that overlap in
 * time on one device (from different host threads, or a batch running
 * meanwhile with this context) race and must be serialized by the caller.
 * The array keeps the last call's contents; each call writes every entry it
 * reads before reading it. */
#define PMX_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int pmx_energy_control_noalloc(pmx_ctx *ctx, unsigned flags, pmx_energies *out);

/* Forces of all beads [P*N] into host memory (ghost forces redistributed). */
int pmx_get_forces(pmx_ctx *ctx, double *fx, double *fy, double *fz);

/* Device pointers of the context's force arrays ([P*N] each). */
int pmx_get_forces_device(pmx_ctx *ctx, const double **fx, const double **fy,
                          const double **fz);

/* Virial-estimator forces fxt of all beads (PIMD only). */
int pmx_get_fxt(pmx_ctx *ctx, double *fx, double *fy, double *fz);

/* Per-bead energies [P*4]: vreal, vrecip(k-sum), vecor, vintra. */
int pmx_get_bead_energies(pmx_ctx *ctx, double *out);

/* Verlet half list in PINY's form (verlist_create.c:373-503): for every
 * site i, nter[i] partners j < i in ascending order, stored at
 * jter[jver_off[i] ...].  Returns the pair count in *nver. */
int pmx_get_verlet_list(pmx_ctx *ctx, int *nter, int *jver_off, int *jter,
                        int cap, int *nver);

/* Setup products for parity checks. */
int pmx_get_nktot(pmx_ctx *ctx, int *nktot, int *ngrid);   /* ngrid[3]   */
int pmx_get_kvectors(pmx_ctx *ctx, int *ka, int *kb, int *kc,
                     double *bweight, int cap);
/* which: 0 cv0, 1 cdv0, 2 cv0_c, 3 cdv0_c for type pair ipair;
 *        4 ecor cv0, 5 ecor cdv0 (ipair ignored). */
int pmx_get_spline(pmx_ctx *ctx, int which, int ipair, double *out, int cap);
int pmx_get_setup_scalars(pmx_ctx *ctx, double *out); /* see setup_scalars */

/* Wall time (ms) of the last pmx_energy_control per stage: [0] list (check,
 * site packing and, when due, the build), [1] real (the pair kernel),
 * [2] recip (the k-space chain, launched before the list decision, on its
 * own stream beside the pair kernel when pmx_autotune finds that faster),
 * [3] intra (everything after the pair kernel: excluded-pair corrections,
 * bonded terms, PIMD finish, ghosts, modes, sums), [4] total. */
int pmx_get_timings(pmx_ctx *ctx, double *ms5);

/* Real-space pairs inside the cutoff evaluated by the last force call,
 * summed over beads (the work unit of the pair kernel). */
int pmx_get_pair_count(pmx_ctx *ctx, double *npair);

/* Velocities of all beads, bead-major [P*N], Bohr per atomic time unit
 * (clatoms_pos[ip].vx); zero after pmx_create. */
int pmx_set_velocities(pmx_ctx *ctx, const double *vx, const double *vy,
                       const double *vz);
int pmx_get_velocities(pmx_ctx *ctx, double *vx, double *vy, double *vz);

/* Current positions, bead-major [P*N], Bohr (ghost sites included). */
int pmx_get_positions(pmx_ctx *ctx, double *x, double *y, double *z);

/* nsteps of velocity Verlet in the NVE ensemble, as int_NVE
 * (md/int_NVE.c:29-125 with int_0_to_dt2_nve and int_dt2_to_dt_nve): per
 * step v += f dt/2 / m, x += v dt, ghost positions, the force call
 * (energy_control), v += f dt/2 / m.  dt in atomic time units (PINY's
 * timeinfo.dt = time_step / TIME_CONV, defines.h:85).  Classical contexts
 * (P = 1) only; forces at the starting positions are computed first when
 * the context has none.  out receives the energies of the last force call,
 * kinetic (if not NULL) 1/2 sum m v^2 at the end (get_tvten). */
int pmx_md_nve(pmx_ctx *ctx, int nsteps, double dt, pmx_energies *out,
               double *kinetic);

/* nsteps of path-integral MD in the NVT ensemble on the device, as
 * int_NVT_pimd (integrate/pimd/int_NVT_pimd.c:36-380) with RESPA off: per
 * step the Nose-Hoover chains and half a kick (int_0_to_dt2_nvt_pimd,
 * int_0_to_dt2_pimd.c:34-196), the drift in mode coordinates with xmod set
 * to mode 1 (control_pimd_trans_mode / _pos), ghost positions, the force
 * call energy_control_pimd with mode forces and springs, half a kick and
 * the chains (int_dt2_to_dt_nvt_pimd, int_dt2_to_dt_pimd.c:31-172).  The
 * velocities of the context (pmx_set_velocities / pmx_get_velocities) are
 * the mode velocities, bead-major [P*N]; positions stay bead Cartesian.  The
 * chain state persists in the context between calls (zero at first;
 * pmx_set_nhc_state / pmx_get_nhc_state).  A call starts with a force call
 * at the current positions.  out receives the energies of the last force
 * call; econs (may be NULL) [4]: mode kinetic energy, the chains' energy,
 * vintert + vintrat + kin_harm, and their sum, the conserved quantity.  The
 * host waits once per step, for the Verlet-list decision.  P > 1. */
int pmx_pimd_nvt(pmx_ctx *ctx, const pmx_nhc_params *nhc, int nsteps, pmx_energies *out, double *econs);

/* Chain state of pmx_pimd_nvt: 2 len doubles for the centroid chain
 * (velocities, then positions), then 2 len per degree of freedom of modes
 * 2..P: mode m, site i, direction d at 2 len + ((m - 2) 3N + 3i + d) 2 len
 * (1-based m; ghost sites' entries unused), as therm_class.v_nhc/x_nhc and
 * therm_bead[m].v_nhc/x_nhc hold them. */
int pmx_set_nhc_state(pmx_ctx *ctx, int len_nhc, const double *state);
int pmx_get_nhc_state(pmx_ctx *ctx, int len_nhc, double *state);

/* New force-field parameters for the context's topology: sys must describe
 * the same sites, types, bonds, bends and ghosts as at pmx_create, with the
 * same charged sites and LJ type pairs; its charges q, LJ inter_sig and
 * inter_eps, bond_c, bond_eq, bend_c and bend_s replace the old ones (the LJ
 * tables, the LJ tail and the self/background sums are rebuilt, as parse()
 * would for an input with these values: control_inter_params.c:643-772,
 * get_clong.c:27-114, ewald3d_self_bgr.c:29-83).  PINY sets these only at
 * parse time; the entry point lets a caller differentiate the energy with
 * respect to them (python/piny/piny.pyx:126-198 builds a new instance per
 * parameter set instead). */
int pmx_set_force_field(pmx_ctx *ctx, const pmx_system *sys);

/* Derivatives of each bead's potential energy
 *     E_b = vreal_b + vrecip_b + vecor_b + vbond_b + vbend_b
 *           + vlong + vself + vbgr
 * (the energy of bead b as a classical configuration; with PMX_EVAL_PIMD
 * the force call's vintert + vintrat is the mean of E_b over beads, and
 * -dE_b/dx_b is P times its bead forces) with respect to the force-field
 * parameters of pmx_system, at the positions and parameters of the last
 * pmx_energy_control (which must have included the k-space term).  This is
 * the backward of the force call; the forces are its position gradient.
 * Each output may be NULL; sizes per bead, bead-major:
 *   energy   [P]                 E_b
 *   dq       [P*natm]            dE_b/dq_i (0 at sites without charge: the
 *                                set of charged sites is fixed)
 *   deps     [P*npair]           dE_b/d inter_eps[t] (LJ tail included)
 *   dsig     [P*npair]           dE_b/d inter_sig[t]
 *   dbond_c  [P*nbond_type*7]    dE_b/d bond_c[ty*7 + m]
 *   dbond_eq [P*nbond_type]      dE_b/d bond_eq[ty]
 *   dbend_c  [P*nbend_type*7]    dE_b/d bend_c[ty*7 + m]
 *   dbend_s  [P*nbend_type*7]    dE_b/d bend_s[ty*7 + m]
 * LJ gradients are formed for up to 4 LJ type pairs.  The reference
 * computes no parameter derivatives; its debug driver checks forces by
 * finite differences (test_energy.c:489-710). */
typedef struct pmx_param_grad {
  double *energy, *dq, *deps, *dsig, *dbond_c, *dbond_eq, *dbend_c, *dbend_s;
} pmx_param_grad;
int pmx_get_param_grad(pmx_ctx *ctx, pmx_param_grad *out);

/* A batch of contexts (independent boxes, each with its own beads) on one
 * device, evaluated together: the list stage of every context, then one
 * launch of a graph that holds every context's force sequence as a
 * parallel branch, then one readback.  PINY evaluates replicas one instance
 * at a time, each on its own host thread (python/piny/piny.pyx:426-436,
 * prange over energy_control); here one call covers them all.  Each
 * context keeps its own state and results (pmx_get_forces etc. work after
 * the call).  flags as for pmx_energy_control; PMX_EVAL_PIMD is implied for
 * contexts with P > 1 and PMX_EVAL_MODE_FORCES is ignored for P = 1.  out
 * receives one pmx_energies per context, in the order given.  The graph is
 * captured on the first call and again when flags change or a context's
 * list grows.  A context must not be used by another call while its batch
 * runs. */
typedef struct pmx_batch pmx_batch;
pmx_batch *pmx_batch_create(pmx_ctx *const *ctx, int n, char *err, size_t errlen);
int pmx_batch_energy_control(pmx_batch *batch, unsigned flags, pmx_energies *out);
const char *pmx_batch_last_error(const pmx_batch *batch);
void pmx_batch_destroy(pmx_batch *batch);

/* Autotuner of the launch shapes (no counterpart in the reference, which
 * has no device): times reps force calls at the current positions for each
 * candidate of the pair kernels' wavefront count, then (mixed precision, a
 * box of one molecule template) the molecule-cluster form of the pair
 * kernel against the per-site one, the PME tile sizes and the
 * list build (plain scan or link cells), keeps the fastest of each, then
 * times whole calls with PME on a second stream beside the pair kernel, for
 * several pair wavefront counts, against the best without, keeps the
 * fastest, and ends with a force call using them.  choice (may be NULL) receives
 * [tile a, tile b, tile c, pair wavefronts, list build 1 plain / 2 cells].
 * Results are unchanged by the shapes up to the order of sums. */
int pmx_autotune(pmx_ctx *ctx, unsigned flags, int reps, int *choice);

/* The autotuner for a batch: the boxes share the device, so candidates are
 * timed as whole batched calls, each applied to every context: the pair
 * kernels' wavefront cap, then the PME tile caps.  choice (may be NULL)
 * receives [pair wavefronts, tile cap a, b, c].  Ends with a batched call. */
int pmx_batch_autotune(pmx_batch *batch, unsigned flags, int reps, int *choice);

/* Coefficient helpers restating intra_coefs.c (units: Hartree, Bohr). */
int pmx_bond_coef_harm(double fk, double *c7);
int pmx_bond_coef_morse(double d0, double alpha, double *c7);
int pmx_bend_coef_harm(double fk, double theta0, double *c7, double *s7);

#ifdef __cplusplus
}
#endif

#endif /* PINY_AMD_H */
