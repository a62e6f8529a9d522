/*
 * pmx_piny.h -- the PINY side of the drop-in: fills the library's
 * pmx_system / pmx_params from the structures PINY's parse() leaves, and
 * makes the force call on PINY's own arrays (the body that replaces
 * energy_control / energy_control_pimd, energy_control.c:33-77 and
 * energy_control_pimd.c:37-278).  It is compiled inside a PINY build,
 * against PINY's headers (-I<piny>/include/<arch>) and include/piny_amd.h.
 */
#ifndef PMX_PINY_H
#define PMX_PINY_H

#include "piny_amd.h"

/* What parse() reads but does not keep: control_inter_params.c:447-470
 * reads each type pair's LJ parameters into arrays that :353-370 frees, and
 * the k-space and shift keywords stay in CLASS_PARSE (typedefs_par.h:26-28).
 * The caller copies them before they go: per type pair in PINY's order
 * (control_inter_params.c:118-126), 0-based. */
typedef struct pmx_piny_parse {
  const int *inter_label;   /* PMX_INTER_LJ or PMX_INTER_NULL             */
  const double *inter_sig;  /* Bohr                                       */
  const double *inter_eps;  /* Hartree                                    */
  int kmax_ewd;             /* class_parse.kmax_ewd                       */
  int ishift;               /* class_parse.ishift_pot (PMX_SHIFT_*)       */
} pmx_piny_parse;

/* The filled descriptions and the 0-based copies they point into. */
typedef struct pmx_piny_sys {
  pmx_system sys;
  pmx_params par;
  void *store;  /* owned copies */
} pmx_piny_sys;

/* Fills out from PINY's structures (flat 0-based copies, PINY units).
 * Returns PMX_OK or PMX_ERR_ARG (a topology the library does not take:
 * more than 4 components of a ghost site). */
int pmx_piny_fill(CLASS *class, BONDED *bonded, GENERAL_DATA *gd, const pmx_piny_parse *pp, pmx_piny_sys *out);
void pmx_piny_free(pmx_piny_sys *s);

/* The force call on PINY's arrays: bead positions and (with beads) xmod
 * in; forces (mode forces with springs when flags has
 * PMX_EVAL_MODE_FORCES), fxt, the stat_avg energies and the pressure
 * tensors out, where the reference's energy_control (P = 1) or
 * energy_control_pimd leaves them.  s is the pmx_piny_fill result the
 * context was created from (its scratch holds the flat copies). */
int pmx_piny_energy_control(pmx_ctx *ctx, pmx_piny_sys *s, CLASS *class, GENERAL_DATA *gd, unsigned flags,
                            pmx_energies *out);

#endif
