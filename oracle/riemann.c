/* oracle/riemann.c -- placeholder until the exact two-material Riemann restatement lands. */
#include "oracle_internal.h"
#include <math.h>
int orc_interface_flux(orc_ctx *c, int dir, const double *vm, int idm, const double *vp, int idp,
                       long lo, long hi, double *lf1, double *lf2) {
  (void)c; (void)dir; (void)vm; (void)idm; (void)vp; (void)idp; (void)lo; (void)hi;
  for (int p = 0; p < 5; p++) lf1[p] = lf2[p] = NAN;
  return -1;
}
void orc_flux_godunov(const orc_ctx *c, int dir, const double *Vm, const double *Vp, int id, double *F) {
  (void)c; (void)dir; (void)Vm; (void)Vp; (void)id;
  for (int p = 0; p < 5; p++) F[p] = NAN;
}
void orc_riemann_begin_stage(orc_ctx *c) { (void)c; }
void orc_riemann_end_stage(orc_ctx *c) { (void)c; }
