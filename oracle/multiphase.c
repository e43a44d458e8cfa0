/* oracle/multiphase.c -- placeholder until the level-set restatement lands. */
#include "oracle_internal.h"
int orc_step_all(orc_ctx *c, double *V, int *ID, double dt, int step) {
  (void)step; (void)ID;
  return orc_advance(c, V, ID, dt);
}
void orc_ls_free(orc_ctx *c) { (void)c; }
