/* placeholder: replaced by the geometric multigrid */
#include "nusif_internal.h"
struct GmgState { int dummy; };
GmgState *gmgCreate(NusifSolverCtx *sc, const Parameter *prm) { (void)sc; (void)prm; return new GmgState(); }
double gmgSolve(NusifSolverCtx *sc, double *p, const double *rhs) { (void)sc; (void)p; (void)rhs; NUSIF_FATAL("gmg not built"); return 0; }
void gmgFree(GmgState *g) { delete g; }
struct UnetModel { int dummy; };
void unetFree(UnetModel *m) { delete m; }
