// nait.hip -- DensityExpression 5 (nAiT) on gfx950.
#include "solver.h"

struct NaitTables {
    int dummy;
};

int nait_init(mpdpft_ctx *c) { return ctx_fail(c, MPDPFT_EINVAL, "nAiT is not built yet"); }
void nait_free(mpdpft_ctx *c) {
    delete c->nait;
    c->nait = nullptr;
}
int nait_density(mpdpft_ctx *c, int, double, double *, double *) { return ctx_fail(c, MPDPFT_EINVAL, "nAiT"); }
int nait_dispersal_energy(mpdpft_ctx *c, int, double *) { return ctx_fail(c, MPDPFT_EINVAL, "nAiT"); }
