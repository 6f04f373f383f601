// gp_nn.hip -- learned warm start / preconditioner for the elliptic solves.
// (network kernels are added in a later unit; until then every entry point
// reports GP_ERR_UNSUPPORTED so a caller cannot mistake it for a result)
#include "gp_internal.h"

struct gp_nn {
    int dummy;
};

int gp_nn_create(gp_ctx *c, int, uint64_t) { (void)c; return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
void gp_nn_destroy(gp_ctx *c) { if (c->nn) { delete c->nn; c->nn = nullptr; } }
int gp_nn_train_impl(gp_ctx *, const double *, const double *, int, int, int, float, float *)
{
    return gp_fail(GP_ERR_UNSUPPORTED, "network not built");
}
int gp_nn_apply(gp_ctx *, gp_solver *, const float *, float *) { return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
int gp_nn_infer_impl(gp_ctx *, const double *, double *) { return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
int gp_nn_save_impl(gp_ctx *, const char *) { return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
int gp_nn_load_impl(gp_ctx *, const char *) { return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
int gp_nn_precorrect_cb(void *, float *, const float *) { return gp_fail(GP_ERR_UNSUPPORTED, "network not built"); }
