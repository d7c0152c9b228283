// pic_unet.hip -- learned warm start / preconditioner for the Poisson solve.
#include "pic_internal.h"

void pic_unet_free(pic_ctx* ctx) { (void)ctx; }

int pic_solve_learned(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st) {
    (void)p;
    (void)st;
    return pic_fail(ctx, PIC_ERR_STATE, "no trained network loaded");
}
