// Learned preconditioner for the pressure CG (placeholder until the U-Net kernels land):
// selecting it reports an error instead of silently running the plain solver.
#include <hip/hip_runtime.h>

#include "efs_internal.h"

int efs_learned_cg_solve(efs_sim* s, efs_stats* st) {
    (void)s;
    (void)st;
    efs_set_error("learned preconditioner: no trained network loaded");
    return -1;
}

void efs_learned_free(efs_sim* s) { (void)s; }
