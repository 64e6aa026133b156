// Batched path: every shot of the job in each launch.
#include "fd_common.h"
#include "fd_kernels.h"

namespace swit {

// Per-plan state of the batched path.
struct BatchedState {
    int ready = 0;
};

}  // namespace swit

struct swit_plan;
static int batched_init(swit_plan *pl);
static void batched_release(swit_plan *pl);
static int batched_set_model(swit_plan *pl, hipStream_t st);
static int batched_forward(swit_plan *pl, const float *src, int save, hipStream_t st);
static int batched_adjoint(swit_plan *pl, const float *src, double *grad, double *fw, double *bk,
                           hipStream_t st);
