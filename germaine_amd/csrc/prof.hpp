// Optional per-kernel-class timing with HIP events recorded on the launch
// stream (off unless enabled through gfdfd_profile), with the algorithmic
// flops and bytes of each launch attached by the caller.
#pragma once
#include <hip/hip_runtime.h>

namespace gfdfd {

enum KClass {
    K_ASSEMBLE, K_FRONT_ASSEMBLE, K_SCATTER, K_SMALL_LU, K_PANEL, K_PANEL_APPLY, K_TRSM, K_GEMM, K_DIAG_INV,
    K_FWD_GATHER, K_FWD_LEAF, K_FWD_SMALL1, K_FWD_FRONT1, K_FWD_FRONT, K_FWD_BLK,
    K_BWD_SMALL1, K_BWD_FRONT1, K_BWD_FRONT, K_BWD_CONTRIB, K_BWD_BLK, K_PIV_UNSWAP,
    K_FWD_VEC, K_BWD_VEC, K_PERMUTE_IN, K_PERMUTE_OUT, K_RESIDUAL, K_NORM_REDUCE, K_AXPY, K_SUBTREE, K_NCLASS
};

extern const char *const kClassNames[K_NCLASS];

bool prof_enabled();

// While alive, this thread's launches record no events (graph capture).
struct ProfPause {
    ProfPause();
    ~ProfPause();
    ProfPause(const ProfPause &) = delete;
    ProfPause &operator=(const ProfPause &) = delete;
};

// Records a start event on construction and an end event on destruction
// (both on `st`) when profiling is enabled; otherwise does nothing.
class ProfScope {
  public:
    ProfScope(KClass c, hipStream_t st, double flops, double bytes);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

  private:
    int slot_ = -1;
    hipStream_t st_;
};

}  // namespace gfdfd
