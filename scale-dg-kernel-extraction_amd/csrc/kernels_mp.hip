// Mixed-precision stage kernel (FP32 MFMA contractions and Lift, FP64 state).
#include "dgk_internal.hpp"

namespace dgk {

int launch_stage_mp(int, const StageArgsMP &, hipStream_t) {
  return set_error(DGK_ERR_UNSUPPORTED, "HIP_MP path not built yet");
}

}  // namespace dgk
