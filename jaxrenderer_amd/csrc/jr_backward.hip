// jr_backward.hip -- reverse-mode pass (placeholder until the kernels land).
#include <hip/hip_runtime.h>
#include "../../include/jrhip.h"

extern "C" int32_t jr_backward(const jr_frame*, const jr_inputs*, const jr_scene*,
                               const float*, const float*, const int32_t*,
                               const float*, jr_grads*, jr_stream_t) {
  return JR_E_UNSUPPORTED;
}
