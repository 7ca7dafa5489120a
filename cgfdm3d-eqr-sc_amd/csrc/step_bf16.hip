// Instantiation of the CGFDM step kernels: wavefield float, coefficients __hip_bfloat16.
#include "launch_impl.h"

namespace cg {
template int launch_interior<float, __hip_bfloat16>(StageArgs<float>, const __hip_bfloat16*, int, int, int, hipStream_t);
template int launch_surface<float, __hip_bfloat16>(StageArgs<float>, const __hip_bfloat16*, SurfArgs<float>, int, int, int, hipStream_t);
}  // namespace cg
