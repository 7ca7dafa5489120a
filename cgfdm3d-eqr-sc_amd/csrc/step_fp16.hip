// Instantiation of the CGFDM step kernels: wavefield float, coefficients __half.
#include "launch_impl.h"

namespace cg {
template int launch_interior<float, __half>(StageArgs<float>, const __half*, int, int, int, hipStream_t);
template int launch_surface<float, __half>(StageArgs<float>, const __half*, SurfArgs<float>, int, int, int, hipStream_t);
}  // namespace cg
