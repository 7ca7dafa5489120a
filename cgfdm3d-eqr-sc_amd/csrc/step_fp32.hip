// Instantiation of the CGFDM step kernels: wavefield float, coefficients float.
#include "launch_impl.h"

namespace cg {
template int launch_interior<float, float>(StageArgs<float>, const float*, int, int, int, hipStream_t);
template int launch_surface<float, float>(StageArgs<float>, const float*, SurfArgs<float>, int, int, int, hipStream_t);
}  // namespace cg

namespace cg {
template void launch_source<float>(float*, long long, int, const long long*, const float*, int, float, float, hipStream_t);
template void launch_record<float>(const float*, long long, int, const long long*, float*, int, int, hipStream_t);
}  // namespace cg
