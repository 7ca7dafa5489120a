// Instantiation of the CGFDM step kernels: wavefield double, coefficients double.
#include "launch_impl.h"

namespace cg {
template int launch_interior<double, double>(StageArgs<double>, const double*, int, int, int, hipStream_t);
template int launch_surface<double, double>(StageArgs<double>, const double*, SurfArgs<double>, int, int, int, hipStream_t);
}  // namespace cg

namespace cg {
template void launch_source<double>(double*, long long, int, const long long*, const double*, int, double, double, hipStream_t);
template void launch_record<double>(const double*, long long, int, const long long*, float*, int, int, hipStream_t);
}  // namespace cg
