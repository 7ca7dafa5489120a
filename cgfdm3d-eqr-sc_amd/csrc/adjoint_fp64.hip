// FP64 instantiation of the CGFDM adjoint kernels.
#include "adjoint_impl.h"

namespace cg {
template int launch_adj_stage<double, double>(AdjArgs<double>, const double*, const double*, const double*, const double*, int, int, int, int,
                                        hipStream_t);
template void launch_adj_inject<double>(double*, long long, int, const long long*, const double*, int, int, hipStream_t);
template void launch_adj_source<double>(const double*, long long, int, const long long*, double*, int, double, double,
                                     hipStream_t);
template void launch_record_vel<double>(const double*, long long, int, const long long*, double*, int, int, hipStream_t);
}  // namespace cg
