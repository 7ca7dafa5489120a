// FP32 instantiation of the CGFDM adjoint kernels.
#include "adjoint_impl.h"

namespace cg {
template int launch_adj_stage<float, float>(AdjArgs<float>, const float*, const float*, const float*, const float*, int, int, int, int,
                                        hipStream_t);
template void launch_adj_inject<float>(float*, long long, int, const long long*, const float*, int, int, hipStream_t);
template void launch_adj_source<float>(const float*, long long, int, const long long*, double*, int, double, double,
                                     hipStream_t);
template void launch_record_vel<float>(const float*, long long, int, const long long*, float*, int, int, hipStream_t);
}  // namespace cg
