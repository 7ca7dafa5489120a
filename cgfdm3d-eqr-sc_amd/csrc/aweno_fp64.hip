// FP64 instantiation of the ME-AWENO step kernels.
#include "aweno_impl.h"

namespace aw {
template void launch_faces<double>(const StageArgs<double>&, hipStream_t);
template void launch_update<double>(const StageArgs<double>&, int, int, hipStream_t, hipStream_t);
template void launch_char_surface<double>(double*, const double*, int, int, int, int, long long, long long,
                                          hipStream_t);
template void launch_sponge<double>(double*, int, int, int, int, long long, long long, int, hipStream_t);
template void launch_source<double>(double*, const double*, long long, int, const long long*, const double*, int,
                                    double, double, hipStream_t);
template void launch_surface_obs<double>(const double*, const double*, float*, double*, int, int, int, int, long long,
                                         long long, double, hipStream_t);
template void launch_record<double>(const double*, const double*, long long, int, const long long*, float*, int, int,
                                    hipStream_t);
}  // namespace aw
