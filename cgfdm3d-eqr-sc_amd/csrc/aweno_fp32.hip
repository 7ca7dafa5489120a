// FP32 instantiation of the ME-AWENO step kernels.
#include "aweno_impl.h"

namespace aw {
template void launch_faces<float>(const StageArgs<float>&, hipStream_t);
template void launch_update<float>(const StageArgs<float>&, int, int, hipStream_t, hipStream_t);
template void launch_char_surface<float>(float*, const float*, int, int, int, int, long long, long long, hipStream_t);
template void launch_sponge<float>(float*, int, int, int, int, long long, long long, int, hipStream_t);
template void launch_source<float>(float*, const float*, long long, int, const long long*, const float*, int, float,
                                   float, hipStream_t);
template void launch_surface_obs<float>(const float*, const float*, float*, float*, int, int, int, int, long long,
                                        long long, float, hipStream_t);
template void launch_record<float>(const float*, const float*, long long, int, const long long*, float*, int, int,
                                   hipStream_t);
}  // namespace aw
