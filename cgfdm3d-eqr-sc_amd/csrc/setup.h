// Setup-kernel launchers (FP64 metrics, packing, surface matrices, CFL ranges).
#pragma once
#include "common.h"

namespace cg {

void launch_metrics(const float* coord, double* met, int nx, int ny, int nz, int NX, int NY, long long N, double dh,
                    bool fs, hipStream_t st);
template <class R, class TM>
void launch_pack(const double* met, const float* med, TM* coef, R* jac, long long N, double stress_scale,
                 hipStream_t st);
// 16-bit coefficient storage: affine table aff[12][NZ][2] and the six pair-word
// planes (see Bits16 in common.h), J as float, and mu / Z as float for the
// moment-rate scaling of the source cells.
template <class TM>
void launch_pack16(const double* met, const float* med, float* aff, unsigned* words, float* jac, float* mu_src, int NX,
                   int NY, int NZ, double stress_scale, hipStream_t st);
template <class R>
void launch_rdz(const double* met, const float* med, R* rdzdx, R* rdzdy, int nx, int ny, int nz, int NX, int NY,
                long long N, hipStream_t st);
void launch_cov_cmm(const float* coord, double* cov, int NX, int NY, int NZ, long long N, double dh, hipStream_t st);
template <class R>
void launch_pack_cmm(const double* cov, const float* med, R* cjm, int NX, int NY, int NZ, long long N, hipStream_t st);
void launch_hrange(const float* coord, double* part, int nblocks, int nx, int ny, int nz, int NX, int NY, long long N,
                   hipStream_t st);
void launch_frange(const float* a, long long n, double* part, int nblocks, hipStream_t st);

// Allocation-free CFL ranges of a device-resident grid: per-block partials in a
// library-owned 64 KiB device array (RANGE_BLOCKS pairs per reduction), the
// eight results written to the device array out8. Asynchronous on st.
constexpr int RANGE_BLOCKS = 1024;
constexpr int kMaxDevices = 64;
hipError_t grid_ranges_noalloc(int nx, int ny, int nz, const float* coord, const float* med, double* out8,
                               hipStream_t st);

}  // namespace cg
