// Setup kernels: curvilinear metrics, Jacobian, free-surface matrices,
// coefficient packing and the CFL grid-spacing reduction, all in FP64 on the
// GPU (the stored coefficients are then rounded to the storage type).
#include <mutex>

#include "common.h"
#include "setup.h"

namespace cg {

__device__ __forceinline__ double lc(const float* c, long long p, long long s, int fb) {
  long long d = fb * s;
  return fb * (AF_1 * (double)c[p - d] + AF0 * (double)c[p] + AF1 * (double)c[p + d] + AF2 * (double)c[p + 2 * d] +
               AF3 * (double)c[p + 3 * d]);
}

// Covariant derivatives as the mean of the forward and backward MacCormack
// differences of the coordinates, J = det, contravariant = cofactor / J
// (reference contravariant.cpp:182-285). Output: 10 planes (xi_x..zt_z, J).
__global__ void k_metrics(const float* __restrict__ coord, double* __restrict__ met, int nx, int ny, int nz, int NX,
                          int NY, long long N, double rdh) {
  int i = H + blockIdx.x * blockDim.x + threadIdx.x;
  int j = H + blockIdx.y;
  int k = H + blockIdx.z;
  if (i >= H + nx) return;
  const long long S = (long long)NX * NY;
  long long p = (long long)i + (long long)NX * j + (long long)k * S;
  const float* X = coord;
  const float* Y = coord + N;
  const float* Z = coord + 2 * N;
  double x_xi = 0.5 * (lc(X, p, 1, 1) + lc(X, p, 1, -1)) * rdh;
  double x_et = 0.5 * (lc(X, p, NX, 1) + lc(X, p, NX, -1)) * rdh;
  double x_zt = 0.5 * (lc(X, p, S, 1) + lc(X, p, S, -1)) * rdh;
  double y_xi = 0.5 * (lc(Y, p, 1, 1) + lc(Y, p, 1, -1)) * rdh;
  double y_et = 0.5 * (lc(Y, p, NX, 1) + lc(Y, p, NX, -1)) * rdh;
  double y_zt = 0.5 * (lc(Y, p, S, 1) + lc(Y, p, S, -1)) * rdh;
  double z_xi = 0.5 * (lc(Z, p, 1, 1) + lc(Z, p, 1, -1)) * rdh;
  double z_et = 0.5 * (lc(Z, p, NX, 1) + lc(Z, p, NX, -1)) * rdh;
  double z_zt = 0.5 * (lc(Z, p, S, 1) + lc(Z, p, S, -1)) * rdh;
  double J = x_xi * y_et * z_zt + x_et * y_zt * z_xi + x_zt * y_xi * z_et - x_zt * y_et * z_xi - x_et * y_xi * z_zt -
             x_xi * z_et * y_zt;
  double r = 1.0 / J;
  met[0 * N + p] = (y_et * z_zt - y_zt * z_et) * r;
  met[1 * N + p] = (x_zt * z_et - x_et * z_zt) * r;
  met[2 * N + p] = (x_et * y_zt - x_zt * y_et) * r;
  met[3 * N + p] = (y_zt * z_xi - y_xi * z_zt) * r;
  met[4 * N + p] = (x_xi * z_zt - x_zt * z_xi) * r;
  met[5 * N + p] = (x_zt * y_xi - x_xi * y_zt) * r;
  met[6 * N + p] = (y_xi * z_et - y_et * z_xi) * r;
  met[7 * N + p] = (x_et * z_xi - x_xi * z_et) * r;
  met[8 * N + p] = (x_xi * y_et - x_et * y_xi) * r;
  met[9 * N + p] = J;
}

// Mirror metrics and J into the three ghost layers above the free surface
// (contravariant.cpp:449-463).
__global__ void k_mirror_top(double* met, int nx, int ny, int nz, int NX, int NY, long long N) {
  int i = H + blockIdx.x * blockDim.x + threadIdx.x;
  int j = H + blockIdx.y;
  int l = blockIdx.z;  // 0..2
  if (i >= H + nx) return;
  const long long S = (long long)NX * NY;
  int ks = H + nz - 1;
  int k = ks + 1 + l;
  long long p = (long long)i + (long long)NX * j + (long long)k * S;
  long long q = (long long)i + (long long)NX * j + (long long)(2 * ks - k) * S;
  for (int c = 0; c < 10; ++c) met[c * N + p] = met[c * N + q];
}

template <class R, class TM>
__global__ void k_pack(const double* __restrict__ met, const float* __restrict__ med, TM* __restrict__ coef,
                       R* __restrict__ jac, long long N, double stress_scale) {
  // (Vs, Vp, rho) -> (mu, lambda, 1/rho) (CJM.cpp:53-75) packed with the metrics.
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  for (int c = 0; c < 9; ++c) coef[c * N + p] = (TM)(float)met[c * N + p];
  double vs = med[p], vp = med[N + p], rho = med[2 * N + p];
  double mu = rho * vs * vs, lam = rho * (vp * vp - 2.0 * vs * vs), b = 1.0 / rho;
  coef[M_MU * N + p] = (TM)(float)(mu / stress_scale);
  coef[M_LAM * N + p] = (TM)(float)(lam / stress_scale);
  coef[M_B * N + p] = (TM)(float)(b * stress_scale);
  jac[p] = (R)met[9 * N + p];
}

template <>
__global__ void k_pack<double, double>(const double* __restrict__ met, const float* __restrict__ med,
                                       double* __restrict__ coef, double* __restrict__ jac, long long N,
                                       double stress_scale) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  for (int c = 0; c < 9; ++c) coef[c * N + p] = met[c * N + p];
  double vs = med[p], vp = med[N + p], rho = med[2 * N + p];
  coef[M_MU * N + p] = rho * vs * vs / stress_scale;
  coef[M_LAM * N + p] = rho * (vp * vp - 2.0 * vs * vs) / stress_scale;
  coef[M_B * N + p] = (1.0 / rho) * stress_scale;
  jac[p] = met[9 * N + p];
}

// Coefficient c of cell p before storage: metrics c < 9, then mu / Z, lambda / Z
// and Z / rho (stress carried as T / Z).
__device__ __forceinline__ double coef_value(const double* __restrict__ met, const float* __restrict__ med, long long N,
                                             long long p, int c, double Z) {
  if (c < 9) return met[c * N + p];
  const double vs = med[p], vp = med[N + p], rho = med[2 * N + p];
  if (c == M_MU) return rho * vs * vs / Z;
  if (c == M_LAM) return rho * (vp * vp - 2.0 * vs * vs) / Z;
  return Z / rho;
}

// Affine table of the 16-bit storage: block (k, c) reduces plane c over layer k
// to (mid, half) = centre and half-width of its value range (FP64), rounded
// to float.
__global__ void k_coef_range(const double* __restrict__ met, const float* __restrict__ med, float* __restrict__ aff,
                             long long S, long long N, int NZ, double Z) {
  __shared__ double smin[256], smax[256];
  const int k = blockIdx.x, c = blockIdx.y;
  double lo = 1e300, hi = -1e300;
  for (long long e = threadIdx.x; e < S; e += blockDim.x) {
    const double v = coef_value(met, med, N, (long long)k * S + e, c, Z);
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    aff[2 * ((long long)c * NZ + k)] = (float)(0.5 * (smin[0] + smax[0]));
    aff[2 * ((long long)c * NZ + k) + 1] = (float)(0.5 * (smax[0] - smin[0]));
  }
}

template <class TM> __device__ __forceinline__ unsigned short bits16(float x);
template <> __device__ __forceinline__ unsigned short bits16<__half>(float x) {
  return __half_as_ushort(__float2half_rn(x));
}
template <> __device__ __forceinline__ unsigned short bits16<__hip_bfloat16>(float x) {
  const __hip_bfloat16 h = __float2bfloat16(x);
  return *reinterpret_cast<const unsigned short*>(&h);
}

// 16-bit pair words: word plane q holds the affine codes of planes 2q (low
// half) and 2q + 1 (high half); J in the wavefield precision.
template <class TM>
__global__ void k_pack16(const double* __restrict__ met, const float* __restrict__ med, const float* __restrict__ aff,
                         unsigned* __restrict__ words, float* __restrict__ jac, float* __restrict__ mu_src, int NX,
                         long long S, long long N, int NZ, double Z) {
  // block (x, j, k): cells i = 256 x + thread of row j in layer k
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
  if (i >= NX) return;
  const long long p = (long long)k * S + (long long)j * NX + i;
  for (int q = 0; q < M_NCOEF / 2; ++q) {
    unsigned short b[2];
    for (int h = 0; h < 2; ++h) {
      const int c = 2 * q + h;
      const double mid = aff[2 * ((long long)c * NZ + k)], half = aff[2 * ((long long)c * NZ + k) + 1];
      const double v = coef_value(met, med, N, p, c, Z);
      b[h] = bits16<TM>(half > 0.0 ? (float)((v - mid) / half) : 0.0f);
    }
    words[q * N + p] = (unsigned)b[0] | ((unsigned)b[1] << 16);
  }
  jac[p] = (float)met[9 * N + p];
  mu_src[p] = (float)coef_value(met, med, N, p, M_MU, Z);
}

// Traction-free condition: dV/dzeta = -DZ^-1 (DX dV/dxi + DY dV/deta) on the
// surface (contravariant.cpp:363-447), with the 3x3 inverse in FP64.
template <class R>
__global__ void k_rdz(const double* __restrict__ met, const float* __restrict__ med, R* rdzdx, R* rdzdy, int nx,
                      int ny, int nz, int NX, int NY, long long N) {
  int i = H + blockIdx.x * blockDim.x + threadIdx.x;
  int j = H + blockIdx.y;
  if (i >= H + nx) return;
  const long long S = (long long)NX * NY;
  int k = H + nz - 1;
  long long p = (long long)i + (long long)NX * j + (long long)k * S;
  double xx = met[0 * N + p], xy = met[1 * N + p], xz = met[2 * N + p];
  double ex = met[3 * N + p], ey = met[4 * N + p], ez = met[5 * N + p];
  double zx = met[6 * N + p], zy = met[7 * N + p], zz = met[8 * N + p];
  double vs = med[p], vp = med[N + p], rho = med[2 * N + p];
  double mu = rho * vs * vs, lam = rho * (vp * vp - 2.0 * vs * vs);
  double chi = 2.0 * lam + mu, lm = lam + mu;
  double DZ[9], DX[9], DY[9], Ri[9];
  DZ[0] = chi * zx * zx + mu * (zy * zy + zz * zz);
  DZ[1] = lm * zx * zy;
  DZ[2] = lm * zx * zz;
  DZ[3] = DZ[1];
  DZ[4] = chi * zy * zy + mu * (zx * zx + zz * zz);
  DZ[5] = lm * zy * zz;
  DZ[6] = DZ[2];
  DZ[7] = DZ[5];
  DZ[8] = chi * zz * zz + mu * (zx * zx + zy * zy);
  const double ax[2] = {xx, ex}, ay[2] = {xy, ey}, az[2] = {xz, ez};
  double* Dm[2] = {DX, DY};
  for (int t = 0; t < 2; ++t) {
    double* D = Dm[t];
    double a1 = ax[t], a2 = ay[t], a3 = az[t];
    D[0] = chi * zx * a1 + mu * (zy * a2 + zz * a3);
    D[1] = lam * zx * a2 + mu * zy * a1;
    D[2] = lam * zx * a3 + mu * zz * a1;
    D[3] = lam * zy * a1 + mu * zx * a2;
    D[4] = chi * zy * a2 + mu * (zx * a1 + zz * a3);
    D[5] = lam * zy * a3 + mu * zz * a2;
    D[6] = lam * zz * a1 + mu * zx * a3;
    D[7] = lam * zz * a2 + mu * zy * a3;
    D[8] = chi * zz * a3 + mu * (zx * a1 + zy * a2);
  }
  double det = DZ[0] * DZ[4] * DZ[8] + DZ[1] * DZ[5] * DZ[6] + DZ[2] * DZ[7] * DZ[3] - DZ[2] * DZ[4] * DZ[6] -
               DZ[1] * DZ[3] * DZ[8] - DZ[0] * DZ[7] * DZ[5];
  Ri[0] = (DZ[4] * DZ[8] - DZ[5] * DZ[7]) / det;
  Ri[1] = (-DZ[3] * DZ[8] + DZ[5] * DZ[6]) / det;
  Ri[2] = (DZ[3] * DZ[7] - DZ[4] * DZ[6]) / det;
  Ri[3] = (-DZ[1] * DZ[8] + DZ[2] * DZ[7]) / det;
  Ri[4] = (DZ[0] * DZ[8] - DZ[2] * DZ[6]) / det;
  Ri[5] = (-DZ[0] * DZ[7] + DZ[1] * DZ[6]) / det;
  Ri[6] = (DZ[1] * DZ[5] - DZ[2] * DZ[4]) / det;
  Ri[7] = (-DZ[0] * DZ[5] + DZ[2] * DZ[3]) / det;
  Ri[8] = (DZ[0] * DZ[4] - DZ[1] * DZ[3]) / det;
  long long s = (long long)i + (long long)NX * j;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      rdzdx[(3 * r + c) * S + s] = (R)(-(Ri[3 * r] * DX[c] + Ri[3 * r + 1] * DX[3 + c] + Ri[3 * r + 2] * DX[6 + c]));
      rdzdy[(3 * r + c) * S + s] = (R)(-(Ri[3 * r] * DY[c] + Ri[3 * r + 1] * DY[3 + c] + Ri[3 * r + 2] * DY[6 + c]));
    }
}

// SCFDM metrics (contravariant.cpp:81-177): covariant basis by second-order
// central differences and J = det on [1, n-1) of every padded axis. Output:
// 10 FP64 planes (x_xi, y_xi, z_xi, x_et, y_et, z_et, x_zt, y_zt, z_zt, J).
__global__ void k_cov_cmm(const float* __restrict__ coord, double* __restrict__ cov, int NX, int NY, int NZ,
                          long long N, double rdh) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  int j = 1 + blockIdx.y;
  int k = 1 + blockIdx.z;
  if (i >= NX - 1) return;
  const long long S = (long long)NX * NY;
  const long long p = (long long)i + (long long)NX * j + (long long)k * S;
  const long long st[3] = {1, NX, S};
  double d[9];
  for (int dir = 0; dir < 3; ++dir)
    for (int a = 0; a < 3; ++a) {
      const float* c = coord + a * N;
      d[3 * dir + a] = 0.5 * ((double)c[p + st[dir]] - (double)c[p - st[dir]]) * rdh;
    }
  for (int q = 0; q < 9; ++q) cov[q * N + p] = d[q];
  cov[9 * N + p] = d[0] * d[4] * d[8] + d[1] * d[5] * d[6] + d[2] * d[3] * d[7] - d[2] * d[4] * d[6] -
                   d[1] * d[3] * d[8] - d[0] * d[7] * d[5];
}

// SCFDM coefficient planes (CJM.cpp:107-147): cofactors J*xi_x .. J*zt_z on
// [2, n-2), J and the tangents x_xi, x_et on [1, n-1), zero elsewhere; medium
// (mu, lambda, 1/rho) everywhere.
template <class R>
__global__ void k_pack_cmm(const double* __restrict__ cov, const float* __restrict__ med, R* __restrict__ cjm, int NX,
                           int NY, int NZ, long long N) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const long long S = (long long)NX * NY;
  const int i = (int)(p % NX), j = (int)((p / NX) % NY), k = (int)(p / S);
  const bool in1 = i >= 1 && i < NX - 1 && j >= 1 && j < NY - 1 && k >= 1 && k < NZ - 1;
  const bool in2 = i >= 2 && i < NX - 2 && j >= 2 && j < NY - 2 && k >= 2 && k < NZ - 2;
  double d[9];
  for (int q = 0; q < 9; ++q) d[q] = in1 ? cov[q * N + p] : 0.0;
  const double cof[9] = {d[4] * d[8] - d[5] * d[7], d[5] * d[6] - d[3] * d[8], d[3] * d[7] - d[4] * d[6],
                         d[7] * d[2] - d[8] * d[1], d[8] * d[0] - d[6] * d[2], d[6] * d[1] - d[7] * d[0],
                         d[1] * d[5] - d[2] * d[4], d[2] * d[3] - d[0] * d[5], d[0] * d[4] - d[1] * d[3]};
  for (int q = 0; q < 9; ++q) cjm[q * N + p] = in2 ? (R)cof[q] : R(0);
  cjm[9 * N + p] = in1 ? (R)cov[9 * N + p] : R(0);
  for (int q = 0; q < 6; ++q) cjm[(13 + q) * N + p] = (R)d[q];
  const double vs = med[p], vp = med[N + p], rho = med[2 * N + p];
  cjm[10 * N + p] = (R)(rho * vs * vs);
  cjm[11 * N + p] = (R)(rho * (vp * vp - 2.0 * vs * vs));
  cjm[12 * N + p] = (R)(1.0 / rho);
}

// Minimum / maximum point-to-plane distance over the interior (calc_CFL.cpp:56-163),
// one partial per block.
__global__ void k_hrange(const float* __restrict__ coord, double* part, int nx, int ny, int nz, int NX, int NY,
                         long long N) {
  __shared__ double smin[256], smax[256];
  long long tot = (long long)nx * ny * nz;
  double lo = 1e300, hi = 0.0;
  const long long S = (long long)NX * NY;
  const float* X = coord;
  const float* Y = coord + N;
  const float* Z = coord + 2 * N;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int i = H + (int)(e % nx);
    int j = H + (int)((e / nx) % ny);
    int k = H + (int)(e / ((long long)nx * ny));
    long long p = (long long)i + (long long)NX * j + (long long)k * S;
    double P[3] = {X[p], Y[p], Z[p]};
    for (int a = -1; a <= 1; a += 2)
      for (int b = -1; b <= 1; b += 2)
        for (int c = -1; c <= 1; c += 2) {
          long long qa = p - a, qb = p - (long long)b * NX, qc = p - (long long)c * S;
          double A[3] = {X[qa], Y[qa], Z[qa]}, B[3] = {X[qb], Y[qb], Z[qb]}, C[3] = {X[qc], Y[qc], Z[qc]};
          double ab0 = B[0] - A[0], ab1 = B[1] - A[1], ab2 = B[2] - A[2];
          double ac0 = C[0] - A[0], ac1 = C[1] - A[1], ac2 = C[2] - A[2];
          double n0 = ab1 * ac2 - ab2 * ac1, n1 = ab2 * ac0 - ab0 * ac2, n2 = ab0 * ac1 - ab1 * ac0;
          double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
          double d = fabs(((P[0] - A[0]) * n0 + (P[1] - A[1]) * n1 + (P[2] - A[2]) * n2) / nn);
          lo = fmin(lo, d);
          hi = fmax(hi, d);
        }
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// min/max of one float array, one partial per block
__global__ void k_frange(const float* __restrict__ a, long long n, double* part) {
  __shared__ double smin[256], smax[256];
  double lo = 1e300, hi = -1e300;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double v = a[e];
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// ---------------------------------------------------------------------------
// Allocation-free CFL ranges: the per-block partials go to `scratch`, a
// file-scope device array (4 reductions x RANGE_BLOCKS (min, max) pairs =
// 64 KiB) that every GPU the library runs on holds until the process ends.
// Host code never passes the array's name to a kernel: its device address is
// looked up once per device with hipGetSymbolAddress and kept per device.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU race; each block overwrites its own partial, so nothing needs clearing.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[8 * RANGE_BLOCKS];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with the device current).
hipError_t range_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

// Block c reduces the nb (min, max) partials of reduction c into out8[2c], out8[2c + 1].
__global__ void k_range_final(const double* __restrict__ part, int nb, double* __restrict__ out8) {
  __shared__ double smin[256], smax[256];
  const double* q = part + 2LL * nb * blockIdx.x;
  double lo = 1e300, hi = -1e300;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    lo = fmin(lo, q[2 * b]);
    hi = fmax(hi, q[2 * b + 1]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out8[2 * blockIdx.x] = smin[0];
    out8[2 * blockIdx.x + 1] = smax[0];
  }
}

hipError_t grid_ranges_noalloc(int nx, int ny, int nz, const float* coord, const float* med, double* out8,
                               hipStream_t st) {
  double* part = nullptr;
  hipError_t e = range_scratch(&part);
  if (e != hipSuccess) return e;
  const int NX = nx + 2 * H, NY = ny + 2 * H, NZ = nz + 2 * H;
  const long long N = (long long)NX * NY * NZ;
  const int nb = RANGE_BLOCKS;
  k_hrange<<<nb, 256, 0, st>>>(coord, part, nx, ny, nz, NX, NY, N);
  for (int c = 0; c < 3; ++c) k_frange<<<nb, 256, 0, st>>>(med + c * N, N, part + 2LL * nb * (c + 1));
  k_range_final<<<4, 256, 0, st>>>(part, nb, out8);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_metrics(const float* coord, double* met, int nx, int ny, int nz, int NX, int NY, long long N, double dh,
                    bool fs, hipStream_t st) {
  dim3 g((nx + 63) / 64, ny, nz);
  k_metrics<<<g, 64, 0, st>>>(coord, met, nx, ny, nz, NX, NY, N, 1.0 / dh);
  if (fs) {
    dim3 g2((nx + 63) / 64, ny, 3);
    k_mirror_top<<<g2, 64, 0, st>>>(met, nx, ny, nz, NX, NY, N);
  }
}

template <class R, class TM>
void launch_pack(const double* met, const float* med, TM* coef, R* jac, long long N, double stress_scale,
                 hipStream_t st) {
  k_pack<R, TM><<<(unsigned)((N + 255) / 256), 256, 0, st>>>(met, med, coef, jac, N, stress_scale);
}

template <class R>
void launch_rdz(const double* met, const float* med, R* rdzdx, R* rdzdy, int nx, int ny, int nz, int NX, int NY,
                long long N, hipStream_t st) {
  dim3 g((nx + 63) / 64, ny, 1);
  k_rdz<R><<<g, 64, 0, st>>>(met, med, rdzdx, rdzdy, nx, ny, nz, NX, NY, N);
}

void launch_cov_cmm(const float* coord, double* cov, int NX, int NY, int NZ, long long N, double dh, hipStream_t st) {
  dim3 g((NX - 2 + 63) / 64, NY - 2, NZ - 2);
  k_cov_cmm<<<g, 64, 0, st>>>(coord, cov, NX, NY, NZ, N, 1.0 / dh);
}

template <class R>
void launch_pack_cmm(const double* cov, const float* med, R* cjm, int NX, int NY, int NZ, long long N, hipStream_t st) {
  k_pack_cmm<R><<<(unsigned)((N + 255) / 256), 256, 0, st>>>(cov, med, cjm, NX, NY, NZ, N);
}

template <class TM>
void launch_pack16(const double* met, const float* med, float* aff, unsigned* words, float* jac, float* mu_src, int NX,
                   int NY, int NZ, double stress_scale, hipStream_t st) {
  const long long S = (long long)NX * NY, N = S * NZ;
  k_coef_range<<<dim3(NZ, M_NCOEF), 256, 0, st>>>(met, med, aff, S, N, NZ, stress_scale);
  k_pack16<TM><<<dim3((NX + 255) / 256, NY, NZ), 256, 0, st>>>(met, med, aff, words, jac, mu_src, NX, S, N, NZ,
                                                                stress_scale);
}
template void launch_pack16<__half>(const double*, const float*, float*, unsigned*, float*, float*, int, int, int,
                                    double, hipStream_t);
template void launch_pack16<__hip_bfloat16>(const double*, const float*, float*, unsigned*, float*, float*, int, int,
                                            int, double, hipStream_t);

void launch_hrange(const float* coord, double* part, int nblocks, int nx, int ny, int nz, int NX, int NY, long long N,
                   hipStream_t st) {
  k_hrange<<<nblocks, 256, 0, st>>>(coord, part, nx, ny, nz, NX, NY, N);
}

void launch_frange(const float* a, long long n, double* part, int nblocks, hipStream_t st) {
  k_frange<<<nblocks, 256, 0, st>>>(a, n, part);
}

template void launch_pack<float, float>(const double*, const float*, float*, float*, long long, double, hipStream_t);
template void launch_pack<float, __half>(const double*, const float*, __half*, float*, long long, double, hipStream_t);
template void launch_pack<float, __hip_bfloat16>(const double*, const float*, __hip_bfloat16*, float*, long long,
                                                  double, hipStream_t);
template void launch_pack<double, double>(const double*, const float*, double*, double*, long long, double,
                                          hipStream_t);
template void launch_rdz<float>(const double*, const float*, float*, float*, int, int, int, int, int, long long,
                                hipStream_t);
template void launch_rdz<double>(const double*, const float*, double*, double*, int, int, int, int, int, long long,
                                 hipStream_t);

template void launch_pack_cmm<float>(const double*, const float*, float*, int, int, int, long long, hipStream_t);
template void launch_pack_cmm<double>(const double*, const float*, double*, int, int, int, long long, hipStream_t);

}  // namespace cg
