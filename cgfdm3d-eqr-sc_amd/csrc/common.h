// Shared definitions of the MI355X CGFDM library (device and host side).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>

namespace cg {

constexpr int H = 3;  // ghost layers on every side of the local domain (reference macro.h:39)

// MacCormack one-sided coefficients (reference macro.h:121-125) and the 3-point
// one-sided stencil one layer under the surface (macro.h:144-146).
constexpr double AF_1 = -0.30874, AF0 = -0.6326, AF1 = 1.2330, AF2 = -0.3334, AF3 = 0.04168;
constexpr double CF1 = -1.16666667, CF2 = 1.33333333, CF3 = -0.16666667;
// Classical RK4 written on (W, t, m) buffers (macro.h:133-142).
constexpr double RK_A2 = 0.5, RK_A3 = 0.5;
constexpr double RK_B1 = 0.16666667, RK_B2 = 0.33333333, RK_B3 = 0.33333333, RK_B4 = 0.16666667;

// Storage type of metric and medium coefficients.
template <class T> struct Store;
template <> struct Store<float> {
  static __device__ __forceinline__ float ld(const float* p) { return *p; }
};
template <> struct Store<double> {
  static __device__ __forceinline__ double ld(const double* p) { return *p; }
};
template <> struct Store<__half> {
  static __device__ __forceinline__ float ld(const __half* p) { return __half2float(*p); }
};
template <> struct Store<__hip_bfloat16> {
  static __device__ __forceinline__ float ld(const __hip_bfloat16* p) { return __bfloat162float(*p); }
};

// 16-bit coefficient storage (fp16 / bf16): the twelve coefficient planes are
// stored in pairs (c = 2q, 2q + 1) as one 32-bit word per cell (code of plane
// 2q in the low half), and each 16-bit value is an affine code over its
// (plane, k) layer: value = mid[c][k] + half[c][k] * code, code in [-1, 1],
// with (mid, half) the centre and half-width of the layer's value range in a
// float table aff[c][k][2]. A layer whose values are all equal is stored
// exactly; otherwise the rounding error is bounded by half * 2^-11 (fp16) or
// half * 2^-8 (bf16) instead of |value| times that.
template <class T> struct Bits16;
template <> struct Bits16<__half> {
  static __device__ __forceinline__ float to_float(unsigned short b) { return __half2float(__ushort_as_half(b)); }
};
template <> struct Bits16<__hip_bfloat16> {
  static __device__ __forceinline__ float to_float(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
};

// Coefficient planes, SoA over the padded grid. Metrics are stored as the
// unit-free products m * DH (i.e. divided by rDH) only when `scaled` is set;
// the default keeps the reference meaning (xi_x ... zt_z in 1/m).
enum { M_XIX = 0, M_XIY, M_XIZ, M_ETX, M_ETY, M_ETZ, M_ZTX, M_ZTY, M_ZTZ, M_MU, M_LAM, M_B, M_NCOEF };

// One PML face slab (reference pml_deriv.cpp:17-37): three RK buffers of 9 SoA
// components over `cnt` cells. a = state at the step start (and the final
// state), b = intermediate stage state (updated in place), t = accumulator.
template <class R>
struct FaceBuf {
  R* a;
  R* b;
  R* t;
  long long cnt;
};

template <class R>
struct PmlArgs {
  int npml;
  int has_z2;         // bottom-only PML in z when a free surface is present
  const R* alpha[3];  // per axis, padded length
  const R* beta[3];   // inverted beta (1 outside the PML)
  const R* d[3];      // d / beta
  FaceBuf<R> f[6];    // x1, y1, z1, x2, y2, z2
};

// Access at a byte offset from a uniform base pointer: the compiler emits the
// global load/store form with a 64-bit SGPR base and a 32-bit VGPR offset, so
// one offset register serves every plane of every grid array (field, RK,
// coefficient). Every grid plane array spans less than 4 GiB (checked on the
// host before a launch), so byte offsets of cells fit 32 bits.
template <class T>
__device__ __forceinline__ const T* at(const T* base, unsigned boff) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <class T>
__device__ __forceinline__ T* at(T* base, unsigned boff) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff);
}

// XCD-aware block order of the interior launch. The hardware hands block b to
// XCD b % 8, so block b is the (b / 8)-th block that XCD runs. The (x, y) plane
// is cut into regions of rtx 64-column tiles by rty 4-row tiles (rtx <= 0 or
// larger than the row: whole rows, so that no region is a partial column); regions are
// dealt to the XCDs in turn and each XCD walks its regions one after another,
// plane by plane in k and, within a plane, row of tiles by row of tiles. The
// x, y and z stencil neighbours of a cell are then read by blocks of the same
// XCD a few planes earlier or later and come from that XCD's L2 instead of
// HBM, except across region edges. Returns false for padding blocks.
struct CellTile {
  int i0, j0, k;
};
__device__ __host__ __forceinline__ int region_extent(int r, int n) { return (r <= 0 || r > n) ? n : r; }

__device__ __forceinline__ bool interior_tile(int b, int nx, int ny, int nk, int rtx, int rty, CellTile& t) {
  const int ntx = (nx + 63) / 64, nty = (ny + 3) / 4;
  rtx = region_extent(rtx, ntx);  // rtx <= 0: whole rows of tiles
  rty = region_extent(rty, nty);
  const int nrx = (ntx + rtx - 1) / rtx, nry = (nty + rty - 1) / rty;
  const int per_plane = rtx * rty;
  const int per_region = per_plane * nk;
  const int xcd = b & 7, seq = b >> 3;
  const int slot = seq / per_region, within = seq - slot * per_region;
  const int r = slot * 8 + xcd;
  const int kk = within / per_plane, q = within - kk * per_plane;
  const int ty = q / rtx, tx = q - ty * rtx;
  const int rx = r % nrx, ry = r / nrx;
  const int txg = rx * rtx + tx, tyg = ry * rty + ty;
  t.i0 = txg * 64;
  t.j0 = tyg * 4;
  t.k = kk;
  return r < nrx * nry && txg < ntx && tyg < nty;
}

// number of blocks of the interior launch for interior_tile()
inline long long interior_blocks(int nx, int ny, int nk, int rtx, int rty) {
  const long long ntx = (nx + 63) / 64, nty = (ny + 3) / 4;
  rtx = region_extent(rtx, (int)ntx);
  rty = region_extent(rty, (int)nty);
  const long long nregions = ((ntx + rtx - 1) / rtx) * ((nty + rty - 1) / rty);
  return 8 * ((nregions + 7) / 8) * (long long)rtx * rty * nk;
}

}  // namespace cg
