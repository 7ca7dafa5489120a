// CGFDM time-step kernels for MI355X (gfx950).
//
// One RK4 stage of the reference (propagate.cpp:212-251: wave_deriv ->
// free_surface_deriv -> pml_deriv -> pml_free_surface -> wave_rk -> pml_rk)
// is done by two fused kernels, one thread per cell, that read the stencil
// input and write the stage output once:
//
//   cg_cell      every cell below the three surface layers: 27 MacCormack
//                derivatives, the right-hand side, the CFS-PML auxiliary ODE of
//                every face the cell belongs to and the RK4 update of W and of
//                the auxiliaries. Rows of 64 cells along x make every load and
//                store one coalesced 256 B (FP32) transaction per wavefront.
//   cg_surface   the top three layers (traction image, reference
//                freeSurface.cpp:243-494 and pml_freeSurface.cpp:17-242),
//                same fusion; on the last stage it also updates the PGV map and
//                the surface displacement.
//
// RK4 buffers (classical RK4 in the reference's (W, t, m) form, wave_rk.cpp):
//   stage 0: in = A (also the RK base m), out = B, t = A + h/6
//   stage 1: in = B, m = A, out = C, t += h/3
//   stage 2: in = C, m = A, out = B, t += h/3
//   stage 3: in = B,         out = A = t + h/6
// Ghost cells of A, B, C are zero and never written (single-rank domain with
// physical boundaries on all sides).
#pragma once
#include "common.h"
#include "kernels.h"

namespace cg {

template <class R>
__device__ __forceinline__ void assemble(const R* D, R m1, R m2, R m3, R mu, R lam, R b, R* Q) {
  // velocity rows: b * (m . T), stress rows: lambda div + 2 mu m_i V_i, mu (m_j V_i + m_i V_j)
  // (reference wave_deriv.cpp:244-272)
  Q[0] = (m1 * D[3] + m2 * D[6] + m3 * D[7]) * b;
  Q[1] = (m1 * D[6] + m2 * D[4] + m3 * D[8]) * b;
  Q[2] = (m1 * D[7] + m2 * D[8] + m3 * D[5]) * b;
  R div = m1 * D[0] + m2 * D[1] + m3 * D[2];
  R tm = R(2) * mu;
  Q[3] = div * lam + tm * (m1 * D[0]);
  Q[4] = div * lam + tm * (m2 * D[1]);
  Q[5] = div * lam + tm * (m3 * D[2]);
  Q[6] = (m2 * D[0] + m1 * D[1]) * mu;
  Q[7] = (m3 * D[0] + m1 * D[2]) * mu;
  Q[8] = (m3 * D[1] + m2 * D[2]) * mu;
}

template <class R>
__device__ __forceinline__ void stress_rows(const R* V, R m1, R m2, R m3, R mu, R lam, R* S) {
  R div = m1 * V[0] + m2 * V[1] + m3 * V[2];
  R tm = R(2) * mu;
  S[0] = div * lam + tm * (m1 * V[0]);
  S[1] = div * lam + tm * (m2 * V[1]);
  S[2] = div * lam + tm * (m3 * V[2]);
  S[3] = (m2 * V[0] + m1 * V[1]) * mu;
  S[4] = (m3 * V[0] + m1 * V[2]) * mu;
  S[5] = (m3 * V[1] + m2 * V[2]) * mu;
}

// 5-point MacCormack difference (without 1/DH) of f around f[0] with stride s.
template <int FB, class R>
__device__ __forceinline__ R mac(const R* f, long long s) {
  return R(FB) * (R(AF_1) * f[-FB * s] + R(AF0) * f[0] + R(AF1) * f[FB * s] + R(AF2) * f[2 * FB * s] +
                  R(AF3) * f[3 * FB * s]);
}

// CFS-PML auxiliary update for one face at slab index pos (reference
// pml_deriv.cpp:235-273 and pml_rk.cpp:19-73): hA = (d Q - (alpha+d) A) DT,
// h -= beta A DT, then the RK4 stage on the auxiliary state. The stage is a
// compile-time parameter: every kernel instance has one straight-line memory
// pattern and no stage-dependent choice of base pointers.
template <class R, int STAGE>
__device__ __forceinline__ void pml_face(const FaceBuf<R>& F, long long pos, const R (&Q)[9], R d, R ad, R be, R DT,
                                         R (&h)[9], bool has_extra, const R (&extra)[6]) {
  // extra: surface stress term added to hA[3..8] on the top layer.
  // The nine fields go in three groups of three, each group's loads before its
  // stores: the buffers are not known not to alias, so a store between two
  // loads would make the compiler wait for each load in turn (three round
  // trips instead of nine, with few registers held).
#pragma unroll
  for (int g = 0; g < 9; g += 3) {
    R cur[3], tv[3], av[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const long long q = (long long)(g + u) * F.cnt + pos;
      cur[u] = STAGE == 0 ? F.a[q] : F.b[q];
      if constexpr (STAGE == 1 || STAGE == 2 || STAGE == 3) tv[u] = F.t[q];
      if constexpr (STAGE == 1 || STAGE == 2) av[u] = F.a[q];
    }
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int v = g + u;
      const long long q = (long long)v * F.cnt + pos;
      R hA = (d * Q[v] - ad * cur[u]) * DT;
      if (has_extra && v >= 3) hA += extra[v - 3] * DT;
      h[v] -= be * cur[u] * DT;
      if constexpr (STAGE == 0) {
        F.t[q] = cur[u] + R(RK_B1) * hA;
        F.b[q] = cur[u] + R(RK_A2) * hA;
      } else if constexpr (STAGE == 1) {
        F.t[q] = tv[u] + R(RK_B2) * hA;
        F.b[q] = av[u] + R(RK_A3) * hA;
      } else if constexpr (STAGE == 2) {
        F.t[q] = tv[u] + R(RK_B3) * hA;
        F.b[q] = av[u] + hA;
      } else {
        F.a[q] = tv[u] + R(RK_B4) * hA;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// interior cell update: 27 derivatives, PML faces, RK4 stage. `xd(v)`, `yd(v)`
// and `zd(v)` return the unscaled xi, eta and zeta differences of field v and
// `wc(v)` the centre value, so the same body serves the thread-per-cell kernel
// (differences from global memory) and the z-marching kernel (x/y from an LDS
// plane, z from a register window).
// ---------------------------------------------------------------------------
// Streaming accesses of a stage (coefficients, RK base and accumulator, the
// two outputs) are used once per cell: they carry the non-temporal hint so
// that they do not displace the stencil rows of W in L2 (measured on the
// 512x512x256 workload: interior kernel 5.20 -> 5.06 ms per launch).
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_stream(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Byte offset of padded cell (i, j, k) in a grid array of R elements (16-bit
// coefficient words are 4 bytes, as R = float in that storage).
template <class R>
__device__ __forceinline__ unsigned cell_off(const StageArgs<R>& a, int i, int j, int k) {
  return (unsigned)i * (unsigned)sizeof(R) + (unsigned)j * a.rowb + (unsigned)k * a.planeb;
}

// 5-point MacCormack difference (without 1/DH) of one plane array around byte
// offset o with a byte stride st between stencil points.
template <int FB, class R>
__device__ __forceinline__ R mac_off(const R* base, unsigned o, unsigned st) {
  const unsigned d = FB > 0 ? st : 0u - st;
  return R(FB) * (R(AF_1) * *at(base, o - d) + R(AF0) * *at(base, o) + R(AF1) * *at(base, o + d) +
                  R(AF2) * *at(base, o + 2 * d) + R(AF3) * *at(base, o + 3 * d));
}

// Coefficient c of the cell at byte offset `off` in layer k. Native storage:
// plane c (with the non-temporal hint when NT is set). 16-bit storage: the half
// of word plane c / 2 that holds plane c, decoded with the layer's affine pair.
template <class R, class TM, bool NT>
__device__ __forceinline__ R coef_off(const TM* __restrict__ coef, const StageArgs<R>& a, int c, unsigned off, int k) {
  if constexpr (sizeof(TM) == 2) {
    static_assert(sizeof(R) == 4, "16-bit coefficient words are addressed like an FP32 grid array");
    const unsigned* w = at(reinterpret_cast<const unsigned*>(coef) + (long long)(c >> 1) * a.N, off);
    const unsigned word = NT ? __builtin_nontemporal_load(w) : *w;
    const unsigned short b = (c & 1) ? (unsigned short)(word >> 16) : (unsigned short)(word & 0xffffu);
    const float* A = a.aff + 2 * ((long long)c * a.nzp + k);
    return R(A[0] + A[1] * Bits16<TM>::to_float(b));
  } else {
    static_assert(sizeof(TM) == sizeof(R), "native coefficients share the grid byte offsets");
    const TM* q = at(coef + (long long)c * a.N, off);
    if constexpr (NT) return R(__builtin_nontemporal_load(q));
    else return R(*q);
  }
}

// Coefficient c of cell q (linear index, in layer k), for the surface kernel.
template <class R, class TM, bool NT>
__device__ __forceinline__ R coef_at(const TM* __restrict__ coef, const StageArgs<R>& a, int c, long long q, int k) {
  if constexpr (sizeof(TM) == 2) {
    const unsigned* w = reinterpret_cast<const unsigned*>(coef) + (long long)(c >> 1) * a.N + q;
    const unsigned word = NT ? __builtin_nontemporal_load(w) : *w;
    const unsigned short b = (c & 1) ? (unsigned short)(word >> 16) : (unsigned short)(word & 0xffffu);
    const float* A = a.aff + 2 * ((long long)c * a.nzp + k);
    return R(A[0] + A[1] * Bits16<TM>::to_float(b));
  } else if constexpr (NT && sizeof(TM) == 4) {
    return R(__builtin_nontemporal_load(coef + (long long)c * a.N + q));
  } else {
    return R(coef[(long long)c * a.N + q]);
  }
}

template <class R, class TM, int FB1, int FB2, int FB3, int STAGE, class XD, class YD, class ZD, class WC>
__device__ __forceinline__ void cell_body(const StageArgs<R>& a, const TM* __restrict__ coef, int i, int j, int k,
                                          unsigned off, XD xd, YD yd, ZD zd, WC wc) {
  const long long N = a.N;
  auto cf = [&](int c) -> R { return coef_off<R, TM, true>(coef, a, c, off, k); };
  const R mu = cf(M_MU), lam = cf(M_LAM), bu = cf(M_B);
  const PmlArgs<R>& P = a.pml;
  R h[9], D[9], Q[9];

  // x: derivatives along xi (wave_deriv.cpp:214-222), scaled by 1/beta_x
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = xd(v);
  assemble(D, cf(M_XIX), cf(M_XIY), cf(M_XIZ), mu, lam, bu, Q);
  const R bx = P.beta[0][i];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= a.rDH;
    h[v] = bx * Q[v] * a.DT;
  }
  const R none[6] = {0, 0, 0, 0, 0, 0};
  if (P.npml > 0) {
    const bool x1 = i < H + P.npml, x2 = i >= H + a.nx - P.npml;
    if (x1 || x2) {
      const R d = P.d[0][i], ad = d + P.alpha[0][i];
      const long long row = (long long)(j - H) + (long long)a.ny * (k - H);
      if (x1) pml_face<R, STAGE>(P.f[0], (long long)(i - H) + (long long)P.npml * row, Q, d, ad, bx, a.DT, h, false, none);
      if (x2) pml_face<R, STAGE>(P.f[3], (long long)(i - H - (a.nx - P.npml)) + (long long)P.npml * row, Q, d, ad, bx, a.DT, h, false, none);
    }
  }
  // y: derivatives along eta (wave_deriv.cpp:224-232)
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = yd(v);
  assemble(D, cf(M_ETX), cf(M_ETY), cf(M_ETZ), mu, lam, bu, Q);
  const R by = P.beta[1][j];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= a.rDH;
    h[v] += by * Q[v] * a.DT;
  }
  if (P.npml > 0) {
    const bool y1 = j < H + P.npml, y2 = j >= H + a.ny - P.npml;
    if (y1 || y2) {
      const R d = P.d[1][j], ad = d + P.alpha[1][j];
      const long long i0 = i - H, k0 = k - H;
      if (y1) pml_face<R, STAGE>(P.f[1], i0 + (long long)a.nx * ((j - H) + (long long)P.npml * k0), Q, d, ad, by, a.DT, h, false, none);
      if (y2) pml_face<R, STAGE>(P.f[4], i0 + (long long)a.nx * ((j - H - (a.ny - P.npml)) + (long long)P.npml * k0), Q, d, ad, by, a.DT, h, false, none);
    }
  }
  // z: derivatives along zeta (wave_deriv.cpp:234-242)
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = zd(v);
  assemble(D, cf(M_ZTX), cf(M_ZTY), cf(M_ZTZ), mu, lam, bu, Q);
  const R bz = P.beta[2][k];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= a.rDH;
    h[v] += bz * Q[v] * a.DT;
  }
  if (P.npml > 0) {
    const bool z1 = k < H + P.npml, z2 = P.has_z2 && k >= H + a.nz - P.npml;
    if (z1 || z2) {
      const R d = P.d[2][k], ad = d + P.alpha[2][k];
      const long long i0 = i - H, j0 = j - H;
      if (z1) pml_face<R, STAGE>(P.f[2], i0 + (long long)a.nx * (j0 + (long long)a.ny * (k - H)), Q, d, ad, bz, a.DT, h, false, none);
      if (z2) pml_face<R, STAGE>(P.f[5], i0 + (long long)a.nx * (j0 + (long long)a.ny * (k - H - (a.nz - P.npml))), Q, d, ad, bz, a.DT, h, false, none);
    }
  }
  // RK4 stage update of W (wave_rk.cpp:18-104). The RK base and accumulator
  // are all loaded before the first store (the arrays are not known not to
  // alias: interleaving would make every load wait for the store before it).
  R tv[9], mv[9];
  if constexpr (STAGE != 0) {
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      tv[v] = ld_stream(at((const R*)a.wt + (long long)v * N, off));
      if constexpr (STAGE == 1 || STAGE == 2) mv[v] = ld_stream(at(a.wm + (long long)v * N, off));
    }
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    R* wt = at(a.wt + (long long)v * N, off);
    R* wo = at(a.wout + (long long)v * N, off);
    if constexpr (STAGE == 0) {
      const R m = wc(v);
      st_stream(wt, R(m + R(RK_B1) * h[v]));
      st_stream(wo, R(m + R(RK_A2) * h[v]));
    } else if constexpr (STAGE == 1) {
      st_stream(wt, R(tv[v] + R(RK_B2) * h[v]));
      st_stream(wo, R(mv[v] + R(RK_A3) * h[v]));
    } else if constexpr (STAGE == 2) {
      st_stream(wt, R(tv[v] + R(RK_B3) * h[v]));
      st_stream(wo, R(mv[v] + h[v]));
    } else {
      st_stream(wo, R(tv[v] + R(RK_B4) * h[v]));
    }
  }
}

// ---------------------------------------------------------------------------
// interior kernel: one thread per cell below the surface layers
// ---------------------------------------------------------------------------

template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_cell(StageArgs<R> a, const TM* __restrict__ coef) {
  CellTile t;
  if (!interior_tile(blockIdx.x, a.nx, a.ny, a.kend - a.kbeg, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = H + t.j0 + (threadIdx.x >> 6);
  const int k = a.kbeg + t.k;
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long N = a.N;
  const unsigned off = cell_off(a, i, j, k);
  const R* W = a.win;
  cell_body<R, TM, FB1, FB2, FB3, STAGE>(
      a, coef, i, j, k, off,
      [&](int v) -> R { return mac_off<FB1>(W + (long long)v * N, off, (unsigned)sizeof(R)); },
      [&](int v) -> R { return mac_off<FB2>(W + (long long)v * N, off, a.rowb); },
      [&](int v) -> R { return mac_off<FB3>(W + (long long)v * N, off, a.planeb); },
      [&](int v) -> R { return *at(W + (long long)v * N, off); });
}

// ---------------------------------------------------------------------------
// register z-marching interior kernel: a block of 64 x 4 threads owns a 64 x 4
// column tile and marches up through a chunk of zc planes. The z stencil of
// every cell comes from a five-plane register window of the thread's own
// column (one new plane of the nine fields per cell, loaded one plane ahead),
// so the z neighbours are read from HBM once instead of once per stencil
// point; the x and y stencils are read from global memory and hit L1/L2.
// No LDS and no barriers. Blocks are dealt to the XCDs in regions of rtx x rty
// tiles; all tiles of a region march the same chunk together, so the y and x
// halo rows of a tile are the rows its neighbours are loading at that moment.
// ---------------------------------------------------------------------------
struct ZregTile {
  int i0, j0, k0, k1;
};
__device__ __forceinline__ bool zreg_tile(int b, int nx, int ny, int kbeg, int kend, int zc, int rtx, int rty,
                                          ZregTile& t) {
  const int ntx = (nx + 63) / 64, nty = (ny + 3) / 4, nch = (kend - kbeg + zc - 1) / zc;
  rtx = region_extent(rtx, ntx);
  rty = region_extent(rty, nty);
  const int nrx = (ntx + rtx - 1) / rtx, nry = (nty + rty - 1) / rty;
  const int per_chunk = rtx * rty, per_region = per_chunk * nch;
  const int xcd = b & 7, seq = b >> 3;
  const int slot = seq / per_region, within = seq - slot * per_region;
  const int r = slot * 8 + xcd;
  const int ch = within / per_chunk, q = within - ch * per_chunk;
  const int ty = q / rtx, tx = q - ty * rtx;
  const int rx = r % nrx, ry = r / nrx;
  const int txg = rx * rtx + tx, tyg = ry * rty + ty;
  t.i0 = txg * 64;
  t.j0 = tyg * 4;
  t.k0 = kbeg + ch * zc;
  t.k1 = min(t.k0 + zc, kend);
  return r < nrx * nry && txg < ntx && tyg < nty;
}

inline long long zreg_blocks(int nx, int ny, int nk, int zc, int rtx, int rty) {
  const long long ntx = (nx + 63) / 64, nty = (ny + 3) / 4, nch = (nk + zc - 1) / zc;
  rtx = region_extent(rtx, (int)ntx);
  rty = region_extent(rty, (int)nty);
  const long long nregions = ((ntx + rtx - 1) / rtx) * ((nty + rty - 1) / rty);
  return 8 * ((nregions + 7) / 8) * (long long)rtx * rty * nch;
}

template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void cg_zreg(StageArgs<R> a, const TM* __restrict__ coef) {
  ZregTile t;
  if (!zreg_tile(blockIdx.x, a.nx, a.ny, a.kbeg, a.kend, a.zreg, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = H + t.j0 + (threadIdx.x >> 6);
  if (i >= H + a.nx || j >= H + a.ny) return;  // no barriers below
  const long long N = a.N;
  const R* W = a.win;
  // window slot m holds plane k - ZL + m (m = 0..4): the 5-point z stencil of
  // plane k spans k - FB3 .. k + 3 FB3
  constexpr int ZL = FB3 > 0 ? 1 : 3;
  R zw[9][5], nxt[9];
  unsigned off = cell_off(a, i, j, t.k0);
  const unsigned pb = a.planeb;
#pragma unroll
  for (int v = 0; v < 9; ++v) {
#pragma unroll
    for (int m = 0; m < 4; ++m) zw[v][m] = *at(W + (long long)v * N, off + (unsigned)(m - ZL) * pb);
    nxt[v] = *at(W + (long long)v * N, off + (unsigned)(4 - ZL) * pb);
  }
  for (int k = t.k0; k < t.k1; ++k, off += pb) {
#pragma unroll
    for (int v = 0; v < 9; ++v) zw[v][4] = nxt[v];
    if (k + 1 < t.k1) {
#pragma unroll
      for (int v = 0; v < 9; ++v) nxt[v] = *at(W + (long long)v * N, off + (unsigned)(5 - ZL) * pb);
    }
    cell_body<R, TM, FB1, FB2, FB3, STAGE>(
        a, coef, i, j, k, off,
        [&](int v) -> R { return mac_off<FB1>(W + (long long)v * N, off, (unsigned)sizeof(R)); },
        [&](int v) -> R { return mac_off<FB2>(W + (long long)v * N, off, a.rowb); },
        [&](int v) -> R {
          return R(FB3) * (R(AF_1) * zw[v][ZL - FB3] + R(AF0) * zw[v][ZL] + R(AF1) * zw[v][ZL + FB3] +
                           R(AF2) * zw[v][ZL + 2 * FB3] + R(AF3) * zw[v][ZL + 3 * FB3]);
        },
        [&](int v) -> R { return zw[v][ZL]; });
#pragma unroll
    for (int v = 0; v < 9; ++v) {
#pragma unroll
      for (int m = 0; m < 4; ++m) zw[v][m] = zw[v][m + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// x-paired interior kernel (FP32 wavefield and coefficients, even NX): each
// thread computes the two x-adjacent cells (i, i + 1), i even, so that every
// field, coefficient and RK array access is one 8-byte load or store per
// thread: half the vector memory instructions of the thread-per-cell kernel,
// and the x stencil of the pair comes from four 8-byte loads per field instead
// of ten 4-byte ones. Pairs that touch a PML slab or a ghost column take the
// per-cell body for each of their interior cells.
// ---------------------------------------------------------------------------
struct __attribute__((aligned(8))) V2 {
  float x, y;
  __device__ V2() {}
  __device__ V2(float s) : x(s), y(s) {}
  __device__ V2(float a, float b) : x(a), y(b) {}
};
__device__ __forceinline__ V2 operator+(V2 a, V2 b) { return V2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ V2 operator-(V2 a, V2 b) { return V2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ V2 operator*(V2 a, V2 b) { return V2(a.x * b.x, a.y * b.y); }
__device__ __forceinline__ V2& operator*=(V2& a, V2 b) { a = a * b; return a; }
__device__ __forceinline__ V2& operator+=(V2& a, V2 b) { a = a + b; return a; }
__device__ __forceinline__ V2 ld2(const float* p) { return *reinterpret_cast<const V2*>(p); }
__device__ __forceinline__ void st2(float* p, V2 v) { *reinterpret_cast<V2*>(p) = v; }

// 5-point difference of a pair along a non-x axis: rows f + m * FB * s, m = -1..3
template <int FB>
__device__ __forceinline__ V2 mac2(const float* f, long long s) {
  return V2(float(FB)) * (V2(float(AF_1)) * ld2(f - FB * s) + V2(float(AF0)) * ld2(f) + V2(float(AF1)) * ld2(f + FB * s) +
                          V2(float(AF2)) * ld2(f + 2 * FB * s) + V2(float(AF3)) * ld2(f + 3 * FB * s));
}

// 5-point difference of the pair (i, i + 1) along x from aligned 8-byte loads
template <int FB>
__device__ __forceinline__ V2 macx2(const float* f) {
  if constexpr (FB > 0) {  // cell i: i-1 .. i+3; cell i+1: i .. i+4
    const V2 A = ld2(f - 2), B = ld2(f), C = ld2(f + 2), D = ld2(f + 4);
    return V2(float(AF_1) * A.y + float(AF0) * B.x + float(AF1) * B.y + float(AF2) * C.x + float(AF3) * C.y,
              float(AF_1) * B.x + float(AF0) * B.y + float(AF1) * C.x + float(AF2) * C.y + float(AF3) * D.x);
  } else {  // cell i: i+1 .. i-3; cell i+1: i+2 .. i-2
    const V2 E = ld2(f - 4), A = ld2(f - 2), B = ld2(f), C = ld2(f + 2);
    return V2(-(float(AF_1) * B.y + float(AF0) * B.x + float(AF1) * A.y + float(AF2) * A.x + float(AF3) * E.y),
              -(float(AF_1) * C.x + float(AF0) * B.y + float(AF1) * B.x + float(AF2) * A.y + float(AF3) * A.x));
  }
}

template <int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_pair(StageArgs<float> a, const float* __restrict__ coef) {
  const int lane = threadIdx.x & 63;
  const int ntx = (a.nx + 1 + 127) / 128;  // 128-cell tiles covering i in [2, H + nx)
  const int nty = (a.ny + 3) / 4;
  const int b = blockIdx.x;
  const int tile_x = b % ntx, tile_y = (b / ntx) % nty;
  const int k = a.kbeg + b / (ntx * nty);
  const int i = 2 + 2 * (64 * tile_x + lane);
  const int j = H + 4 * tile_y + (threadIdx.x >> 6);
  if (k >= a.kend || j >= H + a.ny) return;
  const bool c0 = i >= H && i < H + a.nx, c1 = i + 1 < H + a.nx;
  if (!c0 && !c1) return;
  const long long N = a.N, S = a.S, NXl = a.NX;
  const PmlArgs<float>& P = a.pml;
  const float* W = a.win;
  const bool in_pml = P.npml > 0 && (i < H + P.npml || i + 1 >= H + a.nx - P.npml || j < H + P.npml ||
                                     j >= H + a.ny - P.npml || k < H + P.npml ||
                                     (P.has_z2 && k >= H + a.nz - P.npml));
  if (in_pml || !c0 || !c1) {
#pragma unroll 1
    for (int c = 0; c < 2; ++c) {
      const int ic = i + c;
      if (ic < H || ic >= H + a.nx) continue;
      const long long p = (long long)ic + NXl * j + (long long)k * S;
      cell_body<float, float, FB1, FB2, FB3, STAGE>(
          a, coef, ic, j, k, cell_off(a, ic, j, k), [&](int v) -> float { return mac<FB1>(W + (long long)v * N + p, 1); },
          [&](int v) -> float { return mac<FB2>(W + (long long)v * N + p, NXl); },
          [&](int v) -> float { return mac<FB3>(W + (long long)v * N + p, S); },
          [&](int v) -> float { return W[(long long)v * N + p]; });
    }
    return;
  }
  const long long p = (long long)i + NXl * j + (long long)k * S;  // even: 8-byte aligned
  auto cf = [&](int c) -> V2 { return ld2(coef + (long long)c * N + p); };
  const V2 mu = cf(M_MU), lam = cf(M_LAM), bu = cf(M_B);
  const V2 rdh(a.rDH), dt(a.DT);
  V2 h[9], D[9], Q[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = macx2<FB1>(W + (long long)v * N + p);
  assemble(D, cf(M_XIX), cf(M_XIY), cf(M_XIZ), mu, lam, bu, Q);
  const V2 bx(P.beta[0][i], P.beta[0][i + 1]);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= rdh;
    h[v] = bx * Q[v] * dt;
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = mac2<FB2>(W + (long long)v * N + p, NXl);
  assemble(D, cf(M_ETX), cf(M_ETY), cf(M_ETZ), mu, lam, bu, Q);
  const V2 by(P.beta[1][j]);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= rdh;
    h[v] += by * Q[v] * dt;
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = mac2<FB3>(W + (long long)v * N + p, S);
  assemble(D, cf(M_ZTX), cf(M_ZTY), cf(M_ZTZ), mu, lam, bu, Q);
  const V2 bz(P.beta[2][k]);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Q[v] *= rdh;
    h[v] += bz * Q[v] * dt;
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const long long q = (long long)v * N + p;
    if constexpr (STAGE == 0) {
      const V2 m = ld2(W + q);
      st2(a.wt + q, m + V2(float(RK_B1)) * h[v]);
      st2(a.wout + q, m + V2(float(RK_A2)) * h[v]);
    } else if constexpr (STAGE == 1) {
      st2(a.wt + q, ld2(a.wt + q) + V2(float(RK_B2)) * h[v]);
      st2(a.wout + q, ld2(a.wm + q) + V2(float(RK_A3)) * h[v]);
    } else if constexpr (STAGE == 2) {
      st2(a.wt + q, ld2(a.wt + q) + V2(float(RK_B3)) * h[v]);
      st2(a.wout + q, ld2(a.wm + q) + h[v]);
    } else {
      st2(a.wout + q, ld2(a.wt + q) + V2(float(RK_B4)) * h[v]);
    }
  }
}

// ---------------------------------------------------------------------------
// z-marching interior kernel: a block of 64 x TY threads owns 64 x TY columns
// and walks a chunk of kc planes up z. Per plane it reads the nine fields once:
// the block's own cells into a per-thread register window over z (planes
// k - ZL .. k - ZL + 4, the 5-point zeta stencil of this F/B pattern), the x/y
// halo (1 cell on one side, 3 on the other, per the F/B signs) from global
// memory into a double-buffered LDS plane, which also receives the window's
// centre values. The next plane's window entry and halo are prefetched into
// registers while the current plane is computed. Columns past the interior
// (tile overhang) load nothing and hold zeros, which is what the ghost cells
// they stand for contain.
// ---------------------------------------------------------------------------
template <int FB1, int FB2, int TY>
struct ZTile {
  static constexpr int TX = 64;
  static constexpr int XL = FB1 > 0 ? 1 : 3, XR = 4 - XL;  // x halo left / right
  static constexpr int YL = FB2 > 0 ? 1 : 3, YR = 4 - YL;  // y halo below / above
  static constexpr int LX = TX + 4, LY = TY + 4, LP = LX * LY;
  static constexpr int NT = TX * TY;
  static constexpr int NHALO = 4 * LX + TY * 4;             // halo cells per field
  static constexpr int NHR = (9 * NHALO + NT - 1) / NT;      // halo loads per thread
  // tile-local LDS coordinates (lx, ly) of halo entry q of one field
  static __device__ __forceinline__ void halo_xy(int q, int& lx, int& ly) {
    if (q < YL * LX) {
      ly = q / LX;
      lx = q - ly * LX;
    } else if (q < YL * LX + TY * 4) {
      const int r = q - YL * LX, c = r & 3;
      ly = YL + (r >> 2);
      lx = c < XL ? c : TX + c;
    } else {
      const int r = q - YL * LX - TY * 4;
      ly = YL + TY + r / LX;
      lx = r - (ly - YL - TY) * LX;
    }
  }
};

template <class R, class TM, int FB1, int FB2, int FB3, int STAGE, int TY>
__global__ __launch_bounds__(64 * TY) void cg_zcol(StageArgs<R> a, const TM* __restrict__ coef) {
  using T = ZTile<FB1, FB2, TY>;
  constexpr int ZL = FB3 > 0 ? 1 : 3;  // window slot m holds plane k - ZL + m
  __shared__ R sw[2][9][T::LP];
  const int ntx = (a.nx + 63) / 64, nty = (a.ny + TY - 1) / TY;
  const int b = blockIdx.x;
  const int tile_x = b % ntx, tile_y = (b / ntx) % nty, chunk = b / (ntx * nty);
  const int k0 = a.kbeg + chunk * a.kc;
  const int k1 = min(k0 + a.kc, a.kend);
  if (k0 >= k1) return;  // uniform per block, before any barrier
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i = H + tile_x * 64 + tx, j = H + tile_y * TY + ty;
  const bool active = i < H + a.nx && j < H + a.ny;
  const long long N = a.N, S = a.S, NXl = a.NX;
  const long long col = (long long)i + NXl * j;
  const R* W = a.win;

  // halo entries of this thread: LDS slot (field * LP + ly * LX + lx, or -1)
  // and column offset in the plane (or -1: outside the padded array). The
  // plane size NX * NY fits a 32-bit int.
  int h_lds[T::NHR], h_col[T::NHR];
#pragma unroll
  for (int r = 0; r < T::NHR; ++r) {
    const int e = threadIdx.x + r * T::NT;
    h_lds[r] = -1;
    h_col[r] = -1;
    if (e < 9 * T::NHALO) {
      const int v = e / T::NHALO, q = e - v * T::NHALO;
      int lx, ly;
      T::halo_xy(q, lx, ly);
      const int gi = H + tile_x * 64 + lx - T::XL, gj = H + tile_y * TY + ly - T::YL;
      h_lds[r] = v * T::LP + ly * T::LX + lx;
      if (gi >= 0 && gi < a.NX && gj >= 0 && gj < a.NY) h_col[r] = gi + a.NX * gj;
    }
  }
  auto load_halo = [&](int kk, R (&hr)[T::NHR]) {
#pragma unroll
    for (int r = 0; r < T::NHR; ++r)
      hr[r] = h_col[r] >= 0 ? W[(long long)(h_lds[r] / T::LP) * N + h_col[r] + (long long)kk * S] : R(0);
  };

  // z window: slots 1..4 hold planes k0 - ZL .. k0 - ZL + 3; nxt holds k0 - ZL + 4
  R zq[9][5], nxt[9], hreg[T::NHR];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    zq[v][0] = R(0);
#pragma unroll
    for (int m = 1; m < 5; ++m) zq[v][m] = active ? W[(long long)v * N + col + (long long)(k0 - ZL + m - 1) * S] : R(0);
    nxt[v] = active ? W[(long long)v * N + col + (long long)(k0 - ZL + 4) * S] : R(0);
  }
  load_halo(k0, hreg);

  const int cidx = (ty + T::YL) * T::LX + tx + T::XL;
  for (int k = k0; k < k1; ++k) {
    const int buf = (k - k0) & 1;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
#pragma unroll
      for (int m = 0; m < 4; ++m) zq[v][m] = zq[v][m + 1];
      zq[v][4] = nxt[v];
    }
    // plane k into LDS buffer `buf` (the other buffer may still be read by
    // threads finishing plane k - 1)
#pragma unroll
    for (int v = 0; v < 9; ++v) sw[buf][v][cidx] = zq[v][ZL];
#pragma unroll
    for (int r = 0; r < T::NHR; ++r)
      if (h_lds[r] >= 0) (&sw[buf][0][0])[h_lds[r]] = hreg[r];
    __syncthreads();
    // prefetch plane k + 1: window entry k + 1 - ZL + 4 and halo of k + 1
    if (k + 1 < k1) {
#pragma unroll
      for (int v = 0; v < 9; ++v)
        nxt[v] = active ? W[(long long)v * N + col + (long long)(k + 1 - ZL + 4) * S] : R(0);
      load_halo(k + 1, hreg);
    }
    if (active) {
      const R* L = &sw[buf][0][0];
      cell_body<R, TM, FB1, FB2, FB3, STAGE>(
          a, coef, i, j, k, cell_off(a, i, j, k),
          [&](int v) -> R {
            const R* f = L + v * T::LP + cidx;
            return R(FB1) * (R(AF_1) * f[-FB1] + R(AF0) * f[0] + R(AF1) * f[FB1] + R(AF2) * f[2 * FB1] +
                             R(AF3) * f[3 * FB1]);
          },
          [&](int v) -> R {
            const R* f = L + v * T::LP + cidx;
            constexpr int s = T::LX;
            return R(FB2) * (R(AF_1) * f[-FB2 * s] + R(AF0) * f[0] + R(AF1) * f[FB2 * s] +
                             R(AF2) * f[2 * FB2 * s] + R(AF3) * f[3 * FB2 * s]);
          },
          [&](int v) -> R {
            return R(FB3) * (R(AF_1) * zq[v][ZL - FB3] + R(AF0) * zq[v][ZL] + R(AF1) * zq[v][ZL + FB3] +
                             R(AF2) * zq[v][ZL + 2 * FB3] + R(AF3) * zq[v][ZL + 3 * FB3]);
          },
          [&](int v) -> R { return zq[v][ZL]; });
    }
  }
}

// ---------------------------------------------------------------------------
// plane-tile interior kernel: a block of 64 x TY threads computes one 64 x TY
// tile of one k plane. The nine fields of the tile and its x/y halo (1 cell on
// one side, 3 on the other) are staged once into LDS, so the x and y
// differences read LDS; the z difference reads the four other planes from
// global memory (L2). Per cell this issues about half the vector memory
// instructions of the thread-per-cell kernel.
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE, int TY>
__global__ __launch_bounds__(64 * TY) void cg_plane(StageArgs<R> a, const TM* __restrict__ coef) {
  using T = ZTile<FB1, FB2, TY>;
  __shared__ R sw[9][T::LP];
  const int ntx = (a.nx + 63) / 64, nty = (a.ny + TY - 1) / TY;
  const int b = blockIdx.x;
  const int tile_x = b % ntx, tile_y = (b / ntx) % nty;
  const int k = a.kbeg + b / (ntx * nty);
  if (k >= a.kend) return;  // uniform per block, before the barrier
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int i = H + tile_x * 64 + tx, j = H + tile_y * TY + ty;
  const bool active = i < H + a.nx && j < H + a.ny;
  const long long N = a.N, S = a.S, NXl = a.NX;
  const long long p = (long long)i + NXl * j + (long long)k * S;
  const R* W = a.win;
  const int cidx = (ty + T::YL) * T::LX + tx + T::XL;
  R* L = &sw[0][0];
#pragma unroll
  for (int v = 0; v < 9; ++v) L[v * T::LP + cidx] = active ? W[(long long)v * N + p] : R(0);
  for (int e = threadIdx.x; e < 9 * T::NHALO; e += T::NT) {
    const int v = e / T::NHALO, q = e - v * T::NHALO;
    int lx, ly;
    T::halo_xy(q, lx, ly);
    const int gi = H + tile_x * 64 + lx - T::XL, gj = H + tile_y * TY + ly - T::YL;
    L[v * T::LP + ly * T::LX + lx] =
        (gi >= 0 && gi < a.NX && gj >= 0 && gj < a.NY) ? W[(long long)v * N + gi + NXl * gj + (long long)k * S] : R(0);
  }
  __syncthreads();
  if (!active) return;
  cell_body<R, TM, FB1, FB2, FB3, STAGE>(
      a, coef, i, j, k, cell_off(a, i, j, k),
      [&](int v) -> R {
        const R* f = L + v * T::LP + cidx;
        return R(FB1) * (R(AF_1) * f[-FB1] + R(AF0) * f[0] + R(AF1) * f[FB1] + R(AF2) * f[2 * FB1] +
                         R(AF3) * f[3 * FB1]);
      },
      [&](int v) -> R {
        const R* f = L + v * T::LP + cidx;
        constexpr int s = T::LX;
        return R(FB2) * (R(AF_1) * f[-FB2 * s] + R(AF0) * f[0] + R(AF1) * f[FB2 * s] + R(AF2) * f[2 * FB2 * s] +
                         R(AF3) * f[3 * FB2 * s]);
      },
      [&](int v) -> R {
        const R* f = W + (long long)v * N + p;
        return R(FB3) * (R(AF_1) * f[-FB3 * S] + R(AF0) * L[v * T::LP + cidx] + R(AF1) * f[FB3 * S] +
                         R(AF2) * f[2 * FB3 * S] + R(AF3) * f[3 * FB3 * S]);
      },
      [&](int v) -> R { return L[v * T::LP + cidx]; });
}

// ---------------------------------------------------------------------------
// surface kernel: top three layers, traction image
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_surface(StageArgs<R> a, const TM* __restrict__ coef, SurfArgs<R> s) {
  const int i = H + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = H + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ktop = H + a.nz - 1;
  const int k = ktop - blockIdx.z;  // blockIdx.z = 0,1,2 -> surface, one below, two below
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long N = a.N, S = a.S, NXl = a.NX;
  const long long p = (long long)i + NXl * j + (long long)k * S;
  const R* W = a.win;
  // q lies in layer k (same-plane neighbours) unless kq says otherwise
  auto cfk = [&](int c, long long q, int kq) -> R { return coef_at<R, TM, false>(coef, a, c, q, kq); };
  auto cf = [&](int c, long long q) -> R { return cfk(c, q, k); };
  const R mu = cf(M_MU, p), lam = cf(M_LAM, p), bu = cf(M_B, p);
  const PmlArgs<R>& P = a.pml;
  const R bx = P.beta[0][i], by = P.beta[1][j];
  const int ks = ktop - k + H;  // 3, 4, 5: index of the surface in the 7-point image window

  // J * (m . T) at cell q for metric set c0 (0 = xi, 3 = et, 6 = zt)
  auto jt = [&](long long q, int kq, int c0, R* o) {
    R m1 = cfk(c0, q, kq), m2 = cfk(c0 + 1, q, kq), m3 = cfk(c0 + 2, q, kq), jac = s.jac[q];
    R txx = W[3 * N + q], tyy = W[4 * N + q], tzz = W[5 * N + q];
    R txy = W[6 * N + q], txz = W[7 * N + q], tyz = W[8 * N + q];
    o[0] = (m1 * txx + m2 * txy + m3 * txz) * jac;
    o[1] = (m1 * txy + m2 * tyy + m3 * tyz) * jac;
    o[2] = (m1 * txz + m2 * tyz + m3 * tzz) * jac;
  };
  // conservative velocity derivatives (freeSurface.cpp:357-411)
  R gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
  const R cw[5] = {R(AF_1), R(AF0), R(AF1), R(AF2), R(AF3)};
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    int off1 = (t - 1) * FB1, off2 = (t - 1) * FB2;
    R o[3];
    jt(p + off1, k, 0, o);
#pragma unroll
    for (int c = 0; c < 3; ++c) gx[c] += cw[t] * o[c];
    jt(p + (long long)off2 * NXl, k, 3, o);
#pragma unroll
    for (int c = 0; c < 3; ++c) gy[c] += cw[t] * o[c];
    // z: image about the surface, J T = 0 on it
    int l = 3 + (t - 1) * FB3;
    if (l < ks) {
      jt(p + (long long)(l - 3) * S, k + (l - 3), 6, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) gz[c] += cw[t] * o[c];
    } else if (l > ks) {
      jt(p + (long long)(2 * ks - l - 3) * S, k + (2 * ks - l - 3), 6, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) gz[c] -= cw[t] * o[c];
    }
  }
  const R jinv = R(1) / s.jac[p];
  R h[9];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    h[c] = bu * jinv * ((R(FB1) * gx[c] * a.rDH) * bx + (R(FB2) * gy[c] * a.rDH) * by + (R(FB3) * gz[c] * a.rDH)) * a.DT;

  // stress: x/y derivatives of V, z derivative from the surface condition
  const bool in_x1 = P.npml > 0 && i < H + P.npml;
  const bool in_x2 = P.npml > 0 && i >= H + a.nx - P.npml;
  const bool in_y1 = P.npml > 0 && j < H + P.npml;
  const bool in_y2 = P.npml > 0 && j >= H + a.ny - P.npml;
  const bool in_z1 = P.npml > 0 && k < H + P.npml;
  R Dx[9], Dy[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    Dx[v] = mac<FB1>(W + (long long)v * N + p, 1) * a.rDH;
    Dy[v] = mac<FB2>(W + (long long)v * N + p, NXl) * a.rDH;
  }
  R Vxi[3], Vet[3], Vzt[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    Vxi[v] = Dx[v] * bx;
    Vet[v] = Dy[v] * by;
  }
  const long long sp = (long long)i + NXl * j;
  const long long SS = (long long)a.NX * a.NY;
  if (k == ktop) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
      Vzt[r] = (s.rdzdx[(3 * r) * SS + sp] * Vxi[0] + s.rdzdx[(3 * r + 1) * SS + sp] * Vxi[1] + s.rdzdx[(3 * r + 2) * SS + sp] * Vxi[2]) +
               (s.rdzdy[(3 * r) * SS + sp] * Vet[0] + s.rdzdy[(3 * r + 1) * SS + sp] * Vet[1] + s.rdzdy[(3 * r + 2) * SS + sp] * Vet[2]);
  } else if (k == ktop - 1) {
#pragma unroll
    for (int v = 0; v < 3; ++v) Vzt[v] = R(FB3) * (W[(long long)v * N + p + FB3 * S] - W[(long long)v * N + p]) * a.rDH;
  } else {
#pragma unroll
    for (int v = 0; v < 3; ++v)
      Vzt[v] = R(FB3) * (R(CF1) * W[(long long)v * N + p] + R(CF2) * W[(long long)v * N + p + FB3 * S] +
                         R(CF3) * W[(long long)v * N + p + 2 * FB3 * S]) * a.rDH;
  }
  {
    R s1[6], s2[6], s3[6];
    stress_rows(Vxi, cf(M_XIX, p), cf(M_XIY, p), cf(M_XIZ, p), mu, lam, s1);
    stress_rows(Vet, cf(M_ETX, p), cf(M_ETY, p), cf(M_ETZ, p), mu, lam, s2);
    stress_rows(Vzt, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, s3);
#pragma unroll
    for (int v = 0; v < 6; ++v) h[3 + v] = (s1[v] + s2[v] + s3[v]) * a.DT;
  }
  // PML faces crossing the surface layers: auxiliaries use the ordinary
  // derivatives (pml_deriv.cpp), plus the surface stress term on the top layer
  // (pml_freeSurface.cpp:90-125).
  if (in_x1 || in_x2) {
    R Q[9];
    assemble(Dx, cf(M_XIX, p), cf(M_XIY, p), cf(M_XIZ, p), mu, lam, bu, Q);
    R d = P.d[0][i], ad = d + P.alpha[0][i];
    R ex[6] = {0, 0, 0, 0, 0, 0};
    const bool top = (k == ktop);
    if (top) {
      R V[3] = {Dx[0] * d, Dx[1] * d, Dx[2] * d}, Vz[3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        Vz[r] = s.rdzdx[(3 * r) * SS + sp] * V[0] + s.rdzdx[(3 * r + 1) * SS + sp] * V[1] + s.rdzdx[(3 * r + 2) * SS + sp] * V[2];
      stress_rows(Vz, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, ex);
    }
    long long rowq = (long long)(j - H) + (long long)a.ny * (k - H);
    if (in_x1) pml_face<R, STAGE>(P.f[0], (long long)(i - H) + (long long)P.npml * rowq, Q, d, ad, bx, a.DT, h, top, ex);
    if (in_x2) pml_face<R, STAGE>(P.f[3], (long long)(i - H - (a.nx - P.npml)) + (long long)P.npml * rowq, Q, d, ad, bx, a.DT, h, top, ex);
  }
  if (in_y1 || in_y2) {
    R Q[9];
    assemble(Dy, cf(M_ETX, p), cf(M_ETY, p), cf(M_ETZ, p), mu, lam, bu, Q);
    R d = P.d[1][j], ad = d + P.alpha[1][j];
    R ex[6] = {0, 0, 0, 0, 0, 0};
    const bool top = (k == ktop);
    if (top) {
      R V[3] = {Dy[0] * d, Dy[1] * d, Dy[2] * d}, Vz[3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        Vz[r] = s.rdzdy[(3 * r) * SS + sp] * V[0] + s.rdzdy[(3 * r + 1) * SS + sp] * V[1] + s.rdzdy[(3 * r + 2) * SS + sp] * V[2];
      stress_rows(Vz, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, ex);
    }
    long long i0 = i - H, k0 = k - H;
    if (in_y1) pml_face<R, STAGE>(P.f[1], i0 + (long long)a.nx * ((j - H) + (long long)P.npml * k0), Q, d, ad, by, a.DT, h, top, ex);
    if (in_y2) pml_face<R, STAGE>(P.f[4], i0 + (long long)a.nx * ((j - H - (a.ny - P.npml)) + (long long)P.npml * k0), Q, d, ad, by, a.DT, h, top, ex);
  }
  if (in_z1) {  // only for very shallow grids: bottom PML reaching the surface layers
    R Dz[9], Q[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) Dz[v] = mac<FB3>(W + (long long)v * N + p, S) * a.rDH;
    assemble(Dz, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, bu, Q);
    R d = P.d[2][k], ad = d + P.alpha[2][k], bz = P.beta[2][k];
    const R none[6] = {0, 0, 0, 0, 0, 0};
    pml_face<R, STAGE>(P.f[2], (long long)(i - H) + (long long)a.nx * ((j - H) + (long long)a.ny * (k - H)), Q, d, ad, bz, a.DT, h, false, none);
  }
  // RK4 stage update: every load before the first store (see cell_body)
  R out[9], tv[9], mv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const long long q = (long long)v * N + p;
    if constexpr (STAGE == 0) mv[v] = W[q];
    if constexpr (STAGE != 0) tv[v] = a.wt[q];
    if constexpr (STAGE == 1 || STAGE == 2) mv[v] = a.wm[q];
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const long long q = (long long)v * N + p;
    if constexpr (STAGE == 0) {
      a.wt[q] = mv[v] + R(RK_B1) * h[v];
      out[v] = mv[v] + R(RK_A2) * h[v];
    } else if constexpr (STAGE == 1) {
      a.wt[q] = tv[v] + R(RK_B2) * h[v];
      out[v] = mv[v] + R(RK_A3) * h[v];
    } else if constexpr (STAGE == 2) {
      a.wt[q] = tv[v] + R(RK_B3) * h[v];
      out[v] = mv[v] + h[v];
    } else {
      out[v] = tv[v] + R(RK_B4) * h[v];
    }
    a.wout[q] = out[v];
  }
  if (STAGE == 3 && k == ktop && s.pgv != nullptr) {
    // observers on the final state of the step (PGV.cpp:62-175, solve_displacement.cpp:28-63)
    float fx = (float)out[0], fy = (float)out[1], fz = (float)out[2];
    float vh = sqrtf(fx * fx + fy * fy), va = sqrtf(fx * fx + fy * fy + fz * fz);
    long long q = (long long)(i - H) + (long long)a.nx * (j - H);
    if (s.pgv[2 * q] < vh) s.pgv[2 * q] = vh;
    if (s.pgv[2 * q + 1] < va) s.pgv[2 * q + 1] = va;
    s.dis[sp] += out[0] * a.DT;
    s.dis[SS + sp] += out[1] * a.DT;
    s.dis[2 * SS + sp] += out[2] * a.DT;
  }
}

// ---------------------------------------------------------------------------
// source injection and station recording
// ---------------------------------------------------------------------------
template <class R>
__global__ void cg_add_source(R* W, long long N, int np, const long long* __restrict__ cell, const R* __restrict__ rate,
                              int srcIt, R tw, R DT) {
  // moment-rate interpolation and W -= Mdot DT (addMoment.cpp:101-146,192-260)
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  long long c = cell[p];
  const R* m0 = rate + ((long long)srcIt * np + p) * 6;
  const R* m1 = rate + ((long long)(srcIt + 1) * np + p) * 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    R s = (m1[q] - m0[q]) * tw + m0[q];
    W[(long long)(3 + q) * N + c] -= s * DT;
  }
}

template <class R>
__global__ void cg_record(const R* W, long long N, int ns, const long long* __restrict__ cell, float* rec, int nt, int it) {
  // station.cpp:303-380 layout [ns][NT][9]
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  long long c = cell[s];
  float* o = rec + ((long long)s * nt + it) * 9;
#pragma unroll
  for (int v = 0; v < 9; ++v) o[v] = (float)W[(long long)v * N + c];
}

}  // namespace cg
