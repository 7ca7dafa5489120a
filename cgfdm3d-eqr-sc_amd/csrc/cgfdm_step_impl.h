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
  // extra: surface stress term added to hA[3..8] on the top layer
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    long long q = (long long)v * F.cnt + pos;
    R cur = STAGE == 0 ? F.a[q] : F.b[q];
    R hA = (d * Q[v] - ad * cur) * DT;
    if (has_extra && v >= 3) hA += extra[v - 3] * DT;
    h[v] -= be * cur * DT;
    if constexpr (STAGE == 0) {
      F.t[q] = cur + R(RK_B1) * hA;
      F.b[q] = cur + R(RK_A2) * hA;
    } else if constexpr (STAGE == 1) {
      F.t[q] = F.t[q] + R(RK_B2) * hA;
      F.b[q] = F.a[q] + R(RK_A3) * hA;
    } else if constexpr (STAGE == 2) {
      F.t[q] = F.t[q] + R(RK_B3) * hA;
      F.b[q] = F.a[q] + hA;
    } else {
      F.a[q] = F.t[q] + R(RK_B4) * hA;
    }
  }
}

// ---------------------------------------------------------------------------
// interior cell update: 27 derivatives, PML faces, RK4 stage. `zd(v)` returns
// the unscaled zeta difference of field v and `wc(v)` the centre value, so the
// same body serves the thread-per-cell and the z-marching kernels.
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE, class ZD, class WC>
__device__ __forceinline__ void cell_body(const StageArgs<R>& a, const TM* __restrict__ coef, int i, int j, int k,
                                          ZD zd, WC wc) {
  const long long N = a.N, S = a.S;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * S;
  const R* W = a.win;
  auto cf = [&](int c) -> R { return R(Store<TM>::ld(coef + (long long)c * N + p)); };
  const R mu = cf(M_MU), lam = cf(M_LAM), bu = cf(M_B);
  const PmlArgs<R>& P = a.pml;
  R h[9], D[9], Q[9];

  // x: derivatives along xi (wave_deriv.cpp:214-222), scaled by 1/beta_x
#pragma unroll
  for (int v = 0; v < 9; ++v) D[v] = mac<FB1>(W + (long long)v * N + p, 1);
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
  for (int v = 0; v < 9; ++v) D[v] = mac<FB2>(W + (long long)v * N + p, a.NX);
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
  // RK4 stage update of W (wave_rk.cpp:18-104)
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const long long q = (long long)v * N + p;
    if constexpr (STAGE == 0) {
      const R m = wc(v);
      a.wt[q] = m + R(RK_B1) * h[v];
      a.wout[q] = m + R(RK_A2) * h[v];
    } else if constexpr (STAGE == 1) {
      a.wt[q] = a.wt[q] + R(RK_B2) * h[v];
      a.wout[q] = a.wm[q] + R(RK_A3) * h[v];
    } else if constexpr (STAGE == 2) {
      a.wt[q] = a.wt[q] + R(RK_B3) * h[v];
      a.wout[q] = a.wm[q] + h[v];
    } else {
      a.wout[q] = a.wt[q] + R(RK_B4) * h[v];
    }
  }
}

// ---------------------------------------------------------------------------
// interior kernel: one thread per cell below the surface layers
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_cell(StageArgs<R> a, const TM* __restrict__ coef) {
  const int i = H + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = H + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = a.kbeg + blockIdx.z;
  if (i >= H + a.nx || j >= H + a.ny || k >= a.kend) return;
  const long long N = a.N, S = a.S;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * S;
  const R* W = a.win;
  cell_body<R, TM, FB1, FB2, FB3, STAGE>(
      a, coef, i, j, k, [&](int v) -> R { return mac<FB3>(W + (long long)v * N + p, S); },
      [&](int v) -> R { return W[(long long)v * N + p]; });
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
  auto cf = [&](int c, long long q) -> R { return R(Store<TM>::ld(coef + (long long)c * N + q)); };
  const R mu = cf(M_MU, p), lam = cf(M_LAM, p), bu = cf(M_B, p);
  const PmlArgs<R>& P = a.pml;
  const R bx = P.beta[0][i], by = P.beta[1][j];
  const int ks = ktop - k + H;  // 3, 4, 5: index of the surface in the 7-point image window

  // J * (m . T) at cell q for metric set c0 (0 = xi, 3 = et, 6 = zt)
  auto jt = [&](long long q, int c0, R* o) {
    R m1 = cf(c0, q), m2 = cf(c0 + 1, q), m3 = cf(c0 + 2, q), jac = s.jac[q];
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
    jt(p + off1, 0, o);
#pragma unroll
    for (int c = 0; c < 3; ++c) gx[c] += cw[t] * o[c];
    jt(p + (long long)off2 * NXl, 3, o);
#pragma unroll
    for (int c = 0; c < 3; ++c) gy[c] += cw[t] * o[c];
    // z: image about the surface, J T = 0 on it
    int l = 3 + (t - 1) * FB3;
    if (l < ks) {
      jt(p + (long long)(l - 3) * S, 6, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) gz[c] += cw[t] * o[c];
    } else if (l > ks) {
      jt(p + (long long)(2 * ks - l - 3) * S, 6, o);
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
  R out[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    long long q = (long long)v * N + p;
    if constexpr (STAGE == 0) {
      R m = W[q];
      a.wt[q] = m + R(RK_B1) * h[v];
      out[v] = m + R(RK_A2) * h[v];
    } else if constexpr (STAGE == 1) {
      a.wt[q] = a.wt[q] + R(RK_B2) * h[v];
      out[v] = a.wm[q] + R(RK_A3) * h[v];
    } else if constexpr (STAGE == 2) {
      a.wt[q] = a.wt[q] + R(RK_B3) * h[v];
      out[v] = a.wm[q] + h[v];
    } else {
      out[v] = a.wt[q] + R(RK_B4) * h[v];
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
