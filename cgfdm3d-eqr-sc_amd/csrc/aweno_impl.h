// ME-AWENO (SCFDM) step kernels for gfx950: WENO5 alternative-flux face
// fluxes, flux-difference update fused with TVD-RK3, characteristic free
// surface, exponential sponge, strain-rate source and observers.
//
// Layout: conservative state W = (rho vx, rho vy, rho vz, exx, eyy, ezz, gyz,
// gxz, gxy) as 9 SoA planes over the padded grid (x fastest, 3 ghost layers).
// Coefficients: 19 SoA planes (CJM order of the SCFDM build, CJM.cpp:107-147):
// 0..8 cofactors J*xi_x .. J*zt_z, 9 J, 10 mu, 11 lambda, 12 1/rho,
// 13..15 x_xi, 16..18 x_eta.
#pragma once
#include <algorithm>

#include "aweno.h"

namespace aw {

using cg::H;

// Reciprocal used for the WENO weights: the hardware reciprocal (1 ulp) in
// FP32, an exact division in FP64.
__device__ __forceinline__ float wrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double wrcp(double x) { return 1.0 / x; }

// WENO5-JS value at x_{i+1/2} from u1..u5 = u(i-2..i+2)
// (alternative_flux_FD.cpp:43-59). The smoothness indicators are written as
// the sums of squares they expand to, e.g. (10 u3^2 - 31 u3 u4 + 25 u4^2 +
// 11 u3 u5 - 19 u4 u5 + 4 u5^2) / 3 = 13/12 (u3 - 2 u4 + u5)^2 +
// 1/4 (3 u3 - 4 u4 + u5)^2; the weights are d_k / (beta_k + 1e-6)^2 with
// d = (5/16, 10/16, 1/16), normalised.
template <class R>
__device__ __forceinline__ R weno5(R u1, R u2, R u3, R u4, R u5) {
  // the reference scales the expanded form by the literal 0.3333333333, i.e.
  // the sums of squares by 0.9999999999 (exactly 1 in FP32)
  const R kb = R(3.0 * 0.3333333333);
  const R c13 = kb * R(13.0 / 12.0), c14 = kb * R(0.25);
  const R a1 = u3 - R(2) * u4 + u5, e1 = R(3) * u3 - R(4) * u4 + u5;
  const R a2 = u2 - R(2) * u3 + u4, e2 = u2 - u4;
  const R a3 = u1 - R(2) * u2 + u3, e3 = u1 - R(4) * u2 + R(3) * u3;
  const R b1 = c13 * a1 * a1 + c14 * e1 * e1;
  const R b2 = c13 * a2 * a2 + c14 * e2 * e2;
  const R b3 = c13 * a3 * a3 + c14 * e3 * e3;
  const R eps = R(1e-6);
  const R q1 = b1 + eps, q2 = b2 + eps, q3 = b3 + eps;
  const R w1 = R(0.3125) * wrcp(q1 * q1);
  const R w2 = R(0.625) * wrcp(q2 * q2);
  const R w3 = R(0.0625) * wrcp(q3 * q3);
  const R tot = wrcp(w1 + w2 + w3);
  return (w1 * tot) * (R(0.375) * u3 + R(0.75) * u4 - R(0.125) * u5) +
         (w2 * tot) * (R(-0.125) * u2 + R(0.75) * u3 + R(0.375) * u4) +
         (w3 * tot) * (R(0.375) * u1 - R(1.25) * u2 + R(1.875) * u3);
}

// FP32 form of the same value with fewer operations: the smoothness indicators
// are taken times 12 (13 a^2 + 3 e^2, with eps times 12), which scales every
// weight by 1/144 and cancels in the normalisation; the linear weights d_k are
// folded into the candidate interpolants and into the weight sum, so that the
// value is (sum_k r_k d_k p_k) / (sum_k d_k r_k) with r_k = 1 / q_k^2. Equal to
// the form above up to rounding.
template <>
__device__ __forceinline__ float weno5<float>(float u1, float u2, float u3, float u4, float u5) {
  const float a1 = u3 - 2.0f * u4 + u5, e1 = 3.0f * u3 - 4.0f * u4 + u5;
  const float a2 = u2 - 2.0f * u3 + u4, e2 = u2 - u4;
  const float a3 = u1 - 2.0f * u2 + u3, e3 = u1 - 4.0f * u2 + 3.0f * u3;
  const float q1 = fmaf(13.0f * a1, a1, fmaf(3.0f * e1, e1, 12e-6f));
  const float q2 = fmaf(13.0f * a2, a2, fmaf(3.0f * e2, e2, 12e-6f));
  const float q3 = fmaf(13.0f * a3, a3, fmaf(3.0f * e3, e3, 12e-6f));
  const float r1 = wrcp(q1 * q1), r2 = wrcp(q2 * q2), r3 = wrcp(q3 * q3);
  const float sum = fmaf(0.3125f, r1, fmaf(0.625f, r2, 0.0625f * r3));
  // d_k times the candidate interpolants (0.375 u3 + 0.75 u4 - 0.125 u5 etc.)
  const float p1 = 0.1171875f * u3 + 0.234375f * u4 - 0.0390625f * u5;
  const float p2 = -0.078125f * u2 + 0.46875f * u3 + 0.234375f * u4;
  const float p3 = 0.0234375f * u1 - 0.078125f * u2 + 0.1171875f * u3;
  return fmaf(r1, p1, fmaf(r2, p2, r3 * p3)) * wrcp(sum);
}

// FP32 WENO5 values of both sides of face q+1/2 from w(q-2 .. q+3):
// um = weno5(w(q-2), .., w(q+2)) and up = weno5(w(q+3), .., w(q-1)), the same
// smoothness indicators (times 12) and folded linear weights as weno5<float>.
// The two sides share the second differences of the windows (q-1, q, q+1) and
// (q, q+1, q+2) and two of their three candidate values (the parabola through
// the same three cells at the same face): 4 second differences and 4
// candidates per face instead of 6 and 6. With A_c, C_c the second and the
// centred first difference of the window starting at c, the one-sided
// differences of the indicators are C_c + 2 A_c and 2 A_c - C_c.
__device__ __forceinline__ void weno5_pair(float wm2, float wm1, float w0, float wp1, float wp2, float wp3, float& um,
                                           float& up) {
  const float Am2 = wm2 - 2.0f * wm1 + w0, Am1 = wm1 - 2.0f * w0 + wp1;
  const float A0 = w0 - 2.0f * wp1 + wp2, A1 = wp1 - 2.0f * wp2 + wp3;
  const float Cm2 = wm2 - w0, Cm1 = wm1 - wp1, C0 = w0 - wp2, C1 = wp1 - wp3;
  const float Tm1 = fmaf(13.0f * Am1, Am1, 12e-6f), T0 = fmaf(13.0f * A0, A0, 12e-6f);
  const float e1 = fmaf(2.0f, A0, C0), e3 = fmaf(2.0f, Am2, -Cm2);
  const float f1 = fmaf(2.0f, Am1, -Cm1), f3 = fmaf(2.0f, A1, C1);
  // um: windows (q, q+1, q+2), (q-1, q, q+1), (q-2, q-1, q)
  const float q1 = fmaf(3.0f * e1, e1, T0);
  const float q2 = fmaf(3.0f * Cm1, Cm1, Tm1);
  const float q3 = fmaf(13.0f * Am2, Am2, fmaf(3.0f * e3, e3, 12e-6f));
  // up: windows (q-1, q, q+1), (q, q+1, q+2), (q+1, q+2, q+3)
  const float r1q = fmaf(3.0f * f1, f1, Tm1);
  const float r2q = fmaf(3.0f * C0, C0, T0);
  const float r3q = fmaf(13.0f * A1, A1, fmaf(3.0f * f3, f3, 12e-6f));
  // d1 = 5/16 times the candidates shared by both sides, and the one-sided
  // extrapolations times d3 = 1/16 (d2 = 2 d1)
  const float PA = 0.1171875f * w0 + 0.234375f * wp1 - 0.0390625f * wp2;   // d1 (3 w0 + 6 wp1 - wp2) / 8
  const float PB = -0.0390625f * wm1 + 0.234375f * w0 + 0.1171875f * wp1;  // d1 (-wm1 + 6 w0 + 3 wp1) / 8
  const float Pm = 0.0234375f * wm2 - 0.078125f * wm1 + 0.1171875f * w0;   // d3 (3 wm2 - 10 wm1 + 15 w0) / 8
  const float Pp = 0.0234375f * wp3 - 0.078125f * wp2 + 0.1171875f * wp1;
  // weights r_k = 1 / q_k^2 (hardware reciprocal, 1 ulp). A form with one
  // reciprocal per side (products of the other two squares) measured slower
  // on the MI355X although the reciprocal costs about 3.4 FMAs there.
  {
    const float a1 = __builtin_amdgcn_rcpf(q1 * q1), a2 = __builtin_amdgcn_rcpf(q2 * q2),
                a3 = __builtin_amdgcn_rcpf(q3 * q3);
    const float sum = fmaf(0.3125f, a1, fmaf(0.625f, a2, 0.0625f * a3));
    um = fmaf(a1, PA, fmaf(2.0f * a2, PB, a3 * Pm)) * __builtin_amdgcn_rcpf(sum);
  }
  {
    const float a1 = __builtin_amdgcn_rcpf(r1q * r1q), a2 = __builtin_amdgcn_rcpf(r2q * r2q),
                a3 = __builtin_amdgcn_rcpf(r3q * r3q);
    const float sum = fmaf(0.3125f, a1, fmaf(0.625f, a2, 0.0625f * a3));
    up = fmaf(a1, PB, fmaf(2.0f * a2, PA, a3 * Pp)) * __builtin_amdgcn_rcpf(sum);
  }
}

// Conservative elastic flux through a face of cofactor normal (n1, n2, n3)
// (alternative_flux_FD.cpp:140-158).
template <class R>
__device__ __forceinline__ void phys_flux(const R (&u)[9], R n1, R n2, R n3, R mu, R lam, R b, R (&f)[9]) {
  const R l2m = lam + R(2) * mu;
  f[0] = -(n1 * (lam * u[4] + lam * u[5] + u[3] * l2m) + mu * n2 * u[8] + mu * n3 * u[7]);
  f[1] = -(n2 * (lam * u[3] + lam * u[5] + u[4] * l2m) + mu * n1 * u[8] + mu * n3 * u[6]);
  f[2] = -(n3 * (lam * u[3] + lam * u[4] + u[5] * l2m) + mu * n1 * u[7] + mu * n2 * u[6]);
  f[3] = -((n1 * u[0]) * b);
  f[4] = -((n2 * u[1]) * b);
  f[5] = -((n3 * u[2]) * b);
  f[6] = -((n2 * u[2]) * b + (n3 * u[1]) * b);
  f[7] = -((n1 * u[2]) * b + (n3 * u[0]) * b);
  f[8] = -((n1 * u[1]) * b + (n2 * u[0]) * b);
}

using cg::at;

// value of plane c of a grid array at byte offset o (SGPR base + 32-bit VGPR
// offset addressing, as in the CGFDM stage kernels)
template <class R>
__device__ __forceinline__ R pl(const R* base, long long N, int c, unsigned o) {
  return *at(base + (long long)c * N, o);
}

// Lax-Friedrichs face flux of direction d at face q+1/2, stored at cell q
// (wave_deriv_alternative_flux_FD_{x,y,z}); o is the byte offset of cell q,
// sb the byte stride of direction d.
template <class R>
__device__ __forceinline__ void face_flux(const StageArgs<R>& a, unsigned o, unsigned sb, int d) {
  const long long N = a.N;
  const R n1 = pl(a.cjm, N, 3 * d + 0, o), n2 = pl(a.cjm, N, 3 * d + 1, o), n3 = pl(a.cjm, N, 3 * d + 2, o);
  const R mu = pl(a.cjm, N, C_MU, o), lam = pl(a.cjm, N, C_LAM, o), b = pl(a.cjm, N, C_B, o);
  R um[9], up[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const R* w = a.win + v * N;
    const R wm2 = *at(w, o - 2 * sb), wm1 = *at(w, o - sb), w0 = *at(w, o), wp1 = *at(w, o + sb),
            wp2 = *at(w, o + 2 * sb), wp3 = *at(w, o + 3 * sb);
    if constexpr (sizeof(R) == 4) {
      weno5_pair(wm2, wm1, w0, wp1, wp2, wp3, um[v], up[v]);
    } else {
      um[v] = weno5<R>(wm2, wm1, w0, wp1, wp2);
      up[v] = weno5<R>(wp3, wp2, wp1, w0, wm1);
    }
  }
  R* F = a.F + (long long)(9 * d) * N;
  if constexpr (sizeof(R) == 4) {
    // the flux is linear in the state: F(u+) + F(u-) = F(u+ + u-), one flux
    // evaluation per face instead of two (equal up to rounding)
    R us[9], fs[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) us[v] = up[v] + um[v];
    phys_flux<R>(us, n1, n2, n3, mu, lam, b, fs);
#pragma unroll
    // non-temporal: the flux planes stream out and back in the update kernel
    // without displacing the state rows the neighbouring faces re-read
    for (int v = 0; v < 9; ++v)
      __builtin_nontemporal_store(R(R(0.5) * (fs[v] - a.alpha * (up[v] - um[v]))), at(F + v * N, o));
  } else {
    R fm[9], fp[9];
    phys_flux<R>(um, n1, n2, n3, mu, lam, b, fm);
    phys_flux<R>(up, n1, n2, n3, mu, lam, b, fp);
#pragma unroll
    for (int v = 0; v < 9; ++v) *at(F + v * N, o) = R(0.5) * (fp[v] + fm[v] - a.alpha * (up[v] - um[v]));
  }
}

// Face fluxes of the three directions. Thread (i, j, k) over [2, n-3)^3
// computes the x face when j, k are interior, the y face when i, k are, the
// z face when i, j are: exactly the faces the reference writes, so every other
// entry of F keeps the zero it was allocated with.
template <class R>
__global__ void __launch_bounds__(256) aw_faces(StageArgs<R> a) {
  // 64-column tiles start at the first interior column i = 3 (on a cache line
  // in the row-aligned layout) and cover the nx interior columns: ceil(nx/64)
  // tiles per row, no tile of a few columns. Column 2 has only its x face,
  // which the lane of column 3 computes as a second face. Tiles cover j in
  // [2, NY - 3), k in [2, NZ - 3) in the XCD-aware order of the CGFDM interior
  // kernel.
  cg::CellTile t;
  if (!cg::interior_tile(blockIdx.x, a.XE - 2 * H, a.NY - 5, a.NZ - 5, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = 2 + t.j0 + (threadIdx.x >> 6);
  const int k = 2 + t.k;
  if (i >= a.XE - H || j >= a.NY - 3) return;
  const unsigned o = (unsigned)i * (unsigned)sizeof(R) + (unsigned)j * a.rowb + (unsigned)k * a.planeb;
  const bool ji = j >= H, ki = k >= H;
  if (ji && ki) {
    face_flux<R>(a, o, (unsigned)sizeof(R), 0);
    if (i == H) face_flux<R>(a, o - (unsigned)sizeof(R), (unsigned)sizeof(R), 0);
  }
  if (ki) face_flux<R>(a, o, a.rowb, 1);
  if (ji) face_flux<R>(a, o, a.planeb, 2);
}

template <class R>
__device__ __forceinline__ R d2c(R u1, R u2, R u3, R u4, R u5) {
  return R(1.0 / 12.0) * (-u1 + R(16) * u2 - R(30) * u3 + R(16) * u4 - u5);
}
template <class R>
__device__ __forceinline__ R d4c(R u1, R u2, R u3, R u4, R u5) {
  return u1 - R(4) * u2 + R(6) * u3 - R(4) * u4 + u5;
}
template <class R>
__device__ __forceinline__ R d2f(R u1, R u2, R u3, R u4, R u5, R u6) {
  return R(1.0 / 48.0) * (R(-5) * u1 + R(39) * u2 - R(34) * u3 - R(34) * u4 + R(39) * u5 - R(5) * u6);
}
template <class R>
__device__ __forceinline__ R d4f(R u1, R u2, R u3, R u4, R u5, R u6) {
  return R(0.5) * (u1 - R(3) * u2 + R(2) * u3 + R(2) * u4 - R(3) * u5 + u6);
}

// TVD-RK3 combination (wave_rk_tvd.cpp:18-78); stages 1 and 2 in FP64 as the
// reference's double literals make them.
template <class R>
__device__ __forceinline__ R tvd(int stage, R m, R w, R h) {
  if (stage == 0) return w + h;
  if (stage == 1) return R(0.75 * (double)m + 0.25 * (double)w + 0.25 * (double)h);
  return R((1.0 / 3.0) * (double)m + (2.0 / 3.0) * (double)w + (2.0 / 3.0) * (double)h);
}

// Flux difference with the alternative-flux correction in all three
// directions (cal_du_{x,y,z}, alternative_flux_FD.cpp:1216-1435), fused with
// the TVD-RK3 update, at padded cell (i, j, k): outside the interior the
// right-hand side is zero (the reference's TVD kernels run over the whole
// padded arrays). Cells of the three layers under the top are left to
// aw_update_top.
template <class R>
__device__ __forceinline__ void update_cell(const StageArgs<R>& a, int i, int j, int k) {
  const long long N = a.N;
  const unsigned o = (unsigned)i * (unsigned)sizeof(R) + (unsigned)j * a.rowb + (unsigned)k * a.planeb;
  const bool interior = i >= H && i < a.XE - H && j >= H && j < a.NY - H && k >= H && k < a.NZ - H;
  const int ktop3 = a.NZ - H - 3;
  if (interior && k >= ktop3) return;
  R h[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) h[v] = R(0);
  if (interior) {
    const R jac = R(1.0 / (double)pl(a.cjm, N, C_J, o));
    const R cd = a.DT * a.rDH * jac;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const unsigned sb = d == 0 ? (unsigned)sizeof(R) : d == 1 ? a.rowb : a.planeb;
      const R* F = a.F + (long long)(9 * d) * N;
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        const R* f = F + v * N;
        const R f_3 = *at(f, o - 3 * sb), f_2 = *at(f, o - 2 * sb), f_1 = *at(f, o - sb), f0 = *at(f, o),
                f1 = *at(f, o + sb), f2 = *at(f, o + 2 * sb);
        const R fr = f0 - R(1.0 / 24.0) * d2c<R>(f_2, f_1, f0, f1, f2) + R(7.0 / 5760.0) * d4c<R>(f_2, f_1, f0, f1, f2);
        const R fl = f_1 - R(1.0 / 24.0) * d2c<R>(f_3, f_2, f_1, f0, f1) + R(7.0 / 5760.0) * d4c<R>(f_3, f_2, f_1, f0, f1);
        h[v] += cd * (-(fr - fl));
      }
    }
  }
  // every load before the first store: the state buffers are not known not
  // to alias, and interleaving would serialise the loads
  R mv[9], wv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    // the cell's own state values and its result stream through once:
    // non-temporal, leaving the caches to the flux stencils
    mv[v] = __builtin_nontemporal_load(at(a.wm + (long long)v * N, o));
    wv[v] = __builtin_nontemporal_load(at(a.win + (long long)v * N, o));
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) __builtin_nontemporal_store(tvd<R>(a.stage, mv[v], wv[v], h[v]), at(a.wout + v * N, o));
}

// One thread per interior column: 64-column tiles start at the first interior
// column (on a cache line in the row-aligned layout) over all NY x NZ padded
// rows; the lanes of columns 3..5 and nx..nx+2 also update the ghost column
// three to their left / right (nx >= 3; one lane per ghost column, since the
// two sets of lanes are disjoint or, for nx < 6, reach different columns).
template <class R>
__global__ void __launch_bounds__(256) aw_update(StageArgs<R> a) {
  cg::CellTile t;
  if (!cg::interior_tile(blockIdx.x, a.XE - 2 * H, a.NY, a.NZ, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = t.j0 + (threadIdx.x >> 6);
  const int k = t.k;
  if (i >= a.XE - H || j >= a.NY) return;
  update_cell<R>(a, i, j, k);
  if (i < 2 * H || i >= a.XE - 2 * H) {
    // ghost column: no right-hand side
    const unsigned o = (unsigned)(i < 2 * H ? i - H : i + H) * (unsigned)sizeof(R) + (unsigned)j * a.rowb +
                       (unsigned)k * a.planeb;
    R mv[9], wv[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      mv[v] = pl(a.wm, a.N, v, o);
      wv[v] = pl(a.win, a.N, v, o);
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) *at(a.wout + v * a.N, o) = tvd<R>(a.stage, mv[v], wv[v], R(0));
  }
}

// The three interior layers under the top: the corrections use 6-point
// formulas on the physical fluxes of the 7 cells around the cell
// (alternative_flux_FD.cpp:1257-1278), in every direction.
template <class R>
__global__ void __launch_bounds__(256) aw_update_top(StageArgs<R> a, int nx, int ny) {
  const int i = H + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = H + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = a.NZ - H - 3 + blockIdx.z;
  if (i >= H + nx || j >= H + ny) return;
  const long long N = a.N;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * a.S;
  const R jac = R(1.0 / (double)a.cjm[C_J * N + p]);
  const R cd = a.DT * a.rDH * jac;
  R h[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) h[v] = R(0);
  for (int d = 0; d < 3; ++d) {
    const long long s = d == 0 ? 1 : d == 1 ? (long long)a.NX : a.S;
    R hu[7][9];
#pragma unroll
    for (int n = 0; n < 7; ++n) {
      const long long q = p + (n - 3) * s;
      R u[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) u[v] = a.win[v * N + q];
      phys_flux<R>(u, a.cjm[(3 * d) * N + q], a.cjm[(3 * d + 1) * N + q], a.cjm[(3 * d + 2) * N + q],
                   a.cjm[C_MU * N + q], a.cjm[C_LAM * N + q], a.cjm[C_B * N + q], hu[n]);
    }
    const R* F = a.F + (long long)(9 * d) * N;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      const R fr = F[v * N + p] - R(1.0 / 24.0) * d2f<R>(hu[1][v], hu[2][v], hu[3][v], hu[4][v], hu[5][v], hu[6][v]) +
                   R(7.0 / 5760.0) * d4f<R>(hu[1][v], hu[2][v], hu[3][v], hu[4][v], hu[5][v], hu[6][v]);
      const R fl = F[v * N + p - s] - R(1.0 / 24.0) * d2f<R>(hu[0][v], hu[1][v], hu[2][v], hu[3][v], hu[4][v], hu[5][v]) +
                   R(7.0 / 5760.0) * d4f<R>(hu[0][v], hu[1][v], hu[2][v], hu[3][v], hu[4][v], hu[5][v]);
      h[v] += cd * (-(fr - fl));
    }
  }
  R mv[9], wv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    mv[v] = a.wm[v * N + p];
    wv[v] = a.win[v * N + p];
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) a.wout[v * N + p] = tvd<R>(a.stage, mv[v], wv[v], h[v]);
}

// Characteristic free surface on the top layer and cubic extrapolation into
// the three ghost layers (freeSurface.cpp:41-194). The stress tensor and the
// velocity are expressed in the frame B = [n s t] of the surface normal and
// the tangents x_xi, x_eta: frame components B^-1 S B^-T and B^-1 v, the rows
// of det(B) B^-1 being (s x t, t x n, n x s).
template <class R>
__global__ void __launch_bounds__(256) aw_char_surface(R* __restrict__ W, const R* __restrict__ cjm, int nx, int ny,
                                                       int NX, int NZ, long long N, long long S) {
  const int i = H + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = H + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H + nx || j >= H + ny) return;
  const int k = NZ - H - 1;
  const long long p = (long long)i + (long long)NX * j + (long long)k * S;
  const R mu = cjm[C_MU * N + p], lam = cjm[C_LAM * N + p], b = cjm[C_B * N + p];
  const R rho = R(1) / b;
  const R vs = sqrt(mu * b), vp = sqrt((lam + R(2) * mu) * b);
  const R J = cjm[C_J * N + p];
  R B[3][3];
  {
    R n0 = cjm[6 * N + p] / J, n1 = cjm[7 * N + p] / J, n2 = cjm[8 * N + p] / J;
    R nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    R s0 = cjm[(C_TXI + 0) * N + p], s1 = cjm[(C_TXI + 1) * N + p], s2 = cjm[(C_TXI + 2) * N + p];
    R sn = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
    R t0 = cjm[(C_TET + 0) * N + p], t1 = cjm[(C_TET + 1) * N + p], t2 = cjm[(C_TET + 2) * N + p];
    R tn = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    B[0][0] = n0 / nn; B[0][1] = n1 / nn; B[0][2] = n2 / nn;
    B[1][0] = s0 / sn; B[1][1] = s1 / sn; B[1][2] = s2 / sn;
    B[2][0] = t0 / tn; B[2][1] = t1 / tn; B[2][2] = t2 / tn;
  }
  R Rw[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int u = (r + 1) % 3, v = (r + 2) % 3;
    Rw[r][0] = B[u][1] * B[v][2] - B[u][2] * B[v][1];
    Rw[r][1] = B[u][2] * B[v][0] - B[u][0] * B[v][2];
    Rw[r][2] = B[u][0] * B[v][1] - B[u][1] * B[v][0];
  }
  const R det = B[0][0] * Rw[0][0] + B[0][1] * Rw[0][1] + B[0][2] * Rw[0][2];
  R w[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) w[q] = W[q * N + p];
  const R l2m = lam + R(2) * mu;
  R Sg[3][3];
  Sg[0][0] = lam * w[4] + lam * w[5] + w[3] * l2m;
  Sg[1][1] = lam * w[3] + lam * w[5] + w[4] * l2m;
  Sg[2][2] = lam * w[3] + lam * w[4] + w[5] * l2m;
  Sg[0][1] = Sg[1][0] = mu * w[8];
  Sg[0][2] = Sg[2][0] = mu * w[7];
  Sg[1][2] = Sg[2][1] = mu * w[6];
  R vel[3] = {w[0] * b, w[1] * b, w[2] * b};
  R Sf[3][3], vf[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    R RS[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) RS[q] = Rw[r][0] * Sg[0][q] + Rw[r][1] * Sg[1][q] + Rw[r][2] * Sg[2][q];
#pragma unroll
    for (int c = 0; c < 3; ++c) Sf[r][c] = (RS[0] * Rw[c][0] + RS[1] * Rw[c][1] + RS[2] * Rw[c][2]) / (det * det);
    vf[r] = (Rw[r][0] * vel[0] + Rw[r][1] * vel[1] + Rw[r][2] * vel[2]) / det;
  }
  // traction-free characteristic condition (freeSurface.cpp:145-153)
  Sf[1][1] -= Sf[0][0] * lam / l2m;
  Sf[2][2] -= Sf[0][0] * lam / l2m;
  vf[1] -= Sf[0][1] / (vs * rho);
  vf[2] -= Sf[0][2] / (vs * rho);
  vf[0] -= Sf[0][0] / (vp * rho);
  Sf[0][0] = R(0);
  Sf[0][1] = Sf[1][0] = R(0);
  Sf[0][2] = Sf[2][0] = R(0);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      R acc = R(0);
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc += Sf[u][c] * B[u][r] * B[c][q];
      Sg[r][q] = acc;
    }
    vel[r] = B[0][r] * vf[0] + B[1][r] * vf[1] + B[2][r] * vf[2];
  }
  const R cc = R(1) / (R(2) * mu * (R(3) * lam + R(2) * mu));
  R o[9];
  o[0] = vel[0] * rho;
  o[1] = vel[1] * rho;
  o[2] = vel[2] * rho;
  o[3] = (R(2) * lam * Sg[0][0] - lam * Sg[1][1] - lam * Sg[2][2] + R(2) * mu * Sg[0][0]) * cc;
  o[4] = -(lam * Sg[0][0] - R(2) * lam * Sg[1][1] + lam * Sg[2][2] - R(2) * mu * Sg[1][1]) * cc;
  o[5] = -(lam * Sg[0][0] + lam * Sg[1][1] - R(2) * lam * Sg[2][2] - R(2) * mu * Sg[2][2]) * cc;
  const R rm = R(1) / mu;
  o[6] = Sg[1][2] * rm;
  o[7] = Sg[0][2] * rm;
  o[8] = Sg[0][1] * rm;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    R* f = W + q * N;
    R c0 = o[q], c1 = f[p - S], c2 = f[p - 2 * S], c3 = f[p - 3 * S];
    f[p] = c0;
#pragma unroll
    for (int g = 1; g <= H; ++g) {
      const R e = R(4) * c0 - R(6) * c1 + R(4) * c2 - c3;
      f[p + g * S] = e;
      c3 = c2;
      c2 = c1;
      c1 = c0;
      c0 = e;
    }
  }
}

// Exponential sponge (exp_decay_absorb.cpp:20-106): 20 layers, alpha 0.015,
// factors of the sides a cell lies in applied one after another (x-low,
// x-high, y-low, y-high, z-low, z-high without free surface).
// The sponge frame is launched in two parts: whole 64 x 4 tiles of the planes
// k0 .. k0 + gridDim.z - 1 (slab = 1: the bottom planes, and the top ones
// without a free surface), and in the planes between them only the tiles that
// reach the x or y frame (slab = 0): tile columns [0, xa) and [xb, ntx) in
// every tile row, then tile rows [0, ya) and [yb, nty) of the other columns.
constexpr int SPONGE_L = 20;

template <class R>
__global__ void __launch_bounds__(256) aw_sponge(R* __restrict__ W, int nx, int ny, int nz, int NX, long long N,
                                                 long long S, int fs, int slab, int k0, int xa, int xb, int ya,
                                                 int yb) {
  const int ntx = (nx + 63) / 64, nty = (ny + 3) / 4;
  int tx, ty;
  if (slab) {
    tx = blockIdx.x % ntx;
    ty = blockIdx.x / ntx;
  } else {
    const int nxf = xa + (ntx - xb);  // frame tile columns
    const int b = blockIdx.x;
    if (b < nxf * nty) {
      const int c = b / nty;
      tx = c < xa ? c : xb + (c - xa);
      ty = b % nty;
    } else {
      const int inner = xb - xa;  // > 0 here
      const int b2 = b - nxf * nty;
      const int r = b2 / inner;
      tx = xa + b2 % inner;
      ty = r < ya ? r : yb + (r - ya);
    }
  }
  const int I = tx * 64 + (threadIdx.x & 63);
  const int Jj = ty * 4 + (threadIdx.x >> 6);
  const int K = k0 + blockIdx.z;
  if (I >= nx || Jj >= ny) return;
  const int L = SPONGE_L;
  int pv[6];
  int n = 0;
  if (I <= L) pv[n++] = L - I;
  if (I >= nx - L) pv[n++] = L - (nx - I);
  if (Jj <= L) pv[n++] = L - Jj;
  if (Jj >= ny - L) pv[n++] = L - (ny - Jj);
  if (K <= L) pv[n++] = L - K;
  if (!fs && K >= nz - L) pv[n++] = L - (nz - K);
  if (n == 0) return;
  const long long p = (long long)(I + H) + (long long)NX * (Jj + H) + (long long)(K + H) * S;
  const R al = R(0.015);
  R w[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) w[v] = W[v * N + p];
  for (int q = 0; q < n; ++q) {
    const R x = al * R(pv[q]);
    const R g = exp(-x * x);
#pragma unroll
    for (int v = 0; v < 9; ++v) w[v] *= g;
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) W[v * N + p] = w[v];
}

// Strain-rate source (addMoment.cpp:192-239): the interpolated moment rate
// (already scaled by mu / (J DH^3)) is turned into strain rates by the
// compliance of the cell and subtracted from W[3..8] once per step.
template <class R>
__global__ void aw_source(R* __restrict__ W, const R* __restrict__ cjm, long long N, int np,
                          const long long* __restrict__ cell, const R* __restrict__ rate, int srcIt, R tw, R DT) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const long long c = cell[q];
  R s[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const R m0 = rate[((long long)srcIt * np + q) * 6 + m];
    const R m1 = rate[((long long)(srcIt + 1) * np + q) * 6 + m];
    s[m] = (m1 - m0) * tw + m0;
  }
  const R mu = cjm[C_MU * N + c], lam = cjm[C_LAM * N + c];
  const R den = R(2) * (mu * mu) + R(3) * lam * mu;
  R e[6];
  e[0] = (s[0] * (lam + mu)) / den - (lam * s[1]) / (R(2) * den) - (lam * s[2]) / (R(2) * den);
  e[1] = (s[1] * (lam + mu)) / den - (lam * s[0]) / (R(2) * den) - (lam * s[2]) / (R(2) * den);
  e[2] = (s[2] * (lam + mu)) / den - (lam * s[0]) / (R(2) * den) - (lam * s[1]) / (R(2) * den);
  e[3] = s[5] / mu;
  e[4] = s[4] / mu;
  e[5] = s[3] / mu;
#pragma unroll
  for (int m = 0; m < 6; ++m) W[(3 + m) * N + c] -= e[m] * DT;
}

// Surface observers of the SCFDM build: PGV (PGV.cpp:62-175) and the surface
// displacement (solve_displacement.cpp:25-45) from v = (rho v) / rho.
template <class R>
__global__ void __launch_bounds__(256) aw_surface_obs(const R* __restrict__ W, const R* __restrict__ cjm, float* pgv,
                                                      R* dis, int nx, int ny, int NX, int NZ, long long N, long long S,
                                                      R DT) {
  const int I = blockIdx.x * 64 + (threadIdx.x & 63);
  const int Jj = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (I >= nx || Jj >= ny) return;
  const long long p = (long long)(I + H) + (long long)NX * (Jj + H) + (long long)(NZ - H - 1) * S;
  const R b = cjm[C_B * N + p];
  const R vx = W[p] * b, vy = W[N + p] * b, vz = W[2 * N + p] * b;
  const long long q = (long long)I + (long long)nx * Jj;
  const long long qd = (long long)(I + H) + (long long)NX * (Jj + H);
  dis[qd] += vx * DT;
  dis[S + qd] += vy * DT;
  dis[2 * S + qd] += vz * DT;
  const float fx = (float)vx, fy = (float)vy, fz = (float)vz;
  const float vh = sqrtf(fx * fx + fy * fy), va = sqrtf(fx * fx + fy * fy + fz * fz);
  if (pgv[2 * q] < vh) pgv[2 * q] = vh;
  if (pgv[2 * q + 1] < va) pgv[2 * q + 1] = va;
}

// Station record in physical variables (station.cpp:303-380, SCFDM branch):
// (vx, vy, vz, sxx, syy, szz, sxy, sxz, syz).
template <class R>
__global__ void aw_record(const R* __restrict__ W, const R* __restrict__ cjm, long long N, int ns,
                          const long long* __restrict__ cell, float* rec, int nt, int it) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const long long p = cell[s];
  R u[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) u[v] = W[v * N + p];
  const R mu = cjm[C_MU * N + p], lam = cjm[C_LAM * N + p], b = cjm[C_B * N + p];
  float* o = rec + ((long long)s * nt + it) * 9;
  o[0] = (float)(u[0] * b);
  o[1] = (float)(u[1] * b);
  o[2] = (float)(u[2] * b);
  o[3] = (float)(lam * u[4] + lam * u[5] + u[3] * (lam + R(2) * mu));
  o[4] = (float)(lam * u[3] + lam * u[5] + u[4] * (lam + R(2) * mu));
  o[5] = (float)(lam * u[3] + lam * u[4] + u[5] * (lam + R(2) * mu));
  o[6] = (float)(mu * u[8]);
  o[7] = (float)(mu * u[7]);
  o[8] = (float)(mu * u[6]);
}

// --------------------------------------------------------------- launchers

template <class R>
void launch_faces(const StageArgs<R>& a, hipStream_t st) {
  const long long nb = cg::interior_blocks(a.XE - 2 * H, a.NY - 5, a.NZ - 5, a.rtx, a.rty);
  aw_faces<R><<<dim3((unsigned)nb), dim3(256), 0, st>>>(a);
}

template <class R>
void launch_update(const StageArgs<R>& a, int nx, int ny, hipStream_t st, hipStream_t st_top) {
  // the top three layers (their own kernel, launched first) and the rest
  // write disjoint cells and read only the stage input and the face fluxes
  const dim3 block(256);
  const dim3 gtop((nx + 63) / 64, (ny + 3) / 4, 3);
  aw_update_top<R><<<gtop, block, 0, st_top>>>(a, nx, ny);
  const long long nb = cg::interior_blocks(a.XE - 2 * H, a.NY, a.NZ, a.rtx, a.rty);
  aw_update<R><<<dim3((unsigned)nb), block, 0, st>>>(a);
}

template <class R>
void launch_char_surface(R* W, const R* cjm, int nx, int ny, int NX, int NZ, long long N, long long S, hipStream_t st) {
  aw_char_surface<R><<<dim3((nx + 63) / 64, (ny + 3) / 4, 1), dim3(256), 0, st>>>(W, cjm, nx, ny, NX, NZ, N, S);
}

template <class R>
void launch_sponge(R* W, int nx, int ny, int nz, int NX, long long N, long long S, int fs, hipStream_t st) {
  const int L = SPONGE_L;
  const int ntx = (nx + 63) / 64, nty = (ny + 3) / 4;
  // planes damped in z: K <= L (bottom), and K >= nz - L without a free surface
  const int kb = std::min(L, nz - 1);
  const int kt = fs ? nz : std::max(nz - L, kb + 1);  // first plane of the top slab
  const dim3 block(256);
  aw_sponge<R><<<dim3((unsigned)(ntx * nty), 1, (unsigned)(kb + 1)), block, 0, st>>>(W, nx, ny, nz, NX, N, S, fs, 1, 0,
                                                                                   0, 0, 0, 0);
  if (kt < nz)
    aw_sponge<R><<<dim3((unsigned)(ntx * nty), 1, (unsigned)(nz - kt)), block, 0, st>>>(W, nx, ny, nz, NX, N, S, fs, 1,
                                                                                      kt, 0, 0, 0, 0);
  // planes kb + 1 .. kt - 1: tiles reaching I <= L or I >= nx - L, J <= L or J >= ny - L
  const int nmid = kt - (kb + 1);
  if (nmid <= 0) return;
  int xa = std::min(ntx, L / 64 + 1), xb = std::max(xa, (nx - L) / 64);
  int ya = std::min(nty, L / 4 + 1), yb = std::max(ya, (ny - L) / 4);
  const int nblk = (xa + (ntx - xb)) * nty + (xb - xa) * (ya + (nty - yb));
  aw_sponge<R><<<dim3((unsigned)nblk, 1, (unsigned)nmid), block, 0, st>>>(W, nx, ny, nz, NX, N, S, fs, 0, kb + 1, xa,
                                                                        xb, ya, yb);
}

template <class R>
void launch_source(R* W, const R* cjm, long long N, int np, const long long* cell, const R* rate, int srcIt, R tw, R DT,
                   hipStream_t st) {
  if (np <= 0) return;
  aw_source<R><<<dim3((np + 127) / 128), dim3(128), 0, st>>>(W, cjm, N, np, cell, rate, srcIt, tw, DT);
}

template <class R>
void launch_surface_obs(const R* W, const R* cjm, float* pgv, R* dis, int nx, int ny, int NX, int NZ, long long N,
                        long long S, R DT, hipStream_t st) {
  aw_surface_obs<R><<<dim3((nx + 63) / 64, (ny + 3) / 4, 1), dim3(256), 0, st>>>(W, cjm, pgv, dis, nx, ny, NX, NZ, N,
                                                                                  S, DT);
}

template <class R>
void launch_record(const R* W, const R* cjm, long long N, int ns, const long long* cell, float* rec, int nt, int it,
                   hipStream_t st) {
  if (ns <= 0 || it >= nt) return;
  aw_record<R><<<dim3((ns + 127) / 128), dim3(128), 0, st>>>(W, cjm, N, ns, cell, rec, nt, it);
}

}  // namespace aw
