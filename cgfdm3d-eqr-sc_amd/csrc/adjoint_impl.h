// Discrete adjoint of the CGFDM RK4 stage for gfx950.
//
// The forward stage is h = L X (velocity/stress right-hand side, free-surface
// rows on the top three layers, PML auxiliaries), scaled by DT. For an adjoint
// g of h, the adjoint stage computes l = L^T g, the PML auxiliary adjoint, and
// the medium gradient <g, dL/d(mu, lambda, 1/rho) X_s> at the stage state X_s.
// With the RK4 weights (b1..b4, a2, a3) the transposed step is
//   g3 = b4 mu,        l3 = L^T g3,   out = mu + l3,
//   g2 = b3 mu + l3,   l2 = L^T g2,   out += l2,
//   g1 = b2 mu + a3 l2, l1 = L^T g1,  out += l1,
//   g0 = b1 mu + a2 l1, l0 = L^T g0,  out += l0,
// where mu is the adjoint of the state after the step and out that of the
// state at its start (after the source).
//
// L^T is applied without atomics in the bulk: pass A turns g at every cell p
// into nine products per direction, Z_d(p) = beta_d assemble^T(g DT) (+ the
// PML auxiliary terms of the faces through p), which the forward difference
// at p spreads over its stencil; pass B gathers them, l(j) = sum_d sum_m
// FB_d a_m Z_d(j - FB_d m e_d) / DH, and applies the RK combination. The top
// three layers under a free surface replace the bulk rows by the traction-
// image rows; their transpose touches a 7-cell neighbourhood and is scattered
// with atomics into a separate accumulator that pass B adds and clears.
#pragma once
#include "adjoint.h"
#include "launch_impl.h"

namespace cg {

// RK weight of h_s in the final state, and of h_{s-1} in the stage-s state
constexpr double rk_b(int s) { return s == 0 ? RK_B1 : s == 1 ? RK_B2 : s == 2 ? RK_B3 : RK_B4; }
constexpr double rk_a(int s) { return s == 1 ? RK_A2 : s == 2 ? RK_A3 : 1.0; }

template <class R>
__device__ __forceinline__ void assemble_T(const R* g, R m1, R m2, R m3, R mu, R lam, R b, R* lD) {
  const R gd = lam * (g[3] + g[4] + g[5]);
  lD[0] = m1 * gd + R(2) * mu * m1 * g[3] + mu * (m2 * g[6] + m3 * g[7]);
  lD[1] = m2 * gd + R(2) * mu * m2 * g[4] + mu * (m1 * g[6] + m3 * g[8]);
  lD[2] = m3 * gd + R(2) * mu * m3 * g[5] + mu * (m1 * g[7] + m2 * g[8]);
  lD[3] = b * m1 * g[0];
  lD[4] = b * m2 * g[1];
  lD[5] = b * m3 * g[2];
  lD[6] = b * (m2 * g[0] + m1 * g[1]);
  lD[7] = b * (m3 * g[0] + m1 * g[2]);
  lD[8] = b * (m3 * g[1] + m2 * g[2]);
}

template <class R>
__device__ __forceinline__ void assemble_grad(const R* g, const R* D, R m1, R m2, R m3, double& gmu, double& glam,
                                              double& gb) {
  gb += (double)(g[0] * (m1 * D[3] + m2 * D[6] + m3 * D[7]) + g[1] * (m1 * D[6] + m2 * D[4] + m3 * D[8]) +
                 g[2] * (m1 * D[7] + m2 * D[8] + m3 * D[5]));
  const R div = m1 * D[0] + m2 * D[1] + m3 * D[2];
  glam += (double)(div * (g[3] + g[4] + g[5]));
  gmu += (double)(R(2) * (m1 * D[0] * g[3] + m2 * D[1] * g[4] + m3 * D[2] * g[5]) + (m2 * D[0] + m1 * D[1]) * g[6] +
                  (m3 * D[0] + m1 * D[2]) * g[7] + (m3 * D[1] + m2 * D[2]) * g[8]);
}

template <class R>
__device__ __forceinline__ void stress_T(const R* g, R m1, R m2, R m3, R mu, R lam, R* lV) {
  const R gd = lam * (g[0] + g[1] + g[2]);
  lV[0] = m1 * gd + R(2) * mu * m1 * g[0] + mu * (m2 * g[3] + m3 * g[4]);
  lV[1] = m2 * gd + R(2) * mu * m2 * g[1] + mu * (m1 * g[3] + m3 * g[5]);
  lV[2] = m3 * gd + R(2) * mu * m3 * g[2] + mu * (m1 * g[4] + m2 * g[5]);
}

template <class R>
__device__ __forceinline__ void stress_grad(const R* g, const R* V, R m1, R m2, R m3, double& gmu, double& glam) {
  const R div = m1 * V[0] + m2 * V[1] + m3 * V[2];
  glam += (double)(div * (g[0] + g[1] + g[2]));
  gmu += (double)(R(2) * (m1 * V[0] * g[0] + m2 * V[1] * g[1] + m3 * V[2] * g[2]) + (m2 * V[0] + m1 * V[1]) * g[3] +
                  (m3 * V[0] + m1 * V[2]) * g[4] + (m3 * V[1] + m2 * V[2]) * g[5]);
}

// adjoint of h at cell p for the current stage: b4 mu in stage 3, G otherwise.
// The stage is a compile-time parameter of every adjoint kernel, so no kernel
// chooses base pointers by stage at run time.
template <class R, int STAGE>
__device__ __forceinline__ R g_of(const AdjArgs<R>& a, long long q) {
  if constexpr (STAGE == 3) return R(RK_B4) * a.mu[q];
  else return a.G[q];
}

// ---------------------------------------------------------------------------
// pass A: per interior cell, products Z, PML auxiliary adjoint, bulk gradient
// ---------------------------------------------------------------------------
// Pass A's work at interior cell (i, j, k) (padded index p) with the stage's
// adjoint g of the cell in gv: products into Zout, PML auxiliary adjoint (the
// next stage's g into fa[].gn when GN, else in place into fa[].g), gradient.
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE, bool GN>
__device__ __forceinline__ void adj_a_cell(const AdjArgs<R>& a, const TM* __restrict__ coef, int i, int j, int k,
                                           long long p, const R (&gv)[9], R* __restrict__ Zout) {
  const long long N = a.N, S = a.S;
  auto cf = [&](int c) -> R { return R(Store<TM>::ld(coef + (long long)c * N + p)); };
  const R mu = cf(M_MU), lam = cf(M_LAM), bu = cf(M_B);
  const PmlArgs<R>& P = a.pml;
  const bool top = a.fs && k >= H + a.nz - 3;
  R gd[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) gd[v] = gv[v] * a.DT;
  double gmu = 0, glam = 0, gb = 0;
  const int FBv[3] = {FB1, FB2, FB3};
  const long long sv[3] = {1, (long long)a.NX, S};
  const int pos[3] = {i, j, k};
  const int npos[3] = {a.nx, a.ny, a.nz};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const R m1 = cf(3 * ax), m2 = cf(3 * ax + 1), m3 = cf(3 * ax + 2);
    R Zd[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) Zd[v] = R(0);
    // derivative of the stage state along ax (forward difference of this stage)
    R D[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      const R* f = a.X + (long long)v * N + p;
      const long long s = sv[ax];
      if (FBv[ax] > 0) D[v] = mac<1>(f, s) * a.rDH;
      else D[v] = mac<-1>(f, s) * a.rDH;
    }
    if (!top) {
      const R bet = P.beta[ax][pos[ax]];
      R lD[9];
      assemble_T(gd, m1, m2, m3, mu, lam, bu, lD);
#pragma unroll
      for (int v = 0; v < 9; ++v) Zd[v] = lD[v] * bet;
      R Db[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) Db[v] = D[v] * bet;
      assemble_grad(gd, Db, m1, m2, m3, gmu, glam, gb);
    }
    if (P.npml > 0) {
      const int q0 = pos[ax] - H;
      const bool lo = q0 < P.npml;
      const bool hi = q0 >= npos[ax] - P.npml && (ax < 2 || P.has_z2);
      if (lo || hi) {
        const R d = P.d[ax][pos[ax]], ad = d + P.alpha[ax][pos[ax]], be = P.beta[ax][pos[ax]];
        for (int side = 0; side < 2; ++side) {
          if (side == 0 ? !lo : !hi) continue;
          const int f = ax + 3 * side;
          long long q;
          if (ax == 0) {
            const long long row = (long long)(j - H) + (long long)a.ny * (k - H);
            q = (long long)(side ? q0 - (a.nx - P.npml) : q0) + (long long)P.npml * row;
          } else if (ax == 1) {
            q = (long long)(i - H) + (long long)a.nx * ((long long)(side ? q0 - (a.ny - P.npml) : q0) + (long long)P.npml * (k - H));
          } else {
            q = (long long)(i - H) + (long long)a.nx * ((j - H) + (long long)a.ny * (side ? q0 - (a.nz - P.npml) : q0));
          }
          const FaceAdj<R>& F = a.fa[f];
          R gQ[9];
          // every load of the face adjoints before the first store (the
          // buffers are not known not to alias)
          R fmu[9], fg[9], fo[9];
#pragma unroll
          for (int v = 0; v < 9; ++v) {
            const long long e = (long long)v * F.cnt + q;
            if constexpr (STAGE == 3 || STAGE > 0) fmu[v] = F.mu[e];
            if constexpr (STAGE != 3) {
              fg[v] = F.g[e];
              fo[v] = F.out[e];
            }
          }
#pragma unroll
          for (int v = 0; v < 9; ++v) {
            const long long e = (long long)v * F.cnt + q;
            R gA, base;
            if constexpr (STAGE == 3) {
              gA = R(RK_B4) * fmu[v];
              base = fmu[v];
            } else {
              gA = fg[v];
              base = fo[v];
            }
            const R lA = -ad * a.DT * gA - be * a.DT * gv[v];
            F.out[e] = base + lA;
            if constexpr (STAGE > 0) (GN ? F.gn : F.g)[e] = R(rk_b(STAGE - 1)) * fmu[v] + R(rk_a(STAGE)) * lA;
            gQ[v] = gA * a.DT;
          }
          R lD[9];
          assemble_T(gQ, m1, m2, m3, mu, lam, bu, lD);
#pragma unroll
          for (int v = 0; v < 9; ++v) Zd[v] += lD[v] * d;
          R Dd[9];
#pragma unroll
          for (int v = 0; v < 9; ++v) Dd[v] = D[v] * d;
          assemble_grad(gQ, Dd, m1, m2, m3, gmu, glam, gb);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) Zout[(long long)(9 * ax + v) * N + p] = Zd[v];
  }
  a.gmu[p] += gmu;
  a.glam[p] += glam;
  a.gb[p] += gb;
}

template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_adj_a(AdjArgs<R> a, const TM* __restrict__ coef) {
  CellTile t;
  if (!interior_tile(blockIdx.x, a.nx, a.ny, a.nz, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = H + t.j0 + (threadIdx.x >> 6);
  const int k = H + t.k;
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * a.S;
  R gv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) gv[v] = g_of<R, STAGE>(a, (long long)v * a.N + p);
  adj_a_cell<R, TM, FB1, FB2, FB3, STAGE, false>(a, coef, i, j, k, p, gv, a.Z);
}

// ---------------------------------------------------------------------------
// pass B: gather of the products, surface accumulator, RK combination
// ---------------------------------------------------------------------------
// Pass B's work at interior cell (i, j, k): gather of the products Zin,
// surface accumulator, RK combination into out; the next stage's g of the
// cell into gn (STAGE > 0).
template <class R, int FB1, int FB2, int FB3, int STAGE>
__device__ __forceinline__ void adj_b_cell(const AdjArgs<R>& a, const R* __restrict__ Zin, int k, long long p,
                                           R (&gn)[9]) {
  const long long N = a.N, S = a.S;
  const R cw[5] = {R(AF_1), R(AF0), R(AF1), R(AF2), R(AF3)};
  const int FBv[3] = {FB1, FB2, FB3};
  const long long sv[3] = {1, (long long)a.NX, S};
  // the surface pass scatters into layers ktop - 5 .. ktop only (7-point image
  // window of the three surface rows): the accumulator is read and cleared there
  const bool in_ls = a.fs && k >= H + a.nz - 6;
  // gathers of all nine fields first, then the read-modify-writes: every load
  // comes before the first store (the buffers are not known not to alias)
  R l[9], ls[9], mv[9], ov[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    l[v] = R(0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const R* z = Zin + (long long)(9 * ax + v) * N + p;
      const long long d = FBv[ax] * sv[ax];
      R acc = R(0);
#pragma unroll
      for (int m = 0; m < 5; ++m) acc += cw[m] * z[-(m - 1) * d];
      l[v] += R(FBv[ax]) * acc * a.rDH;
    }
    const long long q = (long long)v * N + p;
    ls[v] = in_ls ? a.Ls[q] : R(0);
    if constexpr (STAGE == 3 || STAGE > 0) mv[v] = a.mu[q];
    if constexpr (STAGE != 3) ov[v] = a.out[q];
  }
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const long long q = (long long)v * N + p;
    const R lv = l[v] + ls[v];
    if (in_ls) a.Ls[q] = R(0);
    if constexpr (STAGE == 3) a.out[q] = mv[v] + lv;
    else a.out[q] = ov[v] + lv;
    if constexpr (STAGE > 0) gn[v] = R(rk_b(STAGE - 1)) * mv[v] + R(rk_a(STAGE)) * lv;
  }
}

template <class R, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_adj_b(AdjArgs<R> a) {
  CellTile t;
  if (!interior_tile(blockIdx.x, a.nx, a.ny, a.nz, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = H + t.j0 + (threadIdx.x >> 6);
  const int k = H + t.k;
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * a.S;
  R gn[9];
  adj_b_cell<R, FB1, FB2, FB3, STAGE>(a, a.Z, k, p, gn);
  if constexpr (STAGE > 0) {
#pragma unroll
    for (int v = 0; v < 9; ++v) a.G[(long long)v * a.N + p] = gn[v];
  }
}

// Pass B of stage STAGE (products in Z, F/B pattern FB) and pass A of stage
// STAGE - 1 (products into Zn, the opposite F/B pattern) in one kernel: the
// next stage's g of a cell goes from B to A in registers. It is stored only in
// the three surface layers, where the surface pass of stage STAGE - 1 reads
// it; that pass runs after this kernel and reads the auxiliary adjoint g of
// stage STAGE - 1, so pass A writes the following one into fa[].gn.
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_adj_ba(AdjArgs<R> a, const TM* __restrict__ coef) {
  static_assert(STAGE >= 1 && STAGE <= 3, "pass A follows pass B of stages 3..1");
  CellTile t;
  if (!interior_tile(blockIdx.x, a.nx, a.ny, a.nz, a.rtx, a.rty, t)) return;
  const int i = H + t.i0 + (threadIdx.x & 63);
  const int j = H + t.j0 + (threadIdx.x >> 6);
  const int k = H + t.k;
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long p = (long long)i + (long long)a.NX * j + (long long)k * a.S;
  R gn[9];
  adj_b_cell<R, FB1, FB2, FB3, STAGE>(a, a.Z, k, p, gn);
  if (a.fs && k >= H + a.nz - 3) {
#pragma unroll
    for (int v = 0; v < 9; ++v) a.G[(long long)v * a.N + p] = gn[v];
  }
  adj_a_cell<R, TM, -FB1, -FB2, -FB3, STAGE - 1, true>(a, coef, i, j, k, p, gn, a.Zn);
}

template <class R>
__device__ __forceinline__ void atomic_add(R* p, R v) {
  atomicAdd(p, v);
}

// scatter of the transposed 5-point difference of one field
template <class R>
__device__ __forceinline__ void scatter_L(R* lf, long long p, long long s, int fb, R g, R rDH) {
  const long long d = fb * s;
  const R c = R(fb) * g * rDH;
  atomic_add(lf + p - d, c * R(AF_1));
  atomic_add(lf + p, c * R(AF0));
  atomic_add(lf + p + d, c * R(AF1));
  atomic_add(lf + p + 2 * d, c * R(AF2));
  atomic_add(lf + p + 3 * d, c * R(AF3));
}

// d(-DZ^-1 DX), d(-DZ^-1 DY) with respect to mu and lambda of the surface cell,
// contracted with gRx, gRy (dR = -DZ^-1 (dDZ R + dD)), FP64; chi = 2 lambda + mu
// as in the forward matrices.
__device__ __forceinline__ void rdz_grad_dev(const double* xm, const double* em, const double* zm, double mu, double lam,
                                    const double* Rx, const double* Ry, const double* gRx, const double* gRy,
                                    double& gmu, double& glam) {
  const double zx = zm[0], zy = zm[1], zz = zm[2];
  auto DZof = [&](double chi, double lm, double m_, double* DZ) {
    DZ[0] = chi * zx * zx + m_ * (zy * zy + zz * zz);
    DZ[1] = lm * zx * zy;
    DZ[2] = lm * zx * zz;
    DZ[3] = DZ[1];
    DZ[4] = chi * zy * zy + m_ * (zx * zx + zz * zz);
    DZ[5] = lm * zy * zz;
    DZ[6] = DZ[2];
    DZ[7] = DZ[5];
    DZ[8] = chi * zz * zz + m_ * (zx * zx + zy * zy);
  };
  auto Dof = [&](double chi, double l_, double m_, const double* ax, double* D) {
    D[0] = chi * zx * ax[0] + m_ * (zy * ax[1] + zz * ax[2]);
    D[1] = l_ * zx * ax[1] + m_ * zy * ax[0];
    D[2] = l_ * zx * ax[2] + m_ * zz * ax[0];
    D[3] = l_ * zy * ax[0] + m_ * zx * ax[1];
    D[4] = chi * zy * ax[1] + m_ * (zx * ax[0] + zz * ax[2]);
    D[5] = l_ * zy * ax[2] + m_ * zz * ax[1];
    D[6] = l_ * zz * ax[0] + m_ * zx * ax[2];
    D[7] = l_ * zz * ax[1] + m_ * zy * ax[2];
    D[8] = chi * zz * ax[2] + m_ * (zx * ax[0] + zy * ax[1]);
  };
  double A[9], Zi[9];
  DZof(2 * lam + mu, lam + mu, mu, A);
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
  Zi[0] = (A[4] * A[8] - A[5] * A[7]) / det;
  Zi[1] = (A[2] * A[7] - A[1] * A[8]) / det;
  Zi[2] = (A[1] * A[5] - A[2] * A[4]) / det;
  Zi[3] = (A[5] * A[6] - A[3] * A[8]) / det;
  Zi[4] = (A[0] * A[8] - A[2] * A[6]) / det;
  Zi[5] = (A[2] * A[3] - A[0] * A[5]) / det;
  Zi[6] = (A[3] * A[7] - A[4] * A[6]) / det;
  Zi[7] = (A[1] * A[6] - A[0] * A[7]) / det;
  Zi[8] = (A[0] * A[4] - A[1] * A[3]) / det;
  double Gx[9], Gy[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      Gx[3 * r + c] = Zi[r] * gRx[c] + Zi[3 + r] * gRx[3 + c] + Zi[6 + r] * gRx[6 + c];
      Gy[3 * r + c] = Zi[r] * gRy[c] + Zi[3 + r] * gRy[3 + c] + Zi[6 + r] * gRy[6 + c];
    }
  for (int which = 0; which < 2; ++which) {
    const double dmu = which == 0 ? 1.0 : 0.0, dlam = which == 0 ? 0.0 : 1.0;
    double dZ[9], dX[9], dY[9];
    DZof(2 * dlam + dmu, dlam + dmu, dmu, dZ);
    Dof(2 * dlam + dmu, dlam, dmu, xm, dX);
    Dof(2 * dlam + dmu, dlam, dmu, em, dY);
    double acc = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double tx = dX[3 * r + c], ty = dY[3 * r + c];
        for (int q = 0; q < 3; ++q) {
          tx += dZ[3 * r + q] * Rx[3 * q + c];
          ty += dZ[3 * r + q] * Ry[3 * q + c];
        }
        acc -= Gx[3 * r + c] * tx + Gy[3 * r + c] * ty;
      }
    if (which == 0) gmu += acc;
    else glam += acc;
  }
}

// ---------------------------------------------------------------------------
// surface pass: transposed traction-image rows of the top three layers and the
// surface term of the x/y PML auxiliaries, scattered into Ls; their gradient
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int STAGE>
__global__ __launch_bounds__(256) void cg_adj_s(AdjArgs<R> a, const TM* __restrict__ coef, const R* __restrict__ jac,
                                                const R* __restrict__ rdzdx, const R* __restrict__ rdzdy) {
  const int i = H + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = H + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ktop = H + a.nz - 1;
  const int k = ktop - blockIdx.z;
  if (i >= H + a.nx || j >= H + a.ny) return;
  const long long N = a.N, S = a.S, NXl = a.NX;
  const long long p = (long long)i + NXl * j + (long long)k * S;
  auto cf = [&](int c, long long q) -> R { return R(Store<TM>::ld(coef + (long long)c * N + q)); };
  const R mu = cf(M_MU, p), lam = cf(M_LAM, p), bu = cf(M_B, p);
  const PmlArgs<R>& P = a.pml;
  const R bx = P.beta[0][i], by = P.beta[1][j];
  const int ks = ktop - k + H;
  const R* X = a.X;
  R* Ls = a.Ls;
  R gv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) gv[v] = g_of<R, STAGE>(a, (long long)v * N + p);
  double gmu = 0, glam = 0, gb = 0;
  const R jinv = R(1) / jac[p];
  const R cw[5] = {R(AF_1), R(AF0), R(AF1), R(AF2), R(AF3)};

  // velocity rows: hv = b jinv (d_xi(J xi.T) bx + d_eta(J eta.T) by + d_zeta(J zeta.T)) DT
  auto jt_T = [&](long long q, int c0, const R* lj) {
    const R m1 = cf(c0, q), m2 = cf(c0 + 1, q), m3 = cf(c0 + 2, q), jq = jac[q];
    atomic_add(Ls + 3 * N + q, m1 * jq * lj[0]);
    atomic_add(Ls + 6 * N + q, m2 * jq * lj[0] + m1 * jq * lj[1]);
    atomic_add(Ls + 7 * N + q, m3 * jq * lj[0] + m1 * jq * lj[2]);
    atomic_add(Ls + 4 * N + q, m2 * jq * lj[1]);
    atomic_add(Ls + 8 * N + q, m3 * jq * lj[1] + m2 * jq * lj[2]);
    atomic_add(Ls + 5 * N + q, m3 * jq * lj[2]);
  };
  auto jt = [&](long long q, int c0, R* o) {
    const R m1 = cf(c0, q), m2 = cf(c0 + 1, q), m3 = cf(c0 + 2, q), jq = jac[q];
    o[0] = (m1 * X[3 * N + q] + m2 * X[6 * N + q] + m3 * X[7 * N + q]) * jq;
    o[1] = (m1 * X[6 * N + q] + m2 * X[4 * N + q] + m3 * X[8 * N + q]) * jq;
    o[2] = (m1 * X[7 * N + q] + m2 * X[8 * N + q] + m3 * X[5 * N + q]) * jq;
  };
  {
    R gx[3], gy[3], gz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const R gvd = gv[c] * a.DT * bu * jinv;
      gx[c] = gvd * bx * R(FB1) * a.rDH;
      gy[c] = gvd * by * R(FB2) * a.rDH;
      gz[c] = gvd * R(FB3) * a.rDH;
    }
    R hx[3] = {0, 0, 0}, hy[3] = {0, 0, 0}, hz[3] = {0, 0, 0};
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const int off1 = (t - 1) * FB1, off2 = (t - 1) * FB2;
      R lj[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) lj[c] = cw[t] * gx[c];
      jt_T(p + off1, 0, lj);
#pragma unroll
      for (int c = 0; c < 3; ++c) lj[c] = cw[t] * gy[c];
      jt_T(p + (long long)off2 * NXl, 3, lj);
      const int l = 3 + (t - 1) * FB3;
      if (l < ks) {
#pragma unroll
        for (int c = 0; c < 3; ++c) lj[c] = cw[t] * gz[c];
        jt_T(p + (long long)(l - 3) * S, 6, lj);
      } else if (l > ks) {
#pragma unroll
        for (int c = 0; c < 3; ++c) lj[c] = -cw[t] * gz[c];
        jt_T(p + (long long)(2 * ks - l - 3) * S, 6, lj);
      }
      // forward values for the buoyancy gradient
      R o[3];
      jt(p + off1, 0, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) hx[c] += cw[t] * o[c];
      jt(p + (long long)off2 * NXl, 3, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) hy[c] += cw[t] * o[c];
      if (l < ks) {
        jt(p + (long long)(l - 3) * S, 6, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) hz[c] += cw[t] * o[c];
      } else if (l > ks) {
        jt(p + (long long)(2 * ks - l - 3) * S, 6, o);
#pragma unroll
        for (int c = 0; c < 3; ++c) hz[c] -= cw[t] * o[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const R hv_b = jinv * ((R(FB1) * hx[c] * a.rDH) * bx + (R(FB2) * hy[c] * a.rDH) * by + (R(FB3) * hz[c] * a.rDH));
      gb += (double)(gv[c] * a.DT * hv_b);
    }
  }
  // stress rows: s(Vxi) + s(Vet) + s(Vzt), DT
  R gs[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) gs[v] = gv[3 + v] * a.DT;
  R lVxi[3], lVet[3], lVzt[3];
  stress_T(gs, cf(M_XIX, p), cf(M_XIY, p), cf(M_XIZ, p), mu, lam, lVxi);
  stress_T(gs, cf(M_ETX, p), cf(M_ETY, p), cf(M_ETZ, p), mu, lam, lVet);
  stress_T(gs, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, lVzt);
  R Vxi[3], Vet[3], Vzt[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    Vxi[v] = mac<FB1>(X + (long long)v * N + p, 1) * a.rDH * bx;
    Vet[v] = mac<FB2>(X + (long long)v * N + p, NXl) * a.rDH * by;
  }
  const long long sp = (long long)i + NXl * j;
  const long long SS = (long long)a.NX * a.NY;
  if (k == ktop) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        lVxi[c] += rdzdx[(3 * r + c) * SS + sp] * lVzt[r];
        lVet[c] += rdzdy[(3 * r + c) * SS + sp] * lVzt[r];
      }
#pragma unroll
    for (int r = 0; r < 3; ++r)
      Vzt[r] = (rdzdx[(3 * r) * SS + sp] * Vxi[0] + rdzdx[(3 * r + 1) * SS + sp] * Vxi[1] + rdzdx[(3 * r + 2) * SS + sp] * Vxi[2]) +
               (rdzdy[(3 * r) * SS + sp] * Vet[0] + rdzdy[(3 * r + 1) * SS + sp] * Vet[1] + rdzdy[(3 * r + 2) * SS + sp] * Vet[2]);
    double xm[3], em[3], zm[3], Rx[9], Ry[9], gRx[9], gRy[9];
    for (int c = 0; c < 3; ++c) {
      xm[c] = (double)cf(M_XIX + c, p);
      em[c] = (double)cf(M_ETX + c, p);
      zm[c] = (double)cf(M_ZTX + c, p);
    }
    for (int q = 0; q < 9; ++q) {
      Rx[q] = (double)rdzdx[q * SS + sp];
      Ry[q] = (double)rdzdy[q * SS + sp];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        gRx[3 * r + c] = (double)lVzt[r] * (double)Vxi[c];
        gRy[3 * r + c] = (double)lVzt[r] * (double)Vet[c];
      }
    rdz_grad_dev(xm, em, zm, (double)mu, (double)lam, Rx, Ry, gRx, gRy, gmu, glam);
  } else if (k == ktop - 1) {
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const R c = R(FB3) * lVzt[v] * a.rDH;
      atomic_add(Ls + (long long)v * N + p + FB3 * S, c);
      atomic_add(Ls + (long long)v * N + p, -c);
      Vzt[v] = R(FB3) * (X[(long long)v * N + p + FB3 * S] - X[(long long)v * N + p]) * a.rDH;
    }
  } else {
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const R c = R(FB3) * lVzt[v] * a.rDH;
      atomic_add(Ls + (long long)v * N + p, c * R(CF1));
      atomic_add(Ls + (long long)v * N + p + FB3 * S, c * R(CF2));
      atomic_add(Ls + (long long)v * N + p + 2 * FB3 * S, c * R(CF3));
      Vzt[v] = R(FB3) * (R(CF1) * X[(long long)v * N + p] + R(CF2) * X[(long long)v * N + p + FB3 * S] +
                         R(CF3) * X[(long long)v * N + p + 2 * FB3 * S]) * a.rDH;
    }
  }
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    scatter_L(Ls + (long long)v * N, p, 1, FB1, lVxi[v] * bx, a.rDH);
    scatter_L(Ls + (long long)v * N, p, NXl, FB2, lVet[v] * by, a.rDH);
  }
  stress_grad(gs, Vxi, cf(M_XIX, p), cf(M_XIY, p), cf(M_XIZ, p), gmu, glam);
  stress_grad(gs, Vet, cf(M_ETX, p), cf(M_ETY, p), cf(M_ETZ, p), gmu, glam);
  stress_grad(gs, Vzt, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), gmu, glam);

  // surface term of the x/y PML auxiliaries on the top layer
  if (k == ktop && P.npml > 0) {
#pragma unroll
    for (int ax = 0; ax < 2; ++ax) {
      const int q0 = (ax == 0 ? i : j) - H;
      const int n = ax == 0 ? a.nx : a.ny;
      const bool lo = q0 < P.npml, hi = q0 >= n - P.npml;
      if (!lo && !hi) continue;
      const R* Rm = ax == 0 ? rdzdx : rdzdy;
      const R d = P.d[ax][q0 + H];
      const long long s = ax == 0 ? 1 : NXl;
      R V[3], Vz[3];
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        const R* f = X + (long long)v * N + p;
        V[v] = (ax == 0 ? (FB1 > 0 ? mac<1>(f, s) : mac<-1>(f, s)) : (FB2 > 0 ? mac<1>(f, s) : mac<-1>(f, s))) * a.rDH * d;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) Vz[r] = Rm[(3 * r) * SS + sp] * V[0] + Rm[(3 * r + 1) * SS + sp] * V[1] + Rm[(3 * r + 2) * SS + sp] * V[2];
      for (int side = 0; side < 2; ++side) {
        if (side == 0 ? !lo : !hi) continue;
        const int f = ax + 3 * side;
        long long q;
        if (ax == 0) {
          const long long row = (long long)(j - H) + (long long)a.ny * (k - H);
          q = (long long)(side ? q0 - (a.nx - P.npml) : q0) + (long long)P.npml * row;
        } else {
          q = (long long)(i - H) + (long long)a.nx * ((long long)(side ? q0 - (a.ny - P.npml) : q0) + (long long)P.npml * (k - H));
        }
        const FaceAdj<R>& F = a.fa[f];
        R ga[6];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
          const long long e = (long long)(3 + v) * F.cnt + q;
          if constexpr (STAGE == 3) ga[v] = R(RK_B4) * F.mu[e] * a.DT;
          else ga[v] = F.g[e] * a.DT;
        }
        R lVz[3], lV[3] = {0, 0, 0};
        stress_T(ga, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), mu, lam, lVz);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) lV[c] += Rm[(3 * r + c) * SS + sp] * lVz[r];
#pragma unroll
        for (int v = 0; v < 3; ++v) scatter_L(Ls + (long long)v * N, p, s, ax == 0 ? FB1 : FB2, lV[v] * d, a.rDH);
        stress_grad(ga, Vz, cf(M_ZTX, p), cf(M_ZTY, p), cf(M_ZTZ, p), gmu, glam);
        double xm[3], em[3], zm[3], Rx[9], Ry[9], gR[9], zero[9];
        for (int c = 0; c < 3; ++c) {
          xm[c] = (double)cf(M_XIX + c, p);
          em[c] = (double)cf(M_ETX + c, p);
          zm[c] = (double)cf(M_ZTX + c, p);
        }
        for (int e = 0; e < 9; ++e) {
          Rx[e] = (double)rdzdx[e * SS + sp];
          Ry[e] = (double)rdzdy[e * SS + sp];
          zero[e] = 0.0;
        }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) gR[3 * r + c] = (double)lVz[r] * (double)V[c];
        if (ax == 0) rdz_grad_dev(xm, em, zm, (double)mu, (double)lam, Rx, Ry, gR, zero, gmu, glam);
        else rdz_grad_dev(xm, em, zm, (double)mu, (double)lam, Rx, Ry, zero, gR, gmu, glam);
      }
    }
  }
  a.gmu[p] += gmu;
  a.glam[p] += glam;
  a.gb[p] += gb;
}

// ---------------------------------------------------------------------------
// small kernels of the adjoint driver
// ---------------------------------------------------------------------------
// adjoint source at the stations: mu[c](station) += src[s][n][c], c < 3;
// stations sharing a cell add in station order
template <class R>
__global__ void cg_adj_inject(R* mu, long long N, int ns, const long long* __restrict__ cell, const R* __restrict__ src,
                              int nsteps, int n) {
  // one thread per cell: the first station of each cell adds the sources of
  // all stations in that cell, in station order
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const long long p = cell[s];
  for (int t = 0; t < s; ++t)
    if (cell[t] == p) return;
  R acc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) acc[c] = mu[(long long)c * N + p];
  for (int t = s; t < ns; ++t)
    if (cell[t] == p)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += src[((long long)t * nsteps + n) * 3 + c];
#pragma unroll
  for (int c = 0; c < 3; ++c) mu[(long long)c * N + p] = acc[c];
}

// gradient of the scaled moment-rate table (time-major [t][p][6]) from the
// adjoint of W after the source injection: W[3+q] -= ((m1 - m0) tw + m0) DT
template <class R>
__global__ void cg_adj_source(const R* __restrict__ out, long long N, int np, const long long* __restrict__ cell,
                              double* grate, int srcIt, double tw, double DT) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const long long c = cell[p];
  for (int q = 0; q < 6; ++q) {
    const double gs = -DT * (double)out[(long long)(3 + q) * N + c];
    grate[((long long)srcIt * np + p) * 6 + q] += (1.0 - tw) * gs;
    grate[((long long)(srcIt + 1) * np + p) * 6 + q] += tw * gs;
  }
}

// velocity at the stations after step n in the working precision
template <class R>
__global__ void cg_record_vel(const R* __restrict__ W, long long N, int ns, const long long* __restrict__ cell, R* rec,
                              int nsteps, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  for (int c = 0; c < 3; ++c) rec[((long long)s * nsteps + n) * 3 + c] = W[(long long)c * N + cell[s]];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// the three passes of one adjoint stage for one F/B pattern
template <class R, class TM, int F1, int F2, int F3, int STAGE>
void adj_passes(const AdjArgs<R>& a, const TM* coef, const R* jac, const R* rdzdx, const R* rdzdy, dim3 grid, dim3 gs,
                dim3 block, int mode, hipStream_t st) {
  const bool s = mode == ADJ_FULL || mode == ADJ_SA || mode == ADJ_S;
  if (a.fs && s) cg_adj_s<R, TM, F1, F2, F3, STAGE><<<gs, block, 0, st>>>(a, coef, jac, rdzdx, rdzdy);
  if (mode == ADJ_FULL || mode == ADJ_SA) cg_adj_a<R, TM, F1, F2, F3, STAGE><<<grid, block, 0, st>>>(a, coef);
  if (mode == ADJ_FULL || mode == ADJ_B) cg_adj_b<R, F1, F2, F3, STAGE><<<grid, block, 0, st>>>(a);
  if constexpr (STAGE > 0)
    if (mode == ADJ_BA) cg_adj_ba<R, TM, F1, F2, F3, STAGE><<<grid, block, 0, st>>>(a, coef);
}

template <class R, class TM, int F1, int F2, int F3>
void adj_stage_fb(const AdjArgs<R>& a, const TM* coef, const R* jac, const R* rdzdx, const R* rdzdy, dim3 grid, dim3 gs,
                  dim3 block, int mode, hipStream_t st) {
  switch (a.stage) {
    case 0: adj_passes<R, TM, F1, F2, F3, 0>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 1: adj_passes<R, TM, F1, F2, F3, 1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 2: adj_passes<R, TM, F1, F2, F3, 2>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    default: adj_passes<R, TM, F1, F2, F3, 3>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
  }
}

template <class R, class TM>
int launch_adj_stage(AdjArgs<R> a, const TM* coef, const R* jac, const R* rdzdx, const R* rdzdy, int fb1, int fb2,
                     int fb3, int mode, hipStream_t st) {
  if (mode < ADJ_FULL || mode > ADJ_B) return LAUNCH_BAD_STAGE;
  if (mode == ADJ_BA && a.stage < 1) return LAUNCH_BAD_STAGE;
  if (a.stage < 0 || a.stage > 3) return LAUNCH_BAD_STAGE;
  if (a.nx <= 0 || a.ny <= 0 || a.nz <= 0) return LAUNCH_EMPTY;
  if ((a.ny + 3) / 4 > 65535) return LAUNCH_TOO_LARGE;
  if (a.rty < 1) a.rty = 1;
  const long long nb = interior_blocks(a.nx, a.ny, a.nz, a.rtx, a.rty);
  if (nb > 0x7fffffffLL) return LAUNCH_TOO_LARGE;
  const dim3 grid((unsigned)nb), block(256);
  const dim3 gs((a.nx + 63) / 64, (a.ny + 3) / 4, 3);
  // the surface pass reads the auxiliary adjoint g of this stage, which pass A
  // then overwrites with that of the next stage: surface, A, B in this order
  switch (fb_index(fb1, fb2, fb3)) {
    case 0: adj_stage_fb<R, TM, -1, -1, -1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 1: adj_stage_fb<R, TM, 1, -1, -1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 2: adj_stage_fb<R, TM, -1, 1, -1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 3: adj_stage_fb<R, TM, 1, 1, -1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 4: adj_stage_fb<R, TM, -1, -1, 1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 5: adj_stage_fb<R, TM, 1, -1, 1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    case 6: adj_stage_fb<R, TM, -1, 1, 1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
    default: adj_stage_fb<R, TM, 1, 1, 1>(a, coef, jac, rdzdx, rdzdy, grid, gs, block, mode, st); break;
  }
  return LAUNCH_OK;
}

template <class R>
void launch_adj_inject(R* mu, long long N, int ns, const long long* cell, const R* src, int nsteps, int n, hipStream_t st) {
  if (ns <= 0) return;
  cg_adj_inject<R><<<(ns + 63) / 64, 64, 0, st>>>(mu, N, ns, cell, src, nsteps, n);
}

template <class R>
void launch_adj_source(const R* out, long long N, int np, const long long* cell, double* grate, int srcIt, double tw,
                       double DT, hipStream_t st) {
  if (np <= 0) return;
  cg_adj_source<R><<<(np + 127) / 128, 128, 0, st>>>(out, N, np, cell, grate, srcIt, tw, DT);
}

template <class R>
void launch_record_vel(const R* W, long long N, int ns, const long long* cell, R* rec, int nsteps, int n, hipStream_t st) {
  if (ns <= 0) return;
  cg_record_vel<R><<<(ns + 127) / 128, 128, 0, st>>>(W, N, ns, cell, rec, nsteps, n);
}

}  // namespace cg
