// DISABLED kernel path, kept for reference and not compiled into the library.
//
// Z-marching interior stage kernel (LDS xy tile with 3 ghost cells, register
// queue in z, XCD-aware tile order). Its first GPU run ended in an illegal
// memory access; a host emulation of this exact source (one CPU thread per GPU
// thread, a real barrier for __syncthreads, every device buffer between
// PROT_NONE guard pages, AddressSanitizer) runs clean and reproduces the FP64
// oracle to 1e-16, so no cause in the indexing has been shown. Until one is,
// this path stays out of the build.
#pragma once
#include "cgfdm_step_impl.h"

namespace cg {

// ---------------------------------------------------------------------------
// interior kernel
// ---------------------------------------------------------------------------
template <class R, class TM, int FB1, int FB2, int FB3, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void cg_interior(StageArgs<R> a, const TM* __restrict__ coef) {
  constexpr int LX = TX + 6, LY = TY + 6, LP = LX * LY;  // LDS plane with 3 ghost cells per side
  constexpr int NT = TX * TY;
  constexpr int HXL = FB1 > 0 ? 1 : 3, HXR = 4 - HXL;     // x halo actually needed (left/right)
  constexpr int HYL = FB2 > 0 ? 1 : 3, HYR = 4 - HYL;
  constexpr int NH = 4 * TY + 4 * TX;                     // halo entries per field
  constexpr int NHR = (9 * NH + NT - 1) / NT;             // halo entries per thread
  constexpr int LO = FB3 > 0 ? -1 : -3;                   // z queue holds planes k+LO .. k+LO+4
  constexpr int CZ = -LO;                                 // queue slot of plane k
  __shared__ R lds[2][9][LP];

  // XCD-aware tile order: hardware hands consecutive block ids to the 8 XCDs in
  // turn; give each XCD a contiguous run of tiles so halo re-reads hit its L2.
  const int nb = gridDim.x;
  int bid = blockIdx.x;
  int logical = (bid & 7) * (nb >> 3) + (bid >> 3);
  if (logical >= a.ntiles) return;  // uniform per block, before any barrier
  const int tile_x = logical % a.ntx;
  const int rest = logical / a.ntx;
  const int tile_y = rest % a.nty;
  const int chunk = rest / a.nty;

  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = H + tile_x * TX + tx;
  const int j = H + tile_y * TY + ty;
  const bool in_pad = (i < a.NX) && (j < a.NY);
  const bool active = (i < H + a.nx) && (j < H + a.ny);
  const int kbeg = a.kbeg + chunk * a.kc;
  const int kend = min(kbeg + a.kc, a.kend);
  const long long N = a.N, S = a.S;
  const long long col = (long long)i + (long long)a.NX * j;

  // halo bookkeeping: entry e -> (field, LDS slot, column offset); -1 = none
  int h_lds[NHR];
  long long h_col[NHR];
  int h_fld[NHR];
#pragma unroll
  for (int r = 0; r < NHR; ++r) {
    int e = threadIdx.x + r * NT;
    h_lds[r] = -1;
    h_col[r] = 0;
    h_fld[r] = 0;
    if (e < 9 * NH) {
      int v = e / NH, q = e % NH;
      int gx, gy;  // tile-local coordinates of the halo cell (may be negative)
      if (q < 4 * TY) {  // x halo: 4 columns x TY rows
        int c = q / TY, rr = q % TY;
        gx = c < HXL ? -HXL + c : TX + (c - HXL);
        gy = rr;
      } else {           // y halo: 4 rows x TX columns
        int q2 = q - 4 * TY;
        int rr = q2 / TX, c = q2 % TX;
        gy = rr < HYL ? -HYL + rr : TY + (rr - HYL);
        gx = c;
      }
      int gi = H + tile_x * TX + gx, gj = H + tile_y * TY + gy;
      h_fld[r] = v;
      h_lds[r] = (gy + 3) * LX + (gx + 3);
      if (gi >= 0 && gi < a.NX && gj >= 0 && gj < a.NY) h_col[r] = (long long)gi + (long long)a.NX * gj;
      else h_col[r] = -1;
    }
  }

  R zq[9][5];
  R nxt[9];
  R hreg[NHR];
  // prologue: planes kbeg+LO .. kbeg+LO+3 into slots 1..4, plane kbeg+LO+4 prefetched
#pragma unroll
  for (int v = 0; v < 9; ++v) {
#pragma unroll
    for (int m = 1; m < 5; ++m) zq[v][m] = in_pad ? a.win[(long long)v * N + col + (long long)(kbeg + LO + m - 1) * S] : R(0);
    nxt[v] = in_pad ? a.win[(long long)v * N + col + (long long)(kbeg + LO + 4) * S] : R(0);
  }
#pragma unroll
  for (int r = 0; r < NHR; ++r)
    hreg[r] = (h_lds[r] >= 0 && h_col[r] >= 0) ? a.win[(long long)h_fld[r] * N + h_col[r] + (long long)kbeg * S] : R(0);

  const PmlArgs<R>& P = a.pml;
  const bool in_x1 = P.npml > 0 && i < H + P.npml;
  const bool in_x2 = P.npml > 0 && i >= H + a.nx - P.npml;
  const bool in_y1 = P.npml > 0 && j < H + P.npml;
  const bool in_y2 = P.npml > 0 && j >= H + a.ny - P.npml;
  const R bx = active ? P.beta[0][i] : R(1);
  const R by = active ? P.beta[1][j] : R(1);

  int buf = 0;
  for (int k = kbeg; k < kend; ++k) {
#pragma unroll
    for (int v = 0; v < 9; ++v) {
#pragma unroll
      for (int m = 0; m < 4; ++m) zq[v][m] = zq[v][m + 1];
      zq[v][4] = nxt[v];
    }
    R* L = &lds[buf][0][0];
#pragma unroll
    for (int v = 0; v < 9; ++v) L[v * LP + (ty + 3) * LX + tx + 3] = zq[v][CZ];
#pragma unroll
    for (int r = 0; r < NHR; ++r)
      if (h_lds[r] >= 0) L[h_fld[r] * LP + h_lds[r]] = hreg[r];
    if (k + 1 < kend) {
#pragma unroll
      for (int v = 0; v < 9; ++v) nxt[v] = in_pad ? a.win[(long long)v * N + col + (long long)(k + 1 + LO + 4) * S] : R(0);
#pragma unroll
      for (int r = 0; r < NHR; ++r)
        hreg[r] = (h_lds[r] >= 0 && h_col[r] >= 0) ? a.win[(long long)h_fld[r] * N + h_col[r] + (long long)(k + 1) * S] : R(0);
    }
    __syncthreads();

    if (active) {
      const long long p = col + (long long)k * S;
      R h[9];
      const R mu = R(Store<TM>::ld(coef + (long long)M_MU * N + p));
      const R lam = R(Store<TM>::ld(coef + (long long)M_LAM * N + p));
      const R bu = R(Store<TM>::ld(coef + (long long)M_B * N + p));
      const R bz = P.beta[2][k];
      const int c0 = (ty + 3) * LX + tx + 3;
      R D[9], Q[9];
      // x direction
#pragma unroll
      for (int v = 0; v < 9; ++v) D[v] = mac<FB1>(L + v * LP + c0, 1);
      assemble(D, R(Store<TM>::ld(coef + (long long)M_XIX * N + p)), R(Store<TM>::ld(coef + (long long)M_XIY * N + p)),
               R(Store<TM>::ld(coef + (long long)M_XIZ * N + p)), mu, lam, bu, Q);
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        Q[v] *= a.rDH;
        h[v] = bx * Q[v] * a.DT;
      }
      if (in_x1 || in_x2) {
        R d = P.d[0][i], ad = d + P.alpha[0][i];
        long long rowq = (long long)(j - H) + (long long)a.ny * (k - H);
        if (in_x1) pml_face<R>(P.f[0], (long long)(i - H) + (long long)P.npml * rowq, Q, d, ad, bx, a.DT, a.stage, h, nullptr);
        if (in_x2) pml_face<R>(P.f[3], (long long)(i - H - (a.nx - P.npml)) + (long long)P.npml * rowq, Q, d, ad, bx, a.DT, a.stage, h, nullptr);
      }
      // y direction
#pragma unroll
      for (int v = 0; v < 9; ++v) D[v] = mac<FB2>(L + v * LP + c0, LX);
      assemble(D, R(Store<TM>::ld(coef + (long long)M_ETX * N + p)), R(Store<TM>::ld(coef + (long long)M_ETY * N + p)),
               R(Store<TM>::ld(coef + (long long)M_ETZ * N + p)), mu, lam, bu, Q);
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        Q[v] *= a.rDH;
        h[v] += by * Q[v] * a.DT;
      }
      if (in_y1 || in_y2) {
        R d = P.d[1][j], ad = d + P.alpha[1][j];
        long long i0 = i - H, k0 = k - H;
        if (in_y1) pml_face<R>(P.f[1], i0 + (long long)a.nx * ((j - H) + (long long)P.npml * k0), Q, d, ad, by, a.DT, a.stage, h, nullptr);
        if (in_y2) pml_face<R>(P.f[4], i0 + (long long)a.nx * ((j - H - (a.ny - P.npml)) + (long long)P.npml * k0), Q, d, ad, by, a.DT, a.stage, h, nullptr);
      }
      // z direction (register queue)
#pragma unroll
      for (int v = 0; v < 9; ++v)
        D[v] = R(FB3) * (R(AF_1) * zq[v][CZ - FB3] + R(AF0) * zq[v][CZ] + R(AF1) * zq[v][CZ + FB3] +
                         R(AF2) * zq[v][CZ + 2 * FB3] + R(AF3) * zq[v][CZ + 3 * FB3]);
      assemble(D, R(Store<TM>::ld(coef + (long long)M_ZTX * N + p)), R(Store<TM>::ld(coef + (long long)M_ZTY * N + p)),
               R(Store<TM>::ld(coef + (long long)M_ZTZ * N + p)), mu, lam, bu, Q);
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        Q[v] *= a.rDH;
        h[v] += bz * Q[v] * a.DT;
      }
      const bool in_z1 = P.npml > 0 && k < H + P.npml;
      const bool in_z2 = P.has_z2 && k >= H + a.nz - P.npml;
      if (in_z1 || in_z2) {
        R d = P.d[2][k], ad = d + P.alpha[2][k];
        long long i0 = i - H, j0 = j - H;
        if (in_z1) pml_face<R>(P.f[2], i0 + (long long)a.nx * (j0 + (long long)a.ny * (k - H)), Q, d, ad, bz, a.DT, a.stage, h, nullptr);
        if (in_z2) pml_face<R>(P.f[5], i0 + (long long)a.nx * (j0 + (long long)a.ny * (k - H - (a.nz - P.npml))), Q, d, ad, bz, a.DT, a.stage, h, nullptr);
      }
      // RK4 stage update of W
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        long long q = (long long)v * N + p;
        if (a.stage == 0) {
          R m = zq[v][CZ];
          a.wt[q] = m + R(RK_B1) * h[v];
          a.wout[q] = m + R(RK_A2) * h[v];
        } else if (a.stage == 1) {
          a.wt[q] = a.wt[q] + R(RK_B2) * h[v];
          a.wout[q] = a.wm[q] + R(RK_A3) * h[v];
        } else if (a.stage == 2) {
          a.wt[q] = a.wt[q] + R(RK_B3) * h[v];
          a.wout[q] = a.wm[q] + h[v];
        } else {
          a.wout[q] = a.wt[q] + R(RK_B4) * h[v];
        }
      }
    }
    buf ^= 1;
  }
}

}  // namespace cg
