// Argument block and launcher declarations of the ME-AWENO (SCFDM) step.
#pragma once
#include "common.h"

namespace aw {

// Coefficient planes (SoA over the padded grid, CJM order of the SCFDM build,
// CJM.cpp:107-147): 0..8 cofactors J*xi_x .. J*zt_z, 9 J, 10 mu, 11 lambda,
// 12 1/rho, 13..15 x_xi, 16..18 x_eta.
enum { C_J = 9, C_MU = 10, C_LAM = 11, C_B = 12, C_TXI = 13, C_TET = 16, C_N = 19 };

template <class R>
struct StageArgs {
  int NX, NY, NZ;      // NX: device row pitch
  int XE;              // padded x extent nx + 6 (NX >= XE in the row-aligned layout)
  long long N, S;
  unsigned rowb, planeb;  // NX and S in bytes of one R element (every plane array < 4 GiB)
  int rtx, rty;        // block order: XCD regions of (64 rtx) x (4 rty) columns, k-major
  int stage;           // TVD-RK3 stage 0..2
  R DT, rDH, alpha;    // alpha: Lax-Friedrichs speed (global Vp max)
  const R* cjm;        // 19 coefficient planes
  const R* win;        // stage input state
  const R* wm;         // state at the start of the step
  R* wout;             // stage output state
  R* F;                // face fluxes: 27 planes, direction-major (3 x 9)
};

template <class R>
void launch_faces(const StageArgs<R>& a, hipStream_t st);
template <class R>
void launch_update(const StageArgs<R>& a, int nx, int ny, hipStream_t st, hipStream_t st_top);
template <class R>
void launch_char_surface(R* W, const R* cjm, int nx, int ny, int NX, int NZ, long long N, long long S, hipStream_t st);
template <class R>
void launch_sponge(R* W, int nx, int ny, int nz, int NX, long long N, long long S, int fs, hipStream_t st);
template <class R>
void launch_source(R* W, const R* cjm, long long N, int np, const long long* cell, const R* rate, int srcIt, R tw, R DT,
                   hipStream_t st);
template <class R>
void launch_surface_obs(const R* W, const R* cjm, float* pgv, R* dis, int nx, int ny, int NX, int NZ, long long N,
                        long long S, R DT, hipStream_t st);
template <class R>
void launch_record(const R* W, const R* cjm, long long N, int ns, const long long* cell, float* rec, int nt, int it,
                   hipStream_t st);

}  // namespace aw
