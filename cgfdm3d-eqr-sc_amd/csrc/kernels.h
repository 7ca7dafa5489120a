// Kernel argument blocks and launcher declarations of the CGFDM step.
#pragma once
#include "common.h"

namespace cg {

template <class R>
struct StageArgs {
  int nx, ny, nz;       // interior cells
  int NX, NY;           // padded x/y extents
  long long N;          // padded cells per field
  long long S;          // NX * NY
  unsigned rowb, planeb;  // NX and S in bytes of one grid element (R); N * sizeof(R) < 4 GiB
  int kbeg, kend;       // k range of the interior kernel [kbeg, kend)
  int stage;            // RK4 stage 0..3
  int rtx, rty;         // interior block order: regions of (64 rtx) x (4 rty) columns per XCD
  int kc;               // > 0: z-marching interior kernel with chunks of kc planes
  int zty;              // z-marching tile height: 8 (512 threads) or 4 (256 threads)
  int zreg;             // > 0: register z-marching interior kernel with chunks of zreg planes
  R DT, rDH;
  const float* aff;     // 16-bit coefficient storage: affine table [12][NZ][2] (mid, half); else unused
  int nzp;              // padded z extent (rows of the affine table)
  const R* win;         // stencil input (9 SoA fields)
  const R* wm;          // RK base (state at the start of the step)
  R* wt;                // RK accumulator
  R* wout;              // stage output
  PmlArgs<R> pml;
};

template <class R>
struct SurfArgs {
  const R* jac;         // Jacobian, padded grid
  const R* rdzdx;       // 9 planes [NY][NX]: -DZ^-1 DX on the surface
  const R* rdzdy;       // 9 planes [NY][NX]: -DZ^-1 DY
  float* pgv;           // [ny][nx][2] or nullptr (updated on stage 3)
  R* dis;               // surface displacement, 3 planes [NY][NX]
};

// launcher results: nothing is launched unless LAUNCH_OK; LAUNCH_EMPTY means
// there is nothing to do (an empty k range), the others are caller errors
constexpr int LAUNCH_OK = 0;
constexpr int LAUNCH_EMPTY = 1;       // empty range: no launch needed
constexpr int LAUNCH_BAD_STAGE = 2;   // RK stage outside 0..3
constexpr int LAUNCH_TOO_LARGE = 3;   // grid beyond the launch limits (2^31 - 1 blocks, 65535 in y)

template <class R, class TM>
int launch_interior(StageArgs<R> a, const TM* coef, int fb1, int fb2, int fb3, hipStream_t st);
template <class R, class TM>
int launch_surface(StageArgs<R> a, const TM* coef, SurfArgs<R> s, int fb1, int fb2, int fb3, hipStream_t st);
template <class R>
void launch_source(R* W, long long N, int np, const long long* cell, const R* rate, int srcIt, R tw, R DT, hipStream_t st);
template <class R>
void launch_record(const R* W, long long N, int ns, const long long* cell, float* rec, int nt, int it, hipStream_t st);

}  // namespace cg
