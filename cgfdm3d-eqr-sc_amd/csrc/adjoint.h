// Argument blocks and launchers of the CGFDM adjoint stage.
#pragma once
#include "kernels.h"

namespace cg {

template <class R>
struct FaceAdj {
  R* mu;    // adjoint of the auxiliary state after the step
  R* out;   // adjoint of the auxiliary state at the step start
  R* g;     // adjoint of the auxiliary right-hand side of the current stage
  long long cnt;
  R* gn;    // that of the next stage, written by the fused pass B + A
};

template <class R>
struct AdjArgs {
  int nx, ny, nz, NX, NY;
  long long N, S;
  int fs, stage;
  int rtx, rty;     // block order of passes A and B: XCD regions of (64 rtx) x (4 rty) columns
  R DT, rDH;
  const R* X;       // stage state (W planes), for the medium gradient
  const R* mu;      // adjoint of W after the step
  R* out;           // adjoint of W at the step start
  R* G;             // adjoint of hW of the current stage
  R* Z;             // 27 planes of per-direction products (read by pass B)
  R* Zn;            // the next stage's products (written by the fused pass B + A)
  R* Ls;            // 9 planes: scattered surface contributions (cleared by pass B)
  double* gmu;      // medium gradient accumulators (padded grid)
  double* glam;
  double* gb;
  PmlArgs<R> pml;   // profiles (the auxiliary buffers are not used)
  FaceAdj<R> fa[6];
};

// passes launched by launch_adj_stage for stage a.stage: surface + A + B;
// surface + A; B of a.stage fused with A of a.stage - 1 (fb the pattern of
// a.stage); surface alone; B alone
enum { ADJ_FULL = 0, ADJ_SA = 1, ADJ_BA = 2, ADJ_S = 3, ADJ_B = 4 };

template <class R, class TM>
int launch_adj_stage(AdjArgs<R> a, const TM* coef, const R* jac, const R* rdzdx, const R* rdzdy, int fb1, int fb2,
                      int fb3, int mode, hipStream_t st);
template <class R>
void launch_adj_inject(R* mu, long long N, int ns, const long long* cell, const R* src, int nsteps, int n, hipStream_t st);
template <class R>
void launch_adj_source(const R* out, long long N, int np, const long long* cell, double* grate, int srcIt, double tw,
                       double DT, hipStream_t st);
template <class R>
void launch_record_vel(const R* W, long long N, int ns, const long long* cell, R* rec, int nsteps, int n, hipStream_t st);

}  // namespace cg
