// Host launchers of the CGFDM step kernels: one direct launch per F/B pattern.
#pragma once
#include "cgfdm_step_impl.h"

namespace cg {

// F/B pattern index: bit 0 = FB1 > 0, bit 1 = FB2 > 0, bit 2 = FB3 > 0
inline int fb_index(int fb1, int fb2, int fb3) { return (fb1 > 0) + 2 * (fb2 > 0) + 4 * (fb3 > 0); }

template <class R, class TM>
void launch_interior(StageArgs<R> a, const TM* coef, int fb1, int fb2, int fb3, hipStream_t st) {
  const int nk = a.kend - a.kbeg;
  if (nk <= 0 || a.nx <= 0 || a.ny <= 0) return;
  const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4, nk), block(256);
  switch (fb_index(fb1, fb2, fb3)) {
    case 0: cg_cell<R, TM, -1, -1, -1><<<grid, block, 0, st>>>(a, coef); break;
    case 1: cg_cell<R, TM, 1, -1, -1><<<grid, block, 0, st>>>(a, coef); break;
    case 2: cg_cell<R, TM, -1, 1, -1><<<grid, block, 0, st>>>(a, coef); break;
    case 3: cg_cell<R, TM, 1, 1, -1><<<grid, block, 0, st>>>(a, coef); break;
    case 4: cg_cell<R, TM, -1, -1, 1><<<grid, block, 0, st>>>(a, coef); break;
    case 5: cg_cell<R, TM, 1, -1, 1><<<grid, block, 0, st>>>(a, coef); break;
    case 6: cg_cell<R, TM, -1, 1, 1><<<grid, block, 0, st>>>(a, coef); break;
    default: cg_cell<R, TM, 1, 1, 1><<<grid, block, 0, st>>>(a, coef); break;
  }
}

template <class R, class TM>
void launch_surface(StageArgs<R> a, const TM* coef, SurfArgs<R> s, int fb1, int fb2, int fb3, hipStream_t st) {
  const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4, 3), block(256);
  switch (fb_index(fb1, fb2, fb3)) {
    case 0: cg_surface<R, TM, -1, -1, -1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 1: cg_surface<R, TM, 1, -1, -1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 2: cg_surface<R, TM, -1, 1, -1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 3: cg_surface<R, TM, 1, 1, -1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 4: cg_surface<R, TM, -1, -1, 1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 5: cg_surface<R, TM, 1, -1, 1><<<grid, block, 0, st>>>(a, coef, s); break;
    case 6: cg_surface<R, TM, -1, 1, 1><<<grid, block, 0, st>>>(a, coef, s); break;
    default: cg_surface<R, TM, 1, 1, 1><<<grid, block, 0, st>>>(a, coef, s); break;
  }
}

template <class R>
void launch_source(R* W, long long N, int np, const long long* cell, const R* rate, int srcIt, R tw, R DT, hipStream_t st) {
  if (np <= 0) return;
  cg_add_source<R><<<dim3((np + 127) / 128), dim3(128), 0, st>>>(W, N, np, cell, rate, srcIt, tw, DT);
}

template <class R>
void launch_record(const R* W, long long N, int ns, const long long* cell, float* rec, int nt, int it, hipStream_t st) {
  if (ns <= 0 || it >= nt) return;
  cg_record<R><<<dim3((ns + 127) / 128), dim3(128), 0, st>>>(W, N, ns, cell, rec, nt, it);
}

}  // namespace cg
