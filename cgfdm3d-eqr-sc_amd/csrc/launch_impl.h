// Host launchers of the CGFDM step kernels: one direct launch per F/B pattern and RK stage.
#pragma once
#include <type_traits>
#include "cgfdm_step_impl.h"

namespace cg {

// F/B pattern index: bit 0 = FB1 > 0, bit 1 = FB2 > 0, bit 2 = FB3 > 0
inline int fb_index(int fb1, int fb2, int fb3) { return (fb1 > 0) + 2 * (fb2 > 0) + 4 * (fb3 > 0); }

// one kernel instance per F/B pattern and RK stage (the stage is a template
// parameter of the kernels: no stage-dependent control flow on the device)
template <class R, class TM, int FB1, int FB2, int FB3>
void cell_stage(const StageArgs<R>& a, const TM* coef, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_cell<R, TM, FB1, FB2, FB3, 0><<<g, b, 0, st>>>(a, coef); break;
    case 1: cg_cell<R, TM, FB1, FB2, FB3, 1><<<g, b, 0, st>>>(a, coef); break;
    case 2: cg_cell<R, TM, FB1, FB2, FB3, 2><<<g, b, 0, st>>>(a, coef); break;
    default: cg_cell<R, TM, FB1, FB2, FB3, 3><<<g, b, 0, st>>>(a, coef); break;
  }
}

// z-marching tiles: 64 x ty columns with ty = 8 (512 threads) or 4 (256)
template <class R, class TM, int FB1, int FB2, int FB3, int TY>
void zcol_stage_ty(const StageArgs<R>& a, const TM* coef, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_zcol<R, TM, FB1, FB2, FB3, 0, TY><<<g, b, 0, st>>>(a, coef); break;
    case 1: cg_zcol<R, TM, FB1, FB2, FB3, 1, TY><<<g, b, 0, st>>>(a, coef); break;
    case 2: cg_zcol<R, TM, FB1, FB2, FB3, 2, TY><<<g, b, 0, st>>>(a, coef); break;
    default: cg_zcol<R, TM, FB1, FB2, FB3, 3, TY><<<g, b, 0, st>>>(a, coef); break;
  }
}

template <class R, class TM, int FB1, int FB2, int FB3, int TY>
void plane_stage_ty(const StageArgs<R>& a, const TM* coef, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_plane<R, TM, FB1, FB2, FB3, 0, TY><<<g, b, 0, st>>>(a, coef); break;
    case 1: cg_plane<R, TM, FB1, FB2, FB3, 1, TY><<<g, b, 0, st>>>(a, coef); break;
    case 2: cg_plane<R, TM, FB1, FB2, FB3, 2, TY><<<g, b, 0, st>>>(a, coef); break;
    default: cg_plane<R, TM, FB1, FB2, FB3, 3, TY><<<g, b, 0, st>>>(a, coef); break;
  }
}

template <class R, class TM, int FB1, int FB2, int FB3>
void zreg_stage(const StageArgs<R>& a, const TM* coef, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_zreg<R, TM, FB1, FB2, FB3, 0><<<g, b, 0, st>>>(a, coef); break;
    case 1: cg_zreg<R, TM, FB1, FB2, FB3, 1><<<g, b, 0, st>>>(a, coef); break;
    case 2: cg_zreg<R, TM, FB1, FB2, FB3, 2><<<g, b, 0, st>>>(a, coef); break;
    default: cg_zreg<R, TM, FB1, FB2, FB3, 3><<<g, b, 0, st>>>(a, coef); break;
  }
}

template <int FB1, int FB2, int FB3>
void pair_stage(const StageArgs<float>& a, const float* coef, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_pair<FB1, FB2, FB3, 0><<<g, b, 0, st>>>(a, coef); break;
    case 1: cg_pair<FB1, FB2, FB3, 1><<<g, b, 0, st>>>(a, coef); break;
    case 2: cg_pair<FB1, FB2, FB3, 2><<<g, b, 0, st>>>(a, coef); break;
    default: cg_pair<FB1, FB2, FB3, 3><<<g, b, 0, st>>>(a, coef); break;
  }
}

// kc > 1: z-marching kernel; kc == 1: plane-tile kernel (one plane per block)
template <class R, class TM, int FB1, int FB2, int FB3>
void zcol_stage(const StageArgs<R>& a, const TM* coef, dim3 g, dim3 b, hipStream_t st) {
  if (a.kc == 1) {
    if (a.zty == 4) plane_stage_ty<R, TM, FB1, FB2, FB3, 4>(a, coef, g, b, st);
    else plane_stage_ty<R, TM, FB1, FB2, FB3, 8>(a, coef, g, b, st);
  } else if (a.zty == 4) {
    zcol_stage_ty<R, TM, FB1, FB2, FB3, 4>(a, coef, g, b, st);
  } else {
    zcol_stage_ty<R, TM, FB1, FB2, FB3, 8>(a, coef, g, b, st);
  }
}

template <class R, class TM, int FB1, int FB2, int FB3>
void surface_stage(const StageArgs<R>& a, const TM* coef, const SurfArgs<R>& s, dim3 g, dim3 b, hipStream_t st) {
  switch (a.stage) {
    case 0: cg_surface<R, TM, FB1, FB2, FB3, 0><<<g, b, 0, st>>>(a, coef, s); break;
    case 1: cg_surface<R, TM, FB1, FB2, FB3, 1><<<g, b, 0, st>>>(a, coef, s); break;
    case 2: cg_surface<R, TM, FB1, FB2, FB3, 2><<<g, b, 0, st>>>(a, coef, s); break;
    default: cg_surface<R, TM, FB1, FB2, FB3, 3><<<g, b, 0, st>>>(a, coef, s); break;
  }
}

template <class R, class TM>
int launch_interior(StageArgs<R> a, const TM* coef, int fb1, int fb2, int fb3, hipStream_t st) {
  const int nk = a.kend - a.kbeg;
  if (a.stage < 0 || a.stage > 3) return LAUNCH_BAD_STAGE;
  if (nk <= 0 || a.nx <= 0 || a.ny <= 0) return LAUNCH_EMPTY;
  if constexpr (std::is_same<R, float>::value && std::is_same<TM, float>::value) {
    // 8-byte pair accesses need every array's even columns on 8-byte boundaries
    auto al8 = [](const void* q) { return (reinterpret_cast<unsigned long long>(q) & 7) == 0; };
    if (a.kc < 0 && a.NX % 2 == 0 && a.N % 2 == 0 && al8(a.win) && al8(a.wout) && al8(a.wm) && al8(a.wt) && al8(coef)) {
      // x-paired kernel: 128 x 4 cells per block, 8-byte accesses
      const long long nb = (long long)((a.nx + 1 + 127) / 128) * ((a.ny + 3) / 4) * nk;
      if (nb > 0x7fffffffLL) return LAUNCH_TOO_LARGE;
      const dim3 grid((unsigned)nb), block(256);
      switch (fb_index(fb1, fb2, fb3)) {
        case 0: pair_stage<-1, -1, -1>(a, coef, grid, block, st); break;
        case 1: pair_stage<1, -1, -1>(a, coef, grid, block, st); break;
        case 2: pair_stage<-1, 1, -1>(a, coef, grid, block, st); break;
        case 3: pair_stage<1, 1, -1>(a, coef, grid, block, st); break;
        case 4: pair_stage<-1, -1, 1>(a, coef, grid, block, st); break;
        case 5: pair_stage<1, -1, 1>(a, coef, grid, block, st); break;
        case 6: pair_stage<-1, 1, 1>(a, coef, grid, block, st); break;
        default: pair_stage<1, 1, 1>(a, coef, grid, block, st); break;
      }
      return LAUNCH_OK;
    }
  }
  if constexpr (std::is_same<R, float>::value) {
   if (a.kc > 0) {
    // z-marching kernel (FP32 wavefield): tiles of 64 x zty columns, chunks of kc planes
    if (a.zty != 4) a.zty = 8;
    const long long nb = (long long)((a.nx + 63) / 64) * ((a.ny + a.zty - 1) / a.zty) * ((nk + a.kc - 1) / a.kc);
    if (nb > 0x7fffffffLL) return LAUNCH_TOO_LARGE;
    const dim3 grid((unsigned)nb), block(64 * a.zty);
    switch (fb_index(fb1, fb2, fb3)) {
      case 0: zcol_stage<R, TM, -1, -1, -1>(a, coef, grid, block, st); break;
      case 1: zcol_stage<R, TM, 1, -1, -1>(a, coef, grid, block, st); break;
      case 2: zcol_stage<R, TM, -1, 1, -1>(a, coef, grid, block, st); break;
      case 3: zcol_stage<R, TM, 1, 1, -1>(a, coef, grid, block, st); break;
      case 4: zcol_stage<R, TM, -1, -1, 1>(a, coef, grid, block, st); break;
      case 5: zcol_stage<R, TM, 1, -1, 1>(a, coef, grid, block, st); break;
      case 6: zcol_stage<R, TM, -1, 1, 1>(a, coef, grid, block, st); break;
      default: zcol_stage<R, TM, 1, 1, 1>(a, coef, grid, block, st); break;
    }
    return LAUNCH_OK;
   }
  }
  if (a.rty < 1) a.rty = 1;
  if (a.zreg > 0) {
    // register z-marching kernel: 64 x 4 column tiles, chunks of zreg planes
    const long long nb = zreg_blocks(a.nx, a.ny, nk, a.zreg, a.rtx, a.rty);
    if (nb > 0x7fffffffLL) return LAUNCH_TOO_LARGE;
    const dim3 grid((unsigned)nb), block(256);
    switch (fb_index(fb1, fb2, fb3)) {
      case 0: zreg_stage<R, TM, -1, -1, -1>(a, coef, grid, block, st); break;
      case 1: zreg_stage<R, TM, 1, -1, -1>(a, coef, grid, block, st); break;
      case 2: zreg_stage<R, TM, -1, 1, -1>(a, coef, grid, block, st); break;
      case 3: zreg_stage<R, TM, 1, 1, -1>(a, coef, grid, block, st); break;
      case 4: zreg_stage<R, TM, -1, -1, 1>(a, coef, grid, block, st); break;
      case 5: zreg_stage<R, TM, 1, -1, 1>(a, coef, grid, block, st); break;
      case 6: zreg_stage<R, TM, -1, 1, 1>(a, coef, grid, block, st); break;
      default: zreg_stage<R, TM, 1, 1, 1>(a, coef, grid, block, st); break;
    }
    return LAUNCH_OK;
  }
  const long long nb = interior_blocks(a.nx, a.ny, nk, a.rtx, a.rty);
  if (nb > 0x7fffffffLL) return LAUNCH_TOO_LARGE;
  const dim3 grid((unsigned)nb), block(256);
  switch (fb_index(fb1, fb2, fb3)) {
    case 0: cell_stage<R, TM, -1, -1, -1>(a, coef, grid, block, st); break;
    case 1: cell_stage<R, TM, 1, -1, -1>(a, coef, grid, block, st); break;
    case 2: cell_stage<R, TM, -1, 1, -1>(a, coef, grid, block, st); break;
    case 3: cell_stage<R, TM, 1, 1, -1>(a, coef, grid, block, st); break;
    case 4: cell_stage<R, TM, -1, -1, 1>(a, coef, grid, block, st); break;
    case 5: cell_stage<R, TM, 1, -1, 1>(a, coef, grid, block, st); break;
    case 6: cell_stage<R, TM, -1, 1, 1>(a, coef, grid, block, st); break;
    default: cell_stage<R, TM, 1, 1, 1>(a, coef, grid, block, st); break;
  }
  return LAUNCH_OK;
}

template <class R, class TM>
int launch_surface(StageArgs<R> a, const TM* coef, SurfArgs<R> s, int fb1, int fb2, int fb3, hipStream_t st) {
  if (a.stage < 0 || a.stage > 3) return LAUNCH_BAD_STAGE;
  if (a.nx <= 0 || a.ny <= 0) return LAUNCH_EMPTY;
  if ((a.ny + 3) / 4 > 65535) return LAUNCH_TOO_LARGE;
  const dim3 grid((a.nx + 63) / 64, (a.ny + 3) / 4, 3), block(256);
  switch (fb_index(fb1, fb2, fb3)) {
    case 0: surface_stage<R, TM, -1, -1, -1>(a, coef, s, grid, block, st); break;
    case 1: surface_stage<R, TM, 1, -1, -1>(a, coef, s, grid, block, st); break;
    case 2: surface_stage<R, TM, -1, 1, -1>(a, coef, s, grid, block, st); break;
    case 3: surface_stage<R, TM, 1, 1, -1>(a, coef, s, grid, block, st); break;
    case 4: surface_stage<R, TM, -1, -1, 1>(a, coef, s, grid, block, st); break;
    case 5: surface_stage<R, TM, 1, -1, 1>(a, coef, s, grid, block, st); break;
    case 6: surface_stage<R, TM, -1, 1, 1>(a, coef, s, grid, block, st); break;
    default: surface_stage<R, TM, 1, 1, 1>(a, coef, s, grid, block, st); break;
  }
  return LAUNCH_OK;
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
