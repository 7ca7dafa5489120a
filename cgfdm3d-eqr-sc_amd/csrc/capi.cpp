// C ABI of the MI355X CGFDM library: solver lifetime, setup, time stepping
// and observers. Host orchestration only; all grid work runs in the HIP
// kernels of setup.hip and step_*.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "cgfdm_amd.h"
#include "common.h"
#include "aweno.h"
#include "kernels.h"
#include "adjoint.h"
#include "setup.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

// hooks of the params.json session code (host/session.cpp), linked into the library
namespace cgfdm_detail {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
void release_session(cgfdm_solver* s);
}  // namespace cgfdm_detail

namespace {

// Code paths that can be switched off as a whole: stepping returns
// CGFDM_ERR_DISABLED without launching a kernel. Both are on; the switch stays
// so that a path can be taken out of service without touching its callers.
constexpr bool kCgfdmStepEnabled = true;
constexpr bool kCgfdmAdjointEnabled = true;
const char* const kCgfdmAdjointOff = "the CGFDM adjoint is switched off in this build";
const char* const kCgfdmStepOff = "the CGFDM step kernels are switched off in this build";

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(CGFDM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A stage launcher that refuses its launch (invalid stage, grid beyond the
// launch limits) is a caller error: the step fails instead of leaving the
// wavefield one stage behind.
#define LAUNCHCHK(x, what)                                                                         \
  do {                                                                                             \
    const int r_ = (x);                                                                            \
    if (r_ == cg::LAUNCH_BAD_STAGE) return fail(CGFDM_ERR_ARG, std::string(what) + ": RK stage out of range"); \
    if (r_ == cg::LAUNCH_TOO_LARGE)                                                                \
      return fail(CGFDM_ERR_ARG, std::string(what) + ": grid beyond the kernel launch limits");     \
  } while (0)

// F/B schedule of the MacCormack operators, selected by it % 8 and flipped
// after every RK stage (reference propagate.cpp:158-168, 195-197, 248-250).
const int FBTAB[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                         {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};

constexpr double REF_PI = 3.141592657;  // reference macro.h:70, enters the PML alpha profile

struct SolverBase {
  virtual ~SolverBase() {
    for (auto& e : ev) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    // teardown: nothing to report to
    for (void* p : allocs) (void)hipFree(p);
  }
  virtual int init(const cgfdm_desc& d, const float* coord, const float* medium) = 0;
  virtual int set_source(int np, int nt, double dt, const long long* cell, const float* rate) = 0;
  virtual int set_stations(int ns, const int* ijk) = 0;
  virtual int step(int n) = 0;
  virtual int get_stations(float* out) = 0;
  virtual int get_pgv(float* out) = 0;
  virtual int get_dis(float* out) = 0;
  virtual int get_field(int comp, double* out) = 0;
  virtual int set_field(int comp, const double* in) = 0;
  virtual int get_setup(int what, double* out) = 0;
  // one component of the current state on an interior plane, physical units
  virtual int get_slice(int comp, int axis, int index, float* out) = 0;
  virtual const char* kernel_name(int which) const = 0;
  virtual int forward_record(int, int, double*) { return fail(CGFDM_ERR_ARG, "the adjoint is available for the CGFDM scheme"); }
  virtual int adjoint(const double*, double*, double*, double*, double*) {
    return fail(CGFDM_ERR_ARG, "the adjoint is available for the CGFDM scheme");
  }
  hipStream_t stream = nullptr;
  int it = 0;
  // padded extents as seen by the plane copies (set by the solvers at init):
  // pNX is the device row pitch, pXe the padded x extent (nx + 6)
  int pNX = 0, pNY = 0, pNZ = 0, pXe = 0;

  // Strided device-to-host copy of the interior of one padded plane of a
  // plane-major array (x fastest): axis 0 = the y-z plane at i = index
  // ([nz][ny]), 1 = the x-z plane at j = index ([nz][nx]), 2 = the x-y plane at
  // k = index ([ny][nx]). No kernel: hipMemcpy2D with the plane's stride.
  template <class T>
  int copy_plane(const T* plane0, int axis, int index, std::vector<T>& out) {
    const int H = 3;
    const size_t e = sizeof(T);
    const long long S2 = (long long)pNX * pNY;
    int na, nb;
    if (axis == 0) {
      na = pNY - 2 * H;
      nb = pNZ - 2 * H;
      out.resize((size_t)na * nb);
      // one element per padded row, rows j = H.. of planes k = H..
      for (int b = 0; b < nb; ++b)
        HIPCHK(hipMemcpy2D(out.data() + (size_t)b * na, e, plane0 + index + (long long)pNX * H + S2 * (b + H),
                           e * pNX, e, na, hipMemcpyDeviceToHost));
    } else if (axis == 1) {
      na = pXe - 2 * H;
      nb = pNZ - 2 * H;
      out.resize((size_t)na * nb);
      HIPCHK(hipMemcpy2D(out.data(), e * na, plane0 + H + (long long)pNX * index + S2 * H, e * S2, e * na, nb,
                         hipMemcpyDeviceToHost));
    } else {
      na = pXe - 2 * H;
      nb = pNY - 2 * H;
      out.resize((size_t)na * nb);
      HIPCHK(hipMemcpy2D(out.data(), e * na, plane0 + H + (long long)pNX * H + S2 * index, e * pNX, e * na, nb,
                         hipMemcpyDeviceToHost));
    }
    return CGFDM_OK;
  }
  // Rows of pXe elements between the tight layout of the C ABI (pitch pXe)
  // and the device layout (pitch pNX); strided hipMemcpy2D, no kernel.
  // repack: device tight <-> device pitched (`to_dev` = tight -> pitched)
  template <class X>
  int repack(X* dev, const X* tight, long long rows, bool to_dev) {
    const size_t e = sizeof(X);
    if (to_dev)
      HIPCHK(hipMemcpy2DAsync(dev, e * pNX, tight, e * pXe, e * pXe, (size_t)rows, hipMemcpyDeviceToDevice, stream));
    else
      HIPCHK(hipMemcpy2DAsync((void*)tight, e * pXe, dev, e * pNX, e * pXe, (size_t)rows, hipMemcpyDeviceToDevice,
                              stream));
    return CGFDM_OK;
  }
  // `planes` device volumes to / from a tight host array
  template <class X>
  int to_host(const X* dev, X* host, int planes) {
    const size_t e = sizeof(X);
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy2D(host, e * pXe, dev, e * pNX, e * pXe, (size_t)planes * pNY * pNZ, hipMemcpyDeviceToHost));
    return CGFDM_OK;
  }
  template <class X>
  int to_dev(X* dev, const X* host, int planes) {
    const size_t e = sizeof(X);
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy2D(dev, e * pNX, host, e * pXe, e * pXe, (size_t)planes * pNY * pNZ, hipMemcpyHostToDevice));
    return CGFDM_OK;
  }
  // tight linear cell index -> device cell index (-1 outside the interior)
  long long dev_cell(long long c) const {
    const long long tS = (long long)pXe * pNY;
    if (c < 0 || c >= tS * pNZ) return -1;
    const int i = (int)(c % pXe), j = (int)((c / pXe) % pNY), k = (int)(c / tS);
    if (i < 3 || i >= pXe - 3 || j < 3 || j >= pNY - 3 || k < 3 || k >= pNZ - 3) return -1;
    return (long long)i + (long long)pNX * j + (long long)k * pNX * pNY;
  }

  int check_plane(int comp, int axis, int index) {
    const int n = axis == 0 ? pXe : axis == 1 ? pNY : pNZ;
    if (comp < 0 || comp > 8 || axis < 0 || axis > 2 || index < 3 || index >= n - 3)
      return fail(CGFDM_ERR_ARG, "slice component, axis or index out of range");
    return CGFDM_OK;
  }
  double ranges[8] = {0};
  long long bytes = 0;

  // device allocations owned by the solver, zero-filled, and their sizes
  std::vector<void*> allocs;
  std::vector<size_t> alloc_bytes;
  template <class X>
  int dalloc(X** p, long long count) {
    size_t nb = (size_t)count * sizeof(X);
    if (nb == 0) nb = 16;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return fail(CGFDM_ERR_ALLOC, "hipMalloc of " + std::to_string(nb) + " bytes failed");
    e = hipMemsetAsync(q, 0, nb, stream);
    if (e != hipSuccess) {
      (void)hipFree(q);  // the memset's error is the one reported
      return fail(CGFDM_ERR_HIP, "hipMemset failed");
    }
    allocs.push_back(q);
    alloc_bytes.push_back(nb);
    bytes += (long long)nb;
    *p = (X*)q;
    return CGFDM_OK;
  }
  // Grid-shaped buffers of the stage kernels in the row-aligned layout:
  // allocated with spare elements and offset so that the first interior
  // column (i = 3) of every row sits on a 128-byte boundary (with a row pitch
  // that is a multiple of 32 elements). `galign` = false keeps offset 0.
  // `gfirst` is the column whose rows start on a line: 3, the first interior
  // cell, where the 64-column tiles of the stage kernels of both schemes start.
  bool galign = true;
  int gfirst = 3;
  template <class X>
  int galloc(X** p, long long count) {
    X* q = nullptr;
    const long long line = 128 / (long long)sizeof(X);
    if (int rc = dalloc(&q, count + 2 * line)) return rc;
    *p = q + (galign ? (line - gfirst) % line : 0);
    return CGFDM_OK;
  }

  void dfree(void* p) {
    for (size_t q = 0; q < allocs.size(); ++q)
      if (allocs[q] == p) {
        (void)hipFree(p);  // a failed free leaves nothing to recover
        bytes -= (long long)alloc_bytes[q];
        allocs.erase(allocs.begin() + q);
        alloc_bytes.erase(alloc_bytes.begin() + q);
        return;
      }
  }

  // kernel timing with HIP events on the solver's stream
  bool timing_on = false;
  struct EvPair {
    hipEvent_t a, b;
    int kind;  // 0 main stencil kernel, 1 second kernel, 2 whole step
  };
  std::vector<EvPair> ev;
  size_t ev_used = 0;
  double tacc[6] = {0};
  size_t next_pair(int kind) {
    if (ev_used >= ev.size()) {
      EvPair e{};
      // a failed creation surfaces at the checked hipEventRecord of this pair
      (void)hipEventCreate(&e.a);
      (void)hipEventCreate(&e.b);
      ev.push_back(e);
    }
    ev[ev_used].kind = kind;
    return ev_used++;
  }
  int timing(int enable, double* out) {
    if (timing_on && !enable) {
      HIPCHK(hipStreamSynchronize(stream));
      for (size_t q = 0; q < ev_used; ++q) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[q].a, ev[q].b));
        tacc[2 * ev[q].kind] += ms;
        tacc[2 * ev[q].kind + 1] += 1;
      }
      ev_used = 0;
      timing_on = false;
    } else if (enable) {
      for (double& t : tacc) t = 0;
      ev_used = 0;
      timing_on = true;
    }
    if (out) std::memcpy(out, tacc, sizeof(tacc));
    return CGFDM_OK;
  }

  // CFL inputs of the grid (calc_CFL.cpp:255-349): ranges = [hmin, hmax,
  // vsmin, vsmax, vpmin, vpmax, rhomin, rhomax]
  int grid_ranges(const float* dcoord, const float* dmed, int nx, int ny, int nz, int NX, int NY, long long N) {
    const int nb = 1024;
    double* part = nullptr;
    int rc;
    if ((rc = dalloc(&part, 2 * nb * 4))) return rc;
    cg::launch_hrange(dcoord, part, nb, nx, ny, nz, NX, NY, N, stream);
    for (int c = 0; c < 3; ++c) cg::launch_frange(dmed + c * N, N, part + 2 * nb * (c + 1), nb, stream);
    HIPCHK(hipGetLastError());
    std::vector<double> h(2 * nb * 4);
    HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int c = 0; c < 4; ++c) {
      double lo = 1e300, hi = -1e300;
      for (int b = 0; b < nb; ++b) {
        lo = std::min(lo, h[2 * nb * c + 2 * b]);
        hi = std::max(hi, h[2 * nb * c + 2 * b + 1]);
      }
      ranges[2 * c] = lo;
      ranges[2 * c + 1] = hi;
    }
    dfree(part);
    return CGFDM_OK;
  }
};

template <class T>
float to_float(T v) { return (float)v; }
template <>
float to_float<__half>(__half v) { return __half2float(v); }
template <>
float to_float<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }

// host decoding of the 16-bit codes (IEEE binary16 / bfloat16 bit patterns)
template <class TM>
float bits16_to_float(unsigned short b);
template <>
float bits16_to_float<__half>(unsigned short b) {
  const unsigned man = b & 0x3ffu, ex = (b >> 10) & 0x1fu;
  float v;
  if (ex == 0) v = std::ldexp((float)man, -24);
  else if (ex == 31) v = man ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
  else v = std::ldexp((float)(man | 0x400u), (int)ex - 25);
  return (b & 0x8000u) ? -v : v;
}
template <>
float bits16_to_float<__hip_bfloat16>(unsigned short b) {
  const unsigned u = (unsigned)b << 16;
  float v;
  std::memcpy(&v, &u, sizeof v);
  return v;
}
template <class TM>
float bits16_to_float(unsigned short) { return 0.f; }

template <class R, class TM>
struct Solver : SolverBase {
  cgfdm_desc d{};
  // surface stage kernels on a second stream, forked from and joined to the
  // solver's stream around every stage (CGFDM_SURFACE_STREAM=0: one stream):
  // they write only the top three layers, which the interior kernel of the
  // same stage neither writes nor reads, so the two run side by side
  int surf_overlap = 1;
  hipStream_t sstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~Solver() override {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (sstream) (void)hipStreamDestroy(sstream);
  }
  // device layout: row pitch NX (nx + 6 rounded up to 32 elements in the
  // row-aligned layout), plane S = NX NY, volume N = S NZ; the host-facing
  // arrays of the C ABI use the tight layout tNX = nx + 6 (tS, tN)
  int NX = 0, NY = 0, NZ = 0;
  long long N = 0, S = 0;
  int tNX = 0;
  long long tS = 0, tN = 0;
  double Z = 1.0;  // stress carried as T / Z
  R *A = nullptr, *B = nullptr, *C = nullptr, *T = nullptr;
  TM* coef = nullptr;
  // 16-bit storage: coef points at six word planes of pair codes, aff at the
  // affine table [12][NZ][2] (common.h, Bits16)
  static constexpr bool kPacked16 = sizeof(TM) == 2;
  float* aff = nullptr;
  // 16-bit storage: mu / Z in FP32 (tight layout), read only by set_source so
  // that the moment-rate scaling mu / (J DH^3) does not carry the 16-bit error
  float* mu_src = nullptr;
  R* jac = nullptr;
  R *rdzdx = nullptr, *rdzdy = nullptr;
  R* pml_par = nullptr;  // alpha, beta, d for x (NX each), then y, then z
  cg::PmlArgs<R> pml{};
  float* pgv = nullptr;
  R* dis = nullptr;
  int ns = 0;
  long long* st_cell = nullptr;
  float* st_rec = nullptr;
  int src_np = 0, src_nt = 0;
  double src_dt = 0;
  long long* src_cell = nullptr;
  R* src_rate = nullptr;
  // host copies for the source gradient: merged unscaled table (time-major),
  // mu and 1/(J DH^3) per merged point, merged index of every input point
  std::vector<double> src_raw, src_muv, src_vol;
  std::vector<int> src_of_point;
  // adjoint: checkpoints of the last forward_record and work buffers
  int fr_n = 0, fr_ck = 1, fr_it0 = 0;
  R* ck_store = nullptr;
  R* rec_vel = nullptr;
  R* adj_buf = nullptr;  // MU, OUT, G (9N each), Z (27N), Ls (9N), X3 (9N): 72N
  R* adj_face = nullptr;
  double* adj_grad = nullptr;  // gmu, glam, gb (N each)
  double* adj_gsrc = nullptr;
  R* adj_seg = nullptr;
  R* adj_src = nullptr;
  R* adj_end = nullptr;
  // stage states X0..X3 of every step of one checkpoint segment (4 x 9N per
  // step), kept when they fit CGFDM_ADJOINT_STORE_GB (default 85 % of the free
  // device memory after the adjoint's work buffers): the adjoint then reads
  // them instead of recomputing three stages per step. stages_seg is the
  // segment whose states the store holds (-1: none): forward_record fills it
  // with its last segment, which the adjoint then does not recompute.
  R* adj_stages = nullptr;
  int stages_seg = -1;
  int stages_ck = 0;  // segment length the store was allocated for

  // interior block order: regions of (64 rtx) x (4 rty) columns per XCD; the
  // environment variables CGFDM_INTERIOR_RTX / _RTY (1..64) override it
  // (rtx 0 = whole rows of tiles, rty 2: 512 x 8-column regions on the 512 x
  // 512 x 256 workload, measured fastest there: 4.43 ms per launch against 5.08
  // ms for 64 x 16 regions; partial-width regions leave XCDs idle when the row
  // of tiles does not divide evenly)
  int rty = 2;
  int rtx = 0;
  // interior kernel choice (FP32 wavefield): 0 thread-per-cell, -1 x-paired
  // (FP32 coefficients, even NX), 1 plane tile, > 1 z-marching with chunks of kc
  // planes. CGFDM_INTERIOR_KC (-1..4096) overrides it.
  int kc = 0;
  int zty = 8;  // z-marching tile height (CGFDM_INTERIOR_ZTY: 4 or 8)
  // adjoint passes: 1 pass B of each stage fused with pass A of the next
  // (the stage's g passed in registers), 0 the three passes per stage;
  // CGFDM_ADJOINT_PIPE overrides it
  int adj_pipe = 1;
  // register z-marching kernel: planes per chunk (0 = off); CGFDM_INTERIOR_ZREG
  // (0..4096) overrides it
  int zreg = 0;

  int init(const cgfdm_desc& dd, const float* coord, const float* medium) override {
    d = dd;
    if (const char* e = std::getenv("CGFDM_INTERIOR_RTY")) {
      const int v = std::atoi(e);
      if (v >= 1 && v <= 64) rty = v;
    }
    if (const char* e = std::getenv("CGFDM_INTERIOR_RTX")) {
      const int v = std::atoi(e);
      if (v >= 0 && v <= 64) rtx = v;
    }
    if (const char* e = std::getenv("CGFDM_INTERIOR_KC")) {
      const int v = std::atoi(e);
      if (v >= -1 && v <= 4096) kc = v;
    }
    if (const char* e = std::getenv("CGFDM_INTERIOR_ZTY")) zty = std::atoi(e) == 4 ? 4 : 8;
    if (const char* e = std::getenv("CGFDM_ADJOINT_PIPE")) adj_pipe = std::atoi(e) != 0 ? 1 : 0;
    if (const char* e = std::getenv("CGFDM_SURFACE_STREAM")) surf_overlap = std::atoi(e) != 0 ? 1 : 0;
    if (const char* e = std::getenv("CGFDM_INTERIOR_ZREG")) {
      const int v = std::atoi(e);
      if (v >= 0 && v <= 4096) zreg = v;
    }
    if (const char* e = std::getenv("CGFDM_ROW_ALIGN")) galign = std::atoi(e) != 0;
    // the x-paired kernel pairs cells (i, i + 1) with even i on 8-byte
    // boundaries, which the tight layout gives and the row-aligned one does not
    if (kc == -1) galign = false;
    // setup runs on the tight layout; its products are repacked into the
    // row-aligned layout below
    tNX = d.nx + 2 * cg::H;
    NX = tNX;
    NY = d.ny + 2 * cg::H;
    NZ = d.nz + 2 * cg::H;
    S = (long long)NX * NY;
    N = S * NZ;
    tS = S;
    tN = N;
    int rc;
    // --- temporaries: coordinates, medium, FP64 metrics
    float *dcoord = nullptr, *dmed = nullptr;
    double* met = nullptr;
    if ((rc = dalloc(&dcoord, 3 * N))) return rc;
    if ((rc = dalloc(&dmed, 3 * N))) return rc;
    if ((rc = dalloc(&met, 10 * N))) return rc;
    HIPCHK(hipMemcpyAsync(dcoord, coord, sizeof(float) * 3 * N, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dmed, medium, sizeof(float) * 3 * N, hipMemcpyHostToDevice, stream));
    cg::launch_metrics(dcoord, met, d.nx, d.ny, d.nz, NX, NY, N, d.dh, d.free_surface != 0, stream);
    HIPCHK(hipGetLastError());

    // --- CFL ranges (calc_CFL.cpp:255-349)
    {
      if ((rc = grid_ranges(dcoord, dmed, d.nx, d.ny, d.nz, NX, NY, N))) return rc;
      // ranges = [hmin, hmax, vsmin, vsmax, vpmin, vpmax, rhomin, rhomax]
      double dtmax = 1.34 * ranges[0] / ranges[5];
      if (d.dt > dtmax)
        return fail(CGFDM_ERR_CFL, "DT " + std::to_string(d.dt) + " exceeds the CFL limit " + std::to_string(dtmax));
    }
    if (d.coef_storage != 0) Z = std::sqrt(ranges[6] * ranges[7]) * std::sqrt(ranges[4] * ranges[5]);

    // --- packed coefficients (tight), then the row-aligned copies
    TM* coef_t = nullptr;
    unsigned* words_t = nullptr;
    R *jac_t = nullptr, *rdzdx_t = nullptr, *rdzdy_t = nullptr;
    if ((rc = dalloc(&jac_t, N))) return rc;
    if constexpr (kPacked16) {
      if ((rc = dalloc(&aff, 2LL * cg::M_NCOEF * NZ))) return rc;
      if ((rc = dalloc(&words_t, (long long)(cg::M_NCOEF / 2) * N))) return rc;
      if ((rc = dalloc(&mu_src, N))) return rc;
      cg::launch_pack16<TM>(met, dmed, aff, words_t, (float*)jac_t, mu_src, NX, NY, NZ, Z, stream);
    } else {
      if ((rc = dalloc(&coef_t, (long long)cg::M_NCOEF * N))) return rc;
      cg::launch_pack<R, TM>(met, dmed, coef_t, jac_t, N, Z, stream);
    }
    if ((rc = dalloc(&rdzdx_t, 9 * S))) return rc;
    if ((rc = dalloc(&rdzdy_t, 9 * S))) return rc;
    if (d.free_surface) cg::launch_rdz<R>(met, dmed, rdzdx_t, rdzdy_t, d.nx, d.ny, d.nz, NX, NY, N, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    dfree(met);
    dfree(dmed);
    dfree(dcoord);
    NX = galign ? (tNX + 31) / 32 * 32 : tNX;
    S = (long long)NX * NY;
    N = S * NZ;
    // the stage kernels address grid arrays with 32-bit byte offsets per plane
    if ((double)N * (double)sizeof(R) >= 4294967296.0)
      return fail(CGFDM_ERR_ARG, "grid too large: one field plane array must stay below 4 GiB (" +
                                     std::to_string(N) + " padded cells)");
    pNX = NX;
    pNY = NY;
    pNZ = NZ;
    pXe = tNX;
    if constexpr (kPacked16) {
      unsigned* words = nullptr;
      if ((rc = galloc(&words, (long long)(cg::M_NCOEF / 2) * N))) return rc;
      if ((rc = repack(words, words_t, (long long)(cg::M_NCOEF / 2) * NY * NZ, true))) return rc;
      coef = reinterpret_cast<TM*>(words);
    } else {
      if ((rc = galloc(&coef, (long long)cg::M_NCOEF * N))) return rc;
      if ((rc = repack(coef, coef_t, (long long)cg::M_NCOEF * NY * NZ, true))) return rc;
    }
    if ((rc = galloc(&jac, N))) return rc;
    if ((rc = galloc(&rdzdx, 9 * S))) return rc;
    if ((rc = galloc(&rdzdy, 9 * S))) return rc;
    if ((rc = repack(jac, jac_t, (long long)NY * NZ, true))) return rc;
    if ((rc = repack(rdzdx, rdzdx_t, 9LL * NY, true))) return rc;
    if ((rc = repack(rdzdy, rdzdy_t, 9LL * NY, true))) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    if (coef_t) dfree(coef_t);
    if (words_t) dfree(words_t);
    dfree(jac_t);
    dfree(rdzdx_t);
    dfree(rdzdy_t);

    // --- PML profiles (init_pml_para.cpp:97-239), host FP64, then upload
    {
      int nn[3] = {tNX, NY, NZ};
      std::vector<R> hp(3 * (size_t)(tNX + NY + NZ));
      double L = d.dh * (d.npml - 1);
      double Rc = d.npml > 0 ? std::pow(10.0, -(std::log10((double)d.npml) - 1.0) / std::log10(2.0) - 3.0) : 1.0;
      double d0 = d.npml > 1 ? -(2.0 + 1.0) * 6000.0 / (2.0 * L) * std::log(Rc) : 0.0;
      double a0 = REF_PI * 0.5, b0 = 3.0;
      size_t off = 0;
      for (int ax = 0; ax < 3; ++ax) {
        int n = nn[ax];
        R* pa = hp.data() + off;
        R* pb = pa + n;
        R* pd = pb + n;
        for (int q = 0; q < n; ++q) {
          pa[q] = R(0);
          pb[q] = R(1);
          pd[q] = R(0);
        }
        int lo = cg::H, hi = n - cg::H;
        for (int q = lo; q < hi; ++q) {
          double l = -1.0;
          if (d.npml > 0) {
            if (q < cg::H + d.npml) l = (cg::H + d.npml - 1 - q) * d.dh;
            if (q >= hi - d.npml && !(ax == 2 && d.free_surface)) l = (q - (hi - d.npml)) * d.dh;
          }
          double al = 0, be = 1, dd = 0;
          if (l >= 0 && L > 0) {
            al = a0 * (1.0 - l / L);
            be = 1.0 + (b0 - 1.0) * (l / L) * (l / L);
            dd = d0 * (l / L) * (l / L);
          } else if (l >= 0) {
            al = a0;
          }
          pa[q] = R(al);
          pd[q] = R(dd / be);
          pb[q] = R(1.0 / be);
        }
        off += 3 * (size_t)n;
      }
      if ((rc = dalloc(&pml_par, (long long)hp.size()))) return rc;
      HIPCHK(hipMemcpyAsync(pml_par, hp.data(), sizeof(R) * hp.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));  // hp is a host temporary
      off = 0;
      for (int ax = 0; ax < 3; ++ax) {
        pml.alpha[ax] = pml_par + off;
        pml.beta[ax] = pml_par + off + nn[ax];
        pml.d[ax] = pml_par + off + 2 * nn[ax];
        off += 3 * (size_t)nn[ax];
      }
      pml.npml = d.npml;
      pml.has_z2 = (d.npml > 0 && !d.free_surface) ? 1 : 0;
      long long cnt[6] = {(long long)d.npml * d.ny * d.nz, (long long)d.nx * d.npml * d.nz,
                          (long long)d.nx * d.ny * d.npml, (long long)d.npml * d.ny * d.nz,
                          (long long)d.nx * d.npml * d.nz, (long long)d.nx * d.ny * d.npml};
      for (int f = 0; f < 6; ++f) {
        pml.f[f] = cg::FaceBuf<R>{nullptr, nullptr, nullptr, 0};
        if (d.npml <= 0 || (f == 5 && !pml.has_z2)) continue;
        R* base = nullptr;
        if ((rc = dalloc(&base, 27 * cnt[f]))) return rc;
        pml.f[f] = cg::FaceBuf<R>{base, base + 9 * cnt[f], base + 18 * cnt[f], cnt[f]};
      }
    }
    // --- wavefield buffers and observers
    if ((rc = galloc(&A, 9 * N))) return rc;
    if ((rc = galloc(&B, 9 * N))) return rc;
    if ((rc = galloc(&C, 9 * N))) return rc;
    if ((rc = galloc(&T, 9 * N))) return rc;
    if ((rc = dalloc(&pgv, 2LL * d.nx * d.ny))) return rc;
    if ((rc = dalloc(&dis, 3 * S))) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    return CGFDM_OK;
  }

  // coefficient c of device cell `cell` as the kernels see it (16-bit storage
  // decoded with its layer's affine pair)
  int coef_host(int c, long long cell, double& out) {
    if constexpr (kPacked16) {
      unsigned w = 0;
      float A[2] = {0.f, 0.f};
      const long long k = cell / S;
      HIPCHK(hipMemcpy(&w, reinterpret_cast<const unsigned*>(coef) + (long long)(c >> 1) * N + cell, sizeof(w),
                       hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(A, aff + 2 * ((long long)c * NZ + k), sizeof(A), hipMemcpyDeviceToHost));
      const unsigned short b = (c & 1) ? (unsigned short)(w >> 16) : (unsigned short)(w & 0xffffu);
      out = (double)(A[0] + A[1] * bits16_to_float<TM>(b));
    } else {
      TM v;
      HIPCHK(hipMemcpy(&v, coef + (long long)c * N + cell, sizeof(TM), hipMemcpyDeviceToHost));
      out = (double)v;
    }
    return CGFDM_OK;
  }

  int set_source(int np, int nt, double dt, const long long* cell, const float* rate) override {
    if (np < 0 || (np > 0 && (nt < 2 || dt <= 0 || !cell || !rate))) return fail(CGFDM_ERR_SOURCE, "bad source table");
    std::map<long long, std::vector<double>> acc;
    for (int p = 0; p < np; ++p) {
      const long long c = dev_cell(cell[p]);
      if (c < 0) return fail(CGFDM_ERR_SOURCE, "source cell outside the interior");
      auto& v = acc[c];
      if (v.empty()) v.assign((size_t)nt * 6, 0.0);
      const float* r = rate + (size_t)p * nt * 6;
      for (size_t q = 0; q < (size_t)nt * 6; ++q) v[q] += r[q];
    }
    int rc;
    if (src_cell) dfree(src_cell);
    if (src_rate) dfree(src_rate);
    src_cell = nullptr;
    src_rate = nullptr;
    src_np = (int)acc.size();
    src_nt = nt;
    src_dt = dt;
    if (src_np == 0) return CGFDM_OK;
    std::vector<long long> hc;
    std::vector<R> hr((size_t)src_np * nt * 6);
    src_raw.assign((size_t)src_np * nt * 6, 0.0);
    src_muv.assign(src_np, 0.0);
    src_vol.assign(src_np, 0.0);
    std::map<long long, int> merged;
    int Jp = 0;
    for (auto& kv : acc) {
      long long c = kv.first;
      merged[c] = Jp;
      hc.push_back(c);
      R jv;
      HIPCHK(hipMemcpy(&jv, jac + c, sizeof(R), hipMemcpyDeviceToHost));
      double mu = 0.0;
      if constexpr (kPacked16) {
        // mu_src keeps the tight layout of the setup stage
        const long long ct = c % NX + (c / NX % NY) * (long long)tNX + c / S * tS;
        float mv = 0.f;
        HIPCHK(hipMemcpy(&mv, mu_src + ct, sizeof(float), hipMemcpyDeviceToHost));
        mu = mv;
      } else if (int rc2 = coef_host(cg::M_MU, c, mu)) {
        return rc2;
      }
      double J = (double)jv;
      double sc = mu / (J * d.dh * d.dh * d.dh);  // mu/(J DH^3), stress in units of Z
      src_muv[Jp] = mu;
      src_vol[Jp] = 1.0 / (J * d.dh * d.dh * d.dh);
      for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 6; ++q) {
          hr[((size_t)t * src_np + Jp) * 6 + q] = R(kv.second[(size_t)t * 6 + q] * sc);
          src_raw[((size_t)t * src_np + Jp) * 6 + q] = kv.second[(size_t)t * 6 + q];
        }
      ++Jp;
    }
    src_of_point.resize(np);
    for (int p = 0; p < np; ++p) src_of_point[p] = merged[dev_cell(cell[p])];
    if ((rc = dalloc(&src_cell, src_np))) return rc;
    if ((rc = dalloc(&src_rate, (long long)hr.size()))) return rc;
    HIPCHK(hipMemcpy(src_cell, hc.data(), sizeof(long long) * src_np, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(src_rate, hr.data(), sizeof(R) * hr.size(), hipMemcpyHostToDevice));
    return CGFDM_OK;
  }

  int set_stations(int n, const int* ijk) override {
    if (n < 0 || (n > 0 && !ijk)) return fail(CGFDM_ERR_ARG, "bad station list");
    std::vector<long long> hc(n);
    for (int s = 0; s < n; ++s) {
      int i = ijk[3 * s], j = ijk[3 * s + 1], k = ijk[3 * s + 2];
      if (i < cg::H || i >= pXe - cg::H || j < cg::H || j >= NY - cg::H || k < cg::H || k >= NZ - cg::H)
        return fail(CGFDM_ERR_ARG, "station outside the interior");
      hc[s] = (long long)i + (long long)NX * j + (long long)k * S;
    }
    int rc;
    if (st_cell) dfree(st_cell);
    if (st_rec) dfree(st_rec);
    st_cell = nullptr;
    st_rec = nullptr;
    ns = n;
    if (n == 0) return CGFDM_OK;
    if ((rc = dalloc(&st_cell, n))) return rc;
    if ((rc = dalloc(&st_rec, (long long)n * d.nt * 9))) return rc;
    HIPCHK(hipMemcpy(st_cell, hc.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    return CGFDM_OK;
  }

  int step(int n) override { return step_impl(n, true); }

  const char* kernel_name(int which) const override {
    if (which == 1) return "cg_surface";
    if (which != 0) return nullptr;
    if (std::is_same<R, float>::value && std::is_same<TM, float>::value && kc < 0 && !galign) return "cg_pair";
    if (std::is_same<R, float>::value && kc == 1) return "cg_plane";
    if (std::is_same<R, float>::value && kc > 1) return "cg_zcol";
    return zreg > 0 ? "cg_zreg" : "cg_cell";
  }

  // observe = false: no station record, PGV or displacement update (used to
  // recompute states for the adjoint)
  int step_impl(int n, bool observe) {
    if (!kCgfdmStepEnabled) return fail(CGFDM_ERR_DISABLED, kCgfdmStepOff);
    const int kend = cg::H + d.nz - (d.free_surface ? 3 : 0);
    const bool fork = d.free_surface && surf_overlap;
    if (fork && !sstream) {
      HIPCHK(hipStreamCreateWithFlags(&sstream, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    for (int s = 0; s < n; ++s) {
      int fb[3] = {FBTAB[it % 8][0], FBTAB[it % 8][1], FBTAB[it % 8][2]};
      size_t sp = 0;
      if (timing_on) {
        sp = next_pair(2);
        HIPCHK(hipEventRecord(ev[sp].a, stream));
      }
      // moment-rate source on W at t = it * DT (addMoment.cpp:262-377), float
      // time arithmetic as in the reference
      if (src_np > 0) {
        float DTf = (float)d.dt;
        float tmpT = ((float)it + 0.0f) * DTf;
        int srcIt = int(tmpT / (float)src_dt);
        if (srcIt + 1 < src_nt) {
          float t1 = float(srcIt) * (float)src_dt;
          float tw = (tmpT - t1) / (float)src_dt;
          cg::launch_source<R>(A, N, src_np, src_cell, src_rate, srcIt, R(tw), R(d.dt), stream);
        }
      }
      for (int irk = 0; irk < 4; ++irk) {
        cg::StageArgs<R> a{};
        a.nx = d.nx;
        a.ny = d.ny;
        a.nz = d.nz;
        a.NX = NX;
        a.NY = NY;
        a.N = N;
        a.S = S;
        a.rowb = (unsigned)(NX * sizeof(R));
        a.planeb = (unsigned)(S * sizeof(R));
        a.kbeg = cg::H;
        a.kend = kend;
        a.stage = irk;
        a.rty = rty;
        a.rtx = rtx;
        a.zreg = zreg;
        a.kc = kc;
        a.zty = zty;
        a.DT = R(d.dt);
        a.rDH = R(1.0 / d.dh);
        a.aff = aff;
        a.nzp = NZ;
        a.win = irk == 0 ? A : irk == 1 ? B : irk == 2 ? C : B;
        a.wout = irk == 0 ? B : irk == 1 ? C : irk == 2 ? B : A;
        a.wm = A;
        a.wt = T;
        a.pml = pml;
        // the surface stage (forked first, so that its few blocks are
        // dispatched before the interior's) and the interior stage
        if (d.free_surface) {
          hipStream_t ss = fork ? sstream : stream;
          if (fork) {
            HIPCHK(hipEventRecord(ev_fork, stream));
            HIPCHK(hipStreamWaitEvent(sstream, ev_fork, 0));
          }
          cg::SurfArgs<R> sa{jac, rdzdx, rdzdy, observe ? pgv : nullptr, dis};
          size_t fp = 0;
          if (timing_on) {
            fp = next_pair(1);
            HIPCHK(hipEventRecord(ev[fp].a, ss));
          }
          LAUNCHCHK((cg::launch_surface<R, TM>(a, coef, sa, fb[0], fb[1], fb[2], ss)), "surface stage");
          if (timing_on) HIPCHK(hipEventRecord(ev[fp].b, ss));
          if (fork) HIPCHK(hipEventRecord(ev_join, sstream));
        }
        size_t ip = 0;
        if (timing_on) {
          ip = next_pair(0);
          HIPCHK(hipEventRecord(ev[ip].a, stream));
        }
        LAUNCHCHK((cg::launch_interior<R, TM>(a, coef, fb[0], fb[1], fb[2], stream)), "interior stage");
        if (timing_on) HIPCHK(hipEventRecord(ev[ip].b, stream));
        if (fork) HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
        fb[0] = -fb[0];
        fb[1] = -fb[1];
        fb[2] = -fb[2];
      }
      if (observe) cg::launch_record<R>(A, N, ns, st_cell, st_rec, d.nt, it, stream);
      if (timing_on) HIPCHK(hipEventRecord(ev[sp].b, stream));
      ++it;
    }
    HIPCHK(hipGetLastError());
    return CGFDM_OK;
  }

  int get_stations(float* out) override {
    if (ns == 0) return CGFDM_OK;
    HIPCHK(hipStreamSynchronize(stream));
    size_t n = (size_t)ns * d.nt * 9;
    HIPCHK(hipMemcpy(out, st_rec, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (Z != 1.0)
      for (size_t q = 0; q < n; ++q)
        if (q % 9 >= 3) out[q] = (float)(out[q] * Z);
    return CGFDM_OK;
  }

  int get_pgv(float* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(out, pgv, sizeof(float) * 2 * (size_t)d.nx * d.ny, hipMemcpyDeviceToHost));
    return CGFDM_OK;
  }

  int get_dis(float* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<R> h(3 * (size_t)S);
    HIPCHK(hipMemcpy(h.data(), dis, sizeof(R) * h.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i)
          out[((size_t)c * d.ny + j) * d.nx + i] = (float)h[(size_t)c * S + (size_t)(j + cg::H) * NX + (i + cg::H)];
    return CGFDM_OK;
  }

  int get_field(int comp, double* out) override {
    if (comp < 0 || comp > 8) return fail(CGFDM_ERR_ARG, "component out of range");
    std::vector<R> h(tN);
    if (int rc = to_host(A + (size_t)comp * N, h.data(), 1)) return rc;
    double sc = comp >= 3 ? Z : 1.0;
    for (long long p = 0; p < tN; ++p) out[p] = (double)h[p] * sc;
    return CGFDM_OK;
  }

  int get_slice(int comp, int axis, int index, float* out) override {
    if (int rc = check_plane(comp, axis, index)) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<R> h;
    if (int rc = copy_plane(A + (size_t)comp * N, axis, index, h)) return rc;
    const double sc = comp >= 3 ? Z : 1.0;
    for (size_t p = 0; p < h.size(); ++p) out[p] = (float)((double)h[p] * sc);
    return CGFDM_OK;
  }

  int set_field(int comp, const double* in) override {
    if (comp < 0 || comp > 8) return fail(CGFDM_ERR_ARG, "component out of range");
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<R> h(tN);
    double sc = comp >= 3 ? 1.0 / Z : 1.0;
    for (long long p = 0; p < tN; ++p) h[p] = R(in[p] * sc);
    return to_dev(A + (size_t)comp * N, h.data(), 1);
  }

  int get_setup(int what, double* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    if (what >= 0 && what <= 12 && what != 9 && kPacked16) {
      const int plane = what < 9 ? what : what - 1;
      std::vector<unsigned> h(tN);
      std::vector<float> ha(2 * (size_t)cg::M_NCOEF * NZ);
      if (int rc = to_host(reinterpret_cast<const unsigned*>(coef) + (size_t)(plane >> 1) * N, h.data(), 1)) return rc;
      HIPCHK(hipMemcpy(ha.data(), aff, sizeof(float) * ha.size(), hipMemcpyDeviceToHost));
      const double sc = what == 10 || what == 11 ? Z : what == 12 ? 1.0 / Z : 1.0;
      for (long long p = 0; p < tN; ++p) {
        const int k = (int)(p / tS);
        const unsigned short b = (plane & 1) ? (unsigned short)(h[p] >> 16) : (unsigned short)(h[p] & 0xffffu);
        const float* A = ha.data() + 2 * ((size_t)plane * NZ + k);
        out[p] = (double)(A[0] + A[1] * bits16_to_float<TM>(b)) * sc;
      }
      return CGFDM_OK;
    }
    if (what >= 0 && what <= 12 && what != 9) {
      int plane = what < 9 ? what : what - 1;  // 10,11,12 -> M_MU.. (9,10,11)
      std::vector<TM> h(tN);
      if (int rc = to_host(coef + (size_t)plane * N, h.data(), 1)) return rc;
      double sc = what == 10 || what == 11 ? Z : what == 12 ? 1.0 / Z : 1.0;
      for (long long p = 0; p < tN; ++p) {
        double v = sizeof(TM) == 8 ? (double)*(const double*)(const void*)&h[p] : (double)to_float(h[p]);
        out[p] = v * sc;
      }
      return CGFDM_OK;
    }
    if (what == 9) {
      std::vector<R> h(tN);
      if (int rc = to_host(jac, h.data(), 1)) return rc;
      for (long long p = 0; p < tN; ++p) out[p] = (double)h[p];
      return CGFDM_OK;
    }
    if ((what >= 20 && what < 29) || (what >= 30 && what < 39)) {
      const R* src = (what < 30 ? rdzdx : rdzdy) + (size_t)(what % 10) * S;
      std::vector<R> h(tS);
      HIPCHK(hipMemcpy2D(h.data(), sizeof(R) * tNX, src, sizeof(R) * NX, sizeof(R) * tNX, NY, hipMemcpyDeviceToHost));
      for (long long p = 0; p < tS; ++p) out[p] = (double)h[p];
      return CGFDM_OK;
    }
    if (what >= 40 && what < 49) {
      int ax = (what - 40) / 3, kind = (what - 40) % 3;
      int n = ax == 0 ? tNX : ax == 1 ? NY : NZ;
      const R* src = kind == 0 ? pml.alpha[ax] : kind == 1 ? pml.beta[ax] : pml.d[ax];
      std::vector<R> h(n);
      HIPCHK(hipMemcpy(h.data(), src, sizeof(R) * n, hipMemcpyDeviceToHost));
      for (int p = 0; p < n; ++p) out[p] = (double)h[p];
      return CGFDM_OK;
    }
    return fail(CGFDM_ERR_ARG, "unknown setup product");
  }

  // ---- adjoint ------------------------------------------------------------
  long long state_size() const {
    long long n = 9 * N;
    for (int f = 0; f < 6; ++f) n += 9 * pml.f[f].cnt;
    return n;
  }
  int copy_state(R* dst, bool save) {
    long long off = 0;
    if (save) HIPCHK(hipMemcpyAsync(dst, A, sizeof(R) * 9 * N, hipMemcpyDeviceToDevice, stream));
    else HIPCHK(hipMemcpyAsync(A, dst, sizeof(R) * 9 * N, hipMemcpyDeviceToDevice, stream));
    off = 9 * N;
    for (int f = 0; f < 6; ++f) {
      const long long c = 9 * pml.f[f].cnt;
      if (c == 0) continue;
      if (save) HIPCHK(hipMemcpyAsync(dst + off, pml.f[f].a, sizeof(R) * c, hipMemcpyDeviceToDevice, stream));
      else HIPCHK(hipMemcpyAsync(pml.f[f].a, dst + off, sizeof(R) * c, hipMemcpyDeviceToDevice, stream));
      off += c;
    }
    return CGFDM_OK;
  }
  // (re)allocates *p for at least `count` elements; *cap is its capacity
  template <class X>
  int dgrow(X** p, long long* cap, long long count) {
    if (*p && *cap >= count) return CGFDM_OK;
    if (*p) dfree(*p);
    *p = nullptr;
    *cap = 0;
    if (int rc = dalloc(p, count)) return rc;
    *cap = count;
    return CGFDM_OK;
  }
  long long ck_cap = 0, rec_cap = 0, seg_cap = 0, src_cap = 0;
  // the adjoint's work buffers, allocated by the first adjoint and kept with
  // the solver, as the checkpoints and the stage store are (reused, and grown
  // when a later run needs more): a hipMalloc of a few GB took over a second on
  // the MI355X where another process had freed the memory shortly before
  int alloc_adjoint_work() {
    if (adj_buf) return CGFDM_OK;
    long long fcnt = 0;
    for (int f = 0; f < 6; ++f) fcnt += 9 * pml.f[f].cnt;
    int rc;
    if ((rc = dalloc(&adj_buf, 99 * N)) || (rc = dalloc(&adj_face, std::max(4 * fcnt, 1LL))) ||
        (rc = dalloc(&adj_grad, 3 * N)) || (rc = dalloc(&adj_gsrc, std::max((long long)src_np * src_nt * 6, 1LL))) ||
        (rc = dalloc(&adj_end, state_size()))) {
      // all or none, so that the next call allocates them again
      for (void* q : {(void*)adj_buf, (void*)adj_face, (void*)adj_grad, (void*)adj_gsrc, (void*)adj_end})
        if (q) dfree(q);
      adj_buf = adj_face = adj_end = nullptr;
      adj_grad = adj_gsrc = nullptr;
      return rc;
    }
    return CGFDM_OK;
  }
  // bytes of the work buffers above still to be allocated
  double adjoint_work_bytes() const {
    if (adj_buf) return 0.0;
    long long fcnt = 0;
    for (int f = 0; f < 6; ++f) fcnt += 9 * pml.f[f].cnt;
    return (double)sizeof(R) * (99.0 * N + 4.0 * std::max(fcnt, 1LL) + (double)state_size()) +
           8.0 * (3.0 * N + 6.0 * std::max((long long)src_np * src_nt, 1LL));
  }
  // allocate the stage store of a ck-step segment if it fits the budget:
  // CGFDM_ADJOINT_STORE_GB, else 85 % of the free device memory less `reserve`
  // bytes still to be allocated (an MI355X's 288 GB hold the stage states of a
  // 100-step segment of a 256 x 256 x 128 grid). The store is optional: on a
  // failed allocation the runtime's sticky error is cleared so that the launch
  // checks of the recompute path do not report it.
  void alloc_stage_store(int ck, double reserve) {
    const double need = (double)sizeof(R) * 36.0 * (double)N * ck;
    const char* env = std::getenv("CGFDM_ADJOINT_STORE_GB");
    const double env_budget = env ? std::atof(env) * 1073741824.0 : 0.0;
    // a store of this size kept from an earlier forward_record
    if (adj_stages && stages_ck == ck && (!env || need <= env_budget)) return;
    stages_seg = -1;
    if (adj_stages) {
      dfree(adj_stages);
      adj_stages = nullptr;
    }
    double budget = env_budget;
    if (!env) {
      size_t fre = 0, tot = 0;
      if (hipMemGetInfo(&fre, &tot) == hipSuccess) budget = 0.85 * ((double)fre - reserve);
      (void)hipGetLastError();
    }
    if (need <= budget && dalloc(&adj_stages, 36LL * N * ck) != CGFDM_OK) {
      adj_stages = nullptr;
      (void)hipGetLastError();
    }
    stages_ck = ck;
  }
  // source table index and weight at step index `step` (addMoment.cpp:262-377), -1 if none
  int source_index(int step, float& tw) const {
    if (src_np <= 0) return -1;
    const float DTf = (float)d.dt;
    const float tmpT = ((float)step + 0.0f) * DTf;
    const int srcIt = int(tmpT / (float)src_dt);
    if (srcIt + 1 >= src_nt) return -1;
    const float t1 = float(srcIt) * (float)src_dt;
    tw = (tmpT - t1) / (float)src_dt;
    return srcIt;
  }
  // one forward step (no observers) that leaves its stage states in x[0..3]:
  // x0 = W after the source, x1..x3 = inputs of stages 1..3
  // One forward step whose stage states stay in the 36N block xs: x0 = the
  // state after the source (the step runs on it in place of A), x1..x3 = the
  // inputs of stages 1..3. The step's result goes to `next` (the x0 of the next
  // step's block) or, when next is null, to A; `first` copies A into x0 first.
  // A chain of steps thus moves the state from block to block without copies.
  int step_store(R* xs, R* next, bool first, bool observe = false) {
    R* x0 = xs;
    R* x1 = xs + 9 * N;
    R* x2 = xs + 18 * N;
    R* x3 = xs + 27 * N;
    const int kend = cg::H + d.nz - (d.free_surface ? 3 : 0);
    int fb[3] = {FBTAB[it % 8][0], FBTAB[it % 8][1], FBTAB[it % 8][2]};
    float tw = 0.f;
    const int srcIt = source_index(it, tw);
    if (first) HIPCHK(hipMemcpyAsync(x0, A, sizeof(R) * 9 * N, hipMemcpyDeviceToDevice, stream));
    if (srcIt >= 0) cg::launch_source<R>(x0, N, src_np, src_cell, src_rate, srcIt, R(tw), R(d.dt), stream);
    R* ins[4] = {x0, x1, x2, x3};
    R* outs[4] = {x1, x2, x3, next ? next : A};
    for (int irk = 0; irk < 4; ++irk) {
      cg::StageArgs<R> a{};
      a.nx = d.nx; a.ny = d.ny; a.nz = d.nz; a.NX = NX; a.NY = NY; a.N = N; a.S = S;
      a.rowb = (unsigned)(NX * sizeof(R)); a.planeb = (unsigned)(S * sizeof(R));
      a.kbeg = cg::H; a.kend = kend; a.stage = irk; a.rty = rty; a.rtx = rtx; a.zreg = zreg; a.kc = kc; a.zty = zty;
      a.DT = R(d.dt); a.rDH = R(1.0 / d.dh); a.aff = aff; a.nzp = NZ;
      a.win = ins[irk]; a.wout = outs[irk]; a.wm = x0; a.wt = T; a.pml = pml;
      LAUNCHCHK((cg::launch_interior<R, TM>(a, coef, fb[0], fb[1], fb[2], stream)), "interior stage");
      if (d.free_surface) {
        cg::SurfArgs<R> sa{jac, rdzdx, rdzdy, observe ? pgv : nullptr, dis};
        LAUNCHCHK((cg::launch_surface<R, TM>(a, coef, sa, fb[0], fb[1], fb[2], stream)), "surface stage");
      }
      fb[0] = -fb[0]; fb[1] = -fb[1]; fb[2] = -fb[2];
    }
    ++it;
    return CGFDM_OK;
  }
  // source on A at step index `step` (addMoment.cpp:262-377); returns srcIt (-1 if none)
  int inject_source(int step, float& tw) {
    if (src_np <= 0) return -1;
    float DTf = (float)d.dt;
    float tmpT = ((float)step + 0.0f) * DTf;
    int srcIt = int(tmpT / (float)src_dt);
    if (srcIt + 1 >= src_nt) return -1;
    float t1 = float(srcIt) * (float)src_dt;
    tw = (tmpT - t1) / (float)src_dt;
    cg::launch_source<R>(A, N, src_np, src_cell, src_rate, srcIt, R(tw), R(d.dt), stream);
    return srcIt;
  }
  // RK stage states of the step at index `step` starting from A (after the
  // source): X1 -> B, X2 -> C, X3 -> x3. Auxiliary states evolve in the face
  // stage buffers; the face state buffers and A are left unchanged.
  int stage_states(int step, R* x3) {
    const int kend = cg::H + d.nz - (d.free_surface ? 3 : 0);
    int fb[3] = {FBTAB[step % 8][0], FBTAB[step % 8][1], FBTAB[step % 8][2]};
    R* ins[3] = {A, B, C};
    R* outs[3] = {B, C, x3};
    for (int irk = 0; irk < 3; ++irk) {
      cg::StageArgs<R> a{};
      a.nx = d.nx; a.ny = d.ny; a.nz = d.nz; a.NX = NX; a.NY = NY; a.N = N; a.S = S;
      a.rowb = (unsigned)(NX * sizeof(R)); a.planeb = (unsigned)(S * sizeof(R));
      a.kbeg = cg::H; a.kend = kend; a.stage = irk; a.rty = rty; a.rtx = rtx; a.zreg = zreg; a.kc = kc; a.zty = zty;
      a.DT = R(d.dt); a.rDH = R(1.0 / d.dh); a.aff = aff; a.nzp = NZ;
      a.win = ins[irk]; a.wout = outs[irk]; a.wm = A; a.wt = T; a.pml = pml;
      LAUNCHCHK((cg::launch_interior<R, TM>(a, coef, fb[0], fb[1], fb[2], stream)), "interior stage");
      if (d.free_surface) {
        cg::SurfArgs<R> sa{jac, rdzdx, rdzdy, nullptr, dis};
        LAUNCHCHK((cg::launch_surface<R, TM>(a, coef, sa, fb[0], fb[1], fb[2], stream)), "surface stage");
      }
      fb[0] = -fb[0]; fb[1] = -fb[1]; fb[2] = -fb[2];
    }
    return CGFDM_OK;
  }

  int forward_record(int nsteps, int ck, double* rec) override {
    if (!kCgfdmAdjointEnabled) return fail(CGFDM_ERR_DISABLED, kCgfdmAdjointOff);
    if (nsteps < 1 || ck < 1) return fail(CGFDM_ERR_ARG, "nsteps and checkpoint_every must be positive");
    if (d.coef_storage != 0) return fail(CGFDM_ERR_ARG, "the adjoint needs native coefficient storage");
    fr_n = 0;
    stages_seg = -1;
    int rc;
    const long long ss = state_size();
    const int nck = (nsteps + ck - 1) / ck;
    if ((rc = dgrow(&ck_store, &ck_cap, ss * nck))) return rc;
    if ((rc = dgrow(&rec_vel, &rec_cap, (long long)std::max(ns, 1) * nsteps * 3))) return rc;
    fr_it0 = it;
    fr_ck = ck;
    // the last segment runs into the stage store when it fits, so that the
    // adjoint starts from its stage states instead of recomputing them
    alloc_stage_store(ck, adjoint_work_bytes());
    const int last0 = (nck - 1) * ck;
    for (int n = 0; n < nsteps; ++n) {
      if (n % ck == 0 && (rc = copy_state(ck_store + (long long)(n / ck) * ss, true))) return rc;
      const R* now = A;  // the state after this step
      if (adj_stages && n >= last0) {
        R* xs = adj_stages + 36LL * N * (n - last0);
        R* next = n + 1 < nsteps ? xs + 36LL * N : nullptr;
        if ((rc = step_store(xs, next, n == last0, true))) return rc;
        if (next) now = next;
      } else if ((rc = step_impl(1, true))) {
        return rc;
      }
      cg::launch_record_vel<R>(now, N, ns, st_cell, rec_vel, nsteps, n, stream);
    }
    if (adj_stages) stages_seg = nck - 1;
    HIPCHK(hipGetLastError());
    fr_n = nsteps;
    if (rec && ns > 0) {
      std::vector<R> h((size_t)ns * nsteps * 3);
      HIPCHK(hipMemcpyAsync(h.data(), rec_vel, sizeof(R) * h.size(), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (size_t q = 0; q < h.size(); ++q) rec[q] = (double)h[q];
    }
    return CGFDM_OK;
  }

  int adjoint(const double* src, double* gmu, double* glam, double* gb, double* grate) override {
    if (!kCgfdmAdjointEnabled) return fail(CGFDM_ERR_DISABLED, kCgfdmAdjointOff);
    if (fr_n <= 0) return fail(CGFDM_ERR_ARG, "no forward_record run to take the adjoint of");
    if (ns > 0 && !src) return fail(CGFDM_ERR_ARG, "null adjoint source");
    int rc;
    const long long ss = state_size();
    long long fcnt = 0;
    for (int f = 0; f < 6; ++f) fcnt += 9 * pml.f[f].cnt;
    if ((rc = alloc_adjoint_work())) return rc;
    if ((rc = dgrow(&adj_src, &src_cap, (long long)std::max(ns, 1) * fr_n * 3))) return rc;
    R* MU = adj_buf;
    R* OUT = adj_buf + 9 * N;
    R* G = adj_buf + 18 * N;
    R* Z = adj_buf + 27 * N;
    R* Ls = adj_buf + 54 * N;
    R* X3 = adj_buf + 63 * N;
    R* Z2 = adj_buf + 72 * N;  // the fused pass B + A writes the next stage's products here
    HIPCHK(hipMemsetAsync(adj_buf, 0, sizeof(R) * 99 * N, stream));
    HIPCHK(hipMemsetAsync(adj_face, 0, sizeof(R) * std::max(4 * fcnt, 1LL), stream));
    HIPCHK(hipMemsetAsync(adj_grad, 0, sizeof(double) * 3 * N, stream));
    HIPCHK(hipMemsetAsync(adj_gsrc, 0, sizeof(double) * std::max((long long)src_np * src_nt * 6, 1LL), stream));
    if (ns > 0) {
      std::vector<R> h((size_t)ns * fr_n * 3);
      for (size_t q = 0; q < h.size(); ++q) h[q] = R(src[q]);
      HIPCHK(hipMemcpyAsync(adj_src, h.data(), sizeof(R) * h.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    cg::FaceAdj<R> fa[6];
    {
      long long off = 0;
      for (int f = 0; f < 6; ++f) {
        const long long c = 9 * pml.f[f].cnt;
        fa[f] = cg::FaceAdj<R>{adj_face + off, adj_face + off + c, adj_face + off + 2 * c, pml.f[f].cnt,
                               adj_face + off + 3 * c};
        off += 4 * c;
      }
    }
    alloc_stage_store(fr_ck, 0.0);
    // without the stage store the segment's step states are kept instead
    if (!adj_stages && (rc = dgrow(&adj_seg, &seg_cap, ss * fr_ck))) return rc;
    if ((rc = copy_state(adj_end, true))) return rc;
    const int nck = (fr_n + fr_ck - 1) / fr_ck;
    for (int seg = nck - 1; seg >= 0; --seg) {
      const int n0 = seg * fr_ck, n1 = std::min(n0 + fr_ck, fr_n);
      if (!(adj_stages && stages_seg == seg)) {
        if ((rc = copy_state(ck_store + (long long)seg * ss, false))) return rc;
        it = fr_it0 + n0;
        stages_seg = -1;
        for (int n = n0; n < n1; ++n) {
          if (adj_stages) {
            R* xs = adj_stages + 36LL * N * (n - n0);
            if ((rc = step_store(xs, n + 1 < n1 ? xs + 36LL * N : nullptr, n == n0))) return rc;
          } else {
            if ((rc = copy_state(adj_seg + (long long)(n - n0) * ss, true))) return rc;
            if ((rc = step_impl(1, false))) return rc;
          }
        }
        if (adj_stages) stages_seg = seg;
      }
      for (int n = n1 - 1; n >= n0; --n) {
        cg::launch_adj_inject<R>(MU, N, ns, st_cell, adj_src, fr_n, n, stream);
        const int step = fr_it0 + n;
        float tw = 0.f;
        int srcIt;
        R* X[4];
        if (adj_stages) {
          R* xs = adj_stages + 36LL * N * (n - n0);
          for (int q = 0; q < 4; ++q) X[q] = xs + 9LL * N * q;
          srcIt = source_index(step, tw);
        } else {
          if ((rc = copy_state(adj_seg + (long long)(n - n0) * ss, false))) return rc;
          srcIt = inject_source(step, tw);
          if ((rc = stage_states(step, X3))) return rc;
          X[0] = A; X[1] = B; X[2] = C; X[3] = X3;
        }
        const int* fb0 = FBTAB[step % 8];
        // one launch of the transposed stage passes: stage s, F/B pattern of
        // stage s, stage state x, products read (zr) and written (zw)
        auto passes = [&](int mode, int s, const R* x, R* zr, R* zw) -> int {
          const int sg = (s % 2 == 0) ? 1 : -1;
          cg::AdjArgs<R> a{};
          a.nx = d.nx; a.ny = d.ny; a.nz = d.nz; a.NX = NX; a.NY = NY; a.N = N; a.S = S;
          a.rtx = rtx; a.rty = rty;
          a.fs = d.free_surface ? 1 : 0;
          a.stage = s;
          a.DT = R(d.dt);
          a.rDH = R(1.0 / d.dh);
          a.X = x;
          a.mu = MU; a.out = OUT; a.G = G; a.Z = zr; a.Zn = zw; a.Ls = Ls;
          a.gmu = adj_grad; a.glam = adj_grad + N; a.gb = adj_grad + 2 * N;
          a.pml = pml;
          for (int f = 0; f < 6; ++f) a.fa[f] = fa[f];
          if constexpr (std::is_same<R, TM>::value)
            LAUNCHCHK((cg::launch_adj_stage<R, TM>(a, coef, jac, rdzdx, rdzdy, sg * fb0[0], sg * fb0[1], sg * fb0[2],
                                                   mode, stream)), "adjoint stage");
          return CGFDM_OK;
        };
        if (adj_pipe) {
          // surface + A of stage 3, then B of stage s with A of stage s - 1
          // (the surface pass of stage s - 1 after it), then B of stage 0
          R* zc = Z;
          R* zo = Z2;
          if ((rc = passes(cg::ADJ_SA, 3, X[3], zc, zo))) return rc;
          for (int s = 3; s >= 1; --s) {
            if ((rc = passes(cg::ADJ_BA, s, X[s - 1], zc, zo))) return rc;
            std::swap(zc, zo);
            if (d.free_surface && (rc = passes(cg::ADJ_S, s - 1, X[s - 1], zc, zo))) return rc;
            if (s - 1 > 0)
              for (int f = 0; f < 6; ++f) std::swap(fa[f].g, fa[f].gn);
          }
          if ((rc = passes(cg::ADJ_B, 0, X[0], zc, zo))) return rc;
        } else {
          for (int s = 3; s >= 0; --s)
            if ((rc = passes(cg::ADJ_FULL, s, X[s], Z, Z2))) return rc;
        }
        if (srcIt >= 0)
          cg::launch_adj_source<R>(OUT, N, src_np, src_cell, adj_gsrc, srcIt, (double)tw, (double)R(d.dt), stream);
        std::swap(MU, OUT);
        for (int f = 0; f < 6; ++f) std::swap(fa[f].mu, fa[f].out);
      }
    }
    HIPCHK(hipGetLastError());
    if ((rc = copy_state(adj_end, false))) return rc;
    it = fr_it0 + fr_n;
    // gradients to the tight layout of the C ABI on the device (strided
    // device copies into the work buffer, no longer needed), then one
    // contiguous copy into each of the caller's arrays
    double* outs[3] = {gmu, glam, gb};
    double* tight = reinterpret_cast<double*>(adj_buf);  // 3 tN doubles fit in its 99N elements of R
    for (int g = 0; g < 3; ++g)
      if (outs[g]) {
        HIPCHK(hipMemcpy2DAsync(tight + (size_t)g * tN, sizeof(double) * tNX, adj_grad + (size_t)g * N,
                                sizeof(double) * NX, sizeof(double) * tNX, (size_t)NY * NZ, hipMemcpyDeviceToDevice,
                                stream));
        HIPCHK(hipMemcpyAsync(outs[g], tight + (size_t)g * tN, sizeof(double) * tN, hipMemcpyDeviceToHost, stream));
      }
    std::vector<double> hs((size_t)src_np * src_nt * 6);
    if (src_np > 0)
      HIPCHK(hipMemcpyAsync(hs.data(), adj_gsrc, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    // scaled rate = raw mu(cell) / (J DH^3): the scaling carries mu of the source
    // cell (device index -> tight index of the same cell)
    std::vector<long long> hcell(src_np);
    if (src_np > 0) HIPCHK(hipMemcpy(hcell.data(), src_cell, sizeof(long long) * src_np, hipMemcpyDeviceToHost));
    if (gmu)
      for (int p = 0; p < src_np; ++p) {
        const long long c = hcell[p], ci = c % NX, cr = c / NX;  // cr = j + NY k
        const long long tc = ci + (long long)tNX * cr;
        for (int t = 0; t < src_nt; ++t)
          for (int q = 0; q < 6; ++q) {
            const size_t e = ((size_t)t * src_np + p) * 6 + q;
            gmu[tc] += hs[e] * src_raw[e] * src_vol[p];
          }
      }
    if (grate)
      for (size_t pi = 0; pi < src_of_point.size(); ++pi) {
        const int p = src_of_point[pi];
        for (int t = 0; t < src_nt; ++t)
          for (int q = 0; q < 6; ++q)
            grate[((size_t)pi * src_nt + t) * 6 + q] = hs[((size_t)t * src_np + p) * 6 + q] * src_muv[p] * src_vol[p];
      }
    return CGFDM_OK;
  }
};

// ME-AWENO (SCFDM) solver: WENO5 alternative-flux finite differences on the
// conservative state, TVD-RK3, characteristic free surface, exponential
// sponge (propagate.cpp:186-321 under SCFDM). The state is kept in three
// buffers that rotate through the stages, so the stage input, the state at
// the step start and the stage output are always distinct arrays.
template <class R>
struct AwenoSolver : SolverBase {
  cgfdm_desc d{};
  // the top-layer update on a second stream beside the bulk update, forked
  // from and joined to the solver's stream (CGFDM_SURFACE_STREAM=0: one stream)
  int surf_overlap = 1;
  hipStream_t sstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~AwenoSolver() override {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (sstream) (void)hipStreamDestroy(sstream);
  }
  // device layout (row pitch NX) and the tight layout of the C ABI (tNX), as
  // in the CGFDM solver
  int NX = 0, NY = 0, NZ = 0;
  long long N = 0, S = 0;
  int tNX = 0;
  long long tS = 0, tN = 0;
  R* buf[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  R* cjm = nullptr;
  R* F = nullptr;
  R alpha = R(0);
  float* pgv = nullptr;
  R* dis = nullptr;
  int ns = 0;
  long long* st_cell = nullptr;
  float* st_rec = nullptr;
  int src_np = 0, src_nt = 0;
  double src_dt = 0;
  long long* src_cell = nullptr;
  R* src_rate = nullptr;
  std::vector<double> jac_host;  // FP64 J for the source scaling

  // block order of the face and update kernels: regions of (64 rtx) x (4 rty)
  // columns per XCD, rtx 0 = whole rows of tiles; the environment variables
  // CGFDM_INTERIOR_RTX (0..64) / _RTY (1..64) override it
  int rty = 2;
  int rtx = 0;

  int init(const cgfdm_desc& dd, const float* coord, const float* medium) override {
    d = dd;
    if (const char* e = std::getenv("CGFDM_INTERIOR_RTY")) {
      const int v = std::atoi(e);
      if (v >= 1 && v <= 64) rty = v;
    }
    if (const char* e = std::getenv("CGFDM_INTERIOR_RTX")) {
      const int v = std::atoi(e);
      if (v >= 0 && v <= 64) rtx = v;
    }
    if (const char* e = std::getenv("CGFDM_SURFACE_STREAM")) surf_overlap = std::atoi(e) != 0 ? 1 : 0;
    if (const char* e = std::getenv("CGFDM_ROW_ALIGN")) galign = std::atoi(e) != 0;
    tNX = d.nx + 2 * cg::H;
    NX = tNX;
    NY = d.ny + 2 * cg::H;
    NZ = d.nz + 2 * cg::H;
    S = (long long)NX * NY;
    N = S * NZ;
    tS = S;
    tN = N;
    int rc;
    float *dcoord = nullptr, *dmed = nullptr;
    double* cov = nullptr;
    if ((rc = dalloc(&dcoord, 3 * N))) return rc;
    if ((rc = dalloc(&dmed, 3 * N))) return rc;
    if ((rc = dalloc(&cov, 10 * N))) return rc;
    HIPCHK(hipMemcpyAsync(dcoord, coord, sizeof(float) * 3 * N, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dmed, medium, sizeof(float) * 3 * N, hipMemcpyHostToDevice, stream));
    if ((rc = grid_ranges(dcoord, dmed, d.nx, d.ny, d.nz, NX, NY, N))) return rc;
    double dtmax = 1.34 * ranges[0] / ranges[5];
    if (d.dt > dtmax)
      return fail(CGFDM_ERR_CFL, "DT " + std::to_string(d.dt) + " exceeds the CFL limit " + std::to_string(dtmax));
    alpha = R((float)ranges[5]);  // vp_max_for_SCFDM (calc_CFL.cpp:317-321)
    cg::launch_cov_cmm(dcoord, cov, NX, NY, NZ, N, d.dh, stream);
    R* cjm_t = nullptr;
    if ((rc = dalloc(&cjm_t, (long long)aw::C_N * N))) return rc;
    cg::launch_pack_cmm<R>(cov, dmed, cjm_t, NX, NY, NZ, N, stream);
    HIPCHK(hipGetLastError());
    jac_host.assign(N, 0.0);  // tight layout
    HIPCHK(hipMemcpyAsync(jac_host.data(), cov + 9 * N, sizeof(double) * N, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    dfree(cov);
    dfree(dmed);
    dfree(dcoord);
    // row-aligned device layout (see the CGFDM solver)
    NX = galign ? (tNX + 31) / 32 * 32 : tNX;
    S = (long long)NX * NY;
    N = S * NZ;
    if ((double)N * (double)sizeof(R) >= 4294967296.0)
      return fail(CGFDM_ERR_ARG, "grid too large: one field plane array must stay below 4 GiB (" +
                                     std::to_string(N) + " padded cells)");
    pNX = NX;
    pNY = NY;
    pNZ = NZ;
    pXe = tNX;
    if ((rc = galloc(&cjm, (long long)aw::C_N * N))) return rc;
    if ((rc = repack(cjm, cjm_t, (long long)aw::C_N * NY * NZ, true))) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    dfree(cjm_t);
    for (int q = 0; q < 3; ++q)
      if ((rc = galloc(&buf[q], 9 * N))) return rc;
    if ((rc = galloc(&F, 27 * N))) return rc;
    if ((rc = dalloc(&pgv, 2LL * d.nx * d.ny))) return rc;
    if ((rc = dalloc(&dis, 3 * S))) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    return CGFDM_OK;
  }

  int set_source(int np, int nt, double dt, const long long* cell, const float* rate) override {
    if (np < 0 || (np > 0 && (nt < 2 || dt <= 0 || !cell || !rate))) return fail(CGFDM_ERR_SOURCE, "bad source table");
    std::map<long long, std::vector<double>> acc;
    for (int p = 0; p < np; ++p) {
      const long long c = cell[p];  // tight index; the device index follows below
      if (dev_cell(c) < 0) return fail(CGFDM_ERR_SOURCE, "source cell outside the interior");
      auto& v = acc[c];
      if (v.empty()) v.assign((size_t)nt * 6, 0.0);
      const float* r = rate + (size_t)p * nt * 6;
      for (size_t q = 0; q < (size_t)nt * 6; ++q) v[q] += r[q];
    }
    int rc;
    if (src_cell) dfree(src_cell);
    if (src_rate) dfree(src_rate);
    src_cell = nullptr;
    src_rate = nullptr;
    src_np = (int)acc.size();
    src_nt = nt;
    src_dt = dt;
    if (src_np == 0) return CGFDM_OK;
    std::vector<long long> hc;
    std::vector<R> hr((size_t)src_np * nt * 6);
    int Jp = 0;
    for (auto& kv : acc) {
      const long long ct = kv.first, c = dev_cell(ct);
      hc.push_back(c);
      R mv;
      HIPCHK(hipMemcpy(&mv, cjm + (long long)aw::C_MU * N + c, sizeof(R), hipMemcpyDeviceToHost));
      double sc = (double)mv / (jac_host[ct] * d.dh * d.dh * d.dh);  // mu/(J DH^3) (addMoment.cpp:27-76)
      for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 6; ++q) hr[((size_t)t * src_np + Jp) * 6 + q] = R(kv.second[(size_t)t * 6 + q] * sc);
      ++Jp;
    }
    if ((rc = dalloc(&src_cell, src_np))) return rc;
    if ((rc = dalloc(&src_rate, (long long)hr.size()))) return rc;
    HIPCHK(hipMemcpy(src_cell, hc.data(), sizeof(long long) * src_np, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(src_rate, hr.data(), sizeof(R) * hr.size(), hipMemcpyHostToDevice));
    return CGFDM_OK;
  }

  int set_stations(int n, const int* ijk) override {
    if (n < 0 || (n > 0 && !ijk)) return fail(CGFDM_ERR_ARG, "bad station list");
    std::vector<long long> hc(n);
    for (int s = 0; s < n; ++s) {
      int i = ijk[3 * s], j = ijk[3 * s + 1], k = ijk[3 * s + 2];
      if (i < cg::H || i >= pXe - cg::H || j < cg::H || j >= NY - cg::H || k < cg::H || k >= NZ - cg::H)
        return fail(CGFDM_ERR_ARG, "station outside the interior");
      hc[s] = (long long)i + (long long)NX * j + (long long)k * S;
    }
    int rc;
    if (st_cell) dfree(st_cell);
    if (st_rec) dfree(st_rec);
    st_cell = nullptr;
    st_rec = nullptr;
    ns = n;
    if (n == 0) return CGFDM_OK;
    if ((rc = dalloc(&st_cell, n))) return rc;
    if ((rc = dalloc(&st_rec, (long long)n * d.nt * 9))) return rc;
    HIPCHK(hipMemcpy(st_cell, hc.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    return CGFDM_OK;
  }

  const char* kernel_name(int which) const override {
    return which == 0 ? "aw_faces" : which == 1 ? "aw_update" : nullptr;
  }

  int step(int n) override {
    if (surf_overlap && !sstream) {
      HIPCHK(hipStreamCreateWithFlags(&sstream, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    for (int s = 0; s < n; ++s) {
      size_t sp = 0;
      if (timing_on) {
        sp = next_pair(2);
        HIPCHK(hipEventRecord(ev[sp].a, stream));
      }
      R* W = buf[cur];
      if (src_np > 0) {
        float DTf = (float)d.dt;
        float tmpT = ((float)it + 0.0f) * DTf;
        int srcIt = int(tmpT / (float)src_dt);
        if (srcIt + 1 < src_nt) {
          float t1 = float(srcIt) * (float)src_dt;
          float tw = (tmpT - t1) / (float)src_dt;
          aw::launch_source<R>(W, cjm, N, src_np, src_cell, src_rate, srcIt, R(tw), R(d.dt), stream);
        }
      }
      const int o1 = (cur + 1) % 3, o2 = (cur + 2) % 3;
      R* ins[3] = {buf[cur], buf[o1], buf[o2]};
      R* outs[3] = {buf[o1], buf[o2], buf[o1]};
      for (int irk = 0; irk < 3; ++irk) {
        aw::StageArgs<R> a{};
        a.NX = NX;
        a.XE = tNX;
        a.NY = NY;
        a.NZ = NZ;
        a.N = N;
        a.S = S;
        a.rowb = (unsigned)(NX * sizeof(R));
        a.planeb = (unsigned)(S * sizeof(R));
        a.rtx = rtx;
        a.rty = rty;
        a.stage = irk;
        a.DT = R(d.dt);
        a.rDH = R(1.0 / d.dh);
        a.alpha = alpha;
        a.cjm = cjm;
        a.win = ins[irk];
        a.wm = buf[cur];
        a.wout = outs[irk];
        a.F = F;
        size_t fp = 0, up = 0;
        if (timing_on) {
          fp = next_pair(0);
          HIPCHK(hipEventRecord(ev[fp].a, stream));
        }
        aw::launch_faces<R>(a, stream);
        if (timing_on) {
          HIPCHK(hipEventRecord(ev[fp].b, stream));
          up = next_pair(1);
          HIPCHK(hipEventRecord(ev[up].a, stream));
        }
        if (surf_overlap) {
          HIPCHK(hipEventRecord(ev_fork, stream));
          HIPCHK(hipStreamWaitEvent(sstream, ev_fork, 0));
        }
        aw::launch_update<R>(a, d.nx, d.ny, stream, surf_overlap ? sstream : stream);
        if (timing_on) HIPCHK(hipEventRecord(ev[up].b, stream));
        if (surf_overlap) {
          HIPCHK(hipEventRecord(ev_join, sstream));
          HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
        }
      }
      cur = o1;
      W = buf[cur];
      if (d.free_surface) aw::launch_char_surface<R>(W, cjm, d.nx, d.ny, NX, NZ, N, S, stream);
      aw::launch_sponge<R>(W, d.nx, d.ny, d.nz, NX, N, S, d.free_surface ? 1 : 0, stream);
      if (d.free_surface) aw::launch_surface_obs<R>(W, cjm, pgv, dis, d.nx, d.ny, NX, NZ, N, S, R(d.dt), stream);
      aw::launch_record<R>(W, cjm, N, ns, st_cell, st_rec, d.nt, it, stream);
      if (timing_on) HIPCHK(hipEventRecord(ev[sp].b, stream));
      ++it;
    }
    HIPCHK(hipGetLastError());
    return CGFDM_OK;
  }

  int get_stations(float* out) override {
    if (ns == 0) return CGFDM_OK;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(out, st_rec, sizeof(float) * (size_t)ns * d.nt * 9, hipMemcpyDeviceToHost));
    return CGFDM_OK;
  }

  int get_pgv(float* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(out, pgv, sizeof(float) * 2 * (size_t)d.nx * d.ny, hipMemcpyDeviceToHost));
    return CGFDM_OK;
  }

  int get_dis(float* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<R> h(3 * (size_t)S);
    HIPCHK(hipMemcpy(h.data(), dis, sizeof(R) * h.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c)
      for (int j = 0; j < d.ny; ++j)
        for (int i = 0; i < d.nx; ++i)
          out[((size_t)c * d.ny + j) * d.nx + i] = (float)h[(size_t)c * S + (size_t)(j + cg::H) * NX + (i + cg::H)];
    return CGFDM_OK;
  }

  int get_field(int comp, double* out) override {
    if (comp < 0 || comp > 8) return fail(CGFDM_ERR_ARG, "component out of range");
    std::vector<R> h(tN);
    if (int rc = to_host(buf[cur] + (size_t)comp * N, h.data(), 1)) return rc;
    for (long long p = 0; p < tN; ++p) out[p] = (double)h[p];
    return CGFDM_OK;
  }

  // physical (Vx, Vy, Vz, Txx, Tyy, Tzz, Txy, Txz, Tyz) from the conservative
  // state (rho v, strains) and mu, lambda, 1/rho (wave_conserv2phy,
  // data_io.cpp:435-483)
  int get_slice(int comp, int axis, int index, float* out) override {
    if (int rc = check_plane(comp, axis, index)) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    const R* W = buf[cur];
    std::vector<R> u0, u1, u2, mu, lam, b;
    int rc;
    if (comp < 3) {
      if ((rc = copy_plane(W + (size_t)comp * N, axis, index, u0))) return rc;
      if ((rc = copy_plane(cjm + (size_t)aw::C_B * N, axis, index, b))) return rc;
      for (size_t p = 0; p < u0.size(); ++p) out[p] = (float)(u0[p] * b[p]);
      return CGFDM_OK;
    }
    if ((rc = copy_plane(cjm + (size_t)aw::C_MU * N, axis, index, mu))) return rc;
    if (comp >= 6) {
      // Txy = mu gxy (8), Txz = mu gxz (7), Tyz = mu gyz (6)
      if ((rc = copy_plane(W + (size_t)(14 - comp) * N, axis, index, u0))) return rc;
      for (size_t p = 0; p < u0.size(); ++p) out[p] = (float)(mu[p] * u0[p]);
      return CGFDM_OK;
    }
    if ((rc = copy_plane(cjm + (size_t)aw::C_LAM * N, axis, index, lam))) return rc;
    if ((rc = copy_plane(W + (size_t)3 * N, axis, index, u0))) return rc;
    if ((rc = copy_plane(W + (size_t)4 * N, axis, index, u1))) return rc;
    if ((rc = copy_plane(W + (size_t)5 * N, axis, index, u2))) return rc;
    for (size_t p = 0; p < u0.size(); ++p) {
      const R l = lam[p], m2 = l + 2 * mu[p];
      const R v = comp == 3 ? l * u1[p] + l * u2[p] + u0[p] * m2
                : comp == 4 ? l * u0[p] + l * u2[p] + u1[p] * m2
                            : l * u0[p] + l * u1[p] + u2[p] * m2;
      out[p] = (float)v;
    }
    return CGFDM_OK;
  }

  int set_field(int comp, const double* in) override {
    if (comp < 0 || comp > 8) return fail(CGFDM_ERR_ARG, "component out of range");
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<R> h(tN);
    for (long long p = 0; p < tN; ++p) h[p] = R(in[p]);
    return to_dev(buf[cur] + (size_t)comp * N, h.data(), 1);
  }

  int get_setup(int what, double* out) override {
    if (what < 0 || what >= aw::C_N) return fail(CGFDM_ERR_ARG, "unknown setup product");
    std::vector<R> h(tN);
    if (int rc = to_host(cjm + (size_t)what * N, h.data(), 1)) return rc;
    for (long long p = 0; p < tN; ++p) out[p] = (double)h[p];
    return CGFDM_OK;
  }
};

}  // namespace

struct cgfdm_solver {
  std::unique_ptr<SolverBase> impl;
};

extern "C" {

const char* cgfdm_last_error(void) { return g_err.c_str(); }

int cgfdm_create(const cgfdm_desc* desc, const float* coord, const float* medium, cgfdm_solver** out) {
  if (!desc || !coord || !medium || !out) return fail(CGFDM_ERR_ARG, "null argument");
  *out = nullptr;
  const cgfdm_desc& d = *desc;
  if (d.nx < 1 || d.ny < 1 || d.nz < 4 || d.dh <= 0 || d.dt <= 0 || d.nt < 0)
    return fail(CGFDM_ERR_ARG, "bad grid description");
  if (d.npml < 0) return fail(CGFDM_ERR_ARG, "negative npml");
  if (d.npml > 0 && (d.npml >= d.nx || d.npml >= d.ny || d.npml >= d.nz))
    return fail(CGFDM_ERR_PML, "nPML must be smaller than nx, ny and nz");
  if (d.scheme != CGFDM_SCHEME_CGFDM && d.scheme != CGFDM_SCHEME_AWENO) return fail(CGFDM_ERR_ARG, "unknown scheme");
  std::unique_ptr<SolverBase> s;
  if (d.scheme == CGFDM_SCHEME_AWENO) {
    if (d.coef_storage != 0) return fail(CGFDM_ERR_ARG, "16-bit coefficient storage is a CGFDM option");
    // the update kernel's lanes of the first and last three interior columns
    // also update the ghost columns
    if (d.nx < 3) return fail(CGFDM_ERR_ARG, "the ME-AWENO scheme needs nx >= 3");
    if (d.precision == 8) s.reset(new AwenoSolver<double>());
    else if (d.precision == 4) s.reset(new AwenoSolver<float>());
    else return fail(CGFDM_ERR_ARG, "precision must be 4 or 8");
  } else if (d.precision == 8) {
    if (d.coef_storage != 0) return fail(CGFDM_ERR_ARG, "16-bit coefficient storage needs the FP32 wavefield");
    s.reset(new Solver<double, double>());
  } else if (d.precision == 4) {
    if (d.coef_storage == 0) s.reset(new Solver<float, float>());
    else if (d.coef_storage == 1) s.reset(new Solver<float, __half>());
    else if (d.coef_storage == 2) s.reset(new Solver<float, __hip_bfloat16>());
    else return fail(CGFDM_ERR_ARG, "bad coef_storage");
  } else {
    return fail(CGFDM_ERR_ARG, "precision must be 4 or 8");
  }
  int rc = s->init(d, coord, medium);
  if (rc) return rc;
  auto* h = new cgfdm_solver();
  h->impl = std::move(s);
  *out = h;
  return CGFDM_OK;
}

int cgfdm_grid_ranges(int nx, int ny, int nz, const float* coord, const float* medium, double* out8) {
  if (nx < 1 || ny < 1 || nz < 1 || !coord || !medium || !out8) return fail(CGFDM_ERR_ARG, "bad arguments");
  const int NX = nx + 2 * cg::H, NY = ny + 2 * cg::H, NZ = nz + 2 * cg::H;
  const long long N = (long long)NX * NY * NZ;
  const int nb = 1024;
  float *dc = nullptr, *dm = nullptr;
  double* part = nullptr;
  auto cleanup = [&]() {
    if (dc) (void)hipFree(dc);
    if (dm) (void)hipFree(dm);
    if (part) (void)hipFree(part);
  };
  if (hipMalloc(&dc, sizeof(float) * 3 * N) != hipSuccess || hipMalloc(&dm, sizeof(float) * 3 * N) != hipSuccess ||
      hipMalloc(&part, sizeof(double) * 8 * nb) != hipSuccess) {
    cleanup();
    return fail(CGFDM_ERR_ALLOC, "hipMalloc failed");
  }
  std::vector<double> h(8 * nb);
  hipError_t e = hipMemcpy(dc, coord, sizeof(float) * 3 * N, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dm, medium, sizeof(float) * 3 * N, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    cg::launch_hrange(dc, part, nb, nx, ny, nz, NX, NY, N, nullptr);
    for (int c = 0; c < 3; ++c) cg::launch_frange(dm + c * N, N, part + 2 * nb * (c + 1), nb, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return fail(CGFDM_ERR_HIP, hipGetErrorString(e));
  for (int c = 0; c < 4; ++c) {
    double lo = 1e300, hi = -1e300;
    for (int b = 0; b < nb; ++b) {
      lo = std::min(lo, h[2 * nb * c + 2 * b]);
      hi = std::max(hi, h[2 * nb * c + 2 * b + 1]);
    }
    out8[2 * c] = lo;
    out8[2 * c + 1] = hi;
  }
  return CGFDM_OK;
}

int cgfdm_grid_ranges_noalloc(int nx, int ny, int nz, const float* coord, const float* medium, double* out8,
                              void* stream) {
  if (!coord || !medium || !out8) return fail(CGFDM_ERR_ARG, "null pointer argument");
  constexpr long long kMaxDim = 1LL << 30, kMaxCells = 1LL << 40;
  if (nx < 1 || ny < 1 || nz < 1 || nx > kMaxDim || ny > kMaxDim || nz > kMaxDim)
    return fail(CGFDM_ERR_ARG, "grid dimensions must lie in 1 .. 2^30");
  const long long cells = (long long)(nx + 2 * cg::H) * (ny + 2 * cg::H);
  if (cells > kMaxCells / (nz + 2 * cg::H))
    return fail(CGFDM_ERR_ARG, "padded grid beyond the 2^40-cell maximum of the allocation-free ranges");
  hipError_t e = cg::grid_ranges_noalloc(nx, ny, nz, coord, medium, out8, (hipStream_t)stream);
  if (e != hipSuccess) return fail(CGFDM_ERR_HIP, hipGetErrorString(e));
  return CGFDM_OK;
}

const char* cgfdm_kernel_name(cgfdm_solver* s, int which) {
  if (!s || !s->impl) return nullptr;
  return s->impl->kernel_name(which);
}

int cgfdm_destroy(cgfdm_solver* s) {
  if (!s) return CGFDM_OK;
  cgfdm_detail::release_session(s);
  if (s->impl && s->impl->stream) (void)hipStreamSynchronize(s->impl->stream);  // drain before the frees
  delete s;
  return CGFDM_OK;
}

#define SOLVER_OR_FAIL(s) \
  if (!(s) || !(s)->impl) return fail(CGFDM_ERR_ARG, "null solver")

int cgfdm_forward_record(cgfdm_solver* s, int nsteps, int checkpoint_every, double* rec) {
  SOLVER_OR_FAIL(s);
  return s->impl->forward_record(nsteps, checkpoint_every, rec);
}

int cgfdm_adjoint(cgfdm_solver* s, const double* src, double* gmu, double* glam, double* gb, double* grate) {
  SOLVER_OR_FAIL(s);
  return s->impl->adjoint(src, gmu, glam, gb, grate);
}

int cgfdm_set_source(cgfdm_solver* s, int npts, int nt, double dt, const long long* cell, const float* rate) {
  SOLVER_OR_FAIL(s);
  return s->impl->set_source(npts, nt, dt, cell, rate);
}

int cgfdm_set_stations(cgfdm_solver* s, int ns, const int* ijk) {
  SOLVER_OR_FAIL(s);
  return s->impl->set_stations(ns, ijk);
}

int cgfdm_step(cgfdm_solver* s, int nsteps) {
  SOLVER_OR_FAIL(s);
  if (nsteps < 0) return fail(CGFDM_ERR_ARG, "negative step count");
  return s->impl->step(nsteps);
}

int cgfdm_set_stream(cgfdm_solver* s, void* stream) {
  SOLVER_OR_FAIL(s);
  // work queued on the old stream finishes before anything is queued on the
  // new one, so the switch keeps the one-stream order of the handle
  if ((hipStream_t)stream != s->impl->stream) {
    hipError_t e = hipStreamSynchronize(s->impl->stream);
    if (e != hipSuccess) return fail(CGFDM_ERR_HIP, hipGetErrorString(e));
  }
  s->impl->stream = (hipStream_t)stream;
  return CGFDM_OK;
}

int cgfdm_synchronize(cgfdm_solver* s) {
  SOLVER_OR_FAIL(s);
  hipError_t e = hipStreamSynchronize(s->impl->stream);
  if (e != hipSuccess) return fail(CGFDM_ERR_HIP, hipGetErrorString(e));
  return CGFDM_OK;
}

int cgfdm_step_index(cgfdm_solver* s) { return (s && s->impl) ? s->impl->it : -1; }

int cgfdm_get_stations(cgfdm_solver* s, float* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->get_stations(out);
}
int cgfdm_get_pgv(cgfdm_solver* s, float* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->get_pgv(out);
}
int cgfdm_get_surface_displacement(cgfdm_solver* s, float* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->get_dis(out);
}
int cgfdm_get_slice(cgfdm_solver* s, int comp, int axis, int index, float* out) {
  SOLVER_OR_FAIL(s);
  if (!out) return fail(CGFDM_ERR_ARG, "null output");
  return s->impl->get_slice(comp, axis, index, out);
}
int cgfdm_get_field(cgfdm_solver* s, int comp, double* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->get_field(comp, out);
}
int cgfdm_set_field(cgfdm_solver* s, int comp, const double* in) {
  SOLVER_OR_FAIL(s);
  return s->impl->set_field(comp, in);
}
int cgfdm_get_setup(cgfdm_solver* s, int what, double* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->get_setup(what, out);
}
int cgfdm_get_ranges(cgfdm_solver* s, double* out8) {
  SOLVER_OR_FAIL(s);
  std::memcpy(out8, s->impl->ranges, sizeof(double) * 8);
  return CGFDM_OK;
}
int cgfdm_timing(cgfdm_solver* s, int enable, double* out) {
  SOLVER_OR_FAIL(s);
  return s->impl->timing(enable, out);
}
long long cgfdm_device_bytes(cgfdm_solver* s) { return (s && s->impl) ? s->impl->bytes : 0; }

}  // extern "C"
