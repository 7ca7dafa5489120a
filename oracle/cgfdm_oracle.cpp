// CPU oracle for the CGFDM3D-EQR-SC time-stepping path (CGFDM scheme).
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain loops, what the
// reference's CUDA kernels compute, so that the HIP product path can be
// checked against it. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never links or calls it.
//
// The restatement keeps the reference's operator order and per-cell
// arithmetic order (single precision instantiation = the reference "FLOAT"
// build; double instantiation = the FP64 oracle). Layout is SoA over the
// padded grid, x fastest: field c of cell (i,j,k) lives at
// c*N + i + NX*(j + NY*k), NX = nx + 6 etc. (HALO = 3, reference macro.h:39).
//
// Pinning: every function below is checked by tests/test_oracle_*.py against
// closed forms (polynomial exactness of the MacCormack operators, RK4
// amplification, flat/affine-grid metrics, traction-free surface rate,
// Aki-Richards moment tensors, PML closed-form profile points, linearity).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>
#include <algorithm>
#include <type_traits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr int H = 3;

// MacCormack one-sided 5-point coefficients (reference macro.h:121-125).
constexpr double AF_1 = -0.30874, AF0 = -0.6326, AF1 = 1.2330, AF2 = -0.3334, AF3 = 0.04168;
// 3-point one-sided coefficients used one layer below the surface (macro.h:144-146).
constexpr double CF1 = -1.16666667, CF2 = 1.33333333, CF3 = -0.16666667;
// Low-storage RK4 weights (macro.h:133-142).
constexpr double RK_A2 = 0.5, RK_A3 = 0.5;
constexpr double RK_B1 = 0.16666667, RK_B2 = 0.33333333, RK_B3 = 0.33333333, RK_B4 = 0.16666667;
// The reference's PI literal (macro.h:70) enters the PML alpha profile.
constexpr double REF_PI = 3.141592657;

// F/B schedule selected by it % 8 (propagate.cpp:158-168).
static const int FBTAB[8][3] = {{-1, -1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, -1, 1},
                                {-1, 1, -1},  {1, -1, -1}, {1, -1, 1}, {-1, 1, 1}};

enum { C_XIX = 0, C_XIY, C_XIZ, C_ETX, C_ETY, C_ETZ, C_ZTX, C_ZTY, C_ZTZ, C_J, C_MU, C_LAM, C_B, C_N };

template <class T>
struct Face {
  int axis, side;      // axis 0/1/2, side 0 = low face (x1,y1,z1), 1 = high face
  int n0, n1;          // extents of the two fastest dims of the slab
  long long cnt;       // cells in the slab
  std::vector<T> h, a, t, m;  // 9*cnt each: RHS, state, RK accumulator, RK base
};

template <class T>
struct Oracle {
  int nx, ny, nz, NX, NY, NZ, npml;
  long long N;
  bool fs;
  double dh;
  T DT, rDH;
  std::vector<T> hW, W, tW, mW;     // 9*N each
  std::vector<T> cjm;               // C_N*N
  std::vector<T> rdzdx, rdzdy;      // 9*NX*NY
  std::vector<T> pa[3], pb[3], pd[3];
  std::vector<Face<T>> faces;
  std::vector<double> metric_d;     // metrics in double (9*N) + J (N) for setup queries
  double hmin = 0, hmax = 0, vpmax = 0, vpmin = 0, vsmin = 0, vsmax = 0, rhomin = 0, rhomax = 0;
  // source (time-major [t][p][6] after scaling)
  int src_np = 0, src_nt = 0;
  double src_dt = 0;
  std::vector<long long> src_idx;
  std::vector<T> src_rate;
  std::vector<double> src_raw;      // merged, unscaled, time-major [t][p][6]
  std::vector<double> src_vol;      // 1 / (J DH^3) per merged point
  std::vector<int> src_of_point;    // merged index of every input point
  // observers
  int ns = 0, NTrec = 0;
  std::vector<int> st_ijk;
  std::vector<float> st_wave;
  std::vector<float> pgv;           // [ny][nx][2]
  std::vector<T> dis;               // surface displacement [3][NY*NX]
  int it = 0;

  long long at(int i, int j, int k) const { return i + (long long)NX * (j + (long long)NY * k); }

  // ---------------------------------------------------------------------
  // setup
  // ---------------------------------------------------------------------
  // Metric terms from coordinates, FP64 (contravariant.cpp:182-285): averaged
  // forward/backward MacCormack differences of x,y,z, J = det, xi_x = cof/J.
  void build_metrics(const float* coord) {
    const double rdh = 1.0 / dh;
    metric_d.assign(10 * (size_t)N, 0.0);
    auto LC = [&](const float* c, long long p, long long s, int fb) {
      long long d = fb * s;
      return fb * (AF_1 * c[p - d] + AF0 * c[p] + AF1 * c[p + d] + AF2 * c[p + 2 * d] + AF3 * c[p + 3 * d]) * rdh;
    };
    const float* X = coord;
    const float* Y = coord + N;
    const float* Z = coord + 2 * N;
    const long long sx = 1, sy = NX, sz = (long long)NX * NY;
#pragma omp parallel for collapse(2)
    for (int k = H; k < NZ - H; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          double x_xi = 0.5 * (LC(X, p, sx, 1) + LC(X, p, sx, -1));
          double x_et = 0.5 * (LC(X, p, sy, 1) + LC(X, p, sy, -1));
          double x_zt = 0.5 * (LC(X, p, sz, 1) + LC(X, p, sz, -1));
          double y_xi = 0.5 * (LC(Y, p, sx, 1) + LC(Y, p, sx, -1));
          double y_et = 0.5 * (LC(Y, p, sy, 1) + LC(Y, p, sy, -1));
          double y_zt = 0.5 * (LC(Y, p, sz, 1) + LC(Y, p, sz, -1));
          double z_xi = 0.5 * (LC(Z, p, sx, 1) + LC(Z, p, sx, -1));
          double z_et = 0.5 * (LC(Z, p, sy, 1) + LC(Z, p, sy, -1));
          double z_zt = 0.5 * (LC(Z, p, sz, 1) + LC(Z, p, sz, -1));
          double J = x_xi * y_et * z_zt + x_et * y_zt * z_xi + x_zt * y_xi * z_et - x_zt * y_et * z_xi -
                     x_et * y_xi * z_zt - x_xi * z_et * y_zt;
          double r = 1.0 / J;
          double* m = metric_d.data();
          m[0 * N + p] = (y_et * z_zt - y_zt * z_et) * r;
          m[1 * N + p] = (x_zt * z_et - x_et * z_zt) * r;
          m[2 * N + p] = (x_et * y_zt - x_zt * y_et) * r;
          m[3 * N + p] = (y_zt * z_xi - y_xi * z_zt) * r;
          m[4 * N + p] = (x_xi * z_zt - x_zt * z_xi) * r;
          m[5 * N + p] = (x_zt * y_xi - x_xi * y_zt) * r;
          m[6 * N + p] = (y_xi * z_et - y_et * z_xi) * r;
          m[7 * N + p] = (x_et * z_xi - x_xi * z_et) * r;
          m[8 * N + p] = (x_xi * y_et - x_et * y_xi) * r;
          m[9 * N + p] = J;
        }
    if (fs) {
      // Metrics and J mirrored into the top halo about the surface layer
      // (contravariant.cpp:449-463).
      int ks = NZ - H - 1;
      for (int k = ks + 1; k < NZ; ++k)
        for (int j = H; j < NY - H; ++j)
          for (int i = H; i < NX - H; ++i) {
            long long p = at(i, j, k), q = at(i, j, 2 * ks - k);
            for (int c = 0; c < 10; ++c) metric_d[c * N + p] = metric_d[c * N + q];
          }
    }
  }

  // Medium (Vs,Vp,rho) -> (mu, lambda, 1/rho) over the whole padded grid
  // (CJM.cpp:53-75) and packing with the metrics (CJM.cpp:107-147).
  void build_cjm(const float* medium) {
    cjm.assign((size_t)C_N * N, T(0));
    const float* VS = medium;
    const float* VP = medium + N;
    const float* RHO = medium + 2 * N;
    for (long long p = 0; p < N; ++p) {
      for (int c = 0; c < 10; ++c) cjm[c * N + p] = T(metric_d[c * N + p]);
      double vs = VS[p], vp = VP[p], rho = RHO[p];
      cjm[C_MU * N + p] = T(rho * vs * vs);
      cjm[C_LAM * N + p] = T(rho * (vp * vp - 2.0 * vs * vs));
      cjm[C_B * N + p] = T(1.0 / rho);
    }
  }

  // Traction-free surface matrices -DZ^-1 DX, -DZ^-1 DY at k = surface
  // (contravariant.cpp:363-447), FP64.
  void build_rdz(const float* medium) {
    rdzdx.assign(9 * (size_t)NX * NY, T(0));
    rdzdy.assign(9 * (size_t)NX * NY, T(0));
    if (!fs) return;
    const float* VS = medium;
    const float* VP = medium ? medium + N : nullptr;
    const float* RHO = medium ? medium + 2 * N : nullptr;
    int k = NZ - H - 1;
    const long long S = (long long)NX * NY;
    for (int j = H; j < NY - H; ++j)
      for (int i = H; i < NX - H; ++i) {
        long long p = at(i, j, k);
        const double* m = metric_d.data();
        double xx = m[0 * N + p], xy = m[1 * N + p], xz = m[2 * N + p];
        double ex = m[3 * N + p], ey = m[4 * N + p], ez = m[5 * N + p];
        double zx = m[6 * N + p], zy = m[7 * N + p], zz = m[8 * N + p];
        double mu, lam;
        if (medium) {
          double vs = VS[p], vp = VP[p], rho = RHO[p];
          mu = rho * vs * vs;
          lam = rho * (vp * vp - 2.0 * vs * vs);
        } else {
          mu = (double)cjm[C_MU * N + p];
          lam = (double)cjm[C_LAM * N + p];
        }
        double chi = 2.0 * lam + mu, lm = lam + mu;
        double DZ[9], DX[9], DY[9], R[9];
        DZ[0] = chi * zx * zx + mu * (zy * zy + zz * zz);
        DZ[1] = lm * zx * zy;
        DZ[2] = lm * zx * zz;
        DZ[3] = DZ[1];
        DZ[4] = chi * zy * zy + mu * (zx * zx + zz * zz);
        DZ[5] = lm * zy * zz;
        DZ[6] = DZ[2];
        DZ[7] = DZ[5];
        DZ[8] = chi * zz * zz + mu * (zx * zx + zy * zy);
        auto fillD = [&](double* D, double ax, double ay, double az) {
          D[0] = chi * zx * ax + mu * (zy * ay + zz * az);
          D[1] = lam * zx * ay + mu * zy * ax;
          D[2] = lam * zx * az + mu * zz * ax;
          D[3] = lam * zy * ax + mu * zx * ay;
          D[4] = chi * zy * ay + mu * (zx * ax + zz * az);
          D[5] = lam * zy * az + mu * zz * ay;
          D[6] = lam * zz * ax + mu * zx * az;
          D[7] = lam * zz * ay + mu * zy * az;
          D[8] = chi * zz * az + mu * (zx * ax + zy * ay);
        };
        fillD(DX, xx, xy, xz);
        fillD(DY, ex, ey, ez);
        double det = DZ[0] * DZ[4] * DZ[8] + DZ[1] * DZ[5] * DZ[6] + DZ[2] * DZ[7] * DZ[3] - DZ[2] * DZ[4] * DZ[6] -
                     DZ[1] * DZ[3] * DZ[8] - DZ[0] * DZ[7] * DZ[5];
        R[0] = (DZ[4] * DZ[8] - DZ[5] * DZ[7]) / det;
        R[1] = (-DZ[3] * DZ[8] + DZ[5] * DZ[6]) / det;
        R[2] = (DZ[3] * DZ[7] - DZ[4] * DZ[6]) / det;
        R[3] = (-DZ[1] * DZ[8] + DZ[2] * DZ[7]) / det;
        R[4] = (DZ[0] * DZ[8] - DZ[2] * DZ[6]) / det;
        R[5] = (-DZ[0] * DZ[7] + DZ[1] * DZ[6]) / det;
        R[6] = (DZ[1] * DZ[5] - DZ[2] * DZ[4]) / det;
        R[7] = (-DZ[0] * DZ[5] + DZ[2] * DZ[3]) / det;
        R[8] = (DZ[0] * DZ[4] - DZ[1] * DZ[3]) / det;
        long long s = i + (long long)NX * j;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            rdzdx[(3 * r + c) * S + s] = T(-(R[3 * r] * DX[c] + R[3 * r + 1] * DX[3 + c] + R[3 * r + 2] * DX[6 + c]));
            rdzdy[(3 * r + c) * S + s] = T(-(R[3 * r] * DY[c] + R[3 * r + 1] * DY[3 + c] + R[3 * r + 2] * DY[6 + c]));
          }
      }
  }

  // CFS-PML profiles (init_pml_para.cpp:97-239): alpha, beta, d on each axis;
  // d is stored divided by beta and beta is stored inverted.
  void build_pml() {
    int nn[3] = {NX, NY, NZ};
    double L = dh * (npml - 1);
    double R = std::pow(10.0, -(std::log10((double)npml) - 1.0) / std::log10(2.0) - 3.0);
    double d0 = -(2.0 + 1.0) * 6000.0 / (2.0 * L) * std::log(R);
    double a0 = REF_PI * 0.5, b0 = 3.0;
    for (int ax = 0; ax < 3; ++ax) {
      int n = nn[ax];
      pa[ax].assign(n, T(0));
      pb[ax].assign(n, T(1));
      pd[ax].assign(n, T(0));
      int lo = H, hi = n - H;  // interior [lo, hi)
      for (int q = lo; q < hi; ++q) {
        double l = -1.0;
        if (npml > 0) {
          if (q < H + npml) l = (H + npml - 1 - q) * dh;
          bool top = (ax == 2);
          if (q >= hi - npml && !(top && fs)) l = (q - (hi - npml)) * dh;
        }
        double al = 0, be = 1, dd = 0;
        if (l >= 0) {
          al = a0 * (1.0 - std::pow(l / L, 1.0));
          be = 1.0 + (b0 - 1.0) * std::pow(l / L, 2.0);
          dd = d0 * std::pow(l / L, 2.0);
        }
        pa[ax][q] = T(al);
        pd[ax][q] = T(dd / be);
        pb[ax][q] = T(1.0 / be);
      }
    }
    faces.clear();
    if (npml <= 0) return;
    // Launch order of the reference (pml_deriv.cpp:648-657): x1, y1, z1, x2, y2, [z2].
    int order[6][2] = {{0, 0}, {1, 0}, {2, 0}, {0, 1}, {1, 1}, {2, 1}};
    for (auto& o : order) {
      if (o[0] == 2 && o[1] == 1 && fs) continue;
      Face<T> f;
      f.axis = o[0];
      f.side = o[1];
      if (f.axis == 0) { f.n0 = npml; f.n1 = ny; f.cnt = (long long)npml * ny * nz; }
      else if (f.axis == 1) { f.n0 = nx; f.n1 = npml; f.cnt = (long long)nx * npml * nz; }
      else { f.n0 = nx; f.n1 = ny; f.cnt = (long long)nx * ny * npml; }
      f.h.assign(9 * f.cnt, T(0));
      f.a.assign(9 * f.cnt, T(0));
      f.t.assign(9 * f.cnt, T(0));
      f.m.assign(9 * f.cnt, T(0));
      faces.push_back(std::move(f));
    }
  }

  // Hmin from point-to-plane distances to the 3 neighbours (calc_CFL.cpp:56-163).
  void build_cfl(const float* coord, const float* medium) {
    const float* X = coord;
    const float* Y = coord + N;
    const float* Z = coord + 2 * N;
    double lo = 1e20, hi = 0.0;
#pragma omp parallel for collapse(2) reduction(min : lo) reduction(max : hi)
    for (int k = H; k < NZ - H; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          double P[3] = {X[p], Y[p], Z[p]};
          for (int a = -1; a <= 1; a += 2)
            for (int b = -1; b <= 1; b += 2)
              for (int c = -1; c <= 1; c += 2) {
                long long qa = at(i - a, j, k), qb = at(i, j - b, k), qc = at(i, j, k - c);
                double A[3] = {X[qa], Y[qa], Z[qa]}, B[3] = {X[qb], Y[qb], Z[qb]}, C[3] = {X[qc], Y[qc], Z[qc]};
                double ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
                double ac[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
                double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
                double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                double d = std::fabs(((P[0] - A[0]) * n[0] + (P[1] - A[1]) * n[1] + (P[2] - A[2]) * n[2]) / nn);
                lo = std::min(lo, d);
                hi = std::max(hi, d);
              }
        }
    hmin = lo;
    hmax = hi;
    // medium ranges over the whole padded grid (calc_CFL.cpp:282-290)
    auto range = [&](const float* a, double& mn, double& mx) {
      mn = 1e20; mx = -1e20;
      for (long long p = 0; p < N; ++p) { mn = std::min(mn, (double)a[p]); mx = std::max(mx, (double)a[p]); }
    };
    range(medium, vsmin, vsmax);
    range(medium + N, vpmin, vpmax);
    range(medium + 2 * N, rhomin, rhomax);
  }

  // Moment-rate table: merge points sharing a cell (multiSource.cpp:537-592),
  // reorder time-major (multiSource.cpp:706-752), scale by mu/(J*DH^3)
  // (addMoment.cpp:27-76).
  void set_source(int np, int nt, double dt, const long long* cell, const float* rate) {
    std::map<long long, std::vector<double>> acc;
    for (int p = 0; p < np; ++p) {
      auto& v = acc[cell[p]];
      if (v.empty()) v.assign((size_t)nt * 6, 0.0);
      for (long long q = 0; q < (long long)nt * 6; ++q) v[q] += rate[(long long)p * nt * 6 + q];
    }
    src_np = (int)acc.size();
    src_nt = nt;
    src_dt = dt;
    src_idx.clear();
    src_raw.assign((size_t)src_np * nt * 6, 0.0);
    src_vol.assign(src_np, 0.0);
    std::map<long long, int> merged;
    int J = 0;
    for (auto& kv : acc) {
      long long c = kv.first;
      merged[c] = J;
      src_idx.push_back(c);
      src_vol[J] = 1.0 / (metric_d[9 * N + c] * dh * dh * dh);
      for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 6; ++q) src_raw[((size_t)t * src_np + J) * 6 + q] = kv.second[(size_t)t * 6 + q];
      ++J;
    }
    src_of_point.resize(np);
    for (int p = 0; p < np; ++p) src_of_point[p] = merged[cell[p]];
    scale_source();
  }
  void scale_source() {
    src_rate.assign((size_t)src_np * src_nt * 6, T(0));
    for (int J = 0; J < src_np; ++J) {
      double mu = (double)cjm[C_MU * N + src_idx[J]];
      for (int t = 0; t < src_nt; ++t)
        for (int q = 0; q < 6; ++q) {
          size_t e = ((size_t)t * src_np + J) * 6 + q;
          src_rate[e] = T(src_raw[e] * mu * src_vol[J]);
        }
    }
  }

  // Medium planes given directly as (mu, lambda, 1/rho) in FP64 (for gradient
  // checks): the surface matrices and the source scaling follow them.
  void set_medium_params(const double* mu, const double* lam, const double* b) {
    for (long long p = 0; p < N; ++p) {
      cjm[C_MU * N + p] = T(mu[p]);
      cjm[C_LAM * N + p] = T(lam[p]);
      cjm[C_B * N + p] = T(b[p]);
    }
    build_rdz(nullptr);
    if (src_np) scale_source();
  }

  // ---------------------------------------------------------------------
  // time stepping
  // ---------------------------------------------------------------------
  T L(const T* f, long long p, long long s, int fb) const {
    long long d = fb * s;
    return T(fb) * (T(AF_1) * f[p - d] + T(AF0) * f[p] + T(AF1) * f[p + d] + T(AF2) * f[p + 2 * d] + T(AF3) * f[p + 3 * d]) * rDH;
  }
  T LJT(const T* a, int fb) const {
    return T(fb) * (T(AF_1) * a[3 - fb] + T(AF0) * a[3] + T(AF1) * a[3 + fb] + T(AF2) * a[3 + 2 * fb] + T(AF3) * a[3 + 3 * fb]) * rDH;
  }
  long long stride(int ax) const { return ax == 0 ? 1 : ax == 1 ? (long long)NX : (long long)NX * NY; }

  // Velocity/stress right-hand side contributed by derivatives along one
  // direction with metric (m1,m2,m3) (wave_deriv.cpp:244-272). D = d/d(dir) of
  // the 9 fields.
  static void assemble(const T* D, T m1, T m2, T m3, T mu, T lam, T b, T* out) {
    out[0] = (m1 * D[3] + m2 * D[6] + m3 * D[7]) * b;
    out[1] = (m1 * D[6] + m2 * D[4] + m3 * D[8]) * b;
    out[2] = (m1 * D[7] + m2 * D[8] + m3 * D[5]) * b;
    T div = m1 * D[0] + m2 * D[1] + m3 * D[2];
    out[3] = div * lam + T(2) * mu * (m1 * D[0]);
    out[4] = div * lam + T(2) * mu * (m2 * D[1]);
    out[5] = div * lam + T(2) * mu * (m3 * D[2]);
    out[6] = (m2 * D[0] + m1 * D[1]) * mu;
    out[7] = (m3 * D[0] + m1 * D[2]) * mu;
    out[8] = (m3 * D[1] + m2 * D[2]) * mu;
  }
  // stress-only part of assemble (used where the velocity RHS is replaced)
  static void stress_of(const T* V, T m1, T m2, T m3, T mu, T lam, T* out) {
    T div = m1 * V[0] + m2 * V[1] + m3 * V[2];
    out[0] = div * lam + T(2) * mu * (m1 * V[0]);
    out[1] = div * lam + T(2) * mu * (m2 * V[1]);
    out[2] = div * lam + T(2) * mu * (m3 * V[2]);
    out[3] = (m2 * V[0] + m1 * V[1]) * mu;
    out[4] = (m3 * V[0] + m1 * V[2]) * mu;
    out[5] = (m3 * V[1] + m2 * V[2]) * mu;
  }

  // wave_deriv.cpp:77-284
  void wave_deriv(const int FB[3]) {
    const T* w = W.data();
    const T* c = cjm.data();
#pragma omp parallel for collapse(2)
    for (int k = H; k < NZ - H; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          T bet[3] = {pb[0][i], pb[1][j], pb[2][k]};
          T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
          T acc[9] = {0};
          T tmp[3][9];
          for (int ax = 0; ax < 3; ++ax) {
            T D[9];
            for (int v = 0; v < 9; ++v) D[v] = L(w + v * N, p, stride(ax), FB[ax]) * bet[ax];
            assemble(D, c[(3 * ax) * N + p], c[(3 * ax + 1) * N + p], c[(3 * ax + 2) * N + p], mu, lam, b, tmp[ax]);
          }
          for (int v = 0; v < 9; ++v) acc[v] = (tmp[0][v] + tmp[1][v] + tmp[2][v]) * DT;
          for (int v = 0; v < 9; ++v) hW[v * N + p] = acc[v];
        }
  }

  // Traction-image free surface on the top three layers (freeSurface.cpp:243-494).
  void free_surface_deriv(const int FB[3]) {
    const T* w = W.data();
    const T* c = cjm.data();
    const int ktop = NZ - H - 1;
    const long long S = (long long)NX * NY;
#pragma omp parallel for collapse(2)
    for (int k = ktop - 2; k <= ktop; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
          T bx = pb[0][i], by = pb[1][j];
          int ks = ktop - k + H;
          T JT[3][3][7];  // [dir][component][l]
          for (int l = 0; l < 7; ++l) {
            for (int ax = 0; ax < 2; ++ax) {
              long long q = ax == 0 ? at(i + l - H, j, k) : at(i, j + l - H, k);
              T m1 = c[(3 * ax) * N + q], m2 = c[(3 * ax + 1) * N + q], m3 = c[(3 * ax + 2) * N + q];
              T jac = c[C_J * N + q];
              JT[ax][0][l] = (m1 * w[3 * N + q] + m2 * w[6 * N + q] + m3 * w[7 * N + q]) * jac;
              JT[ax][1][l] = (m1 * w[6 * N + q] + m2 * w[4 * N + q] + m3 * w[8 * N + q]) * jac;
              JT[ax][2][l] = (m1 * w[7 * N + q] + m2 * w[8 * N + q] + m3 * w[5 * N + q]) * jac;
            }
          }
          for (int l = 0; l < ks; ++l) {
            long long q = at(i, j, k + l - H);
            T m1 = c[C_ZTX * N + q], m2 = c[C_ZTY * N + q], m3 = c[C_ZTZ * N + q];
            T jac = c[C_J * N + q];
            JT[2][0][l] = (m1 * w[3 * N + q] + m2 * w[6 * N + q] + m3 * w[7 * N + q]) * jac;
            JT[2][1][l] = (m1 * w[6 * N + q] + m2 * w[4 * N + q] + m3 * w[8 * N + q]) * jac;
            JT[2][2][l] = (m1 * w[7 * N + q] + m2 * w[8 * N + q] + m3 * w[5 * N + q]) * jac;
          }
          for (int cc = 0; cc < 3; ++cc) {
            JT[2][cc][ks] = T(0);
            for (int l = ks + 1; l < 7; ++l) JT[2][cc][l] = -JT[2][cc][2 * ks - l];
          }
          T jinv = T(1) / c[C_J * N + p];
          T hv[3];
          for (int cc = 0; cc < 3; ++cc)
            hv[cc] = b * jinv * (LJT(JT[0][cc], FB[0]) * bx + LJT(JT[1][cc], FB[1]) * by + LJT(JT[2][cc], FB[2]));
          T Vxi[3], Vet[3], Vzt[3];
          for (int v = 0; v < 3; ++v) {
            Vxi[v] = L(w + v * N, p, 1, FB[0]) * bx;
            Vet[v] = L(w + v * N, p, NX, FB[1]) * by;
          }
          if (k == ktop) {
            long long s = i + (long long)NX * j;
            for (int r = 0; r < 3; ++r)
              Vzt[r] = (rdzdx[(3 * r) * S + s] * Vxi[0] + rdzdx[(3 * r + 1) * S + s] * Vxi[1] + rdzdx[(3 * r + 2) * S + s] * Vxi[2]) +
                       (rdzdy[(3 * r) * S + s] * Vet[0] + rdzdy[(3 * r + 1) * S + s] * Vet[1] + rdzdy[(3 * r + 2) * S + s] * Vet[2]);
          } else if (k == ktop - 1) {
            long long d = FB[2] * (long long)S;
            for (int v = 0; v < 3; ++v) Vzt[v] = T(FB[2]) * (w[v * N + p + d] - w[v * N + p]) * rDH;
          } else {
            long long d = FB[2] * (long long)S;
            for (int v = 0; v < 3; ++v)
              Vzt[v] = T(FB[2]) * (T(CF1) * w[v * N + p] + T(CF2) * w[v * N + p + d] + T(CF3) * w[v * N + p + 2 * d]) * rDH;
          }
          T s1[6], s2[6], s3[6];
          stress_of(Vxi, c[C_XIX * N + p], c[C_XIY * N + p], c[C_XIZ * N + p], mu, lam, s1);
          stress_of(Vet, c[C_ETX * N + p], c[C_ETY * N + p], c[C_ETZ * N + p], mu, lam, s2);
          stress_of(Vzt, c[C_ZTX * N + p], c[C_ZTY * N + p], c[C_ZTZ * N + p], mu, lam, s3);
          for (int v = 0; v < 3; ++v) hW[v * N + p] = hv[v] * DT;
          for (int v = 0; v < 6; ++v) hW[(3 + v) * N + p] = (s1[v] + s2[v] + s3[v]) * DT;
        }
  }

  void slab_cell(const Face<T>& f, long long q, int& i, int& j, int& k) const {
    int i0, j0, k0;
    i0 = (int)(q % f.n0);
    j0 = (int)((q / f.n0) % f.n1);
    k0 = (int)(q / ((long long)f.n0 * f.n1));
    i = i0 + H; j = j0 + H; k = k0 + H;
    if (f.axis == 0) i += f.side * (nx - npml);
    if (f.axis == 1) j += f.side * (ny - npml);
    if (f.axis == 2) k += f.side * (nz - npml);
  }

  // pml_deriv.cpp:126-584 (one slab per face, launch order kept).
  void pml_deriv(const int FB[3]) {
    const T* w = W.data();
    const T* c = cjm.data();
    for (auto& f : faces) {
      const int ax = f.axis;
#pragma omp parallel for
      for (long long q = 0; q < f.cnt; ++q) {
        int i, j, k;
        slab_cell(f, q, i, j, k);
        int qq = ax == 0 ? i : ax == 1 ? j : k;
        long long p = at(i, j, k);
        T be = pb[ax][qq], d = pd[ax][qq], ad = d + pa[ax][qq];
        T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
        T D[9], Q[9];
        for (int v = 0; v < 9; ++v) D[v] = L(w + v * N, p, stride(ax), FB[ax]) * d;
        assemble(D, c[(3 * ax) * N + p], c[(3 * ax + 1) * N + p], c[(3 * ax + 2) * N + p], mu, lam, b, Q);
        for (int v = 0; v < 9; ++v) {
          T a = f.a[v * f.cnt + q];
          f.h[v * f.cnt + q] = (Q[v] - ad * a) * DT;
          hW[v * N + p] += -be * a * DT;
        }
      }
    }
  }

  // pml_freeSurface.cpp:17-242
  void pml_free_surface(const int FB[3]) {
    const T* w = W.data();
    const T* c = cjm.data();
    const long long S = (long long)NX * NY;
    for (auto& f : faces) {
      if (f.axis == 2) continue;
      const int ax = f.axis;
      const T* R = ax == 0 ? rdzdx.data() : rdzdy.data();
      long long plane = (long long)f.n0 * f.n1;
      long long base = (long long)(nz - 1) * plane;
      for (long long q2 = 0; q2 < plane; ++q2) {
        long long q = base + q2;
        int i, j, k;
        slab_cell(f, q, i, j, k);
        long long p = at(i, j, k);
        long long s = i + (long long)NX * j;
        T d = pd[ax][ax == 0 ? i : j];
        T mu = c[C_MU * N + p], lam = c[C_LAM * N + p];
        T V[3], Vz[3], st[6];
        for (int v = 0; v < 3; ++v) V[v] = L(w + v * N, p, stride(ax), FB[ax]) * d;
        for (int r = 0; r < 3; ++r) Vz[r] = R[(3 * r) * S + s] * V[0] + R[(3 * r + 1) * S + s] * V[1] + R[(3 * r + 2) * S + s] * V[2];
        stress_of(Vz, c[C_ZTX * N + p], c[C_ZTY * N + p], c[C_ZTZ * N + p], mu, lam, st);
        for (int v = 0; v < 6; ++v) f.h[(3 + v) * f.cnt + q] += st[v] * DT;
      }
    }
  }

  // wave_rk.cpp:18-104 and pml_rk.cpp:19-73 (the same low-storage RK4 form).
  static void rk(int irk, long long n, const T* h, T* w, T* t, T* m) {
#pragma omp parallel for
    for (long long q = 0; q < n; ++q) {
      if (irk == 0) {
        T mw = w[q];
        m[q] = mw;
        t[q] = mw + T(RK_B1) * h[q];
        w[q] = mw + T(RK_A2) * h[q];
      } else if (irk == 1) {
        t[q] += T(RK_B2) * h[q];
        w[q] = m[q] + T(RK_A3) * h[q];
      } else if (irk == 2) {
        t[q] += T(RK_B3) * h[q];
        w[q] = m[q] + h[q];
      } else {
        w[q] = t[q] + T(RK_B4) * h[q];
      }
    }
  }

  // addMoment.cpp:101-146,192-260,262-377: applied to W once per step.
  void add_source() {
    if (src_np == 0) return;
    float tmpT = (float)((it + 0.0f) * (float)DT);
    int srcIt = int(tmpT / (float)src_dt);
    if (srcIt + 1 >= src_nt) return;
    float t1 = float(srcIt) * (float)src_dt;
    T tw = T((tmpT - t1) / (float)src_dt);
    for (int p = 0; p < src_np; ++p) {
      long long cidx = src_idx[p];
      for (int q = 0; q < 6; ++q) {
        T m0 = src_rate[((size_t)srcIt * src_np + p) * 6 + q];
        T m1 = src_rate[((size_t)(srcIt + 1) * src_np + p) * 6 + q];
        T s = (m1 - m0) * tw + m0;
        W[(3 + q) * N + cidx] -= s * DT;
      }
    }
  }

  // station.cpp:303-380, PGV.cpp:62-175, solve_displacement.cpp:28-63
  // (displacement kept on the surface plane, the only part ever output).
  void observe() {
    if (fs) {
      int k = NZ - H - 1;
      const long long S = (long long)NX * NY;
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k), s = i + (long long)NX * j;
          T vx = W[p], vy = W[N + p], vz = W[2 * N + p];
          dis[s] += vx * DT;
          dis[S + s] += vy * DT;
          dis[2 * S + s] += vz * DT;
          float fx = (float)vx, fy = (float)vy, fz = (float)vz;
          float vh = std::sqrt(fx * fx + fy * fy), va = std::sqrt(fx * fx + fy * fy + fz * fz);
          long long q = (long long)(i - H) + (long long)nx * (j - H);
          if (pgv[2 * q] < vh) pgv[2 * q] = vh;
          if (pgv[2 * q + 1] < va) pgv[2 * q + 1] = va;
        }
    }
    if (it < NTrec)
      for (int s = 0; s < ns; ++s) {
        long long p = at(st_ijk[3 * s], st_ijk[3 * s + 1], st_ijk[3 * s + 2]);
        for (int v = 0; v < 9; ++v) st_wave[((size_t)s * NTrec + it) * 9 + v] = (float)W[v * N + p];
      }
  }

  // propagate.cpp:186-321 (CGFDM branch).
  void step() {
    int FB[3] = {FBTAB[it % 8][0], FBTAB[it % 8][1], FBTAB[it % 8][2]};
    add_source();
    for (int irk = 0; irk < 4; ++irk) {
      wave_deriv(FB);
      if (fs) free_surface_deriv(FB);
      pml_deriv(FB);
      if (fs) pml_free_surface(FB);
      rk(irk, 9 * N, hW.data(), W.data(), tW.data(), mW.data());
      for (auto& f : faces) rk(irk, 9 * f.cnt, f.h.data(), f.a.data(), f.t.data(), f.m.data());
      FB[0] = -FB[0]; FB[1] = -FB[1]; FB[2] = -FB[2];
    }
    observe();
    ++it;
  }
};

}  // namespace orc

#include "cgfdm_adjoint.h"

// ---------------------------------------------------------------------------
// C ABI used by the tests (ctypes).
// ---------------------------------------------------------------------------
struct OrcHandle {
  int prec;
  void* p;
};

template <class T>
static orc::Oracle<T>* make(int nx, int ny, int nz, int npml, int fs, double dh, double dt, int ntrec,
                            const float* coord, const float* medium) {
  auto* o = new orc::Oracle<T>();
  o->nx = nx; o->ny = ny; o->nz = nz;
  o->NX = nx + 2 * orc::H; o->NY = ny + 2 * orc::H; o->NZ = nz + 2 * orc::H;
  o->N = (long long)o->NX * o->NY * o->NZ;
  o->npml = npml; o->fs = fs != 0; o->dh = dh;
  o->DT = T(dt); o->rDH = T(1.0 / dh);
  o->hW.assign(9 * o->N, T(0)); o->W.assign(9 * o->N, T(0));
  o->tW.assign(9 * o->N, T(0)); o->mW.assign(9 * o->N, T(0));
  o->build_metrics(coord);
  o->build_cjm(medium);
  o->build_rdz(medium);
  o->build_pml();
  o->build_cfl(coord, medium);
  o->NTrec = ntrec;
  o->pgv.assign(2 * (size_t)nx * ny, 0.f);
  o->dis.assign(3 * (size_t)o->NX * o->NY, T(0));
  return o;
}

#define DISPATCH(h, ...)                                                      \
  do {                                                                        \
    if ((h)->prec == 8) { auto* o = (orc::Oracle<double>*)(h)->p; __VA_ARGS__; } \
    else { auto* o = (orc::Oracle<float>*)(h)->p; __VA_ARGS__; }              \
  } while (0)

extern "C" {

void* orc_create(int prec, int nx, int ny, int nz, int npml, int free_surface, double dh, double dt, int nt_record,
                 const float* coord, const float* medium) {
  auto* h = new OrcHandle();
  h->prec = prec;
  if (prec == 8) h->p = make<double>(nx, ny, nz, npml, free_surface, dh, dt, nt_record, coord, medium);
  else h->p = make<float>(nx, ny, nz, npml, free_surface, dh, dt, nt_record, coord, medium);
  return h;
}

void orc_destroy(void* hv) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, delete o);
  delete h;
}

void orc_set_source(void* hv, int np, int nt, double dt, const long long* cell, const float* rate) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, o->set_source(np, nt, dt, cell, rate));
}

void orc_set_stations(void* hv, int ns, const int* ijk) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    o->ns = ns;
    o->st_ijk.assign(ijk, ijk + 3 * ns);
    o->st_wave.assign((size_t)ns * o->NTrec * 9, 0.f);
  });
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_step(void* hv, int nsteps) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, for (int s = 0; s < nsteps; ++s) o->step());
}

// which: 0 = W, 1 = t_W, 2 = m_W, 3 = h_W ; comp 0..8 ; out has N doubles
void orc_get_field(void* hv, int which, int comp, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    auto& v = which == 0 ? o->W : which == 1 ? o->tW : which == 2 ? o->mW : o->hW;
    for (long long p = 0; p < o->N; ++p) out[p] = (double)v[comp * o->N + p];
  });
}

void orc_set_field(void* hv, int comp, const double* in) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference<decltype(o->W[0])>::type;
    for (long long p = 0; p < o->N; ++p) o->W[comp * o->N + p] = (T)in[p];
  });
}

// what: 0..9 metrics+J (double), 10..12 mu,lam,b (stored precision), 20..28 rDZ_DX, 30..38 rDZ_DY
//       40+3*axis+{0,1,2} = pml alpha, beta(inverted), d(divided)
void orc_get_setup(void* hv, int what, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    long long S = (long long)o->NX * o->NY;
    if (what < 10) for (long long p = 0; p < o->N; ++p) out[p] = o->metric_d[what * o->N + p];
    else if (what < 13) for (long long p = 0; p < o->N; ++p) out[p] = (double)o->cjm[what * o->N + p];
    else if (what >= 20 && what < 29) for (long long s = 0; s < S; ++s) out[s] = (double)o->rdzdx[(what - 20) * S + s];
    else if (what >= 30 && what < 39) for (long long s = 0; s < S; ++s) out[s] = (double)o->rdzdy[(what - 30) * S + s];
    else if (what >= 40 && what < 49) {
      int ax = (what - 40) / 3, kind = (what - 40) % 3;
      auto& v = kind == 0 ? o->pa[ax] : kind == 1 ? o->pb[ax] : o->pd[ax];
      for (size_t q = 0; q < v.size(); ++q) out[q] = (double)v[q];
    }
  });
}

// [hmin, hmax, vsmin, vsmax, vpmin, vpmax, rhomin, rhomax]
void orc_get_ranges(void* hv, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    double r[8] = {o->hmin, o->hmax, o->vsmin, o->vsmax, o->vpmin, o->vpmax, o->rhomin, o->rhomax};
    std::memcpy(out, r, sizeof(r));
  });
}

void orc_get_stations(void* hv, float* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, std::memcpy(out, o->st_wave.data(), o->st_wave.size() * sizeof(float)));
}

void orc_get_pgv(void* hv, float* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, std::memcpy(out, o->pgv.data(), o->pgv.size() * sizeof(float)));
}

// surface displacement [3][NY][NX]
void orc_get_surface_dis(void* hv, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, for (size_t q = 0; q < o->dis.size(); ++q) out[q] = (double)o->dis[q]);
}

// Aux state of face f (0-based in launch order), component v: cnt doubles
long long orc_face_count(void* hv, int f) {
  auto* h = (OrcHandle*)hv;
  long long r = 0;
  DISPATCH(h, r = f < (int)o->faces.size() ? o->faces[f].cnt : 0);
  return r;
}
void orc_get_face(void* hv, int f, int comp, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    auto& F = o->faces[f];
    for (long long q = 0; q < F.cnt; ++q) out[q] = (double)F.a[comp * F.cnt + q];
  });
}
// Right-hand side h_Aux of face f (filled by orc_rhs with boundaries)
void orc_get_face_rhs(void* hv, int f, int comp, double* out) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    auto& F = o->faces[f];
    for (long long q = 0; q < F.cnt; ++q) out[q] = (double)F.h[comp * F.cnt + q];
  });
}
void orc_set_face(void* hv, int f, int comp, const double* in) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    auto& F = o->faces[f];
    using T = std::remove_reference<decltype(F.a[0])>::type;
    for (long long q = 0; q < F.cnt; ++q) F.a[comp * F.cnt + q] = (T)in[q];
  });
}
// Axis (0..2) and side (0 low, 1 high) of face f
void orc_face_info(void* hv, int f, int* axis_side) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, {
    axis_side[0] = o->faces[f].axis;
    axis_side[1] = o->faces[f].side;
  });
}
// Source injection at the current step index, without stepping
void orc_source_only(void* hv) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, o->add_source());
}

// Single operator applications for the closed-form tests: fills h_W with the
// interior RHS (+ free surface + PML terms) for the current W without
// advancing time.
void orc_rhs(void* hv, int fb1, int fb2, int fb3, int with_boundaries) {
  auto* h = (OrcHandle*)hv;
  int FB[3] = {fb1, fb2, fb3};
  DISPATCH(h, {
    o->wave_deriv(FB);
    if (with_boundaries) {
      if (o->fs) o->free_surface_deriv(FB);
      o->pml_deriv(FB);
      if (o->fs) o->pml_free_surface(FB);
    }
  });
}

// RK combination on a raw array (for the amplification-factor test).
void orc_rk_stage(int prec, int irk, long long n, void* h, void* w, void* t, void* m) {
  if (prec == 8) orc::Oracle<double>::rk(irk, n, (double*)h, (double*)w, (double*)t, (double*)m);
  else orc::Oracle<float>::rk(irk, n, (float*)h, (float*)w, (float*)t, (float*)m);
}

int orc_step_index(void* hv) {
  auto* h = (OrcHandle*)hv;
  int r = 0;
  DISPATCH(h, r = o->it);
  return r;
}

// ---- adjoint (FP64 oracle only) --------------------------------------------
static double rnd_fill(std::vector<double>& v, unsigned long long& st) {
  double s = 0;
  for (auto& x : v) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    x = (double)((st >> 11) & ((1ULL << 53) - 1)) / (double)(1ULL << 53) - 0.5;
    s += x * x;
  }
  return s;
}

// Dot-product test of the transposed operators on random vectors:
// which = 0: stage right-hand side L with the F/B pattern of the current step;
// which = 1: the whole RK4 step (no source). out2 = {<Op X, Y>, <X, Op^T Y>}.
int orc_adjoint_dot(void* hv, int which, unsigned long long seed, double* out2) {
  auto* h = (OrcHandle*)hv;
  if (h->prec != 8) return 2;
  auto* o = (orc::Oracle<double>*)h->p;
  orc::Adjoint<double> ad(*o);
  unsigned long long st = seed * 2654435761ULL + 1;
  orc::AdjVec<double> X = ad.zeros(), Y = ad.zeros();
  rnd_fill(X.W, st);
  rnd_fill(Y.W, st);
  for (auto& a : X.A) rnd_fill(a, st);
  for (auto& a : Y.A) rnd_fill(a, st);
  int FB[3] = {orc::FBTAB[o->it % 8][0], orc::FBTAB[o->it % 8][1], orc::FBTAB[o->it % 8][2]};
  orc::AdjVec<double> OX, OTY = ad.zeros();
  if (which == 0) {
    OX = ad.rhs(FB, X);
    ad.rhs_T(FB, Y, OTY, nullptr, nullptr, nullptr, nullptr);
  } else {
    orc::AdjVec<double> Xs[4];
    ad.stages(FB, X, Xs);
    // X' = X0 + sum b_s h_s
    OX = X;
    const double bw[4] = {orc::RK_B1, orc::RK_B2, orc::RK_B3, orc::RK_B4};
    int F[3] = {FB[0], FB[1], FB[2]};
    for (int s = 0; s < 4; ++s) {
      orc::AdjVec<double> hs = ad.rhs(F, Xs[s]);
      ad.axpy(bw[s], hs, OX);
      F[0] = -F[0]; F[1] = -F[1]; F[2] = -F[2];
    }
    OTY = Y;
    ad.step_T(o->it, X, OTY, nullptr, nullptr, nullptr);
  }
  out2[0] = ad.dot(OX, Y);
  out2[1] = ad.dot(X, OTY);
  return 0;
}

void orc_set_step_index(void* hv, int it) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, o->it = it);
}

void orc_set_medium_params(void* hv, const double* mu, const double* lam, const double* b) {
  auto* h = (OrcHandle*)hv;
  DISPATCH(h, o->set_medium_params(mu, lam, b));
}

// Misfit J = 1/2 sum_it sum_s sum_{c<3} (W_c(station s) after step it - data[s][it][c])^2
// over nsteps steps from the current state, and its gradient by the adjoint:
// gmu, glam, gb per padded cell (may be null) and grate per input source point
// [np][nt][6] in the units of the table given to set_source (may be null).
// The oracle is left at the state after the forward run. Returns J (NaN on error).
double orc_adjoint_misfit(void* hv, int nsteps, const double* data, double* gmu, double* glam, double* gb,
                          double* grate) {
  auto* h = (OrcHandle*)hv;
  if (h->prec != 8) return std::nan("");
  auto* o = (orc::Oracle<double>*)h->p;
  orc::Adjoint<double> ad(*o);
  const long long N = o->N;
  const int it0 = o->it;
  std::vector<orc::AdjVec<double>> Xpre;
  std::vector<double> rec((size_t)o->ns * nsteps * 3);
  std::vector<long long> scell(o->ns);
  for (int s = 0; s < o->ns; ++s) scell[s] = o->at(o->st_ijk[3 * s], o->st_ijk[3 * s + 1], o->st_ijk[3 * s + 2]);
  double J = 0;
  for (int n = 0; n < nsteps; ++n) {
    Xpre.push_back(ad.state());
    o->step();
    for (int s = 0; s < o->ns; ++s)
      for (int c = 0; c < 3; ++c) {
        double r = o->W[c * N + scell[s]] - data[((size_t)s * nsteps + n) * 3 + c];
        rec[((size_t)s * nsteps + n) * 3 + c] = r;
        J += 0.5 * r * r;
      }
  }
  orc::AdjVec<double> endstate = ad.state();
  std::vector<double> gm(N, 0.0), gl(N, 0.0), gbb(N, 0.0);
  std::vector<double> gscaled((size_t)o->src_np * o->src_nt * 6, 0.0);
  orc::AdjVec<double> mu = ad.zeros();
  for (int n = nsteps - 1; n >= 0; --n) {
    for (int s = 0; s < o->ns; ++s)
      for (int c = 0; c < 3; ++c) mu.W[c * N + scell[s]] += rec[((size_t)s * nsteps + n) * 3 + c];
    // state after the source injection of step n
    ad.load(Xpre[n]);
    o->it = it0 + n;
    o->add_source();
    orc::AdjVec<double> X0 = ad.state();
    ad.step_T(it0 + n, X0, mu, gm.data(), gl.data(), gbb.data());
    // source: W[3+q](c) -= ((m1 - m0) tw + m0) DT
    if (o->src_np) {
      float tmpT = (float)((o->it + 0.0f) * (float)o->DT);
      int srcIt = int(tmpT / (float)o->src_dt);
      if (srcIt + 1 < o->src_nt) {
        float t1 = float(srcIt) * (float)o->src_dt;
        double tw = (double)(float)((tmpT - t1) / (float)o->src_dt);
        for (int p = 0; p < o->src_np; ++p)
          for (int q = 0; q < 6; ++q) {
            double gs = -o->DT * mu.W[(3 + q) * N + o->src_idx[p]];
            gscaled[((size_t)srcIt * o->src_np + p) * 6 + q] += (1.0 - tw) * gs;
            gscaled[((size_t)(srcIt + 1) * o->src_np + p) * 6 + q] += tw * gs;
          }
      }
    }
  }
  // scaled rate = raw * mu(cell) / (J DH^3)
  for (int p = 0; p < o->src_np; ++p) {
    long long c = o->src_idx[p];
    double muc = o->cjm[orc::C_MU * N + c];
    for (int t = 0; t < o->src_nt; ++t)
      for (int q = 0; q < 6; ++q) {
        size_t e = ((size_t)t * o->src_np + p) * 6 + q;
        gm[c] += gscaled[e] * o->src_raw[e] * o->src_vol[p];
      }
  }
  if (grate)
    for (size_t pi = 0; pi < o->src_of_point.size(); ++pi) {
      int p = o->src_of_point[pi];
      double muc = o->cjm[orc::C_MU * N + o->src_idx[p]];
      for (int t = 0; t < o->src_nt; ++t)
        for (int q = 0; q < 6; ++q)
          grate[((size_t)pi * o->src_nt + t) * 6 + q] = gscaled[((size_t)t * o->src_np + p) * 6 + q] * muc * o->src_vol[p];
    }
  if (gmu) std::memcpy(gmu, gm.data(), sizeof(double) * N);
  if (glam) std::memcpy(glam, gl.data(), sizeof(double) * N);
  if (gb) std::memcpy(gb, gbb.data(), sizeof(double) * N);
  ad.load(endstate);
  o->it = it0 + nsteps;
  return J;
}
}
