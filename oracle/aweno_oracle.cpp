// CPU oracle for the ME-AWENO (SCFDM build) time step of CGFDM3D-EQR-SC.
//
// TEST INFRASTRUCTURE ONLY (see cgfdm_oracle.cpp). Restates the reference's
// SCFDM path: covariant/contravariant metrics by central differences
// (contravariant.cpp:81-177), WENO5-JS reconstruction and Lax-Friedrichs
// alternative flux (alternative_flux_FD.cpp:43-174), the flux-difference
// update with the high-order correction (alternative_flux_FD.cpp:1216-1435),
// TVD-RK3 (wave_rk_tvd.cpp:18-78), the characteristic free surface
// (freeSurface.cpp:41-194), the exponential sponge (exp_decay_absorb.cpp:20-106)
// and the strain-rate source (addMoment.cpp:192-239). The float instantiation
// keeps the reference's FP64 promotions (WENO weights, TVD combination, 1/J).
//
// State: W = (rho vx, rho vy, rho vz, exx, eyy, ezz, gyz, gxz, gxy).
// Layout: SoA over the padded grid, x fastest, HALO = 3.
#include <cmath>
#include <cstring>
#include <map>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace awo {

constexpr int H = 3;
// CJM planes of the SCFDM build (CJM.cpp:107-147): 0..8 J*(xi_x..zt_z)
// cofactors, 9 J, 10 mu, 11 lambda, 12 1/rho, 13..18 x_xi,y_xi,z_xi,x_et,y_et,z_et
enum { NC = 19 };

template <class T>
struct Aweno {
  int nx, ny, nz, NX, NY, NZ;
  long long N;
  bool fs;
  double dh;
  T DT, rDH;
  T alpha;  // Lax-Friedrichs speed = global Vp max (calc_CFL.cpp:321)
  std::vector<T> hW, W, mW, F;
  std::vector<T> cjm;
  double ranges[8] = {0};
  int src_np = 0, src_nt = 0;
  double src_dt = 0;
  std::vector<long long> src_idx;
  std::vector<T> src_rate;
  int ns = 0, NTrec = 0;
  std::vector<int> st_ijk;
  std::vector<float> st_wave, pgv;
  std::vector<T> dis;          // surface displacement [3][ny][nx]
  std::vector<double> jacd;    // J in double, for the source scaling
  int it = 0;

  long long at(int i, int j, int k) const { return i + (long long)NX * (j + (long long)NY * k); }
  T c(int comp, long long p) const { return cjm[(size_t)comp * N + p]; }

  // contravariant.cpp:81-177 (second-order central covariant basis, J = det,
  // cofactors J*xi_x ...); ranges as in the reference kernel
  void build_metrics(const float* coord) {
    cjm.assign((size_t)NC * N, T(0));
    std::vector<double> cov(9 * (size_t)N, 0.0), jac(N, 0.0);
    const double rdh = 1.0 / dh;
    const float* X = coord;
    const float* Y = coord + N;
    const float* Z = coord + 2 * N;
    const float* C[3] = {X, Y, Z};
    long long st[3] = {1, NX, (long long)NX * NY};
#pragma omp parallel for collapse(2)
    for (int k = 1; k < NZ - 1; ++k)
      for (int j = 1; j < NY - 1; ++j)
        for (int i = 1; i < NX - 1; ++i) {
          long long p = at(i, j, k);
          double d[9];
          for (int dir = 0; dir < 3; ++dir)
            for (int a = 0; a < 3; ++a)
              d[3 * dir + a] = 0.5 * ((double)C[a][p + st[dir]] - (double)C[a][p - st[dir]]) * rdh;
          // d = (x_xi, y_xi, z_xi, x_et, y_et, z_et, x_zt, y_zt, z_zt)
          for (int q = 0; q < 9; ++q) cov[(size_t)q * N + p] = d[q];
          jac[p] = d[0] * d[4] * d[8] + d[1] * d[5] * d[6] + d[2] * d[3] * d[7] - d[2] * d[4] * d[6] -
                   d[1] * d[3] * d[8] - d[0] * d[7] * d[5];
        }
    jacd = jac;
    for (int k = 1; k < NZ - 1; ++k)
      for (int j = 1; j < NY - 1; ++j)
        for (int i = 1; i < NX - 1; ++i) {
          long long p = at(i, j, k);
          cjm[9 * (size_t)N + p] = T(jac[p]);
          for (int q = 0; q < 6; ++q) cjm[(size_t)(13 + q) * N + p] = T(cov[(size_t)q * N + p]);
          if (i >= 2 && i < NX - 2 && j >= 2 && j < NY - 2 && k >= 2 && k < NZ - 2) {
            double d[9];
            for (int q = 0; q < 9; ++q) d[q] = cov[(size_t)q * N + p];
            double cof[9] = {d[4] * d[8] - d[5] * d[7], d[5] * d[6] - d[3] * d[8], d[3] * d[7] - d[4] * d[6],
                             d[7] * d[2] - d[8] * d[1], d[8] * d[0] - d[6] * d[2], d[6] * d[1] - d[7] * d[0],
                             d[1] * d[5] - d[2] * d[4], d[2] * d[3] - d[0] * d[5], d[0] * d[4] - d[1] * d[3]};
            for (int q = 0; q < 9; ++q) cjm[(size_t)q * N + p] = T(cof[q]);
          }
        }
  }

  void build_medium(const float* medium) {
    for (long long p = 0; p < N; ++p) {
      double vs = medium[p], vp = medium[N + p], rho = medium[2 * N + p];
      cjm[10 * (size_t)N + p] = T(rho * vs * vs);
      cjm[11 * (size_t)N + p] = T(rho * (vp * vp - 2.0 * vs * vs));
      cjm[12 * (size_t)N + p] = T(1.0 / rho);
    }
    double lo = 1e300, hi = -1e300;
    for (long long p = 0; p < N; ++p) { lo = std::min(lo, (double)medium[N + p]); hi = std::max(hi, (double)medium[N + p]); }
    ranges[4] = lo;
    ranges[5] = hi;
    alpha = T((float)hi);
  }

  // WENO5-JS value at x_{i+1/2} from (u1..u5) = (i-2..i+2) (alternative_flux_FD.cpp:43-59);
  // the nonlinear weights are evaluated in double as the reference's 1e-6 literal makes them
  static T weno(T u1, T u2, T u3, T u4, T u5) {
    T b1 = T(0.3333333333) * (T(10) * u3 * u3 - T(31) * u3 * u4 + T(25) * u4 * u4 + T(11) * u3 * u5 - T(19) * u4 * u5 + T(4) * u5 * u5);
    T b2 = T(0.3333333333) * (T(4) * u2 * u2 - T(13) * u2 * u3 + T(13) * u3 * u3 + T(5) * u2 * u4 - T(13) * u3 * u4 + T(4) * u4 * u4);
    T b3 = T(0.3333333333) * (T(4) * u1 * u1 - T(19) * u1 * u2 + T(25) * u2 * u2 + T(11) * u1 * u3 - T(31) * u2 * u3 + T(10) * u3 * u3);
    T w1 = T(0.3125 / (((double)b1 + 1e-6) * ((double)b1 + 1e-6)));
    T w2 = T(0.625 / (((double)b2 + 1e-6) * ((double)b2 + 1e-6)));
    T w3 = T(0.0625 / (((double)b3 + 1e-6) * ((double)b3 + 1e-6)));
    T tot = T(1) / (w1 + w2 + w3);
    return (w1 * tot) * (T(0.375) * u3 + T(0.75) * u4 - T(0.125) * u5) + (w2 * tot) * (T(-0.125) * u2 + T(0.75) * u3 + T(0.375) * u4) +
           (w3 * tot) * (T(0.375) * u1 - T(1.25) * u2 + T(1.875) * u3);
  }

  // conservative flux along a direction with cofactor metric (n1,n2,n3)
  static void flux(const T* u, T n1, T n2, T n3, T mu, T lam, T b, T* f) {
    T l2m = lam + T(2) * mu;
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

  static T d2(T u1, T u2, T u3, T u4, T u5) { return T(1.0 / 12.0) * (-u1 + T(16) * u2 - T(30) * u3 + T(16) * u4 - u5); }
  static T d4(T u1, T u2, T u3, T u4, T u5) { return u1 - T(4) * u2 + T(6) * u3 - T(4) * u4 + u5; }
  static T d2f(const T* u) { return T(1.0 / 48.0) * (T(-5) * u[0] + T(39) * u[1] - T(34) * u[2] - T(34) * u[3] + T(39) * u[4] - T(5) * u[5]); }
  static T d4f(const T* u) { return T(0.5) * (u[0] - T(3) * u[1] + T(2) * u[2] + T(2) * u[3] - T(3) * u[4] + u[5]); }

  long long stride(int d) const { return d == 0 ? 1 : d == 1 ? (long long)NX : (long long)NX * NY; }

  // Face fluxes of one direction at faces q+1/2, q in [H-1, n+H) along d, the
  // other two indices interior (wave_deriv_alternative_flux_FD_{x,y,z}); every
  // other entry of the flux buffer is zero.
  void face_fluxes(int d) {
    std::fill(F.begin(), F.end(), T(0));
    const long long s = stride(d);
    int lo[3] = {H, H, H}, hi[3] = {NX - H, NY - H, NZ - H};
    lo[d] = H - 1;
#pragma omp parallel for collapse(2)
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) {
          long long p = at(i, j, k);
          T n1 = c(3 * d, p), n2 = c(3 * d + 1, p), n3 = c(3 * d + 2, p);
          T mu = c(10, p), lam = c(11, p), b = c(12, p);
          T um[9], up[9], fm[9], fp[9];
          for (int v = 0; v < 9; ++v) {
            const T* w = W.data() + (size_t)v * N;
            um[v] = weno(w[p - 2 * s], w[p - s], w[p], w[p + s], w[p + 2 * s]);
            up[v] = weno(w[p + 3 * s], w[p + 2 * s], w[p + s], w[p], w[p - s]);
          }
          flux(um, n1, n2, n3, mu, lam, b, fm);
          flux(up, n1, n2, n3, mu, lam, b, fp);
          for (int v = 0; v < 9; ++v) F[(size_t)v * N + p] = T(0.5) * (fp[v] + fm[v] - alpha * (up[v] - um[v]));
        }
  }

  // Flux difference with the alternative-flux correction (cal_du_{x,y,z}):
  // direction 0 assigns h_W, 1 and 2 accumulate.
  void flux_difference(int d) {
    const long long s = stride(d);
    const int ktop = NZ - H;  // top three layers are [ktop-3, ktop)
#pragma omp parallel for collapse(2)
    for (int k = H; k < NZ - H; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          T jac = T(1.0 / (double)c(9, p));
          T cd = DT * rDH * jac;
          if (k >= ktop - 3) {
            T hu[7][9];
            for (int n = 0; n < 7; ++n) {
              long long q = p + (n - 3) * s;
              T u[9];
              for (int v = 0; v < 9; ++v) u[v] = W[(size_t)v * N + q];
              flux(u, c(3 * d, q), c(3 * d + 1, q), c(3 * d + 2, q), c(10, q), c(11, q), c(12, q), hu[n]);
            }
            for (int v = 0; v < 9; ++v) {
              T a6[6], b6[6];
              for (int t = 0; t < 6; ++t) { a6[t] = hu[t + 1][v]; b6[t] = hu[t][v]; }
              T fr = F[(size_t)v * N + p] - T(1.0 / 24.0) * d2f(a6) + T(7.0 / 5760.0) * d4f(a6);
              T fl = F[(size_t)v * N + p - s] - T(1.0 / 24.0) * d2f(b6) + T(7.0 / 5760.0) * d4f(b6);
              T val = cd * (-(fr - fl));
              if (d == 0) hW[(size_t)v * N + p] = val;
              else hW[(size_t)v * N + p] += val;
            }
          } else {
            for (int v = 0; v < 9; ++v) {
              const T* f = F.data() + (size_t)v * N;
              T f_3 = f[p - 3 * s], f_2 = f[p - 2 * s], f_1 = f[p - s], f0 = f[p], f1 = f[p + s], f2 = f[p + 2 * s];
              T fr = f0 - T(1.0 / 24.0) * d2(f_2, f_1, f0, f1, f2) + T(7.0 / 5760.0) * d4(f_2, f_1, f0, f1, f2);
              T fl = f_1 - T(1.0 / 24.0) * d2(f_3, f_2, f_1, f0, f1) + T(7.0 / 5760.0) * d4(f_3, f_2, f_1, f0, f1);
              T val = cd * (-(fr - fl));
              if (d == 0) hW[(size_t)v * N + p] = val;
              else hW[(size_t)v * N + p] += val;
            }
          }
        }
  }

  void rhs() {
    for (int d = 0; d < 3; ++d) {
      face_fluxes(d);
      flux_difference(d);
    }
  }

  // TVD-RK3 in the reference's double arithmetic (wave_rk_tvd.cpp:18-78)
  void tvd(int irk) {
    const long long n = 9 * N;
#pragma omp parallel for
    for (long long q = 0; q < n; ++q) {
      if (irk == 0) {
        mW[q] = W[q];
        W[q] = W[q] + hW[q];
      } else if (irk == 1) {
        W[q] = T(0.75 * (double)mW[q] + 0.25 * (double)W[q] + 0.25 * (double)hW[q]);
      } else {
        W[q] = T((1.0 / 3.0) * (double)mW[q] + (2.0 / 3.0) * (double)W[q] + (2.0 / 3.0) * (double)hW[q]);
      }
    }
  }

  // Characteristic free surface at the top layer and cubic extrapolation into
  // the ghost layers (freeSurface.cpp:41-194). The stress tensor and the
  // velocity are expressed in the frame (n, s, t) of the surface normal and the
  // two tangents x_xi, x_eta: with B = [n s t] the frame components are
  // B^-1 S B^-T and B^-1 v, whose rows are (s x t, t x n, n x s) / det(B).
  void char_free_surface() {
    const int k = NZ - H - 1;
    for (int j = H; j < NY - H; ++j)
      for (int i = H; i < NX - H; ++i) {
        long long p = at(i, j, k);
        T mu = c(10, p), lam = c(11, p), b = c(12, p);
        T rho = T(1) / b;
        T vs = std::sqrt(mu * b), vp = std::sqrt((lam + T(2) * mu) * b);
        T J = c(9, p);
        T nv[3] = {c(6, p) / J, c(7, p) / J, c(8, p) / J};
        T nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        T sn = std::sqrt(c(13, p) * c(13, p) + c(14, p) * c(14, p) + c(15, p) * c(15, p));
        T tn = std::sqrt(c(16, p) * c(16, p) + c(17, p) * c(17, p) + c(18, p) * c(18, p));
        T B[3][3];  // B[a] = a-th frame vector (n, s, t)
        for (int a = 0; a < 3; ++a) {
          B[0][a] = nv[a] / nn;
          B[1][a] = c(13 + a, p) / sn;
          B[2][a] = c(16 + a, p) / tn;
        }
        T R[3][3];  // rows of det(B) * B^-1
        for (int a = 0; a < 3; ++a) {
          const T* u = B[(a + 1) % 3];
          const T* v = B[(a + 2) % 3];
          R[a][0] = u[1] * v[2] - u[2] * v[1];
          R[a][1] = u[2] * v[0] - u[0] * v[2];
          R[a][2] = u[0] * v[1] - u[1] * v[0];
        }
        T det = B[0][0] * R[0][0] + B[0][1] * R[0][1] + B[0][2] * R[0][2];
        T w[9];
        for (int q = 0; q < 9; ++q) w[q] = W[(size_t)q * N + p];
        T l2m = lam + T(2) * mu;
        T S[3][3];
        S[0][0] = lam * w[4] + lam * w[5] + w[3] * l2m;
        S[1][1] = lam * w[3] + lam * w[5] + w[4] * l2m;
        S[2][2] = lam * w[3] + lam * w[4] + w[5] * l2m;
        S[0][1] = S[1][0] = mu * w[8];
        S[0][2] = S[2][0] = mu * w[7];
        S[1][2] = S[2][1] = mu * w[6];
        T v[3] = {w[0] * b, w[1] * b, w[2] * b};
        T Sf[3][3], vf[3];
        for (int a = 0; a < 3; ++a) {
          T RS[3];
          for (int q = 0; q < 3; ++q) RS[q] = R[a][0] * S[0][q] + R[a][1] * S[1][q] + R[a][2] * S[2][q];
          for (int c2 = 0; c2 < 3; ++c2) Sf[a][c2] = (RS[0] * R[c2][0] + RS[1] * R[c2][1] + RS[2] * R[c2][2]) / (det * det);
          vf[a] = (R[a][0] * v[0] + R[a][1] * v[1] + R[a][2] * v[2]) / det;
        }
        // traction-free characteristic condition (freeSurface.cpp:145-153)
        Sf[1][1] -= Sf[0][0] * lam / l2m;
        Sf[2][2] -= Sf[0][0] * lam / l2m;
        vf[1] -= Sf[0][1] / (vs * rho);
        vf[2] -= Sf[0][2] / (vs * rho);
        vf[0] -= Sf[0][0] / (vp * rho);
        Sf[0][0] = T(0);
        Sf[0][1] = Sf[1][0] = T(0);
        Sf[0][2] = Sf[2][0] = T(0);
        // back to the Cartesian frame: S = B S' B^T, v = B v'
        for (int r = 0; r < 3; ++r) {
          for (int q = 0; q < 3; ++q) {
            T acc = T(0);
            for (int a = 0; a < 3; ++a)
              for (int c2 = 0; c2 < 3; ++c2) acc += Sf[a][c2] * B[a][r] * B[c2][q];
            S[r][q] = acc;
          }
          v[r] = B[0][r] * vf[0] + B[1][r] * vf[1] + B[2][r] * vf[2];
        }
        // physical -> conservative (compliance, freeSurface.cpp:167-177)
        T cc = T(1) / (T(2) * mu * (T(3) * lam + T(2) * mu));
        T o[9];
        o[0] = v[0] * rho;
        o[1] = v[1] * rho;
        o[2] = v[2] * rho;
        o[3] = (T(2) * lam * S[0][0] - lam * S[1][1] - lam * S[2][2] + T(2) * mu * S[0][0]) * cc;
        o[4] = -(lam * S[0][0] - T(2) * lam * S[1][1] + lam * S[2][2] - T(2) * mu * S[1][1]) * cc;
        o[5] = -(lam * S[0][0] + lam * S[1][1] - T(2) * lam * S[2][2] - T(2) * mu * S[2][2]) * cc;
        T rm = T(1) / mu;
        o[6] = S[1][2] * rm;
        o[7] = S[0][2] * rm;
        o[8] = S[0][1] * rm;
        const long long S2 = (long long)NX * NY;
        for (int q = 0; q < 9; ++q) {
          T* f = W.data() + (size_t)q * N;
          f[p] = o[q];
          for (int h = 1; h <= H; ++h) {
            long long g = p + h * S2;
            f[g] = T(4) * f[g - S2] - T(6) * f[g - 2 * S2] + T(4) * f[g - 3 * S2] - f[g - 4 * S2];
          }
        }
      }
  }

  // exp_decay_absorb.cpp:20-106 (20 layers, alpha = 0.015, one domain): the
  // factors of the sides a cell lies in are applied one after another in the
  // order x-low, x-high, y-low, y-high, z-low[, z-high without free surface]
  void sponge() {
    const int L = 20;
    const T a = T(0.015);
    auto G = [&](int p) { return std::exp(-(a * T(p)) * (a * T(p))); };
#pragma omp parallel for collapse(2)
    for (int k = H; k < NZ - H; ++k)
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          int I = i - H, Jj = j - H, K = k - H;
          T g[6];
          int n = 0;
          if (I <= L) g[n++] = G(L - I);
          if (I >= nx - L) g[n++] = G(L - (nx - I));
          if (Jj <= L) g[n++] = G(L - Jj);
          if (Jj >= ny - L) g[n++] = G(L - (ny - Jj));
          if (K <= L) g[n++] = G(L - K);
          if (!fs && K >= nz - L) g[n++] = G(L - (nz - K));
          long long p = at(i, j, k);
          for (int q = 0; q < n; ++q)
            for (int v = 0; v < 9; ++v) W[(size_t)v * N + p] *= g[q];
        }
  }

  void add_source() {
    if (src_np == 0) return;
    float tmpT = (float)((it + 0.0f) * (float)DT);
    int srcIt = int(tmpT / (float)src_dt);
    if (srcIt + 1 >= src_nt) return;
    float t1 = float(srcIt) * (float)src_dt;
    T tw = T((tmpT - t1) / (float)src_dt);
    for (int p = 0; p < src_np; ++p) {
      long long cidx = src_idx[p];
      T s[6];
      for (int q = 0; q < 6; ++q) {
        T m0 = src_rate[((size_t)srcIt * src_np + p) * 6 + q];
        T m1 = src_rate[((size_t)(srcIt + 1) * src_np + p) * 6 + q];
        s[q] = (m1 - m0) * tw + m0;
      }
      T mu = c(10, cidx), lam = c(11, cidx);
      T den = T(2) * (mu * mu) + T(3) * lam * mu;
      T e[6];
      e[0] = (s[0] * (lam + mu)) / den - (lam * s[1]) / (T(2) * den) - (lam * s[2]) / (T(2) * den);
      e[1] = (s[1] * (lam + mu)) / den - (lam * s[0]) / (T(2) * den) - (lam * s[2]) / (T(2) * den);
      e[2] = (s[2] * (lam + mu)) / den - (lam * s[0]) / (T(2) * den) - (lam * s[1]) / (T(2) * den);
      e[3] = s[5] / mu;
      e[4] = s[4] / mu;
      e[5] = s[3] / mu;
      for (int q = 0; q < 6; ++q) W[(size_t)(3 + q) * N + cidx] -= e[q] * DT;
    }
  }

  void physical(long long p, T* o) const {
    T u[9];
    for (int v = 0; v < 9; ++v) u[v] = W[(size_t)v * N + p];
    T mu = c(10, p), lam = c(11, p), b = c(12, p);
    o[0] = u[0] * b;
    o[1] = u[1] * b;
    o[2] = u[2] * b;
    o[3] = lam * u[4] + lam * u[5] + u[3] * (lam + T(2) * mu);
    o[4] = lam * u[3] + lam * u[5] + u[4] * (lam + T(2) * mu);
    o[5] = lam * u[3] + lam * u[4] + u[5] * (lam + T(2) * mu);
    o[6] = mu * u[8];
    o[7] = mu * u[7];
    o[8] = mu * u[6];
  }

  // station.cpp:303-380, PGV.cpp:62-175 and solve_displacement.cpp (SCFDM branches)
  void observe() {
    if (fs) {
      int k = NZ - H - 1;
      for (int j = H; j < NY - H; ++j)
        for (int i = H; i < NX - H; ++i) {
          long long p = at(i, j, k);
          T b = c(12, p);
          float fx = (float)(W[p] * b), fy = (float)(W[N + p] * b), fz = (float)(W[2 * N + p] * b);
          float vh = std::sqrt(fx * fx + fy * fy), va = std::sqrt(fx * fx + fy * fy + fz * fz);
          long long q = (long long)(i - H) + (long long)nx * (j - H);
          // solve_displacement.cpp:25-45, kept on the surface plane
          for (int a = 0; a < 3; ++a) dis[(size_t)a * nx * ny + q] += (W[(size_t)a * N + p] * b) * DT;
          if (pgv[2 * q] < vh) pgv[2 * q] = vh;
          if (pgv[2 * q + 1] < va) pgv[2 * q + 1] = va;
        }
    }
    if (it < NTrec)
      for (int s = 0; s < ns; ++s) {
        T o[9];
        physical(at(st_ijk[3 * s], st_ijk[3 * s + 1], st_ijk[3 * s + 2]), o);
        for (int v = 0; v < 9; ++v) st_wave[((size_t)s * NTrec + it) * 9 + v] = (float)o[v];
      }
  }

  // propagate.cpp:186-321, SCFDM branch
  void step() {
    add_source();
    for (int irk = 0; irk < 3; ++irk) {
      rhs();
      tvd(irk);
    }
    if (fs) char_free_surface();
    sponge();
    observe();
    ++it;
  }
};

}  // namespace awo

struct AwoHandle {
  int prec;
  void* p;
};

#define ADISPATCH(h, ...)                                                          \
  do {                                                                             \
    if ((h)->prec == 8) { auto* o = (awo::Aweno<double>*)(h)->p; __VA_ARGS__; }    \
    else { auto* o = (awo::Aweno<float>*)(h)->p; __VA_ARGS__; }                    \
  } while (0)

template <class T>
static awo::Aweno<T>* amake(int nx, int ny, int nz, int fs, double dh, double dt, int ntrec, const float* coord,
                             const float* medium) {
  auto* o = new awo::Aweno<T>();
  o->nx = nx; o->ny = ny; o->nz = nz;
  o->NX = nx + 6; o->NY = ny + 6; o->NZ = nz + 6;
  o->N = (long long)o->NX * o->NY * o->NZ;
  o->fs = fs != 0;
  o->dh = dh;
  o->DT = T(dt);
  o->rDH = T(1.0 / dh);
  o->hW.assign(9 * o->N, T(0));
  o->W.assign(9 * o->N, T(0));
  o->mW.assign(9 * o->N, T(0));
  o->F.assign(9 * o->N, T(0));
  o->build_metrics(coord);
  o->build_medium(medium);
  o->NTrec = ntrec;
  o->pgv.assign(2 * (size_t)nx * ny, 0.f);
  o->dis.assign(3 * (size_t)nx * ny, T(0));
  return o;
}

extern "C" {

void* awo_create(int prec, int nx, int ny, int nz, int free_surface, double dh, double dt, int nt_record,
                 const float* coord, const float* medium) {
  auto* h = new AwoHandle();
  h->prec = prec;
  if (prec == 8) h->p = amake<double>(nx, ny, nz, free_surface, dh, dt, nt_record, coord, medium);
  else h->p = amake<float>(nx, ny, nz, free_surface, dh, dt, nt_record, coord, medium);
  return h;
}

void awo_destroy(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, delete o);
  delete h;
}

void awo_set_source(void* hv, int np, int nt, double dt, const long long* cell, const float* rate) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, {
    using T = std::remove_reference<decltype(o->W[0])>::type;
    std::map<long long, std::vector<double>> acc;
    for (int p = 0; p < np; ++p) {
      auto& v = acc[cell[p]];
      if (v.empty()) v.assign((size_t)nt * 6, 0.0);
      for (long long q = 0; q < (long long)nt * 6; ++q) v[q] += rate[(long long)p * nt * 6 + q];
    }
    o->src_np = (int)acc.size();
    o->src_nt = nt;
    o->src_dt = dt;
    o->src_idx.clear();
    o->src_rate.assign((size_t)o->src_np * nt * 6, T(0));
    int J = 0;
    for (auto& kv : acc) {
      long long c = kv.first;
      o->src_idx.push_back(c);
      double jac = o->jacd[c], mu = (double)o->c(10, c);
      double V = 1.0 / (jac * o->dh * o->dh * o->dh);
      for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 6; ++q) o->src_rate[((size_t)t * o->src_np + J) * 6 + q] = T(kv.second[(size_t)t * 6 + q] * mu * V);
      ++J;
    }
  });
}

void awo_set_stations(void* hv, int ns, const int* ijk) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, {
    o->ns = ns;
    o->st_ijk.assign(ijk, ijk + 3 * ns);
    o->st_wave.assign((size_t)ns * o->NTrec * 9, 0.f);
  });
}

void awo_step(void* hv, int nsteps) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, for (int s = 0; s < nsteps; ++s) o->step());
}

// which: 0 = W, 1 = h_W, 2 = flux buffer; comp 0..8
void awo_get_field(void* hv, int which, int comp, double* out) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, {
    auto& v = which == 0 ? o->W : which == 1 ? o->hW : o->F;
    for (long long p = 0; p < o->N; ++p) out[p] = (double)v[(size_t)comp * o->N + p];
  });
}

void awo_set_field(void* hv, int comp, const double* in) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, {
    using T = std::remove_reference<decltype(o->W[0])>::type;
    for (long long p = 0; p < o->N; ++p) o->W[(size_t)comp * o->N + p] = (T)in[p];
  });
}

// CJM plane 0..18
void awo_get_setup(void* hv, int what, double* out) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, for (long long p = 0; p < o->N; ++p) out[p] = (double)o->c(what, p));
}

void awo_rhs(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, o->rhs());
}

// right-hand side of one direction only (0 = x, 1 = y, 2 = z), assigned to h_W
void awo_rhs_dir(void* hv, int d) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, {
    std::fill(o->hW.begin(), o->hW.end(), 0);  // y and z accumulate onto h_W
    o->face_fluxes(d);
    o->flux_difference(d);
  });
}

// injection of the source at the current step index, without stepping
void awo_source_only(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, o->add_source());
}

// observers (stations, PGV, displacement) at the current step index
void awo_observe(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, o->observe());
}

void awo_sponge(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, o->sponge());
}

// one TVD-RK3 stage on caller arrays (wave_rk_tvd.cpp:18-78)
void awo_tvd_stage(int prec, int irk, long long n, double* hW, double* W, double* mW) {
  if (prec == 8) {
    awo::Aweno<double> o;
    o.N = 0;
    o.hW.assign(hW, hW + n);
    o.W.assign(W, W + n);
    o.mW.assign(mW, mW + n);
    o.N = n / 9;
    o.tvd(irk);
    std::copy(o.W.begin(), o.W.end(), W);
    std::copy(o.mW.begin(), o.mW.end(), mW);
  } else {
    awo::Aweno<float> o;
    o.hW.assign(hW, hW + n);
    o.W.assign(W, W + n);
    o.mW.assign(mW, mW + n);
    o.N = n / 9;
    o.tvd(irk);
    for (long long q = 0; q < n; ++q) { W[q] = o.W[q]; mW[q] = o.mW[q]; }
  }
}

void awo_char_free_surface(void* hv) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, o->char_free_surface());
}

double awo_weno(double u1, double u2, double u3, double u4, double u5, int prec) {
  if (prec == 8) return awo::Aweno<double>::weno(u1, u2, u3, u4, u5);
  return (double)awo::Aweno<float>::weno((float)u1, (float)u2, (float)u3, (float)u4, (float)u5);
}

void awo_get_stations(void* hv, float* out) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, std::memcpy(out, o->st_wave.data(), o->st_wave.size() * sizeof(float)));
}

void awo_get_pgv(void* hv, float* out) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, std::memcpy(out, o->pgv.data(), o->pgv.size() * sizeof(float)));
}

void awo_get_surface_dis(void* hv, float* out) {
  auto* h = (AwoHandle*)hv;
  ADISPATCH(h, for (size_t q = 0; q < o->dis.size(); ++q) out[q] = (float)o->dis[q]);
}

int awo_step_index(void* hv) {
  auto* h = (AwoHandle*)hv;
  int r = 0;
  ADISPATCH(h, r = o->it);
  return r;
}

double awo_alpha(void* hv) {
  auto* h = (AwoHandle*)hv;
  double r = 0;
  ADISPATCH(h, r = (double)o->alpha);
  return r;
}
}
