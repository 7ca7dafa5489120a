// Discrete adjoint of the CGFDM step in the CPU oracle (FP64).
//
// TEST INFRASTRUCTURE ONLY (see cgfdm_oracle.cpp). The forward step of the
// oracle is affine in the state X = (W, PML auxiliaries) for fixed medium:
// per RK stage s the right-hand side is h_s = L_s X_s (wave_deriv, then the
// free-surface rows, then the PML terms, all scaled by DT), and RK4 gives
//   X1 = X0 + a2 h0, X2 = X0 + a3 h1, X3 = X0 + h2,
//   X' = X0 + b1 h0 + b2 h1 + b3 h2 + b4 h3.
// Its transpose, for an adjoint state mu of X':
//   g3 = b4 mu,            l3 = L3^T g3
//   g2 = b3 mu + l3,       l2 = L2^T g2
//   g1 = b2 mu + a3 l2,    l1 = L1^T g1
//   g0 = b1 mu + a2 l1,
//   adjoint of X0 = mu + l1 + l2 + l3 + L0^T g0,
// and the medium gradient is sum_s <g_s, dL_s/d(mu, lambda, b) X_s>.
// L^T is written here as the scatter form of each forward kernel: every
// forward product h_i += c W_j becomes lW_j += c g_i, in the forward order.
#pragma once

namespace orc {

template <class T>
struct AdjVec {
  std::vector<T> W;                    // 9*N
  std::vector<std::vector<T>> A;       // per face, 9*cnt
};

template <class T>
struct Adjoint {
  Oracle<T>& o;
  explicit Adjoint(Oracle<T>& oo) : o(oo) {}

  AdjVec<T> zeros() const {
    AdjVec<T> v;
    v.W.assign(9 * (size_t)o.N, T(0));
    for (auto& f : o.faces) v.A.emplace_back(9 * (size_t)f.cnt, T(0));
    return v;
  }
  AdjVec<T> state() const {
    AdjVec<T> v;
    v.W = o.W;
    for (auto& f : o.faces) v.A.push_back(f.a);
    return v;
  }
  void load(const AdjVec<T>& v) {
    o.W = v.W;
    for (size_t q = 0; q < o.faces.size(); ++q) o.faces[q].a = v.A[q];
  }
  static void axpy(T a, const AdjVec<T>& x, AdjVec<T>& y) {
    for (size_t q = 0; q < y.W.size(); ++q) y.W[q] += a * x.W[q];
    for (size_t f = 0; f < y.A.size(); ++f)
      for (size_t q = 0; q < y.A[f].size(); ++q) y.A[f][q] += a * x.A[f][q];
  }
  static double dot(const AdjVec<T>& x, const AdjVec<T>& y) {
    double s = 0;
    for (size_t q = 0; q < y.W.size(); ++q) s += (double)x.W[q] * (double)y.W[q];
    for (size_t f = 0; f < y.A.size(); ++f)
      for (size_t q = 0; q < y.A[f].size(); ++q) s += (double)x.A[f][q] * (double)y.A[f][q];
    return s;
  }

  // forward right-hand side of an arbitrary state, through the oracle's own kernels
  AdjVec<T> rhs(const int FB[3], const AdjVec<T>& X) {
    std::swap(o.W, const_cast<std::vector<T>&>(X.W));
    for (size_t q = 0; q < o.faces.size(); ++q) std::swap(o.faces[q].a, const_cast<std::vector<T>&>(X.A[q]));
    std::fill(o.hW.begin(), o.hW.end(), T(0));
    o.wave_deriv(FB);
    if (o.fs) o.free_surface_deriv(FB);
    o.pml_deriv(FB);
    if (o.fs) o.pml_free_surface(FB);
    std::swap(o.W, const_cast<std::vector<T>&>(X.W));
    for (size_t q = 0; q < o.faces.size(); ++q) std::swap(o.faces[q].a, const_cast<std::vector<T>&>(X.A[q]));
    AdjVec<T> h;
    h.W = o.hW;
    for (auto& f : o.faces) h.A.push_back(f.h);
    return h;
  }

  // ---- transposes of the building blocks --------------------------------
  // L(f, p, s, fb) = fb (a_1 f[p-d] + a0 f[p] + a1 f[p+d] + a2 f[p+2d] + a3 f[p+3d]) rDH, d = fb s
  void L_T(T* lf, long long p, long long s, int fb, T g) const {
    const long long d = fb * s;
    const T c = T(fb) * g * o.rDH;
    lf[p - d] += c * T(AF_1);
    lf[p] += c * T(AF0);
    lf[p + d] += c * T(AF1);
    lf[p + 2 * d] += c * T(AF2);
    lf[p + 3 * d] += c * T(AF3);
  }
  // LJT on a 7-sample array centred at index 3
  void LJT_T(T* la, int fb, T g) const {
    const T c = T(fb) * g * o.rDH;
    la[3 - fb] += c * T(AF_1);
    la[3] += c * T(AF0);
    la[3 + fb] += c * T(AF1);
    la[3 + 2 * fb] += c * T(AF2);
    la[3 + 3 * fb] += c * T(AF3);
  }
  static void assemble_T(const T* g, T m1, T m2, T m3, T mu, T lam, T b, T* lD) {
    const T gd = lam * (g[3] + g[4] + g[5]);
    lD[0] = m1 * gd + T(2) * mu * m1 * g[3] + mu * (m2 * g[6] + m3 * g[7]);
    lD[1] = m2 * gd + T(2) * mu * m2 * g[4] + mu * (m1 * g[6] + m3 * g[8]);
    lD[2] = m3 * gd + T(2) * mu * m3 * g[5] + mu * (m1 * g[7] + m2 * g[8]);
    lD[3] = b * m1 * g[0];
    lD[4] = b * m2 * g[1];
    lD[5] = b * m3 * g[2];
    lD[6] = b * (m2 * g[0] + m1 * g[1]);
    lD[7] = b * (m3 * g[0] + m1 * g[2]);
    lD[8] = b * (m3 * g[1] + m2 * g[2]);
  }
  // <g, d assemble / d(mu, lambda, b)>
  static void assemble_grad(const T* g, const T* D, T m1, T m2, T m3, T& gmu, T& glam, T& gb) {
    gb += g[0] * (m1 * D[3] + m2 * D[6] + m3 * D[7]) + g[1] * (m1 * D[6] + m2 * D[4] + m3 * D[8]) +
          g[2] * (m1 * D[7] + m2 * D[8] + m3 * D[5]);
    const T div = m1 * D[0] + m2 * D[1] + m3 * D[2];
    glam += div * (g[3] + g[4] + g[5]);
    gmu += T(2) * (m1 * D[0] * g[3] + m2 * D[1] * g[4] + m3 * D[2] * g[5]) + (m2 * D[0] + m1 * D[1]) * g[6] +
           (m3 * D[0] + m1 * D[2]) * g[7] + (m3 * D[1] + m2 * D[2]) * g[8];
  }
  static void stress_T(const T* g, T m1, T m2, T m3, T mu, T lam, T* lV) {
    const T gd = lam * (g[0] + g[1] + g[2]);
    lV[0] = m1 * gd + T(2) * mu * m1 * g[0] + mu * (m2 * g[3] + m3 * g[4]);
    lV[1] = m2 * gd + T(2) * mu * m2 * g[1] + mu * (m1 * g[3] + m3 * g[5]);
    lV[2] = m3 * gd + T(2) * mu * m3 * g[2] + mu * (m1 * g[4] + m2 * g[5]);
  }
  static void stress_grad(const T* g, const T* V, T m1, T m2, T m3, T& gmu, T& glam) {
    const T div = m1 * V[0] + m2 * V[1] + m3 * V[2];
    glam += div * (g[0] + g[1] + g[2]);
    gmu += T(2) * (m1 * V[0] * g[0] + m2 * V[1] * g[1] + m3 * V[2] * g[2]) + (m2 * V[0] + m1 * V[1]) * g[3] +
           (m3 * V[0] + m1 * V[2]) * g[4] + (m3 * V[1] + m2 * V[2]) * g[5];
  }

  // Derivatives of the surface matrices -DZ^-1 DX, -DZ^-1 DY at surface column
  // s with respect to mu and lambda of that cell, contracted with gRx, gRy
  // (R = -DZ^-1 D  =>  dR = -DZ^-1 (dDZ R + dD)), FP64.
  void rdz_grad(long long p, long long s, const double* gRx, const double* gRy, T& gmu, T& glam) const {
    const double* m = o.metric_d.data();
    const long long N = o.N, S = (long long)o.NX * o.NY;
    const double xx = m[0 * N + p], xy = m[1 * N + p], xz = m[2 * N + p];
    const double ex = m[3 * N + p], ey = m[4 * N + p], ez = m[5 * N + p];
    const double zx = m[6 * N + p], zy = m[7 * N + p], zz = m[8 * N + p];
    const double mu = (double)o.cjm[C_MU * N + p], lam = (double)o.cjm[C_LAM * N + p];
    // D matrices as functions of (chi, lm, lam, mu) with chi = 2 lam + mu, lm = lam + mu
    auto DZof = [&](double chi, double lm, double mu_, double* DZ) {
      DZ[0] = chi * zx * zx + mu_ * (zy * zy + zz * zz);
      DZ[1] = lm * zx * zy;
      DZ[2] = lm * zx * zz;
      DZ[3] = DZ[1];
      DZ[4] = chi * zy * zy + mu_ * (zx * zx + zz * zz);
      DZ[5] = lm * zy * zz;
      DZ[6] = DZ[2];
      DZ[7] = DZ[5];
      DZ[8] = chi * zz * zz + mu_ * (zx * zx + zy * zy);
    };
    auto Dof = [&](double chi, double lam_, double mu_, double ax, double ay, double az, double* D) {
      D[0] = chi * zx * ax + mu_ * (zy * ay + zz * az);
      D[1] = lam_ * zx * ay + mu_ * zy * ax;
      D[2] = lam_ * zx * az + mu_ * zz * ax;
      D[3] = lam_ * zy * ax + mu_ * zx * ay;
      D[4] = chi * zy * ay + mu_ * (zx * ax + zz * az);
      D[5] = lam_ * zy * az + mu_ * zz * ay;
      D[6] = lam_ * zz * ax + mu_ * zx * az;
      D[7] = lam_ * zz * ay + mu_ * zy * az;
      D[8] = chi * zz * az + mu_ * (zx * ax + zy * ay);
    };
    double DZ[9], Zi[9];
    DZof(2 * lam + mu, lam + mu, mu, DZ);
    inv3(DZ, Zi);
    double Rx[9], Ry[9];
    for (int q = 0; q < 9; ++q) {
      Rx[q] = (double)o.rdzdx[q * S + s];
      Ry[q] = (double)o.rdzdy[q * S + s];
    }
    // G = Zi^T gR; contribution = -<G, dDZ R + dD>
    double Gx[9], Gy[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        Gx[3 * r + c] = Zi[0 * 3 + r] * gRx[0 * 3 + c] + Zi[1 * 3 + r] * gRx[1 * 3 + c] + Zi[2 * 3 + r] * gRx[2 * 3 + c];
        Gy[3 * r + c] = Zi[0 * 3 + r] * gRy[0 * 3 + c] + Zi[1 * 3 + r] * gRy[1 * 3 + c] + Zi[2 * 3 + r] * gRy[2 * 3 + c];
      }
    // the matrices are linear in (mu, lam): derivative = value at unit parameter
    for (int which = 0; which < 2; ++which) {
      const double dmu = which == 0 ? 1.0 : 0.0, dlam = which == 0 ? 0.0 : 1.0;
      double dZ[9], dX[9], dY[9];
      DZof(2 * dlam + dmu, dlam + dmu, dmu, dZ);
      Dof(2 * dlam + dmu, dlam, dmu, xx, xy, xz, dX);
      Dof(2 * dlam + dmu, dlam, dmu, ex, ey, ez, dY);
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
      if (which == 0) gmu += T(acc);
      else glam += T(acc);
    }
  }
  static void inv3(const double* A, double* B) {
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
    B[0] = (A[4] * A[8] - A[5] * A[7]) / det;
    B[1] = (A[2] * A[7] - A[1] * A[8]) / det;
    B[2] = (A[1] * A[5] - A[2] * A[4]) / det;
    B[3] = (A[5] * A[6] - A[3] * A[8]) / det;
    B[4] = (A[0] * A[8] - A[2] * A[6]) / det;
    B[5] = (A[2] * A[3] - A[0] * A[5]) / det;
    B[6] = (A[3] * A[7] - A[4] * A[6]) / det;
    B[7] = (A[1] * A[6] - A[0] * A[7]) / det;
    B[8] = (A[0] * A[4] - A[1] * A[3]) / det;
  }

  // ---- transpose of the stage right-hand side -----------------------------
  // (lW, lA) += L^T (g) and, when X is given, (gmu, glam, gb) += <g, dL/dtheta X>.
  void rhs_T(const int FB[3], const AdjVec<T>& g, AdjVec<T>& l, const AdjVec<T>* X, T* gmu, T* glam, T* gb) {
    const long long N = o.N, S = (long long)o.NX * o.NY;
    const T* c = o.cjm.data();
    const int ktop = o.NZ - H - 1;
    const T* w = X ? X->W.data() : nullptr;
    T* lw = l.W.data();
    // wave_deriv (interior; its top-3-layer rows are replaced by the surface rows)
    for (int k = H; k < o.NZ - H; ++k)
      for (int j = H; j < o.NY - H; ++j)
        for (int i = H; i < o.NX - H; ++i) {
          if (o.fs && k >= ktop - 2) continue;
          const long long p = o.at(i, j, k);
          const T bet[3] = {o.pb[0][i], o.pb[1][j], o.pb[2][k]};
          const T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
          T gg[9];
          for (int v = 0; v < 9; ++v) gg[v] = g.W[v * N + p] * o.DT;
          for (int ax = 0; ax < 3; ++ax) {
            const T m1 = c[(3 * ax) * N + p], m2 = c[(3 * ax + 1) * N + p], m3 = c[(3 * ax + 2) * N + p];
            T lD[9];
            assemble_T(gg, m1, m2, m3, mu, lam, b, lD);
            for (int v = 0; v < 9; ++v) L_T(lw + v * N, p, o.stride(ax), FB[ax], lD[v] * bet[ax]);
            if (X) {
              T D[9];
              for (int v = 0; v < 9; ++v) D[v] = o.L(w + v * N, p, o.stride(ax), FB[ax]) * bet[ax];
              assemble_grad(gg, D, m1, m2, m3, gmu[p], glam[p], gb[p]);
            }
          }
        }
    if (o.fs) free_surface_T(FB, g, l, X, gmu, glam, gb);
    // PML faces: h_A = (Q - (alpha + d) A) DT, hW -= beta A DT
    for (size_t fi = 0; fi < o.faces.size(); ++fi) {
      auto& f = o.faces[fi];
      const int ax = f.axis;
      for (long long q = 0; q < f.cnt; ++q) {
        int i, j, k;
        o.slab_cell(f, q, i, j, k);
        const int qq = ax == 0 ? i : ax == 1 ? j : k;
        const long long p = o.at(i, j, k);
        const T be = o.pb[ax][qq], d = o.pd[ax][qq], ad = d + o.pa[ax][qq];
        const T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
        const T m1 = c[(3 * ax) * N + p], m2 = c[(3 * ax + 1) * N + p], m3 = c[(3 * ax + 2) * N + p];
        T gQ[9], lD[9];
        for (int v = 0; v < 9; ++v) {
          const T gA = g.A[fi][v * f.cnt + q];
          l.A[fi][v * f.cnt + q] += -ad * o.DT * gA - be * o.DT * g.W[v * N + p];
          gQ[v] = gA * o.DT;
        }
        assemble_T(gQ, m1, m2, m3, mu, lam, b, lD);
        for (int v = 0; v < 9; ++v) L_T(lw + v * N, p, o.stride(ax), FB[ax], lD[v] * d);
        if (X) {
          T D[9];
          for (int v = 0; v < 9; ++v) D[v] = o.L(w + v * N, p, o.stride(ax), FB[ax]) * d;
          assemble_grad(gQ, D, m1, m2, m3, gmu[p], glam[p], gb[p]);
        }
      }
    }
    // surface term of the x/y PML auxiliaries (pml_freeSurface.cpp:17-242)
    if (o.fs)
      for (size_t fi = 0; fi < o.faces.size(); ++fi) {
        auto& f = o.faces[fi];
        if (f.axis == 2) continue;
        const int ax = f.axis;
        const T* R = ax == 0 ? o.rdzdx.data() : o.rdzdy.data();
        const long long plane = (long long)f.n0 * f.n1, base = (long long)(o.nz - 1) * plane;
        for (long long q2 = 0; q2 < plane; ++q2) {
          const long long q = base + q2;
          int i, j, k;
          o.slab_cell(f, q, i, j, k);
          const long long p = o.at(i, j, k), s = i + (long long)o.NX * j;
          const T d = o.pd[ax][ax == 0 ? i : j];
          const T mu = c[C_MU * N + p], lam = c[C_LAM * N + p];
          const T zx = c[C_ZTX * N + p], zy = c[C_ZTY * N + p], zz = c[C_ZTZ * N + p];
          T gs[6], lVz[3], lV[3] = {0, 0, 0};
          for (int v = 0; v < 6; ++v) gs[v] = g.A[fi][(3 + v) * f.cnt + q] * o.DT;
          stress_T(gs, zx, zy, zz, mu, lam, lVz);
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) lV[cc] += R[(3 * r + cc) * S + s] * lVz[r];
          for (int v = 0; v < 3; ++v) L_T(lw + v * N, p, o.stride(ax), FB[ax], lV[v] * d);
          if (X) {
            T V[3], Vz[3];
            for (int v = 0; v < 3; ++v) V[v] = o.L(w + v * N, p, o.stride(ax), FB[ax]) * d;
            for (int r = 0; r < 3; ++r)
              Vz[r] = R[(3 * r) * S + s] * V[0] + R[(3 * r + 1) * S + s] * V[1] + R[(3 * r + 2) * S + s] * V[2];
            stress_grad(gs, Vz, zx, zy, zz, gmu[p], glam[p]);
            double gR[9] = {0}, zero[9] = {0};
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc) gR[3 * r + cc] = (double)lVz[r] * (double)V[cc];
            if (ax == 0) rdz_grad(p, s, gR, zero, gmu[p], glam[p]);
            else rdz_grad(p, s, zero, gR, gmu[p], glam[p]);
          }
        }
      }
  }

  // transpose of free_surface_deriv (freeSurface.cpp:243-494)
  void free_surface_T(const int FB[3], const AdjVec<T>& g, AdjVec<T>& l, const AdjVec<T>* X, T* gmu, T* glam, T* gb) {
    const long long N = o.N, S = (long long)o.NX * o.NY;
    const T* c = o.cjm.data();
    const int ktop = o.NZ - H - 1;
    const T* w = X ? X->W.data() : nullptr;
    T* lw = l.W.data();
    for (int k = ktop - 2; k <= ktop; ++k)
      for (int j = H; j < o.NY - H; ++j)
        for (int i = H; i < o.NX - H; ++i) {
          const long long p = o.at(i, j, k);
          const T mu = c[C_MU * N + p], lam = c[C_LAM * N + p], b = c[C_B * N + p];
          const T bx = o.pb[0][i], by = o.pb[1][j];
          const int ks = ktop - k + H;
          const T jinv = T(1) / c[C_J * N + p];
          // velocity rows: hv = b jinv (LJT(JT_x) bx + LJT(JT_y) by + LJT(JT_z)), hW = hv DT
          T lJT[3][3][7] = {};
          for (int cc = 0; cc < 3; ++cc) {
            const T gv = g.W[cc * N + p] * o.DT * b * jinv;
            LJT_T(lJT[0][cc], FB[0], gv * bx);
            LJT_T(lJT[1][cc], FB[1], gv * by);
            LJT_T(lJT[2][cc], FB[2], gv);
          }
          if (X) {
            // gb: <g, hv / b>
            T JT[3][3][7];
            for (int l2 = 0; l2 < 7; ++l2)
              for (int ax = 0; ax < 2; ++ax) {
                const long long q = ax == 0 ? o.at(i + l2 - H, j, k) : o.at(i, j + l2 - H, k);
                const T m1 = c[(3 * ax) * N + q], m2 = c[(3 * ax + 1) * N + q], m3 = c[(3 * ax + 2) * N + q];
                const T jac = c[C_J * N + q];
                JT[ax][0][l2] = (m1 * w[3 * N + q] + m2 * w[6 * N + q] + m3 * w[7 * N + q]) * jac;
                JT[ax][1][l2] = (m1 * w[6 * N + q] + m2 * w[4 * N + q] + m3 * w[8 * N + q]) * jac;
                JT[ax][2][l2] = (m1 * w[7 * N + q] + m2 * w[8 * N + q] + m3 * w[5 * N + q]) * jac;
              }
            for (int l2 = 0; l2 < ks; ++l2) {
              const long long q = o.at(i, j, k + l2 - H);
              const T m1 = c[C_ZTX * N + q], m2 = c[C_ZTY * N + q], m3 = c[C_ZTZ * N + q];
              const T jac = c[C_J * N + q];
              JT[2][0][l2] = (m1 * w[3 * N + q] + m2 * w[6 * N + q] + m3 * w[7 * N + q]) * jac;
              JT[2][1][l2] = (m1 * w[6 * N + q] + m2 * w[4 * N + q] + m3 * w[8 * N + q]) * jac;
              JT[2][2][l2] = (m1 * w[7 * N + q] + m2 * w[8 * N + q] + m3 * w[5 * N + q]) * jac;
            }
            for (int cc = 0; cc < 3; ++cc) {
              JT[2][cc][ks] = T(0);
              for (int l2 = ks + 1; l2 < 7; ++l2) JT[2][cc][l2] = -JT[2][cc][2 * ks - l2];
            }
            for (int cc = 0; cc < 3; ++cc) {
              const T hv_b = jinv * (o.LJT(JT[0][cc], FB[0]) * bx + o.LJT(JT[1][cc], FB[1]) * by + o.LJT(JT[2][cc], FB[2]));
              gb[p] += g.W[cc * N + p] * o.DT * hv_b;
            }
          }
          // images: JT_z[l] = -JT_z[2 ks - l] for l > ks, JT_z[ks] = 0
          for (int cc = 0; cc < 3; ++cc) {
            for (int l2 = 6; l2 > ks; --l2) lJT[2][cc][2 * ks - l2] -= lJT[2][cc][l2];
            lJT[2][cc][ks] = T(0);
          }
          auto scatterJT = [&](const T* lj, long long q, int ax) {
            // JT = (m . T) jac with metrics of direction ax at q
            const T m1 = c[(3 * ax) * N + q], m2 = c[(3 * ax + 1) * N + q], m3 = c[(3 * ax + 2) * N + q];
            const T jac = c[C_J * N + q];
            lw[3 * N + q] += m1 * jac * lj[0];
            lw[6 * N + q] += m2 * jac * lj[0];
            lw[7 * N + q] += m3 * jac * lj[0];
            lw[6 * N + q] += m1 * jac * lj[1];
            lw[4 * N + q] += m2 * jac * lj[1];
            lw[8 * N + q] += m3 * jac * lj[1];
            lw[7 * N + q] += m1 * jac * lj[2];
            lw[8 * N + q] += m2 * jac * lj[2];
            lw[5 * N + q] += m3 * jac * lj[2];
          };
          for (int l2 = 0; l2 < 7; ++l2) {
            const T ljx[3] = {lJT[0][0][l2], lJT[0][1][l2], lJT[0][2][l2]};
            const T ljy[3] = {lJT[1][0][l2], lJT[1][1][l2], lJT[1][2][l2]};
            scatterJT(ljx, o.at(i + l2 - H, j, k), 0);
            scatterJT(ljy, o.at(i, j + l2 - H, k), 1);
          }
          for (int l2 = 0; l2 < ks; ++l2) {
            const T ljz[3] = {lJT[2][0][l2], lJT[2][1][l2], lJT[2][2][l2]};
            scatterJT(ljz, o.at(i, j, k + l2 - H), 2);
          }
          // stress rows: hW[3+v] = (s(Vxi) + s(Vet) + s(Vzt)) DT
          T gs[6];
          for (int v = 0; v < 6; ++v) gs[v] = g.W[(3 + v) * N + p] * o.DT;
          T lVxi[3], lVet[3], lVzt[3];
          stress_T(gs, c[C_XIX * N + p], c[C_XIY * N + p], c[C_XIZ * N + p], mu, lam, lVxi);
          stress_T(gs, c[C_ETX * N + p], c[C_ETY * N + p], c[C_ETZ * N + p], mu, lam, lVet);
          stress_T(gs, c[C_ZTX * N + p], c[C_ZTY * N + p], c[C_ZTZ * N + p], mu, lam, lVzt);
          T Vxi[3] = {0, 0, 0}, Vet[3] = {0, 0, 0}, Vzt[3] = {0, 0, 0};
          if (X)
            for (int v = 0; v < 3; ++v) {
              Vxi[v] = o.L(w + v * N, p, 1, FB[0]) * bx;
              Vet[v] = o.L(w + v * N, p, o.NX, FB[1]) * by;
            }
          const long long s = i + (long long)o.NX * j;
          if (k == ktop) {
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc) {
                lVxi[cc] += o.rdzdx[(3 * r + cc) * S + s] * lVzt[r];
                lVet[cc] += o.rdzdy[(3 * r + cc) * S + s] * lVzt[r];
              }
            if (X) {
              for (int r = 0; r < 3; ++r)
                Vzt[r] = (o.rdzdx[(3 * r) * S + s] * Vxi[0] + o.rdzdx[(3 * r + 1) * S + s] * Vxi[1] + o.rdzdx[(3 * r + 2) * S + s] * Vxi[2]) +
                         (o.rdzdy[(3 * r) * S + s] * Vet[0] + o.rdzdy[(3 * r + 1) * S + s] * Vet[1] + o.rdzdy[(3 * r + 2) * S + s] * Vet[2]);
              double gRx[9], gRy[9];
              for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) {
                  gRx[3 * r + cc] = (double)lVzt[r] * (double)Vxi[cc];
                  gRy[3 * r + cc] = (double)lVzt[r] * (double)Vet[cc];
                }
              rdz_grad(p, s, gRx, gRy, gmu[p], glam[p]);
            }
          } else if (k == ktop - 1) {
            const long long d = FB[2] * S;
            for (int v = 0; v < 3; ++v) {
              const T cc = T(FB[2]) * lVzt[v] * o.rDH;
              lw[v * N + p + d] += cc;
              lw[v * N + p] -= cc;
              if (X) Vzt[v] = T(FB[2]) * (w[v * N + p + d] - w[v * N + p]) * o.rDH;
            }
          } else {
            const long long d = FB[2] * S;
            for (int v = 0; v < 3; ++v) {
              const T cc = T(FB[2]) * lVzt[v] * o.rDH;
              lw[v * N + p] += cc * T(CF1);
              lw[v * N + p + d] += cc * T(CF2);
              lw[v * N + p + 2 * d] += cc * T(CF3);
              if (X) Vzt[v] = T(FB[2]) * (T(CF1) * w[v * N + p] + T(CF2) * w[v * N + p + d] + T(CF3) * w[v * N + p + 2 * d]) * o.rDH;
            }
          }
          for (int v = 0; v < 3; ++v) {
            L_T(lw + v * N, p, 1, FB[0], lVxi[v] * bx);
            L_T(lw + v * N, p, o.NX, FB[1], lVet[v] * by);
          }
          if (X) {
            stress_grad(gs, Vxi, c[C_XIX * N + p], c[C_XIY * N + p], c[C_XIZ * N + p], gmu[p], glam[p]);
            stress_grad(gs, Vet, c[C_ETX * N + p], c[C_ETY * N + p], c[C_ETZ * N + p], gmu[p], glam[p]);
            stress_grad(gs, Vzt, c[C_ZTX * N + p], c[C_ZTY * N + p], c[C_ZTZ * N + p], gmu[p], glam[p]);
          }
        }
  }

  // ---- one time step -------------------------------------------------------
  // RK4 stage states of the step that starts from X0 (after the source).
  void stages(const int FB0[3], const AdjVec<T>& X0, AdjVec<T> (&Xs)[4]) {
    int FB[3] = {FB0[0], FB0[1], FB0[2]};
    Xs[0] = X0;
    const T a[3] = {T(RK_A2), T(RK_A3), T(1)};
    for (int s = 0; s < 3; ++s) {
      AdjVec<T> h = rhs(FB, Xs[s]);
      Xs[s + 1] = X0;
      axpy(a[s], h, Xs[s + 1]);
      FB[0] = -FB[0];
      FB[1] = -FB[1];
      FB[2] = -FB[2];
    }
  }

  // mu: adjoint of the state after the step (in/out: becomes the adjoint of
  // the state after the source injection at the step start).
  void step_T(int step_it, const AdjVec<T>& X0, AdjVec<T>& mu, T* gmu, T* glam, T* gb) {
    int FB0[3] = {FBTAB[step_it % 8][0], FBTAB[step_it % 8][1], FBTAB[step_it % 8][2]};
    AdjVec<T> Xs[4];
    const bool want_grad = gmu != nullptr;
    if (want_grad) stages(FB0, X0, Xs);
    int FBs[4][3];
    for (int s = 0; s < 4; ++s)
      for (int d = 0; d < 3; ++d) FBs[s][d] = (s % 2 == 0) ? FB0[d] : -FB0[d];
    const T bw[4] = {T(RK_B1), T(RK_B2), T(RK_B3), T(RK_B4)};
    const T aw[4] = {T(0), T(RK_A2), T(RK_A3), T(1)};  // X_s = X0 + aw[s] h_{s-1}
    AdjVec<T> out = mu;   // adjoint of X0 accumulates here
    AdjVec<T> lnext;      // l_{s+1} = L_{s+1}^T g_{s+1}
    for (int s = 3; s >= 0; --s) {
      AdjVec<T> g = zeros();
      axpy(bw[s], mu, g);
      if (s < 3) axpy(aw[s + 1], lnext, g);
      AdjVec<T> l = zeros();
      rhs_T(FBs[s], g, l, want_grad ? &Xs[s] : nullptr, gmu, glam, gb);
      axpy(T(1), l, out);
      lnext = std::move(l);
    }
    mu = std::move(out);
  }
};

}  // namespace orc
