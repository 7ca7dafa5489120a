#include "source_build.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace cgh {

namespace {

// reference macro.h:70-73: PI is the float literal 3.141592657f and
// DEGREE2RADIAN = PI / 180.0 in double
const double REF_PI = (double)3.141592657f;
const double REF_DEG2RAD = REF_PI / 180.0;

// GAUSS_FUN(t, a, t0) of macro.h:37
inline double gauss_fun(double t, double a) { return std::exp(-((t / a) * (t / a))) / (a * 1.772453850905516); }

float ricker_value(float fc, int it, float DT) {
  // sourceFunction, singleSource.cpp:42-68 (irk = 0)
  float t = ((float)it + 0.0f) * DT;
  float tdelay = 1.0f / fc;
  float r = (float)REF_PI * fc * (t - tdelay);
  float rr = r * r;
  float s = (float)(r * (3.0 - 2.0f * rr) * std::exp(-rr) * std::sqrt(REF_PI) * 0.5 * fc * REF_PI);
  return s * 1e9f;
}

float gauss_value(float duration, int it, float DT) {
  // gaussFunction, singleSource.cpp:71-92 (irk = 0)
  float t = ((float)it + 0.0f) * DT;
  float tdelay = duration / 2 + 1;
  return (float)(std::exp(-(t - tdelay) * (t - tdelay) / (duration * duration)) / (std::sqrt(REF_PI) * duration));
}

}  // namespace

void double_couple(float strike, float dip, float rake, float m[6]) {
  const double s = strike, d = dip, r = rake;
  const float M11 = (float)(-(std::sin(d) * std::cos(r) * std::sin(2.0 * s) +
                              std::sin(2.0 * d) * std::sin(r) * std::sin(s) * std::sin(s)));
  const float M22 = (float)(std::sin(d) * std::cos(r) * std::sin(2.0 * s) -
                            std::sin(2.0 * d) * std::sin(r) * std::cos(s) * std::cos(s));
  const float M33 = -(M11 + M22);
  const float M12 = (float)(std::sin(d) * std::cos(r) * std::cos(2.0 * s) +
                            0.5 * std::sin(2.0 * d) * std::sin(r) * std::sin(2.0 * s));
  const float M13 = (float)(-(std::cos(d) * std::cos(r) * std::cos(s) + std::cos(2.0 * d) * std::sin(r) * std::sin(s)));
  const float M23 = (float)(-(std::cos(d) * std::cos(r) * std::sin(s) - std::cos(2.0 * d) * std::sin(r) * std::cos(s)));
  m[0] = M22;
  m[1] = M11;
  m[2] = M33;
  m[3] = M12;
  m[4] = -M23;
  m[5] = -M13;
}

int single_source_table(const Params& p, const Grid& g, int nsteps, const double* jac, const double* mu,
                        SourceTable& t, std::string& msg) {
  t = SourceTable();
  const bool ricker = p.useSingleSource_ricker != 0, dc = p.useSingleSource_double_couple != 0;
  if (!ricker && !dc) return 0;
  if (ricker && dc) {
    // modelChecking.cpp:23-31
    msg = "useSingleSource(ricker) and useSingleSource(double_couple) are both set";
    return 180;
  }
  // locateSource (singleSource.cpp:25-40), one rank: front = 0
  const int SX = p.sourceX + HALO, SY = p.sourceY + HALO, SZ = p.sourceZ + HALO;
  if (SX < HALO || SX >= g.NX - HALO || SY < HALO || SY >= g.NY - HALO || SZ < HALO || SZ >= g.NZ - HALO) {
    // the reference launches nothing for a source outside the rank
    msg = "single source outside the grid";
    return 180;
  }
  const float DT = (float)p.DT;
  const double Jc = jac[g.idx(SX, SY, SZ)];
  const int nGauss = 3;
  const double ra = nGauss * 0.5;
  float mdc[6] = {0, 0, 0, 0, 0, 0};
  double m0 = 0.0;
  if (dc) {
    double_couple((float)(p.strike * REF_DEG2RAD), (float)(p.dip * REF_DEG2RAD), (float)(p.rake * REF_DEG2RAD), mdc);
    m0 = 1.5 * std::pow(10.0, (double)p.Mw + 6.0);
  }
  t.nt = nsteps + 1;
  t.dt = DT;
  std::vector<float> stf((size_t)t.nt);
  for (int n = 0; n < t.nt; ++n) stf[(size_t)n] = ricker ? ricker_value(p.rickerfc, n, DT) : gauss_value(p.duration, n, DT);
  for (int k = SZ - nGauss; k <= SZ + nGauss; ++k)
    for (int j = SY - nGauss; j <= SY + nGauss; ++j)
      for (int i = SX - nGauss; i <= SX + nGauss; ++i) {
        if (i < HALO || i >= g.NX - HALO || j < HALO || j >= g.NY - HALO || k < HALO || k >= g.NZ - HALO) continue;
        const long long c = g.idx(i, j, k);
        const double amp = gauss_fun(i - SX, ra) * gauss_fun(j - SY, ra) * gauss_fun(k - SZ, ra) / 0.998125703461425;
        // the library injects table * mu/(J DH^3) * DT; the reference adds
        // -s amp / (J_c DH^3) DT (ricker, to Txx..Tzz) or subtracts m * rate
        // with rate = -s amp m0 / (J_c DH^3) DT (double couple)
        const double undo = jac[c] / mu[c];  // (J DH^3 / mu) * (1/DH^3)
        t.cell.push_back(c);
        const size_t base = t.rate.size();
        t.rate.resize(base + (size_t)t.nt * 6, 0.f);
        for (int n = 0; n < t.nt; ++n) {
          float* r = &t.rate[base + (size_t)n * 6];
          const double s = stf[(size_t)n];
          if (ricker) {
            const double v = s * amp / Jc * undo;  // minus the reference's increment / (DT sc)
            r[0] = r[1] = r[2] = (float)v;
          } else {
            const double rate = -s * amp * m0 / Jc * undo;
            for (int q = 0; q < 6; ++q) r[q] = (float)(mdc[q] * rate);
          }
        }
      }
  return 0;
}

// ---------------------------------------------------------------------------
int read_fault_header(const std::string& path, FaultHeader& h, std::string& msg) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    msg = "can't open source file! Please check if there is a source named " + path;
    return 1001;
  }
  int npts = 0, nt = 0;
  float dt = 0;
  bool ok = std::fread(&npts, sizeof(int), 1, f) == 1 && std::fread(&nt, sizeof(int), 1, f) == 1 &&
            std::fread(&dt, sizeof(float), 1, f) == 1;
  std::fclose(f);
  if (!ok || npts < 0 || nt < 2 || !(dt > 0)) {
    msg = "bad source file header in " + path;
    return 1001;
  }
  h.npts = npts;
  h.nt = nt;
  h.dt = dt;
  return 0;
}

int read_fault_locations(const std::string& path, const FaultHeader& h, std::vector<float>& lon,
                         std::vector<float>& lat, std::vector<float>& z, std::string& msg) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    msg = "can't open source file " + path;
    return 1001;
  }
  lon.assign((size_t)h.npts, 0.f);
  lat.assign((size_t)h.npts, 0.f);
  z.assign((size_t)h.npts, 0.f);
  const long long rec = 6LL * 4 + 2LL * 4 * h.nt;
  bool ok = true;
  for (long long i = 0; ok && i < h.npts; ++i) {
    ok = std::fseek(f, (long)(12 + rec * i), SEEK_SET) == 0;
    float v[3];
    ok = ok && std::fread(v, sizeof(float), 3, f) == 3;
    if (ok) {
      lon[(size_t)i] = v[0];
      lat[(size_t)i] = v[1];
      z[(size_t)i] = v[2];
    }
  }
  std::fclose(f);
  if (!ok) {
    msg = "source file " + path + " is truncated";
    return 1001;
  }
  return 0;
}

bool aeqd_forward(double lon0, double lat0, double lon, double lat, double a, double f, double& x, double& y) {
  const double D2R = M_PI / 180.0;
  if (std::fabs(lon - lon0) < 1e-14 && std::fabs(lat - lat0) < 1e-14) {
    x = y = 0.0;
    return true;
  }
  const double b = a * (1.0 - f);
  const double L = (lon - lon0) * D2R;
  const double U1 = std::atan((1.0 - f) * std::tan(lat0 * D2R));
  const double U2 = std::atan((1.0 - f) * std::tan(lat * D2R));
  const double sU1 = std::sin(U1), cU1 = std::cos(U1), sU2 = std::sin(U2), cU2 = std::cos(U2);
  double lam = L, lamP, sinS = 0, cosS = 0, sig = 0, cos2a = 0, cos2sm = 0, sinL = 0, cosL = 0;
  int iter = 0;
  do {
    sinL = std::sin(lam);
    cosL = std::cos(lam);
    const double t1 = cU2 * sinL, t2 = cU1 * sU2 - sU1 * cU2 * cosL;
    sinS = std::sqrt(t1 * t1 + t2 * t2);
    if (sinS == 0.0) {
      x = y = 0.0;
      return true;
    }
    cosS = sU1 * sU2 + cU1 * cU2 * cosL;
    sig = std::atan2(sinS, cosS);
    const double sinA = cU1 * cU2 * sinL / sinS;
    cos2a = 1.0 - sinA * sinA;
    cos2sm = cos2a != 0.0 ? cosS - 2.0 * sU1 * sU2 / cos2a : 0.0;
    const double C = f / 16.0 * cos2a * (4.0 + f * (4.0 - 3.0 * cos2a));
    lamP = lam;
    lam = L + (1.0 - C) * f * sinA * (sig + C * sinS * (cos2sm + C * cosS * (-1.0 + 2.0 * cos2sm * cos2sm)));
  } while (std::fabs(lam - lamP) > 1e-13 && ++iter < 500);
  if (iter >= 500) return false;
  const double u2 = cos2a * (a * a - b * b) / (b * b);
  const double A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)));
  const double B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)));
  const double dS = B * sinS * (cos2sm + B / 4.0 * (cosS * (-1.0 + 2.0 * cos2sm * cos2sm) -
                                                      B / 6.0 * cos2sm * (-3.0 + 4.0 * sinS * sinS) *
                                                          (-3.0 + 4.0 * cos2sm * cos2sm)));
  const double s = b * A * (sig - dS);
  const double az = std::atan2(cU2 * sinL, cU1 * sU2 - sU1 * cU2 * cosL);
  x = s * std::sin(az);
  y = s * std::cos(az);
  return true;
}

std::map<long long, long long> fault_point_cells(const Grid& g, const Params& p, const std::vector<float>& coord,
                                                 const std::vector<float>& x, const std::vector<float>& y,
                                                 std::vector<float>& z) {
  std::map<long long, long long> out;
  const long long N = g.cells();
  const float Depth = p.Depth * 1000.f;
  const float DH = (float)p.DH;
  const long long npts = (long long)z.size();
  float maxZ = -Depth;
  for (long long q = 0; q < npts; ++q)
    if (maxZ < z[(size_t)q]) maxZ = z[(size_t)q];
  for (long long q = 0; q < npts; ++q) z[(size_t)q] = z[(size_t)q] - maxZ;
  const int nxe = g.NX - HALO, nye = g.NY - HALO, nze = g.NZ - HALO;
  auto cz = [&](int i, int j, int k) { return coord[(size_t)(2 * N + g.idx(i, j, k))]; };
  for (long long q = 0; q < npts; ++q) {
    // int() truncates toward zero, as the reference's conversion does
    const int sx = (int)(x[(size_t)q] / DH + 0.5) + p.centerX + HALO;
    const int sy = (int)(y[(size_t)q] / DH + 0.5) + p.centerY + HALO;
    if (!(sx >= HALO && sx < nxe && sy >= HALO && sy < nye)) continue;
    const double DZ = (double)(float)(coord[(size_t)(2 * N + g.idx(sx, sy, 4))] - coord[(size_t)(2 * N + g.idx(sx, sy, 3))]);
    const double terrain = DZ * double(p.NZ - 1) - std::fabs(Depth);
    z[(size_t)q] = (float)(terrain + z[(size_t)q]);
    for (int k = HALO; k < nze; ++k) {
      // float differences halved in double, as the reference evaluates them
      const double dz1 = (double)(float)(cz(sx, sy, k) - cz(sx, sy, k - 1)) * 0.5;
      const double dz2 = (double)(float)(cz(sx, sy, k + 1) - cz(sx, sy, k)) * 0.5;
      const double zk = cz(sx, sy, k), zz = z[(size_t)q];
      if (zk - dz1 <= zz && zz < dz2 + zk) out[q] = g.idx(sx, sy, k);
    }
  }
  return out;
}

int fault_source_table(const Params& p, const Grid& g, const std::vector<float>& coord, SourceTable& t,
                       std::vector<float>* xyz, std::string& msg) {
  t = SourceTable();
  const std::string path = p.sourceDir + "/" + p.sourceFile;
  FaultHeader h;
  int rc = read_fault_header(path, h, msg);
  if (rc) return rc;
  std::vector<float> x, y, z;
  if ((rc = read_fault_locations(path, h, x, y, z, msg))) return rc;
  // projTrans (multiSource.cpp:122-171): centre from the float parameters,
  // printed with %lf into the PROJ string; angles scaled by the reference's
  // DEGREE2RADIAN on the way in and by the exact factor inside PROJ
  char b0[64], b1[64];
  std::snprintf(b0, sizeof b0, "%lf", (double)(float)p.centerLongitude);
  std::snprintf(b1, sizeof b1, "%lf", (double)(float)p.centerLatitude);
  const double lon0 = std::strtod(b0, nullptr), lat0 = std::strtod(b1, nullptr);
  const double scale = REF_DEG2RAD / (M_PI / 180.0);
  for (size_t q = 0; q < x.size(); ++q) {
    double px, py;
    if (!aeqd_forward(lon0, lat0, (double)x[q] * scale, (double)y[q] * scale, WGS84_A, WGS84_F, px, py)) {
      msg = "projection of source point " + std::to_string(q) + " did not converge";
      return 180;
    }
    x[q] = (float)px;
    y[q] = (float)py;
  }
  std::map<long long, long long> pc = fault_point_cells(g, p, coord, x, y, z);
  if (xyz) {
    xyz->clear();
    xyz->insert(xyz->end(), x.begin(), x.end());
    xyz->insert(xyz->end(), y.begin(), y.end());
    xyz->insert(xyz->end(), z.begin(), z.end());
  }
  // readSourceParams + solveMomentRate (multiSource.cpp:345-535)
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    msg = "can't open source file " + path;
    return 1001;
  }
  const int nt = h.nt;
  const long long rec = 6LL * 4 + 2LL * 4 * nt;
  t.nt = nt;
  t.dt = h.dt;
  std::vector<float> rake((size_t)nt), rt((size_t)nt);
  for (auto& kv : pc) {
    float v3[3];
    bool ok = std::fseek(f, (long)(12 + rec * kv.first + 12), SEEK_SET) == 0 &&
              std::fread(v3, sizeof(float), 3, f) == 3 && std::fread(rake.data(), sizeof(float), (size_t)nt, f) == (size_t)nt &&
              std::fread(rt.data(), sizeof(float), (size_t)nt, f) == (size_t)nt;
    if (!ok) {
      std::fclose(f);
      msg = "source file " + path + " is truncated";
      return 1001;
    }
    const float area = v3[0];
    const float s = (float)(v3[1] * REF_DEG2RAD), d = (float)(v3[2] * REF_DEG2RAD);
    t.cell.push_back(kv.second);
    const size_t base = t.rate.size();
    t.rate.resize(base + (size_t)nt * 6);
    for (int n = 0; n < nt; ++n) {
      const float r = p.degree2radian == 0 ? rake[(size_t)n] : (float)(rake[(size_t)n] * REF_DEG2RAD);
      float m[6];
      double_couple(s, d, r, m);
      for (int q = 0; q < 6; ++q) t.rate[base + (size_t)n * 6 + q] = m[q] * area * rt[(size_t)n];
    }
  }
  std::fclose(f);
  return 0;
}

}  // namespace cgh
