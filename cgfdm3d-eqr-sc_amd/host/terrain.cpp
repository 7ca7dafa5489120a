#include "terrain.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "source_build.h"

namespace cgh {

namespace {

const double REF_PI = (double)3.141592657f;  // macro.h:70
// PROJ returns radians; the reference turns them into degrees with
// RADIAN2DEGREE = 180 / PI of its float PI (macro.h:72)
const double TO_REF_DEG = M_PI / REF_PI;

inline double gauss_fun(double t, double a) { return std::exp(-((t / a) * (t / a))) / (a * 1.772453850905516); }

// the centre printed with "%lf" into the PROJ string
double fmt6(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%lf", v);
  return std::strtod(b, nullptr);
}

// 11 x 11 Gaussian smoothing weights (ra = 2.5) over 0.996768 (terrain.cpp:414-431, 726-748)
struct Smooth {
  double Dg[11];
  Smooth() {
    for (int n = 0; n < 11; ++n) Dg[n] = gauss_fun(n - 5, 2.5);
  }
  double w(int a, int b) const { return Dg[a] * Dg[b] / 0.996768; }
};

}  // namespace

bool aeqd_inverse(double lon0, double lat0, double x, double y, double a, double f, double& lon, double& lat) {
  const double D2R = M_PI / 180.0;
  const double s = std::hypot(x, y);
  if (s == 0.0) {
    lon = lon0;
    lat = lat0;
    return true;
  }
  const double az = std::atan2(x, y);
  const double b = a * (1.0 - f);
  const double sa1 = std::sin(az), ca1 = std::cos(az);
  const double tU1 = (1.0 - f) * std::tan(lat0 * D2R);
  const double cU1 = 1.0 / std::sqrt(1.0 + tU1 * tU1), sU1 = tU1 * cU1;
  const double s1 = std::atan2(tU1, ca1);
  const double sinA = cU1 * sa1;
  const double cos2a = 1.0 - sinA * sinA;
  const double u2 = cos2a * (a * a - b * b) / (b * b);
  const double A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)));
  const double B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)));
  double sig = s / (b * A), sigP, c2sm = 0, sinS = 0, cosS = 0;
  int iter = 0;
  do {
    c2sm = std::cos(2.0 * s1 + sig);
    sinS = std::sin(sig);
    cosS = std::cos(sig);
    const double dS = B * sinS * (c2sm + B / 4.0 * (cosS * (-1.0 + 2.0 * c2sm * c2sm) -
                                                   B / 6.0 * c2sm * (-3.0 + 4.0 * sinS * sinS) *
                                                       (-3.0 + 4.0 * c2sm * c2sm)));
    sigP = sig;
    sig = s / (b * A) + dS;
  } while (std::fabs(sig - sigP) > 1e-13 && ++iter < 500);
  if (iter >= 500) return false;
  c2sm = std::cos(2.0 * s1 + sig);
  sinS = std::sin(sig);
  cosS = std::cos(sig);
  const double tmp = sU1 * sinS - cU1 * cosS * ca1;
  const double phi2 = std::atan2(sU1 * cosS + cU1 * sinS * ca1, (1.0 - f) * std::sqrt(sinA * sinA + tmp * tmp));
  const double lam = std::atan2(sinS * sa1, cU1 * cosS - sU1 * sinS * ca1);
  const double C = f / 16.0 * cos2a * (4.0 + f * (4.0 - 3.0 * cos2a));
  const double L = lam - (1.0 - C) * f * sinA * (sig + C * sinS * (c2sm + C * cosS * (-1.0 + 2.0 * c2sm * c2sm)));
  lat = phi2 / D2R;
  lon = lon0 + L / D2R;
  return true;
}

int build_srtm_terrain(const Params& p, const Grid& g, SrtmResult& out, std::string& msg) {
  const int P = SRTM_POINTS;
  const int TX = (P - 1) * p.blockX + 1, TY = (P - 1) * p.blockY + 1;
  if (p.blockX < 1 || p.blockY < 1) {
    msg = "blockX and blockY must be at least 1";
    return 1001;
  }
  out.totalX = TX;
  out.totalY = TY;
  std::vector<float>& T = out.total;
  T.assign((size_t)TX * TY, 0.f);
  // readSRTM90 (terrain.cpp:206-289): tiles overlap by one row and column; each
  // contributes its first (P-1) x (P-1) points
  {
    std::vector<float> tile((size_t)P * P);
    const int lonEnd = p.lonStart + (p.blockX - 1) * SRTM_DEGREES, latEnd = p.latStart + (p.blockY - 1) * SRTM_DEGREES;
    int j = 0;
    for (int lat = p.latStart; lat <= latEnd; lat += 5, ++j) {
      int i = 0;
      for (int lon = p.lonStart; lon <= lonEnd; lon += 5, ++i) {
        char name[1024];
        std::snprintf(name, sizeof name, "%s/srtm_%dN%dE.bin", p.TerrainDir.c_str(), lat, lon);
        FILE* f = std::fopen(name, "rb");
        if (!f) {
          msg = std::string("There is no such file ") + name;
          return 1001;
        }
        size_t got = std::fread(tile.data(), sizeof(float), tile.size(), f);
        std::fclose(f);
        if (got != tile.size()) {
          msg = std::string("terrain tile ") + name + " is truncated";
          return 1001;
        }
        const int sI = i * (P - 1), sJ = j * (P - 1);
        for (int jj = 0; jj < P - 1; ++jj)
          for (int ii = 0; ii < P - 1; ++ii) T[(size_t)(sI + ii) + (size_t)(sJ + jj) * TX] = tile[(size_t)ii + (size_t)jj * P];
      }
    }
  }
  // resetTotalTerrain (ERROR_ELEVATION, terrain.cpp:291-305)
  for (float& v : T)
    if (v < -32768 + 1.0 || v > 32768 - 1.0) v = 0.f;

  const double deltaLon = double(SRTM_DEGREES * p.blockX) / double(TX);
  const double deltaLat = double(SRTM_DEGREES * p.blockY) / double(TY);
  const double lonStart = p.lonStart, latStart = p.latStart;
  const float DH = (float)p.DH;
  const Smooth sm;

  // callibrateTotalTerrain (terrain.cpp:307-448): despike where the central
  // gradient exceeds 100 m, then smooth, over the tile points under the grid
  {
    const double c0 = fmt6((double)(float)p.centerLongitude), c1 = fmt6((double)(float)p.centerLatitude);
    const double xs[2] = {((0 - HALO) * (double)DH - p.centerX * (double)DH), ((g.NX - HALO) * (double)DH - p.centerX * (double)DH)};
    const double ys[2] = {((0 - HALO) * (double)DH - p.centerY * (double)DH), ((g.NY - HALO) * (double)DH - p.centerY * (double)DH)};
    float maxLon = -1000.f, maxLat = -1000.f, minLon = 1000.f, minLat = 1000.f;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        double lo, la;
        if (!aeqd_inverse(c0, c1, xs[a], ys[b], WGS84_A, WGS84_F, lo, la)) {
          msg = "inverse projection did not converge";
          return 1001;
        }
        const float L0 = (float)(lo * TO_REF_DEG), L1 = (float)(la * TO_REF_DEG);
        maxLon = std::max(maxLon, L0);
        maxLat = std::max(maxLat, L1);
        minLon = std::min(minLon, L0);
        minLat = std::min(minLat, L1);
      }
    // region of the grid, widened by 10 points (clamped to the map, which the
    // reference does not check)
    int I0 = int((minLon - lonStart) / deltaLon) - 10, J0 = int((minLat - latStart) / deltaLat) - 10;
    int I1 = int((maxLon - lonStart) / deltaLon) + 10, J1 = int((maxLat - latStart) / deltaLat) + 10;
    I0 = std::max(I0, 5);
    J0 = std::max(J0, 5);
    I1 = std::min(I1, TX - 6);
    J1 = std::min(J1, TY - 6);
    std::vector<float> N(T);
    for (int j = J0; j < J1; ++j)
      for (int i = I0; i < I1; ++i) {
        const float a = T[(size_t)(i + 1) + (size_t)j * TX], b = T[(size_t)(i - 1) + (size_t)j * TX];
        const float c = T[(size_t)i + (size_t)(j + 1) * TX], d = T[(size_t)i + (size_t)(j - 1) * TX];
        const float gx = a - b, gy = c - d;
        if (std::sqrt(gx * gx + gy * gy) > 100.0f) N[(size_t)i + (size_t)j * TX] = (a + b + c + d) * 0.25f;
      }
    T = N;
    for (int j = J0; j < J1; ++j)
      for (int i = I0; i < I1; ++i) {
        double sum = 0.0;
        for (int jj = j - 5; jj <= j + 5; ++jj)
          for (int ii = i - 5; ii <= i + 5; ++ii) sum += sm.w(ii + 5 - i, jj + 5 - j) * T[(size_t)ii + (size_t)jj * TX];
        N[(size_t)i + (size_t)j * TX] = (float)sum;
      }
    T.swap(N);
  }

  // projTrans (terrain.cpp:490-546): surface longitude / latitude, and the
  // grid's corners (widened by 2 cells) must lie on the tiles
  {
    const double c0 = fmt6(p.centerLongitude), c1 = fmt6(p.centerLatitude);
    const int oX = p.centerX + HALO, oY = p.centerY + HALO;
    const double xs[2] = {-(oX + 2) * (double)DH, ((g.NX - oX) + 2) * (double)DH};
    const double ys[2] = {-(oY + 2) * (double)DH, ((g.NY - oY) + 2) * (double)DH};
    const float lonMax = (float)(p.lonStart + p.blockX * SRTM_DEGREES), latMax = (float)(p.latStart + p.blockY * SRTM_DEGREES);
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        double lo, la;
        aeqd_inverse(c0, c1, xs[a], ys[b], WGS84_A, WGS84_F, lo, la);
        const float L0 = (float)(lo * TO_REF_DEG), L1 = (float)(la * TO_REF_DEG);
        if (L0 > lonMax || L0 < (float)p.lonStart || L1 > latMax || L1 < (float)p.latStart) {
          msg = "The longitude or latitude range of the calculation extends beyond the terrain tiles";
          return 1001;
        }
      }
    out.lon.assign((size_t)g.plane(), 0.f);
    out.lat.assign((size_t)g.plane(), 0.f);
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const double x = (double)((float)(i - HALO) * DH - (float)p.centerX * DH);
        const double y = (double)((float)(j - HALO) * DH - (float)p.centerY * DH);
        double lo, la;
        aeqd_inverse(c0, c1, x, y, WGS84_A, WGS84_F, lo, la);
        out.lon[(size_t)g.idx(i, j, 0)] = (float)(lo * TO_REF_DEG);
        out.lat[(size_t)g.idx(i, j, 0)] = (float)(la * TO_REF_DEG);
      }
  }

  // terrainInterp (terrain.cpp:602-770): the map at the surface points and 5
  // cells beyond, then the 11 x 11 smoothing onto the grid
  {
    const double c0 = fmt6((double)(float)p.centerLongitude), c1 = fmt6((double)(float)p.centerLatitude);
    const int nG = 5, ex = g.NX + 2 * nG, ey = g.NY + 2 * nG;
    std::vector<float> E((size_t)ex * ey);
    for (int j = 0; j < ey; ++j)
      for (int i = 0; i < ex; ++i) {
        const int I = i - nG, J = j - nG;
        const double x = (double)((float)((I - HALO) * DH) - (float)(p.centerX * DH));
        const double y = (double)((float)((J - HALO) * DH) - (float)(p.centerY * DH));
        double lo, la;
        aeqd_inverse(c0, c1, x, y, WGS84_A, WGS84_F, lo, la);
        const double xq = lo * TO_REF_DEG, yq = la * TO_REF_DEG;
        int Ii = int((xq - lonStart) / deltaLon), Jj = int((yq - latStart) / deltaLat);
        Ii = std::max(0, std::min(Ii, TX - 2));
        Jj = std::max(0, std::min(Jj, TY - 2));
        const double x1 = Ii * deltaLon + lonStart, y1 = Jj * deltaLat + latStart;
        const double x2 = x1 + deltaLon, y2 = y1 + deltaLat;
        // corner values as the reference loads them: A = (I, J), B = (I+1, J)
        // and D = (I+1, J+1); its (x1, y2) corner also reads (I+1, J+1)
        const double zA = T[(size_t)Ii + (size_t)Jj * TX];
        const double zB = T[(size_t)(Ii + 1) + (size_t)Jj * TX];
        const double zD = T[(size_t)(Ii + 1) + (size_t)(Jj + 1) * TX];
        const double Lx0 = (xq - x2) / (x1 - x2), Lx1 = (xq - x1) / (x2 - x1);
        const double Ly0 = (yq - y2) / (y1 - y2), Ly1 = (yq - y1) / (y2 - y1);
        // interp2d: sum over z[i*2+j] Lx[i] Ly[j] with z = {A, D(as x1,y2), B, D}
        E[(size_t)i + (size_t)j * ex] = (float)(Lx0 * Ly0 * zA + Lx0 * Ly1 * zD + Lx1 * Ly0 * zB + Lx1 * Ly1 * zD);
      }
    out.terrain.assign((size_t)g.plane(), 0.f);
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        double sum = 0.0;
        for (int jj = j; jj <= j + 2 * nG; ++jj)
          for (int ii = i; ii <= i + 2 * nG; ++ii) sum += sm.w(ii - i, jj - j) * E[(size_t)ii + (size_t)jj * ex];
        out.terrain[(size_t)g.idx(i, j, 0)] = (float)sum;
      }
  }
  return 0;
}

int surface_lonlat(const Params& p, const Grid& g, float lon0, float lon1, float lat0, float lat1,
                   std::vector<double>& lon, std::vector<double>& lat, std::string& msg) {
  const double c0 = fmt6(p.centerLongitude), c1 = fmt6(p.centerLatitude);
  const float DH = (float)p.DH;
  const int oX = p.centerX + HALO, oY = p.centerY + HALO;
  const double xs[2] = {-(oX + 2) * (double)DH, ((g.NX - oX) + 2) * (double)DH};
  const double ys[2] = {-(oY + 2) * (double)DH, ((g.NY - oY) + 2) * (double)DH};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      double lo, la;
      aeqd_inverse(c0, c1, xs[a], ys[b], WGS84_A, WGS84_F, lo, la);
      const float L0 = (float)(lo * TO_REF_DEG), L1 = (float)(la * TO_REF_DEG);
      if (L0 > lon1 || L0 < lon0 || L1 > lat1 || L1 < lat0) {
        msg = "The longitude or latitude range of the calculation extends beyond the terrain or medium data";
        return 1001;
      }
    }
  lon.assign((size_t)g.plane(), 0.0);
  lat.assign((size_t)g.plane(), 0.0);
  for (int j = 0; j < g.NY; ++j)
    for (int i = 0; i < g.NX; ++i) {
      const double x = (double)((float)(i - HALO) * DH - (float)p.centerX * DH);
      const double y = (double)((float)(j - HALO) * DH - (float)p.centerY * DH);
      double lo, la;
      aeqd_inverse(c0, c1, x, y, WGS84_A, WGS84_F, lo, la);
      lon[(size_t)g.idx(i, j, 0)] = lo * TO_REF_DEG;
      lat[(size_t)g.idx(i, j, 0)] = la * TO_REF_DEG;
    }
  return 0;
}

int build_shen_medium(const Params& p, const Grid& g, const std::vector<float>& coord, std::vector<float>& medium,
                      std::string& msg) {
  const float LonStart = p.MLonStart, LatStart = p.MLatStart, LonEnd = p.MLonEnd, LatEnd = p.MLatEnd;
  const float LonStep = p.MLonStep, LatStep = p.MLatStep;
  if (!(LonStep > 0) || !(LatStep > 0) || !(p.MVeticalStep > 0)) {
    msg = "MLonStep, MLatStep and MVeticalStep must be positive";
    return 1001;
  }
  const int nLon = int((LonEnd - LonStart) / LonStep + 0.5 + 1), nLat = int((LatEnd - LatStart) / LatStep + 0.5 + 1);
  const float depthAll = p.Depth * 1e3f + 3e4f;  // Deeper (dealMedium.cpp:201)
  const int NLayer = int(depthAll / p.MVeticalStep);
  std::vector<float> VM((size_t)nLon * nLat * NLayer);
  for (int j = 0; j < nLat; ++j)
    for (int i = 0; i < nLon; ++i) {
      char ln[64], la[64], name[1024];
      std::snprintf(ln, sizeof ln, "%.1f", LonStart + i * LonStep);
      std::snprintf(la, sizeof la, "%.1f", LatStart + j * LatStep);
      std::snprintf(name, sizeof name, "%s/%s_%s.mod", p.MediumDir.c_str(), ln, la);
      FILE* f = std::fopen(name, "r");
      if (!f) {
        msg = std::string("Can't open file: ") + name;
        return 1001;
      }
      for (int K = 0; K < NLayer; ++K) {
        double dep, vs, err;
        if (std::fscanf(f, "%lf %lf %lf", &dep, &vs, &err) != 3) {
          std::fclose(f);
          msg = std::string("model file ") + name + " has fewer than " + std::to_string(NLayer) + " layers";
          return 1001;
        }
        VM[(size_t)K + (size_t)(i + j * nLon) * NLayer] = (float)(vs * 1e3);
      }
      std::fclose(f);
    }
  std::vector<double> lon, lat;
  if (int rc = surface_lonlat(p, g, LonStart, LonEnd, LatStart, LatEnd, lon, lat, msg)) return rc;
  // structureInterp: bilinear per layer at the surface points
  const size_t P2 = (size_t)g.plane();
  std::vector<float> Vs(P2 * NLayer);
  for (int K = 0; K < NLayer; ++K)
    for (size_t q = 0; q < P2; ++q) {
      const double xq = lon[q], yq = lat[q];
      int I = int((xq - LonStart) / LonStep), J = int((yq - LatStart) / LatStep);
      I = std::max(0, std::min(I, nLon - 2));
      J = std::max(0, std::min(J, nLat - 2));
      const double x1 = I * LonStep + LonStart, y1 = J * LatStep + LatStart, x2 = x1 + LonStep, y2 = y1 + LatStep;
      auto V = [&](int a, int b) { return (double)VM[(size_t)K + (size_t)(a + b * nLon) * NLayer]; };
      const double Lx0 = (xq - x2) / (x1 - x2), Lx1 = (xq - x1) / (x2 - x1);
      const double Ly0 = (yq - y2) / (y1 - y2), Ly1 = (yq - y1) / (y2 - y1);
      Vs[(size_t)K * P2 + q] =
          (float)(Lx0 * Ly0 * V(I, J) + Lx0 * Ly1 * V(I, J + 1) + Lx1 * Ly0 * V(I + 1, J) + Lx1 * Ly1 * V(I + 1, J + 1));
    }
  if (NLayer > 1)  // DealWithFirstLayer
    for (size_t q = 0; q < P2; ++q) Vs[q] = Vs[P2 + q];
  const long long N = g.cells();
  const double Depth = (double)p.Depth * 1e3;
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const long long c = g.idx(i, j, k);
        // StructureTerrain: depth below the column's surface DZ (NZ - 1) - |Depth|
        const double DZ = (double)(float)(coord[(size_t)(2 * N + g.idx(i, j, 4))] - coord[(size_t)(2 * N + g.idx(i, j, 3))]);
        const double ter = DZ * double(p.NZ - 1) - std::fabs(Depth);
        int K = (int)((float)int(-(coord[(size_t)(2 * N + c)] - ter)) / p.MVeticalStep);
        if (K <= 0) K = 0;
        if (K >= NLayer) K = NLayer - 1;
        const float VS = Vs[(size_t)K * P2 + (size_t)g.idx(i, j, 0)] * 1e-3f;
        const float VP = (float)(0.9409 + 2.0947 * VS - 0.8206 * std::pow(VS, 2) + 0.2683 * std::pow(VS, 3) -
                                 0.0251 * std::pow(VS, 4));
        const float RHO = (float)(1.6612 * VP - 0.4721 * std::pow(VP, 2) + 0.0671 * std::pow(VP, 3) -
                                  0.0043 * std::pow(VP, 4) + 0.000106 * std::pow(VP, 5));
        medium[(size_t)c] = VS * 1e3f;
        medium[(size_t)(N + c)] = VP * 1e3f;
        medium[(size_t)(2 * N + c)] = RHO * 1e3f;
      }
  return 0;
}

int build_crust1_medium(const Params& p, const Grid& g, const std::vector<float>& coord, std::vector<float>& medium,
                        std::string& msg) {
  const int nLon = 360, nLat = 180, nL = 9, nLL = nLon * nLat;
  const char* names[4] = {"crust1.vs", "crust1.vp", "crust1.rho", "crust1.bnds"};
  std::vector<float> F[4];
  for (int q = 0; q < 4; ++q) {
    const std::string path = p.crustDir + "/" + names[q];
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) {
      msg = "Can't open file: " + path;
      return 1001;
    }
    F[q].assign((size_t)nLL * nL, 0.f);
    for (int i = 0; i < nLL; ++i)
      for (int n = 0; n < nL; ++n)
        if (std::fscanf(f, "%f", &F[q][(size_t)i + (size_t)n * nLL]) != 1) {
          std::fclose(f);
          msg = path + " is truncated";
          return 1001;
        }
    std::fclose(f);
  }
  std::vector<double> lon, lat;
  if (int rc = surface_lonlat(p, g, -180.f, 180.f, -90.f, 90.f, lon, lat, msg)) return rc;
  // crustInterpFunc: bilinear between the cell centres around each surface point
  const float LonStart = -179.5f, LatStart = 89.5f, LonStep = p.CrustLonStep, LatStep = p.CrustLatStep;
  if (!(LonStep > 0) || !(LatStep > 0)) {
    msg = "CrustLonStep and CrustLatStep must be positive";
    return 1001;
  }
  const size_t P2 = (size_t)g.plane();
  std::vector<float> I4[4];
  for (int q = 0; q < 4; ++q) I4[q].assign(P2 * nL, 0.f);
  for (int K = 0; K < nL; ++K)
    for (size_t s2 = 0; s2 < P2; ++s2) {
      const double xq = lon[s2], yq = lat[s2];
      int I = int((xq - LonStart) / LonStep), J = int((LatStart - yq) / LatStep) + 1;
      I = std::max(0, std::min(I, nLon - 2));
      J = std::max(1, std::min(J, nLat - 1));
      const double x1 = I * LonStep + LonStart, y1 = LatStart - J * LatStep, x2 = x1 + LonStep, y2 = y1 + LatStep;
      const double Lx0 = (xq - x2) / (x1 - x2), Lx1 = (xq - x1) / (x2 - x1);
      const double Ly0 = (yq - y2) / (y1 - y2), Ly1 = (yq - y1) / (y2 - y1);
      for (int q = 0; q < 4; ++q) {
        auto V = [&](int a, int b) { return (double)F[q][(size_t)(a + b * nLon) + (size_t)K * nLL]; };
        I4[q][(size_t)K * P2 + s2] =
            (float)(Lx0 * Ly0 * V(I, J) + Lx0 * Ly1 * V(I, J - 1) + Lx1 * Ly0 * V(I + 1, J) + Lx1 * Ly1 * V(I + 1, J - 1));
      }
    }
  // solveStructure: boundaries in metres; the properties carry over from the
  // previous cell where no rule applies, as in the reference
  for (float& b : I4[3]) b *= 1e3f;
  const long long N = g.cells();
  float vs = 0, vp = 0, rho = 0;
  auto bnd = [&](int l, size_t s2) { return I4[3][(size_t)l * P2 + s2]; };
  auto take = [&](int l, size_t s2) {
    vs = I4[0][(size_t)l * P2 + s2];
    vp = I4[1][(size_t)l * P2 + s2];
    rho = I4[2][(size_t)l * P2 + s2];
  };
  // first layer below `m` whose top differs from layer m's by 1 km or more, minus one
  auto skip_thin = [&](int m, size_t s2) {
    int l = m;
    while (l < nL && std::fabs(bnd(m, s2) - bnd(l, s2)) < 1000.0f) ++l;
    return l - 1;
  };
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const size_t s2 = (size_t)g.idx(i, j, 0);
        const long long c = g.idx(i, j, k);
        const float z = coord[(size_t)(2 * N + c)];
        if (z > bnd(0, s2)) take(skip_thin(0, s2), s2);
        if (z <= bnd(nL - 1, s2)) take(nL - 1, s2);
        if (z <= bnd(0, s2) && z > bnd(nL - 1, s2))
          for (int m = 0; m < nL - 1; ++m)
            if (z <= bnd(m, s2) && z > bnd(m + 1, s2)) take(skip_thin(m, s2), s2);
        medium[(size_t)c] = vs * 1e3f;
        medium[(size_t)(N + c)] = vp * 1e3f;
        medium[(size_t)(2 * N + c)] = rho * 1e3f;
      }
  return 0;
}

void terrain_following_z(const Params& p, const Grid& g, const std::vector<float>& terrain, std::vector<float>& coord) {
  const long long N = g.cells();
  const double Depth = (double)p.Depth * 1e3;
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const double DZ = double(terrain[(size_t)g.idx(i, j, 0)] + std::fabs(Depth)) / double(p.NZ - 1);
        coord[(size_t)(2 * N + g.idx(i, j, k))] = (float)(-std::fabs(Depth) + DZ * (k - HALO));
      }
}

}  // namespace cgh
