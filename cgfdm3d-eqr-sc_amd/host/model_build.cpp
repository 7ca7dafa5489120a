#include "model_build.h"

#include "terrain.h"

#include <cmath>
#include <cstdio>

namespace cgh {

Span rank_span(int N, int P, int r) {
  int n = N / P, res = N % P;
  if (r < res) return {r * (n + 1), n + 1};
  return {res * (n + 1) + (r - res) * n, n};
}

Grid make_grid(const Params& p) {
  Grid g;
  g.nx = p.NX;
  g.ny = p.NY;
  g.nz = p.NZ;
  g.NX = p.NX + 2 * HALO;
  g.NY = p.NY + 2 * HALO;
  g.NZ = p.NZ + 2 * HALO;
  return g;
}

int build_coordinates(const Params& p, const Grid& g, std::vector<float>& coord, std::vector<float>& terrain,
                      std::string& msg, SrtmResult* srtm) {
  const long long N = g.cells();
  coord.assign((size_t)(3 * N), 0.f);
  terrain.assign((size_t)g.plane(), 0.f);
  const float DH = (float)p.DH;
  const int NZ = p.NZ;
  // flat grid in float arithmetic, as the reference kernel computes it
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        long long c = g.idx(i, j, k);
        coord[(size_t)c] = (float)(i - HALO) * DH - (float)p.centerX * DH;
        coord[(size_t)(N + c)] = (float)(j - HALO) * DH - (float)p.centerY * DH;
        coord[(size_t)(2 * N + c)] = (float)(k - HALO + 1) * DH - (float)NZ * DH;
        if (k == NZ + HALO - 1) terrain[(size_t)g.idx(i, j, 0)] = (float)(k - HALO + 1) * DH - (float)NZ * DH;
      }
  if (p.gauss_hill) {
    const float cal_depth = p.Depth * 1000.f;
    const float h = 0.2f * cal_depth, a = 0.1f * cal_depth, b = 0.1f * cal_depth;
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        float x = (float)(i - HALO) * DH - (float)p.centerX * DH;
        float y = (float)(j - HALO) * DH - (float)p.centerY * DH;
        terrain[(size_t)g.idx(i, j, 0)] = h * std::exp(-0.5f * (x * x / (a * a) + y * y / (b * b)));
      }
    for (int k = 0; k < g.NZ; ++k)
      for (int j = 0; j < g.NY; ++j)
        for (int i = 0; i < g.NX; ++i) {
          float height = terrain[(size_t)g.idx(i, j, 0)];
          float DZ = (float)((double)(height + std::fabs(cal_depth)) / (double)(NZ - 1));
          coord[(size_t)(2 * N + g.idx(i, j, k))] = -std::fabs(cal_depth) + DZ * (float)(k - HALO);
        }
  }
  if (p.useTerrain) {
    // preprocessTerrain (terrain.cpp:772-860)
    if (p.SRTM90 == 1) {
      SrtmResult local;
      SrtmResult& r = srtm ? *srtm : local;
      if (int rc = build_srtm_terrain(p, g, r, msg)) return rc;
      terrain = r.terrain;
    }
    terrain_following_z(p, g, terrain, coord);
  }
  return 0;
}

int read_layered_file(const std::string& path, std::vector<float>& depth, std::vector<float>& vs,
                      std::vector<float>& vp, std::vector<float>& rho, std::string& msg) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    msg = "Can't open Layered Model file " + path;
    return 2;
  }
  int layers = 0;
  bool ok = std::fread(&layers, sizeof(int), 1, f) == 1 && layers > 0 && layers < (1 << 20);
  depth.clear();
  vs.clear();
  vp.clear();
  rho.clear();
  for (int F = 0; ok && F < layers; ++F) {
    float v[4];
    ok = std::fread(v, sizeof(float), 4, f) == 4;
    if (ok) {
      depth.push_back(v[0] * 1000.f);
      vs.push_back(v[1] * 1000.f);
      vp.push_back(v[2] * 1000.f);
      rho.push_back(v[3] * 1000.f);
    }
  }
  std::fclose(f);
  if (!ok) {
    msg = "Layered Model file " + path + " is truncated or has no layers";
    return 2;
  }
  return 0;
}

int build_medium(const Params& p, const Grid& g, const std::vector<float>& coord, const std::vector<float>& terrain,
                 std::vector<float>& medium, std::string& msg) {
  const long long N = g.cells();
  medium.assign((size_t)(3 * N), 0.f);
  for (long long c = 0; c < N; ++c) {
    medium[(size_t)c] = 3300.0f;
    medium[(size_t)(N + c)] = 5716.0f;
    medium[(size_t)(2 * N + c)] = 2700.0f;
  }
  if (!p.useMedium) return 0;
  if (p.Crust_1Model) {
    if (int rc = build_crust1_medium(p, g, coord, medium, msg)) return rc;
  }
  if (p.ShenModel) {
    if (int rc = build_shen_medium(p, g, coord, medium, msg)) return rc;
  }
  if (!p.LayeredModel) return 0;
  std::vector<float> H, VS, VP, RHO;
  int rc = read_layered_file(p.LayeredFileName, H, VS, VP, RHO, msg);
  if (rc) return rc;
  const int layers = (int)H.size();
  const float tolerance = 1.0f;
  // the assignment order of generate_1D_medium: later tests overwrite earlier ones
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        long long c = g.idx(i, j, k);
        const double ter = (double)terrain[(size_t)g.idx(i, j, 0)];
        const float coordz = coord[(size_t)(2 * N + c)];
        for (int F = 1; F < layers; ++F) {
          float h0 = (float)(-H[0] + ter), h1 = (float)(-H[F - 1] + ter), h2 = (float)(-H[F] + ter);
          int pick = -1;
          if (coordz > (h0 - tolerance)) pick = 0;
          if ((h1 - tolerance) > coordz && coordz > (h2 - tolerance)) pick = F - 1;
          if ((h2 - tolerance) > coordz) pick = F;
          if (pick >= 0) {
            medium[(size_t)c] = VS[(size_t)pick];
            medium[(size_t)(N + c)] = VP[(size_t)pick];
            medium[(size_t)(2 * N + c)] = RHO[(size_t)pick];
          }
        }
      }
  return 0;
}

}  // namespace cgh
