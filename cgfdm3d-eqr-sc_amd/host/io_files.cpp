#include "io_files.h"

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

#include "json_lite.h"

namespace cgh {

namespace {

bool write_file(const std::string& path, const void* a, size_t na, const void* b = nullptr, size_t nb = 0) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = (na == 0 || std::fwrite(a, 1, na, f) == na) && (nb == 0 || std::fwrite(b, 1, nb, f) == nb);
  return std::fclose(f) == 0 && ok;
}

}  // namespace

std::string rank_suffix(const Rank& r) {
  return "_mpi_" + std::to_string(r.X) + "_" + std::to_string(r.Y) + "_" + std::to_string(r.Z) + ".bin";
}

std::vector<Rank> ranks_of(const Params& p) {
  std::vector<Rank> out;
  const int PX = p.PX > 0 ? p.PX : 1, PY = p.PY > 0 ? p.PY : 1, PZ = p.PZ > 0 ? p.PZ : 1;
  for (int X = 0; X < PX; ++X)
    for (int Y = 0; Y < PY; ++Y)
      for (int Z = 0; Z < PZ; ++Z)
        out.push_back(Rank{X, Y, Z, rank_span(p.NX, PX, X), rank_span(p.NY, PY, Y), rank_span(p.NZ, PZ, Z)});
  return out;
}

int read_stations(const std::string& path, const Grid& g, std::vector<int>& gijk, std::string& msg) {
  gijk.clear();
  std::ifstream f(path, std::ios::binary);
  if (!f) return 0;  // "There is not station.json file!": no stations
  std::stringstream ss;
  ss << f.rdbuf();
  JValue root;
  std::string err;
  if (!parse_json(ss.str(), root, err)) {
    msg = "Can't parse " + path + ": " + err;
    return 1001;
  }
  const JValue* arr = root.get("station(point)");
  if (!arr) return 0;
  for (int q = 0; q < arr->size(); ++q) {
    const JValue* s = arr->at(q);
    if (s->size() != 3) {
      msg = "In file " + path + ", the coordinate index count is " + std::to_string(s->size()) + ", not 3";
      return 1001;
    }
    int X = s->at(0)->as_int(), Y = s->at(1)->as_int(), Z = s->at(2)->as_int();
    if (X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz) {
      gijk.push_back(X);
      gijk.push_back(Y);
      gijk.push_back(Z);
    }
  }
  return 0;
}

int write_station_files(const Params& p, const std::vector<int>& gijk, int NT, const float* rec, std::string& msg) {
  const int ns = (int)gijk.size() / 3;
  for (const Rank& r : ranks_of(p)) {
    std::vector<int> xyz;
    std::vector<float> wave;
    for (int s = 0; s < ns; ++s) {
      if (!r.has(gijk[3 * s], gijk[3 * s + 1], gijk[3 * s + 2])) continue;
      xyz.insert(xyz.end(), &gijk[3 * s], &gijk[3 * s] + 3);
      wave.insert(wave.end(), rec + (size_t)s * NT * 9, rec + (size_t)(s + 1) * NT * 9);
    }
    if (xyz.empty()) continue;
    std::string path = p.OUT + "/station" + rank_suffix(r);
    if (!write_file(path, xyz.data(), xyz.size() * sizeof(int), wave.data(), wave.size() * sizeof(float))) {
      msg = "cannot write " + path;
      return 2;
    }
  }
  return 0;
}

int write_pgv_files(const Params& p, const float* pgv, std::string& msg) {
  const int PZ = p.PZ > 0 ? p.PZ : 1;
  for (const Rank& r : ranks_of(p)) {
    if (r.Z != PZ - 1) continue;
    std::vector<float> v((size_t)r.sx.n * r.sy.n * 2);
    for (int j = 0; j < r.sy.n; ++j)
      for (int i = 0; i < r.sx.n; ++i)
        for (int c = 0; c < 2; ++c)
          v[((size_t)j * r.sx.n + i) * 2 + c] = pgv[(((size_t)(r.sy.front + j)) * p.NX + (r.sx.front + i)) * 2 + c];
    std::string path = p.OUT + "/PGV_Z" + rank_suffix(r);
    if (!write_file(path, v.data(), v.size() * sizeof(float))) {
      msg = "cannot write " + path;
      return 2;
    }
  }
  return 0;
}

int write_slices(const Params& p, const std::string& name, int gx, int gy, int gz, PlaneFn get, void* ctx,
                 std::string& msg) {
  const int g[3] = {gx, gy, gz};
  const char tag[3] = {'X', 'Y', 'Z'};
  const int NN[3] = {p.NX, p.NY, p.NZ};
  std::vector<Rank> rk = ranks_of(p);
  for (int ax = 0; ax < 3; ++ax) {
    if (g[ax] < 0 || g[ax] >= NN[ax]) continue;
    std::vector<float> plane;
    if (!get(ctx, ax, g[ax], plane)) {
      msg = "cannot read the plane of slice " + name;
      return 2;
    }
    // in-plane axes (fast, slow) and their whole-domain extents
    const int a0 = ax == 0 ? 1 : 0, a1 = ax == 2 ? 1 : 2;
    const int W = NN[a0];
    for (const Rank& r : rk) {
      const Span sp[3] = {r.sx, r.sy, r.sz};
      if (g[ax] < sp[ax].front || g[ax] >= sp[ax].front + sp[ax].n) continue;
      std::vector<float> v((size_t)sp[a0].n * sp[a1].n);
      for (int b = 0; b < sp[a1].n; ++b)
        for (int a = 0; a < sp[a0].n; ++a)
          v[(size_t)b * sp[a0].n + a] = plane[(size_t)(sp[a1].front + b) * W + (sp[a0].front + a)];
      std::string path = name + "_" + tag[ax] + rank_suffix(r);
      if (!write_file(path, v.data(), v.size() * sizeof(float))) {
        msg = "cannot write " + path;
        return 2;
      }
    }
  }
  return 0;
}

int write_source_files(const Params& p, const Grid& g, const std::vector<long long>& cell,
                       const std::vector<float>& rate, int nt, float dt, std::string& msg) {
  for (const Rank& r : ranks_of(p)) {
    // merge duplicate cells by summation, keyed by the rank-local padded index
    const int lnx = r.sx.n + 2 * HALO, lny = r.sy.n + 2 * HALO;
    std::map<long long, std::vector<float>> m;
    for (size_t q = 0; q < cell.size(); ++q) {
      long long c = cell[q];
      int i = (int)(c % g.NX), j = (int)((c / g.NX) % g.NY), k = (int)(c / g.plane());
      int gi = i - HALO, gj = j - HALO, gk = k - HALO;
      if (!r.has(gi, gj, gk)) continue;
      long long li = (long long)(gi - r.sx.front + HALO) + (long long)lnx * ((gj - r.sy.front + HALO) +
                                                                              (long long)lny * (gk - r.sz.front + HALO));
      auto& v = m[li];
      if (v.empty()) v.assign((size_t)nt * 6, 0.f);
      for (int t = 0; t < nt; ++t)
        for (int c6 = 0; c6 < 6; ++c6) v[(size_t)c6 * nt + t] += rate[(q * nt + t) * 6 + c6];
    }
    if (m.empty()) continue;
    std::string path = p.OUT + "/source" + rank_suffix(r);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
      msg = "cannot write " + path;
      return 2;
    }
    int n = (int)m.size();
    bool ok = std::fwrite(&n, sizeof(int), 1, f) == 1 && std::fwrite(&nt, sizeof(int), 1, f) == 1 &&
              std::fwrite(&dt, sizeof(float), 1, f) == 1;
    for (auto& kv : m) {
      ok = ok && std::fwrite(&kv.first, sizeof(long long), 1, f) == 1 &&
           std::fwrite(kv.second.data(), sizeof(float), kv.second.size(), f) == kv.second.size();
    }
    if (std::fclose(f) != 0 || !ok) {
      msg = "cannot write " + path;
      return 2;
    }
  }
  return 0;
}

int write_surface_files(const Params& p, const std::string& name, const std::vector<float>& plane, std::string& msg) {
  const int PZ = p.PZ > 0 ? p.PZ : 1, NXp = p.NX + 2 * HALO;
  for (const Rank& r : ranks_of(p)) {
    if (r.Z != PZ - 1) continue;
    std::vector<float> v((size_t)r.sx.n * r.sy.n);
    for (int j = 0; j < r.sy.n; ++j)
      for (int i = 0; i < r.sx.n; ++i)
        v[(size_t)j * r.sx.n + i] = plane[(size_t)(r.sx.front + i + HALO) + (size_t)NXp * (r.sy.front + j + HALO)];
    std::string path = p.OUT + "/" + name + rank_suffix(r);
    if (!write_file(path, v.data(), v.size() * sizeof(float))) {
      msg = "cannot write " + path;
      return 2;
    }
  }
  return 0;
}

int write_total_terrain(const Params& p, const std::vector<float>& total, std::string& msg) {
  std::string path = p.OUT + "/totalTerrain.bin";
  if (!write_file(path, total.data(), total.size() * sizeof(float))) {
    msg = "cannot write " + path;
    return 2;
  }
  return 0;
}

int write_source_coords(const Params& p, const std::vector<float>& xyz, std::string& msg) {
  const size_t n = xyz.size() / 3;
  const char* tag[3] = {"X", "Y", "Z"};
  for (int c = 0; c < 3; ++c) {
    std::string path = p.OUT + "/source_coord_" + tag[c] + ".bin";
    if (!write_file(path, xyz.data() + c * n, n * sizeof(float))) {
      msg = "cannot write " + path;
      return 2;
    }
  }
  return 0;
}

}  // namespace cgh
