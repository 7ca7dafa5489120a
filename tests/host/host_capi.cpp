// C entry points over the host-side sources of cgfdm_run (parameter and station
// files, grid and medium builders, sources, output writers), for the CPU tests.
// Built by tests/test_host_driver.py with g++; no GPU and no HIP library.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "io_files.h"
#include "model_build.h"
#include "params.h"
#include "source_build.h"
#include "terrain.h"

using namespace cgh;

namespace {
std::string g_msg;
int put(const std::string& s, char* buf, int n) {
  if (n > 0) {
    std::snprintf(buf, (size_t)n, "%s", s.c_str());
  }
  return (int)s.size();
}
}  // namespace

extern "C" {

const char* h_last_msg() { return g_msg.c_str(); }

// parsed parameters as a JSON object (numbers and strings)
int h_params(const char* path, char* buf, int n) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return -rc;
  char t[8192];
  std::snprintf(t, sizeof t,
                "{\"TMAX\":%.17g,\"DT\":%.17g,\"DH\":%.17g,\"NX\":%d,\"NY\":%d,\"NZ\":%d,\"PX\":%d,\"PY\":%d,\"PZ\":%d,"
                "\"centerX\":%d,\"centerY\":%d,\"centerLongitude\":%.17g,\"centerLatitude\":%.17g,\"IT_SKIP\":%d,"
                "\"sliceX\":%d,\"sliceY\":%d,\"sliceZ\":%d,\"sliceFreeSurf\":%d,\"nPML\":%d,\"gauss_hill\":%d,"
                "\"useTerrain\":%d,\"useMedium\":%d,\"useMultiSource\":%d,\"LayeredModel\":%d,"
                "\"LayeredFileName\":\"%s\",\"ricker\":%d,\"double_couple\":%d,\"strike\":%.9g,\"dip\":%.9g,"
                "\"rake\":%.9g,\"Mw\":%.9g,\"duration\":%.9g,\"rickerfc\":%.9g,\"OUT\":\"%s\",\"SRTM90\":%d,"
                "\"Depth\":%.9g,\"sourceX\":%d,\"sourceY\":%d,\"sourceZ\":%d,\"sourceFile\":\"%s\","
                "\"sourceDir\":\"%s\",\"degree2radian\":%d,\"NT\":%d,\"MLonStep\":%.9g,\"CrustDir\":\"%s\"}",
                p.TMAX, p.DT, p.DH, p.NX, p.NY, p.NZ, p.PX, p.PY, p.PZ, p.centerX, p.centerY, p.centerLongitude,
                p.centerLatitude, p.IT_SKIP, p.sliceX, p.sliceY, p.sliceZ, p.sliceFreeSurf, p.nPML, p.gauss_hill,
                p.useTerrain, p.useMedium, p.useMultiSource, p.LayeredModel, p.LayeredFileName.c_str(),
                p.useSingleSource_ricker, p.useSingleSource_double_couple, p.strike, p.dip, p.rake, p.Mw, p.duration,
                p.rickerfc, p.OUT.c_str(), p.SRTM90, p.Depth, p.sourceX, p.sourceY, p.sourceZ, p.sourceFile.c_str(),
                p.sourceDir.c_str(), p.degree2radian, step_count(p), p.MLonStep, p.crustDir.c_str());
  return put(t, buf, n);
}

// coordinates [3][NZ][NY][NX], terrain [NY][NX], medium [3][NZ][NY][NX]
int h_model(const char* path, float* coord, float* terrain, float* medium) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  Grid g = make_grid(p);
  std::vector<float> c, t, m;
  if ((rc = build_coordinates(p, g, c, t, g_msg))) return rc;
  if ((rc = build_medium(p, g, c, t, m, g_msg))) return rc;
  std::memcpy(coord, c.data(), c.size() * 4);
  std::memcpy(terrain, t.data(), t.size() * 4);
  std::memcpy(medium, m.data(), m.size() * 4);
  return 0;
}

int h_aeqd(double lon0, double lat0, double lon, double lat, double a, double f, double* xy) {
  return aeqd_forward(lon0, lat0, lon, lat, a, f, xy[0], xy[1]) ? 0 : 1;
}

int h_aeqd_inv(double lon0, double lat0, double x, double y, double a, double f, double* ll) {
  return aeqd_inverse(lon0, lat0, x, y, a, f, ll[0], ll[1]) ? 0 : 1;
}

void h_double_couple(float s, float d, float r, float* m) { double_couple(s, d, r, m); }

// single point source table; cells/rate sized for 343 points of nsteps+1 samples
int h_single_source(const char* path, int nsteps, const double* jac, const double* mu, long long* cells, float* rate,
                    int* npts, int* nt, double* dt) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  Grid g = make_grid(p);
  SourceTable t;
  if ((rc = single_source_table(p, g, nsteps, jac, mu, t, g_msg))) return rc;
  *npts = t.npts();
  *nt = t.nt;
  *dt = t.dt;
  std::memcpy(cells, t.cell.data(), t.cell.size() * 8);
  std::memcpy(rate, t.rate.data(), t.rate.size() * 4);
  return 0;
}

// finite-fault table; xyz receives 3*npts_file floats
int h_fault_source(const char* path, long long* cells, float* rate, int* npts, int* nt, double* dt, float* xyz,
                   int cap_pts) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  Grid g = make_grid(p);
  std::vector<float> c, t, m;
  if ((rc = build_coordinates(p, g, c, t, g_msg))) return rc;
  SourceTable st;
  std::vector<float> x;
  if ((rc = fault_source_table(p, g, c, st, &x, g_msg))) return rc;
  if (st.npts() > cap_pts) return -1;
  *npts = st.npts();
  *nt = st.nt;
  *dt = st.dt;
  std::memcpy(cells, st.cell.data(), st.cell.size() * 8);
  std::memcpy(rate, st.rate.data(), st.rate.size() * 4);
  std::memcpy(xyz, x.data(), x.size() * 4);
  return 0;
}

int h_stations(const char* params_path, const char* path, int* gijk, int cap) {
  Params p;
  int rc = read_params(params_path, p, g_msg);
  if (rc) return -rc;
  Grid g = make_grid(p);
  std::vector<int> v;
  if ((rc = read_stations(path, g, v, g_msg))) return -rc;
  if ((int)v.size() > 3 * cap) return -1;
  std::memcpy(gijk, v.data(), v.size() * 4);
  return (int)v.size() / 3;
}

// writers: station files from rec[ns][NT][9], PGV from pgv[NY][NX][2],
// a slice set from a whole-domain volume vol[NZ][NY][NX] (interior only)
int h_write_stations(const char* path, const int* gijk, int ns, int NT, const float* rec) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  std::vector<int> v(gijk, gijk + 3 * ns);
  return write_station_files(p, v, NT, rec, g_msg);
}

int h_write_pgv(const char* path, const float* pgv) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  return write_pgv_files(p, pgv, g_msg);
}

struct Vol {
  const float* v;
  int nx, ny, nz;
};
static bool vol_plane(void* ctx, int axis, int gi, std::vector<float>& out) {
  const Vol& V = *(const Vol*)ctx;
  auto at = [&](int i, int j, int k) { return V.v[(size_t)i + (size_t)V.nx * (j + (size_t)V.ny * k)]; };
  if (axis == 0) {
    out.resize((size_t)V.ny * V.nz);
    for (int k = 0; k < V.nz; ++k)
      for (int j = 0; j < V.ny; ++j) out[(size_t)k * V.ny + j] = at(gi, j, k);
  } else if (axis == 1) {
    out.resize((size_t)V.nx * V.nz);
    for (int k = 0; k < V.nz; ++k)
      for (int i = 0; i < V.nx; ++i) out[(size_t)k * V.nx + i] = at(i, gi, k);
  } else {
    out.resize((size_t)V.nx * V.ny);
    for (int j = 0; j < V.ny; ++j)
      for (int i = 0; i < V.nx; ++i) out[(size_t)j * V.nx + i] = at(i, j, gi);
  }
  return true;
}

int h_write_slices(const char* path, const char* name, const float* vol) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  Vol V{vol, p.NX, p.NY, p.NZ};
  return write_slices(p, name, p.sliceX, p.sliceY, p.sliceZ, vol_plane, &V, g_msg);
}

int h_write_source_files(const char* path, const long long* cells, const float* rate, int npts, int nt, float dt) {
  Params p;
  int rc = read_params(path, p, g_msg);
  if (rc) return rc;
  Grid g = make_grid(p);
  std::vector<long long> c(cells, cells + npts);
  std::vector<float> r(rate, rate + (size_t)npts * nt * 6);
  return write_source_files(p, g, c, r, nt, dt, g_msg);
}

}  // extern "C"
