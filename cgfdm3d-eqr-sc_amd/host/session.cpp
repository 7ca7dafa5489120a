// The reference's setup from its input files and its output files behind the C
// ABI: cgfdm_create_from_params does what run.cpp:28-138 does before the time
// loop (params.json, station.json, coordinates, terrain, medium, source, the
// setup output files) and cgfdm_write_outputs writes what the time loop and
// its end write (slices every IT_SKIP steps, data_io.cpp:485-633; stations,
// station.cpp:422-443; PGV, PGV.cpp:177-185), in the reference's byte layouts
// split over its PX x PY x PZ ranks. The description a solver was created from
// is kept beside the solver until cgfdm_destroy.
#include <sys/stat.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cgfdm_amd.h"
#include "io_files.h"
#include "model_build.h"
#include "params.h"
#include "source_build.h"
#include "terrain.h"

namespace cgfdm_detail {
int set_error(int code, const std::string& msg);  // capi.cpp: the thread's last error
void release_session(cgfdm_solver* s);
}  // namespace cgfdm_detail

namespace {

using namespace cgh;

struct Session {
  Params p;
  Grid g;
  std::vector<int> gijk;  // stations in global indices, station.json order
  int nt = 0;             // steps the station record holds
};

std::mutex g_mu;
std::map<const cgfdm_solver*, std::unique_ptr<Session>> g_sessions;

Session* session_of(const cgfdm_solver* s) {
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_sessions.find(s);
  return it == g_sessions.end() ? nullptr : it->second.get();
}

int err(int code, const std::string& msg) { return cgfdm_detail::set_error(code, msg); }

struct HostPlanes {
  const Grid* g;
  const std::vector<float>* a;  // [3][NZ][NY][NX]
  int comp;
};

// whole-domain interior plane of a host array, in the slice layouts
bool host_plane(void* ctx, int axis, int gindex, std::vector<float>& out) {
  const HostPlanes& h = *(const HostPlanes*)ctx;
  const Grid& g = *h.g;
  const float* base = h.a->data() + (size_t)h.comp * g.cells();
  const int c = gindex + HALO;
  if (axis == 0) {
    out.resize((size_t)g.ny * g.nz);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j) out[(size_t)k * g.ny + j] = base[g.idx(c, j + HALO, k + HALO)];
  } else if (axis == 1) {
    out.resize((size_t)g.nx * g.nz);
    for (int k = 0; k < g.nz; ++k)
      for (int i = 0; i < g.nx; ++i) out[(size_t)k * g.nx + i] = base[g.idx(i + HALO, c, k + HALO)];
  } else {
    out.resize((size_t)g.nx * g.ny);
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) out[(size_t)j * g.nx + i] = base[g.idx(i + HALO, j + HALO, c)];
  }
  return true;
}

// whole-domain interior plane of the solver's current state
struct DevPlanes {
  cgfdm_solver* s;
  const Grid* g;
  int comp;
};

bool dev_plane(void* ctx, int axis, int gindex, std::vector<float>& out) {
  const DevPlanes& d = *(const DevPlanes*)ctx;
  const Grid& G = *d.g;
  out.resize(axis == 0 ? (size_t)G.ny * G.nz : axis == 1 ? (size_t)G.nx * G.nz : (size_t)G.nx * G.ny);
  return cgfdm_get_slice(d.s, d.comp, axis, gindex + HALO, out.data()) == CGFDM_OK;
}

struct DisPlane {
  const std::vector<float>* dis;  // [3][ny][nx]
  int comp, nx, ny;
};

bool dis_plane(void* ctx, int axis, int, std::vector<float>& out) {
  const DisPlane& d = *(const DisPlane*)ctx;
  if (axis != 2) return false;
  out.assign(d.dis->begin() + (size_t)d.comp * d.nx * d.ny, d.dis->begin() + (size_t)(d.comp + 1) * d.nx * d.ny);
  return true;
}

int create(const char* params_path, const char* stations_path, int scheme, int precision, int coef_storage,
           int nsteps, cgfdm_solver** out) {
  if (!out) return err(CGFDM_ERR_ARG, "null output handle");
  *out = nullptr;
  auto ses = std::make_unique<Session>();
  Params& p = ses->p;
  std::string msg;
  int rc = read_params(params_path ? params_path : "params.json", p, msg);
  if (rc) return err(rc, msg);
  // modelChecking (modelChecking.cpp:17-68)
  if (p.useSingleSource_ricker && p.useSingleSource_double_couple)
    return err(CGFDM_ERR_SOURCE, "useSingleSource(ricker) and useSingleSource(double_couple) are both set");
  if (!(p.useMultiSource || p.useSingleSource_ricker || p.useSingleSource_double_couple))
    return err(CGFDM_ERR_SOURCE, "You did not set any source");
  if (p.useMultiSource && (p.useSingleSource_ricker || p.useSingleSource_double_couple))
    return err(CGFDM_ERR_SOURCE, "useMultiSource and useSingleSource are both set: give one source description");
  if (p.NX <= 0 || p.NY <= 0 || p.NZ <= 1 || !(p.DH > 0) || !(p.DT > 0) || p.IT_SKIP <= 0)
    return err(CGFDM_ERR_PARSE, "NX, NY, NZ, DH, DT and IT_SKIP must be positive");
  ses->nt = nsteps >= 0 ? nsteps : step_count(p);
  mkdir(p.OUT.c_str(), 0777);  // createDir (create_dir.cpp:17-32)

  Grid& g = ses->g;
  g = make_grid(p);
  std::vector<float> coord, terrain, medium;
  {
    SrtmResult srtm;
    if ((rc = build_coordinates(p, g, coord, terrain, msg, &srtm))) return err(rc, msg);
    if (p.useTerrain && p.SRTM90 == 1) {
      // outputTotalTerrain, outputLonLat, outputTerrain (terrain.cpp:28-118, 825-833)
      if ((rc = write_total_terrain(p, srtm.total, msg)) || (rc = write_surface_files(p, "lon", srtm.lon, msg)) ||
          (rc = write_surface_files(p, "lat", srtm.lat, msg)) || (rc = write_surface_files(p, "terrain", terrain, msg)))
        return err(rc, msg);
    }
  }
  if ((rc = build_medium(p, g, coord, terrain, medium, msg))) return err(rc, msg);

  // data2D_Model_out (data_io.cpp:589-618): coordinates and medium on the slice planes
  const char* cn[3] = {"coordX", "coordY", "coordZ"};
  const char* mn[3] = {"Vs", "Vp", "rho"};
  for (int c = 0; c < 3; ++c) {
    HostPlanes hc{&g, &coord, c}, hm{&g, &medium, c};
    if ((rc = write_slices(p, p.OUT + "/" + cn[c], p.sliceX, p.sliceY, p.sliceZ, host_plane, &hc, msg)))
      return err(rc, msg);
    if ((rc = write_slices(p, p.OUT + "/" + mn[c], p.sliceX, p.sliceY, p.sliceZ, host_plane, &hm, msg)))
      return err(rc, msg);
  }

  cgfdm_desc d{};
  d.nx = p.NX;
  d.ny = p.NY;
  d.nz = p.NZ;
  d.npml = p.nPML;
  d.free_surface = 1;  // FREE_SURFACE is on in both reference builds
  d.scheme = scheme;
  d.precision = precision;
  d.coef_storage = coef_storage;
  d.nt = ses->nt > 0 ? ses->nt : 1;
  d.dh = (double)(float)p.DH;  // GRID.DH is a float in the reference
  d.dt = (double)(float)p.DT;  // the reference steps with DT held in a float
  cgfdm_solver* s = nullptr;
  if ((rc = cgfdm_create(&d, coord.data(), medium.data(), &s))) return rc;  // message set by cgfdm_create
  auto fail_destroy = [&](int code, const std::string& m) {
    cgfdm_destroy(s);
    return m.empty() ? code : err(code, m);
  };

  // source (init_MultiSource, multiSource.cpp:780-868; singleSource.cpp)
  SourceTable st;
  if (p.useMultiSource) {
    std::vector<float> xyz;
    if ((rc = fault_source_table(p, g, coord, st, &xyz, msg))) return fail_destroy(rc, msg);
    if ((rc = write_source_coords(p, xyz, msg))) return fail_destroy(rc, msg);
    if ((rc = write_source_files(p, g, st.cell, st.rate, st.nt, (float)st.dt, msg))) return fail_destroy(rc, msg);
  } else {
    std::vector<double> jac((size_t)g.cells()), mu((size_t)g.cells());
    if ((rc = cgfdm_get_setup(s, 9, jac.data())) || (rc = cgfdm_get_setup(s, 10, mu.data())))
      return fail_destroy(rc, "");
    if ((rc = single_source_table(p, g, ses->nt, jac.data(), mu.data(), st, msg))) return fail_destroy(rc, msg);
  }
  if (st.npts() > 0 && (rc = cgfdm_set_source(s, st.npts(), st.nt, st.dt, st.cell.data(), st.rate.data())))
    return fail_destroy(rc, "");

  // stations (station.cpp:92-301)
  if ((rc = read_stations(stations_path ? stations_path : "station.json", g, ses->gijk, msg)))
    return fail_destroy(rc, msg);
  const int ns = (int)ses->gijk.size() / 3;
  if (ns > 0) {
    std::vector<int> ijk(ses->gijk);
    for (int& v : ijk) v += HALO;
    if ((rc = cgfdm_set_stations(s, ns, ijk.data()))) return fail_destroy(rc, "");
  }
  {
    std::lock_guard<std::mutex> lock(g_mu);
    g_sessions[s] = std::move(ses);
  }
  *out = s;
  return CGFDM_OK;
}

int write_outputs(cgfdm_solver* s, int what) {
  Session* ses = session_of(s);
  if (!ses) return err(CGFDM_ERR_ARG, "the solver was not created by cgfdm_create_from_params");
  const Params& p = ses->p;
  const Grid& g = ses->g;
  std::string msg;
  int rc;
  if (what & CGFDM_OUT_SLICES) {
    const std::string sfx = "_" + std::to_string(cgfdm_step_index(s) - 1);
    const char* vn[3] = {"Vx", "Vy", "Vz"};
    for (int c = 0; c < 3; ++c) {
      DevPlanes dp{s, &g, c};
      if ((rc = write_slices(p, p.OUT + "/" + vn[c] + sfx, p.sliceX, p.sliceY, p.sliceZ, dev_plane, &dp, msg)))
        return err(rc, msg);
      if (p.sliceFreeSurf &&
          (rc = write_slices(p, p.OUT + "/FreeSurf" + vn[c] + sfx, -1, -1, p.NZ - 1, dev_plane, &dp, msg)))
        return err(rc, msg);
    }
    if (p.sliceFreeSurf) {
      // SOLVE_DISPLACEMENT output (data_io.cpp:620-633)
      std::vector<float> dis((size_t)3 * p.NX * p.NY);
      if ((rc = cgfdm_get_surface_displacement(s, dis.data()))) return rc;
      const char* un[3] = {"FreeSurfUx", "FreeSurfUy", "FreeSurfUz"};
      for (int c = 0; c < 3; ++c) {
        DisPlane dpl{&dis, c, p.NX, p.NY};
        if ((rc = write_slices(p, p.OUT + "/" + un[c] + sfx, -1, -1, p.NZ - 1, dis_plane, &dpl, msg)))
          return err(rc, msg);
      }
    }
  }
  if (what & CGFDM_OUT_STATIONS) {
    const int ns = (int)ses->gijk.size() / 3;
    if (ns > 0) {
      const int nt = ses->nt > 0 ? ses->nt : 1;
      std::vector<float> rec((size_t)ns * nt * 9);
      if ((rc = cgfdm_get_stations(s, rec.data()))) return rc;
      if ((rc = write_station_files(p, ses->gijk, nt, rec.data(), msg))) return err(rc, msg);
    }
  }
  if (what & CGFDM_OUT_PGV) {
    std::vector<float> pgv((size_t)2 * p.NX * p.NY);
    if ((rc = cgfdm_get_pgv(s, pgv.data()))) return rc;
    if ((rc = write_pgv_files(p, pgv.data(), msg))) return err(rc, msg);
  }
  return CGFDM_OK;
}

}  // namespace

namespace cgfdm_detail {
void release_session(cgfdm_solver* s) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_sessions.erase(s);
}
}  // namespace cgfdm_detail

extern "C" {

int cgfdm_create_from_params(const char* params_path, const char* stations_path, int scheme, int precision,
                             int coef_storage, int nsteps, cgfdm_solver** out) {
  try {
    return create(params_path, stations_path, scheme, precision, coef_storage, nsteps, out);
  } catch (const std::exception& e) {
    return err(CGFDM_ERR_ALLOC, std::string("host setup failed: ") + e.what());
  }
}

int cgfdm_write_outputs(cgfdm_solver* s, int what) {
  if (!s) return err(CGFDM_ERR_ARG, "null solver");
  try {
    return write_outputs(s, what);
  } catch (const std::exception& e) {
    return err(CGFDM_ERR_ALLOC, std::string("output failed: ") + e.what());
  }
}

int cgfdm_run_info(cgfdm_solver* s, int* nt, int* it_skip) {
  Session* ses = session_of(s);
  if (!ses) return err(CGFDM_ERR_ARG, "the solver was not created by cgfdm_create_from_params");
  if (nt) *nt = ses->nt;
  if (it_skip) *it_skip = ses->p.IT_SKIP;
  return CGFDM_OK;
}

}  // extern "C"
