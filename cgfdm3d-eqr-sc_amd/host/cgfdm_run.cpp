// cgfdm_run: the reference's executable flow on one GPU (main.cpp:17-45,
// run.cpp:28-138, propagate.cpp:76-350): params.json and station.json from the
// working directory, coordinates and medium, source, time loop on the HIP
// library through its C ABI, and the reference's output files.
//
// usage: cgfdm_run [--scheme cgfdm|aweno] [--precision 4|8] [--coef 0|1|2]
//                  [--params FILE] [--stations FILE] [--steps N] [--quiet]
//
// --scheme picks the build the reference makes with Makefile_CGFDM (default) or
// Makefile_SCFDM. --steps replaces NT = TMAX/DT (for short checks). Errors print
// a message and end the process with the reference's MPI_Abort code (100 no
// params.json, 1001 parse error, 110 CFL, 130 PML, 180 source), as an exit
// status modulo 256 and in the message.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cgfdm_amd.h"
#include "io_files.h"
#include "model_build.h"
#include "params.h"
#include "source_build.h"
#include "terrain.h"

using namespace cgh;

namespace {

int die(int code, const std::string& msg) {
  std::fprintf(stderr, "cgfdm_run: %s (code %d)\n", msg.c_str(), code);
  return code & 0xFF ? code & 0xFF : 1;
}

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

}  // namespace

int main(int argc, char** argv) {
  std::string params_path = "params.json", stations_path = "station.json";
  int scheme = CGFDM_SCHEME_CGFDM, precision = 4, coef = 0, steps_override = -1;
  bool quiet = false;
  for (int a = 1; a < argc; ++a) {
    std::string o = argv[a];
    auto next = [&](const char* what) -> const char* {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "cgfdm_run: %s needs a value\n", what);
        std::exit(2);
      }
      return argv[++a];
    };
    if (o == "--scheme") {
      std::string v = next("--scheme");
      if (v == "cgfdm") scheme = CGFDM_SCHEME_CGFDM;
      else if (v == "aweno" || v == "scfdm") scheme = CGFDM_SCHEME_AWENO;
      else return die(2, "unknown scheme " + v);
    } else if (o == "--precision") {
      precision = std::atoi(next("--precision"));
    } else if (o == "--coef") {
      coef = std::atoi(next("--coef"));
    } else if (o == "--params") {
      params_path = next("--params");
    } else if (o == "--stations") {
      stations_path = next("--stations");
    } else if (o == "--steps") {
      steps_override = std::atoi(next("--steps"));
    } else if (o == "--quiet") {
      quiet = true;
    } else {
      std::fprintf(stderr, "usage: cgfdm_run [--scheme cgfdm|aweno] [--precision 4|8] [--coef 0|1|2] "
                           "[--params FILE] [--stations FILE] [--steps N] [--quiet]\n");
      return 2;
    }
  }
  Params p;
  std::string msg;
  int rc = read_params(params_path, p, msg);
  if (rc) return die(rc, msg);
  // modelChecking (modelChecking.cpp:17-68)
  if (p.useSingleSource_ricker && p.useSingleSource_double_couple)
    return die(180, "useSingleSource(ricker) and useSingleSource(double_couple) are both set");
  if (!(p.useMultiSource || p.useSingleSource_ricker || p.useSingleSource_double_couple))
    return die(180, "You did not set any source");
  if (p.useMultiSource && (p.useSingleSource_ricker || p.useSingleSource_double_couple))
    return die(180, "useMultiSource and useSingleSource are both set: give one source description");
  if (p.NX <= 0 || p.NY <= 0 || p.NZ <= 1 || !(p.DH > 0) || !(p.DT > 0) || p.IT_SKIP <= 0)
    return die(1001, "NX, NY, NZ, DH, DT and IT_SKIP must be positive");
  const int NT = steps_override >= 0 ? steps_override : step_count(p);
  mkdir(p.OUT.c_str(), 0777);  // createDir (create_dir.cpp:17-32)

  Grid g = make_grid(p);
  std::vector<float> coord, terrain, medium;
  {
    SrtmResult srtm;
    if ((rc = build_coordinates(p, g, coord, terrain, msg, &srtm))) return die(rc, msg);
    if (p.useTerrain && p.SRTM90 == 1) {
      // outputTotalTerrain, outputLonLat, outputTerrain (terrain.cpp:28-118, 825-833)
      if ((rc = write_total_terrain(p, srtm.total, msg)) || (rc = write_surface_files(p, "lon", srtm.lon, msg)) ||
          (rc = write_surface_files(p, "lat", srtm.lat, msg)) || (rc = write_surface_files(p, "terrain", terrain, msg)))
        return die(rc, msg);
    }
  }
  if ((rc = build_medium(p, g, coord, terrain, medium, msg))) return die(rc, msg);
  if (!quiet) std::printf("grid %d x %d x %d, DH %g, DT %g, NT %d, nPML %d\n", p.NX, p.NY, p.NZ, p.DH, p.DT, NT, p.nPML);

  // data2D_Model_out (data_io.cpp:589-618): coordinates and medium on the slice planes
  {
    const char* cn[3] = {"coordX", "coordY", "coordZ"};
    const char* mn[3] = {"Vs", "Vp", "rho"};
    for (int c = 0; c < 3; ++c) {
      HostPlanes hc{&g, &coord, c}, hm{&g, &medium, c};
      if ((rc = write_slices(p, p.OUT + "/" + cn[c], p.sliceX, p.sliceY, p.sliceZ, host_plane, &hc, msg)))
        return die(rc, msg);
      if ((rc = write_slices(p, p.OUT + "/" + mn[c], p.sliceX, p.sliceY, p.sliceZ, host_plane, &hm, msg)))
        return die(rc, msg);
    }
  }

  cgfdm_desc d{};
  d.nx = p.NX;
  d.ny = p.NY;
  d.nz = p.NZ;
  d.npml = p.nPML;
  d.free_surface = 1;  // FREE_SURFACE is on in both reference builds
  d.scheme = scheme;
  d.precision = precision;
  d.coef_storage = coef;
  d.nt = NT > 0 ? NT : 1;
  d.dh = (double)(float)p.DH;  // GRID.DH is a float in the reference
  d.dt = (double)(float)p.DT;  // the reference steps with DT held in a float
  cgfdm_solver* s = nullptr;
  if ((rc = cgfdm_create(&d, coord.data(), medium.data(), &s))) return die(rc, cgfdm_last_error());

  // source
  SourceTable st;
  if (p.useMultiSource) {
    std::vector<float> xyz;
    if ((rc = fault_source_table(p, g, coord, st, &xyz, msg))) return die(rc, msg);
    if ((rc = write_source_coords(p, xyz, msg))) return die(rc, msg);
    if ((rc = write_source_files(p, g, st.cell, st.rate, st.nt, (float)st.dt, msg))) return die(rc, msg);
  } else {
    std::vector<double> jac((size_t)g.cells()), mu((size_t)g.cells());
    if ((rc = cgfdm_get_setup(s, 9, jac.data())) || (rc = cgfdm_get_setup(s, 10, mu.data())))
      return die(rc, cgfdm_last_error());
    if ((rc = single_source_table(p, g, NT, jac.data(), mu.data(), st, msg))) return die(rc, msg);
  }
  if (st.npts() > 0 &&
      (rc = cgfdm_set_source(s, st.npts(), st.nt, st.dt, st.cell.data(), st.rate.data())))
    return die(rc, cgfdm_last_error());
  if (!quiet) std::printf("source: %d cells, %d samples at dt %g\n", st.npts(), st.nt, st.dt);
  coord.clear();
  coord.shrink_to_fit();
  medium.clear();
  medium.shrink_to_fit();

  // stations
  std::vector<int> gijk;
  if ((rc = read_stations(stations_path, g, gijk, msg))) return die(rc, msg);
  const int ns = (int)gijk.size() / 3;
  if (ns > 0) {
    std::vector<int> ijk(gijk);
    for (int& v : ijk) v += HALO;
    if ((rc = cgfdm_set_stations(s, ns, ijk.data()))) return die(rc, cgfdm_last_error());
  }
  if (!quiet) std::printf("stations: %d\n", ns);

  // time loop (propagate.cpp:186-321): output after step `it` when it % IT_SKIP == 0
  auto t0 = std::chrono::steady_clock::now();
  std::vector<float> dis;
  int it = 0;
  while (it < NT) {
    const int o = ((it + p.IT_SKIP - 1) / p.IT_SKIP) * p.IT_SKIP;  // next output step
    const int last = o < NT ? o : NT - 1;
    if ((rc = cgfdm_step(s, last - it + 1))) return die(rc, cgfdm_last_error());
    it = last + 1;
    if (last != o) break;
    const std::string sfx = "_" + std::to_string(o);
    const char* vn[3] = {"Vx", "Vy", "Vz"};
    for (int c = 0; c < 3; ++c) {
      DevPlanes dp{s, &g, c};
      if ((rc = write_slices(p, p.OUT + "/" + vn[c] + sfx, p.sliceX, p.sliceY, p.sliceZ, dev_plane, &dp, msg)))
        return die(rc, msg);
      if (p.sliceFreeSurf &&
          (rc = write_slices(p, p.OUT + "/FreeSurf" + vn[c] + sfx, -1, -1, p.NZ - 1, dev_plane, &dp, msg)))
        return die(rc, msg);
    }
    if (p.sliceFreeSurf) {
      // SOLVE_DISPLACEMENT output (data_io.cpp:620-633)
      dis.resize((size_t)3 * p.NX * p.NY);
      if ((rc = cgfdm_get_surface_displacement(s, dis.data()))) return die(rc, cgfdm_last_error());
      const char* un[3] = {"FreeSurfUx", "FreeSurfUy", "FreeSurfUz"};
      for (int c = 0; c < 3; ++c) {
        DisPlane dpl{&dis, c, p.NX, p.NY};
        if ((rc = write_slices(p, p.OUT + "/" + un[c] + sfx, -1, -1, p.NZ - 1, dis_plane, &dpl, msg)))
          return die(rc, msg);
      }
    }
    if (!quiet) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("it = %8d. Total time: %8.3lfs.\n", o, el);
      std::fflush(stdout);
    }
  }
  if ((rc = cgfdm_synchronize(s))) return die(rc, cgfdm_last_error());
  double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // stationWrite and outputPGV (station.cpp:422-443, PGV.cpp:177-185)
  if (ns > 0) {
    std::vector<float> rec((size_t)ns * d.nt * 9);
    if ((rc = cgfdm_get_stations(s, rec.data()))) return die(rc, cgfdm_last_error());
    if ((rc = write_station_files(p, gijk, d.nt, rec.data(), msg))) return die(rc, msg);
  }
  std::vector<float> pgv((size_t)2 * p.NX * p.NY);
  if ((rc = cgfdm_get_pgv(s, pgv.data()))) return die(rc, cgfdm_last_error());
  if ((rc = write_pgv_files(p, pgv.data(), msg))) return die(rc, msg);
  cgfdm_destroy(s);
  if (!quiet)
    std::printf("done: %d steps in %.3f s, %.3f Gcell-updates/s\n", NT, el,
                NT > 0 && el > 0 ? (double)p.NX * p.NY * p.NZ * NT / el * 1e-9 : 0.0);
  return 0;
}
