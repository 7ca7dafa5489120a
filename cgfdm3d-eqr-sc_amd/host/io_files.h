// station.json reading and the output files of a run, in the byte layouts the
// reference writes, split over the PX x PY x PZ ranks of its decomposition so
// that its post-processing scripts read them unchanged.
#ifndef CGFDM_HOST_IO_FILES_H
#define CGFDM_HOST_IO_FILES_H

#include <string>
#include <vector>

#include "model_build.h"
#include "params.h"

namespace cgh {

// Stations of station.json ("station(point)": members or elements of three
// global indices X, Y, Z), kept when inside the grid (readStationIndex /
// initStationIndex, station.cpp:92-301). Returns 0 (also when the file is
// missing: the run then has no stations, as in the reference), 1001 when it
// does not parse or an entry does not have three indices.
int read_stations(const std::string& path, const Grid& g, std::vector<int>& gijk, std::string& msg);

// One rank of the reference decomposition.
struct Rank {
  int X, Y, Z;      // MPI coordinates
  Span sx, sy, sz;  // global interior ranges
  bool has(int gi, int gj, int gk) const {
    return gi >= sx.front && gi < sx.front + sx.n && gj >= sy.front && gj < sy.front + sy.n && gk >= sz.front &&
           gk < sz.front + sz.n;
  }
};
std::vector<Rank> ranks_of(const Params& p);

// station_mpi_X_Y_Z.bin of every rank holding stations (stationWrite,
// station.cpp:422-443): int XYZ[n][3] global indices, then float
// wave[n][NT][9] (Vx, Vy, Vz, Txx, Tyy, Tzz, Txy, Txz, Tyz). `rec` is the whole
// run's [ns][NT][9] record in the order of `gijk`.
int write_station_files(const Params& p, const std::vector<int>& gijk, int NT, const float* rec, std::string& msg);

// PGV_Z_mpi_X_Y_Z.bin of the surface ranks (outputPgvData, PGV.cpp:177-185):
// float[ny][nx][2] (PGVh, PGV) of the rank's columns. `pgv` is the whole
// surface, float[NY][NX][2] of the interior.
int write_pgv_files(const Params& p, const float* pgv, std::string& msg);

// Slice files "<name>_{X,Y,Z}_mpi_X_Y_Z.bin" (data2D_output_bin,
// data_io.cpp:332-433) of the ranks that contain the slice planes: X slice
// float[nz][ny], Y slice float[nz][nx], Z slice float[ny][nx]. `get` returns
// the whole-domain plane of the interior for an axis (0, 1, 2) at a global
// index, in the same layouts. gx, gy, gz < 0 skip that axis.
using PlaneFn = bool (*)(void* ctx, int axis, int gindex, std::vector<float>& out);
int write_slices(const Params& p, const std::string& name, int gx, int gy, int gz, PlaneFn get, void* ctx,
                 std::string& msg);

// source_mpi_X_Y_Z.bin of every rank holding source cells (outputSourceData,
// multiSource.cpp:594-639): int n, int nt, float dt, then per merged cell (in
// increasing rank-local padded index) long long index and float
// Mxx[nt], Myy[nt], Mzz[nt], Mxy[nt], Mxz[nt], Myz[nt].
int write_source_files(const Params& p, const Grid& g, const std::vector<long long>& cell,
                       const std::vector<float>& rate, int nt, float dt, std::string& msg);

// source_coord_{X,Y,Z}.bin (verifyLocation, multiSource.cpp:641-672): float[npts] each.
int write_source_coords(const Params& p, const std::vector<float>& xyz, std::string& msg);

// "<OUT>/<name>_mpi_X_Y_Z.bin" of the surface ranks: float[ny][nx] of the rank's
// interior columns from a padded plane [NY][NX] (outputTerrain / outputLonLat,
// terrain.cpp:42-118), and "<OUT>/totalTerrain.bin" (outputTotalTerrain :28-40).
int write_surface_files(const Params& p, const std::string& name, const std::vector<float>& plane, std::string& msg);
int write_total_terrain(const Params& p, const std::vector<float>& total, std::string& msg);

std::string rank_suffix(const Rank& r);  // "_mpi_X_Y_Z.bin"

}  // namespace cgh

#endif
