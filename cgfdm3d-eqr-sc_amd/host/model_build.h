// Grid, coordinates and medium from the run parameters, for one GPU holding
// the whole domain (the reference's P = 1 grid), plus the reference's block
// decomposition so that output files can be split as PX x PY x PZ ranks write
// them.
#ifndef CGFDM_HOST_MODEL_BUILD_H
#define CGFDM_HOST_MODEL_BUILD_H

#include <string>
#include <vector>

#include "params.h"

namespace cgh {

constexpr int HALO = 3;

// Extent of one rank of the reference's decomposition along one axis
// (init_grid.cpp:17-68): N / P cells, the remainder going to the first ranks.
struct Span {
  int front;  // first global interior index of the rank
  int n;      // interior cells of the rank
};
Span rank_span(int N, int P, int r);

// Padded whole-domain grid of the run.
struct Grid {
  int nx = 0, ny = 0, nz = 0;  // interior cells
  int NX = 0, NY = 0, NZ = 0;  // padded: n + 2*HALO
  long long plane() const { return (long long)NX * NY; }
  long long cells() const { return plane() * NZ; }
  long long idx(int i, int j, int k) const { return (long long)i + (long long)NX * j + plane() * k; }
};
Grid make_grid(const Params& p);

struct SrtmResult;

// Coordinates float[3][NZ][NY][NX] (x, y, z in metres) and the surface height
// float[NY][NX] of the padded grid (coord.cpp:259-333): the flat grid
// (construct_flat_coord, :100-135), optionally bent to the Gaussian hill
// (construct_gauss_hill_surface :137-170 and construct_terrain_coord :205-257),
// and with useTerrain the terrain-following grid of preprocessTerrain
// (terrain.cpp:772-860) over the SRTM surface when SRTM90 is 1 (`srtm`, when
// given, receives the surface longitude/latitude and the tile map). Returns 0
// or the reference's abort code with a message.
int build_coordinates(const Params& p, const Grid& g, std::vector<float>& coord, std::vector<float>& terrain,
                      std::string& msg, SrtmResult* srtm = nullptr);

// Medium float[3][NZ][NY][NX] (Vs, Vp, rho) (medium.cpp:251-317): homogeneous
// Vs 3300, Vp 5716, rho 2700 (construct_homo_medium :84-105), then with
// useMedium the CRUST1.0 model, the Shen model and the layered model file in
// the reference's order, each overwriting the previous (readCrustal_1,
// readWeisenShenModel, generate_1D_medium :107-213 with layers measured from the
// local surface height as the reference build does with LayeredStructureTerrain).
// Returns 0 or the reference's abort code with a message.
int build_medium(const Params& p, const Grid& g, const std::vector<float>& coord, const std::vector<float>& terrain,
                 std::vector<float>& medium, std::string& msg);

// Layered model file: int layers, then per layer float (top depth km, Vs km/s,
// Vp km/s, rho g/cm^3) (medium.cpp:122-146). Values returned in m, m/s, kg/m^3.
int read_layered_file(const std::string& path, std::vector<float>& depth, std::vector<float>& vs,
                      std::vector<float>& vp, std::vector<float>& rho, std::string& msg);

}  // namespace cgh

#endif
