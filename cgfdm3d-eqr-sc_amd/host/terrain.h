// SRTM terrain for the terrain-following grid (reference terrain.cpp:206-860):
// 5-degree tiles of 6000 x 6000 float elevations, calibration and smoothing
// of the assembled map, interpolation onto the grid's surface through the
// inverse azimuthal-equidistant projection, and the terrain-following
// vertical coordinate.
#ifndef CGFDM_HOST_TERRAIN_H
#define CGFDM_HOST_TERRAIN_H

#include <string>
#include <vector>

#include "model_build.h"
#include "params.h"

namespace cgh {

constexpr int SRTM_POINTS = 6000;  // pointsPerBlockX/Y (terrain.cpp:20-21)
constexpr int SRTM_DEGREES = 5;    // degreePerBlcokX/Y (terrain.cpp:17-18)

// Inverse of aeqd_forward: the point at distance hypot(x, y) from (lon0, lat0)
// along the azimuth atan2(x, y), from Vincenty's direct method (T. Vincenty,
// Survey Review 23(176), 1975), iterated to 1e-13 rad. Degrees in and out.
bool aeqd_inverse(double lon0, double lat0, double x, double y, double a, double f, double& lon, double& lat);

struct SrtmResult {
  std::vector<float> terrain;       // surface height [NY][NX] of the padded grid
  std::vector<float> lon, lat;      // surface longitude / latitude [NY][NX] (projTrans, terrain.cpp:490-546)
  std::vector<float> total;         // calibrated tile map [totalY][totalX] (totalTerrain.bin)
  int totalX = 0, totalY = 0;
};

// readSRTM90 + resetTotalTerrain + callibrateTotalTerrain + projTrans +
// terrainInterp (terrain.cpp:206-770) for the one-rank grid. Tiles are read
// from "<TerrainDir>/srtm_<lat>N<lon>E.bin". Returns 0, or 1001 with a message
// for a missing tile or a grid outside the tiles (the reference's abort code).
int build_srtm_terrain(const Params& p, const Grid& g, SrtmResult& out, std::string& msg);

// Surface longitude/latitude of the grid points (projTrans / cart2LonLat,
// terrain.cpp:450-546) in the reference's degree convention, and its check that
// the grid's corners, widened by 2 cells, lie inside [lon0, lon1] x [lat0, lat1]
// (1001 otherwise).
int surface_lonlat(const Params& p, const Grid& g, float lon0, float lon1, float lat0, float lat1,
                   std::vector<double>& lon, std::vector<double>& lat, std::string& msg);

// Shen model medium (readWeisenShenModel / structureInterp, dealMedium.cpp:49-360):
// "<MediumDir>/<lon>_<lat>.mod" text profiles of (depth, Vs, error) every
// MVeticalStep metres on the MLonStart..MLonEnd x MLatStart..MLatEnd node grid,
// interpolated to the surface points layer by layer, the first layer replaced by
// the second (DealWithFirstLayer), looked up by the depth below the local surface
// (StructureTerrain) and turned into Vp and rho by the empirical relations of the
// reference. Returns 0, or 1001 with a message (missing file, grid outside).
int build_shen_medium(const Params& p, const Grid& g, const std::vector<float>& coord, std::vector<float>& medium,
                      std::string& msg);

// CRUST1.0 medium (readCrustal_1 / crustInterpFunc / solveStructure,
// crustMedium.cpp): "<CrustDir>/crust1.{vs,vp,rho,bnds}" text files of 9 layer
// values per 1-degree cell (360 x 180, longitude fastest from -179.5, latitude
// rows from 89.5 southwards), bilinear at the surface points, then per grid
// cell the properties of the layer whose top lies above it, skipping layers
// thinner than 1 km as the reference does. Returns 0 or 1001 with a message.
int build_crust1_medium(const Params& p, const Grid& g, const std::vector<float>& coord, std::vector<float>& medium,
                        std::string& msg);

// Terrain-following z of preprocessTerrain (terrain.cpp:838-858):
// z = -|Depth| + (h + |Depth|) / (NZ - 1) * (k - 3), with h the surface height.
void terrain_following_z(const Params& p, const Grid& g, const std::vector<float>& terrain, std::vector<float>& coord);

}  // namespace cgh

#endif
