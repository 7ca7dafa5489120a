// Run parameters read from params.json, with the reference's key names,
// field types and defaults (reference struct.h:21-125, getParams.cpp:17-235).
#ifndef CGFDM_HOST_PARAMS_H
#define CGFDM_HOST_PARAMS_H

#include <string>

namespace cgh {

// Field types follow the reference PARAMS struct: a key the reference stores in
// a float is rounded to float here too (rickerfc, strike, Depth ...). Missing
// keys keep the zero defaults below (the reference leaves them uninitialised).
struct Params {
  double TMAX = 0, DT = 0, DH = 0;
  int NX = 0, NY = 0, NZ = 0;
  int PX = 1, PY = 1, PZ = 1;
  int centerX = 0, centerY = 0;
  double centerLatitude = 0, centerLongitude = 0;
  int sourceX = 0, sourceY = 0, sourceZ = 0;
  int IT_SKIP = 1;
  int sliceX = 0, sliceY = 0, sliceZ = 0, sliceFreeSurf = 0;
  int nPML = 0;
  int gauss_hill = 0, useTerrain = 0, useMedium = 0, useMultiSource = 0;
  int useSingleSource_ricker = 0;
  float rickerfc = 0;
  int useSingleSource_double_couple = 0;
  float strike = 0, dip = 0, rake = 0, Mw = 0, duration = 0;
  int ShenModel = 0, Crust_1Model = 0, LayeredModel = 0;
  std::string LayeredFileName;
  int itSlice = 0, itStep = 0;
  std::string waveOutput, sliceName;
  int itStart = 0, itEnd = 0, igpu = 0;
  std::string OUT;
  std::string TerrainDir;
  int SRTM90 = 0;
  int lonStart = 0, latStart = 0, blockX = 0, blockY = 0;
  float Depth = 0;  // km
  float MLonStart = 0, MLatStart = 0, MLonEnd = 0, MLatEnd = 0, MLonStep = 0, MLatStep = 0;
  float CrustLonStep = 0, CrustLatStep = 0, MVeticalStep = 0;
  std::string MediumDir, crustDir;
  std::string sourceFile, sourceDir;
  int degree2radian = 0;
};

// Read a params.json file. Returns 0, 100 when the file cannot be opened
// (getParams.cpp:25-29) or 1001 when it does not parse (getParams.cpp:48-53);
// `msg` receives the reason.
int read_params(const std::string& path, Params& p, std::string& msg);
// Same from the text of the file.
int parse_params(const std::string& text, Params& p, std::string& msg);

// Number of time steps, NT = TMAX / DT with DT held in a float (propagate.cpp:86-88).
int step_count(const Params& p);

}  // namespace cgh

#endif
