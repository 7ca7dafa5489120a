// Source descriptions of the run parameters turned into the moment-rate table
// that cgfdm_set_source takes: rate[npts][nt][6] (Mxx, Myy, Mzz, Mxy, Mxz, Myz)
// at padded cells, injected as W[3+q] -= rate * mu / (J DH^3) * DT at every step.
#ifndef CGFDM_HOST_SOURCE_BUILD_H
#define CGFDM_HOST_SOURCE_BUILD_H

#include <map>
#include <string>
#include <vector>

#include "model_build.h"
#include "params.h"

namespace cgh {

struct SourceTable {
  int nt = 0;
  double dt = 0;
  std::vector<long long> cell;  // padded linear index of the whole-domain grid
  std::vector<float> rate;      // [npts][nt][6]
  int npts() const { return (int)cell.size(); }
};

// Reference double-couple radiation pattern in its axis order (singleSource.cpp:273-285,
// multiSource.cpp:476-494): m = (M22, M11, M33, M12, -M23, -M13) of the
// Aki-Richards tensor, i.e. (Mxx, Myy, Mzz, Mxy, Mxz, Myz) with x east, y north, z up.
// Angles in radians; evaluated in float as the reference does.
void double_couple(float strike, float dip, float rake, float m[6]);

// Single point sources (singleSource.cpp:226-431). Both spread the source over the
// 7x7x7 cells around (sourceX, sourceY, sourceZ) with the Gaussian weights
// GAUSS_FUN(d, 1.5, 0) per axis / 0.998125703461425 and scale by 1/(J_c DH^3) with
// the Jacobian J_c of the centre cell. Ricker (load_smooth_source_ricker): an
// isotropic source s(t) = Ricker(fc) * 1e9 on Txx, Tyy, Tzz. Double couple
// (load_smooth_source_double_couple): m0 = 1.5 * 10^(Mw + 6) times a Gaussian
// time function of the given duration, radiation pattern from strike/dip/rake.
// The reference evaluates the source at t = it * DT and adds it to W before the
// RK stages; the table holds those values at t_n = n * DT, n = 0..nsteps, so the
// step's interpolation reads them unchanged. `jac` and `mu` are the setup's J
// and mu planes (double[NZ][NY][NX]); mu enters only to cancel the factor that
// the library applies to every table entry. Cells outside the interior are left
// out (the reference writes them into the ghost layer). Returns 0 or 180 with
// a message when the source lies outside the grid.
int single_source_table(const Params& p, const Grid& g, int nsteps, const double* jac, const double* mu,
                        SourceTable& t, std::string& msg);

// ---- finite fault file (multiSource.cpp) ----------------------------------
// File: int npts, int nt, float dt; then per point float lon, lat, z, area,
// strike, dip, rake[nt], rate[nt] (multiSource.cpp:24-94, 345-396).
struct FaultHeader {
  long long npts = 0;
  int nt = 0;
  float dt = 0;
};
int read_fault_header(const std::string& path, FaultHeader& h, std::string& msg);
// lon, lat, z of every point (readSourceLocation)
int read_fault_locations(const std::string& path, const FaultHeader& h, std::vector<float>& lon,
                         std::vector<float>& lat, std::vector<float>& z, std::string& msg);

// Azimuthal equidistant projection centred on (lon0, lat0) on an ellipsoid of
// semi-major axis a and flattening f (the reference calls PROJ with
// "+proj=aeqd +lon_0 +lat_0 +x_0=0 +y_0=0 +ellps=WGS84", multiSource.cpp:122-171;
// PROJ is not available here). x = s sin(az), y = s cos(az) with s and az the
// geodesic distance and forward azimuth from the centre, from Vincenty's inverse
// method (T. Vincenty, Survey Review 23(176), 1975), iterated to 1e-13 rad.
// Angles in degrees, output in metres. Returns false when the iteration does not
// converge (nearly antipodal points).
bool aeqd_forward(double lon0, double lat0, double lon, double lat, double a, double f, double& x, double& y);
constexpr double WGS84_A = 6378137.0;
constexpr double WGS84_F = 1.0 / 298.257223563;

// Point -> padded cell of the whole-domain grid (srcCoord2Index,
// multiSource.cpp:173-314, the NO_SOURCE_SMOOTH branch that the reference's
// macro.h selects): depths are shifted so the shallowest point is at 0, then
// lifted by the column's surface height DZ*(NZ-1) - |Depth|, and matched to the
// cell whose half-spacing interval holds them. x, y are the projected
// coordinates (modified in place for z, as the reference does). Returns the map
// point number -> cell for points inside the grid.
std::map<long long, long long> fault_point_cells(const Grid& g, const Params& p, const std::vector<float>& coord,
                                                 const std::vector<float>& x, const std::vector<float>& y,
                                                 std::vector<float>& z);

// Whole reader: file -> projection -> cells -> moment rates (solveMomentRate,
// multiSource.cpp:450-535: radiation pattern * area * rate(t), rake in degrees
// unless degree2radian is 0). Duplicate cells are kept as separate points (the
// library sums them, multiSource.cpp:537-592). `xyz` (optional) receives the
// projected and shifted point coordinates that the reference writes to
// source_coord_{X,Y,Z}.bin.
int fault_source_table(const Params& p, const Grid& g, const std::vector<float>& coord, SourceTable& t,
                       std::vector<float>* xyz, std::string& msg);

}  // namespace cgh

#endif
