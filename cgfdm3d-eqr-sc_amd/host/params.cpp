#include "params.h"

#include <fstream>
#include <sstream>

#include "json_lite.h"

namespace cgh {

namespace {

struct Fill {
  const JValue& o;
  std::string bad;
  void i(const char* k, int& v) {
    if (const JValue* x = o.get(k)) v = x->as_int();
  }
  void d(const char* k, double& v) {
    if (const JValue* x = o.get(k)) v = x->as_double();
  }
  void f(const char* k, float& v) {
    if (const JValue* x = o.get(k)) v = (float)x->as_double();
  }
  void s(const char* k, std::string& v) {
    if (const JValue* x = o.get(k)) {
      // the reference copies valuestring, which a non-string member does not have
      if (x->kind != JValue::Str) bad = std::string("key \"") + k + "\" must be a string";
      else v = x->str;
    }
  }
};

}  // namespace

int parse_params(const std::string& text, Params& p, std::string& msg) {
  JValue root;
  std::string err;
  if (!parse_json(text, root, err) || root.kind != JValue::Obj) {
    msg = "Can't parse json file! " + err;
    return 1001;
  }
  Fill F{root, {}};
  // key order and names of getParams.cpp:58-232
  F.d("TMAX", p.TMAX);
  F.d("DT", p.DT);
  F.d("DH", p.DH);
  F.i("NX", p.NX);
  F.i("NY", p.NY);
  F.i("NZ", p.NZ);
  F.i("PX", p.PX);
  F.i("PY", p.PY);
  F.i("PZ", p.PZ);
  F.i("centerX", p.centerX);
  F.i("centerY", p.centerY);
  F.d("centerLongitude", p.centerLongitude);
  F.d("centerLatitude", p.centerLatitude);
  F.i("IT_SKIP", p.IT_SKIP);
  F.i("sliceX", p.sliceX);
  F.i("sliceY", p.sliceY);
  F.i("sliceZ", p.sliceZ);
  F.i("sliceFreeSurf", p.sliceFreeSurf);
  F.i("nPML", p.nPML);
  F.i("gauss_hill", p.gauss_hill);
  F.i("useTerrain", p.useTerrain);
  F.i("useMedium", p.useMedium);
  F.i("useMultiSource", p.useMultiSource);
  F.i("LayeredModel", p.LayeredModel);
  F.s("LayeredFileName", p.LayeredFileName);
  F.i("useSingleSource(ricker)", p.useSingleSource_ricker);
  F.i("useSingleSource(double_couple)", p.useSingleSource_double_couple);
  F.f("strike", p.strike);
  F.f("dip", p.dip);
  F.f("rake", p.rake);
  F.f("Mw", p.Mw);
  F.f("duration", p.duration);
  F.f("rickerfc", p.rickerfc);
  F.i("ShenModel", p.ShenModel);
  F.i("Crust_1Model", p.Crust_1Model);
  F.i("itSlice", p.itSlice);
  F.i("itStep", p.itStep);
  F.s("waveOutput", p.waveOutput);
  F.s("sliceName", p.sliceName);
  F.i("itStart", p.itStart);
  F.i("itEnd", p.itEnd);
  F.i("igpu", p.igpu);
  F.s("OUT", p.OUT);
  F.s("TerrainDir", p.TerrainDir);
  F.i("SRTM90", p.SRTM90);
  F.i("lonStart", p.lonStart);
  F.i("latStart", p.latStart);
  F.i("blockX", p.blockX);
  F.i("blockY", p.blockY);
  F.f("Depth(km)", p.Depth);
  F.f("MLonStart", p.MLonStart);
  F.f("MLatStart", p.MLatStart);
  F.f("MLonEnd", p.MLonEnd);
  F.f("MLatEnd", p.MLatEnd);
  F.f("MLonStep", p.MLonStep);
  F.f("MLatStep", p.MLatStep);
  F.f("CrustLonStep", p.CrustLonStep);
  F.f("CrustLatStep", p.CrustLatStep);
  F.f("MVeticalStep", p.MVeticalStep);
  F.s("MediumDir", p.MediumDir);
  F.s("CrustDir", p.crustDir);
  F.i("sourceX", p.sourceX);
  F.i("sourceY", p.sourceY);
  F.i("sourceZ", p.sourceZ);
  F.s("sourceFile", p.sourceFile);
  F.s("sourceDir", p.sourceDir);
  F.i("degree2radian", p.degree2radian);
  if (!F.bad.empty()) {
    msg = F.bad;
    return 1001;
  }
  // without terrain the depth is the grid's own (getParams.cpp:228-231)
  if (!p.useTerrain) p.Depth = (float)(p.DH * p.NZ * 1e-3);
  return 0;
}

int read_params(const std::string& path, Params& p, std::string& msg) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    msg = "There is not " + path + " file!";
    return 100;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_params(ss.str(), p, msg);
}

int step_count(const Params& p) {
  const float DT = (float)p.DT;
  return (int)(p.TMAX / DT);
}

}  // namespace cgh
