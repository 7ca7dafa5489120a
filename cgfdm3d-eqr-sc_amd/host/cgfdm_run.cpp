// cgfdm_run: the reference's executable flow on one GPU (main.cpp:17-45,
// run.cpp:28-138, propagate.cpp:76-350): params.json and station.json from the
// working directory, coordinates and medium, source, time loop on the HIP
// library through its C ABI, and the reference's output files. Setup and
// output files go through cgfdm_create_from_params / cgfdm_write_outputs.
//
// usage: cgfdm_run [--scheme cgfdm|aweno] [--precision 4|8] [--coef 0|1|2]
//                  [--params FILE] [--stations FILE] [--steps N] [--quiet]
//
// --scheme picks the build the reference makes with Makefile_CGFDM (default) or
// Makefile_SCFDM. --steps replaces NT = TMAX/DT (for short checks). Errors print
// a message and end the process with the reference's MPI_Abort code (100 no
// params.json, 1001 parse error, 110 CFL, 130 PML, 180 source), as an exit
// status modulo 256 and in the message.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "cgfdm_amd.h"

namespace {

int die(int code, const std::string& msg) {
  std::fprintf(stderr, "cgfdm_run: %s (code %d)\n", msg.c_str(), code);
  return code & 0xFF ? code & 0xFF : 1;
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
  cgfdm_solver* s = nullptr;
  int rc = cgfdm_create_from_params(params_path.c_str(), stations_path.c_str(), scheme, precision, coef,
                                    steps_override, &s);
  if (rc) return die(rc, cgfdm_last_error());
  int NT = 0, it_skip = 1;
  if ((rc = cgfdm_run_info(s, &NT, &it_skip))) return die(rc, cgfdm_last_error());
  if (!quiet) std::printf("NT %d, IT_SKIP %d\n", NT, it_skip);

  // time loop (propagate.cpp:186-321): output after step `it` when it % IT_SKIP == 0
  auto t0 = std::chrono::steady_clock::now();
  int it = 0;
  while (it < NT) {
    const int o = ((it + it_skip - 1) / it_skip) * it_skip;  // next output step
    const int last = o < NT ? o : NT - 1;
    if ((rc = cgfdm_step(s, last - it + 1))) return die(rc, cgfdm_last_error());
    it = last + 1;
    if (last != o) break;
    if ((rc = cgfdm_write_outputs(s, CGFDM_OUT_SLICES))) return die(rc, cgfdm_last_error());
    if (!quiet) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("it = %8d. Total time: %8.3lfs.\n", o, el);
      std::fflush(stdout);
    }
  }
  if ((rc = cgfdm_synchronize(s))) return die(rc, cgfdm_last_error());
  double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // stationWrite and outputPGV (station.cpp:422-443, PGV.cpp:177-185)
  if ((rc = cgfdm_write_outputs(s, CGFDM_OUT_STATIONS | CGFDM_OUT_PGV))) return die(rc, cgfdm_last_error());
  cgfdm_destroy(s);
  if (!quiet) std::printf("done: %d steps in %.3f s\n", NT, el);
  return 0;
}
