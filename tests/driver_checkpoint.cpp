// Checkpoint/resume of the lambda-queue driver (run_until_no_criticals.hpp)
// with the synthetic launcher of driver_trace.cpp.
//   driver_checkpoint run  FILE STOP   save after every launch, stop (exit 3) after STOP launches (-1: never)
//   driver_checkpoint resume FILE      continue from FILE
//   driver_checkpoint roundtrip        read "[lb, ub]" lines from stdin, write them back
// Prints one line per launch: <lb> <ub> <first or inf>, then "result <0|1>".
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../circlepacking_amd/csrc/run_until_no_criticals.hpp"

struct Stop {};

int main(int argc, char** argv) {
  using circlepacking::IV;
  std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "roundtrip") {
    IV iv;
    while (std::cin >> iv) std::cout << iv << "\n";
    return 0;
  }
  if (argc < 3) return 2;
  const std::string path = argv[2];
  const long stop = (mode == "run" && argc > 3) ? std::atol(argv[3]) : -1;
  long n = 0;
  auto launch = [&](IV la, bool) {
    if (stop >= 0 && n >= stop) throw Stop{};
    ++n;
    double first = DBL_MAX;
    for (int i = 0; i < 16; ++i) {
      double d = (la.ub - la.lb) / 16;
      double l = la.lb + d * i, u = (i == 15) ? la.ub : la.lb + d * (i + 1);
      if (u - l > 0.002 + 0.02 * (l - 0.5)) { first = l; break; }
    }
    std::printf("%.17g %.17g %.17g\n", la.lb, la.ub, first == DBL_MAX ? INFINITY : first);
    return first;
  };
  circlepacking::QueueCheckpoint cp;
  cp.save = [&](const std::vector<IV>& q) { circlepacking::write_queue_file(path, "stage 0", q); };
  if (mode == "resume") {
    std::string tag;
    if (!circlepacking::read_queue_file(path, tag, cp.resume_from) || tag != "stage 0") return 2;
    cp.resume = true;
  }
  try {
    bool ok = circlepacking::run_kernel_until_no_criticals(launch, IV(0.5, 0.9900000000000001), std::cerr, &cp);
    std::printf("result %d\n", ok ? 1 : 0);
  } catch (const Stop&) {
    std::fflush(stdout);
    return 3;
  }
  return 0;
}
