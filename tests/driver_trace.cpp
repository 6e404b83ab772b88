// Trace test of the lambda-queue driver (run_until_no_criticals.hpp) with a
// synthetic launcher: lambda cells fail while their width exceeds w(lb), a
// function of the cell's position.  Prints one line per launch:
//   <lb> <ub> <first-problematic or inf> <output_criticals>
// and finally "result <0|1>".  argv[1] = launch index at which to raise
// SIGINT (-1: never).
#include <cfloat>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>

#include "../circlepacking_amd/csrc/run_until_no_criticals.hpp"

int main(int argc, char** argv) {
  long interrupt_at = argc > 1 ? std::atol(argv[1]) : -1;
  long n = 0;
  auto launch = [&](circlepacking::IV la, bool out) {
    if (n == interrupt_at) std::raise(SIGINT);
    ++n;
    // 16 sub-cells per launch as a kernel would evaluate; a sub-cell [l,u]
    // fails if u - l > 0.002 + 0.02 * (l - 0.5) (narrower cells needed near 0.5).
    double first = DBL_MAX;
    for (int i = 0; i < 16; ++i) {
      double d = (la.ub - la.lb) / 16;
      double l = la.lb + d * i, u = (i == 15) ? la.ub : la.lb + d * (i + 1);
      if (u - l > 0.002 + 0.02 * (l - 0.5)) { first = l; break; }
    }
    std::printf("%.17g %.17g %.17g %d\n", la.lb, la.ub, first == DBL_MAX ? INFINITY : first, out ? 1 : 0);
    return first;
  };
  bool ok = circlepacking::run_kernel_until_no_criticals(launch, circlepacking::IV(0.5, 0.9900000000000001));
  std::printf("result %d\n", ok ? 1 : 0);
  return 0;
}
