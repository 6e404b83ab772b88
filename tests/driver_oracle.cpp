// The lambda-queue driver (run_until_no_criticals.hpp) with the CPU oracle's
// reference-grid search as launcher: the CPU side of the lambda-trace parity
// test against cpl_prover.  Usage: driver_oracle F a,b,c,d lo,hi max_launches
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "../circlepacking_amd/csrc/run_until_no_criticals.hpp"

extern "C" int cplo_first_problematic(int family, double outer_lb, double outer_ub, const int* subdiv,
                                      int output_criticals, double* first_out, unsigned long long* evaluated_out,
                                      int* trap_code_out, double* recs, unsigned long long rec_cap,
                                      unsigned long long* n_recs_out, unsigned long long* nofit_flags_out,
                                      int nthreads);

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  int family = std::atoi(argv[1]);
  int sd[4];
  double lo, hi;
  if (std::sscanf(argv[2], "%d,%d,%d,%d", &sd[0], &sd[1], &sd[2], &sd[3]) != 4) return 2;
  if (std::sscanf(argv[3], "%lf,%lf", &lo, &hi) != 2) return 2;
  long max_launches = std::atol(argv[4]), n = 0;
  struct Limit {};
  auto launch = [&](circlepacking::IV la, bool) {
    if (n++ >= max_launches) throw Limit{};
    double first = DBL_MAX;
    unsigned long long ev = 0;
    int trap = 0;
    int rc = cplo_first_problematic(family, la.lb, la.ub, sd, 0, &first, &ev, &trap, nullptr, 0, nullptr, nullptr, 0);
    if (rc != 0) {
      std::cerr << "error " << rc << std::endl;
      std::exit(4);
    }
    return first;
  };
  try {
    circlepacking::run_kernel_until_no_criticals(launch, circlepacking::IV(lo, hi));
  } catch (const Limit&) {
    std::cerr << "Launch limit reached" << std::endl;
    return 3;
  }
  return 0;
}
