// The lane-packing prover executable (cpl_prover).
//
// Same driver contract as the reference's prover.cpp:32-36 ->
// lane_packing::find_critical_intervals (lane_packing.cu:32-37) ->
// run_kernel_until_no_criticals per configuration family, with the same stage
// order and stdout/stderr banners; the kernel launches go through the C ABI of
// libcpl_hip.so instead of CUDA.  With no arguments it runs the whole proof on
// the reference grids.  Options (all optional, for bounded runs):
//   --family F         run only family F (0 start .. 5 end_3d)
//   --subdiv a,b,c,d   subdivision counts instead of the reference's
//   --lambda lo,hi     outer lambda interval instead of [lambda_min, lambda_max]
//   --max-launches N   stop a stage after N launches (reports it and exits 3)
//   --checkpoint FILE  after every launch, save the stage and its lambda queue
//   --resume FILE      continue a run from a saved queue (earlier stages are
//                      skipped; the saved stage resumes with its queue)
//   --stage-times      after each stage, print its launch count, wall time and
//                      cells evaluated to stderr (no counterpart in the reference)
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <system_error>
#include <vector>

#include "cpl.h"
#include "run_until_no_criticals.hpp"

namespace {

// Errors of the C ABI become std::system_error, as the reference's
// throw_if_cuda_error does (algcuda/cuda-cpp-wrapper/include/algcuda/error.hpp:84-98).
struct CplCategory : std::error_category {
  const char* name() const noexcept override { return "cpl"; }
  std::string message(int c) const override {
    switch (c) {
      case CPL_ERR_HIP: return "HIP runtime error";
      case CPL_ERR_ARG: return "invalid argument";
      case CPL_ERR_TRAP: return "device consistency check failed (trap)";
      case CPL_ERR_NOFIT: return "circle possibly does not fit (trap)";
      case CPL_ERR_CAPACITY: return "worklist capacity exceeded";
      case CPL_ERR_NO_DEVICE: return "no HIP device";
      default: return "unknown error";
    }
  }
};
const CplCategory& cpl_category() {
  static CplCategory c;
  return c;
}
void throw_if_cpl_error(int rc) {
  if (rc != CPL_OK) throw std::system_error(rc, cpl_category(), cpl_last_error());
}

struct Options {
  int family = -1;
  bool have_subdiv = false;
  int32_t subdiv[4] = {0, 0, 0, 0};
  double lo = 0.5, hi = 0.9900000000000001;  // values.hpp:31-32
  long max_launches = -1;
  std::string checkpoint;
  bool resume = false;
  int resume_family = -1;
  std::vector<circlepacking::IV> resume_queue;
  bool stage_times = false;
};
Options g_opt;

struct StageLimit {};

struct HipLauncher {
  int family;
  long launches = 0;
  uint64_t cells = 0, rows_skipped = 0;
  double operator()(circlepacking::IV la, bool output_criticals) {
    if (g_opt.max_launches >= 0 && launches >= g_opt.max_launches) throw StageLimit{};
    ++launches;
    double first = DBL_MAX;
    cpl_mode_r_info info;
    // every line kept on the device fits (a line of the reference's formats is < 384 bytes)
    std::vector<char> text(output_criticals ? (size_t)CPL_MR_MAX_RECORDS * 384 : 0);
    cpl_iv iv{la.lb, la.ub};
    int rc = cpl_first_problematic_ex(family, iv, g_opt.have_subdiv ? g_opt.subdiv : nullptr,
                                      output_criticals ? CPL_MR_OUTPUT_CRITICALS : 0u, &first, &info,
                                      output_criticals ? text.data() : nullptr, text.size());
    cells += info.cells_evaluated;
    rows_skipped += info.rows_skipped;
    if (output_criticals && !text.empty()) {
      std::fputs(text.data(), stdout);
      std::fflush(stdout);
      if (info.lines_written < info.lines_found)
        std::cerr << "Critical-cell output truncated: " << info.lines_written << " of " << info.lines_found
                  << " lines shown" << std::endl;
    }
    throw_if_cpl_error(rc);
    return first;
  }
};

bool run_stage(int family) {
  HipLauncher launch{family};
  circlepacking::QueueCheckpoint cp;
  const std::string tag = "stage " + std::to_string(family);
  if (!g_opt.checkpoint.empty())
    cp.save = [&](const std::vector<circlepacking::IV>& q) {
      if (!circlepacking::write_queue_file(g_opt.checkpoint, tag, q))
        std::cerr << "cannot write checkpoint " << g_opt.checkpoint << std::endl;
    };
  if (g_opt.resume && family == g_opt.resume_family) {
    cp.resume = true;
    cp.resume_from = g_opt.resume_queue;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const bool done =
      circlepacking::run_kernel_until_no_criticals(launch, circlepacking::IV(g_opt.lo, g_opt.hi), std::cerr, &cp);
  if (g_opt.stage_times) {
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cerr << "Stage " << family << ": " << launch.launches << " launches, " << sec << " s, " << launch.cells
              << " cells evaluated, " << launch.rows_skipped << " rows skipped" << std::endl;
  }
  return done;
}

// Stage selection: --family, and on resume every stage before the saved one is done.
bool want(int f) { return (g_opt.family < 0 || g_opt.family == f) && (!g_opt.resume || f >= g_opt.resume_family); }

}  // namespace

namespace circlepacking {
namespace lane_packing {

// start_configuration.cu:217-220
void find_critical_intervals_start_configuration() {
  if (!want(CPL_START)) return;
  std::cout << "Start configuration..." << std::endl;
  run_stage(CPL_START);
}
// start_configuration.cu:222-225, 300-308
void find_critical_intervals_start_end_configuration() {
  if (want(CPL_START_END_2D)) {
    std::cout << "Start & end configuration, 2 disks... " << std::endl;
    run_stage(CPL_START_END_2D);
  }
  if (want(CPL_START_END_3D)) {
    std::cerr << "Start & end configuration, 3 disks ..." << std::endl;
    run_stage(CPL_START_END_3D);
  }
}
// middle_configuration.cu:207-212
void find_critical_intervals_middle_configuration() {
  if (!want(CPL_MIDDLE)) return;
  std::cout << "Middle configuration ..." << std::endl;
  run_stage(CPL_MIDDLE);
}
// end_configuration.cu:270-286
void find_critical_intervals_end_configuration() {
  if (want(CPL_END_2D)) {
    std::cout << "End configuration (2 disks) ..." << std::endl;
    run_stage(CPL_END_2D);
  }
  if (want(CPL_END_3D)) {
    std::cout << "End configuration (3 disks) ..." << std::endl;
    run_stage(CPL_END_3D);
  }
}
// lane_packing.cu:32-37
void find_critical_intervals() {
  find_critical_intervals_start_configuration();
  find_critical_intervals_start_end_configuration();
  find_critical_intervals_middle_configuration();
  find_critical_intervals_end_configuration();
}

}  // namespace lane_packing
}  // namespace circlepacking

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--family") {
      g_opt.family = std::atoi(next());
    } else if (a == "--subdiv") {
      g_opt.have_subdiv = std::sscanf(next(), "%d,%d,%d,%d", &g_opt.subdiv[0], &g_opt.subdiv[1], &g_opt.subdiv[2],
                                      &g_opt.subdiv[3]) == 4;
      if (!g_opt.have_subdiv) return 2;
    } else if (a == "--lambda") {
      if (std::sscanf(next(), "%lf,%lf", &g_opt.lo, &g_opt.hi) != 2) return 2;
    } else if (a == "--max-launches") {
      g_opt.max_launches = std::atol(next());
    } else if (a == "--stage-times") {
      g_opt.stage_times = true;
    } else if (a == "--checkpoint") {
      g_opt.checkpoint = next();
    } else if (a == "--resume") {
      std::string tag;
      const char* path = next();
      if (!circlepacking::read_queue_file(path, tag, g_opt.resume_queue) ||
          std::sscanf(tag.c_str(), "stage %d", &g_opt.resume_family) != 1) {
        std::fprintf(stderr, "cannot read checkpoint %s\n", path);
        return 2;
      }
      g_opt.resume = true;
    } else {
      std::fprintf(stderr,
                   "usage: %s [--family F] [--subdiv a,b,c,d] [--lambda lo,hi] [--max-launches N] "
                   "[--checkpoint FILE] [--resume FILE] [--stage-times]\n",
                   argv[0]);
      return 2;
    }
  }
  throw_if_cpl_error(cpl_init());
  try {
    circlepacking::lane_packing::find_critical_intervals();
  } catch (const StageLimit&) {
    std::cerr << "Launch limit reached" << std::endl;
    cpl_shutdown();
    return 3;
  } catch (const std::system_error& e) {
    // the reference's kernels trap here and its next synchronising copy throws
    std::cerr << "error: " << e.code().message() << ": " << e.what() << std::endl;
    cpl_shutdown();
    return 1;
  }
  cpl_shutdown();
  return 0;  // prover.cpp:35: the result of each stage is not inspected
}
