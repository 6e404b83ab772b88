// Host lambda-refinement driver of the lane-packing proof.
//
// Same contract as the reference's run_kernel_until_no_criticals
// (lane_packing/run_kernel.cuh:61-132): a min-heap of lambda intervals ordered
// by lower bound; each interval is handed to one kernel launch that returns the
// smallest la.lb of a failing cell (DBL_MAX if none); the certified prefix
// [current.lb, first) is dropped and the rest is re-queued whole or split in
// half; the first SIGINT makes the driver re-run the first remaining interval
// with critical-cell output and return false; a second SIGINT terminates.
//
// The launcher is a template parameter so the queue logic is shared by the
// MI355X prover (HIP launches through libcpl_hip.so) and by a trace test that
// drives it with a synthetic failing-lambda function.
//
// Checkpoint/resume (no counterpart in the reference): the heap is a vector
// kept with std::push_heap / std::pop_heap, exactly what std::priority_queue
// does, so its contents can be saved after every launch in the reference's
// interval text format (19 significant digits: every double round-trips) and
// a run can resume from the saved queue with the same processing order.
#pragma once

#include <algorithm>
#include <cfloat>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace circlepacking {

struct IV {
  double lb, ub;
  IV() = default;
  IV(double l, double u) : lb(l), ub(u) {}
  double get_lb() const { return lb; }
  double get_ub() const { return ub; }
};

// Text form of algcuda::Interval (interval.hpp:294-303): right-aligned,
// precision 19, width 26.
inline std::ostream& operator<<(std::ostream& out, const IV& iv) {
  std::ios_base::fmtflags f = out.flags(std::ios::right);
  std::streamsize p = out.precision(19);
  out << '[' << std::setw(26) << iv.lb << ", " << std::setw(26) << iv.ub << ']';
  out.flags(f);
  out.precision(p);
  return out;
}

// Interval text input of algcuda::Interval (interval.hpp:305-349): "[lb, ub]"
// (',' or ';' between the bounds) or a single number for a point interval.
template <class CharType>
std::basic_istream<CharType>& operator>>(std::basic_istream<CharType>& in, IV& iv) {
  CharType c;
  if (!(in >> c)) return in;
  if (c == '[') {
    double a, b;
    CharType comma, end;
    if (!(in >> a) || !(in >> comma)) return in;
    if (comma != ',' && comma != ';') {
      in.setstate(std::ios_base::failbit);
      return in;
    }
    if (!(in >> b) || !(in >> end)) return in;
    if (end != ']') {
      in.setstate(std::ios_base::failbit);
      return in;
    }
    iv = IV(a, b);
  } else {
    if (!in.putback(c)) return in;
    double v;
    if (!(in >> v)) return in;
    iv = IV(v, v);
  }
  return in;
}

// Saved state of one driver run: the queue in heap order.
struct QueueCheckpoint {
  std::function<void(const std::vector<IV>&)> save;  // called after every launch (may be empty)
  std::vector<IV> resume_from;                       // non-empty: start from this queue instead of outer_la
  bool resume = false;
};

// Queue file: a header line, the stage tag, the count, one interval per line.
inline bool write_queue_file(const std::string& path, const std::string& tag, const std::vector<IV>& q) {
  const std::string tmp = path + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) return false;
    out << "cpl-queue 1\n" << tag << "\n" << q.size() << "\n";
    for (const IV& iv : q) out << iv << "\n";
    out.flush();
    if (!out) return false;
  }
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}
inline bool read_queue_file(const std::string& path, std::string& tag, std::vector<IV>& q) {
  std::ifstream in(path);
  std::string magic;
  size_t n = 0;
  if (!std::getline(in, magic) || magic != "cpl-queue 1" || !std::getline(in, tag) || !(in >> n)) return false;
  q.clear();
  for (size_t k = 0; k < n; ++k) {
    IV iv;
    if (!(in >> iv)) return false;
    q.push_back(iv);
  }
  return true;
}

// SIGINT protocol state (run_kernel.cuh:38-59).  Global and not reentrant, as
// in the reference.
inline volatile std::sig_atomic_t& was_interrupted() {
  static volatile std::sig_atomic_t flag = 0;
  return flag;
}
inline void (*&stored_handler())(int) {
  static void (*h)(int) = SIG_DFL;
  return h;
}
inline void interrupt_handler(int) {
  if (was_interrupted()) {
    std::signal(SIGINT, stored_handler());
    std::raise(SIGINT);
    std::abort();
  }
  was_interrupted() = 1;
}
inline void start_interruptible_computation() {
  was_interrupted() = 0;
  stored_handler() = std::signal(SIGINT, &interrupt_handler);
}
inline void stop_interruptible_computation() {
  std::signal(SIGINT, stored_handler());
  stored_handler() = SIG_DFL;
}

// Launcher: double launch(IV lambda, bool output_criticals) returning the
// first problematic la.lb or DBL_MAX.  Returns false if interrupted.
template <class Launcher>
bool run_kernel_until_no_criticals(Launcher&& launch, IV outer_la, std::ostream& log = std::cerr,
                                   QueueCheckpoint* cp = nullptr) {
  struct Comp_lower {
    bool operator()(const IV& a, const IV& b) const noexcept { return a.lb > b.lb; }
  };
  // the std::priority_queue<IV, std::vector<IV>, Comp_lower> of the reference, as a visible heap
  struct Heap {
    std::vector<IV> c;
    bool empty() const { return c.empty(); }
    const IV& top() const { return c.front(); }
    void push(const IV& v) {
      c.push_back(v);
      std::push_heap(c.begin(), c.end(), Comp_lower());
    }
    void pop() {
      std::pop_heap(c.begin(), c.end(), Comp_lower());
      c.pop_back();
    }
  } q;
  if (cp && cp->resume) q.c = cp->resume_from;  // saved in heap order
  else q.push(outer_la);

  start_interruptible_computation();
  while (!q.empty()) {
    IV current = q.top();
    q.pop();
    log << "Handling lambda interval " << current << std::endl;
    double first = launch(current, false);

    if (was_interrupted()) {
      was_interrupted() = 0;
      IV last_run;
      if (first != DBL_MAX) {
        last_run = current;
      } else {
        if (q.empty()) {
          stop_interruptible_computation();
          return true;
        }
        last_run = q.top();
      }
      log << "We were interrupted - running the first remaining interval with output once!" << std::endl;
      launch(last_run, true);
      stop_interruptible_computation();
      return false;
    }

    if (first != DBL_MAX) {
      double width_before = current.ub - current.lb;
      double width_after = current.ub - first;
      if (width_before >= 2 * width_after) {
        q.push(IV(first, current.ub));
      } else {
        double mid = 0.5 * (first + current.ub);
        if (first < mid) q.push(IV(first, mid));
        if (mid <= current.ub) q.push(IV(mid, current.ub));
      }
    }
    if (cp && cp->save) cp->save(q.c);
  }
  stop_interruptible_computation();
  return true;
}

}  // namespace circlepacking
