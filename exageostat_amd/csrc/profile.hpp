// Optional per-kernel timing with HIP events recorded on the launching
// stream, used by the benchmark to price the dominant kernel against its
// roofline.  Disabled by default (no events are recorded then).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace exa {

struct KernelTotals {
  long launches = 0;
  double ms = 0.0;
  double flops = 0.0;  // algorithmic flops
  double bytes = 0.0;  // algorithmic HBM bytes
};

class Profiler {
 public:
  static Profiler& get();
  bool enabled() const { return enabled_; }
  void enable(bool on);
  // Returns a handle for end(); records the start event on s.
  int begin(const char* name, double flops, double bytes, hipStream_t s);
  void end(int handle, hipStream_t s);
  // Adds flops and bytes to a group without timing (a share of work done
  // inside another group's launch).
  void count(const char* name, double flops, double bytes);
  // Resolves pending events (requires them to have completed).
  void collect();
  void reset();
  KernelTotals totals(const std::string& name);

 private:
  struct Pending {
    std::string name;
    int ev_a, ev_b;
    double flops, bytes;
  };
  hipEvent_t event(int i);
  bool enabled_ = false;
  std::vector<hipEvent_t> pool_;
  int next_ = 0;
  std::vector<Pending> pending_;
  std::map<std::string, KernelTotals> totals_;
};

// Runs `launch` between two timing events when profiling is on.
template <class F>
inline void profiled(const char* name, double flops, double bytes, hipStream_t s, F&& launch) {
  Profiler& p = Profiler::get();
  if (!p.enabled()) {
    launch();
    return;
  }
  const int h = p.begin(name, flops, bytes, s);
  launch();
  p.end(h, s);
}

}  // namespace exa
