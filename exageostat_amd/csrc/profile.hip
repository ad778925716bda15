#include "profile.hpp"

#include "common.hpp"

namespace exa {

Profiler& Profiler::get() {
  static Profiler p;
  return p;
}

void Profiler::enable(bool on) { enabled_ = on; }

hipEvent_t Profiler::event(int i) {
  while ((int)pool_.size() <= i) {
    hipEvent_t e;
    EXA_HIP(hipEventCreate(&e));
    pool_.push_back(e);
  }
  return pool_[i];
}

int Profiler::begin(const char* name, double flops, double bytes, hipStream_t s) {
  Pending p;
  p.name = name;
  p.ev_a = next_++;
  p.ev_b = next_++;
  p.flops = flops;
  p.bytes = bytes;
  EXA_HIP(hipEventRecord(event(p.ev_a), s));
  pending_.push_back(p);
  return (int)pending_.size() - 1;
}

void Profiler::end(int handle, hipStream_t s) { EXA_HIP(hipEventRecord(event(pending_[handle].ev_b), s)); }

void Profiler::count(const char* name, double flops, double bytes) {
  if (!enabled_) return;
  KernelTotals& t = totals_[name];
  t.launches += 1;
  t.flops += flops;
  t.bytes += bytes;
}

void Profiler::collect() {
  for (const Pending& p : pending_) {
    EXA_HIP(hipEventSynchronize(event(p.ev_b)));
    float ms = 0.0f;
    EXA_HIP(hipEventElapsedTime(&ms, event(p.ev_a), event(p.ev_b)));
    KernelTotals& t = totals_[p.name];
    t.launches += 1;
    t.ms += ms;
    t.flops += p.flops;
    t.bytes += p.bytes;
  }
  pending_.clear();
  next_ = 0;
}

void Profiler::reset() {
  pending_.clear();
  next_ = 0;
  totals_.clear();
}

KernelTotals Profiler::totals(const std::string& name) {
  auto it = totals_.find(name);
  return it == totals_.end() ? KernelTotals() : it->second;
}

}  // namespace exa
