// Phase wall-clock timer with the reference's keys and CSV semantics
// (src/time_durations.cuh:7-39): Tic/Toc per named phase;
// TotalExecutionTime() sums every recorded phase (so a merge_phase nested in
// sort_phase is counted twice, as in the reference; TotalOuterTime() does
// not double count).
#pragma once

#include <chrono>
#include <map>
#include <mutex>
#include <string>

namespace mgsort {

class TimeDurations {
 public:
  static TimeDurations* Get() {
    static TimeDurations instance;
    return &instance;
  }
  TimeDurations(const TimeDurations&) = delete;
  void operator=(const TimeDurations&) = delete;

  void Tic(const std::string& key) {
    std::lock_guard<std::mutex> g(mu_);
    begin_[key] = std::chrono::high_resolution_clock::now();
  }
  void Toc(const std::string& key) {
    const auto now = std::chrono::high_resolution_clock::now();
    std::lock_guard<std::mutex> g(mu_);
    durations[key] = now - begin_.at(key);
  }
  double Seconds(const std::string& key) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = durations.find(key);
    return it == durations.end() ? 0.0 : it->second.count();
  }
  std::chrono::duration<double> TotalExecutionTime() {
    std::lock_guard<std::mutex> g(mu_);
    std::chrono::duration<double> t(0);
    for (const auto& d : durations) t += d.second;
    return t;
  }
  double TotalOuterTime() {
    return Seconds("memory_allocation") + Seconds("sort_phase") + Seconds("memory_deallocation");
  }
  void Reset() {
    std::lock_guard<std::mutex> g(mu_);
    durations.clear();
    begin_.clear();
  }

  std::map<std::string, std::chrono::duration<double>> durations;

 private:
  TimeDurations() = default;
  std::mutex mu_;
  std::map<std::string, std::chrono::time_point<std::chrono::high_resolution_clock>> begin_;
};

}  // namespace mgsort
