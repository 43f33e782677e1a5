// Timing statistics, following the reference's include/time_metrics.h:26-153:
// per timing name it emits mean, sample standard deviation (n-1), median
// (element n/2 of the sorted samples), min, the quoted sample list and
// throughput = metric / min; timings that are unavailable still emit every
// column as N/A so all rows share one CSV header.  Values are seconds printed
// with std::to_string (six decimals), as the reference does.
//
// Difference: with --warmup-run the first sample of each timing is dropped
// whenever that timing has more than one sample (the reference tests the
// number of timing *names*, :80-82, which makes the option inert).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "benchmark_traits.h"
#include "command_line.h"
#include "result_consumer.h"

// Returned by a benchmark's optional static getThroughputMetric(): the amount
// of work (e.g. FLOP), not a rate.
struct ThroughputMetric {
  double metric = 0.0;
  std::string unit;
};

template <class Benchmark>
class TimeMetricsProcessor {
public:
  explicit TimeMetricsProcessor(const BenchmarkArgs& args) : args_(args) {}

  void addTimingResult(const std::string& name, std::chrono::nanoseconds t) {
    if(unavailable_.count(name)) throw std::invalid_argument("Cannot add result for unavailable timing " + name);
    samples_[name].push_back(t);
  }

  void markAsUnavailable(const std::string& name) {
    if(samples_.count(name)) throw std::invalid_argument("Cannot mark timing " + name + " with results as unavailable");
    unavailable_.insert(name);
  }

  void emitResults(ResultConsumer& out) const {
    ThroughputMetric tpm;
    bool have_tpm = false;
    if constexpr(detail::BenchmarkTraits<Benchmark>::hasGetThroughputMetric) {
      tpm = Benchmark::getThroughputMetric(args_);
      have_tpm = true;
      out.consumeResult("throughput-metric", std::to_string(tpm.metric), tpm.unit);
    } else {
      out.consumeResult("throughput-metric", "N/A", "");
    }

    std::set<std::string> names(unavailable_.begin(), unavailable_.end());
    for(const auto& kv : samples_) names.insert(kv.first);

    for(const auto& name : names) {
      if(unavailable_.count(name)) {
        for(const char* suffix : {"-mean", "-stddev", "-median", "-min", "-samples", "-throughput"})
          out.consumeResult(name + suffix, "N/A");
        continue;
      }
      const auto& raw = samples_.at(name);
      const std::size_t skip = (args_.warmup_run && raw.size() > 1) ? 1 : 0;
      std::vector<double> s;
      for(std::size_t i = skip; i < raw.size(); ++i) s.push_back(raw[i].count() / 1.0e9);
      std::sort(s.begin(), s.end());
      const double n = static_cast<double>(s.size());
      const double mean = std::accumulate(s.begin(), s.end(), 0.0) / n;
      double var = 0.0;
      for(double x : s) var += (x - mean) * (x - mean);
      const double stddev = s.size() > 1 ? std::sqrt(var / (n - 1.0)) : 0.0;
      out.consumeResult(name + "-mean", std::to_string(mean), "s");
      out.consumeResult(name + "-stddev", std::to_string(stddev), "s");
      out.consumeResult(name + "-median", std::to_string(s[s.size() / 2]), "s");
      out.consumeResult(name + "-min", std::to_string(s.front()), "s");
      std::ostringstream list;
      list << '"';
      for(std::size_t i = 0; i < s.size(); ++i) list << (i ? " " : "") << std::to_string(s[i]);
      list << '"';
      out.consumeResult(name + "-samples", list.str());
      if(have_tpm && tpm.metric > 0.0)
        out.consumeResult(name + "-throughput", std::to_string(tpm.metric / s.front()), tpm.unit + "/s");
      else
        out.consumeResult(name + "-throughput", "N/A", "");
    }
  }

private:
  BenchmarkArgs args_;
  std::map<std::string, std::vector<std::chrono::nanoseconds>> samples_;
  std::set<std::string> unavailable_;
};
