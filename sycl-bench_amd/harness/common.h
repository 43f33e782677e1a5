// BenchmarkManager and BenchmarkApp: the run loop of the suite, following the
// reference's include/common.h:29-214.
//
// Per benchmark: the name (a temporary object's getBenchmarkName), metadata
// columns, then num_runs iterations each on a fresh object -- setup() and a
// device synchronisation untimed, run() plus synchronisation timed with the
// host clock ("run-time"), per-kernel device times when the benchmark takes a
// KernelTiming vector and queue profiling is on ("kernel-time",
// "submit-time", "system-time" = run-time - kernel-time; N/A otherwise), and
// verify() after every run while all runs pass.  A throwing run discards the
// benchmark's results.  Verification is reported as N/A / FAIL / PASS.
//
// Differences: queue profiling is switched on by --queue-profiling (or by
// building with SYCL_BENCH_ENABLE_QUEUE_PROFILING, the reference's switch);
// BenchmarkApp::exitCode() is non-zero after a command-line error or a
// failed / throwing benchmark, so scripts can detect it (the reference's
// programs always return 0).
#pragma once

#include <chrono>
#include <exception>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "benchmark_hook.h"
#include "benchmark_traits.h"
#include "command_line.h"
#include "result_consumer.h"
#include "time_metrics.h"
#include "type_traits.h"

inline bool queueProfilingEnabled(const BenchmarkArgs& args) {
#if defined(SYCL_BENCH_ENABLE_QUEUE_PROFILING)
  return !args.cli.isFlagSet("--no-queue-profiling");
#else
  return args.cli.isFlagSet("--queue-profiling");
#endif
}

template <class Benchmark>
class BenchmarkManager {
public:
  explicit BenchmarkManager(const BenchmarkArgs& args) : args_(args) {}

  void addHook(BenchmarkHook& h) { hooks_.push_back(&h); }

  // Returns false when verification failed.
  template <class... Extra>
  bool run(Extra&&... extra) {
    using Traits = detail::BenchmarkTraits<Benchmark>;
    ResultConsumer& out = *args_.result_consumer;
    out.proceedToBenchmark(Benchmark{args_, extra...}.getBenchmarkName(args_));
    out.consumeResult("problem-size", std::to_string(args_.problem_size));
    out.consumeResult("local-size", std::to_string(args_.local_size));
    out.consumeResult("device-name", args_.device->name());
    out.consumeResult("sycl-implementation", args_.device->implementation());

    TimeMetricsProcessor<Benchmark> times(args_);
    std::map<std::string, std::string> extra_results;
    const bool profiling = Traits::supportsQueueProfiling && queueProfilingEnabled(args_);
    const bool extended = args_.cli.getOrDefault<std::string>("--metrics", "reference") == "extended";
    for(auto* h : hooks_) h->atInit();

    bool all_pass = true;
    try {
      for(std::size_t r = 0; r < args_.num_runs && all_pass; ++r) {
        Benchmark b(args_, extra...);
        for(auto* h : hooks_) h->preSetup();
        b.setup();
        args_.device->synchronize();
        for(auto* h : hooks_) h->postSetup();

        std::vector<KernelTiming> kernels;
        kernels.reserve(1024);
        for(auto* h : hooks_) h->preKernel();
        const auto t0 = std::chrono::high_resolution_clock::now();
        if constexpr(Traits::supportsQueueProfiling)
          b.run(kernels);
        else
          b.run();
        args_.device->synchronize();
        const auto t1 = std::chrono::high_resolution_clock::now();
        for(auto* h : hooks_) h->postKernel();

        const auto run_time = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0);
        times.addTimingResult("run-time", run_time);
        if(profiling) {
          std::chrono::nanoseconds kernel{0}, submit{0};
          for(auto& k : kernels) {
            if(k.duration) kernel += k.duration();
            submit += k.submit;
          }
          times.addTimingResult("kernel-time", kernel);
          times.addTimingResult("submit-time", submit);
          times.addTimingResult("system-time", run_time - kernel);
        } else {
          times.markAsUnavailable("kernel-time");
          times.markAsUnavailable("submit-time");
          times.markAsUnavailable("system-time");
        }

        if constexpr(Traits::hasVerify) {
          if(args_.verification.size() > 0 && args_.verification.enabled) {
            VerificationSetting v = args_.verification;
            if(!b.verify(v)) all_pass = false;
          }
        }
        if constexpr(Traits::hasExtraResults) {
          if(extended) b.extraResults(extra_results);
        }
      }
    } catch(...) {
      out.discard();
      throw;
    }

    times.emitResults(out);
    for(const auto& kv : extra_results) out.consumeResult(kv.first, kv.second);
    for(auto* h : hooks_) h->emitResults(out);

    if(!Traits::hasVerify || args_.verification.size() == 0 || !args_.verification.enabled)
      out.consumeResult("Verification", "N/A");
    else
      out.consumeResult("Verification", all_pass ? "PASS" : "FAIL");
    out.flush();
    return all_pass;
  }

private:
  BenchmarkArgs args_;
  std::vector<BenchmarkHook*> hooks_;
};

class BenchmarkApp {
public:
  BenchmarkApp(int argc, char** argv) {
    try {
      args_ = BenchmarkCommandLine(argc, argv).getBenchmarkArgs();
      ok_ = true;
    } catch(const std::exception& e) {
      std::cerr << "Error while parsing command lines: " << e.what() << std::endl;
      failures_ = 1;
    }
  }

  const BenchmarkArgs& getArgs() const { return args_; }
  bool shouldRunNDRangeKernels() const { return !args_.cli.isFlagSet("--no-ndrange-kernels"); }

  template <class Benchmark, class... Extra>
  void run(Extra&&... extra) {
    if(!ok_) return;
    try {
      const std::string name = Benchmark{args_, extra...}.getBenchmarkName(args_);
      if(!names_.insert(name).second) {
        std::cerr << "Benchmark with name '" << name << "' has already been run\n";
        throw std::runtime_error("Duplicate benchmark name");
      }
      BenchmarkManager<Benchmark> mgr(args_);
      if(!mgr.run(extra...)) ++failures_;
    } catch(const std::exception& e) {
      std::cerr << "Error: " << e.what() << std::endl;
      ++failures_;
    }
  }

  // 0 when every benchmark ran and verified (or verification was off).
  int exitCode() const { return failures_ ? 1 : 0; }

private:
  BenchmarkArgs args_;
  bool ok_ = false;
  int failures_ = 0;
  std::unordered_set<std::string> names_;
};
