// CPU-only program for the harness tests: fake device and benchmarks that
// exercise BenchmarkApp / BenchmarkManager / TimeMetricsProcessor /
// ResultConsumer without a GPU.  The scenario comes from HARNESS_SCENARIO.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

class FakeDevice : public Device {
public:
  std::string name() const override { return "Fake Device"; }
  std::string implementation() const override { return "fake"; }
  std::string kind() const override { return "cpu"; }
  void synchronize() override {}
};

std::shared_ptr<Device> makeDevice(const std::string& type) {
  if(type == "default" || type == "cpu") return std::make_shared<FakeDevice>();
  throw std::invalid_argument("unknown device type: " + type);
}

static int g_setups = 0, g_runs = 0, g_verifies = 0;

// Without KernelTiming run(): kernel/submit/system times are N/A.
class PlainBench {
public:
  explicit PlainBench(const BenchmarkArgs& a) : args_(a) {}
  void setup() { ++g_setups; }
  void run() {
    ++g_runs;
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  bool verify(VerificationSetting&) {
    ++g_verifies;
    return args_.cli.getOrDefault<std::string>("--fail-at", "") != std::to_string(g_verifies);
  }
  static std::string getBenchmarkName(const BenchmarkArgs&) { return "Plain"; }

private:
  BenchmarkArgs args_;
};

// With KernelTiming run() and a throughput metric.
class TimedBench {
public:
  explicit TimedBench(const BenchmarkArgs&) {}
  void setup() {}
  void run(std::vector<KernelTiming>& k) {
    KernelTiming t;
    t.submit = std::chrono::nanoseconds(1000);
    t.duration = [] { return std::chrono::nanoseconds(100000); };
    k.push_back(t);
  }
  static ThroughputMetric getThroughputMetric(const BenchmarkArgs&) { return {1e9, "FLOP"}; }
  static std::string getBenchmarkName(const BenchmarkArgs&) { return "Timed"; }
};

// No verify(): Verification is N/A.
class NoVerifyBench {
public:
  explicit NoVerifyBench(const BenchmarkArgs&) {}
  void setup() {}
  void run() {}
  static std::string getBenchmarkName(const BenchmarkArgs&) { return "NoVerify"; }
};

class ThrowingBench {
public:
  explicit ThrowingBench(const BenchmarkArgs&) {}
  void setup() { throw std::runtime_error("setup exploded"); }
  void run() {}
  static std::string getBenchmarkName(const BenchmarkArgs&) { return "Throwing"; }
};

static void stats_scenario() {
  // Known samples through TimeMetricsProcessor: 1, 2, 3, 4, 10 ms.
  BenchmarkArgs a;
  a.warmup_run = std::getenv("HARNESS_WARMUP") != nullptr;
  TimeMetricsProcessor<TimedBench> t(a);
  for(int ms : {4, 2, 10, 1, 3}) t.addTimingResult("run-time", std::chrono::milliseconds(ms));
  t.markAsUnavailable("kernel-time");
  OstreamResultConsumer out(std::cout);
  t.emitResults(out);
}

int main(int argc, char** argv) {
  const std::string s = std::getenv("HARNESS_SCENARIO") ? std::getenv("HARNESS_SCENARIO") : "";
  if(s == "stats") {
    stats_scenario();
    return 0;
  }
  BenchmarkApp app(argc, argv);
  if(s == "plain") app.run<PlainBench>();
  else if(s == "timed") app.run<TimedBench>();
  else if(s == "noverify") app.run<NoVerifyBench>();
  else if(s == "throwing") {
    app.run<ThrowingBench>();
    app.run<PlainBench>();
  } else if(s == "duplicate") {
    app.run<PlainBench>();
    app.run<PlainBench>();
  } else if(s == "all") {
    app.run<PlainBench>();
    app.run<TimedBench>();
    app.run<NoVerifyBench>();
  }
  std::cerr << "setups=" << g_setups << " runs=" << g_runs << " verifies=" << g_verifies << std::endl;
  return app.exitCode();
}
