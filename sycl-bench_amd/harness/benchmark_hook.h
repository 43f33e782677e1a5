// Lifecycle hooks around setup and the timed kernel section, as in the
// reference's include/benchmark_hook.h:6-16; the attach point for profiler or
// power-measurement hooks.
#pragma once

#include "result_consumer.h"

class BenchmarkHook {
public:
  virtual ~BenchmarkHook() = default;
  virtual void atInit() = 0;
  virtual void preSetup() = 0;
  virtual void postSetup() = 0;
  virtual void preKernel() = 0;
  virtual void postKernel() = 0;
  virtual void emitResults(ResultConsumer&) {}
};
