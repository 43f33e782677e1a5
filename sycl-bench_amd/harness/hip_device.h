// The HIP device behind every benchmark executable: one stream and one mfx
// handle on one gfx950 GPU.  --device=gpu / default select it; --device=cpu
// is rejected (the suite's CPU baseline is the OpenMP restatement under
// oracle/, which is test infrastructure and not part of these programs).
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "benchmark_traits.h"
#include "command_line.h"
#include "device.h"
#include "mfx.h"

inline void hipCheck(hipError_t e, const char* what) {
  if(e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void mfxCheck(mfx_status s, const char* what) {
  if(s != MFX_SUCCESS)
    throw std::runtime_error(std::string(what) + ": " + mfx_status_string(s) + ": " + mfx_last_error());
}

class HipDevice : public Device {
public:
  explicit HipDevice(int id) : id_(id) {
    hipCheck(hipSetDevice(id_), "hipSetDevice");
    hipDeviceProp_t prop;
    hipCheck(hipGetDeviceProperties(&prop, id_), "hipGetDeviceProperties");
    name_ = prop.name;
    arch_ = prop.gcnArchName;
    // The marketing name can be empty when the driver's id table is missing.
    if(name_.empty()) name_ = "AMD GPU " + arch_.substr(0, arch_.find(':'));
    hipCheck(hipStreamCreate(&stream_), "hipStreamCreate");
    mfxCheck(mfx_create(&handle_, id_), "mfx_create");
  }
  ~HipDevice() override {
    // destructors cannot report errors; teardown failures are ignored
    if(handle_) (void)mfx_destroy(handle_);
    if(stream_) (void)hipStreamDestroy(stream_);
  }
  HipDevice(const HipDevice&) = delete;
  HipDevice& operator=(const HipDevice&) = delete;

  std::string name() const override { return name_; }
  std::string implementation() const override { return "mfx HIP " + arch_.substr(0, arch_.find(':')); }
  std::string kind() const override { return "gpu"; }
  void synchronize() override { hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

  hipStream_t stream() const { return stream_; }
  mfx_handle handle() const { return handle_; }
  int id() const { return id_; }

private:
  int id_;
  std::string name_, arch_;
  hipStream_t stream_ = nullptr;
  mfx_handle handle_ = nullptr;
};

inline std::shared_ptr<Device> makeDevice(const std::string& type) {
  if(type == "gpu" || type == "default") {
    const char* env = std::getenv("MFX_DEVICE");
    return std::make_shared<HipDevice>(env ? std::atoi(env) : 0);
  }
  if(type == "cpu")
    throw std::invalid_argument(
        "--device=cpu is not provided by these programs: the CPU baseline is the OpenMP restatement of the reference "
        "kernels under oracle/ (test infrastructure)");
  throw std::invalid_argument("unknown device type: " + type);
}

inline HipDevice& hipDevice(const BenchmarkArgs& args) {
  auto* d = dynamic_cast<HipDevice*>(args.device.get());
  if(!d) throw std::runtime_error("benchmark needs a HIP device");
  return *d;
}

// Runs `submit` (which enqueues one or more kernels on the device stream) and
// records a hipEvent pair around it as one KernelTiming entry.
template <class F>
void timedSubmit(HipDevice& d, std::vector<KernelTiming>& out, F&& submit) {
  struct Events {
    hipEvent_t start = nullptr, stop = nullptr;
    Events() {
      hipCheck(hipEventCreate(&start), "hipEventCreate");
      hipCheck(hipEventCreate(&stop), "hipEventCreate");
    }
    ~Events() {
      (void)hipEventDestroy(start);
      (void)hipEventDestroy(stop);
    }
  };
  auto ev = std::make_shared<Events>();
  const auto t0 = std::chrono::high_resolution_clock::now();
  hipCheck(hipEventRecord(ev->start, d.stream()), "hipEventRecord");
  submit();
  hipCheck(hipEventRecord(ev->stop, d.stream()), "hipEventRecord");
  const auto t1 = std::chrono::high_resolution_clock::now();
  KernelTiming k;
  k.submit = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0);
  k.duration = [ev]() {
    float ms = 0.0f;
    hipCheck(hipEventElapsedTime(&ms, ev->start, ev->stop), "hipEventElapsedTime");
    return std::chrono::nanoseconds(static_cast<long long>(ms * 1.0e6));
  };
  out.push_back(std::move(k));
}
