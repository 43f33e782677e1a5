// The execution device behind BenchmarkArgs.  Replaces the sycl::queue pair of
// the reference's BenchmarkArgs (include/command_line.h:129-140): what the
// harness needs from it is a name, an implementation label, and a way to wait
// for all submitted work (the reference's queue.wait_and_throw()).
#pragma once

#include <memory>
#include <string>

class Device {
public:
  virtual ~Device() = default;
  // "device-name" result column.
  virtual std::string name() const = 0;
  // "sycl-implementation" result column (kept under the reference's column name).
  virtual std::string implementation() const = 0;
  // Blocks until all work submitted to the device has finished; throws on a
  // device error (like wait_and_throw()).
  virtual void synchronize() = 0;
  // "gpu", "cpu", ...
  virtual std::string kind() const = 0;
};

// Creates the device selected by --device=<type>.  Defined by the program:
// hip_device.h provides the HIP one used by every benchmark executable.
std::shared_ptr<Device> makeDevice(const std::string& type);
