// Detection of the optional parts of the benchmark concept, as in the
// reference's include/benchmark_traits.h:7-35:
//   verify(VerificationSetting&)            -> hasVerify
//   static getThroughputMetric(args)        -> hasGetThroughputMetric
//   run(std::vector<KernelTiming>&)         -> supportsQueueProfiling
// KernelTiming takes the place of sycl::event: one entry per device kernel
// the run submitted.
#pragma once

#include <chrono>
#include <functional>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

struct KernelTiming {
  // Host time spent submitting the kernel.
  std::chrono::nanoseconds submit{0};
  // Device execution time; valid once the device has been synchronised.
  std::function<std::chrono::nanoseconds()> duration;
};

namespace detail {

template <class T, class = void>
struct SupportsQueueProfiling : std::false_type {};
template <class T>
struct SupportsQueueProfiling<T, std::void_t<decltype(std::declval<T&>().run(std::declval<std::vector<KernelTiming>&>()))>>
    : std::true_type {};

template <class T, class = void>
struct HasVerify : std::false_type {};
template <class T>
struct HasVerify<T, std::void_t<decltype(&T::verify)>> : std::true_type {};

template <class T, class = void>
struct HasThroughputMetric : std::false_type {};
template <class T>
struct HasThroughputMetric<T, std::void_t<decltype(&T::getThroughputMetric)>> : std::true_type {};

// Optional extension: void extraResults(std::map<std::string, std::string>&)
// adds columns (precision, error, TFLOP/s) when --metrics=extended is given.
template <class T, class = void>
struct HasExtraResults : std::false_type {};
template <class T>
struct HasExtraResults<T,
    std::void_t<decltype(std::declval<T&>().extraResults(std::declval<std::map<std::string, std::string>&>()))>>
    : std::true_type {};

template <class T>
struct BenchmarkTraits {
  static constexpr bool hasVerify = HasVerify<T>::value;
  static constexpr bool hasGetThroughputMetric = HasThroughputMetric<T>::value;
  static constexpr bool supportsQueueProfiling = SupportsQueueProfiling<T>::value;
  static constexpr bool hasExtraResults = HasExtraResults<T>::value;
};

}  // namespace detail
