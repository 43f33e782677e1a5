// Command line and per-run arguments, following the reference's
// include/command_line.h:73-205: "--key=value" arguments (a key given twice is
// an error), bare flags, BenchmarkArgs with the same fields and defaults
// (--size=3072, --local=256, --num-runs=5), and --output=stdio|<csv file>.
//
// Deliberate differences (the reference's behaviour is a bug in each case):
//   - --warmup-run is parsed into BenchmarkArgs::warmup_run (the reference
//     documents it but never sets it, include/command_line.h:166-167).
//   - --verification-begin / --verification-range are parsed by element
//     count (the reference checks the string length, :44-51, so "1,1,1" throws).
#pragma once

#include <array>
#include <cstddef>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "device.h"
#include "result_consumer.h"

namespace detail {

template <class T>
inline T parse_value(const std::string& s) {
  std::istringstream in(s);
  T v{};
  in >> v;
  if(in.fail() || !(in >> std::ws).eof()) throw std::invalid_argument("cannot parse '" + s + "'");
  return v;
}

template <>
inline std::string parse_value<std::string>(const std::string& s) {
  return s;
}

template <>
inline bool parse_value<bool>(const std::string& s) {
  if(s == "1" || s == "true" || s == "yes" || s == "on") return true;
  if(s == "0" || s == "false" || s == "no" || s == "off") return false;
  throw std::invalid_argument("cannot parse '" + s + "' as a boolean");
}

// "a,b,c" -> up to three sizes; missing trailing entries take `fill`.
inline std::array<std::size_t, 3> parse_triple(const std::string& s, std::size_t fill) {
  std::vector<std::size_t> v;
  std::stringstream in(s);
  std::string item;
  while(std::getline(in, item, ',')) v.push_back(parse_value<std::size_t>(item));
  if(v.empty() || v.size() > 3) throw std::invalid_argument("expected 1 to 3 comma-separated values: " + s);
  std::array<std::size_t, 3> out{fill, fill, fill};
  for(std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  return out;
}

}  // namespace detail

class CommandLine {
public:
  CommandLine() = default;

  CommandLine(int argc, char** argv) {
    for(int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      const auto eq = a.find('=');
      if(eq == std::string::npos) {
        flags_.insert(a);
        continue;
      }
      const std::string key = a.substr(0, eq);
      if(args_.count(key)) throw std::invalid_argument("Encountered command line argument several times: " + key);
      args_[key] = a.substr(eq + 1);
    }
  }

  bool isArgSet(const std::string& key) const { return args_.count(key) != 0; }
  bool isFlagSet(const std::string& flag) const { return flags_.count(flag) != 0; }

  template <class T>
  T getOrDefault(const std::string& key, const T& fallback) const {
    const auto it = args_.find(key);
    if(it == args_.end()) return fallback;
    try {
      return detail::parse_value<T>(it->second);
    } catch(const std::invalid_argument& e) {
      throw std::invalid_argument(key + ": " + e.what());
    }
  }

  template <class T>
  T get(const std::string& key) const {
    if(!isArgSet(key)) throw std::invalid_argument("Command line argument was requested but missing: " + key);
    return getOrDefault<T>(key, T{});
  }

private:
  std::unordered_map<std::string, std::string> args_;
  std::unordered_set<std::string> flags_;
};

struct VerificationSetting {
  bool enabled = true;
  std::array<std::size_t, 3> begin{0, 0, 0};
  std::array<std::size_t, 3> range{1, 1, 1};
  // Number of elements of the verification window (0 disables verification).
  std::size_t size() const { return range[0] * range[1] * range[2]; }
};

struct BenchmarkArgs {
  std::size_t problem_size = 0;
  std::size_t local_size = 0;
  std::size_t num_runs = 0;
  std::shared_ptr<Device> device;
  VerificationSetting verification;
  CommandLine cli;
  std::shared_ptr<ResultConsumer> result_consumer;
  bool warmup_run = false;
};

class BenchmarkCommandLine {
public:
  BenchmarkCommandLine(int argc, char** argv) : cli_(argc, argv) {}

  BenchmarkArgs getBenchmarkArgs() const {
    BenchmarkArgs a;
    a.problem_size = cli_.getOrDefault<std::size_t>("--size", 3072);
    a.local_size = cli_.getOrDefault<std::size_t>("--local", 256);
    a.num_runs = cli_.getOrDefault<std::size_t>("--num-runs", 5);
    if(a.num_runs == 0) throw std::invalid_argument("--num-runs must be at least 1");
    a.device = makeDevice(cli_.getOrDefault<std::string>("--device", "default"));
    a.verification.enabled = !cli_.isFlagSet("--no-verification");
    if(cli_.isArgSet("--verification-begin"))
      a.verification.begin = detail::parse_triple(cli_.get<std::string>("--verification-begin"), 0);
    if(cli_.isArgSet("--verification-range"))
      a.verification.range = detail::parse_triple(cli_.get<std::string>("--verification-range"), 1);
    a.cli = cli_;
    a.result_consumer = makeConsumer(cli_.getOrDefault<std::string>("--output", "stdio"));
    a.warmup_run = cli_.isFlagSet("--warmup-run");
    return a;
  }

private:
  static std::shared_ptr<ResultConsumer> makeConsumer(const std::string& out) {
    if(out == "stdio") return std::make_shared<OstreamResultConsumer>(std::cout);
    return std::make_shared<AppendingCsvResultConsumer>(out);
  }

  CommandLine cli_;
};
