// Result sinks of the benchmark harness, output-compatible with the
// reference's include/result_consumer.h:12-106:
//   - stdout: "********** Results for <name>**********" then "key: value [unit]"
//   - CSV (appending): on every flush, the union of all keys sorted byte-wise,
//     one "# Benchmark name,<keys...>" header line, then one row per benchmark.
#pragma once

#include <algorithm>
#include <fstream>
#include <map>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

class ResultConsumer {
public:
  virtual ~ResultConsumer() = default;
  virtual void proceedToBenchmark(const std::string& name) = 0;
  virtual void consumeResult(const std::string& result_name, const std::string& result, const std::string& unit = "") = 0;
  // Makes the results of all benchmarks seen so far visible in the output.
  virtual void flush() = 0;
  // Drops the current benchmark's results (used when a run throws).
  virtual void discard() {}
};

class OstreamResultConsumer : public ResultConsumer {
public:
  explicit OstreamResultConsumer(std::ostream& out) : out_(out) {}

  void proceedToBenchmark(const std::string& name) override {
    out_ << "********** Results for " << name << "**********" << std::endl;
  }

  void consumeResult(const std::string& key, const std::string& value, const std::string& unit = "") override {
    out_ << key << ": " << value;
    if(!unit.empty()) out_ << " [" << unit << "]";
    out_ << std::endl;
  }

  void flush() override { out_.flush(); }

private:
  std::ostream& out_;
};

class AppendingCsvResultConsumer : public ResultConsumer {
public:
  explicit AppendingCsvResultConsumer(const std::string& path) : file_(path, std::ios::app) {
    if(!file_) throw std::runtime_error("cannot open result file " + path);
  }

  void proceedToBenchmark(const std::string& name) override {
    current_ = name;
    if(rows_.find(name) == rows_.end()) order_.push_back(name);
    rows_[name];
  }

  void consumeResult(const std::string& key, const std::string& value, const std::string& /*unit*/ = "") override {
    rows_[current_][key] = value;
  }

  void flush() override {
    std::set<std::string> keys;  // std::string ordering == std::sort of the column names
    for(const auto& row : rows_)
      for(const auto& kv : row.second) keys.insert(kv.first);
    if(rows_.empty()) return;
    file_ << "# Benchmark name";
    for(const auto& k : keys) file_ << "," << k;
    file_ << "\n";
    for(const auto& name : order_) {
      const auto it = rows_.find(name);
      if(it == rows_.end()) continue;
      file_ << name;
      for(const auto& k : keys) {
        const auto v = it->second.find(k);
        if(v == it->second.end()) throw std::runtime_error("benchmark " + name + " has no value for column " + k);
        file_ << "," << v->second;
      }
      file_ << "\n";
    }
    file_.flush();
    rows_.clear();
    order_.clear();
  }

  void discard() override {
    rows_.erase(current_);
    order_.erase(std::remove(order_.begin(), order_.end(), current_), order_.end());
    current_.clear();
  }

private:
  std::ofstream file_;
  std::string current_;
  std::vector<std::string> order_;
  std::map<std::string, std::map<std::string, std::string>> rows_;
};
