// C++ host API of the tree-ensemble path, over the libdte C ABI.
//
// The public surface keeps the reference's names and meaning so callers of
// lib_ensembles / lib_algorithms / lib_models / lib_data / lib_parsing /
// lib_core switch without code changes:
//   EnsemblesInterface::CreateGpuRfAlgorithm<T>, CreateGpuErtAlgorithm<T>,
//     CreateGpuRfParamPack, CreateGpuErtParamPack and the kDevId .. kBagging
//     keys (lib_ensembles/include/ensembles_interface.h:7-38);
//   MlAlgorithm<T>::Fit / Predict / AggregateModels / SplitModel /
//     AggregateResults / SplitParameterPack (lib_algorithms/include/
//     ml_algorithm.h:29-51);
//   AlgorithmsInterface::CreateAlgorithmParams, CreateGpuAlgorithm<T>,
//     kClassification / kRegression (algorithms_interface.h:7-21);
//   MlModel Add / Get<T> with the GpuDteModel keys (lib_models/include/
//     models_interface.h:9-19; ml_model.h);
//   MlDataFrame<T> (feature-major samples) and MlResultData<T>
//     (AddPredictions, GetAccuracy, operator+=) (lib_data);
//   ParsingInterface::ParseData / ParseFile with kCsv (lib_parsing);
//   CoreInterface::ThrowException / ParallelFor (lib_core).
// Errors of the C ABI are rethrown through CoreInterface::ThrowException,
// as the reference's code throws.
//
// T is float or double; the device computes in fp32, so double samples and
// thresholds are rounded to fp32 at the boundary.
#pragma once

#include <any>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

template <typename T>
using sp = std::shared_ptr<T>;
template <typename T>
using col_array = std::vector<T>;
template <typename K, typename V>
using col_map = std::map<K, V>;
using std::string;

struct dte_forest;

namespace sutil {
// Type-erased parameter / model value.  Integral and enum types convert
// into each other on read (the reference stores kAlgoType as an enum and
// reads it both as the enum and as int).
struct any_type {
  any_type() = default;
  template <typename T, typename = std::enable_if_t<!std::is_same_v<std::decay_t<T>, any_type>>>
  any_type(T&& v) : value(std::forward<T>(v)) {}  // NOLINT: implicit like the reference

  template <typename T>
  T& get_value() {
    if (T* p = std::any_cast<T>(&value)) return *p;
    if constexpr (std::is_integral_v<T> || std::is_enum_v<T>) {
      long long v;
      if (read_integral(v)) {
        value = static_cast<T>(v);
        return *std::any_cast<T>(&value);
      }
    }
    throw std::bad_any_cast();
  }
  bool empty() const { return !value.has_value(); }
  std::any value;

 private:
  bool read_integral(long long& out) const;
};
}  // namespace sutil

namespace lib_core {
class CoreInterface {
 public:
  static CoreInterface& GetInstance();
  [[noreturn]] void ThrowException(string msg);
  void DebugMessage(string msg);

  // Runs func(i) for i in [start, end) on hardware_concurrency() threads in
  // contiguous slices (core_interface.h:22-61).
  template <typename I, typename F>
  void ParallelFor(I start, I end, F func) {
    const unsigned hint = std::thread::hardware_concurrency();
    const I threads = (I)(hint == 0 ? 8 : hint);
    const I n = end - start;
    if (n <= 0) return;
    const I slice = std::max<I>((n + threads - 1) / threads, (I)1);
    std::vector<std::thread> pool;
    for (I a = start; a < end; a += slice) {
      const I b = std::min<I>(a + slice, end);
      pool.emplace_back([&func, a, b] {
        for (I k = a; k < b; ++k) func(k);
      });
    }
    for (auto& t : pool) t.join();
  }

 private:
  CoreInterface() = default;
};
}  // namespace lib_core
using CoreLib = lib_core::CoreInterface;

namespace lib_algorithms {
class MlAlgorithmParams {
 public:
  explicit MlAlgorithmParams(int size) : params_(size) {}
  template <typename T>
  void Set(int param, T value) {
    at(param) = sutil::any_type(value);
  }
  template <typename T>
  const T& Get(const int param) {
    return at(param).get_value<T>();
  }
  bool Has(int param) const { return param >= 0 && param < (int)params_.size() && !params_[param].empty(); }
  int Size() const { return (int)params_.size(); }

 private:
  sutil::any_type& at(int param);
  col_array<sutil::any_type> params_;
};
}  // namespace lib_algorithms

namespace lib_models {
class MlModel {
 public:
  template <typename T>
  void Add(const int id, const T& data) {
    data_[id] = sutil::any_type(data);
  }
  template <typename T>
  T& Get(const int id);
  bool Has(int id) const { return data_.count(id) != 0; }
  void Remove(int id) { data_.erase(id); }

 private:
  std::map<int, sutil::any_type> data_;
};

class ModelsInterface {
 public:
  static ModelsInterface& GetInstance();
  sp<MlModel> CreateModel();
  enum GpuDteModel {
    kNrTrees,
    kNrTargets,
    kNrFeatures,
    kModelType,
    kNodeArray,
    kProbArray,
    kGpuDteModelEndMarker
  };
  // Beyond the reference's keys: the forest kept resident in HBM by Fit
  // (a std::shared_ptr<dte_forest>), reused by Predict.
  static constexpr int kResidentForest = kGpuDteModelEndMarker;

 private:
  ModelsInterface() = default;
};

template <typename T>
T& MlModel::Get(const int id) {
  auto it = data_.find(id);
  if (it == data_.end()) CoreLib::GetInstance().ThrowException("model has no entry " + std::to_string(id));
  return it->second.get_value<T>();
}
}  // namespace lib_models
using ModelsLib = lib_models::ModelsInterface;

namespace lib_data {
template <typename T>
class MlDataFrame {
 public:
  void AddTargetData(col_array<T> targets) { targets_ = std::move(targets); }
  void AddSampleData(col_array<T> samples) { samples_ = std::move(samples); }
  void AddFeatureMap(col_array<col_map<string, T>> m) { feature_maps_ = std::move(m); }
  void AddTargetMap(col_map<string, T> m) { target_map_ = std::move(m); }
  int GetNrSamples() { return (int)targets_.size(); }
  int GetNrFeatures() { return (int)feature_maps_.size(); }
  int GetNrTargets() { return (int)target_map_.size(); }
  const col_array<T>& GetTargets() { return targets_; }
  const col_array<T>& GetSamples() { return samples_; }

 private:
  col_array<T> targets_;
  col_array<T> samples_;  // feature-major: samples_[f * n + row]
  col_array<col_map<string, T>> feature_maps_;
  col_map<string, T> target_map_;
};

template <typename T>
class MlResultData {
 public:
  void AddPredictions(col_array<col_array<T>> predictions) {
    if (predictions_.empty()) predictions_ = std::move(predictions);
  }
  const col_array<col_array<T>>& GetPredictions() const { return predictions_; }
  // Share of rows whose highest vote (first on ties, class 0 when all are
  // zero) equals the target (ml_resultdata_impl.cc:24-39).
  T GetAccuracy(const col_array<T>& targets);
  MlResultData<T>& operator+=(const MlResultData<T>& rhs);
  void AddSingleValue(string name, T value) { single_[name] = value; }
  T GetSingleValue(string name) { return single_[name]; }

 private:
  col_array<col_array<T>> predictions_;
  col_map<string, T> single_;
};

class DataInterface {
 public:
  static DataInterface& GetInstance();
  template <typename T>
  sp<MlDataFrame<T>> CreateDataFrame() {
    return std::make_shared<MlDataFrame<T>>();
  }
  template <typename T>
  sp<MlResultData<T>> CreateResultData() {
    return std::make_shared<MlResultData<T>>();
  }

 private:
  DataInterface() = default;
};
}  // namespace lib_data
using DataLib = lib_data::DataInterface;

namespace lib_parsing {
class ParsingInterface {
 public:
  static ParsingInterface& GetInstance();
  enum FileFormat { kCsv = 0 };
  // csv_parser.cc:36-104: header line of column names; ',' and blanks
  // separate fields; a field that reads as a number is that number, any
  // other becomes its column's next dense id in order of first appearance;
  // the last column is the target.  Returns nullptr on a malformed input.
  template <typename T>
  sp<lib_data::MlDataFrame<T>> ParseData(FileFormat format, col_array<char>& raw_data);
  template <typename T>
  sp<lib_data::MlDataFrame<T>> ParseFile(FileFormat format, string filepath);

 private:
  ParsingInterface() = default;
};
}  // namespace lib_parsing
using ParsingLib = lib_parsing::ParsingInterface;

namespace lib_algorithms {
template <typename T>
class MlAlgorithm {
 public:
  virtual ~MlAlgorithm() = default;
  virtual sp<lib_models::MlModel> Fit(sp<lib_data::MlDataFrame<T>> data, sp<MlAlgorithmParams> params) = 0;
  virtual sp<lib_data::MlResultData<T>> Predict(sp<lib_data::MlDataFrame<T>> data, sp<lib_models::MlModel> model,
                                                sp<MlAlgorithmParams> params) = 0;
  virtual sp<lib_models::MlModel> AggregateModels(col_array<sp<lib_models::MlModel>> models) = 0;
  virtual col_array<sp<lib_models::MlModel>> SplitModel(sp<lib_models::MlModel> model, const int parts) = 0;
  virtual sp<lib_data::MlResultData<T>> AggregateResults(col_array<sp<lib_data::MlResultData<T>>> results) = 0;
  virtual col_array<sp<MlAlgorithmParams>> SplitParameterPack(sp<MlAlgorithmParams> params, const int parts) = 0;
};

class AlgorithmsInterface {
 public:
  static AlgorithmsInterface& GetInstance();
  enum AlgorithmType { kClassification = 0, kRegression };
  sp<MlAlgorithmParams> CreateAlgorithmParams(int size);
  // Wraps a device algorithm in the multi-device driver (gpu_algorithm.cc:
  // 17-50): parameters split over the devices, one host thread per device,
  // models aggregated.
  template <typename T>
  sp<MlAlgorithm<T>> CreateGpuAlgorithm(sp<MlAlgorithm<T>> algo);

 private:
  AlgorithmsInterface() = default;
};
}  // namespace lib_algorithms
using AlgorithmsLib = lib_algorithms::AlgorithmsInterface;

namespace lib_ensembles {
// gpuDTE_NodeHeader_Classify<T> (gpudte_algorithm_shared.h:82-90): the
// exported model node.
namespace GpuDteAlgorithmShared {
template <typename T>
struct gpuDTE_NodeHeader_Classify {
  int child_start;
  int child_count;
  int probability_start;
  int attribute;
  T split_point;
};
}  // namespace GpuDteAlgorithmShared

class EnsemblesInterface {
 public:
  static EnsemblesInterface& GetInstance();
  template <typename T>
  sp<lib_algorithms::MlAlgorithm<T>> CreateGpuRfAlgorithm();
  template <typename T>
  sp<lib_algorithms::MlAlgorithm<T>> CreateGpuErtAlgorithm();
  sp<lib_algorithms::MlAlgorithmParams> CreateGpuRfParamPack();
  sp<lib_algorithms::MlAlgorithmParams> CreateGpuErtParamPack();
  enum CommonParams { kDevId = 0, kCommonEndMarker };
  enum DteParams {
    kNrTrees = kCommonEndMarker,
    kAlgoType,
    kNrFeatures,
    kMaxDepth,
    kMinNodeSize,
    kEasyEnsemble,
    kBagging,
    kDteEndMarker
  };
  // Keys beyond the reference's pack: RNG seed (uint64), the build driver
  // (int: 0 batched, 1 plain per-node) and the index of the part's first
  // tree (int; set by SplitParameterPack so that parts concatenate to the
  // forest of one build).
  static constexpr int kSeed = kDteEndMarker;
  static constexpr int kBuildMode = kDteEndMarker + 1;
  static constexpr int kTreeOffset = kDteEndMarker + 2;
  static constexpr int kParamPackSize = kDteEndMarker + 3;

 private:
  EnsemblesInterface() = default;
};

// The device algorithm: GpuDteAlgorithm<T> (gpudte_algorithm.h) with the
// variant (random forest / extra trees) fixed at construction.
template <typename T>
class GpuDteAlgorithm : public lib_algorithms::MlAlgorithm<T> {
 public:
  explicit GpuDteAlgorithm(int algorithm) : algorithm_(algorithm) {}
  sp<lib_models::MlModel> Fit(sp<lib_data::MlDataFrame<T>> data,
                              sp<lib_algorithms::MlAlgorithmParams> params) override;
  sp<lib_data::MlResultData<T>> Predict(sp<lib_data::MlDataFrame<T>> data, sp<lib_models::MlModel> model,
                                        sp<lib_algorithms::MlAlgorithmParams> params) override;
  sp<lib_models::MlModel> AggregateModels(col_array<sp<lib_models::MlModel>> models) override;
  col_array<sp<lib_models::MlModel>> SplitModel(sp<lib_models::MlModel> model, const int parts) override;
  sp<lib_data::MlResultData<T>> AggregateResults(col_array<sp<lib_data::MlResultData<T>>> results) override;
  col_array<sp<lib_algorithms::MlAlgorithmParams>> SplitParameterPack(sp<lib_algorithms::MlAlgorithmParams> params,
                                                                      const int parts) override;
  int algorithm() const { return algorithm_; }

 private:
  int algorithm_;
};
}  // namespace lib_ensembles
using EnsemblesLib = lib_ensembles::EnsemblesInterface;
