// Implementation of the C++ host API (include/lib_dte.h) over libdte.
#include "include/lib_dte.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <queue>
#include <sstream>
#include <stdexcept>

#include "../../include/dte.h"

// ------------------------------------------------------------ sutil ----
namespace sutil {
bool any_type::read_integral(long long& out) const {
#define DTE_TRY(TYPE)                                      \
  if (const TYPE* p = std::any_cast<TYPE>(&value)) {       \
    out = (long long)*p;                                   \
    return true;                                           \
  }
  DTE_TRY(int)
  DTE_TRY(long)
  DTE_TRY(long long)
  DTE_TRY(unsigned)
  DTE_TRY(unsigned long)
  DTE_TRY(unsigned long long)
  DTE_TRY(short)
  DTE_TRY(bool)
  DTE_TRY(lib_algorithms::AlgorithmsInterface::AlgorithmType)
  DTE_TRY(lib_ensembles::EnsemblesInterface::DteParams)
#undef DTE_TRY
  return false;
}
}  // namespace sutil

// ---------------------------------------------------------- lib_core ----
namespace lib_core {
namespace {
struct LibException : std::exception {
  std::string msg;
  explicit LibException(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};
}  // namespace

CoreInterface& CoreInterface::GetInstance() {
  static CoreInterface instance;
  return instance;
}
void CoreInterface::ThrowException(string msg) { throw LibException(std::move(msg)); }
void CoreInterface::DebugMessage(string) {}
}  // namespace lib_core

namespace {
void check(int rc) {
  if (rc == DTE_OK) return;
  char buf[1024];
  dte_last_error(buf, sizeof(buf));
  CoreLib::GetInstance().ThrowException("libdte error " + std::to_string(rc) + ": " + buf);
}
}  // namespace

// ---------------------------------------------------- lib_algorithms ----
namespace lib_algorithms {
sutil::any_type& MlAlgorithmParams::at(int param) {
  if (param < 0 || param >= (int)params_.size())
    CoreLib::GetInstance().ThrowException("parameter id out of range: " + std::to_string(param));
  return params_[param];
}

AlgorithmsInterface& AlgorithmsInterface::GetInstance() {
  static AlgorithmsInterface instance;
  return instance;
}

sp<MlAlgorithmParams> AlgorithmsInterface::CreateAlgorithmParams(int size) {
  return std::make_shared<MlAlgorithmParams>(size);
}

namespace {
// Multi-device driver (gpu_algorithm.cc:17-76): one host thread per device.
template <typename T>
class GpuAlgorithm : public MlAlgorithm<T> {
 public:
  explicit GpuAlgorithm(sp<MlAlgorithm<T>> algo) : dev_(std::move(algo)) {}

  static int device_count() {
    int32_t n = 0;
    check(dte_device_count(&n));
    if (n <= 0) CoreLib::GetInstance().ThrowException("no HIP device available");
    return n;
  }

  sp<lib_models::MlModel> Fit(sp<lib_data::MlDataFrame<T>> data, sp<MlAlgorithmParams> params) override {
    const int n = device_count();
    if (n == 1) {
      auto p = SplitParameterPack(params, 1)[0];
      p->Set(EnsemblesLib::kDevId, 0);
      return dev_->Fit(data, p);
    }
    auto parts = SplitParameterPack(params, n);
    col_array<sp<lib_models::MlModel>> models(n);
    CoreLib::GetInstance().ParallelFor(0, n, [&](int i) {
      parts[i]->Set(EnsemblesLib::kDevId, i);
      models[i] = dev_->Fit(data, parts[i]);
    });
    return AggregateModels(models);
  }

  sp<lib_data::MlResultData<T>> Predict(sp<lib_data::MlDataFrame<T>> data, sp<lib_models::MlModel> model,
                                        sp<MlAlgorithmParams> params) override {
    const int n = device_count();
    if (n == 1) return dev_->Predict(data, model, params);
    auto parts = SplitParameterPack(params, n);
    auto models = SplitModel(model, n);
    col_array<sp<lib_data::MlResultData<T>>> results(n);
    CoreLib::GetInstance().ParallelFor(0, n, [&](int i) {
      parts[i]->Set(EnsemblesLib::kDevId, i);
      results[i] = dev_->Predict(data, models[i], parts[i]);
    });
    return AggregateResults(results);
  }
  sp<lib_models::MlModel> AggregateModels(col_array<sp<lib_models::MlModel>> models) override {
    return dev_->AggregateModels(models);
  }
  col_array<sp<lib_models::MlModel>> SplitModel(sp<lib_models::MlModel> model, const int parts) override {
    return dev_->SplitModel(model, parts);
  }
  sp<lib_data::MlResultData<T>> AggregateResults(col_array<sp<lib_data::MlResultData<T>>> results) override {
    return dev_->AggregateResults(results);
  }
  col_array<sp<MlAlgorithmParams>> SplitParameterPack(sp<MlAlgorithmParams> params, const int parts) override {
    return dev_->SplitParameterPack(params, parts);
  }

 private:
  sp<MlAlgorithm<T>> dev_;
};
}  // namespace

template <typename T>
sp<MlAlgorithm<T>> AlgorithmsInterface::CreateGpuAlgorithm(sp<MlAlgorithm<T>> algo) {
  return std::make_shared<GpuAlgorithm<T>>(std::move(algo));
}
template sp<MlAlgorithm<float>> AlgorithmsInterface::CreateGpuAlgorithm(sp<MlAlgorithm<float>>);
template sp<MlAlgorithm<double>> AlgorithmsInterface::CreateGpuAlgorithm(sp<MlAlgorithm<double>>);
}  // namespace lib_algorithms

// -------------------------------------------------------- lib_models ----
namespace lib_models {
ModelsInterface& ModelsInterface::GetInstance() {
  static ModelsInterface instance;
  return instance;
}
sp<MlModel> ModelsInterface::CreateModel() { return std::make_shared<MlModel>(); }
}  // namespace lib_models

// ---------------------------------------------------------- lib_data ----
namespace lib_data {
DataInterface& DataInterface::GetInstance() {
  static DataInterface instance;
  return instance;
}

template <typename T>
T MlResultData<T>::GetAccuracy(const col_array<T>& targets) {
  if (targets.empty()) return T(0);
  T hits = 0;
  for (size_t i = 0; i < targets.size() && i < predictions_.size(); ++i) {
    T high = 0;
    T label = 0;
    for (size_t c = 0; c < predictions_[i].size(); ++c)
      if (predictions_[i][c] > high) {
        high = predictions_[i][c];
        label = T(c);
      }
    if (std::abs(label - targets[i]) < T(0.001)) hits += 1;
  }
  return hits / T(targets.size());
}

template <typename T>
MlResultData<T>& MlResultData<T>::operator+=(const MlResultData<T>& rhs) {
  if (predictions_.size() != rhs.predictions_.size())
    CoreLib::GetInstance().ThrowException("result sets of different row counts");
  for (size_t i = 0; i < predictions_.size(); ++i)
    for (size_t c = 0; c < predictions_[i].size(); ++c) predictions_[i][c] += rhs.predictions_[i][c];
  return *this;
}
template class MlResultData<float>;
template class MlResultData<double>;
}  // namespace lib_data

// ------------------------------------------------------- lib_parsing ----
namespace lib_parsing {
ParsingInterface& ParsingInterface::GetInstance() {
  static ParsingInterface instance;
  return instance;
}

namespace {
std::vector<std::string> fields(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : line) {
    if (ch == ',' || ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

template <typename T>
sp<lib_data::MlDataFrame<T>> parse_csv(std::istream& in) {
  std::string line;
  if (!std::getline(in, line)) return nullptr;
  const auto names = fields(line);
  if (names.size() < 2) return nullptr;
  const size_t ncol = names.size();
  std::vector<col_map<string, T>> maps(ncol);
  std::vector<T> next_id(ncol, T(0));
  std::vector<col_array<T>> cols(ncol);
  while (std::getline(in, line)) {
    const auto f = fields(line);
    if (f.empty()) continue;
    if (f.size() != ncol) return nullptr;
    for (size_t j = 0; j < ncol; ++j) {
      std::istringstream ps(f[j]);
      T v;
      ps >> v;
      if (ps.fail()) {
        auto it = maps[j].find(f[j]);
        if (it == maps[j].end()) it = maps[j].emplace(f[j], next_id[j]++).first;
        v = it->second;
      } else {
        maps[j][f[j]] = v;
      }
      cols[j].push_back(v);
    }
  }
  const size_t n = cols[0].size();
  if (n == 0) return nullptr;
  col_array<T> samples;
  samples.reserve((ncol - 1) * n);
  for (size_t j = 0; j + 1 < ncol; ++j) samples.insert(samples.end(), cols[j].begin(), cols[j].end());
  auto df = DataLib::GetInstance().CreateDataFrame<T>();
  df->AddTargetData(std::move(cols[ncol - 1]));
  df->AddSampleData(std::move(samples));
  df->AddTargetMap(maps[ncol - 1]);
  maps.pop_back();
  df->AddFeatureMap(std::move(maps));
  return df;
}
}  // namespace

template <typename T>
sp<lib_data::MlDataFrame<T>> ParsingInterface::ParseData(FileFormat format, col_array<char>& raw_data) {
  if (format != kCsv) CoreLib::GetInstance().ThrowException("unsupported format");
  std::string text(raw_data.begin(), raw_data.end());
  text = text.c_str();  // the reference's raw arrays end in a NUL
  std::istringstream in(text);
  return parse_csv<T>(in);
}

template <typename T>
sp<lib_data::MlDataFrame<T>> ParsingInterface::ParseFile(FileFormat format, string filepath) {
  if (format != kCsv) CoreLib::GetInstance().ThrowException("unsupported format");
  std::ifstream in(filepath);
  if (!in) return nullptr;
  return parse_csv<T>(in);
}
template sp<lib_data::MlDataFrame<float>> ParsingInterface::ParseData(FileFormat, col_array<char>&);
template sp<lib_data::MlDataFrame<double>> ParsingInterface::ParseData(FileFormat, col_array<char>&);
template sp<lib_data::MlDataFrame<float>> ParsingInterface::ParseFile(FileFormat, string);
template sp<lib_data::MlDataFrame<double>> ParsingInterface::ParseFile(FileFormat, string);
}  // namespace lib_parsing

// ----------------------------------------------------- lib_ensembles ----
namespace lib_ensembles {
using lib_algorithms::MlAlgorithmParams;
using lib_models::MlModel;
template <typename T>
using Node = GpuDteAlgorithmShared::gpuDTE_NodeHeader_Classify<T>;

EnsemblesInterface& EnsemblesInterface::GetInstance() {
  static EnsemblesInterface instance;
  return instance;
}

namespace {
// ensembles_interface.cc:28-58 defaults.
sp<MlAlgorithmParams> default_pack(bool bagging) {
  auto p = AlgorithmsLib::GetInstance().CreateAlgorithmParams(EnsemblesInterface::kParamPackSize);
  p->Set(EnsemblesInterface::kNrTrees, 100);
  p->Set(EnsemblesInterface::kNrFeatures, 0);
  p->Set(EnsemblesInterface::kMaxDepth, 1000);
  p->Set(EnsemblesInterface::kMinNodeSize, 5);
  p->Set(EnsemblesInterface::kAlgoType, AlgorithmsLib::kClassification);
  p->Set(EnsemblesInterface::kBagging, bagging);
  p->Set(EnsemblesInterface::kEasyEnsemble, false);
  return p;
}

int get_int(const sp<MlAlgorithmParams>& p, int key, int dflt) { return p->Has(key) ? p->Get<int>(key) : dflt; }

struct ForestDeleter {
  void operator()(dte_forest* f) const { dte_forest_free(f); }
};

template <typename T>
std::vector<float> to_float(const col_array<T>& v) {
  return std::vector<float>(v.begin(), v.end());
}

// The nodes of trees [t0, t1) of a model, re-linked into a fresh array:
// roots first (as the reference's models hold them), then each tree's
// nodes in level order, leaf payloads copied in the same order.
template <typename T>
void gather_trees(const col_array<Node<T>>& nodes, const col_array<T>& probs, int width, int t0, int t1,
                  col_array<Node<T>>& out_nodes, col_array<T>& out_probs) {
  const size_t root0 = out_nodes.size();
  for (int t = t0; t < t1; ++t) out_nodes.push_back(nodes.at(t));
  for (int t = t0; t < t1; ++t) {
    std::queue<size_t> q;  // positions in out_nodes whose children are still to be copied
    q.push(root0 + (t - t0));
    while (!q.empty()) {
      const size_t pos = q.front();
      q.pop();
      Node<T>& nd = out_nodes[pos];
      if (nd.child_count > 0) {
        const int src = nd.child_start;
        const int cnt = nd.child_count;
        nd.child_start = (int)out_nodes.size();
        for (int c = 0; c < cnt; ++c) out_nodes.push_back(nodes.at(src + c));
        const size_t first = out_nodes.size() - cnt;
        for (int c = 0; c < cnt; ++c) q.push(first + c);
      } else {
        const int src = out_nodes[pos].probability_start;
        out_nodes[pos].probability_start = (int)out_probs.size();
        for (int k = 0; k < width; ++k) out_probs.push_back(probs.at(src + k));
      }
    }
  }
}

int model_width(MlModel& m) {
  return m.Get<AlgorithmsLib::AlgorithmType>(ModelsLib::kModelType) == AlgorithmsLib::kRegression
             ? 1
             : m.Get<int>(ModelsLib::kNrTargets);
}
}  // namespace

sp<MlAlgorithmParams> EnsemblesInterface::CreateGpuRfParamPack() { return default_pack(true); }
sp<MlAlgorithmParams> EnsemblesInterface::CreateGpuErtParamPack() { return default_pack(false); }

template <typename T>
sp<lib_algorithms::MlAlgorithm<T>> EnsemblesInterface::CreateGpuRfAlgorithm() {
  return AlgorithmsLib::GetInstance().CreateGpuAlgorithm<T>(std::make_shared<GpuDteAlgorithm<T>>(DTE_RANDOM_FOREST));
}
template <typename T>
sp<lib_algorithms::MlAlgorithm<T>> EnsemblesInterface::CreateGpuErtAlgorithm() {
  return AlgorithmsLib::GetInstance().CreateGpuAlgorithm<T>(std::make_shared<GpuDteAlgorithm<T>>(DTE_EXTRA_TREES));
}
template sp<lib_algorithms::MlAlgorithm<float>> EnsemblesInterface::CreateGpuRfAlgorithm();
template sp<lib_algorithms::MlAlgorithm<double>> EnsemblesInterface::CreateGpuRfAlgorithm();
template sp<lib_algorithms::MlAlgorithm<float>> EnsemblesInterface::CreateGpuErtAlgorithm();
template sp<lib_algorithms::MlAlgorithm<double>> EnsemblesInterface::CreateGpuErtAlgorithm();

template <typename T>
sp<MlModel> GpuDteAlgorithm<T>::Fit(sp<lib_data::MlDataFrame<T>> data, sp<MlAlgorithmParams> params) {
  auto& core = CoreLib::GetInstance();
  if (!data || !params) core.ThrowException("Fit: data and params are required");
  dte_params p;
  check(dte_params_default(algorithm_, &p));
  p.n_trees = params->Get<int>(EnsemblesInterface::kNrTrees);
  p.task = params->Has(EnsemblesInterface::kAlgoType) ? params->Get<int>(EnsemblesInterface::kAlgoType)
                                                      : DTE_CLASSIFICATION;
  p.n_features_per_node = get_int(params, EnsemblesInterface::kNrFeatures, 0);
  p.max_depth = get_int(params, EnsemblesInterface::kMaxDepth, 1000);
  p.min_node_size = get_int(params, EnsemblesInterface::kMinNodeSize, 5);
  p.device_id = get_int(params, EnsemblesInterface::kDevId, 0);
  if (params->Has(EnsemblesInterface::kSeed)) p.seed = params->Get<uint64_t>(EnsemblesInterface::kSeed);
  p.build_mode = get_int(params, EnsemblesInterface::kBuildMode, DTE_BUILD_BATCHED);
  p.tree_offset = get_int(params, EnsemblesInterface::kTreeOffset, 0);
  const int64_t N = data->GetNrSamples();
  const int F = data->GetNrFeatures();
  if ((int64_t)data->GetSamples().size() != N * F) core.ThrowException("Fit: samples do not match the frame's shape");
  p.n_classes = p.task == DTE_CLASSIFICATION ? data->GetNrTargets() : 0;
  const std::vector<float> X = to_float(data->GetSamples());
  const std::vector<float> y = to_float(data->GetTargets());
  dte_forest* raw = nullptr;
  check(dte_fit(X.data(), y.data(), N, F, &p, &raw));
  std::shared_ptr<dte_forest> forest(raw, ForestDeleter());
  dte_forest_info info;
  check(dte_forest_info_get(raw, &info));
  std::vector<dte_node> nodes(info.n_nodes);
  std::vector<float> probs(info.n_probabilities);
  check(dte_forest_export(raw, nodes.data(), info.n_nodes, probs.data(), info.n_probabilities));
  col_array<Node<T>> node_array(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i)
    node_array[i] = Node<T>{nodes[i].child_start, nodes[i].child_count, nodes[i].probability_start,
                            nodes[i].attribute, (T)nodes[i].split_point};
  auto model = ModelsLib::GetInstance().CreateModel();
  model->Add(ModelsLib::kNrTrees, (int)info.n_trees);
  model->Add(ModelsLib::kNrTargets, p.task == DTE_CLASSIFICATION ? (int)info.n_classes : 1);
  model->Add(ModelsLib::kNrFeatures, F);
  model->Add(ModelsLib::kModelType, (AlgorithmsLib::AlgorithmType)p.task);
  model->Add(ModelsLib::kNodeArray, node_array);
  model->Add(ModelsLib::kProbArray, col_array<T>(probs.begin(), probs.end()));
  model->Add(ModelsLib::kResidentForest, forest);
  return model;
}

template <typename T>
sp<lib_data::MlResultData<T>> GpuDteAlgorithm<T>::Predict(sp<lib_data::MlDataFrame<T>> data, sp<MlModel> model,
                                                          sp<MlAlgorithmParams> params) {
  auto& core = CoreLib::GetInstance();
  if (!data || !model) core.ThrowException("Predict: data and model are required");
  const int device = params ? get_int(params, EnsemblesInterface::kDevId, 0) : 0;
  const int width = model_width(*model);
  const bool regression =
      model->Get<AlgorithmsLib::AlgorithmType>(ModelsLib::kModelType) == AlgorithmsLib::kRegression;
  std::shared_ptr<dte_forest> forest;
  if (model->Has(ModelsLib::kResidentForest))
    forest = model->Get<std::shared_ptr<dte_forest>>(ModelsLib::kResidentForest);
  if (!forest) {
    auto& na = model->Get<col_array<Node<T>>>(ModelsLib::kNodeArray);
    auto& pa = model->Get<col_array<T>>(ModelsLib::kProbArray);
    std::vector<dte_node> nodes(na.size());
    for (size_t i = 0; i < na.size(); ++i)
      nodes[i] = dte_node{na[i].child_start, na[i].child_count, na[i].probability_start, na[i].attribute,
                          (float)na[i].split_point};
    const std::vector<float> probs(pa.begin(), pa.end());
    dte_forest* raw = nullptr;
    check(dte_forest_import(nodes.data(), (int64_t)nodes.size(), probs.data(), (int64_t)probs.size(),
                            model->Get<int>(ModelsLib::kNrTrees), regression ? 1 : width,
                            model->Get<int>(ModelsLib::kNrFeatures), regression ? DTE_REGRESSION : DTE_CLASSIFICATION,
                            device, &raw));
    forest.reset(raw, ForestDeleter());
    model->Add(ModelsLib::kResidentForest, forest);
  }
  const int64_t n = (int64_t)data->GetTargets().size() > 0
                        ? data->GetNrSamples()
                        : (int64_t)(data->GetSamples().size() / std::max(1, model->Get<int>(ModelsLib::kNrFeatures)));
  if ((int64_t)data->GetSamples().size() != n * model->Get<int>(ModelsLib::kNrFeatures))
    core.ThrowException("Predict: samples do not match the model's feature count");
  const std::vector<float> X = to_float(data->GetSamples());
  std::vector<float> out((size_t)n * width);
  check(dte_predict(forest.get(), X.data(), n, out.data()));
  col_array<col_array<T>> predictions(n);
  for (int64_t i = 0; i < n; ++i) predictions[i].assign(out.begin() + i * width, out.begin() + (i + 1) * width);
  auto result = DataLib::GetInstance().CreateResultData<T>();
  result->AddPredictions(std::move(predictions));
  return result;
}

template <typename T>
sp<MlModel> GpuDteAlgorithm<T>::AggregateModels(col_array<sp<MlModel>> models) {
  if (models.empty()) CoreLib::GetInstance().ThrowException("AggregateModels: no models");
  if (models.size() == 1) return models[0];
  const int width = model_width(*models[0]);
  int trees = 0;
  for (auto& m : models) {
    if (model_width(*m) != width) CoreLib::GetInstance().ThrowException("AggregateModels: models of different width");
    trees += m->Get<int>(ModelsLib::kNrTrees);
  }
  // Layout of gpudte_algorithm.cc:333-397: the roots of every model first,
  // then each model's other nodes.  gather_trees gives one model as
  // [t roots][other nodes] with child links into its own other nodes.
  col_array<Node<T>> all(trees);
  col_array<T> probs;
  int root_base = 0;
  for (auto& m : models) {
    const int t = m->Get<int>(ModelsLib::kNrTrees);
    col_array<Node<T>> part;
    gather_trees(m->Get<col_array<Node<T>>>(ModelsLib::kNodeArray), m->Get<col_array<T>>(ModelsLib::kProbArray),
                 width, 0, t, part, probs);
    const int rest_base = (int)all.size();
    for (auto& nd : part)
      if (nd.child_count > 0) nd.child_start = rest_base + (nd.child_start - t);
    std::copy(part.begin(), part.begin() + t, all.begin() + root_base);
    all.insert(all.end(), part.begin() + t, part.end());
    root_base += t;
  }
  auto out = ModelsLib::GetInstance().CreateModel();
  out->Add(ModelsLib::kNrTrees, trees);
  out->Add(ModelsLib::kNrTargets, models[0]->Get<int>(ModelsLib::kNrTargets));
  out->Add(ModelsLib::kNrFeatures, models[0]->Get<int>(ModelsLib::kNrFeatures));
  out->Add(ModelsLib::kModelType, models[0]->Get<AlgorithmsLib::AlgorithmType>(ModelsLib::kModelType));
  out->Add(ModelsLib::kNodeArray, all);
  out->Add(ModelsLib::kProbArray, probs);
  return out;
}

template <typename T>
col_array<sp<MlModel>> GpuDteAlgorithm<T>::SplitModel(sp<MlModel> model, const int parts) {
  if (parts <= 0) CoreLib::GetInstance().ThrowException("SplitModel: parts must be positive");
  const int trees = model->Get<int>(ModelsLib::kNrTrees);
  const int width = model_width(*model);
  auto& na = model->Get<col_array<Node<T>>>(ModelsLib::kNodeArray);
  auto& pa = model->Get<col_array<T>>(ModelsLib::kProbArray);
  col_array<sp<MlModel>> out;
  int t0 = 0;
  for (int i = 0; i < parts; ++i) {
    const int cnt = trees / parts + (i < trees % parts ? 1 : 0);
    col_array<Node<T>> nodes;
    col_array<T> probs;
    gather_trees(na, pa, width, t0, t0 + cnt, nodes, probs);
    auto m = ModelsLib::GetInstance().CreateModel();  // an independent model per part
    m->Add(ModelsLib::kNrTrees, cnt);
    m->Add(ModelsLib::kNrTargets, model->Get<int>(ModelsLib::kNrTargets));
    m->Add(ModelsLib::kNrFeatures, model->Get<int>(ModelsLib::kNrFeatures));
    m->Add(ModelsLib::kModelType, model->Get<AlgorithmsLib::AlgorithmType>(ModelsLib::kModelType));
    m->Add(ModelsLib::kNodeArray, nodes);
    m->Add(ModelsLib::kProbArray, probs);
    out.push_back(m);
    t0 += cnt;
  }
  return out;
}

template <typename T>
sp<lib_data::MlResultData<T>> GpuDteAlgorithm<T>::AggregateResults(col_array<sp<lib_data::MlResultData<T>>> results) {
  if (results.empty()) CoreLib::GetInstance().ThrowException("AggregateResults: no results");
  for (size_t i = 1; i < results.size(); ++i) *results[0] += *results[i];
  return results[0];
}

template <typename T>
col_array<sp<MlAlgorithmParams>> GpuDteAlgorithm<T>::SplitParameterPack(sp<MlAlgorithmParams> params,
                                                                        const int parts) {
  if (parts <= 0) CoreLib::GetInstance().ThrowException("SplitParameterPack: parts must be positive");
  const int trees = params->Get<int>(EnsemblesInterface::kNrTrees);
  const int offset0 = get_int(params, EnsemblesInterface::kTreeOffset, 0);
  col_array<sp<MlAlgorithmParams>> out;
  int t0 = 0;
  for (int i = 0; i < parts; ++i) {
    const int cnt = trees / parts + (i < trees % parts ? 1 : 0);
    auto p = AlgorithmsLib::GetInstance().CreateAlgorithmParams(EnsemblesInterface::kParamPackSize);
    for (int k = 0; k < std::min(params->Size(), EnsemblesInterface::kParamPackSize); ++k)
      if (params->Has(k)) {
        // copy every parameter the caller set (the reference keeps only kNrTrees)
        if (k == EnsemblesInterface::kSeed)
          p->Set(k, params->Get<uint64_t>(k));
        else if (k == EnsemblesInterface::kBagging || k == EnsemblesInterface::kEasyEnsemble)
          p->Set(k, params->Get<bool>(k));
        else
          p->Set(k, params->Get<int>(k));
      }
    p->Set(EnsemblesInterface::kNrTrees, cnt);
    p->Set(EnsemblesInterface::kTreeOffset, offset0 + t0);  // first tree of this part
    out.push_back(p);
    t0 += cnt;
  }
  return out;
}

template class GpuDteAlgorithm<float>;
template class GpuDteAlgorithm<double>;
}  // namespace lib_ensembles
