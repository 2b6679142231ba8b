// C++ API tests in the shape of the reference's lib_ensembles tests
// (lib_ensembles/test/test_gpurf.h, test_gpuert.h: fit / predict on the toy
// raw data and on spambase.csv, float and double), plus checks of the host
// logic (parameter packs, model split / aggregate, result aggregation).
//
//   test_host_api cpu <golden dir>             host logic, no device needed
//   test_host_api gpu <golden dir> <out dir>   fit / predict on device 0
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/dte.h"
#include "lib_dte.h"

namespace {
int g_failed = 0;
std::string g_golden, g_out;

#define EXPECT(cond)                                                                       \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      std::cout << "  FAILED " << __FILE__ << ":" << __LINE__ << ": " #cond << std::endl;  \
      throw std::runtime_error("expectation failed");                                      \
    }                                                                                      \
  } while (0)

struct Case {
  const char* name;
  std::function<void()> fn;
};

void run(const std::vector<Case>& cases) {
  for (const auto& c : cases) {
    try {
      c.fn();
      std::cout << "[       OK ] " << c.name << std::endl;
    } catch (const std::exception& e) {
      ++g_failed;
      std::cout << "[  FAILED  ] " << c.name << ": " << e.what() << std::endl;
    }
  }
}

col_array<char> read_raw(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  col_array<char> raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  raw.push_back('\0');  // the reference's raw arrays are C strings
  return raw;
}

using Node = lib_ensembles::GpuDteAlgorithmShared::gpuDTE_NodeHeader_Classify<float>;

auto& ensembles_face = EnsemblesLib::GetInstance();
auto& parser_face = ParsingLib::GetInstance();

template <typename T>
sp<lib_data::MlDataFrame<T>> toy_fit() {
  auto raw = read_raw(g_golden + "/toy_fit.csv");
  return parser_face.ParseData<T>(ParsingLib::kCsv, raw);
}
template <typename T>
sp<lib_data::MlDataFrame<T>> toy_predict() {
  auto raw = read_raw(g_golden + "/toy_predict.csv");
  return parser_face.ParseData<T>(ParsingLib::kCsv, raw);
}

// ---------------------------------------------------------------- CPU ----
void parse_toy() {
  auto f = toy_fit<float>();
  EXPECT(f && f->GetNrSamples() == 36 && f->GetNrFeatures() == 9 && f->GetNrTargets() == 2);
  const auto& s = f->GetSamples();
  for (int i = 0; i < 9; ++i) EXPECT(s[i] == (float)i);  // string column -> ids in order of appearance
  EXPECT(s[36 * 1 + 5] == 1.0f && s[36 * 8 + 35] == 8.0f);
  auto d = toy_predict<double>();
  EXPECT(d && d->GetNrSamples() == 9 && d->GetTargets()[1] == 1.0);
}

void parse_spambase() {
  auto f = parser_face.ParseFile<float>(ParsingLib::kCsv, g_golden + "/spambase.csv");
  EXPECT(f && f->GetNrSamples() == 4601 && f->GetNrFeatures() == 57 && f->GetNrTargets() == 2);
  double pos = 0;
  for (float t : f->GetTargets()) pos += t;
  EXPECT(pos == 1813);
  EXPECT(parser_face.ParseFile<float>(ParsingLib::kCsv, g_golden + "/missing.csv") == nullptr);
}

void param_packs() {
  auto rf = ensembles_face.CreateGpuRfParamPack();
  auto ert = ensembles_face.CreateGpuErtParamPack();
  EXPECT(rf->Get<int>(EnsemblesLib::kNrTrees) == 100 && rf->Get<int>(EnsemblesLib::kMaxDepth) == 1000);
  EXPECT(rf->Get<int>(EnsemblesLib::kMinNodeSize) == 5 && rf->Get<int>(EnsemblesLib::kNrFeatures) == 0);
  EXPECT(rf->Get<bool>(EnsemblesLib::kBagging) && !ert->Get<bool>(EnsemblesLib::kBagging));
  rf->Set(EnsemblesLib::kAlgoType, AlgorithmsLib::kRegression);
  EXPECT(rf->Get<int>(EnsemblesLib::kAlgoType) == 1);  // enum read back as int
  EXPECT(rf->Get<AlgorithmsLib::AlgorithmType>(EnsemblesLib::kAlgoType) == AlgorithmsLib::kRegression);
  rf->Set(EnsemblesLib::kMaxDepth, 7);
  rf->Set(EnsemblesLib::kSeed, (uint64_t)99);
  lib_ensembles::GpuDteAlgorithm<float> algo(0);
  auto parts = algo.SplitParameterPack(rf, 3);
  EXPECT(parts.size() == 3);
  EXPECT(parts[0]->Get<int>(EnsemblesLib::kNrTrees) == 34 && parts[2]->Get<int>(EnsemblesLib::kNrTrees) == 33);
  EXPECT(parts[1]->Get<int>(EnsemblesLib::kTreeOffset) == 34 && parts[2]->Get<int>(EnsemblesLib::kTreeOffset) == 67);
  for (auto& p : parts) {
    EXPECT(p->Get<int>(EnsemblesLib::kMaxDepth) == 7);  // every parameter survives the split
    EXPECT(p->Get<uint64_t>(EnsemblesLib::kSeed) == 99);
  }
  bool threw = false;
  try {
    rf->Get<int>(1000);
  } catch (const std::exception&) {
    threw = true;
  }
  EXPECT(threw);
}

sp<lib_models::MlModel> tiny_model() {
  // tree 0: x0 < 0.5 -> leaf A else (x1 < 2 -> leaf B, leaf C); tree 1: leaf D
  col_array<Node> nodes = {
      {2, 2, 0, 0, 0.5f},  // root 0
      {0, 0, 0, -1, 0},    // root 1 (leaf D, probs at 0)
      {0, 0, 2, -1, 0},    // leaf A
      {4, 2, 0, 1, 2.0f},  // inner
      {0, 0, 4, -1, 0},    // leaf B
      {0, 0, 6, -1, 0},    // leaf C
  };
  col_array<float> probs = {0, 5, 3, 1, 2, 2, 0, 7};
  auto m = ModelsLib::GetInstance().CreateModel();
  m->Add(ModelsLib::kNrTrees, 2);
  m->Add(ModelsLib::kNrTargets, 2);
  m->Add(ModelsLib::kNrFeatures, 2);
  m->Add(ModelsLib::kModelType, AlgorithmsLib::kClassification);
  m->Add(ModelsLib::kNodeArray, nodes);
  m->Add(ModelsLib::kProbArray, probs);
  return m;
}

void split_and_aggregate_models() {
  lib_ensembles::GpuDteAlgorithm<float> algo(0);
  auto m = tiny_model();
  auto parts = algo.SplitModel(m, 2);
  EXPECT(parts.size() == 2 && parts[0] != parts[1]);  // independent parts
  auto& n0 = parts[0]->Get<col_array<Node>>(ModelsLib::kNodeArray);
  auto& p0 = parts[0]->Get<col_array<float>>(ModelsLib::kProbArray);
  EXPECT(n0.size() == 5 && parts[0]->Get<int>(ModelsLib::kNrTrees) == 1);
  EXPECT(n0[0].child_start == 1 && n0[2].child_start == 3);  // level order
  EXPECT(p0.size() == 6 && p0[0] == 3 && p0[1] == 1 && p0[4] == 0 && p0[5] == 7);
  auto& n1 = parts[1]->Get<col_array<Node>>(ModelsLib::kNodeArray);
  EXPECT(n1.size() == 1 && n1[0].child_count == 0);
  auto back = algo.AggregateModels(parts);
  auto& nb = back->Get<col_array<Node>>(ModelsLib::kNodeArray);
  auto& pb = back->Get<col_array<float>>(ModelsLib::kProbArray);
  EXPECT(back->Get<int>(ModelsLib::kNrTrees) == 2 && nb.size() == 6);
  EXPECT(nb[0].attribute == 0 && nb[1].child_count == 0);  // roots first
  EXPECT(nb[0].child_start == 2 && nb[3].child_start == 4);
  // every leaf keeps its payload
  auto leaf_payload = [&](int node) { return std::vector<float>(pb.begin() + nb[node].probability_start,
                                                                pb.begin() + nb[node].probability_start + 2); };
  EXPECT((leaf_payload(1) == std::vector<float>{0, 5}) && (leaf_payload(2) == std::vector<float>{3, 1}));
  EXPECT((leaf_payload(4) == std::vector<float>{2, 2}) && (leaf_payload(5) == std::vector<float>{0, 7}));
}

void aggregate_results_sums_votes() {
  lib_ensembles::GpuDteAlgorithm<double> algo(0);
  auto a = DataLib::GetInstance().CreateResultData<double>();
  auto b = DataLib::GetInstance().CreateResultData<double>();
  a->AddPredictions({{1, 2}, {3, 0}});
  b->AddPredictions({{0, 4}, {1, 1}});
  auto r = algo.AggregateResults({a, b});
  EXPECT(r->GetPredictions()[0][1] == 6 && r->GetPredictions()[1][0] == 4);
  EXPECT(r->GetAccuracy({1, 0}) == 1.0);
  EXPECT(r->GetAccuracy({0, 1}) == 0.0);
}

void fit_without_device_throws() {
  int32_t n = 0;
  dte_device_count(&n);
  if (n > 0) return;  // only meaningful on a machine without a GPU
  auto rf = ensembles_face.CreateGpuRfAlgorithm<float>();
  bool threw = false;
  try {
    rf->Fit(toy_fit<float>(), ensembles_face.CreateGpuRfParamPack());
  } catch (const std::exception& e) {
    threw = std::string(e.what()).find("no HIP device") != std::string::npos;
  }
  EXPECT(threw);
}

// ---------------------------------------------------------------- GPU ----
template <typename T>
void votes_sum_to_trees(const sp<lib_data::MlResultData<T>>& r, int trees) {
  for (const auto& row : r->GetPredictions()) {
    T s = 0;
    for (T v : row) s += v;
    EXPECT(s == (T)trees);
  }
}

template <typename T>
void fit_predict_rawdata(bool rf, int trees) {
  auto algo = rf ? ensembles_face.CreateGpuRfAlgorithm<T>() : ensembles_face.CreateGpuErtAlgorithm<T>();
  auto params = rf ? ensembles_face.CreateGpuRfParamPack() : ensembles_face.CreateGpuErtParamPack();
  params->Set(EnsemblesLib::kNrTrees, trees);
  params->Set(EnsemblesLib::kAlgoType, AlgorithmsLib::kClassification);
  auto model = algo->Fit(toy_fit<T>(), params);
  auto data = toy_predict<T>();
  auto results = algo->Predict(data, model, params);
  const T acc = results->GetAccuracy(data->GetTargets());
  std::cout << "[          ] accuracy = " << acc << std::endl;
  votes_sum_to_trees(results, trees);
}

void write_votes(const std::string& name, const sp<lib_data::MlResultData<float>>& r) {
  std::ofstream out(g_out + "/" + name, std::ios::binary);
  for (const auto& row : r->GetPredictions())
    for (float v : row) {
      const int32_t i = (int32_t)v;
      out.write(reinterpret_cast<const char*>(&i), 4);
    }
}

void spambase_rf_400() {
  auto data = parser_face.ParseFile<float>(ParsingLib::kCsv, g_golden + "/spambase.csv");
  auto algo = ensembles_face.CreateGpuRfAlgorithm<float>();
  auto params = ensembles_face.CreateGpuRfParamPack();
  params->Set(EnsemblesLib::kNrTrees, 400);
  params->Set(EnsemblesLib::kAlgoType, AlgorithmsLib::kClassification);
  params->Set(EnsemblesLib::kSeed, (uint64_t)0);
  auto model = algo->Fit(data, params);
  auto results = algo->Predict(data, model, params);
  std::cout << "[          ] accuracy = " << results->GetAccuracy(data->GetTargets()) << std::endl;
  votes_sum_to_trees(results, 400);
  write_votes("spambase_rf_votes.bin", results);
}

void parts_concatenate_to_one_forest() {
  auto data = parser_face.ParseFile<float>(ParsingLib::kCsv, g_golden + "/spambase.csv");
  lib_ensembles::GpuDteAlgorithm<float> algo(0);
  auto params = ensembles_face.CreateGpuRfParamPack();
  params->Set(EnsemblesLib::kNrTrees, 30);
  params->Set(EnsemblesLib::kMaxDepth, 12);
  auto whole = algo.Fit(data, params);
  auto parts = algo.SplitParameterPack(params, 3);
  col_array<sp<lib_models::MlModel>> models;
  for (auto& p : parts) models.push_back(algo.Fit(data, p));
  auto agg = algo.AggregateModels(models);
  auto& a = whole->Get<col_array<Node>>(ModelsLib::kNodeArray);
  auto& b = agg->Get<col_array<Node>>(ModelsLib::kNodeArray);
  EXPECT(a.size() == b.size());
  EXPECT(std::memcmp(a.data(), b.data(), sizeof(Node) * a.size()) == 0);
  EXPECT(whole->Get<col_array<float>>(ModelsLib::kProbArray) == agg->Get<col_array<float>>(ModelsLib::kProbArray));
  auto r_whole = algo.Predict(data, whole, params);
  auto r_agg = algo.Predict(data, agg, params);  // imported from the arrays
  EXPECT(r_whole->GetPredictions() == r_agg->GetPredictions());
  // split the model, predict the parts, sum the votes
  auto split = algo.SplitModel(whole, 4);
  col_array<sp<lib_data::MlResultData<float>>> rs;
  for (auto& m : split) rs.push_back(algo.Predict(data, m, params));
  auto summed = algo.AggregateResults(rs);
  EXPECT(summed->GetPredictions() == r_whole->GetPredictions());
}

void regression_double() {
  const int n = 2000, F = 4;
  auto df = DataLib::GetInstance().CreateDataFrame<double>();
  col_array<double> X((size_t)n * F), y(n);
  uint64_t s = 12345;
  auto u = [&] {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) / 9007199254740992.0;
  };
  for (int f = 0; f < F; ++f)
    for (int i = 0; i < n; ++i) X[(size_t)f * n + i] = u();
  for (int i = 0; i < n; ++i) y[i] = std::sin(4 * X[i]) + X[n + i] * X[2 * n + i];
  df->AddSampleData(X);
  df->AddTargetData(y);
  df->AddFeatureMap(col_array<col_map<string, double>>(F));
  auto algo = ensembles_face.CreateGpuRfAlgorithm<double>();
  auto params = ensembles_face.CreateGpuRfParamPack();
  params->Set(EnsemblesLib::kNrTrees, 20);
  params->Set(EnsemblesLib::kAlgoType, AlgorithmsLib::kRegression);
  params->Set(EnsemblesLib::kNrFeatures, 2);
  auto model = algo->Fit(df, params);
  auto r = algo->Predict(df, model, params);
  double se = 0;
  for (int i = 0; i < n; ++i) se += std::pow(r->GetPredictions()[i][0] / 20 - y[i], 2);
  const double rmse = std::sqrt(se / n);
  std::cout << "[          ] rmse = " << rmse << std::endl;
  EXPECT(rmse < 0.2);
}
// T = double: the device computes in fp32, so a double data frame is
// rounded to fp32 at the boundary (the reference's ERT computes in double,
// gpuert.cu:383-384).  Rows that differ only below fp32 resolution land
// together: a double frame whose feature 0 holds v + 1e-12 (label 1) and
// v - 1e-12 (label 0) for fp32 values v cannot be split on that difference.
// The caller gets exactly the model of the fp32-rounded frame, every split
// point is an fp32 value, and each pair of rows gets the same votes.
void double_frame_rounds_to_fp32() {
  const int n = 1200, F = 3;
  col_array<double> Xd((size_t)n * F), y(n);
  uint64_t s = 777;
  auto u = [&] {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) / 9007199254740992.0;
  };
  for (int i = 0; i < n; i += 2) {  // rows i, i + 1: a pair equal in fp32
    const double v = (double)(float)(0.25 + 0.5 * u());  // an fp32 value
    const double x1 = u(), x2 = u() < 0.5 ? 0.0 : 1.0;
    for (int j = 0; j < 2; ++j) {
      Xd[i + j] = v + (j ? 1e-12 : -1e-12);  // not fp32-representable
      y[i + j] = j;
      Xd[(size_t)n + i + j] = x1;
      Xd[(size_t)2 * n + i + j] = x2;
    }
  }
  for (int i = 0; i < n; ++i) EXPECT((double)(float)Xd[i] != Xd[i]);
  col_array<float> Xf(Xd.begin(), Xd.end());
  col_array<float> yf(y.begin(), y.end());
  auto dd = DataLib::GetInstance().CreateDataFrame<double>();
  dd->AddSampleData(Xd);
  dd->AddTargetData(y);
  dd->AddFeatureMap(col_array<col_map<string, double>>(F));
  dd->AddTargetMap(col_map<string, double>{{"0", 0.0}, {"1", 1.0}});  // two classes
  auto df = DataLib::GetInstance().CreateDataFrame<float>();
  df->AddSampleData(Xf);
  df->AddTargetData(yf);
  df->AddFeatureMap(col_array<col_map<string, float>>(F));
  df->AddTargetMap(col_map<string, float>{{"0", 0.0f}, {"1", 1.0f}});
  auto params = ensembles_face.CreateGpuRfParamPack();
  params->Set(EnsemblesLib::kNrTrees, 16);
  params->Set(EnsemblesLib::kAlgoType, AlgorithmsLib::kClassification);
  auto md = ensembles_face.CreateGpuRfAlgorithm<double>()->Fit(dd, params);
  auto mf = ensembles_face.CreateGpuRfAlgorithm<float>()->Fit(df, params);
  using NodeD = lib_ensembles::GpuDteAlgorithmShared::gpuDTE_NodeHeader_Classify<double>;
  auto& a = md->Get<col_array<NodeD>>(ModelsLib::kNodeArray);
  auto& b = mf->Get<col_array<Node>>(ModelsLib::kNodeArray);
  EXPECT(a.size() == b.size());
  for (size_t i = 0; i < a.size(); ++i) {
    EXPECT(a[i].child_start == b[i].child_start && a[i].child_count == b[i].child_count &&
           a[i].probability_start == b[i].probability_start && a[i].attribute == b[i].attribute);
    EXPECT(a[i].split_point == (double)b[i].split_point);  // fp32 split points, widened
  }
  auto algo = ensembles_face.CreateGpuRfAlgorithm<double>();
  auto r = algo->Predict(dd, md, params);
  const auto& pr = r->GetPredictions();
  for (int i = 0; i + 1 < n; i += 2) EXPECT(pr[i] == pr[i + 1]);  // each pair routed together
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: test_host_api cpu|gpu <golden dir> [out dir]" << std::endl;
    return 2;
  }
  const std::string mode = argv[1];
  g_golden = argv[2];
  g_out = argc > 3 ? argv[3] : ".";
  if (mode == "cpu") {
    run({{"parse_toy", parse_toy},
         {"parse_spambase", parse_spambase},
         {"param_packs", param_packs},
         {"split_and_aggregate_models", split_and_aggregate_models},
         {"aggregate_results_sums_votes", aggregate_results_sums_votes},
         {"fit_without_device_throws", fit_without_device_throws}});
  } else {
    run({{"gpurf fit_predict_float_rawdata", [] { fit_predict_rawdata<float>(true, 100); }},
         {"gpurf fit_predict_double_rawdata", [] { fit_predict_rawdata<double>(true, 200); }},
         {"gpuert fit_predict_float_rawdata", [] { fit_predict_rawdata<float>(false, 100); }},
         {"gpuert fit_predict_double_rawdata", [] { fit_predict_rawdata<double>(false, 200); }},
         {"gpurf fit_predict_float_csvdata", spambase_rf_400},
         {"parts_concatenate_to_one_forest", parts_concatenate_to_one_forest},
         {"regression_double", regression_double},
         {"double_frame_rounds_to_fp32", double_frame_rounds_to_fp32}});
  }
  std::cout << (g_failed ? "FAILED " : "PASSED ") << g_failed << std::endl;
  return g_failed ? 1 : 0;
}
