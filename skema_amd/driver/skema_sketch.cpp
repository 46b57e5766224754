// Command-line driver for the sketch route of Skema's `skema` binary
// (driver/main.cpp:149-224 -> sketchysvd, src/Skema_SketchySVD.cpp:1243-1577),
// on one MI355X through the C++ layer in include/skema_mi355x.hpp.
//
//   skema_sketch --input A.mtx|A.bin [--sparse] --rank r [--range k] [--core s]
//                [--window w] [--model gauss|sparse-sign] [--seeds a,b,c,d]
//                [--eta e] [--nu v] [--kernel gaussrbf --gamma g]
//                [--precision fp64|fp32|bf16|bf16x3|fp8] [--history-file h.json]
//                [--symmetric | --asymmetric] [--force-three-sketch]
//                [--normalize-matrix x] [--sketch-compute-svals-iters]
//                [--debug-file f] [--debug] [--print-level p] [--map-dist uniform|normal]
//
// Flags are the reference's (src/Skema_AlgParams.cpp:242-330, driver/main.cpp:
// 149-224), parsed against an explicit table: a flag that takes no value
// never swallows the next token, the iSVD / PRIMME flags are accepted and
// reported as ignored, and an unknown flag or one this build does not
// provide (--rr, --solver other than sketch) is an error.  Like the
// reference, a flag may be written without its leading "--".
//
// --symmetric (and kernel input unless --force-three-sketch) takes the
// SketchySPD route (src/Skema_SketchySVD.cpp:1243-1277): eigenvalues as
// "svals" and eigenpair residuals as "rnrms" in the history.
//
// Inputs: MatrixMarket "array" (dense, column-major values) or "coordinate"
// (read as CSR with --sparse, densified otherwise), and the Kokkos-Kernels
// binary CRS (.bin: int32 nrows, uint64 nnz, uint64 rowptr[n+1], int32
// col[nnz], double val[nnz]).  With --kernel gaussrbf the input rows are the
// points (n x d) and the sketched matrix is the n x n RBF kernel, as the
// reference's dense driver arranges (driver/main.cpp:99-105).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/skema_mi355x.hpp"

namespace {

struct Matrix {
  int64_t m = 0, n = 0;
  bool sparse = false;
  std::vector<double> dense;  // column-major
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

[[noreturn]] void fail(const std::string& msg) {
  std::cerr << "skema_sketch: " << msg << std::endl;
  std::exit(1);
}

Matrix read_mtx(const std::string& path, bool want_sparse) {
  std::ifstream f(path);
  if (!f) fail("cannot open " + path);
  std::string line;
  std::getline(f, line);
  std::string lower = line;
  std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
  if (lower.rfind("%%matrixmarket", 0) != 0) fail("not a MatrixMarket file: " + path);
  const bool array = lower.find("array") != std::string::npos;
  const bool symmetric = lower.find("symmetric") != std::string::npos;
  while (std::getline(f, line) && (line.empty() || line[0] == '%')) {
  }
  std::istringstream hdr(line);
  Matrix A;
  int64_t nnz = 0;
  hdr >> A.m >> A.n;
  if (!array) hdr >> nnz;
  if (array) {
    A.dense.resize(A.m * A.n);
    for (auto& x : A.dense) f >> x;  // column-major order
    if (!f) fail("truncated array data in " + path);
    return A;
  }
  std::vector<std::tuple<int64_t, int64_t, double>> t;
  for (int64_t e = 0; e < nnz; ++e) {
    int64_t i, j;
    double v;
    f >> i >> j >> v;
    if (!f) fail("truncated coordinate data in " + path);
    t.emplace_back(i - 1, j - 1, v);
    if (symmetric && i != j) t.emplace_back(j - 1, i - 1, v);
  }
  std::sort(t.begin(), t.end());
  if (want_sparse) {
    A.sparse = true;
    A.rowptr.assign(A.m + 1, 0);
    for (auto& e : t) A.rowptr[std::get<0>(e) + 1]++;
    for (int64_t r = 0; r < A.m; ++r) A.rowptr[r + 1] += A.rowptr[r];
    for (auto& e : t) {
      A.col.push_back((int32_t)std::get<1>(e));
      A.val.push_back(std::get<2>(e));
    }
  } else {
    A.dense.assign(A.m * A.n, 0.0);
    for (auto& e : t) A.dense[std::get<0>(e) + std::get<1>(e) * A.m] += std::get<2>(e);
  }
  return A;
}

Matrix read_bin(const std::string& path, bool want_sparse) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail("cannot open " + path);
  int32_t n = 0;
  uint64_t nnz = 0;
  f.read(reinterpret_cast<char*>(&n), 4);
  f.read(reinterpret_cast<char*>(&nnz), 8);
  Matrix A;
  A.m = A.n = n;
  std::vector<uint64_t> rp(n + 1);
  f.read(reinterpret_cast<char*>(rp.data()), 8 * (n + 1));
  A.col.resize(nnz);
  A.val.resize(nnz);
  f.read(reinterpret_cast<char*>(A.col.data()), 4 * nnz);
  f.read(reinterpret_cast<char*>(A.val.data()), 8 * nnz);
  if (!f) fail("truncated binary CRS " + path);
  A.rowptr.assign(rp.begin(), rp.end());
  A.sparse = true;
  if (!want_sparse) {
    A.dense.assign((size_t)n * n, 0.0);
    for (int64_t r = 0; r < n; ++r)
      for (int64_t e = A.rowptr[r]; e < A.rowptr[r + 1]; ++e) A.dense[r + (int64_t)A.col[e] * n] += A.val[e];
    A.sparse = false;
  }
  return A;
}

std::vector<uint64_t> parse_seeds(std::string s) {
  std::replace(s.begin(), s.end(), '[', ' ');
  std::replace(s.begin(), s.end(), ']', ' ');
  std::replace(s.begin(), s.end(), ',', ' ');
  std::istringstream is(s);
  std::vector<uint64_t> v;
  uint64_t x;
  while (is >> x) v.push_back(x);
  if (v.size() != 4) fail("--seeds needs four values");
  return v;
}

void write_history(const std::string& history, const std::string& js) {
  if (history.empty()) {
    std::cout << js << std::endl;
  } else {
    std::ofstream o(history);
    o << js << std::endl;
  }
}

void json_array(std::ostringstream& js, const char* key, const std::vector<double>& v) {
  js << ",\"" << key << "\":[";
  for (size_t i = 0; i < v.size(); ++i) js << (i ? "," : "") << v[i];
  js << "]";
}

// The final factors under --debug-file (src/Skema_SketchySVD.cpp:1387-1404).
void dump_factors(const skema_mi355x::AlgParams& p, const skema_mi355x::DeviceMatrix* U, const std::vector<double>& S,
                  const skema_mi355x::DeviceMatrix* V) {
  using namespace skema_mi355x::detail;
  if (p.debug_filename.empty()) return;
  if (U) write_matrix(debug_name(p.debug_filename, "U.txt"), U->to_host().data(), U->extent(0), U->extent(1));
  write_vector(debug_name(p.debug_filename, "S.txt"), S);
  if (V) write_matrix(debug_name(p.debug_filename, "V.txt"), V->to_host().data(), V->extent(0), V->extent(1));
}

template <class M, class D>
int run(const skema_mi355x::AlgParams& p, const M& A, const std::string& history, const skema_mi355x::Context& ctx) {
  skema_mi355x::SketchySVD<M, D> sk(p, ctx);
  sk.linear_update(A);
  auto usv = sk.low_rank_approx();
  const std::vector<double>& S = std::get<1>(usv);
  // compute_residuals, then the history {timings, traces, svals, rnrms}
  // (src/Skema_SketchySVD.cpp:729-755)
  const std::vector<double>& R = sk.compute_residuals(A, std::get<0>(usv), S, std::get<2>(usv));
  dump_factors(p, &std::get<0>(usv), S, &std::get<2>(usv));
  std::ostringstream js;
  js.precision(17);
  const std::string t = sk.timings_json();
  js << "{" << t.substr(1, t.size() - 2) << ",\"traces\":" << sk.traces_json();
  json_array(js, "svals", S);
  json_array(js, "rnrms", R);
  js << "}";
  write_history(history, js.str());
  return 0;
}

// The SketchySPD route (sketchyspd, src/Skema_SketchySVD.cpp:1243-1277):
// history with its timings, the eigenvalues as svals and the eigenpair
// residuals.
template <class M, class D>
int run_spd(const skema_mi355x::AlgParams& p, const M& A, const std::string& history,
            const skema_mi355x::Context& ctx) {
  skema_mi355x::SketchySPD<M, D> sk(p, ctx);
  sk.linear_update(A);
  if (p.print_level > 0) std::cout << "\nComputing fixed-rank PSD approximation" << std::endl;
  auto us = sk.low_rank_approx();
  const std::vector<double>& S = std::get<1>(us);
  const std::vector<double>& R = sk.compute_residuals(A, std::get<0>(us), S);
  dump_factors(p, &std::get<0>(us), S, nullptr);
  std::ostringstream js;
  js.precision(17);
  const std::string t = sk.timings_json();
  js << "{" << t.substr(1, t.size() - 2) << ",\"traces\":{}";
  json_array(js, "svals", S);
  json_array(js, "rnrms", R);
  js << "}";
  write_history(history, js.str());
  return 0;
}

enum class Arg { Value, Flag };
struct ArgSpec {
  Arg kind;
  bool used;  // read by the sketch route; otherwise accepted and ignored
};

// Every flag of AlgParams::parse (src/Skema_AlgParams.cpp:242-330) and of
// main (driver/main.cpp:156-170), plus this build's --precision / --map-dist.
const std::map<std::string, ArgSpec>& arg_table() {
  static const std::map<std::string, ArgSpec> t = {
      {"--input", {Arg::Value, true}}, {"--output", {Arg::Value, false}},
      {"--debug-file", {Arg::Value, true}}, {"--history-file", {Arg::Value, true}},
      {"--sparse", {Arg::Flag, true}}, {"--dense", {Arg::Flag, true}},
      {"--symmetric", {Arg::Flag, true}}, {"--asymmetric", {Arg::Flag, true}},
      {"--m", {Arg::Value, false}}, {"--n", {Arg::Value, false}},
      {"--rank", {Arg::Value, true}}, {"--num-passes", {Arg::Value, false}},
      {"--solver", {Arg::Value, true}}, {"--print-level", {Arg::Value, true}},
      {"--debug", {Arg::Flag, true}}, {"--debug-off", {Arg::Flag, true}},
      {"--rr", {Arg::Flag, true}}, {"--rr-off", {Arg::Flag, true}},
      {"--normalize-matrix", {Arg::Value, true}}, {"--window", {Arg::Value, true}},
      {"--model", {Arg::Value, true}}, {"--range", {Arg::Value, true}}, {"--core", {Arg::Value, true}},
      {"--seeds", {Arg::Value, true}}, {"--eta", {Arg::Value, true}}, {"--nu", {Arg::Value, true}},
      {"--force-three-sketch", {Arg::Flag, true}}, {"--force-three-sketch-off", {Arg::Flag, true}},
      {"--sketch-compute-svals-iters", {Arg::Flag, true}}, {"--sketch-compute-svals-iters-off", {Arg::Flag, true}},
      {"--svd", {Arg::Flag, false}}, {"--svds", {Arg::Flag, false}},
      {"--isvd-compute-residual-iters", {Arg::Flag, false}}, {"--isvd-compute-residual-iters-off", {Arg::Flag, false}},
      {"--isvd-sampler", {Arg::Value, false}}, {"--isvd-num-samples", {Arg::Value, false}},
      {"--isvd-convtest-eps", {Arg::Value, false}}, {"--isvd-convtest-skip", {Arg::Value, false}},
      {"--isvd-initial-guess", {Arg::Flag, false}}, {"--isvd-initial-guess-off", {Arg::Flag, false}},
      {"--isvd-rank-add-factor", {Arg::Value, false}},
      {"--primme_outputFile", {Arg::Value, false}}, {"--primme_printLevel", {Arg::Value, false}},
      {"--primme_eps", {Arg::Value, false}}, {"--primme_initSize", {Arg::Value, false}},
      {"--primme_maxBasisSize", {Arg::Value, false}}, {"--primme_minRestartSize", {Arg::Value, false}},
      {"--primme_maxBlockSize", {Arg::Value, false}}, {"--primme_maxMatvecs", {Arg::Value, false}},
      {"--primme_maxIters", {Arg::Value, false}}, {"--primme_locking", {Arg::Flag, false}},
      {"--primme_locking-off", {Arg::Flag, false}}, {"--primme_method", {Arg::Value, false}},
      {"--prime_methodStage2", {Arg::Value, false}},
      {"--kernel", {Arg::Value, true}}, {"--gamma", {Arg::Value, true}},
      {"--help", {Arg::Flag, true}}, {"--no-help", {Arg::Flag, true}},
      {"--precision", {Arg::Value, true}}, {"--map-dist", {Arg::Value, true}},
  };
  return t;
}

void usage() {
  std::cout << "Usage: skema_sketch --input A.mtx|A.bin [options]\n"
               "  --sparse | --dense, --symmetric | --asymmetric, --rank r, --range k, --core s,\n"
               "  --window w, --model gauss|sparse-sign, --seeds [a,b,c,d], --eta e, --nu v,\n"
               "  --kernel gaussrbf, --gamma g, --force-three-sketch, --normalize-matrix x,\n"
               "  --sketch-compute-svals-iters, --history-file f, --debug-file f, --debug,\n"
               "  --print-level p, --precision fp64|fp32|bf16x3|bf16|fp8, --map-dist uniform|normal\n";
}

int64_t to_int(const std::string& flag, const std::string& v, int64_t lo) {
  size_t pos = 0;
  int64_t x = 0;
  try {
    x = std::stoll(v, &pos);
  } catch (const std::exception&) {
    fail("Bad input: " + flag + " " + v + ", must be an integer");
  }
  if (pos != v.size() || x < lo) fail("Bad input: " + flag + " " + v + ", must be an integer >= " + std::to_string(lo));
  return x;
}
double to_real(const std::string& flag, const std::string& v) {
  size_t pos = 0;
  double x = 0.0;
  try {
    x = std::stod(v, &pos);
  } catch (const std::exception&) {
    fail("Bad input: " + flag + " " + v + ", must be a number");
  }
  if (pos != v.size() || !(x >= 0.0)) fail("Bad input: " + flag + " " + v + ", must be a number >= 0");
  return x;
}

}  // namespace

int main(int argc, char** argv) {
  using namespace skema_mi355x;
  std::map<std::string, std::string> kv;
  const auto& table = arg_table();
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0 && table.count("--" + a)) a = "--" + a;  // "rank" for "--rank"
    auto it = table.find(a);
    if (it == table.end()) fail("unknown argument " + a);
    std::string v = "1";
    if (it->second.kind == Arg::Value) {
      if (i + 1 >= argc) fail("missing value for " + a);
      v = argv[++i];
    }
    if (!it->second.used) {
      std::cerr << "skema_sketch: " << a << " applies to the isvd / primme solvers; ignored\n";
      continue;
    }
    kv[a] = v;
  }
  if (kv.count("--help")) {
    usage();
    return 0;
  }
  if (!kv.count("--input")) fail("must provide --input");
  if (kv.count("--solver") && kv["--solver"] != "sketch") fail("only --solver sketch is provided (got " + kv["--solver"] + ")");
  if (kv.count("--rr")) fail("--rr (the Rayleigh-Ritz pass) needs PRIMME and is not provided");
  AlgParams p;
  // of an on/off pair the on form wins when both are given, as in
  // parse_bool (src/Skema_AlgParams.cpp:333-360)
  const bool sparse = kv.count("--sparse") > 0;
  p.rank = kv.count("--rank") ? to_int("--rank", kv["--rank"], 1) : 1;
  p.sketch_range = kv.count("--range") ? to_int("--range", kv["--range"], 0) : 1;
  p.sketch_core = kv.count("--core") ? to_int("--core", kv["--core"], 0) : 1;
  p.window = kv.count("--window") ? to_int("--window", kv["--window"], 1) : 1;
  p.sketch_eta = kv.count("--eta") ? to_real("--eta", kv["--eta"]) : 1.0;
  p.sketch_nu = kv.count("--nu") ? to_real("--nu", kv["--nu"]) : 1.0;
  p.print_level = kv.count("--print-level") ? (int)to_int("--print-level", kv["--print-level"], 0) : 0;
  if (p.print_level > 5) fail("Bad input: --print-level must be in [0, 5]");
  p.debug = kv.count("--debug") > 0;
  p.sketch_compute_svals_iters = kv.count("--sketch-compute-svals-iters") > 0;
  if (kv.count("--debug-file")) p.debug_filename = kv["--debug-file"];
  const double normalize = kv.count("--normalize-matrix") ? to_real("--normalize-matrix", kv["--normalize-matrix"]) : 0.0;
  if (kv.count("--seeds")) {
    auto s = parse_seeds(kv["--seeds"]);
    std::copy(s.begin(), s.end(), p.seeds.begin());
  }
  const std::string model = kv.count("--model") ? kv["--model"] : "gauss";
  if (model != "gauss" && model != "sparse-sign") fail("Bad input: --model " + model + ", must be one of: gauss, sparse-sign");
  p.dim_redux = model == "gauss" ? SKEMA_MAP_GAUSS : SKEMA_MAP_SPARSE_SIGN;
  if (kv.count("--map-dist")) {
    const std::string d = kv["--map-dist"];
    if (d != "uniform" && d != "normal") fail("Bad input: --map-dist " + d + ", must be uniform or normal");
    p.gauss_dist = d == "uniform" ? SKEMA_DIST_UNIFORM : SKEMA_DIST_NORMAL;
  }
  const std::map<std::string, int> precs = {{"fp64", SKEMA_PREC_FP64}, {"fp32", SKEMA_PREC_FP32},
                                            {"bf16", SKEMA_PREC_BF16}, {"bf16x3", SKEMA_PREC_BF16X3},
                                            {"fp8", SKEMA_PREC_FP8}};
  if (kv.count("--precision")) {
    if (!precs.count(kv["--precision"])) fail("unknown --precision " + kv["--precision"]);
    p.precision = precs.at(kv["--precision"]);
  }
  if (kv.count("--kernel") && kv["--kernel"] != "gaussrbf" && kv["--kernel"] != "none")
    fail("Bad input: --kernel " + kv["--kernel"] + ", must be one of: none, gaussrbf");
  const bool rbf = kv.count("--kernel") && kv["--kernel"] == "gaussrbf";
  p.kernel_gamma = kv.count("--gamma") ? to_real("--gamma", kv["--gamma"]) : 1.0;
  const std::string in = kv["--input"];
  const bool is_bin = in.size() > 4 && in.substr(in.size() - 4) == ".bin";
  Matrix A = is_bin ? read_bin(in, sparse && !rbf) : read_mtx(in, sparse && !rbf);
  // --normalize-matrix x divides A (for kernel input: the points) by x
  // before sketching (driver/main.cpp:110-114, :139-141)
  if (normalize > 0.0) {
    for (auto& x : A.dense) x /= normalize;
    for (auto& x : A.val) x /= normalize;
  }
  const std::string history = kv.count("--history-file") ? kv["--history-file"] : "";
  // --symmetric selects SketchySPD; kernel input is symmetric unless
  // --force-three-sketch (driver/main.cpp:99-105, src/Skema_SketchySVD.cpp:1249)
  const bool force3 = kv.count("--force-three-sketch") > 0;
  bool symmetric = kv.count("--symmetric") > 0;
  if (rbf) symmetric = !force3;
  const bool spd = symmetric && !force3;
  const bool gauss = p.dim_redux == SKEMA_MAP_GAUSS;
  if (p.print_level > 0) {
    std::cout << "\n===============================================\n==================== Skema ===================="
              << "\n===============================================\nOptions: \n  input = " << in
              << "\n  rank = " << p.rank << "\n  window = " << p.window << "\n  model = " << model
              << "\n  range = " << p.range() << "\n  core = " << p.core() << "\n  eta = " << p.sketch_eta
              << "\n  nu = " << p.sketch_nu << "\n  precision = "
              << (kv.count("--precision") ? kv["--precision"] : "fp64")
              << "\n===============================================" << std::endl;
  }
  try {
    Context ctx(0);
    if (rbf) {
      // rows are points; the sketched matrix is n x n (driver/main.cpp:99-105)
      p.kernel_func = SKEMA_KERNEL_GAUSSRBF;
      p.matrix_m = p.matrix_n = A.m;
      p.nfeat = A.n;
      PointSet P{DeviceMatrix::from_host(A.dense.data(), A.m, A.n, A.m)};
      if (spd) return gauss ? run_spd<PointSet, GaussDimRedux>(p, P, history, ctx)
                            : run_spd<PointSet, SparseSignDimRedux>(p, P, history, ctx);
      if (gauss) return run<PointSet, GaussDimRedux>(p, P, history, ctx);
      return run<PointSet, SparseSignDimRedux>(p, P, history, ctx);
    }
    p.matrix_m = A.m;
    p.matrix_n = A.n;
    p.issparse = A.sparse;
    if (spd && A.m != A.n) fail("--symmetric needs a square matrix");
    if (A.sparse) {
      CrsMatrix C(A.m, A.n, A.rowptr, A.col, A.val);
      if (spd) return gauss ? run_spd<CrsMatrix, GaussDimRedux>(p, C, history, ctx)
                            : run_spd<CrsMatrix, SparseSignDimRedux>(p, C, history, ctx);
      if (gauss) return run<CrsMatrix, GaussDimRedux>(p, C, history, ctx);
      return run<CrsMatrix, SparseSignDimRedux>(p, C, history, ctx);
    }
    DeviceMatrix D = DeviceMatrix::from_host(A.dense.data(), A.m, A.n, A.m);
    if (spd) return gauss ? run_spd<DeviceMatrix, GaussDimRedux>(p, D, history, ctx)
                          : run_spd<DeviceMatrix, SparseSignDimRedux>(p, D, history, ctx);
    if (gauss) return run<DeviceMatrix, GaussDimRedux>(p, D, history, ctx);
    return run<DeviceMatrix, SparseSignDimRedux>(p, D, history, ctx);
  } catch (const Error& e) {
    fail(e.what());
  }
}
