// Command-line driver for the sketch route of Skema's `skema` binary
// (driver/main.cpp:149-224 -> sketchysvd, src/Skema_SketchySVD.cpp:1243-1577),
// on one MI355X through the C++ layer in include/skema_mi355x.hpp.
//
//   skema_sketch --input A.mtx|A.bin [--sparse] --rank r [--range k] [--core s]
//                [--window w] [--model gauss|sparse-sign] [--seeds a,b,c,d]
//                [--eta e] [--nu v] [--kernel gaussrbf --gamma g]
//                [--precision fp64|fp32|bf16|bf16x3|fp8] [--history-file h.json]
//                [--symmetric | --asymmetric] [--force-three-sketch]
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

template <class M, class D>
int run(const skema_mi355x::AlgParams& p, const M& A, const std::string& history, const skema_mi355x::Context& ctx) {
  skema_mi355x::SketchySVD<M, D> sk(p, ctx);
  sk.linear_update(A);
  auto usv = sk.low_rank_approx();
  const std::vector<double>& S = std::get<1>(usv);
  // compute_residuals, then the history's svals and rnrms (src/Skema_SketchySVD.cpp:729-755)
  const std::vector<double>& R = sk.compute_residuals(A, std::get<0>(usv), S, std::get<2>(usv));
  std::ostringstream js;
  js.precision(17);
  std::string t = sk.timings_json();
  js << "{" << t.substr(1, t.size() - 2) << ",\"svals\":[";
  for (size_t i = 0; i < S.size(); ++i) js << (i ? "," : "") << S[i];
  js << "],\"rnrms\":[";
  for (size_t i = 0; i < R.size(); ++i) js << (i ? "," : "") << R[i];
  js << "]}";
  if (history.empty()) {
    std::cout << js.str() << std::endl;
  } else {
    std::ofstream o(history);
    o << js.str() << std::endl;
  }
  return 0;
}

// The SketchySPD route (sketchyspd, src/Skema_SketchySVD.cpp:1243-1277):
// history with the eigenvalues as svals and the eigenpair residuals.
template <class M, class D>
int run_spd(const skema_mi355x::AlgParams& p, const M& A, const std::string& history,
            const skema_mi355x::Context& ctx) {
  skema_mi355x::SketchySPD<M, D> sk(p, ctx);
  sk.linear_update(A);
  auto us = sk.low_rank_approx();
  const std::vector<double>& S = std::get<1>(us);
  const std::vector<double>& R = sk.compute_residuals(A, std::get<0>(us), S);
  std::ostringstream js;
  js.precision(17);
  js << "{\"svals\":[";
  for (size_t i = 0; i < S.size(); ++i) js << (i ? "," : "") << S[i];
  js << "],\"rnrms\":[";
  for (size_t i = 0; i < R.size(); ++i) js << (i ? "," : "") << R[i];
  js << "]}";
  if (history.empty()) {
    std::cout << js.str() << std::endl;
  } else {
    std::ofstream o(history);
    o << js.str() << std::endl;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  using namespace skema_mi355x;
  std::map<std::string, std::string> kv;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) fail("unexpected argument " + a);
    const bool flag = (a == "--sparse" || a == "--dense" || a == "--force-three-sketch" || a == "--symmetric" ||
                       a == "--asymmetric");
    if (flag) kv[a] = "1";
    else if (i + 1 < argc) kv[a] = argv[++i];
    else fail("missing value for " + a);
  }
  if (!kv.count("--input")) fail("must provide --input");
  if (kv.count("--solver") && kv["--solver"] != "sketch") fail("only --solver sketch is provided");
  AlgParams p;
  const bool sparse = kv.count("--sparse") > 0;
  p.rank = kv.count("--rank") ? std::stoll(kv["--rank"]) : 1;
  p.sketch_range = kv.count("--range") ? std::stoll(kv["--range"]) : 1;
  p.sketch_core = kv.count("--core") ? std::stoll(kv["--core"]) : 1;
  p.window = kv.count("--window") ? std::stoll(kv["--window"]) : 1;
  p.sketch_eta = kv.count("--eta") ? std::stod(kv["--eta"]) : 1.0;
  p.sketch_nu = kv.count("--nu") ? std::stod(kv["--nu"]) : 1.0;
  if (kv.count("--seeds")) {
    auto s = parse_seeds(kv["--seeds"]);
    std::copy(s.begin(), s.end(), p.seeds.begin());
  }
  const std::string model = kv.count("--model") ? kv["--model"] : "gauss";
  if (model != "gauss" && model != "sparse-sign") fail("--model must be gauss or sparse-sign");
  p.dim_redux = model == "gauss" ? SKEMA_MAP_GAUSS : SKEMA_MAP_SPARSE_SIGN;
  const std::map<std::string, int> precs = {{"fp64", SKEMA_PREC_FP64}, {"fp32", SKEMA_PREC_FP32},
                                            {"bf16", SKEMA_PREC_BF16}, {"bf16x3", SKEMA_PREC_BF16X3},
                                            {"fp8", SKEMA_PREC_FP8}};
  if (kv.count("--precision")) {
    if (!precs.count(kv["--precision"])) fail("unknown --precision");
    p.precision = precs.at(kv["--precision"]);
  }
  const bool rbf = kv.count("--kernel") && kv["--kernel"] == "gaussrbf";
  p.kernel_gamma = kv.count("--gamma") ? std::stod(kv["--gamma"]) : 1.0;
  const std::string in = kv["--input"];
  const bool is_bin = in.size() > 4 && in.substr(in.size() - 4) == ".bin";
  Matrix A = is_bin ? read_bin(in, sparse && !rbf) : read_mtx(in, sparse && !rbf);
  const std::string history = kv.count("--history-file") ? kv["--history-file"] : "";
  // --symmetric selects SketchySPD; kernel input is symmetric unless
  // --force-three-sketch (driver/main.cpp:99-105, src/Skema_SketchySVD.cpp:1249)
  const bool force3 = kv.count("--force-three-sketch") > 0;
  bool symmetric = kv.count("--symmetric") > 0 && !kv.count("--asymmetric");
  if (rbf) symmetric = !force3;
  const bool spd = symmetric && !force3;
  const bool gauss = p.dim_redux == SKEMA_MAP_GAUSS;
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
