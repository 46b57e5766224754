// Materialising RBF window: 64 x 64 output tiles, features staged through
// LDS in slices of 32, each thread owning a 4 x 4 block of the tile.
#include "rbf.hpp"

namespace skema {
namespace {

constexpr int kT = 64, kF = 32;

__global__ __launch_bounds__(256) void k_rbf_window(const double* X, int64_t n, int64_t d, int64_t ldp, double gamma,
                                                    int64_t row0, int64_t rows, double* K, int64_t ldk) {
  __shared__ double xi[kF][kT + 1];
  __shared__ double xj[kF][kT + 1];
  const int tid = threadIdx.x;
  const int ti = tid & 15, tj = tid >> 4;  // 16 x 16 threads, 4 x 4 each
  const int64_t i0 = (int64_t)blockIdx.x * kT, j0 = (int64_t)blockIdx.y * kT;
  double acc[4][4] = {};
  for (int64_t f0 = 0; f0 < d; f0 += kF) {
    for (int e = tid; e < kT * kF; e += 256) {
      const int r = e % kT, f = e / kT;
      const int64_t gi = i0 + r, gj = j0 + r, gf = f0 + f;
      xi[f][r] = (gi < rows && gf < d) ? X[(row0 + gi) + gf * ldp] : 0.0;
      xj[f][r] = (gj < n && gf < d) ? X[gj + gf * ldp] : 0.0;
    }
    __syncthreads();
    const int fe = (d - f0 < kF) ? (int)(d - f0) : kF;
    for (int f = 0; f < fe; ++f) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = xi[f][ti + 16 * u];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = xj[f][tj + 16 * u];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const double df = a[u] - b[w];
          acc[u][w] += df * df;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int64_t i = i0 + ti + 16 * u, j = j0 + tj + 16 * w;
      if (i < rows && j < n) K[i + j * ldk] = (j == row0 + i) ? 1.0 : exp(-1.0 * gamma * acc[u][w]);
    }
}

__global__ void k_row_norms2(const double* X, int64_t n, int64_t d, int64_t ldp, double* nrm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t f = 0; f < d; ++f) {
    const double x = X[i + f * ldp];
    s += x * x;
  }
  nrm[i] = s;
}

}  // namespace

void rbf_window(hipStream_t st, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, int64_t row0,
                int64_t rows, double* K, int64_t ldk) {
  SKEMA_REQUIRE(n > 0 && d > 0 && rows >= 0 && row0 >= 0 && row0 + rows <= n, kErrShape, "rbf_window: bad range");
  SKEMA_REQUIRE(ldp >= n && ldk >= rows, kErrShape, "rbf_window: leading dimension too small");
  if (rows == 0) return;
  dim3 grid((unsigned)cdiv(rows, kT), (unsigned)cdiv(n, kT));
  hipLaunchKernelGGL(k_rbf_window, grid, dim3(256), 0, st, X, n, d, ldp, gamma, row0, rows, K, ldk);
  check_launch("k_rbf_window");
}

void row_norms2(hipStream_t st, const double* X, int64_t n, int64_t d, int64_t ldp, double* nrm) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_row_norms2, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, X, n, d, ldp, nrm);
  check_launch("k_row_norms2");
}

}  // namespace skema
