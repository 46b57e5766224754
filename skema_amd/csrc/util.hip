#include "util.hpp"

namespace skema {
namespace {

__global__ void k_fill_f64(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i32(int* p, int64_t n, int v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_copy_matrix(int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    B[i + j * ldb] = A[i + j * lda];
  }
}
__global__ void k_scale_matrix(int64_t rows, int64_t cols, double a, double* A, int64_t lda) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    A[i + j * lda] = (a == 0.0) ? 0.0 : a * A[i + j * lda];
  }
}
__global__ void k_scale_columns(int64_t rows, int64_t cols, const double* A, int64_t lda, const double* s, double* B,
                                int64_t ldb) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, j = e / rows;
    B[i + j * ldb] = A[i + j * lda] * s[j];
  }
}
__global__ __launch_bounds__(256) void k_sumsq_partial(int64_t rows, int64_t cols, const double* A, int64_t lda,
                                                       double* part) {
  __shared__ double red[4];
  const int64_t total = rows * cols;
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const double x = A[(e % rows) + (e / rows) * lda];
    s += x * x;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void k_sum_serial(const double* part, int n, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s;
  }
}
__global__ void k_sum_serial_acc(const double* part, int n, double* acc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *acc += s;
  }
}

unsigned grid_for(int64_t n) {
  int64_t b = cdiv(n, 256);
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

void fill_f64(hipStream_t st, double* p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
  check_launch("k_fill_f64");
}
void fill_i32(hipStream_t st, int* p, int64_t n, int v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
  check_launch("k_fill_i32");
}
void copy_matrix(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_copy_matrix, dim3(grid_for(rows * cols)), dim3(256), 0, st, rows, cols, A, lda, B, ldb);
  check_launch("k_copy_matrix");
}
void scale_matrix(hipStream_t st, int64_t rows, int64_t cols, double a, double* A, int64_t lda) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_scale_matrix, dim3(grid_for(rows * cols)), dim3(256), 0, st, rows, cols, a, A, lda);
  check_launch("k_scale_matrix");
}
void scale_columns(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, const double* s,
                   double* B, int64_t ldb) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_scale_columns, dim3(grid_for(rows * cols)), dim3(256), 0, st, rows, cols, A, lda, s, B, ldb);
  check_launch("k_scale_columns");
}
void sum_squares(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* out,
                 DevBuf* scratch) {
  const int nb = 1024;
  scratch->reserve(nb * sizeof(double));
  if (rows <= 0 || cols <= 0) {
    fill_f64(st, out, 1, 0.0);
    return;
  }
  hipLaunchKernelGGL(k_sumsq_partial, dim3(nb), dim3(256), 0, st, rows, cols, A, lda, scratch->as<double>());
  hipLaunchKernelGGL(k_sum_serial, dim3(1), dim3(64), 0, st, scratch->as<double>(), nb, out);
  check_launch("sum_squares");
}
void sum_squares_acc(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* part,
                     double* acc) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_sumsq_partial, dim3(kSumSqPartials), dim3(256), 0, st, rows, cols, A, lda, part);
  hipLaunchKernelGGL(k_sum_serial_acc, dim3(1), dim3(64), 0, st, part, kSumSqPartials, acc);
  check_launch("sum_squares_acc");
}

}  // namespace skema
