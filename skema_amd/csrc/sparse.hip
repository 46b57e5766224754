// Sparse kernels of the sketch path (sparse-sign maps, CSR windows).
// Bandwidth-bound gathers: consecutive lanes own consecutive output rows of a
// column-major result, so stores and the dense-operand loads coalesce, and
// no two threads ever accumulate into the same element.
#include <vector>

#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

unsigned blocks_for(int64_t n, int bs = 256) {
  int64_t b = cdiv(n, bs);
  if (b > (int64_t)1 << 30) b = (int64_t)1 << 30;
  return (unsigned)(b < 1 ? 1 : b);
}

__device__ __forceinline__ int64_t lower_bound_col(const int32_t* ci, int64_t lo, int64_t hi, int64_t key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ci[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void put(const GemmOut& o, int64_t i, int64_t p, double v) {
  if (p < o.nsplit) {
    double* c = o.C1 + i + p * o.ldc1;
    *c = (o.beta1 != 0.0) ? o.alpha1 * v + o.beta1 * *c : o.alpha1 * v;
  } else {
    double* c = o.C2 + i + (p - o.nsplit) * o.ldc2;
    *c = (o.beta2 != 0.0) ? o.alpha2 * v + o.beta2 * *c : o.alpha2 * v;
  }
}

__global__ void k_spmm_csr_rows(const int64_t* rp, const int32_t* ci, const double* cv, int64_t r0, int64_t R,
                                int64_t c0, int64_t c1, bool windowed, double alpha, const double* B, int64_t ldb,
                                int64_t N, double beta, double* C, int64_t ldc) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= R * N) return;
  const int64_t p = g % R, j = g / R;
  int64_t lo = rp[r0 + p], hi = rp[r0 + p + 1];
  if (windowed) {
    const int64_t a = lower_bound_col(ci, lo, hi, c0);
    hi = lower_bound_col(ci, a, hi, c1);
    lo = a;
  }
  const double* b = B + j * ldb - c0;
  double acc = 0.0;
  for (int64_t e = lo; e < hi; ++e) acc += cv[e] * b[ci[e]];
  double* c = C + p + j * ldc;
  *c = (beta != 0.0) ? alpha * acc + beta * *c : alpha * acc;
}

// Split form for short M and long rows of S: block (p, z) sums the z-th
// chunk of row p's nonzeros for all M rows of B (thread = row i), partials
// part[z][i + p M] reduced in a fixed order by k_spmm_split_reduce.
__global__ __launch_bounds__(256) void k_spmm_dense_csrT_split(int64_t M, const double* B, int64_t ldb,
                                                               const int64_t* rp, const int32_t* ci, const double* cv,
                                                               int64_t R, int64_t chunk, double* part) {
  const int64_t p = blockIdx.x, z = blockIdx.y;
  const int64_t lo = rp[p] + z * chunk, hi0 = rp[p + 1];
  const int64_t hi = (lo + chunk < hi0) ? lo + chunk : hi0;
  for (int64_t i = threadIdx.x; i < M; i += 256) {
    double acc = 0.0;
    for (int64_t e = lo; e < hi; ++e) acc += cv[e] * B[i + (int64_t)ci[e] * ldb];
    part[z * M * R + i + p * M] = acc;
  }
}
__global__ void k_spmm_split_reduce(int64_t M, int64_t R, int splits, const double* part, GemmOut o) {
  const int64_t total = M * R;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int z = 0; z < splits; ++z) v += part[(int64_t)z * total + e];
    put(o, e % M, e / M, v);
  }
}

__global__ void k_spmm_dense_csrT(int64_t M, const double* B, int64_t ldb, const int64_t* rp, const int32_t* ci,
                                  const double* cv, int64_t R, GemmOut o) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= M * R) return;
  const int64_t i = g % M, p = g / M;
  double acc = 0.0;
  for (int64_t e = rp[p]; e < rp[p + 1]; ++e) acc += cv[e] * B[i + (int64_t)ci[e] * ldb];
  put(o, i, p, acc);
}

// ---- CSR window -> CSC ----------------------------------------------------
// flag bit 0: rowptr decreasing, bit 1: column index outside [0, cols).
__global__ void k_validate_csr(const int64_t* rowptr, int64_t rows, const int32_t* col, int64_t cols, int* flag) {
  const int64_t base = rowptr[0];
  const int64_t nnz = rowptr[rows] - base;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x)
    if (rowptr[i + 1] < rowptr[i]) atomicOr(flag, 1);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = col[base + e];
    if (c < 0 || c >= cols) atomicOr(flag, 2);
  }
}

__global__ void k_count_cols(const int64_t* rowptr, int64_t rows, const int32_t* col, int* counts) {
  const int64_t base = rowptr[0];
  const int64_t nnz = rowptr[rows] - base;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&counts[col[base + e]], 1);
}

// Column counts per row block in LDS (cols <= kHistCols, the histogram sized
// to cols in dynamic LDS): block b takes rows [b R, (b+1) R), walks its
// nonzeros in storage order (coalesced), counts them per column in LDS and
// writes the histogram to T[b][0..cols).  No global atomics.
constexpr int kHistCols = 32768;
__global__ __launch_bounds__(1024) void k_count_cols_blocked(const int64_t* rowptr, int64_t rows, const int32_t* col,
                                                            int64_t cols, int64_t R, int* T) {
  extern __shared__ int hist[];
  const int64_t b = blockIdx.x;
  for (int64_t c = threadIdx.x; c < cols; c += 1024) hist[c] = 0;
  __syncthreads();
  const int64_t r0 = b * R, r1 = (r0 + R < rows) ? r0 + R : rows;
  if (r0 < r1) {
    const int64_t e1 = rowptr[r1];
    for (int64_t e = rowptr[r0] + threadIdx.x; e < e1; e += 1024) atomicAdd(&hist[col[e]], 1);
  }
  __syncthreads();
  for (int64_t c = threadIdx.x; c < cols; c += 1024) T[b * cols + c] = hist[c];
}
// Per column: T[b][c] <- sum_{b' < b} T[b'][c] (exclusive, in block order),
// counts[c] <- the column total.
// Sixteen row blocks' counts are loaded before any is overwritten, so the
// loads of a batch are in flight together.
__global__ void k_scan_blocks(int64_t nb, int64_t cols, int* T, int* counts) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  constexpr int kB = 16;
  int run = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kB) {
    int t[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j) t[j] = (b0 + j < nb) ? T[(b0 + j) * cols + c] : 0;
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      if (b0 + j < nb) T[(b0 + j) * cols + c] = run;
      run += t[j];
    }
  }
  counts[c] = run;
}
// Scatter of block b's rows into its own sub-range of every column:
// cursor[c] = ptr[c] + T[b][c] in LDS, one LDS atomic per nonzero.  The
// block's nonzeros are walked in storage order (coalesced reads); the row of
// nonzero e is found by binary search in the block's slice of rowptr, which
// is staged in LDS after the cursors (R + 1 entries, relative to the slice
// start).
__global__ __launch_bounds__(1024) void k_fill_csc_blocked(const int64_t* rowptr, int64_t rows, const int32_t* col,
                                                          const double* val, int64_t cols, int64_t R, const int* T,
                                                          const int64_t* ptr, int32_t* orow, double* oval) {
  extern __shared__ int lds[];
  int* cur = lds;
  int* rp = lds + cols;  // rowptr[r0 .. r1] - rowptr[r0]
  const int64_t b = blockIdx.x;
  const int64_t r0 = b * R, r1 = (r0 + R < rows) ? r0 + R : rows;
  if (r0 >= r1) return;  // whole block: no barrier is skipped by part of it
  const int64_t e0 = rowptr[r0];
  const int nr = (int)(r1 - r0);
  for (int64_t c = threadIdx.x; c < cols; c += 1024) cur[c] = (int)ptr[c] + T[b * cols + c];
  for (int i = threadIdx.x; i <= nr; i += 1024) rp[i] = (int)(rowptr[r0 + i] - e0);
  __syncthreads();
  const int ne = rp[nr];
  for (int e = threadIdx.x; e < ne; e += 1024) {
    // last row start <= e
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const int32_t c = col[e0 + e];
    const int pos = atomicAdd(&cur[c], 1);
    orow[pos] = (int32_t)(r0 + lo);
    oval[pos] = val[e0 + e];
  }
}

// Exclusive scan of counts[0..n) into ptr[0..n] by one block of 1024 threads.
// Tiles of 8192 counts: coalesced loads into LDS, each thread sums its 8
// consecutive entries, a block scan of the 1024 sums, coalesced writes of the
// exclusive prefixes, and the tile total carried to the next tile.
__global__ __launch_bounds__(1024) void k_scan_counts(const int* counts, int64_t n, int64_t* ptr) {
  constexpr int kPer = 8, kTile = 1024 * kPer;
  __shared__ int64_t part[1024];
  __shared__ int64_t tile[kTile];
  const int t = threadIdx.x;
  int64_t carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += kTile) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t i = t0 + j * 1024 + t;
      tile[j * 1024 + t] = i < n ? counts[i] : 0;
    }
    __syncthreads();
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) s += tile[t * kPer + j];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t v = (t >= off) ? part[t - off] : 0;
      __syncthreads();
      part[t] += v;
      __syncthreads();
    }
    int64_t run = carry + part[t] - s;  // exclusive prefix of this thread's 8 entries
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t c = tile[t * kPer + j];
      tile[t * kPer + j] = run;
      run += c;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t i = t0 + j * 1024 + t;
      if (i < n) ptr[i] = tile[j * 1024 + t];
    }
    carry += part[1023];
    __syncthreads();
  }
  if (t == 0) ptr[n] = carry;
}

__global__ void k_fill_csc(const int64_t* rowptr, int64_t rows, const int32_t* col, const double* val, int* cursor,
                           const int64_t* ptr, int32_t* orow, double* oval) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
    const int32_t c = col[e];
    const int64_t pos = ptr[c] + atomicAdd(&cursor[c], 1);
    orow[pos] = (int32_t)i;
    oval[pos] = val[e];
  }
}

// Sort of each column segment by (row, value): makes the order of summation
// independent of the atomic placement above.  One 64-thread block per column:
// a bitonic sort in LDS for up to 1024 entries, insertion sort by one thread
// above that.
constexpr int kSortMax = 1024;
__global__ __launch_bounds__(64) void k_sort_columns(const int64_t* ptr, int32_t* row, double* val) {
  __shared__ int32_t kr[kSortMax];
  __shared__ double kv[kSortMax];
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x;
  const int64_t lo = ptr[c], hi = ptr[c + 1], len = hi - lo;
  if (len <= 1) return;
  if (len > kSortMax) {
    if (tid != 0) return;
    for (int64_t a = lo + 1; a < hi; ++a) {
      const int32_t r = row[a];
      const double v = val[a];
      int64_t b = a - 1;
      while (b >= lo && (row[b] > r || (row[b] == r && val[b] > v))) {
        row[b + 1] = row[b];
        val[b + 1] = val[b];
        --b;
      }
      row[b + 1] = r;
      val[b + 1] = v;
    }
    return;
  }
  int n = 1;
  while (n < len) n <<= 1;
  for (int t = tid; t < n; t += 64) {
    const bool in = t < len;
    kr[t] = in ? row[lo + t] : INT32_MAX;
    kv[t] = in ? val[lo + t] : INFINITY;
  }
  __syncthreads();
  for (int k = 2; k <= n; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = tid; t < n; t += 64) {
        const int u = t ^ jj;
        if (u > t) {
          const bool up = (t & k) == 0;
          const bool gt = kr[t] > kr[u] || (kr[t] == kr[u] && kv[t] > kv[u]);
          if (gt == up) {
            const int32_t r = kr[t];
            kr[t] = kr[u];
            kr[u] = r;
            const double v = kv[t];
            kv[t] = kv[u];
            kv[u] = v;
          }
        }
      }
      __syncthreads();
    }
  for (int t = tid; t < len; t += 64) {
    row[lo + t] = kr[t];
    val[lo + t] = kv[t];
  }
}

__global__ void k_csc_left_dense(const int64_t* ptr, const int32_t* row, const double* val, int64_t cols,
                                 const double* M, int64_t ldm, int64_t row0, int64_t k, double eta, double nu,
                                 double* X, int64_t ldx) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= k * cols) return;
  const int64_t p = g % k, j = g / k;
  double acc = 0.0;
  for (int64_t e = ptr[j]; e < ptr[j + 1]; ++e) acc += val[e] * M[p + (row0 + row[e]) * ldm];
  double* x = X + p + j * ldx;
  *x = (nu != 0.0) ? nu * *x + eta * acc : eta * acc;
}

__global__ void k_csc_left_sparse(const int64_t* ptr, const int32_t* row, const double* val, int64_t cols,
                                  const int64_t* mrp, const int32_t* mci, const double* mv, int64_t row0, int64_t k,
                                  double eta, double nu, double* X, int64_t ldx) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double* x = X + j * ldx;
  if (nu != 1.0)
    for (int64_t p = 0; p < k; ++p) x[p] = (nu != 0.0) ? nu * x[p] : 0.0;
  for (int64_t e = ptr[j]; e < ptr[j + 1]; ++e) {
    const double a = eta * val[e];
    const int64_t c = row0 + row[e];
    for (int64_t f = mrp[c]; f < mrp[c + 1]; ++f) x[mci[f]] += a * mv[f];
  }
}

__global__ void k_csr_right_dense(const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                                  const double* M, int64_t ldm, int64_t kk, GemmOut o) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= kk * rows) return;
  const int64_t p = g % kk, i = g / kk;
  double acc = 0.0;
  for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += val[e] * M[p + (int64_t)col[e] * ldm];
  put(o, i, p, acc);
}

__global__ void k_csr_right_sparse(const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                                   const int64_t* mrp, const int32_t* mci, const double* mv, int64_t kk, GemmOut o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  // beta pass
  for (int64_t p = 0; p < kk; ++p) {
    if (p < o.nsplit) {
      double* c = o.C1 + i + p * o.ldc1;
      *c = (o.beta1 != 0.0) ? o.beta1 * *c : 0.0;
    } else {
      double* c = o.C2 + i + (p - o.nsplit) * o.ldc2;
      *c = (o.beta2 != 0.0) ? o.beta2 * *c : 0.0;
    }
  }
  for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
    const double a = val[e];
    const int32_t j = col[e];
    for (int64_t f = mrp[j]; f < mrp[j + 1]; ++f) {
      const int64_t p = mci[f];
      if (p < o.nsplit) o.C1[i + p * o.ldc1] += o.alpha1 * (a * mv[f]);
      else o.C2[i + (p - o.nsplit) * o.ldc2] += o.alpha2 * (a * mv[f]);
    }
  }
}

}  // namespace

void spmm_csr_rows(hipStream_t st, const Csr& S, int64_t r0, int64_t r1, int64_t c0, int64_t c1, double alpha,
                   const double* B, int64_t ldb, int64_t N, double beta, double* C, int64_t ldc) {
  const int64_t R = r1 - r0;
  SKEMA_REQUIRE(r0 >= 0 && r1 <= S.rows && R >= 0 && c0 >= 0 && c1 <= S.cols && c0 <= c1, kErrShape,
                "spmm_csr_rows: range outside the map");
  if (R == 0 || N == 0) return;
  const bool windowed = !(c0 == 0 && c1 == S.cols);
  hipLaunchKernelGGL(k_spmm_csr_rows, dim3(blocks_for(R * N)), dim3(256), 0, st, S.rp(), S.ci(), S.v(), r0, R, c0,
                     c1, windowed, alpha, B, ldb, N, beta, C, ldc);
  check_launch("k_spmm_csr_rows");
}

void spmm_dense_csrT(hipStream_t st, int64_t M, const double* B, int64_t ldb, const Csr& S, const GemmOut& out,
                     DevBuf* ws) {
  if (M == 0 || S.rows == 0) return;
  // few output elements and long rows of S: split each row's nonzeros so the
  // grid covers the chip (partials then a fixed-order sum)
  const int64_t avg = S.nnz / S.rows;
  if (ws && M <= 1024 && M * S.rows < 256 * 1024 && avg >= 64) {
    const int64_t splits = std::min<int64_t>(64, std::max<int64_t>(1, avg / 32));
    // the longest row bounds the chunk (read once per CSR and kept)
    if (S.max_row < 0) {
      std::vector<int64_t> h((size_t)S.rows + 1);
      SKEMA_HIP(hipMemcpyAsync(h.data(), S.rp(), h.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      SKEMA_HIP(hipStreamSynchronize(st));
      int64_t mr = 0;
      for (int64_t r = 0; r < S.rows; ++r) mr = std::max(mr, h[(size_t)r + 1] - h[(size_t)r]);
      S.max_row = mr;
    }
    const int64_t maxrow = S.max_row;
    const int64_t chunk = std::max<int64_t>(1, cdiv(maxrow, splits));
    const int64_t zs = cdiv(maxrow, chunk);
    ws->reserve((size_t)zs * M * S.rows * sizeof(double));
    double* part = ws->as<double>();
    hipLaunchKernelGGL(k_spmm_dense_csrT_split, dim3((unsigned)S.rows, (unsigned)zs), dim3(256), 0, st, M, B, ldb,
                       S.rp(), S.ci(), S.v(), S.rows, chunk, part);
    check_launch("k_spmm_dense_csrT_split");
    const unsigned nb = (unsigned)std::min<int64_t>(cdiv(M * S.rows, 256), 4096);
    hipLaunchKernelGGL(k_spmm_split_reduce, dim3(nb), dim3(256), 0, st, M, S.rows, (int)zs, part, out);
    check_launch("k_spmm_split_reduce");
    return;
  }
  hipLaunchKernelGGL(k_spmm_dense_csrT, dim3(blocks_for(M * S.rows)), dim3(256), 0, st, M, B, ldb, S.rp(), S.ci(),
                     S.v(), S.rows, out);
  check_launch("k_spmm_dense_csrT");
}

int64_t validate_csr_window(hipStream_t st, const int64_t* rowptr, const int32_t* col, int64_t rows, int64_t cols,
                            int* flag, const int64_t* known) {
  int64_t ends[2] = {0, 0};
  if (known) {
    ends[0] = known[0];
    ends[1] = known[1];
  } else {
    SKEMA_HIP(hipMemcpyAsync(&ends[0], rowptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipMemcpyAsync(&ends[1], rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
  }
  const int64_t nnz = ends[1] - ends[0];
  SKEMA_REQUIRE(nnz >= 0 && nnz < ((int64_t)1 << 31), kErrArg, "CSR window: bad rowptr range");
  // Reject malformed windows before any kernel indexes with them.
  fill_i32(st, flag, 1, 0);
  const int64_t work = nnz > rows ? nnz : rows;
  const unsigned nb = blocks_for(work) > 65536 ? 65536 : blocks_for(work);
  hipLaunchKernelGGL(k_validate_csr, dim3(nb), dim3(256), 0, st, rowptr, rows, col, cols, flag);
  check_launch("k_validate_csr");
  int h = 0;
  SKEMA_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  SKEMA_REQUIRE(!(h & 1), kErrArg, "CSR window: rowptr is not non-decreasing");
  SKEMA_REQUIRE(!(h & 2), kErrArg, "CSR window: column index outside [0, n)");
  return nnz;
}

CscWindow csr_window_to_csc(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val,
                            int64_t rows, int64_t cols, DevBuf& scratch, bool sorted) {
  int64_t nnz = 0;
  {
    DevBuf flag(sizeof(int));
    nnz = validate_csr_window(st, rowptr, col, rows, cols, flag.as<int>());
  }
  const size_t o_cnt = 0;
  const size_t o_ptr = align_up((size_t)cols * sizeof(int) * 2, 256);
  const size_t o_row = o_ptr + align_up((size_t)(cols + 1) * sizeof(int64_t), 256);
  const size_t o_val = o_row + align_up((size_t)nnz * sizeof(int32_t), 256);
  // Row-blocked counts and placement in LDS when the columns fit (512 row
  // blocks; no global atomics); global atomic counters otherwise.
  const int64_t nbk = 512, R = cdiv(rows, nbk);
  // (the fill keeps cols cursors and the R + 1 row starts of its slice in LDS)
  const bool blocked =
      cols <= kHistCols && rows >= 4096 && nnz > 0 && (size_t)(cols + R + 1) * sizeof(int) <= 160 * 1024;
  const size_t o_T = o_val + align_up((size_t)nnz * sizeof(double), 256);
  const size_t total = o_T + (blocked ? align_up((size_t)nbk * cols * sizeof(int), 256) : 0);
  scratch.reserve(total);
  char* base = scratch.as<char>();
  int* counts = reinterpret_cast<int*>(base + o_cnt);
  int* cursor = counts + cols;
  CscWindow w;
  w.rows = rows;
  w.cols = cols;
  w.nnz = nnz;
  w.ptr = reinterpret_cast<int64_t*>(base + o_ptr);
  w.row = reinterpret_cast<int32_t*>(base + o_row);
  w.val = reinterpret_cast<double*>(base + o_val);
  int* T = blocked ? reinterpret_cast<int*>(base + o_T) : nullptr;
  if (blocked) {
    hipLaunchKernelGGL(k_count_cols_blocked, dim3((unsigned)nbk), dim3(1024), (size_t)cols * sizeof(int), st, rowptr,
                       rows, col, cols, R, T);
    hipLaunchKernelGGL(k_scan_blocks, dim3(blocks_for(cols)), dim3(256), 0, st, nbk, cols, T, counts);
  } else {
    fill_i32(st, counts, 2 * cols, 0);
    if (nnz > 0) {
      const unsigned nb = blocks_for(nnz) > 65536 ? 65536 : blocks_for(nnz);
      hipLaunchKernelGGL(k_count_cols, dim3(nb), dim3(256), 0, st, rowptr, rows, col, counts);
    }
  }
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, cols, w.ptr);
  if (nnz > 0) {
    if (blocked)
      hipLaunchKernelGGL(k_fill_csc_blocked, dim3((unsigned)nbk), dim3(1024), (size_t)(cols + R + 1) * sizeof(int), st,
                         rowptr, rows, col, val, cols, R, T, w.ptr, w.row, w.val);
    else
      hipLaunchKernelGGL(k_fill_csc, dim3(blocks_for(rows)), dim3(256), 0, st, rowptr, rows, col, val, cursor, w.ptr,
                         w.row, w.val);
    if (sorted) hipLaunchKernelGGL(k_sort_columns, dim3((unsigned)cols), dim3(64), 0, st, w.ptr, w.row, w.val);
  }
  check_launch("csr_window_to_csc");
  return w;
}

void csc_left_dense(hipStream_t st, const CscWindow& H, const double* M, int64_t ldm, int64_t row0, int64_t k,
                    double eta, double nu, double* X, int64_t ldx) {
  if (k == 0 || H.cols == 0) return;
  hipLaunchKernelGGL(k_csc_left_dense, dim3(blocks_for(k * H.cols)), dim3(256), 0, st, H.ptr, H.row, H.val, H.cols,
                     M, ldm, row0, k, eta, nu, X, ldx);
  check_launch("k_csc_left_dense");
}

void csc_left_sparse(hipStream_t st, const CscWindow& H, const Csr& Mt, int64_t row0, int64_t k, double eta,
                     double nu, double* X, int64_t ldx) {
  if (k == 0 || H.cols == 0) return;
  hipLaunchKernelGGL(k_csc_left_sparse, dim3(blocks_for(H.cols)), dim3(256), 0, st, H.ptr, H.row, H.val, H.cols,
                     Mt.rp(), Mt.ci(), Mt.v(), row0, k, eta, nu, X, ldx);
  check_launch("k_csc_left_sparse");
}

void csr_right_dense(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                     const double* M, int64_t ldm, int64_t kk, const GemmOut& out) {
  if (rows == 0 || kk == 0) return;
  hipLaunchKernelGGL(k_csr_right_dense, dim3(blocks_for(kk * rows)), dim3(256), 0, st, rowptr, col, val, rows, M,
                     ldm, kk, out);
  check_launch("k_csr_right_dense");
}

void csr_right_sparse(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                      const Csr& Mt, int64_t kk, const GemmOut& out) {
  if (rows == 0 || kk == 0) return;
  hipLaunchKernelGGL(k_csr_right_sparse, dim3(blocks_for(rows)), dim3(256), 0, st, rowptr, col, val, rows, Mt.rp(),
                     Mt.ci(), Mt.v(), kk, out);
  check_launch("k_csr_right_sparse");
}

}  // namespace skema
