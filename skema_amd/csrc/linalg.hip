// FP64 finalize kernels: Householder QR (dgeqr2/dorg2r structure, which is
// what the reference's dgeqrf/dorgqr calls with lwork = n execute,
// src/Skema_BlasLapack.hpp:168-217), upper-triangular solve (dgels on an
// already-triangular matrix, :219-243) and a one-sided Jacobi SVD standing in
// for dgesvd (:71-97).
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kQrThreads = 512;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Deterministic block sum; red must hold blockDim.x/64 doubles.
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// dlarfg on column `col` of length len (col[0] = alpha, col[1..] = x).
// The thread->element assignment (i = tid + k*bs) is the one every caller
// uses for its own writes, so no cross-thread global visibility is needed.
__device__ void make_reflector(double* col, int64_t len, double* tau, double* red, double* bcast) {
  const int tid = threadIdx.x, bs = blockDim.x;
  double ss = 0.0;
  for (int64_t i = 1 + tid; i < len; i += bs) ss += col[i] * col[i];
  const double xnorm2 = block_sum(ss, red);
  if (tid == 0) bcast[0] = col[0];
  __syncthreads();
  const double alpha = bcast[0];
  if (xnorm2 == 0.0) {
    if (tid == 0) *tau = 0.0;
    return;
  }
  const double beta = -copysign(hypot(alpha, sqrt(xnorm2)), alpha);
  const double scal = 1.0 / (alpha - beta);
  for (int64_t i = 1 + tid; i < len; i += bs) col[i] *= scal;
  if (tid == 0) {
    *tau = (beta - alpha) / beta;
    col[0] = beta;
  }
}

__global__ __launch_bounds__(kQrThreads) void k_qr_first(double* A, int64_t m, double* tau) {
  __shared__ double red[kQrThreads / 64];
  __shared__ double bcast[1];
  make_reflector(A, m, tau, red, bcast);
}

// Step j: apply H_j to columns c = j+1+blockIdx.x (rows j..m-1); the block that
// owns column j+1 then forms the next reflector from the values it wrote.
__global__ __launch_bounds__(kQrThreads) void k_qr_step(double* A, int64_t lda, int64_t m, int64_t n,
                                                        int64_t j, double* tau) {
  __shared__ double red[kQrThreads / 64];
  __shared__ double bcast[1];
  const int tid = threadIdx.x, bs = blockDim.x;
  const int64_t c = j + 1 + blockIdx.x;
  const double* v = A + j + j * lda;  // v[0] = 1 implicitly
  double* col = A + j + c * lda;
  const int64_t len = m - j;
  const double tj = tau[j];
  if (tj != 0.0) {
    double d = 0.0;
    for (int64_t i = tid; i < len; i += bs) d += (i == 0 ? 1.0 : v[i]) * col[i];
    const double w = block_sum(d, red) * tj;
    for (int64_t i = tid; i < len; i += bs) col[i] -= (i == 0 ? 1.0 : v[i]) * w;
  }
  if (c == j + 1) {
    // Reflector for column j+1 on rows j+1..m-1: element i of `col` is row j+i,
    // so the sub-column starts at col+1.  Keep the same thread mapping: thread
    // tid wrote rows j+tid+k*bs; make_reflector reads col+1 at offsets
    // 1+tid+k*bs -> rows j+2+tid+k*bs, a different assignment, so fence.
    __syncthreads();
    __threadfence_block();
    make_reflector(col + 1, len - 1, tau + j + 1, red, bcast);
  }
}

// dorg2r step j: column j of Q from reflector j, and H_j applied to the
// columns c > j already formed.
__global__ __launch_bounds__(kQrThreads) void k_org_step(const double* A, int64_t lda, const double* tau,
                                                         double* Q, int64_t ldq, int64_t m, int64_t j) {
  __shared__ double red[kQrThreads / 64];
  const int tid = threadIdx.x, bs = blockDim.x;
  const int64_t c = j + blockIdx.x;
  const double* v = A + j * lda;  // full column; rows > j hold v
  const double tj = tau[j];
  double* q = Q + c * ldq;
  if (c == j) {
    for (int64_t i = tid; i < m; i += bs)
      q[i] = (i < j) ? 0.0 : (i == j ? 1.0 - tj : -tj * v[i]);
    return;
  }
  if (tj == 0.0) return;
  double d = 0.0;
  for (int64_t i = j + tid; i < m; i += bs) d += (i == j ? 1.0 : v[i]) * q[i];
  const double w = block_sum(d, red) * tj;
  for (int64_t i = j + tid; i < m; i += bs) q[i] -= (i == j ? 1.0 : v[i]) * w;
}

__global__ void k_extract_r(int64_t n, const double* A, int64_t lda, double* R, int64_t ldr) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, c = e / n;
  R[i + c * ldr] = (i <= c) ? A[i + c * lda] : 0.0;
}

// dtrsm Left/Upper/NoTrans/NonUnit column loop; a zero on the diagonal leaves
// B untouched (dgels returns early with info > 0 after dtrtrs' check).
__global__ __launch_bounds__(256) void k_trsm_upper(int64_t n, const double* T, int64_t ldt, double* B,
                                                    int64_t ldb) {
  extern __shared__ double b[];
  __shared__ int singular;
  const int tid = threadIdx.x, bs = blockDim.x;
  double* bj = B + (int64_t)blockIdx.x * ldb;
  if (tid == 0) singular = 0;
  __syncthreads();
  for (int64_t i = tid; i < n; i += bs)
    if (T[i + i * ldt] == 0.0) singular = 1;
  __syncthreads();
  if (singular) return;
  for (int64_t i = tid; i < n; i += bs) b[i] = bj[i];
  __syncthreads();
  for (int64_t k = n - 1; k >= 0; --k) {
    double bk = b[k];
    if (bk != 0.0) {
      bk = bk / T[k + k * ldt];
      for (int64_t i = tid; i < k; i += bs) b[i] -= bk * T[i + k * ldt];
    }
    __syncthreads();
    if (tid == 0) b[k] = bk;
  }
  __syncthreads();
  for (int64_t i = tid; i < n; i += bs) bj[i] = b[i];
}

__global__ void k_transpose(int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb) {
  __shared__ double tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + tx, c = c0 + k;
    if (r < rows && c < cols) tile[k][tx] = A[r + c * lda];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t c = c0 + tx, r = r0 + k;  // B(c, r) = A(r, c)
    if (r < rows && c < cols) B[c + r * ldb] = tile[tx][k];
  }
}

// One Jacobi round: every pair (p, q) of the round is orthogonalised by its
// own block (pairs of a round are disjoint).
__global__ __launch_bounds__(256) void k_jacobi_round(double* A, int64_t lda, double* V, int64_t ldv, int64_t n,
                                                      const int2* pairs, double tol, int* rotations) {
  __shared__ double red[3][4];
  const int tid = threadIdx.x;
  const int2 pq = pairs[blockIdx.x];
  double* ap = A + pq.x * lda;
  double* aq = A + pq.y * lda;
  double sa = 0.0, sb = 0.0, sg = 0.0;
  for (int64_t i = tid; i < n; i += 256) {
    const double x = ap[i], y = aq[i];
    sa += x * x;
    sb += y * y;
    sg += x * y;
  }
  sa = wave_sum(sa);
  sb = wave_sum(sb);
  sg = wave_sum(sg);
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0) { red[0][w] = sa; red[1][w] = sb; red[2][w] = sg; }
  __syncthreads();
  const double alpha = red[0][0] + red[0][1] + red[0][2] + red[0][3];
  const double beta = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  const double gamma = red[2][0] + red[2][1] + red[2][2] + red[2][3];
  if (!(fabs(gamma) > tol * sqrt(alpha * beta))) return;
  const double zeta = (beta - alpha) / (2.0 * gamma);
  const double az = fabs(zeta);
  const double t = (az > 1e150) ? 0.5 / zeta : copysign(1.0, zeta) / (az + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  for (int64_t i = tid; i < n; i += 256) {
    const double x = ap[i], y = aq[i];
    ap[i] = c * x - s * y;
    aq[i] = s * x + c * y;
  }
  double* vp = V + pq.x * ldv;
  double* vq = V + pq.y * ldv;
  for (int64_t i = tid; i < n; i += 256) {
    const double x = vp[i], y = vq[i];
    vp[i] = c * x - s * y;
    vq[i] = s * x + c * y;
  }
  if (tid == 0) atomicAdd(rotations, 1);
}

__global__ __launch_bounds__(256) void k_col_norms(const double* A, int64_t lda, int64_t n, double* S) {
  __shared__ double red[4];
  const double* a = A + (int64_t)blockIdx.x * lda;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += a[i] * a[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) S[blockIdx.x] = sqrt(s);
}

__global__ __launch_bounds__(256) void k_jacobi_finish(const double* A, int64_t lda, const double* V, int64_t ldv,
                                                       int64_t n, const int* perm, const double* Sraw, double* U,
                                                       int64_t ldu, double* S, double* Vout, int64_t ldvo) {
  const int64_t t = blockIdx.x;
  const int src = perm[t];
  const double sig = Sraw[src];
  const double inv = (sig > 0.0) ? 1.0 / sig : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    U[i + t * ldu] = A[i + (int64_t)src * lda] * inv;
    Vout[i + t * ldvo] = V[i + (int64_t)src * ldv];
  }
  if (threadIdx.x == 0) S[t] = sig;
}

__global__ void k_set_identity(int64_t n, double* V, int64_t ldv) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, c = e / n;
  V[i + c * ldv] = (i == c) ? 1.0 : 0.0;
}

// V (rows x jb, ld rows): the reflectors of a panel whose top-left element is
// A[0]: unit diagonal, zeros above it, the stored vectors below.
__global__ void k_build_v(const double* A, int64_t lda, int64_t rows, int64_t jb, double* V) {
  const int64_t total = rows * jb;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, c = e / rows;
    V[e] = (i < c) ? 0.0 : (i == c ? 1.0 : A[i + c * lda]);
  }
}

// dlarft (forward, columnwise) from G = V^T V: T upper triangular (jb x jb,
// jb <= 64) with H(0) H(1) ... H(jb-1) = I - V T V^T, i.e. T(i, i) = tau_i and
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).  One 64-thread block.
__global__ __launch_bounds__(64) void k_build_t(const double* G, int64_t jb, const double* tau, double* T) {
  const int t = threadIdx.x;
  for (int64_t i = 0; i < jb; ++i) {
    if (t < i) {
      double w = 0.0;
      for (int64_t l = t; l < i; ++l) w += T[t + l * jb] * G[l + i * jb];
      T[t + i * jb] = -tau[i] * w;
    } else if (t == i) {
      T[i + i * jb] = tau[i];
    } else if (t < jb) {
      T[t + i * jb] = 0.0;
    }
    __syncthreads();
  }
}

// Q (rows x cols) = the first cols columns of the identity.
__global__ void k_eye_rect(int64_t rows, int64_t cols, double* Q, int64_t ldq) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % rows, c = e / rows;
    Q[i + c * ldq] = (i == c) ? 1.0 : 0.0;
  }
}

constexpr int64_t kQrPanel = 32;

unsigned grid_for_elems(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, 256), 65536)); }

// Workspace of the blocked QR: V (m x nb), G and T (nb x nb), two nb x n.
struct QrWork {
  double *V, *G, *T, *W1, *W2;
};
QrWork qr_work(DevBuf* work, int64_t m, int64_t n) {
  const int64_t nb = kQrPanel;
  const size_t oV = 0, oG = align_up((size_t)m * nb * sizeof(double), 256);
  const size_t oT = oG + align_up((size_t)nb * nb * sizeof(double), 256);
  const size_t o1 = oT + align_up((size_t)nb * nb * sizeof(double), 256);
  const size_t o2 = o1 + align_up((size_t)nb * n * sizeof(double), 256);
  work->reserve(o2 + (size_t)nb * n * sizeof(double));
  char* b = work->as<char>();
  return QrWork{reinterpret_cast<double*>(b + oV), reinterpret_cast<double*>(b + oG),
                reinterpret_cast<double*>(b + oT), reinterpret_cast<double*>(b + o1),
                reinterpret_cast<double*>(b + o2)};
}

// V and T of the panel at (j0, j0) with jb columns (rows j0 .. m-1).
void panel_vt(hipStream_t st, const double* A, int64_t lda, int64_t m, int64_t j0, int64_t jb, const double* tau,
              const QrWork& w, GemmWorkspace* gws) {
  const int64_t mp = m - j0;
  hipLaunchKernelGGL(k_build_v, dim3(grid_for_elems(mp * jb)), dim3(256), 0, st, A + j0 + j0 * lda, lda, mp, jb, w.V);
  check_launch("k_build_v");
  gemm_f64(st, 'T', 'N', jb, jb, mp, 1.0, w.V, mp, w.V, mp, 0.0, w.G, jb, gws);
  hipLaunchKernelGGL(k_build_t, dim3(1), dim3(64), 0, st, w.G, jb, tau + j0, w.T);
  check_launch("k_build_t");
}

}  // namespace

void householder_qr(hipStream_t st, int64_t m, int64_t n, double* A, int64_t lda, double* tau, DevBuf* work,
                    GemmWorkspace* gws) {
  SKEMA_REQUIRE(m >= n && n >= 0, kErrShape, "householder_qr: needs rows >= cols");
  SKEMA_REQUIRE(lda >= m, kErrShape, "householder_qr: lda < rows");
  if (n == 0) return;
  if (!work || n <= kQrPanel) {
    hipLaunchKernelGGL(k_qr_first, dim3(1), dim3(kQrThreads), 0, st, A, m, tau);
    check_launch("k_qr_first");
    for (int64_t j = 0; j + 1 < n; ++j) {
      hipLaunchKernelGGL(k_qr_step, dim3((unsigned)(n - j - 1)), dim3(kQrThreads), 0, st, A, lda, m, n, j, tau);
    }
    check_launch("k_qr_step");
    return;
  }
  const QrWork w = qr_work(work, m, n);
  for (int64_t j0 = 0; j0 < n; j0 += kQrPanel) {
    const int64_t jb = std::min(kQrPanel, n - j0), mp = m - j0;
    // the panel, column by column (dgeqr2 on rows j0..m-1 of columns j0..j0+jb-1)
    hipLaunchKernelGGL(k_qr_first, dim3(1), dim3(kQrThreads), 0, st, A + j0 + j0 * lda, mp, tau + j0);
    for (int64_t j = j0; j + 1 < j0 + jb; ++j)
      hipLaunchKernelGGL(k_qr_step, dim3((unsigned)(j0 + jb - j - 1)), dim3(kQrThreads), 0, st, A, lda, m, n, j, tau);
    check_launch("k_qr_step");
    if (j0 + jb >= n) break;
    // trailing columns: A2 <- (I - V T V^T)^T A2 = A2 - V (T^T (V^T A2))  (dlarfb 'L','T','F','C')
    const int64_t n2 = n - j0 - jb;
    double* A2 = A + j0 + (j0 + jb) * lda;
    panel_vt(st, A, lda, m, j0, jb, tau, w, gws);
    gemm_f64(st, 'T', 'N', jb, n2, mp, 1.0, w.V, mp, A2, lda, 0.0, w.W1, jb, gws);
    gemm_f64(st, 'T', 'N', jb, n2, jb, 1.0, w.T, jb, w.W1, jb, 0.0, w.W2, jb, gws);
    gemm_f64(st, 'N', 'N', mp, n2, jb, -1.0, w.V, mp, w.W2, jb, 1.0, A2, lda, gws);
  }
}

void householder_q(hipStream_t st, int64_t m, int64_t n, const double* A, int64_t lda, const double* tau,
                   double* Q, int64_t ldq, DevBuf* work, GemmWorkspace* gws) {
  SKEMA_REQUIRE(m >= n && n >= 0 && lda >= m && ldq >= m, kErrShape, "householder_q: bad shape");
  if (n == 0) return;
  if (!work || n <= kQrPanel) {
    for (int64_t j = n - 1; j >= 0; --j) {
      hipLaunchKernelGGL(k_org_step, dim3((unsigned)(n - j)), dim3(kQrThreads), 0, st, A, lda, tau, Q, ldq, m, j);
    }
    check_launch("k_org_step");
    return;
  }
  const QrWork w = qr_work(work, m, n);
  hipLaunchKernelGGL(k_eye_rect, dim3(grid_for_elems(m * n)), dim3(256), 0, st, m, n, Q, ldq);
  check_launch("k_eye_rect");
  // Q = H(0) ... H(n-1) [I; 0], one panel at a time from the last: panel p
  // acts on rows >= j0 only, where the columns before j0 are still zero
  for (int64_t j0 = ((n - 1) / kQrPanel) * kQrPanel; j0 >= 0; j0 -= kQrPanel) {
    const int64_t jb = std::min(kQrPanel, n - j0), mp = m - j0, nq = n - j0;
    double* Q2 = Q + j0 + j0 * ldq;
    panel_vt(st, A, lda, m, j0, jb, tau, w, gws);
    gemm_f64(st, 'T', 'N', jb, nq, mp, 1.0, w.V, mp, Q2, ldq, 0.0, w.W1, jb, gws);
    gemm_f64(st, 'N', 'N', jb, nq, jb, 1.0, w.T, jb, w.W1, jb, 0.0, w.W2, jb, gws);
    gemm_f64(st, 'N', 'N', mp, nq, jb, -1.0, w.V, mp, w.W2, jb, 1.0, Q2, ldq, gws);
  }
}

void extract_r(hipStream_t st, int64_t n, const double* A, int64_t lda, double* R, int64_t ldr) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_extract_r, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, st, n, A, lda, R, ldr);
  check_launch("k_extract_r");
}

void trsm_upper(hipStream_t st, int64_t n, int64_t nrhs, const double* T, int64_t ldt, double* B, int64_t ldb) {
  SKEMA_REQUIRE(n <= 8192, kErrUnsupported, "trsm_upper: order above 8192");
  if (n == 0 || nrhs == 0) return;
  hipLaunchKernelGGL(k_trsm_upper, dim3((unsigned)nrhs), dim3(256), n * sizeof(double), st, n, T, ldt, B, ldb);
  check_launch("k_trsm_upper");
}

void transpose(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb) {
  if (rows == 0 || cols == 0) return;
  dim3 grid((unsigned)cdiv(rows, 32), (unsigned)cdiv(cols, 32));
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, st, rows, cols, A, lda, B, ldb);
  check_launch("k_transpose");
}

int jacobi_svd(hipStream_t st, int64_t n, double* A, int64_t lda, double* U, int64_t ldu, double* S, double* V,
               int64_t ldv, DevBuf* scratch) {
  if (n == 0) return 0;
  SKEMA_REQUIRE(n < (1 << 20), kErrUnsupported, "jacobi_svd: order too large");
  // Round-robin (circle method) schedule: n' - 1 rounds of disjoint pairs.
  const int np = (int)(n + (n & 1));
  std::vector<int> arr(np);
  std::iota(arr.begin(), arr.end(), 0);
  std::vector<int2> pairs;
  std::vector<int> round_start;
  for (int r = 0; r < np - 1; ++r) {
    round_start.push_back((int)pairs.size());
    for (int i = 0; i < np / 2; ++i) {
      int p = arr[i], q = arr[np - 1 - i];
      if (p > q) std::swap(p, q);
      if (q < n) pairs.push_back(make_int2(p, q));
    }
    std::rotate(arr.begin() + 1, arr.end() - 1, arr.end());
  }
  round_start.push_back((int)pairs.size());

  const size_t off_pairs = 0;
  const size_t off_vw = align_up(pairs.size() * sizeof(int2), 256);
  const size_t off_cnt = off_vw + align_up((size_t)n * n * sizeof(double), 256);
  const size_t off_sraw = off_cnt + 256;
  const size_t off_perm = off_sraw + align_up((size_t)n * sizeof(double), 256);
  const size_t total = off_perm + align_up((size_t)n * sizeof(int), 256);
  scratch->reserve(total);
  char* base = scratch->as<char>();
  int2* d_pairs = reinterpret_cast<int2*>(base + off_pairs);
  double* Vw = reinterpret_cast<double*>(base + off_vw);
  int* d_cnt = reinterpret_cast<int*>(base + off_cnt);
  double* Sraw = reinterpret_cast<double*>(base + off_sraw);
  int* d_perm = reinterpret_cast<int*>(base + off_perm);

  SKEMA_HIP(hipMemcpyAsync(d_pairs, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_set_identity, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, st, n, Vw, n);
  check_launch("k_set_identity");

  const double tol = 1.0e-15;
  constexpr int kMaxSweeps = 100;
  int sweeps = 0;
  bool converged = false;
  for (; sweeps < kMaxSweeps; ++sweeps) {
    fill_i32(st, d_cnt, 1, 0);
    for (int r = 0; r + 1 < (int)round_start.size(); ++r) {
      const int cnt = round_start[r + 1] - round_start[r];
      if (cnt == 0) continue;
      hipLaunchKernelGGL(k_jacobi_round, dim3(cnt), dim3(256), 0, st, A, lda, Vw, n, n, d_pairs + round_start[r],
                         tol, d_cnt);
    }
    check_launch("k_jacobi_round");
    int h_cnt = 0;
    SKEMA_HIP(hipMemcpyAsync(&h_cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
    if (h_cnt == 0) {
      converged = true;
      break;
    }
  }
  SKEMA_REQUIRE(converged, kErrNumeric,
                "jacobi_svd: still rotating after " + std::to_string(kMaxSweeps) + " sweeps (order " + std::to_string(n) +
                    "); the factors would not be orthogonal");
  hipLaunchKernelGGL(k_col_norms, dim3((unsigned)n), dim3(256), 0, st, A, lda, n, Sraw);
  check_launch("k_col_norms");
  std::vector<double> hs(n);
  SKEMA_HIP(hipMemcpyAsync(hs.data(), Sraw, n * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return hs[a] > hs[b]; });
  SKEMA_HIP(hipMemcpyAsync(d_perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_jacobi_finish, dim3((unsigned)n), dim3(256), 0, st, A, lda, Vw, n, n, d_perm, Sraw, U, ldu, S,
                     V, ldv);
  check_launch("k_jacobi_finish");
  SKEMA_HIP(hipStreamSynchronize(st));
  return sweeps + 1;
}

}  // namespace skema
