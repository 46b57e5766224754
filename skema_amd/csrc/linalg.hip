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

// Sum of v[k] over the 64 lanes of a wave for 32 values per lane; lane l
// returns the total of value (l >> 1) & 31.  Reduce-scatter by halving
// (16 + 8 + 4 + 2 + 1 exchanges), then one exchange with the neighbouring
// lane: 32 exchanges instead of 32 x 6.
template <int H, int O, int N>
__device__ inline void halve(double (&v)[N], bool up) {
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const double send = up ? v[k] : v[k + H];
    const double keep = up ? v[k + H] : v[k];
    v[k] = keep + __shfl_xor(send, O);
  }
}
__device__ inline double wave_sum32(double (&v)[32]) {
  const int lane = threadIdx.x & 63;
  halve<16, 32>(v, lane & 32);
  halve<8, 16>(v, lane & 16);
  halve<4, 8>(v, lane & 8);
  halve<2, 4>(v, lane & 4);
  halve<1, 2>(v, lane & 2);
  return v[0] + __shfl_xor(v[0], 1);
}
// Same for 16 values per lane: lane l returns the total of value (l >> 2) & 15.
__device__ inline double wave_sum16(double (&v)[16]) {
  const int lane = threadIdx.x & 63;
  halve<8, 32>(v, lane & 32);
  halve<4, 16>(v, lane & 16);
  halve<2, 8>(v, lane & 8);
  halve<1, 4>(v, lane & 4);
  const double t = v[0] + __shfl_xor(v[0], 2);
  return t + __shfl_xor(t, 1);
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

// Block-pair Jacobi round in registers.  The columns come in blocks of
// kJacobiBlock = 8; a round's block pairs are disjoint and each takes one
// 512-thread workgroup in which thread i holds row i of the pair's 16
// columns of A and of V (orders up to 512), so a rotation is local to every
// thread and only the inner products need a reduction.  A local round
// rotates 8 disjoint column pairs at fixed register positions: one
// block-wide reduction of their 24 inner products (fixed order), then each
// thread applies the 8 rotations to its row.  Between local rounds the
// columns move instead of the pair indices (so no register is indexed at run
// time):
//   intra (the first round of each sweep, every block once): pairs
//     (p, 7 - p) within each block, then positions 1..7 of each block
//     rotate by one (the circle method; 7 rounds, back in place after);
//   cross: pairs (p, 8 + p), then the second block shifts by one position
//     (8 rounds, back in place after).
// Padding columns (>= n) are zero and never rotate.
constexpr int kJacobiBlock = 8;

// Rotation of columns x, y from their inner products (the k_jacobi_round
// formula); false when they are already orthogonal to tol.
__device__ inline bool jacobi_cs(double alpha, double beta, double gamma, double tol, double& c, double& s) {
  if (!(fabs(gamma) > tol * sqrt(alpha * beta))) return false;
  const double zeta = (beta - alpha) / (2.0 * gamma);
  const double az = fabs(zeta);
  const double t = (az > 1e150) ? 0.5 / zeta : copysign(1.0, zeta) / (az + sqrt(1.0 + zeta * zeta));
  c = 1.0 / sqrt(1.0 + t * t);
  s = c * t;
  return true;
}

__global__ __launch_bounds__(512) void k_jacobi_rows(double* A, int64_t lda, double* V, int64_t ldv, int64_t n,
                                                     const int2* bpairs, int intra, double tol, int* rotations) {
  constexpr int NB = kJacobiBlock, NC = 2 * NB;
  __shared__ double red[8][24];
  __shared__ double par[24];
  __shared__ double cs[8][2];
  __shared__ int rotp[8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int2 bp = bpairs[blockIdx.x];
  const int64_t i = tid;
  double a[NC], v[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int64_t c = (int64_t)(t < NB ? bp.x : bp.y) * NB + (t % NB);
    const bool ok = i < n && c < n;
    a[t] = ok ? A[i + c * lda] : 0.0;
    v[t] = ok ? V[i + c * ldv] : 0.0;
  }
  int count = 0;
  // one local round on the pairs (pos_a(k), pos_b(k)), k < 8 (register
  // positions known at compile time)
  auto round = [&](auto pos_a, auto pos_b) {
    double w[32];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double x = a[pos_a(k)], y = a[pos_b(k)];
      w[3 * k] = x * x;
      w[3 * k + 1] = y * y;
      w[3 * k + 2] = x * y;
    }
#pragma unroll
    for (int k = 24; k < 32; ++k) w[k] = 0.0;
    const double tot = wave_sum32(w);
    if ((lane & 1) == 0 && (lane >> 1) < 24) red[wave][lane >> 1] = tot;
    __syncthreads();
    if (tid < 24) {
      double t = 0.0;
      for (int q = 0; q < 8; ++q) t += red[q][tid];
      par[tid] = t;
    }
    __syncthreads();
    if (tid < 8) {  // the 8 rotations, once each (FP64 divisions and square roots)
      double c = 1.0, s = 0.0;
      const bool rot = jacobi_cs(par[3 * tid], par[3 * tid + 1], par[3 * tid + 2], tol, c, s);
      cs[tid][0] = c;
      cs[tid][1] = rot ? s : 0.0;
      rotp[tid] = rot ? 1 : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (rotp[k]) {
        const double c = cs[k][0], s = cs[k][1];
        const int pa = pos_a(k), pb = pos_b(k);
        const double x = a[pa], y = a[pb];
        a[pa] = c * x - s * y;
        a[pb] = s * x + c * y;
        const double vx = v[pa], vy = v[pb];
        v[pa] = c * vx - s * vy;
        v[pb] = s * vx + c * vy;
        ++count;
      }
    }
    __syncthreads();  // par, red and cs are rewritten by the next round
  };
  if (intra) {
    for (int r = 0; r < NB - 1; ++r) {
      // pairs (p, 7 - p) in block X (k < 4) and block Y (k >= 4)
      round([](int k) { return k < 4 ? k : NB + (k - 4); }, [](int k) { return k < 4 ? NB - 1 - k : NC - 1 - (k - 4); });
      // circle method: positions 1 .. 7 of each block rotate by one
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int o = NB * b;
        const double ta = a[o + NB - 1], tv = v[o + NB - 1];
#pragma unroll
        for (int p = NB - 1; p > 1; --p) {
          a[o + p] = a[o + p - 1];
          v[o + p] = v[o + p - 1];
        }
        a[o + 1] = ta;
        v[o + 1] = tv;
      }
    }
  }
  for (int r = 0; r < NB; ++r) {
    round([](int k) { return k; }, [](int k) { return NB + k; });
    // the second block shifts by one position
    const double ta = a[NB], tv = v[NB];
#pragma unroll
    for (int p = NB; p < NC - 1; ++p) {
      a[p] = a[p + 1];
      v[p] = v[p + 1];
    }
    a[NC - 1] = ta;
    v[NC - 1] = tv;
  }
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int64_t c = (int64_t)(t < NB ? bp.x : bp.y) * NB + (t % NB);
    if (i < n && c < n) {
      A[i + c * lda] = a[t];
      V[i + c * ldv] = v[t];
    }
  }
  if (tid == 0 && count) atomicAdd(rotations, count);
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
// jb <= 32) with H(0) H(1) ... H(jb-1) = I - V T V^T, i.e. T(i, i) = tau_i and
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i).  One 64-thread block, G and T in
// LDS.
__global__ __launch_bounds__(64) void k_build_t(const double* G, int64_t jb, const double* tau, double* T) {
  __shared__ double Gs[32][33], Ts[32][33];
  const int t = threadIdx.x;
  for (int e = t; e < jb * jb; e += 64) {
    Gs[e % jb][e / jb] = G[e];
    Ts[e % jb][e / jb] = 0.0;
  }
  __syncthreads();
  for (int64_t i = 0; i < jb; ++i) {
    if (t < i) {
      double w = 0.0;
      for (int64_t l = t; l < i; ++l) w += Ts[t][l] * Gs[l][i];
      Ts[t][i] = -tau[i] * w;
    } else if (t == i) {
      Ts[i][i] = tau[i];
    }
    __syncthreads();
  }
  for (int e = t; e < jb * jb; e += 64) T[e] = Ts[e % jb][e / jb];
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

// ---- panel factorization (dgeqr2 on rows j0..m-1 of a <= 32-column panel) ----
// Step j of Householder QR needs, over the rows i > j of the panel, the
// squares of column j and its products with the later columns: with
// s_c = sum_{i>j} x_i a_ic (x = column j, s_j its squared norm), alpha =
// a_jj, beta = -sign(alpha) sqrt(alpha^2 + s_j), tau = (beta - alpha)/beta,
// v = [1; x / (alpha - beta)], the update of column c is
// a_c -= v tau (a_jc + s_c / (alpha - beta)).  So one reduction per column
// (the row of products) drives the whole step.

// Reflector of step j from the reduced row of products ss (ss[j] = s_j) and
// the row j of the panel: returns false when x = 0 (tau = 0, no update).
__device__ inline bool reflector(double alpha, double xnorm2, double& beta, double& scal, double& tau) {
  if (xnorm2 == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 0.0;
    return false;
  }
  beta = -copysign(hypot(alpha, sqrt(xnorm2)), alpha);
  scal = 1.0 / (alpha - beta);
  tau = (beta - alpha) / beta;
  return true;
}

// Short panels (mp <= 1024 rows): one block of 512 threads, two rows per
// thread held in registers, the jb steps in one launch.
__global__ __launch_bounds__(512) void k_panel_small(double* P, int64_t lda, int64_t mp, int jb, double* tau) {
  constexpr int RPT = 2;
  __shared__ double red[8][32];
  __shared__ double ss[32], row[32], wv[32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double a[RPT][32];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = tid + 512 * q;
#pragma unroll
    for (int c = 0; c < 32; ++c) a[q][c] = (i < mp && c < jb) ? P[i + c * lda] : 0.0;
  }
  for (int j = 0; j < jb; ++j) {
    double x[RPT];  // a[q][j] without a dynamic register index
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      x[q] = 0.0;
#pragma unroll
      for (int c = 0; c < 32; ++c)
        if (c == j) x[q] = a[q][c];
    }
    // products with the columns c >= j, 16 columns at a time (register budget)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (16 * h + 15 < j) continue;  // wave-uniform: the whole chunk is before j
      double v[16];
#pragma unroll
      for (int cc = 0; cc < 16; ++cc) {
        const int c = 16 * h + cc;
        v[cc] = 0.0;
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (tid + 512 * q > j && c >= j) v[cc] += x[q] * a[q][c];
      }
      const double tot = wave_sum16(v);
      if ((lane & 3) == 0) red[wave][16 * h + (lane >> 2)] = tot;
    }
    if (tid == j) {  // row j < 32: thread j, q = 0
#pragma unroll
      for (int c = 0; c < 32; ++c) row[c] = a[0][c];
    }
    __syncthreads();
    if (tid >= j && tid < 32) {  // red[.][c] for c < j is stale: never read
      double t = 0.0;
      for (int w = 0; w < 8; ++w) t += red[w][tid];
      ss[tid] = t;
    }
    __syncthreads();
    double beta, scal, tj;
    const bool act = reflector(row[j], ss[j], beta, scal, tj);
    if (tid == 0) tau[j] = tj;
    if (tid > j && tid < 32) wv[tid] = tj * (row[tid] + ss[tid] * scal);  // update coefficient of column tid
    __syncthreads();
    if (act) {
      // 16 columns at a time: the memory clobber keeps the coefficient loads
      // of the second half from being hoisted next to the first (registers)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        asm volatile("" ::: "memory");
        if (16 * h + 15 < j) continue;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          const int64_t i = tid + 512 * q;
          if (i > j && i < mp) {
            const double vi = x[q] * scal;
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) {
              const int c = 16 * h + cc;
              if (c == j) a[q][c] = vi;
              else if (c > j) a[q][c] -= vi * wv[c];
            }
          } else if (i == j) {
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) {
              const int c = 16 * h + cc;
              if (c == j) a[q][c] = beta;
              else if (c > j) a[q][c] -= wv[c];
            }
          }
        }
      }
    }
    __syncthreads();  // row, ss and wv are rewritten by the next step
  }
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = tid + 512 * q;
    if (i < mp) {
#pragma unroll
      for (int c = 0; c < 32; ++c)
        if (c < jb) P[i + c * lda] = a[q][c];
    }
  }
}

// Tall panels: one launch per step j (j = -1 .. jb-1), the rows spread over
// nblk blocks of 256 threads (RPT rows per thread, in registers).  Launch j
// reduces the partial rows of products that launch j-1 left in part[j & 1]
// (every block sums all blocks' partials in the same fixed order), applies
// step j to its rows, and leaves the partials of step j+1 in
// part[(j+1) & 1] and row j+1 of the panel in rowbuf[(j+1) & 1] (row j+1 <
// 32 always lies in block 0; the copy keeps other blocks from reading it
// while block 0 rewrites it).
constexpr int kPanelThreads = 256;

template <int RPT>
__global__ __launch_bounds__(kPanelThreads) void k_panel_step(double* P, int64_t lda, int64_t mp, int jb, int j,
                                                              double* tau, double* part, double* rowbuf, int nblk) {
  __shared__ double red[8][32];
  __shared__ double ss[32], row[32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t i0 = (int64_t)blockIdx.x * (kPanelThreads * RPT);
  const int c0 = j > 0 ? j : 0;  // columns before j are final
  double a[RPT][32];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int64_t i = i0 + tid + kPanelThreads * q;
#pragma unroll
    for (int c = 0; c < 32; ++c) a[q][c] = (i < mp && c >= c0 && c < jb) ? P[i + c * lda] : 0.0;
  }
  if (j >= 0) {
    {  // reduce part[j & 1]: 8 groups of 32 threads over the blocks, then the groups in order
      const double* pj = part + (int64_t)(j & 1) * nblk * 32;
      const int g = tid >> 5, c = tid & 31;
      double t = 0.0;
      for (int b = g; b < nblk; b += 8) t += pj[(int64_t)b * 32 + c];
      red[g][c] = t;
      if (tid < 32) row[tid] = rowbuf[(j & 1) * 32 + tid];
    }
    __syncthreads();
    if (tid < 32) {
      double t = 0.0;
      for (int g = 0; g < 8; ++g) t += red[g][tid];
      ss[tid] = t;
    }
    __syncthreads();
    double beta, scal, tj;
    const bool act = reflector(row[j], ss[j], beta, scal, tj);
    if (blockIdx.x == 0 && tid == 0) tau[j] = tj;
    if (act) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int64_t i = i0 + tid + kPanelThreads * q;
        if (i > j && i < mp) {
          double vi = 0.0;
#pragma unroll
          for (int c = 0; c < 32; ++c)
            if (c == j) vi = a[q][c] * scal;
#pragma unroll
          for (int c = 0; c < 32; ++c) {
            if (c == j) a[q][c] = vi;
            else if (c > j) a[q][c] -= vi * (tj * (row[c] + ss[c] * scal));
          }
        } else if (i == j) {
#pragma unroll
          for (int c = 0; c < 32; ++c) {
            if (c == j) a[q][c] = beta;
            else if (c > j) a[q][c] -= tj * (row[c] + ss[c] * scal);
          }
        }
        if (i < mp) {
#pragma unroll
          for (int c = 0; c < 32; ++c)
            if (c >= j && c < jb) P[i + c * lda] = a[q][c];
        }
      }
    }
  }
  const int jn = j + 1;
  if (jn < jb) {
    double v[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) v[c] = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int64_t i = i0 + tid + kPanelThreads * q;
      double x = 0.0;
#pragma unroll
      for (int c = 0; c < 32; ++c)
        if (c == jn) x = a[q][c];
      if (i > jn && i < mp) {
#pragma unroll
        for (int c = 0; c < 32; ++c)
          if (c >= jn) v[c] += x * a[q][c];
      }
    }
    const double tot = wave_sum32(v);
    if ((lane & 1) == 0) red[wave][lane >> 1] = tot;
    if (blockIdx.x == 0 && tid == jn) {  // row jn: block 0, thread jn, q = 0
#pragma unroll
      for (int c = 0; c < 32; ++c) rowbuf[(jn & 1) * 32 + c] = a[0][c];
    }
    __syncthreads();
    if (tid < 32) {
      double t = 0.0;
      for (int w = 0; w < kPanelThreads / 64; ++w) t += red[w][tid];
      part[(int64_t)(jn & 1) * nblk * 32 + (int64_t)blockIdx.x * 32 + tid] = t;
    }
  }
}

unsigned grid_for_elems(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, 256), 65536)); }

// Panel of jb <= 32 columns at P (rows 0..mp-1): reflectors below the
// diagonal, R on and above it, tau[0..jb).
constexpr int64_t kPanelSmallRows = 1024;
int64_t panel_rpt(int64_t mp) { return mp > 32768 ? 2 : 1; }
int64_t panel_blocks(int64_t mp) { return cdiv(mp, kPanelThreads * panel_rpt(mp)); }
// Most blocks any panel of a matrix with m rows takes (panels have m - j0
// rows; below 32769 rows a panel takes one row per thread).
int64_t panel_blocks_max(int64_t m) { return std::max(panel_blocks(m), std::min<int64_t>(cdiv(m, kPanelThreads), 128)); }

void factor_panel(hipStream_t st, double* P, int64_t lda, int64_t mp, int64_t jb, double* tau, double* part,
                  double* rowbuf) {
  if (mp <= kPanelSmallRows) {
    hipLaunchKernelGGL(k_panel_small, dim3(1), dim3(512), 0, st, P, lda, mp, (int)jb, tau);
    check_launch("k_panel_small");
    return;
  }
  const int nblk = (int)panel_blocks(mp);
  for (int j = -1; j < (int)jb; ++j) {
    if (panel_rpt(mp) == 2)
      hipLaunchKernelGGL(k_panel_step<2>, dim3(nblk), dim3(kPanelThreads), 0, st, P, lda, mp, (int)jb, j, tau, part,
                         rowbuf, nblk);
    else
      hipLaunchKernelGGL(k_panel_step<1>, dim3(nblk), dim3(kPanelThreads), 0, st, P, lda, mp, (int)jb, j, tau, part,
                         rowbuf, nblk);
  }
  check_launch("k_panel_step");
}

// Workspace of the blocked QR: V (m x nb), G (nb x nb), the T factor of
// every panel (Tall, ceil(n / nb) x nb x nb, so the Q formation can reuse
// them), two nb x n products, and the panel step's partial rows and row copy.
struct QrWork {
  double *V, *G, *Tall, *W1, *W2, *part, *rowbuf;
};
QrWork qr_work(DevBuf* work, int64_t m, int64_t n) {
  const int64_t nb = kQrPanel, np = cdiv(n, nb);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(bytes, 256);
    return at;
  };
  const size_t oV = take((size_t)m * nb * sizeof(double));
  const size_t oG = take((size_t)nb * nb * sizeof(double));
  const size_t oT = take((size_t)np * nb * nb * sizeof(double));
  const size_t o1 = take((size_t)nb * n * sizeof(double));
  const size_t o2 = take((size_t)nb * n * sizeof(double));
  const size_t oP = take((size_t)2 * panel_blocks_max(m) * 32 * sizeof(double));
  const size_t oR = take((size_t)2 * 32 * sizeof(double));
  work->reserve(o);
  char* b = work->as<char>();
  auto at = [&](size_t off) { return reinterpret_cast<double*>(b + off); };
  return QrWork{at(oV), at(oG), at(oT), at(o1), at(o2), at(oP), at(oR)};
}

// V of the panel at (j0, j0) with jb columns (rows j0 .. m-1).
void panel_v(hipStream_t st, const double* A, int64_t lda, int64_t m, int64_t j0, int64_t jb, const QrWork& w) {
  const int64_t mp = m - j0;
  hipLaunchKernelGGL(k_build_v, dim3(grid_for_elems(mp * jb)), dim3(256), 0, st, A + j0 + j0 * lda, lda, mp, jb, w.V);
  check_launch("k_build_v");
}

// T of the panel (from G = V^T V) into its slot of Tall.
double* panel_t(hipStream_t st, int64_t m, int64_t j0, int64_t jb, const double* tau, const QrWork& w,
                GemmWorkspace* gws) {
  const int64_t mp = m - j0;
  double* T = w.Tall + (j0 / kQrPanel) * kQrPanel * kQrPanel;
  gemm_f64(st, 'T', 'N', jb, jb, mp, 1.0, w.V, mp, w.V, mp, 0.0, w.G, jb, gws);
  hipLaunchKernelGGL(k_build_t, dim3(1), dim3(64), 0, st, w.G, jb, tau + j0, T);
  check_launch("k_build_t");
  return T;
}

}  // namespace

void householder_qr(hipStream_t st, int64_t m, int64_t n, double* A, int64_t lda, double* tau, DevBuf* work,
                    GemmWorkspace* gws) {
  SKEMA_REQUIRE(m >= n && n >= 0, kErrShape, "householder_qr: needs rows >= cols");
  SKEMA_REQUIRE(lda >= m, kErrShape, "householder_qr: lda < rows");
  if (n == 0) return;
  if (!work) {
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
    factor_panel(st, A + j0 + j0 * lda, lda, mp, jb, tau + j0, w.part, w.rowbuf);
    // T of every panel (the Q formation reuses them), then the trailing
    // columns: A2 <- (I - V T V^T)^T A2 = A2 - V (T^T (V^T A2))  (dlarfb 'L','T','F','C')
    panel_v(st, A, lda, m, j0, jb, w);
    const double* T = panel_t(st, m, j0, jb, tau, w, gws);
    if (j0 + jb >= n) break;
    const int64_t n2 = n - j0 - jb;
    double* A2 = A + j0 + (j0 + jb) * lda;
    gemm_f64(st, 'T', 'N', jb, n2, mp, 1.0, w.V, mp, A2, lda, 0.0, w.W1, jb, gws);
    gemm_f64(st, 'T', 'N', jb, n2, jb, 1.0, T, jb, w.W1, jb, 0.0, w.W2, jb, gws);
    gemm_f64(st, 'N', 'N', mp, n2, jb, -1.0, w.V, mp, w.W2, jb, 1.0, A2, lda, gws);
  }
}

void householder_q(hipStream_t st, int64_t m, int64_t n, const double* A, int64_t lda, const double* tau,
                   double* Q, int64_t ldq, DevBuf* work, GemmWorkspace* gws, bool reuse_t) {
  SKEMA_REQUIRE(m >= n && n >= 0 && lda >= m && ldq >= m, kErrShape, "householder_q: bad shape");
  if (n == 0) return;
  if (!work) {
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
    panel_v(st, A, lda, m, j0, jb, w);
    const double* T = reuse_t ? w.Tall + (j0 / kQrPanel) * kQrPanel * kQrPanel : panel_t(st, m, j0, jb, tau, w, gws);
    gemm_f64(st, 'T', 'N', jb, nq, mp, 1.0, w.V, mp, Q2, ldq, 0.0, w.W1, jb, gws);
    gemm_f64(st, 'N', 'N', jb, nq, jb, 1.0, T, jb, w.W1, jb, 0.0, w.W2, jb, gws);
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
  // Block-pair rounds in registers for orders up to 512 (one launch per
  // round of block pairs, ~n/16 per sweep), else scalar rounds (~n).
  const bool blocked = n > kJacobiBlock && n <= 512;
  const int64_t units = blocked ? cdiv(n, kJacobiBlock) : n;
  // Round-robin (circle method) schedule over the units: units' - 1 rounds
  // of disjoint pairs.
  const int np = (int)(units + (units & 1));
  std::vector<int> arr(np);
  std::iota(arr.begin(), arr.end(), 0);
  std::vector<int2> pairs;
  std::vector<int> round_start;
  for (int r = 0; r < np - 1; ++r) {
    round_start.push_back((int)pairs.size());
    for (int i = 0; i < np / 2; ++i) {
      int p = arr[i], q = arr[np - 1 - i];
      if (p > q) std::swap(p, q);
      if (blocked) pairs.push_back(make_int2(p, q));  // a block beyond the last holds only padding
      else if (q < n) pairs.push_back(make_int2(p, q));
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
      if (blocked)
        hipLaunchKernelGGL(k_jacobi_rows, dim3(cnt), dim3(512), 0, st, A, lda, Vw, n, n, d_pairs + round_start[r],
                           r == 0 ? 1 : 0, tol, d_cnt);
      else
        hipLaunchKernelGGL(k_jacobi_round, dim3(cnt), dim3(256), 0, st, A, lda, Vw, n, n, d_pairs + round_start[r],
                           tol, d_cnt);
    }
    check_launch(blocked ? "k_jacobi_rows" : "k_jacobi_round");
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
