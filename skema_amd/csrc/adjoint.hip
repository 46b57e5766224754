// Adjoint of the fused RBF window + Gauss sketch.
//
// Loss over the n points X (n x d), K = exp(-gamma D), D_ij = ||x_i - x_j||^2:
//   L = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2,   Y = K Omega^T, W = Phi K Psi^T.
// dL/dK = Gbar = 2 (Y Omega + Phi^T W Psi) = 2 B OP with B = [Y | Phi^T W]
// (n x h) and OP = [Omega; Psi] (h x n), h = k + s.  With H = (Gbar + Gbar^T)
// o K (elementwise):
//   dL/dx_a   = -2 gamma (sum_j H_aj x_a - sum_j H_aj x_j)
//   dL/dgamma = -1/2 sum_ij H_ij D_ij.
// Gbar + Gbar^T = 2 F R^T with F = [B | OP^T] (n x 2h) and R the same rows
// with the two column halves swapped, so a tile (I, J) of H is
// 2 (F_I R_J^T) o K_IJ: one GEMM tile with inner dimension 2h, the RBF tile
// recomputed from its Gram on the matrix cores, and the product H X_J folded
// into the same tile loop.  Neither K nor H is ever written out.
#include <vector>

#include "adjoint.hpp"
#include "linalg.hpp"
#include "lowp.hpp"
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "rbf_lowp.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kThreads = 512;

struct AdjArgs {
  const double* X;
  int64_t ldp, d, n;
  const double* nrm;
  double gamma;
  const double* F;  // n x 2h, ld = n
  int64_t h;
  int64_t jblocks_per_split;
  double* part_hx;  // [split][n x d]
  double* part_rs;  // [split][n]
  double* part_dg;  // [block]
};

// Tile pass over (I, J) pairs of 64 rows x 64 points (512 threads).  The block owns I and
// walks the J tiles of its split; everything is formed transposed (rows j,
// columns i) because the f64 16x16x4 result layout (lane l: column l%16,
// rows l/16 + 4v) then serves directly as the B operand of the H X product
// with K = j.  Per J tile:
//   S^T = R_J F_I^T over the 2h columns of F, streamed through
//         double-buffered LDS chunks of 16 (one barrier per chunk); wave w
//         owns j tile w%4 and i tiles 2(w/4), 2(w/4)+1,
//   the Gram^T tile over the d features (X_I in registers, X_J in LDS),
//   H^T = 2 S^T o K^T with its row sums and sum H o D in registers,
//   H^T through LDS, and (H X)^T += X_J^T H^T with wave w owning feature
//   tiles w%4 (+4, ...) and the same two i tiles, accumulated over all J.
// Neither K nor H is ever written to HBM.
template <int DMAX>
__global__ __launch_bounds__(kThreads) void k_rbf_adjoint_f64(AdjArgs g) {
  constexpr int BM = 64, BJ = 64, BKC = 16, TF = DMAX / 4, FT = DMAX / 16;
  constexpr int FTW = (FT + 3) / 4;  // feature tiles per wave in the H X product
  __shared__ double Rc[2][BKC][BJ + 2];
  __shared__ double Fc[2][BKC][BM + 2];
  __shared__ double XJ[DMAX][BJ + 1];
  __shared__ double Hs[BJ][BM + 2];
  __shared__ double red[BM];
  __shared__ double redg[8];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int jt = wave & 3, ih = wave >> 2;
  const int64_t i0 = (int64_t)blockIdx.x * BM;
  const int split = blockIdx.y;
  const int64_t nblk = cdiv(g.n, BJ);
  const int64_t jb0 = (int64_t)split * g.jblocks_per_split;
  const int64_t jb1 = (jb0 + g.jblocks_per_split < nblk) ? jb0 + g.jblocks_per_split : nblk;
  const int64_t two_h = 2 * g.h;
  const int nch = (int)((two_h + BKC - 1) / BKC);

  // Gram B operand: X_I rows i = i0 + 16 it + l16 (it = 2 ih, 2 ih + 1),
  // features 4t + lq; held in registers up to 64 features, re-read from
  // global memory (cache-resident, once per J tile) above that.
  constexpr int TFR = DMAX <= 64 ? TF : 1;
  double xi[2][TFR];
  double ni[2];
  int64_t irow[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t i = i0 + 16 * (2 * ih + u) + l16;
    irow[u] = i;
    if constexpr (DMAX <= 64) {
#pragma unroll
      for (int t = 0; t < TF; ++t) {
        const int64_t f = 4 * t + lq;
        xi[u][t] = (i < g.n && f < g.d) ? g.X[i + f * g.ldp] : 0.0;
      }
    }
    ni[u] = (i < g.n) ? g.nrm[i] : 0.0;
  }
  auto xi_at = [&](int u, int t) -> double {
    if constexpr (DMAX <= 64) {
      return xi[u][t];
    } else {
      const int64_t f = 4 * t + lq;
      return (irow[u] < g.n && f < g.d) ? g.X[irow[u] + f * g.ldp] : 0.0;
    }
  };
  d4 hx[FTW][2];
#pragma unroll
  for (int q = 0; q < FTW; ++q)
#pragma unroll
    for (int u = 0; u < 2; ++u) hx[q][u] = d4{0.0, 0.0, 0.0, 0.0};
  double rs[2] = {0.0, 0.0};
  double dg = 0.0;

  // chunk c of R_J (columns of F with halves swapped) and F_I: 4 doubles per thread
  double cr[2], cf[2];
  auto load_chunk = [&](int64_t j0, int t) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + q * kThreads;  // 0 .. 1023
      const int cc = e / 64, r = e % 64;
      const int64_t c = (int64_t)t * BKC + cc;
      const int64_t crs = (c < g.h) ? c + g.h : c - g.h;
      const int64_t j = j0 + r, i = i0 + r;
      cr[q] = (c < two_h && j < g.n) ? g.F[j + crs * g.n] : 0.0;
      cf[q] = (c < two_h && i < g.n) ? g.F[i + c * g.n] : 0.0;
    }
  };
  auto store_chunk = [&](int b) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + q * kThreads;
      Rc[b][e / 64][e % 64] = cr[q];
      Fc[b][e / 64][e % 64] = cf[q];
    }
  };

  for (int64_t jb = jb0; jb < jb1; ++jb) {
    const int64_t j0 = jb * BJ;
    __syncthreads();  // previous tile's readers of XJ / Hs / Rc / Fc are done
    for (int e = tid; e < DMAX * BJ; e += kThreads) {
      const int jj = e % BJ, f = e / BJ;
      const int64_t j = j0 + jj;
      XJ[f][jj] = (j < g.n && f < g.d) ? g.X[j + (int64_t)f * g.ldp] : 0.0;
    }
    load_chunk(j0, 0);
    store_chunk(0);
    if (nch > 1) load_chunk(j0, 1);
    __syncthreads();

    d4 sa[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
    for (int t = 0; t < nch; ++t) {
      const int b = t & 1;
#pragma unroll
      for (int kk = 0; kk < BKC; kk += 4) {
        const double a = Rc[b][kk + lq][16 * jt + l16];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          sa[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Fc[b][kk + lq][16 * (2 * ih + u) + l16], sa[u], 0, 0, 0);
      }
      if (t + 1 < nch) {
        store_chunk(b ^ 1);
        if (t + 2 < nch) load_chunk(j0, t + 2);
      }
      __syncthreads();
    }

    // Gram^T (rows j of tile jt, columns i of tiles 2ih, 2ih+1)
    d4 gr[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int t = 0; t < TF; ++t) {
      if (4 * t < g.d) {
        const double a = XJ[4 * t + lq][16 * jt + l16];
#pragma unroll
        for (int u = 0; u < 2; ++u) gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xi_at(u, t), gr[u], 0, 0, 0);
      }
    }
    // H^T = 2 S^T o K^T, row sums over j, sum H o D; H^T -> LDS
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int il = 16 * (2 * ih + u) + l16;
      const int64_t i = i0 + il;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jl = 16 * jt + lq + 4 * v;
        const int64_t j = j0 + jl;
        double hv = 0.0;
        if (i < g.n && j < g.n) {
          double D = ni[u] + g.nrm[j] - 2.0 * gr[u][v];
          D = (D > 0.0 && i != j) ? D : 0.0;
          const double kv = (i == j) ? 1.0 : exp(-g.gamma * D);
          hv = 2.0 * sa[u][v] * kv;
          dg += hv * D;
        }
        rs[u] += hv;
        Hs[jl][il] = hv;
      }
    }
    __syncthreads();
    // (H X)^T(f, i) += sum_j X_J(j, f) H^T(j, i)
#pragma unroll
    for (int kk = 0; kk < BJ; kk += 4) {
      double bh[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) bh[u] = Hs[kk + lq][16 * (2 * ih + u) + l16];
#pragma unroll
      for (int q = 0; q < FTW; ++q) {
        const int ft = jt + 4 * q;
        if (ft < FT) {
          const double a = XJ[16 * ft + l16][kk + lq];
#pragma unroll
          for (int u = 0; u < 2; ++u) hx[q][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bh[u], hx[q][u], 0, 0, 0);
        }
      }
    }
  }

  // (H X)^T result: lane l holds column i (l%16) and feature rows 16 ft + l/16 + 4v
#pragma unroll
  for (int q = 0; q < FTW; ++q) {
    const int ft = jt + 4 * q;
    if (ft >= FT) continue;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * (2 * ih + u) + l16;
        const int64_t f = 16 * ft + lq + 4 * v;
        if (i < g.n && f < g.d) g.part_hx[(int64_t)split * g.n * g.d + i + f * g.n] = hx[q][u][v];
      }
  }
  // row sums: over the 4 lane groups of a column, then over the 4 j-tile waves
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    double r = rs[u];
    r += __shfl_xor(r, 16);
    r += __shfl_xor(r, 32);
    rs[u] = r;
  }
  __syncthreads();
  if (jt == 0 && lq == 0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) red[16 * (2 * ih + u) + l16] = 0.0;
  }
  __syncthreads();
  for (int w = 0; w < 4; ++w) {  // fixed order over the j-tile waves
    if (jt == w && lq == 0) {
#pragma unroll
      for (int u = 0; u < 2; ++u) red[16 * (2 * ih + u) + l16] += rs[u];
    }
    __syncthreads();
  }
  if (tid < BM && i0 + tid < g.n) g.part_rs[(int64_t)split * g.n + i0 + tid] = red[tid];
  // sum H o D over the block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dg += __shfl_xor(dg, o);
  if (lane == 0) redg[wave] = dg;
  __syncthreads();
  if (tid == 0) {
    double sdg = 0.0;
    for (int w = 0; w < 8; ++w) sdg += redg[w];
    g.part_dg[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sdg;
  }
}

// dX = -2 gamma (rowsum o X - HX), deterministic sums over the splits.
__global__ void k_adjoint_finish(const double* X, int64_t ldp, int64_t n, int64_t d, int splits,
                                 const double* part_hx, const double* part_rs, double gamma, double* dX, int64_t ldg) {
  const int64_t total = n * d;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, f = e / n;
    double hx = 0.0, rs = 0.0;
    for (int s = 0; s < splits; ++s) {
      hx += part_hx[(int64_t)s * total + e];
      rs += part_rs[(int64_t)s * n + i];
    }
    dX[i + f * ldg] = -2.0 * gamma * (rs * X[i + f * ldp] - hx);
  }
}

// ---- BF16 form of the tile pass ------------------------------------------------
// The same tiles with S^T = R_J F_I^T and the Gram on v_mfma_f32_16x16x32_bf16
// (F and the points rounded to BF16, FP32 accumulation), K and H in FP32 and
// (H X)^T on the same instruction.  F is row-major n x 2hp: [B | OP^T] with
// each half zero-padded to hp (a multiple of 64), so R (halves swapped) is F
// read at column (c + hp) mod 2hp, one contiguous 64-column chunk at a time.
// 256 threads, I = 64 rows, J tiles of 64 points; wave w owns j rows 32(w&1)
// .. +31 (two 16-row tiles) and i columns 32(w>>1) .. +31 (two tiles).  The
// f32 result layout gives lane l rows 4(l/16)+v of each j tile, so the eight
// H values of a lane over its two j tiles are directly the B operand of the
// (H X)^T MFMA (K = the 32 j rows) with K slot q of lane l standing for
// j = 4(l/16)+q (q < 4) or 16+4(l/16)+q-4; X_J^T is read from LDS in that
// order.  Each wave keeps its own (H X)^T partial over its j rows; the two
// j halves go out as separate partial sums.
typedef __bf16 bf16x8a __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4a __attribute__((ext_vector_type(4)));
typedef float f4a __attribute__((ext_vector_type(4)));
typedef int i4a __attribute__((ext_vector_type(4)));

struct AdjLpArgs {
  const __bf16* P;   // rounded points, n x dp row-major
  const float* nrm;  // |x~_j|^2
  float gamma;
  int64_t n, d;
  const __bf16* F;   // n x 2hp row-major
  int64_t hp;
  int64_t jblocks_per_split;
  double* part_hx;  // [2 split + half][n x d]
  double* part_rs;  // [2 split + half][n]
  double* part_dg;  // [block]
};

template <int DP>
__global__ __launch_bounds__(256) void k_rbf_adjoint_bf16(AdjLpArgs g) {
  constexpr int BM = 64, BJ = 64, CK = 64, LDC = CK + 8, LDX = DP + 8, FT = DP / 16;
  __shared__ __attribute__((aligned(16))) __bf16 Rs[2][BJ][LDC];
  __shared__ __attribute__((aligned(16))) __bf16 Fs[2][BM][LDC];
  __shared__ __attribute__((aligned(16))) __bf16 XJ[BJ][LDX];  // rows j, features
  __shared__ __attribute__((aligned(16))) __bf16 XT[DP][BJ + 8];  // features, rows j
  __shared__ float nj[BJ];
  __shared__ float redg[4];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, q = lane >> 4;
  const int jg = wave & 1, ih = wave >> 1;
  const int64_t i0 = (int64_t)blockIdx.x * BM;
  const int split = blockIdx.y;
  const int64_t nblk = (g.n + BJ - 1) / BJ;
  const int64_t jb0 = (int64_t)split * g.jblocks_per_split;
  const int64_t jb1 = (jb0 + g.jblocks_per_split < nblk) ? jb0 + g.jblocks_per_split : nblk;
  const int64_t w2 = 2 * g.hp;
  const int nch = (int)(w2 / CK);

  // Gram B operand: X_I rows i = i0 + 16(2ih+u) + l16, features 32 ks + 8q .. +7
  bf16x8a xi[2][DP / 32];
  float ni[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t i = i0 + 16 * (2 * ih + u) + l16;
#pragma unroll
    for (int ks = 0; ks < DP / 32; ++ks) {
      if (i < g.n) xi[u][ks] = *reinterpret_cast<const bf16x8a*>(g.P + i * DP + 32 * ks + 8 * q);
      else for (int e = 0; e < 8; ++e) xi[u][ks][e] = (__bf16)0.f;
    }
    ni[u] = (i < g.n) ? g.nrm[i] : 0.f;
  }
  f4a hx[FT][2];
#pragma unroll
  for (int f = 0; f < FT; ++f)
#pragma unroll
    for (int u = 0; u < 2; ++u) hx[f][u] = f4a{0.f, 0.f, 0.f, 0.f};
  float rs[2] = {0.f, 0.f};
  float dg = 0.f;

  // 64-column chunk t of R_J and F_I: two 16-byte pieces of each per thread
  i4a cr[2], cf[2];
  auto load_chunk = [&](int64_t j0, int t) {
    const int64_t cF = (int64_t)t * CK, cR = (cF + g.hp) % w2;
#pragma unroll
    for (int e2 = 0; e2 < 2; ++e2) {
      const int e = tid + 256 * e2;  // 0 .. 511: row e/8, piece e%8
      const int r = e >> 3, pc = (e & 7) * 8;
      const int64_t j = j0 + r, i = i0 + r;
      cr[e2] = (j < g.n) ? *reinterpret_cast<const i4a*>(g.F + j * w2 + cR + pc) : i4a{0, 0, 0, 0};
      cf[e2] = (i < g.n) ? *reinterpret_cast<const i4a*>(g.F + i * w2 + cF + pc) : i4a{0, 0, 0, 0};
    }
  };
  auto store_chunk = [&](int b) {
#pragma unroll
    for (int e2 = 0; e2 < 2; ++e2) {
      const int e = tid + 256 * e2;
      const int r = e >> 3, pc = (e & 7) * 8;
      *reinterpret_cast<i4a*>(&Rs[b][r][pc]) = cr[e2];
      *reinterpret_cast<i4a*>(&Fs[b][r][pc]) = cf[e2];
    }
  };

  for (int64_t jb = jb0; jb < jb1; ++jb) {
    const int64_t j0 = jb * BJ;
    __syncthreads();  // previous tile's readers of XJ / XT / Rs / Fs are done
    for (int e = tid; e < BJ * (DP / 8); e += 256) {
      const int jj = e / (DP / 8), f0 = (e % (DP / 8)) * 8;
      const int64_t j = j0 + jj;
      bf16x8a v;
      if (j < g.n) v = *reinterpret_cast<const bf16x8a*>(g.P + j * DP + f0);
      else for (int t = 0; t < 8; ++t) v[t] = (__bf16)0.f;
      *reinterpret_cast<bf16x8a*>(&XJ[jj][f0]) = v;
#pragma unroll
      for (int t = 0; t < 8; ++t) XT[f0 + t][jj] = v[t];
    }
    if (tid < BJ) nj[tid] = (j0 + tid < g.n) ? g.nrm[j0 + tid] : 0.f;
    load_chunk(j0, 0);
    store_chunk(0);
    if (nch > 1) load_chunk(j0, 1);
    __syncthreads();

    f4a sacc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int u = 0; u < 2; ++u) sacc[a][u] = f4a{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < nch; ++t) {
      const int b = t & 1;
#pragma unroll
      for (int ks = 0; ks < CK / 32; ++ks) {
        bf16x8a ra[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
          ra[a] = *reinterpret_cast<const bf16x8a*>(&Rs[b][16 * (2 * jg + a) + l16][32 * ks + 8 * q]);
#pragma unroll
        for (int u = 0; u < 2; ++u)
          fb[u] = *reinterpret_cast<const bf16x8a*>(&Fs[b][16 * (2 * ih + u) + l16][32 * ks + 8 * q]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int u = 0; u < 2; ++u)
            sacc[a][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ra[a], fb[u], sacc[a][u], 0, 0, 0);
      }
      if (t + 1 < nch) {
        store_chunk(b ^ 1);
        if (t + 2 < nch) load_chunk(j0, t + 2);
      }
      __syncthreads();
    }

    // Gram^T: rows j (tiles 2jg, 2jg+1), columns i (tiles 2ih, 2ih+1)
    f4a gacc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int u = 0; u < 2; ++u) gacc[a][u] = f4a{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < DP / 32; ++ks) {
      if (32 * ks < g.d) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const bf16x8a xa = *reinterpret_cast<const bf16x8a*>(&XJ[16 * (2 * jg + a) + l16][32 * ks + 8 * q]);
#pragma unroll
          for (int u = 0; u < 2; ++u)
            gacc[a][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, xi[u][ks], gacc[a][u], 0, 0, 0);
        }
      }
    }
    // H^T = 2 S^T o K^T (FP32), row sums and sum H o D; H^T as the BF16 B
    // operand of the (H X)^T product
    bf16x8a hb[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int64_t i = i0 + 16 * (2 * ih + u) + l16;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int jl = 16 * (2 * jg + a) + 4 * q + v;
          const int64_t j = j0 + jl;
          float hv = 0.f;
          if (i < g.n && j < g.n) {
            float D = ni[u] + nj[jl] - 2.f * gacc[a][u][v];
            D = (D > 0.f && i != j) ? D : 0.f;
            const float kv = (i == j) ? 1.f : __expf(-g.gamma * D);
            hv = 2.f * sacc[a][u][v] * kv;
            dg += hv * D;
          }
          rs[u] += hv;
          hb[u][4 * a + v] = (__bf16)hv;
        }
    }
    // (H X)^T (f, i) += sum over the wave's 32 j of X_J^T (f, j) H^T (j, i)
#pragma unroll
    for (int ft = 0; ft < FT; ++ft) {
      if (16 * ft < g.d) {
        const __bf16* xr = &XT[16 * ft + l16][32 * jg];
        const bf16x4a lo = *reinterpret_cast<const bf16x4a*>(xr + 4 * q);
        const bf16x4a hi = *reinterpret_cast<const bf16x4a*>(xr + 16 + 4 * q);
        const bf16x8a xa = bf16x8a{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
        for (int u = 0; u < 2; ++u) hx[ft][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, hb[u], hx[ft][u], 0, 0, 0);
      }
    }
  }

  // (H X)^T partial of this wave: lane l holds column i (l%16), feature rows 16 ft + 4(l/16) + v
  const int64_t slot = 2 * (int64_t)split + jg;
#pragma unroll
  for (int ft = 0; ft < FT; ++ft)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * (2 * ih + u) + l16;
        const int64_t f = 16 * ft + 4 * q + v;
        if (i < g.n && f < g.d) g.part_hx[slot * g.n * g.d + i + f * g.n] = (double)hx[ft][u][v];
      }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    float r = rs[u];
    r += __shfl_xor(r, 16);
    r += __shfl_xor(r, 32);
    const int64_t i = i0 + 16 * (2 * ih + u) + l16;
    if (q == 0 && i < g.n) g.part_rs[slot * g.n + i] = (double)r;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dg += __shfl_xor(dg, o);
  if (lane == 0) redg[wave] = dg;
  __syncthreads();
  if (tid == 0) {
    double sdg = 0.0;
    for (int w = 0; w < 4; ++w) sdg += (double)redg[w];
    g.part_dg[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sdg;
  }
}

// F (n x 2hp, BF16 row-major) = [Y | PhiTW | 0 | OP^T | 0]: Y n x k and PhiTW
// n x s column-major (ld n), OP h x n column-major (ld h).
__global__ void k_pack_f_bf16(int64_t n, int64_t k, int64_t s, int64_t hp, const double* Y, const double* PhiTW,
                              const double* OP, __bf16* F) {
  const int64_t h = k + s, w2 = 2 * hp, total = n * w2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / w2, c = e % w2;
    double v = 0.0;
    if (c < hp) {
      if (c < k) v = Y[j + c * n];
      else if (c < h) v = PhiTW[j + (c - k) * n];
    } else if (c - hp < h) {
      v = OP[(c - hp) + j * h];
    }
    F[e] = (__bf16)(float)v;
  }
}
// Logical map (rows x cols, column-major FP64, ld) -> BF16 row-major (ld ldo).
__global__ void k_map_bf16_rows(const double* src, int64_t ld, int64_t rows, int64_t cols, __bf16* dst, int64_t ldo) {
  const int64_t total = rows * ldo;  // the padding columns [cols, ldo) are zero
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / ldo, c = e % ldo;
    dst[r * ldo + c] = (c < cols) ? (__bf16)(float)src[r + c * ld] : (__bf16)0.f;
  }
}

// Logical (rows x cols) view of a dense map into dst (ld = ldd).
void logical_copy(hipStream_t st, const Map* m, double* dst, int64_t ldd) {
  if (!m->transposed) copy_matrix(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, ldd);
  else transpose(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, ldd);
}

}  // namespace

namespace {

// BF16 form: forward through the BF16 fused generator (Y in FP64, G in
// BF16), W = Phi G on the BF16 NT GEMM, the FP64 loss of those, then the
// BF16 tile pass.
void loss_grad_bf16(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, const Map* omega,
                    const Map* phi, const Map* psi, int64_t window, double* loss, double* dX, int64_t ldg,
                    double* dgamma) {
  hipStream_t st = c->stream;
  const int64_t k = omega->lrows(), s = phi->lrows(), h = k + s;
  const int dp = rbf_bf16_padded_features(d);
  const int64_t hp = cdiv(h, 64) * 64, ldn = cdiv(n, 128) * 128;
  DevBuf P((size_t)n * dp * 2), nrm((size_t)n * 4), OP((size_t)h * n * sizeof(double)), OPb((size_t)h * ldn * 2),
      PhiL((size_t)s * n * sizeof(double)), Phib((size_t)s * ldn * 2), Y((size_t)n * k * sizeof(double)),
      Gb((size_t)s * ldn * 2), W((size_t)s * s * sizeof(double)), PhiTW((size_t)n * s * sizeof(double)),
      F((size_t)n * 2 * hp * 2), sc(2 * sizeof(double)), red;
  const unsigned ge = (unsigned)std::min<int64_t>(cdiv(h * n, 256), 65536);
  {
    Timed t(c, "rbf_forward");
    rbf_points_bf16(st, X, ldp, n, d, dp, P.as<__bf16>(), nrm.as<float>());
    logical_copy(st, omega, OP.as<double>(), h);
    logical_copy(st, psi, OP.as<double>() + k, h);
    logical_copy(st, phi, PhiL.as<double>(), s);
    hipLaunchKernelGGL(k_map_bf16_rows, dim3(ge), dim3(256), 0, st, OP.as<double>(), h, h, n, OPb.as<__bf16>(), ldn);
    hipLaunchKernelGGL(k_map_bf16_rows, dim3(ge), dim3(256), 0, st, PhiL.as<double>(), s, s, n, Phib.as<__bf16>(),
                       ldn);
    check_launch("k_map_bf16_rows");
    for (int64_t r0 = 0; r0 < n; r0 += window) {
      const int64_t rows = std::min(window, n - r0);
      GemmOut y;
      y.C1 = Y.as<double>() + r0; y.ldc1 = n; y.alpha1 = 1.0; y.beta1 = 0.0;
      rbf_right_bf16(c, P.as<__bf16>(), dp, nrm.as<float>(), (float)gamma, n, r0, rows, OPb.as<__bf16>(), ldn, k, s,
                     y, Gb.as<__bf16>() + r0, ldn);
    }
    // W = Phi G (s x s): rows of Phi and columns of G are both n-contiguous
    gemm_nt_bf16(st, s, s, n, Phib.as<__bf16>(), ldn, Gb.as<__bf16>(), ldn, 1.0, W.as<double>(), s, &c->gws.buf);
    sum_squares(st, n, k, Y.as<double>(), n, sc.as<double>(), &red);
    sum_squares(st, s, s, W.as<double>(), s, sc.as<double>() + 1, &red);
  }
  double hs[2];
  SKEMA_HIP(hipMemcpyAsync(hs, sc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  *loss = hs[0] + hs[1];
  if (!dX) {
    *dgamma = 0.0;
    return;
  }
  // B = [Y | Phi^T W] and OP^T, rounded into F
  gemm_f64(st, 'T', 'N', n, s, s, 1.0, PhiL.as<double>(), s, W.as<double>(), s, 0.0, PhiTW.as<double>(), n, &c->gws);
  {
    const unsigned gf = (unsigned)std::min<int64_t>(cdiv(n * 2 * hp, 256), 65536);
    hipLaunchKernelGGL(k_pack_f_bf16, dim3(gf), dim3(256), 0, st, n, k, s, hp, Y.as<double>(), PhiTW.as<double>(),
                       OP.as<double>(), F.as<__bf16>());
    check_launch("k_pack_f_bf16");
  }
  // two 256-thread blocks per CU: split the J tiles until the grid covers them
  const int64_t iblocks = cdiv(n, 64), jblocks = cdiv(n, 64);
  int splits = (int)std::max<int64_t>(1, std::min<int64_t>(jblocks, cdiv(512, iblocks)));
  const int64_t per = cdiv(jblocks, splits);
  splits = (int)cdiv(jblocks, per);
  DevBuf phx((size_t)2 * splits * n * d * sizeof(double)), prs((size_t)2 * splits * n * sizeof(double)),
      pdg((size_t)splits * iblocks * sizeof(double));
  AdjLpArgs g;
  g.P = P.as<__bf16>(); g.nrm = nrm.as<float>(); g.gamma = (float)gamma; g.n = n; g.d = d; g.F = F.as<__bf16>();
  g.hp = hp; g.jblocks_per_split = per;
  g.part_hx = phx.as<double>(); g.part_rs = prs.as<double>(); g.part_dg = pdg.as<double>();
  dim3 grid((unsigned)iblocks, (unsigned)splits);
  {
    Timed t(c, "rbf_adjoint");
    if (dp == 32) hipLaunchKernelGGL(k_rbf_adjoint_bf16<32>, grid, dim3(256), 0, st, g);
    else if (dp == 64) hipLaunchKernelGGL(k_rbf_adjoint_bf16<64>, grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL(k_rbf_adjoint_bf16<128>, grid, dim3(256), 0, st, g);
    check_launch("k_rbf_adjoint_bf16");
  }
  const int nb = (int)std::min<int64_t>(cdiv(n * d, 256), 4096);
  hipLaunchKernelGGL(k_adjoint_finish, dim3(nb), dim3(256), 0, st, X, ldp, n, d, 2 * splits, phx.as<double>(),
                     prs.as<double>(), gamma, dX, ldg);
  check_launch("k_adjoint_finish");
  std::vector<double> hdg((size_t)splits * iblocks);
  SKEMA_HIP(hipMemcpyAsync(hdg.data(), pdg.p, hdg.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  double sdg = 0.0;
  for (double v : hdg) sdg += v;
  *dgamma = -0.5 * sdg;
}

}  // namespace

void rbf_sketch_loss_grad(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, const Map* omega,
                          const Map* phi, const Map* psi, int64_t window, int precision, double* loss, double* dX,
                          int64_t ldg, double* dgamma) {
  SKEMA_REQUIRE(precision == SKEMA_PREC_FP64 || precision == SKEMA_PREC_BF16, kErrUnsupported,
                "rbf_loss_grad: FP64 and BF16 forms are built");
  SKEMA_REQUIRE(!omega->sparse && !phi->sparse && !psi->sparse, kErrUnsupported, "rbf_loss_grad: dense maps only");
  SKEMA_REQUIRE(d <= kRbfMaxFeat, kErrUnsupported, "rbf_loss_grad: d above 128");
  const int64_t k = omega->lrows(), s = phi->lrows();
  SKEMA_REQUIRE(omega->lcols() == n && phi->lcols() == n && psi->lcols() == n && psi->lrows() == s, kErrShape,
                "rbf_loss_grad: maps must be Omega (k x n), Phi (s x n), Psi (s x n)");
  if (precision == SKEMA_PREC_BF16) {
    loss_grad_bf16(c, X, n, d, ldp, gamma, omega, phi, psi, window, loss, dX, ldg, dgamma);
    return;
  }
  hipStream_t st = c->stream;
  const int64_t h = k + s;
  DevBuf nrm(n * sizeof(double)), OP((size_t)h * n * sizeof(double)), F((size_t)n * 2 * h * sizeof(double)),
      PhiL((size_t)s * n * sizeof(double)), W((size_t)s * s * sizeof(double)), sc(2 * sizeof(double)), red;
  Timed tf(c, "rbf_forward");
  row_norms2(st, X, n, d, ldp, nrm.as<double>());
  logical_copy(st, omega, OP.as<double>(), h);
  logical_copy(st, psi, OP.as<double>() + k, h);
  logical_copy(st, phi, PhiL.as<double>(), s);
  double* Fp = F.as<double>();
  double* Y = Fp;           // F[:, 0:k]   = Y = K Omega^T
  double* G = Fp + k * n;   // F[:, k:h]   = G = K Psi^T, then Phi^T W
  // forward: [Y | G] window by window through the fused generator
  for (int64_t r0 = 0; r0 < n; r0 += window) {
    const int64_t rows = std::min(window, n - r0);
    RbfArgs a;
    a.X = X; a.ldp = ldp; a.d = d; a.nrm = nrm.as<double>(); a.gamma = gamma; a.row0 = r0;
    a.M = rows; a.N = h; a.K = n; a.OP = OP.as<double>(); a.ldop = h;
    a.out.C1 = Y + r0; a.out.ldc1 = n; a.out.alpha1 = 1.0; a.out.beta1 = 0.0; a.out.nsplit = k;
    a.out.C2 = G + r0; a.out.ldc2 = n; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    rbf_fused_right_f64(st, a, &c->gws.buf);
  }
  // W = Phi G (s x s); loss = ||Y||^2 + ||W||^2
  gemm_f64(st, 'N', 'N', s, s, n, 1.0, PhiL.as<double>(), s, G, n, 0.0, W.as<double>(), s, &c->gws);
  sum_squares(st, n, k, Y, n, sc.as<double>(), &red);
  sum_squares(st, s, s, W.as<double>(), s, sc.as<double>() + 1, &red);
  tf.stop();
  double hs[2];
  SKEMA_HIP(hipMemcpyAsync(hs, sc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  // B = [Y | Phi^T W] in place of G; F[:, h:2h] = OP^T
  gemm_f64(st, 'T', 'N', n, s, s, 1.0, PhiL.as<double>(), s, W.as<double>(), s, 0.0, G, n, &c->gws);
  transpose(st, h, n, OP.as<double>(), h, Fp + h * n, n);
  SKEMA_HIP(hipStreamSynchronize(st));
  *loss = hs[0] + hs[1];
  if (!dX) {
    *dgamma = 0.0;
    return;
  }
  // tile pass
  // one 512-thread block per CU: split the J tiles until the grid covers
  // the 256 CUs
  const int64_t iblocks = cdiv(n, 64), jblocks = cdiv(n, 64);
  int splits = (int)std::max<int64_t>(1, std::min<int64_t>(jblocks, cdiv(256, iblocks)));
  const int64_t per = cdiv(jblocks, splits);
  splits = (int)cdiv(jblocks, per);
  DevBuf phx((size_t)splits * n * d * sizeof(double)), prs((size_t)splits * n * sizeof(double)),
      pdg((size_t)splits * iblocks * sizeof(double));
  AdjArgs g;
  g.X = X; g.ldp = ldp; g.d = d; g.n = n; g.nrm = nrm.as<double>(); g.gamma = gamma; g.F = Fp; g.h = h;
  g.jblocks_per_split = per; g.part_hx = phx.as<double>(); g.part_rs = prs.as<double>(); g.part_dg = pdg.as<double>();
  dim3 grid((unsigned)iblocks, (unsigned)splits);
  {
    Timed t(c, "rbf_adjoint");
    if (d <= 32) hipLaunchKernelGGL(k_rbf_adjoint_f64<32>, grid, dim3(kThreads), 0, st, g);
    else if (d <= 64) hipLaunchKernelGGL(k_rbf_adjoint_f64<64>, grid, dim3(kThreads), 0, st, g);
    else hipLaunchKernelGGL(k_rbf_adjoint_f64<128>, grid, dim3(kThreads), 0, st, g);
    check_launch("k_rbf_adjoint_f64");
  }
  const int nb = (int)std::min<int64_t>(cdiv(n * d, 256), 4096);
  hipLaunchKernelGGL(k_adjoint_finish, dim3(nb), dim3(256), 0, st, X, ldp, n, d, splits, phx.as<double>(),
                     prs.as<double>(), gamma, dX, ldg);
  check_launch("k_adjoint_finish");
  std::vector<double> hdg((size_t)splits * iblocks);
  SKEMA_HIP(hipMemcpyAsync(hdg.data(), pdg.p, hdg.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  double sdg = 0.0;
  for (double v : hdg) sdg += v;
  *dgamma = -0.5 * sdg;
}

}  // namespace skema
