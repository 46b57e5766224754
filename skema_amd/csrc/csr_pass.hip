// Whole-matrix pass of linear_update over CSR input.
//
// linear_update starts from a zero state, so the reference's window loop
// (src/Skema_SketchySVD.cpp:167-233: X <- nu X + eta x_w, Y[idx] <- nu Y[idx]
// + eta y_w, Z <- nu Z + eta z_w over W windows) has the closed form
//   X = Upsilon diag(c) A,   Z = (Phi diag(c) A) Psi^T,   Y[i, :] = eta A[i, :] Omega^T
// with c_i = eta nu^(W-1-w(i)) for row i in window w(i).  The whole CSR
// (48 MB at 200000 x 20000, nnz 4e6) fits in HBM many times over, so the
// pass runs as a handful of large launches instead of W rounds of small
// ones, with no host synchronisation between windows.
//
// Sparse (sparse-sign) maps:
//   CSC of A in two bucketing levels (column tiles, then columns) with LDS
//   cursors, so the entries of a column come in no fixed order;
//   left  (X and W = Phi diag(c) A in one pass): one wave per column of A,
//         lanes over its nonzeros, each nonzero's term c_i a added into the
//         rows p of the maps' column i;
//   right (Y): one block per 32 rows of A, threads over their nonzeros.
//   Both accumulate in LDS in int64 fixed point (fixedpoint.hpp),
//   so every sum is the same whatever order the atomics land in: X, Y, Z
//   are bitwise reproducible.
// Dense maps: one wave per column (left) or row (right), lanes own map rows,
//   nonzeros in ascending row order over a sorted CSC (deterministic).
//   core  Z = W Psi^T: dense x CSR^T gather (sparse Psi) or the FP64 GEMM.
#include <vector>

#include "csr_pass.hpp"
#include "fixedpoint.hpp"
#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

// ---- CSC of A by two bucketing levels (sparse-map path) --------------------
// Level 1: blocks of kL1Rows rows count their nonzeros per column tile (nt
// tiles of tc columns) in LDS; an exclusive scan over (tile, block) gives
// each block a contiguous run in every tile's bucket; the scatter pass walks
// the nonzeros again (coalesced) and writes (row, column, value) into the
// runs.  Level 2: one block per tile counts the tile's nonzeros per column,
// scans the counts in LDS into the column pointers and places each nonzero.
// Runs are filled through LDS cursors, so the entries of a column come in no
// fixed order; the products below do not depend on it.
constexpr int kL1Rows = 256;
constexpr int kMaxTiles = 256;
constexpr int kMaxTileCols = 4096;

__global__ __launch_bounds__(256) void k_l1_count(const int64_t* rowptr, int64_t rows, const int32_t* col, int tc,
                                                 int nt, int nb, int* cnt) {
  __shared__ int hist[kMaxTiles];
  for (int t = threadIdx.x; t < nt; t += 256) hist[t] = 0;
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * kL1Rows;
  const int64_t r1 = (r0 + kL1Rows < rows) ? r0 + kL1Rows : rows;
  for (int64_t e = rowptr[r0] + threadIdx.x; e < rowptr[r1]; e += 256) atomicAdd(&hist[col[e] / tc], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < nt; t += 256) cnt[(int64_t)t * nb + blockIdx.x] = hist[t];
}

// Per tile t: exclusive scan of its counts over the row blocks (contiguous,
// cnt[t nb .. t nb + nb)) into off, and the tile's total into tot[t].
__global__ __launch_bounds__(1024) void k_tile_scan(const int* cnt, int nb, int* off, int* tot) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int* c = cnt + (int64_t)t * nb;
  int* o = off + (int64_t)t * nb;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int v = (b0 + tid < nb) ? c[b0 + tid] : 0;
    int x = v;  // inclusive scan within the wave
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = carry;
    for (int i = 0; i < w; ++i) before += wsum[i];
    if (b0 + tid < nb) o[b0 + tid] = before + x - v;
    __syncthreads();
    if (tid == 1023) carry = before + x;
    __syncthreads();
  }
  if (tid == 0) tot[t] = carry;
}

// base[t] = sum of tot[0 .. t) (nt <= 1024, one block); base[nt] = nnz.
__global__ __launch_bounds__(1024) void k_tile_base(const int* tot, int nt, int* base) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  part[tid] = tid < nt ? tot[tid] : 0;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  if (tid <= nt) base[tid] = tid ? part[tid - 1] : 0;
}

// The block's nonzeros go out in batches of kL1Batch: each batch is
// counting-sorted by column tile in LDS (entries of a tile contiguous), then
// written as 16-byte records (row, column, value) with consecutive threads
// on consecutive positions of each tile's run, so the writes arrive as runs
// of whole cache lines instead of one scattered word per nonzero (which left
// lines partly written and multiplied the HBM write traffic).
constexpr int kL1Batch = 4096;

struct CooRec {
  int32_t row, col;
  double val;
};
static_assert(sizeof(CooRec) == 16, "16-byte records");

__global__ __launch_bounds__(256) void k_l1_scatter(const int64_t* rowptr, int64_t rows, const int32_t* col,
                                                   const double* val, int tc, int nt, int nb, const int* off,
                                                   const int* base, CooRec* brec) {
  constexpr int E = kL1Batch / 256;
  __shared__ int cur[kMaxTiles];  // next global position of each tile's run
  __shared__ int cnt[kMaxTiles];  // entries of the batch per tile
  __shared__ int lst[kMaxTiles];  // LDS start of each tile in the batch
  __shared__ int wsum[4];
  __shared__ int64_t rp[kL1Rows + 1];
  __shared__ CooRec sb[kL1Batch];
  __shared__ uint8_t rowof[kL1Batch];  // row (within the block) of each entry of the batch
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * kL1Rows;
  const int nr = (int)((r0 + kL1Rows < rows) ? kL1Rows : rows - r0);
  static_assert(kMaxTiles == 256, "one thread per tile");
  cur[tid] = (tid < nt) ? base[tid] + off[(int64_t)tid * nb + blockIdx.x] : 0;
  for (int i = tid; i <= nr; i += 256) rp[i] = rowptr[r0 + i];
  __syncthreads();
  const int64_t e_end = rp[nr];
  static_assert(kL1Rows <= 256, "rows of a block fit a byte");
  for (int64_t b0 = rp[0]; b0 < e_end; b0 += kL1Batch) {
    const int nbat = (int)((e_end - b0 < kL1Batch) ? e_end - b0 : kL1Batch);
    cnt[tid] = 0;  // nt <= 256 = blockDim
    if (tid < nr) {  // thread t labels the entries of row t that fall in this batch
      const int64_t a = rp[tid] > b0 ? rp[tid] : b0;
      const int64_t b = rp[tid + 1] < b0 + nbat ? rp[tid + 1] : b0 + nbat;
      for (int64_t e = a; e < b; ++e) rowof[e - b0] = (uint8_t)tid;
    }
    __syncthreads();
    CooRec rec[E];
    int slot[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int x = tid + 256 * q;
      if (x < nbat) {
        const int64_t e = b0 + x;
        rec[q].row = (int32_t)(r0 + rowof[x]);
        rec[q].col = col[e];
        rec[q].val = val[e];
      }
    }
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (tid + 256 * q < nbat) slot[q] = atomicAdd(&cnt[rec[q].col / tc], 1);
    __syncthreads();
    {  // exclusive scan of cnt over the tiles, one thread per tile
      const int v = cnt[tid];
      int x = v;
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      if (lane == 63) wsum[w] = x;
      __syncthreads();
      int before = 0;
      for (int i = 0; i < w; ++i) before += wsum[i];
      lst[tid] = before + x - v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (tid + 256 * q < nbat) sb[lst[rec[q].col / tc] + slot[q]] = rec[q];
    __syncthreads();
    for (int x = tid; x < nbat; x += 256) {
      const CooRec e = sb[x];
      const int t = e.col / tc;
      brec[cur[t] + (x - lst[t])] = e;
    }
    __syncthreads();
    cur[tid] += cnt[tid];  // cur has kMaxTiles = blockDim entries; tiles >= nt unused
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_l2_place(const int* base, int nt, int tc, int64_t cols, int64_t nnz,
                                                  const CooRec* brec, int64_t* cptr, int32_t* crow, double* cval) {
  __shared__ int cnt[kMaxTileCols];
  __shared__ int part[1024];
  const int t = blockIdx.x, tid = threadIdx.x;
  const int64_t j0 = (int64_t)t * tc;
  const int w = (int)((j0 + tc < cols) ? tc : cols - j0);
  const int64_t b0 = base[t];
  const int64_t b1 = (t + 1 < nt) ? base[t + 1] : nnz;
  for (int i = tid; i < w; i += 1024) cnt[i] = 0;
  __syncthreads();
  for (int64_t e = b0 + tid; e < b1; e += 1024) atomicAdd(&cnt[brec[e].col - j0], 1);
  __syncthreads();
  // exclusive scan of cnt[0 .. w) into cursors (chunk sums, then a scan of those)
  const int chunk = (w + 1023) / 1024;
  const int cb = tid * chunk, ce = (cb + chunk < w) ? cb + chunk : w;
  int sum = 0;
  for (int i = cb; i < ce; ++i) sum += cnt[i];
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = (int)b0 + (tid ? part[tid - 1] : 0);
  for (int i = cb; i < ce; ++i) {
    const int v = cnt[i];
    cnt[i] = run;  // now the cursor of column j0 + i
    cptr[j0 + i] = run;
    run += v;
  }
  if (t == nt - 1 && tid == 0) cptr[cols] = nnz;
  __syncthreads();
  for (int64_t e = b0 + tid; e < b1; e += 1024) {
    const CooRec x = brec[e];
    const int pos = atomicAdd(&cnt[x.col - j0], 1);
    crow[pos] = x.row;
    cval[pos] = x.val;
  }
}

// ---- sparse-map products in fixed point --------------------------------------
// A sparse logical map seen by its columns: the ELL form of its transpose
// (z entries of one magnitude per column: ell[c z + f] = row | sign << 31),
// or the CSR of its transpose (rp, ci, v).  Either way at most zmax entries
// per column (rows may repeat) and |v| <= vmax.
struct SpMapCols {
  const uint32_t* ell = nullptr;
  const uint16_t* ell16 = nullptr;  // the same entries in 16 bits, when the rows fit
  int z = 0;
  double mag = 1.0;
  const int64_t* rp = nullptr;
  const int32_t* ci = nullptr;
  const double* v = nullptr;
  double vmax = 1.0;
  int64_t zmax = 1;
};

__device__ inline double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  return x;
}

// ELL column c of M (z <= ZM entries): the ZM loads are issued together so
// one memory latency covers the whole column.
// E16: the map has the 16-bit copy (chosen at launch, so each kernel
// instance carries one load path); its entries stay packed two to a word in
// u[0 .. ZM/2) and ell_entry unpacks them at use.
template <int ZM, bool E16 = false>
__device__ inline void ell_load(const SpMapCols& M, int64_t c, uint32_t (&u)[ZM]) {
  if (E16 && M.z == ZM) {  // ZM / 8 aligned 16-byte words of 16-bit entries
    const uint4* src = reinterpret_cast<const uint4*>(M.ell16 + c * ZM);
#pragma unroll
    for (int w = 0; w < ZM / 8; ++w) {
      const uint4 x = src[w];
      u[4 * w] = x.x;
      u[4 * w + 1] = x.y;
      u[4 * w + 2] = x.z;
      u[4 * w + 3] = x.w;
    }
  } else if (!E16 && M.z == ZM) {  // the column is ZM / 4 aligned 16-byte words: one request each
    const uint4* src = reinterpret_cast<const uint4*>(M.ell + c * ZM);
#pragma unroll
    for (int w = 0; w < ZM / 4; ++w) {
      const uint4 x = src[w];
      u[4 * w] = x.x;
      u[4 * w + 1] = x.y;
      u[4 * w + 2] = x.z;
      u[4 * w + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int f = 0; f < ZM; ++f) u[f] = (f < M.z) ? M.ell[c * M.z + f] : 0u;
  }
}

// Entry f of a loaded column as row | sign << 31 (E16 only when the column
// was loaded packed, i.e. z == ZM).
template <int ZM, bool E16>
__device__ inline uint32_t ell_entry(const uint32_t (&u)[ZM], int z, int f) {
  if (E16 && z == ZM) {
    const uint32_t h = (u[f / 2] >> (16 * (f % 2))) & 0xffffu;
    return (h & 0x7fffu) | ((h & 0x8000u) << 16);
  }
  return u[f];
}

// Adds the fixed-point term V, signed by each entry, into the accumulators
// acc[p - p0] of the entries' rows p in [p0, p0 + pn).  Branch-free: an
// entry outside the range (or past z) adds into acc[dummy], a slot nobody
// reads, so every lane issues every atomic with 32-bit index arithmetic.
template <int ZM, bool E16>
__device__ inline void ell_add(const uint32_t (&u)[ZM], int z, long long V, int p0, int pn, int dummy,
                               unsigned long long* acc) {
  const unsigned long long vp = (unsigned long long)V, vn = (unsigned long long)(-V);
#pragma unroll
  for (int f = 0; f < ZM; ++f) {
    const uint32_t e = ell_entry<ZM, E16>(u, z, f);
    const int p = (int)(e & 0x7fffffffu) - p0;
    const bool ok = f < z && (unsigned)p < (unsigned)pn;
    atomicAdd(acc + (ok ? p : dummy), (e >> 31) ? vn : vp);
  }
}

// CSR form: term t times each entry of column c of M.
__device__ inline void csr_add(const SpMapCols& M, int64_t c, double t, const FxScale& fs, int64_t p0, int64_t pn,
                               unsigned long long* acc) {
  for (int64_t f = M.rp[c]; f < M.rp[c + 1]; ++f) {
    const int64_t p = (int64_t)M.ci[f] - p0;
    if (p >= 0 && p < pn) {
      fx_add_lds(acc + p, fx_term(t * M.v[f], fs));
    }
  }
}

// Left products over the CSC of A: out1 rows [p1, p1 + n1) of M1 diag(c) A
// and out2 rows [p2, p2 + n2) of M2 diag(c) A (n2 may be 0), with c_i =
// cw[i / window].  One wave per column j: a first pass takes the column's
// terms c_i a (up to kLeftCache per lane kept in registers) and their
// largest magnitude (the fixed-point scale of the column's accumulators); a
// second adds every term into the n1 + n2 accumulators of the wave in
// LDS, the map columns of all cached terms loaded before any is added.  ELL
// terms carry the map magnitude only at the end.  ZM: ELL entries per map
// column at most (both maps), 0 for maps given as CSR.  Four waves per
// block, each on its own columns with its own accumulators: a wave's LDS
// operations complete in issue order, so it needs no barrier between
// clearing, adding and reading back (only a compiler fence), and the waves
// of a block never wait for each other.
constexpr int kLeftCache = 4;

template <int ZM, bool E16>
__global__ __launch_bounds__(256) void k_left_fx(const int64_t* cptr, const int32_t* crow, const double* cval,
                                                int64_t ncols, const double* cw, int64_t window, SpMapCols M1,
                                                int64_t p1, int64_t n1, double* out1, int64_t ld1, SpMapCols M2,
                                                int64_t p2, int64_t n2, double* out2, int64_t ld2) {
  extern __shared__ unsigned long long fxacc[];  // 4 waves x (n1 + n2 + 1): the last slot of a wave is its dummy
  constexpr bool ELL = ZM > 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t na = n1 + n2;
  unsigned long long* acc = fxacc + (int64_t)wave * (na + 1);
  const int dummy1 = (int)na, dummy2 = (int)n2;  // acc[na], seen from acc and from acc + n1
  const uint32_t wdw = (uint32_t)window;
  for (int64_t j = 4 * (int64_t)blockIdx.x + wave; j < ncols; j += 4 * (int64_t)gridDim.x) {
    for (int64_t q = lane; q < na; q += 64) acc[q] = 0ull;
    asm volatile("" ::: "memory");
    FxScale fs;
    {
      const int64_t e0 = cptr[j], e1 = cptr[j + 1];
      double tq[kLeftCache];
      int32_t iq[kLeftCache];
      double mx = 0.0;
      bool bad = false;
#pragma unroll
      for (int q = 0; q < kLeftCache; ++q) {
        const int64_t e = e0 + lane + 64 * q;
        iq[q] = (e < e1) ? crow[e] : -1;
        tq[q] = (e < e1) ? cval[e] : 0.0;
      }
      double cq[kLeftCache];  // window weights: unconditional loads (cw[0] always exists), all in flight
#pragma unroll
      for (int q = 0; q < kLeftCache; ++q) cq[q] = cw[(uint32_t)(iq[q] > 0 ? iq[q] : 0) / wdw];
      // the map columns of the cached terms, in flight with the weights and
      // across the scale reduction below
      uint32_t u1[kLeftCache][ZM > 0 ? ZM : 1], u2[kLeftCache][ZM > 0 ? ZM : 1];
      if constexpr (ELL) {
#pragma unroll
        for (int q = 0; q < kLeftCache; ++q) {
          if (iq[q] >= 0) {
            ell_load<ZM, E16>(M1, iq[q], u1[q]);
            if (n2) ell_load<ZM, E16>(M2, iq[q], u2[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kLeftCache; ++q) {
        tq[q] *= cq[q];  // 0 for the lanes past the column's end
        mx = fmax(mx, fabs(tq[q]));
        bad |= isnan(tq[q]);
      }
      for (int64_t e = e0 + 64 * kLeftCache + lane; e < e1; e += 64) {
        const double t = cval[e] * cw[(uint32_t)crow[e] / wdw];
        mx = fmax(mx, fabs(t));
        bad |= isnan(t);
      }
      if (bad) mx = __builtin_inf();
      mx = wave_max(mx);
      // a target receives at most zmax terms per nonzero (entries of one map
      // column may repeat a row)
      const int64_t terms = (e1 - e0) * (M1.zmax > M2.zmax ? M1.zmax : M2.zmax);
      fs = fx_scale(ELL ? mx : mx * (M1.vmax > M2.vmax ? M1.vmax : M2.vmax), terms);
      if (fs.finite) {
        if constexpr (ELL) {
#pragma unroll
          for (int q = 0; q < kLeftCache; ++q) {
            if (iq[q] >= 0) {
              const long long V = fx_term(tq[q], fs);
              ell_add<ZM, E16>(u1[q], M1.z, V, (int)p1, (int)n1, dummy1, acc);
              if (n2) ell_add<ZM, E16>(u2[q], M2.z, V, (int)p2, (int)n2, dummy2, acc + n1);
            }
          }
          for (int64_t e = e0 + 64 * kLeftCache + lane; e < e1; e += 64) {
            const int32_t i = crow[e];
            const double t = cval[e] * cw[(uint32_t)i / wdw];
            uint32_t v1[ZM], v2[ZM];
            ell_load<ZM, E16>(M1, i, v1);
            if (n2) ell_load<ZM, E16>(M2, i, v2);
            const long long V = fx_term(t, fs);
            ell_add<ZM, E16>(v1, M1.z, V, (int)p1, (int)n1, dummy1, acc);
            if (n2) ell_add<ZM, E16>(v2, M2.z, V, (int)p2, (int)n2, dummy2, acc + n1);
          }
        } else {
#pragma unroll
          for (int q = 0; q < kLeftCache; ++q) {
            if (iq[q] >= 0) {
              csr_add(M1, iq[q], tq[q], fs, p1, n1, acc);
              if (n2) csr_add(M2, iq[q], tq[q], fs, p2, n2, acc + n1);
            }
          }
          for (int64_t e = e0 + 64 * kLeftCache + lane; e < e1; e += 64) {
            const int32_t i = crow[e];
            const double t = cval[e] * cw[(uint32_t)i / wdw];
            csr_add(M1, i, t, fs, p1, n1, acc);
            if (n2) csr_add(M2, i, t, fs, p2, n2, acc + n1);
          }
        }
      }
    }
    asm volatile("" ::: "memory");  // the wave's atomics precede its reads (LDS order within a wave)
    for (int64_t q = lane; q < n1; q += 64)
      out1[(p1 + q) + j * ld1] = (ELL ? M1.mag : 1.0) * fx_value((long long)acc[q], fs);
    for (int64_t q = lane; q < n2; q += 64)
      out2[(p2 + q) + j * ld2] = (ELL ? M2.mag : 1.0) * fx_value((long long)acc[n1 + q], fs);
    asm volatile("" ::: "memory");  // read back before the next column clears the accumulators
  }
}

// Y(i, p) = eta * sum over row i's nonzeros (j, a) of a * M(p, j) for p in
// [p0, p0 + pn), pn <= kRightSlots.  One block per kRightRows rows; each
// thread takes the block's nonzeros tid, tid + 256, ... (kRightCache of them
// held in registers, with their map columns loaded at once), the rows'
// largest |a| (their fixed-point scales) by integer atomicMax on the IEEE
// bits, then adds every term into acc[slot][row]; the rows are written
// out coalesced (consecutive threads, consecutive rows).  ZM as for
// k_left_fx.
constexpr int kRightRows = 32;
constexpr int kRightSlots = 128;
constexpr int kRightCache = 3;

template <int ZM, bool E16>
__global__ __launch_bounds__(256) void k_right_fx(const int64_t* rowptr, const int32_t* col, const double* val,
                                                 int64_t rows, SpMapCols M, int64_t p0, int64_t pn, double eta,
                                                 double* Y, int64_t ldy) {
  // kRightRows x ps accumulators, row-major with an odd stride: the lanes of
  // a wave mostly work on one row, so their slots p fall in distinct banks
  // (a slot-major layout put every lane of a row in the same bank)
  extern __shared__ unsigned long long fxacc[];
  const int64_t ps = pn | 1;
  const int dummy = kRightRows * (int)ps;  // one slot past the rows, never read
  __shared__ int64_t rp[kRightRows + 1];
  __shared__ unsigned long long rmax[kRightRows];  // bits of the row's largest |a| (non-negative doubles order as integers)
  __shared__ FxScale rs[kRightRows];
  __shared__ uint8_t rowof[256 * kRightCache];  // row of each cached entry
  constexpr bool ELL = ZM > 0;
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kRightRows;
  const int nr = (int)((r0 + kRightRows < rows) ? kRightRows : rows - r0);
  for (int q = tid; q < ps * kRightRows; q += 256) fxacc[q] = 0ull;
  for (int i = tid; i <= nr; i += 256) rp[i] = rowptr[r0 + i];
  if (tid < kRightRows) rmax[tid] = 0ull;
  __syncthreads();
  const int64_t e_end = rp[nr];
  if (tid < nr) {  // thread t labels the cached entries of row t
    const int64_t c1 = rp[0] + 256 * kRightCache;
    const int64_t b = rp[tid + 1] < c1 ? rp[tid + 1] : c1;
    for (int64_t e = rp[tid]; e < b; ++e) rowof[e - rp[0]] = (uint8_t)tid;
  }
  __syncthreads();
  int64_t jq[kRightCache];
  double tq[kRightCache];
  int rq[kRightCache];
#pragma unroll
  for (int q = 0; q < kRightCache; ++q) {
    const int64_t e = rp[0] + tid + 256 * q;
    if (e < e_end) {
      jq[q] = col[e];
      tq[q] = val[e];
      rq[q] = rowof[tid + 256 * q];
    } else {
      jq[q] = -1;
      tq[q] = 0.0;
      rq[q] = 0;
    }
  }
  uint32_t u[kRightCache][ZM > 0 ? ZM : 1];
  if constexpr (ELL) {
#pragma unroll
    for (int q = 0; q < kRightCache; ++q)
      if (jq[q] >= 0) ell_load<ZM, E16>(M, jq[q], u[q]);
  }
#pragma unroll
  for (int q = 0; q < kRightCache; ++q)
    if (jq[q] >= 0) atomicMax(&rmax[rq[q]], (unsigned long long)__double_as_longlong(fabs(tq[q])));  // NaN orders above inf
  int r = rq[kRightCache - 1];  // row of the thread's last cached entry (rows past it follow in order)
  {
    int rr = r;
    for (int64_t e = rp[0] + tid + 256 * kRightCache; e < e_end; e += 256) {
      while (rp[rr + 1] <= e) ++rr;
      atomicMax(&rmax[rr], (unsigned long long)__double_as_longlong(fabs(val[e])));
    }
  }
  __syncthreads();
  if (tid < nr) {
    const double mx = __longlong_as_double((long long)rmax[tid]);
    const int64_t cnt = rp[tid + 1] - rp[tid];
    rs[tid] = fx_scale(ELL ? mx : mx * M.vmax, cnt * M.zmax);
  }
  __syncthreads();
  // adds the term t of row rr (nonzero in column j, map column entries uu)
  auto add_ell = [&](const uint32_t (&uu)[ZM > 0 ? ZM : 1], double t, int rr) {
    const FxScale fs = rs[rr];
    if (!fs.finite) return;
    const long long V = fx_term(t, fs);
    ell_add<(ZM > 0 ? ZM : 1), E16>(uu, M.z, V, (int)p0, (int)pn, dummy - rr * (int)ps, fxacc + rr * ps);
  };
  auto add_csr = [&](int64_t j, double t, int rr) {
    const FxScale fs = rs[rr];
    if (!fs.finite) return;
    for (int64_t f = M.rp[j]; f < M.rp[j + 1]; ++f) {
      const int64_t p = (int64_t)M.ci[f] - p0;
      if (p >= 0 && p < pn) {
        fx_add_lds(fxacc + rr * ps + p, fx_term(t * M.v[f], fs));
      }
    }
  };
#pragma unroll
  for (int q = 0; q < kRightCache; ++q) {
    if (jq[q] < 0) continue;
    if constexpr (ELL) add_ell(u[q], tq[q], rq[q]);
    else add_csr(jq[q], tq[q], rq[q]);
  }
  for (int64_t e = rp[0] + tid + 256 * kRightCache; e < e_end; e += 256) {  // rows denser than the cache
    while (rp[r + 1] <= e) ++r;
    if constexpr (ELL) {
      uint32_t uu[ZM > 0 ? ZM : 1];
      ell_load<ZM, E16>(M, col[e], uu);
      add_ell(uu, val[e], r);
    } else {
      add_csr(col[e], val[e], r);
    }
  }
  __syncthreads();
  const double scale = eta * (ELL ? M.mag : 1.0);
  for (int q = tid; q < pn * kRightRows; q += 256) {
    const int rr = q % kRightRows;
    const int64_t p = q / kRightRows;
    if (rr < nr)
      Y[(r0 + rr) + (p0 + p) * ldy] = scale * fx_value((long long)fxacc[rr * ps + p], rs[rr]);
  }
}

// out(p, j) = sum over column j's nonzeros (i, a) of a cw[i / window] M(p, i),
// M dense (P x m logical, column-major, ld ldm); one wave per column, lanes
// own rows p (PR per lane).
template <int PR>
__global__ __launch_bounds__(256) void k_left_all_dense(const int64_t* cptr, const int32_t* crow, const double* cval,
                                                        int64_t ncols, const double* M, int64_t ldm, int64_t P,
                                                        const double* cw, int64_t window, double* out, int64_t ldo) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= ncols) return;
  double acc[PR];
#pragma unroll
  for (int r = 0; r < PR; ++r) acc[r] = 0.0;
  for (int64_t e = cptr[j]; e < cptr[j + 1]; ++e) {
    const int64_t i = crow[e];
    const double a = cval[e] * cw[(uint32_t)i / (uint32_t)window];
    const double* mcol = M + i * ldm;
#pragma unroll
    for (int r = 0; r < PR; ++r) {
      const int64_t p = lane + 64 * r;
      if (p < P) acc[r] += a * mcol[p];
    }
  }
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    const int64_t p = lane + 64 * r;
    if (p < P) out[p + j * ldo] = acc[r];
  }
}

// Y(i, p) = nu Y + eta * sum over row i's nonzeros (j, a) of a M(p, j), M dense
// (P x n logical, ld ldm); one wave per row, lanes own p.
template <int PR>
__global__ __launch_bounds__(256) void k_right_all_dense(const int64_t* rowptr, const int32_t* col, const double* val,
                                                         int64_t rows, const double* M, int64_t ldm, int64_t P,
                                                         double eta, double nu, double* Y, int64_t ldy) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= rows) return;
  double acc[PR];
#pragma unroll
  for (int r = 0; r < PR; ++r) acc[r] = 0.0;
  for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
    const double a = val[e];
    const double* mcol = M + (int64_t)col[e] * ldm;
#pragma unroll
    for (int r = 0; r < PR; ++r) {
      const int64_t p = lane + 64 * r;
      if (p < P) acc[r] += a * mcol[p];
    }
  }
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    const int64_t p = lane + 64 * r;
    if (p < P) {
      double* y = Y + i + p * ldy;
      *y = (nu != 0.0) ? nu * *y + eta * acc[r] : eta * acc[r];
    }
  }
}

unsigned nblocks(int64_t work, int64_t per) { return (unsigned)std::max<int64_t>(1, cdiv(work, per)); }

SpMapCols map_cols(const Map& M) {
  SpMapCols c;
  if (M.ell_zeta > 0) {
    c.ell = M.ell.as<uint32_t>();
    c.ell16 = M.ell16.p ? M.ell16.as<uint16_t>() : nullptr;
    c.z = M.ell_zeta;
    c.mag = M.ell_mag;
    c.vmax = M.ell_mag;
    c.zmax = M.ell_zeta;
  } else {
    const Csr& t = M.logical_csr_t();
    c.rp = t.rp();
    c.ci = t.ci();
    c.v = t.v();
    c.vmax = M.sp_vmax;
    c.zmax = M.sp_zmax > 0 ? M.sp_zmax : 1;
  }
  return c;
}

// Accumulators per wave of k_left_fx (4 waves: 16 KB of LDS).
constexpr int64_t kLeftSlots = 512;

// ELL entries per map column the kernels are compiled for (0: CSR maps).
int zm_class(int z) { return z <= 0 ? 0 : (z <= 8 ? 8 : 16); }

void launch_left_fx(hipStream_t st, int zm, const CscWindow& H, const double* cw, int64_t window,
                    const SpMapCols& M1, int64_t p1, int64_t n1, double* o1, int64_t ld1, const SpMapCols& M2,
                    int64_t p2, int64_t n2, double* o2, int64_t ld2) {
  const int64_t groups = cdiv(H.cols, 4);
  const unsigned grid = (unsigned)std::min<int64_t>(groups, 8192);
  const size_t lds = (size_t)4 * (n1 + n2 + 1) * sizeof(unsigned long long);
#define SKEMA_LEFT_FX(ZM, E16)                                                                                    \
  hipLaunchKernelGGL((k_left_fx<ZM, E16>), dim3(grid), dim3(256), lds, st, H.ptr, H.row, H.val, H.cols, cw, window, \
                     M1, p1, n1, o1, ld1, M2, p2, n2, o2, ld2)
  // the 32-bit entries here: the packed ones cost this kernel registers
  // (its cached map columns) below the four waves per SIMD it runs at
  if (zm == 8) SKEMA_LEFT_FX(8, false);
  else if (zm == 16) SKEMA_LEFT_FX(16, false);
  else SKEMA_LEFT_FX(0, false);
#undef SKEMA_LEFT_FX
  check_launch("k_left_fx");
}

// X = M1 diag(c) A (k1 rows) and W = M2 diag(c) A (k2 rows): one pass over the
// CSC when both maps have the same form and their rows fit one wave's
// accumulators, otherwise one pass per map and chunk of rows.
void left_sparse_fx(hipStream_t st, const CscWindow& H, const Map& Ma, int64_t ka, double* oa, int64_t lda,
                    const Map& Mb, int64_t kb, double* ob, int64_t ldb, const double* cw, int64_t window) {
  if (H.cols == 0) return;
  const SpMapCols ca = map_cols(Ma), cb = map_cols(Mb);
  const int za = zm_class(Ma.ell_zeta), zb = zm_class(Mb.ell_zeta);
  if ((za > 0) == (zb > 0) && ka + kb <= kLeftSlots) {
    launch_left_fx(st, std::max(za, zb), H, cw, window, ca, 0, ka, oa, lda, cb, 0, kb, ob, ldb);
    return;
  }
  for (int64_t p0 = 0; p0 < ka; p0 += kLeftSlots)
    launch_left_fx(st, za, H, cw, window, ca, p0, std::min(kLeftSlots, ka - p0), oa, lda, ca, 0, 0, nullptr, 1);
  for (int64_t p0 = 0; p0 < kb; p0 += kLeftSlots)
    launch_left_fx(st, zb, H, cw, window, cb, p0, std::min(kLeftSlots, kb - p0), ob, ldb, cb, 0, 0, nullptr, 1);
}

// Y = eta A M^T by row blocks, in chunks of kRightSlots map rows.
void right_sparse_fx(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t m,
                     const Map& M, int64_t k, double eta, double* Y, int64_t ldy) {
  const SpMapCols c = map_cols(M);
  const dim3 grid((unsigned)cdiv(m, kRightRows));
  for (int64_t p0 = 0; p0 < k; p0 += kRightSlots) {
    const int64_t pn = std::min<int64_t>(kRightSlots, k - p0);
    const size_t lds = ((size_t)(pn | 1) * kRightRows + 1) * sizeof(unsigned long long);
    const int zm = zm_class(M.ell_zeta);
#define SKEMA_RIGHT_FX(ZM, E16) \
  hipLaunchKernelGGL((k_right_fx<ZM, E16>), grid, dim3(256), lds, st, rowptr, col, val, m, c, p0, pn, eta, Y, ldy)
    if (zm == 8 && c.ell16) SKEMA_RIGHT_FX(8, true);
    else if (zm == 8) SKEMA_RIGHT_FX(8, false);
    else if (zm == 16 && c.ell16) SKEMA_RIGHT_FX(16, true);
    else if (zm == 16) SKEMA_RIGHT_FX(16, false);
    else SKEMA_RIGHT_FX(0, false);
#undef SKEMA_RIGHT_FX
    check_launch("k_right_fx");
  }
}

// CSC of the whole CSR by the two bucketing levels (entries of a column in no
// fixed order); false when a column tile would exceed kMaxTileCols.
bool csc_by_buckets(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t m,
                    int64_t n, int64_t nnz, DevBuf& scratch, CscWindow& H) {
  int64_t nt = std::min<int64_t>(kMaxTiles, n);
  const int64_t tc = cdiv(n, nt);
  nt = cdiv(n, tc);
  if (tc > kMaxTileCols || nnz <= 0 || nnz >= ((int64_t)1 << 31)) return false;  // int positions
  const int64_t nb = cdiv(m, kL1Rows);
  SKEMA_REQUIRE(nt * nb < ((int64_t)1 << 31), kErrShape, "linear_update_csr: too many row blocks");
  const size_t o_off = align_up((size_t)nt * nb * sizeof(int), 256);
  const size_t o_tot = o_off + align_up((size_t)nt * nb * sizeof(int), 256);
  const size_t o_base = o_tot + align_up((size_t)nt * sizeof(int), 256);
  const size_t o_brec = o_base + align_up((size_t)(nt + 1) * sizeof(int), 256);
  const size_t o_ptr = o_brec + align_up((size_t)nnz * sizeof(CooRec), 256);
  const size_t o_row = o_ptr + align_up((size_t)(n + 1) * sizeof(int64_t), 256);
  const size_t o_val = o_row + align_up((size_t)nnz * sizeof(int32_t), 256);
  scratch.reserve(o_val + (size_t)nnz * sizeof(double));
  char* b = scratch.as<char>();
  int* cnt = reinterpret_cast<int*>(b);
  int* off = reinterpret_cast<int*>(b + o_off);
  int* tot = reinterpret_cast<int*>(b + o_tot);
  int* base = reinterpret_cast<int*>(b + o_base);
  CooRec* brec = reinterpret_cast<CooRec*>(b + o_brec);
  H.rows = m;
  H.cols = n;
  H.nnz = nnz;
  H.ptr = reinterpret_cast<int64_t*>(b + o_ptr);
  H.row = reinterpret_cast<int32_t*>(b + o_row);
  H.val = reinterpret_cast<double*>(b + o_val);
  hipLaunchKernelGGL(k_l1_count, dim3((unsigned)nb), dim3(256), 0, st, rowptr, m, col, (int)tc, (int)nt, (int)nb, cnt);
  hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nt), dim3(1024), 0, st, cnt, (int)nb, off, tot);
  hipLaunchKernelGGL(k_tile_base, dim3(1), dim3(1024), 0, st, tot, (int)nt, base);
  hipLaunchKernelGGL(k_l1_scatter, dim3((unsigned)nb), dim3(256), 0, st, rowptr, m, col, val, (int)tc, (int)nt,
                     (int)nb, off, base, brec);
  hipLaunchKernelGGL(k_l2_place, dim3((unsigned)nt), dim3(1024), 0, st, base, (int)nt, (int)tc, n, nnz, brec, H.ptr,
                     H.row, H.val);
  check_launch("csc_by_buckets");
  return true;
}

void left_dense(hipStream_t st, const CscWindow& H, const double* M, int64_t ldm, int64_t P, const double* cw,
                int64_t window, double* out, int64_t ldo) {
  for (int64_t p0 = 0; p0 < P; p0 += 512) {
    const int64_t pc = std::min<int64_t>(512, P - p0);
    const dim3 grid(nblocks(H.cols, 4));
    const double* Mp = M + p0;
    double* op = out + p0;
    const int pr = (int)cdiv(pc, 64);
    if (pr <= 2)
      hipLaunchKernelGGL(k_left_all_dense<2>, grid, dim3(256), 0, st, H.ptr, H.row, H.val, H.cols, Mp, ldm, pc, cw,
                         window, op, ldo);
    else if (pr <= 4)
      hipLaunchKernelGGL(k_left_all_dense<4>, grid, dim3(256), 0, st, H.ptr, H.row, H.val, H.cols, Mp, ldm, pc, cw,
                         window, op, ldo);
    else
      hipLaunchKernelGGL(k_left_all_dense<8>, grid, dim3(256), 0, st, H.ptr, H.row, H.val, H.cols, Mp, ldm, pc, cw,
                         window, op, ldo);
    check_launch("k_left_all_dense");
  }
}

}  // namespace

void update_csr_all(Sketch& sk, const int64_t* rowptr, const int32_t* col, const double* val, const int64_t* ends) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  // window partition of for_each_window: windows of p.window rows (the last
  // one shorter), or one window when window == m
  // rows and windows fit 32 bits (CSR column indices are int32 and the
  // window lookups below divide in 32 bits)
  SKEMA_REQUIRE(m < ((int64_t)1 << 31), kErrShape, "linear_update_csr: more than 2^31 rows");
  const int64_t window = (sk.p.window >= m) ? m : sk.p.window;
  const int64_t W = cdiv(m, window);
  std::vector<double> hw((size_t)W);
  double f = eta;
  for (int64_t w = W - 1; w >= 0; --w) {  // c_w = eta nu^(W-1-w), by repeated products like the loop
    hw[(size_t)w] = f;
    f *= nu;
  }
  DevBuf& scr = c->scratch;
  CscWindow H;
  double* cw = nullptr;
  double* Wm = nullptr;
  {
    Timed t(c, "window");
    scr.reserve(align_up((size_t)W * sizeof(double), 256) + align_up((size_t)s * n * sizeof(double), 256) + 256);
    cw = scr.as<double>();
    Wm = reinterpret_cast<double*>(scr.as<char>() + align_up((size_t)W * sizeof(double), 256));
    int* flag = reinterpret_cast<int*>(scr.as<char>() + align_up((size_t)W * sizeof(double), 256) +
                                       align_up((size_t)s * n * sizeof(double), 256));
    SKEMA_HIP(hipMemcpyAsync(cw, hw.data(), (size_t)W * sizeof(double), hipMemcpyHostToDevice, st));
    bool done = false;
    if (sk.sparse_maps) {
      // the fixed-point products take the column entries in any order
      const int64_t nnz = validate_csr_window(st, rowptr, col, m, n, flag, ends);
      done = csc_by_buckets(st, rowptr, col, val, m, n, nnz, sk.win, H);
    }
    if (!done) H = csr_window_to_csc(st, rowptr, col, val, m, n, sk.win, /*sorted=*/true);
  }
  double* X = sk.X.as<double>();
  double* Y = sk.Y.as<double>();
  double* Z = sk.Z.as<double>();
  if (sk.sparse_maps) {
    {
      Timed t(c, "sketch_left");
      left_sparse_fx(st, H, *sk.maps[0], k, X, k, *sk.maps[2], s, Wm, s, cw, window);
    }
    {
      Timed t(c, "sketch_right");
      right_sparse_fx(st, rowptr, col, val, m, *sk.maps[1], k, eta, Y, m);
    }
    {
      Timed t(c, "sketch_core");
      GemmOut o;
      o.C1 = Z; o.ldc1 = s; o.alpha1 = 1.0; o.beta1 = 0.0; o.nsplit = s;
      o.C2 = Z; o.ldc2 = s;
      spmm_dense_csrT(st, s, Wm, s, sk.maps[3]->logical_csr(), o, &c->gws.buf);
    }
    return;
  }
  {
    Timed t(c, "sketch_left");
    left_dense(st, H, sk.upsL, sk.ldups, k, cw, window, X, k);
    left_dense(st, H, sk.phiL, sk.ldphi, s, cw, window, Wm, s);
  }
  {
    Timed t(c, "sketch_right");
    const double* Om = sk.omps.as<double>();  // Omega = rows 0..k-1 of [Omega; Psi], ld k+s
    for (int64_t p0 = 0; p0 < k; p0 += 512) {
      const int64_t pc = std::min<int64_t>(512, k - p0);
      const dim3 grid(nblocks(m, 4));
      const int pr = (int)cdiv(pc, 64);
      if (pr <= 2)
        hipLaunchKernelGGL(k_right_all_dense<2>, grid, dim3(256), 0, st, rowptr, col, val, m, Om + p0, k + s, pc, eta,
                           0.0, Y + p0 * m, m);
      else if (pr <= 4)
        hipLaunchKernelGGL(k_right_all_dense<4>, grid, dim3(256), 0, st, rowptr, col, val, m, Om + p0, k + s, pc, eta,
                           0.0, Y + p0 * m, m);
      else
        hipLaunchKernelGGL(k_right_all_dense<8>, grid, dim3(256), 0, st, rowptr, col, val, m, Om + p0, k + s, pc, eta,
                           0.0, Y + p0 * m, m);
      check_launch("k_right_all_dense");
    }
  }
  {
    Timed t(c, "sketch_core");
    // Z = W Psi^T, Psi = rows k..k+s-1 of [Omega; Psi] (ld k+s)
    gemm_f64(st, 'N', 'T', s, s, n, 1.0, Wm, s, sk.omps.as<double>() + k, k + s, 0.0, Z, s, &c->gws);
  }
}

}  // namespace skema
