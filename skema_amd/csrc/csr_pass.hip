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
//   left  (X and W = Phi diag(c) A): sparse maps -> one 64-thread block per
//         column j of A (CSC), lanes over its nonzeros, FP64 LDS atomics
//         into the column's accumulators; dense maps -> one wave per column,
//         lanes own map rows, nonzeros in ascending row order
//         (deterministic).
//   right (Y): sparse maps -> one block per 64 rows of A, lanes over their
//         nonzeros, LDS atomics; dense maps -> one wave per row, lanes own
//         map rows (deterministic).
// The sparse-map products are exact to rounding but, through the atomics,
// not bitwise reproducible from run to run.
//   core  Z = W Psi^T: dense x CSR^T gather (sparse Psi) or the FP64 GEMM.
#include <vector>

#include "csr_pass.hpp"
#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kPad = 65;  // LDS accumulator stride (doubles) per map row: lanes spread over banks

// out(p, j) for p in [p0, p0 + P): sum over column j's nonzeros (i, a) of
// a * cw[i / window] * Mt(i, p), Mt = CSR of the map's transpose.  One
// 256-thread block per column: threads take the column's nonzeros, walk their
// map entries, and add into the column's accumulators in LDS (FP64 LDS
// atomics: the order of the additions, and so the last bits, may vary from
// run to run); the column is then written contiguously.
__global__ __launch_bounds__(256) void k_left_all_sparse(const int64_t* cptr, const int32_t* crow, const double* cval,
                                                        const int64_t* mrp, const int32_t* mci, const double* mv,
                                                        const double* cw, int64_t window, int64_t p0, int64_t P,
                                                        double* out, int64_t ldo) {
  extern __shared__ double acc[];  // P doubles
  const int tid = threadIdx.x;
  const int64_t j = blockIdx.x;
  for (int64_t p = tid; p < P; p += 256) acc[p] = 0.0;
  __syncthreads();
  for (int64_t e = cptr[j] + tid; e < cptr[j + 1]; e += 256) {
    const int64_t i = crow[e];
    const double a = cval[e] * cw[(uint32_t)i / (uint32_t)window];
    for (int64_t f = mrp[i]; f < mrp[i + 1]; ++f) {
      const int64_t p = (int64_t)mci[f] - p0;
      if (p >= 0 && p < P) atomicAdd(&acc[p], a * mv[f]);
    }
  }
  __syncthreads();
  for (int64_t p = tid; p < P; p += 256) out[(p0 + p) + j * ldo] = acc[p];
}

// The z entries of ELL row c: with ZC = 8 (the sparse-sign default) two
// 16-byte loads issued together, otherwise one dword per entry.
template <int ZC>
__device__ __forceinline__ void ell_row(const uint32_t* ell, int64_t c, int z, uint32_t (&u)[16]) {
  if constexpr (ZC == 8) {
    const uint4 a = *reinterpret_cast<const uint4*>(ell + c * 8);
    const uint4 b = *reinterpret_cast<const uint4*>(ell + c * 8 + 4);
    u[0] = a.x; u[1] = a.y; u[2] = a.z; u[3] = a.w;
    u[4] = b.x; u[5] = b.y; u[6] = b.z; u[7] = b.w;
  } else {
    for (int f = 0; f < z; ++f) u[f] = ell[c * z + f];
  }
}

// The same with the map's transpose in ELL form (z entries per row,
// u = row | sign << 31, magnitude mag): one contiguous 4z-byte read per
// nonzero.  One wave per column (four per block, each with its own LDS
// accumulator of P doubles): a column holds a few hundred nonzeros, so a
// block per column would leave most of its lanes idle.
template <int ZC>
__global__ __launch_bounds__(256) void k_left_all_ell(const int64_t* cptr, const int32_t* crow, const double* cval,
                                                      int64_t ncols, const uint32_t* ell, int z, double mag,
                                                      const double* cw, int64_t window, int64_t p0, int64_t P,
                                                      double* out, int64_t ldo) {
  extern __shared__ double acc_all[];  // 4 x P doubles
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t j = (int64_t)blockIdx.x * 4 + wave;
  double* acc = acc_all + wave * P;
  for (int64_t p = lane; p < P; p += 64) acc[p] = 0.0;
  if (j < ncols) {
    for (int64_t e = cptr[j] + lane; e < cptr[j + 1]; e += 64) {
      const int64_t i = crow[e];
      const double a = cval[e] * cw[(uint32_t)i / (uint32_t)window] * mag;
      uint32_t uu[16];
      ell_row<ZC>(ell, i, z, uu);
      const int zz = ZC ? ZC : z;
#pragma unroll
      for (int f = 0; f < (ZC ? ZC : 16); ++f) {
        if (f >= zz) break;
        const uint32_t u = uu[f];
        const int64_t p = (int64_t)(u & 0x7fffffffu) - p0;
        if (p >= 0 && p < P) atomicAdd(&acc[p], (u >> 31) ? -a : a);
      }
    }
  }
  __syncthreads();  // the wave's LDS atomics are complete before it reads back
  if (j < ncols)
    for (int64_t p = lane; p < P; p += 64) out[(p0 + p) + j * ldo] = acc[p];
}

// Y(i, p) = nu Y(i, p) + eta * sum over row i's nonzeros (j, a) of a * Mt(j, p)
// for p in [p0, p0 + P), P <= PCH.  One 256-thread block per 64 rows:
// threads take the block's nonzeros and add into acc[p][row] (LDS atomics),
// then the first wave writes the rows, consecutive lanes consecutive rows.
template <int PCH>
__global__ __launch_bounds__(256) void k_right_all_sparse(const int64_t* rowptr, const int32_t* col, const double* val,
                                                         int64_t rows, const int64_t* mrp, const int32_t* mci,
                                                         const double* mv, int64_t p0, int64_t P, double eta,
                                                         double nu, double* Y, int64_t ldy) {
  __shared__ double acc[PCH * kPad];
  const int tid = threadIdx.x, lane = tid & 63;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int64_t r1 = (r0 + 64 < rows) ? r0 + 64 : rows;
  for (int e = tid; e < PCH * kPad; e += 256) acc[e] = 0.0;
  __syncthreads();
  const int64_t e0 = rowptr[r0], e1 = rowptr[r1];
  int64_t r = r0;  // row of the thread's current nonzero (rows advance monotonically)
  for (int64_t e = e0 + tid; e < e1; e += 256) {
    while (rowptr[r + 1] <= e) ++r;
    const int64_t jc = col[e];
    const double a = val[e];
    for (int64_t f = mrp[jc]; f < mrp[jc + 1]; ++f) {
      const int64_t p = (int64_t)mci[f] - p0;
      if (p >= 0 && p < P) atomicAdd(&acc[p * kPad + (r - r0)], a * mv[f]);
    }
  }
  __syncthreads();
  const int64_t i = r0 + lane;
  if (tid < 64 && i < rows) {
    for (int64_t p = 0; p < P; ++p) {
      double* y = Y + i + (p0 + p) * ldy;
      *y = (nu != 0.0) ? nu * *y + eta * acc[p * kPad + lane] : eta * acc[p * kPad + lane];
    }
  }
}

template <int PCH, int ZC>
__global__ __launch_bounds__(256) void k_right_all_ell(const int64_t* rowptr, const int32_t* col, const double* val,
                                                       int64_t rows, const uint32_t* ell, int z, double mag,
                                                       int64_t p0, int64_t P, double eta, double nu, double* Y,
                                                       int64_t ldy) {
  __shared__ double acc[PCH * kPad];
  __shared__ int64_t rp[65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int64_t r1 = (r0 + 64 < rows) ? r0 + 64 : rows;
  for (int e = tid; e < PCH * kPad; e += 256) acc[e] = 0.0;
  if (tid <= r1 - r0) rp[tid] = rowptr[r0 + tid];
  __syncthreads();
  const int64_t e0 = rp[0], e1 = rp[r1 - r0];
  int r = 0;  // block-relative row of the thread's current nonzero (advances monotonically, LDS reads)
  for (int64_t e = e0 + tid; e < e1; e += 256) {
    while (rp[r + 1] <= e) ++r;
    const double a = val[e] * mag;
    uint32_t uu[16];
    ell_row<ZC>(ell, (int64_t)col[e], z, uu);
    const int zz = ZC ? ZC : z;
#pragma unroll
    for (int f = 0; f < (ZC ? ZC : 16); ++f) {
      if (f >= zz) break;
      const uint32_t u = uu[f];
      const int64_t p = (int64_t)(u & 0x7fffffffu) - p0;
      if (p >= 0 && p < P) atomicAdd(&acc[p * kPad + r], (u >> 31) ? -a : a);
    }
  }
  __syncthreads();
  // wave w writes columns p = w, w + 4, ...: 64 consecutive rows per store
  const int64_t i = r0 + lane;
  if (i < rows) {
    for (int64_t p = wave; p < P; p += 4) {
      double* y = Y + i + (p0 + p) * ldy;
      *y = (nu != 0.0) ? nu * *y + eta * acc[p * kPad + lane] : eta * acc[p * kPad + lane];
    }
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

// Sparse left product over map rows [0, P) in LDS chunks of up to 4096 rows
// (ELL form of the map's transpose when it has one).
void left_sparse(hipStream_t st, const CscWindow& H, const Map& M, int64_t P, const double* cw, int64_t window,
                 double* out, int64_t ldo) {
  if (H.cols == 0) return;
  const Csr& Mt = M.logical_csr_t();
  const int64_t chunk = M.ell_zeta > 0 ? 1024 : 4096;
  for (int64_t p0 = 0; p0 < P; p0 += chunk) {
    const int64_t pc = std::min<int64_t>(chunk, P - p0);
    if (M.ell_zeta == 8)
      hipLaunchKernelGGL(k_left_all_ell<8>, dim3(nblocks(H.cols, 4)), dim3(256), (size_t)4 * pc * sizeof(double), st,
                         H.ptr, H.row, H.val, H.cols, M.ell.as<uint32_t>(), M.ell_zeta, M.ell_mag, cw, window, p0, pc,
                         out, ldo);
    else if (M.ell_zeta > 0)
      hipLaunchKernelGGL(k_left_all_ell<0>, dim3(nblocks(H.cols, 4)), dim3(256), (size_t)4 * pc * sizeof(double), st,
                         H.ptr, H.row, H.val, H.cols, M.ell.as<uint32_t>(), M.ell_zeta, M.ell_mag, cw, window, p0, pc,
                         out, ldo);
    else
      hipLaunchKernelGGL(k_left_all_sparse, dim3((unsigned)H.cols), dim3(256), (size_t)pc * sizeof(double), st,
                         H.ptr, H.row, H.val, Mt.rp(), Mt.ci(), Mt.v(), cw, window, p0, pc, out, ldo);
    check_launch("k_left_all_sparse");
  }
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

void update_csr_all(Sketch& sk, const int64_t* rowptr, const int32_t* col, const double* val) {
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
    // the sparse-map left products accumulate with atomics: no column sort
    H = csr_window_to_csc(st, rowptr, col, val, m, n, sk.win, /*sorted=*/!sk.sparse_maps);
    scr.reserve(align_up((size_t)W * sizeof(double), 256) + (size_t)s * n * sizeof(double));
    cw = scr.as<double>();
    Wm = reinterpret_cast<double*>(scr.as<char>() + align_up((size_t)W * sizeof(double), 256));
    SKEMA_HIP(hipMemcpyAsync(cw, hw.data(), (size_t)W * sizeof(double), hipMemcpyHostToDevice, st));
  }
  double* X = sk.X.as<double>();
  double* Y = sk.Y.as<double>();
  double* Z = sk.Z.as<double>();
  if (sk.sparse_maps) {
    {
      Timed t(c, "sketch_left");
      left_sparse(st, H, *sk.maps[0], k, cw, window, X, k);
      left_sparse(st, H, *sk.maps[2], s, cw, window, Wm, s);
    }
    {
      Timed t(c, "sketch_right");
      const Map& Om = *sk.maps[1];
      const Csr& Ot = Om.logical_csr_t();
      for (int64_t p0 = 0; p0 < k; p0 += 128) {
        const int64_t pc = std::min<int64_t>(128, k - p0);
        const dim3 grid(nblocks(m, 64));
        if (Om.ell_zeta == 8 && pc <= 64)
          hipLaunchKernelGGL((k_right_all_ell<64, 8>), grid, dim3(256), 0, st, rowptr, col, val, m,
                             Om.ell.as<uint32_t>(), Om.ell_zeta, Om.ell_mag, p0, pc, eta, 0.0, Y, m);
        else if (Om.ell_zeta == 8)
          hipLaunchKernelGGL((k_right_all_ell<128, 8>), grid, dim3(256), 0, st, rowptr, col, val, m,
                             Om.ell.as<uint32_t>(), Om.ell_zeta, Om.ell_mag, p0, pc, eta, 0.0, Y, m);
        else if (Om.ell_zeta > 0 && pc <= 64)
          hipLaunchKernelGGL((k_right_all_ell<64, 0>), grid, dim3(256), 0, st, rowptr, col, val, m,
                             Om.ell.as<uint32_t>(), Om.ell_zeta, Om.ell_mag, p0, pc, eta, 0.0, Y, m);
        else if (Om.ell_zeta > 0)
          hipLaunchKernelGGL((k_right_all_ell<128, 0>), grid, dim3(256), 0, st, rowptr, col, val, m,
                             Om.ell.as<uint32_t>(), Om.ell_zeta, Om.ell_mag, p0, pc, eta, 0.0, Y, m);
        else if (pc <= 64)
          hipLaunchKernelGGL(k_right_all_sparse<64>, grid, dim3(256), 0, st, rowptr, col, val, m, Ot.rp(), Ot.ci(),
                             Ot.v(), p0, pc, eta, 0.0, Y, m);
        else
          hipLaunchKernelGGL(k_right_all_sparse<128>, grid, dim3(256), 0, st, rowptr, col, val, m, Ot.rp(), Ot.ci(),
                             Ot.v(), p0, pc, eta, 0.0, Y, m);
        check_launch("k_right_all_sparse");
      }
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
