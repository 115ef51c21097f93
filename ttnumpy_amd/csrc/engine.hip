// Batched, device-resident AMEn engine.
//
// Every request ("slot") owns its trains, environments and scratch in HBM and
// a small descriptor (include/ttx_engine.h).  A tick launches four kernels over all slots,
// one workgroup per slot:
//
//   env   : the environment updates of the slot's current site (grid slots x 5,
//           one workgroup per environment kind);
//   solve : the one-site local solve (LU on the explicit matrix, or MINRES
//           with the true-residual restart loop);
//   post  : residual core, SVD truncation with the residual-budget rank rule,
//           enrichment, QR centre shifts (or the QR sweeps of set-up and the
//           residual-train fit);
//   check : the per-sweep relative residual in TT arithmetic and the stopping
//           test, then the advance of the slot's phase machine.
//
// All rank decisions and the stopping test run on the device, so the host
// only replays ticks and polls the slot status; requests at different sites,
// sweeps and ranks share every launch (continuous batching).  The algorithm
// and its order of operations are those of ttnumpy/solvers/amen.py:395-725
// with the helpers of one_site.py and common.py; each phase below cites the
// reference lines it follows.

#include "block_ops.h"
#include "gmres_blk.h"
#include "ttx_engine.h"

namespace ttx {

// One 256-thread workgroup per request slot with 75 KiB of LDS, so two slots
// share a CU (160 KiB of LDS, 2 waves per SIMD at <= 256 VGPRs): while one
// slot waits at a barrier or on memory, the other one issues.
constexpr int kEngThreads = 256;
constexpr size_t kEngLds = 75 * 1024;

// local solve in flight (slot.sstage): a direct solve is handed from the
// solve kernel to the Cholesky or LU kernel of the same tick; MINRES resumes
// over ticks, slot.minres_slice iterations at a time, so that a long solve of
// one request does not hold up the tick of every other request
enum { kSolveNone = 0, kSolveChol = 1, kSolveLu = 2, kSolveMinres = 3, kSolveChol32 = 4,
       kSolveMinres32 = 5, kSolveGmres = 6 };
constexpr int64_t kMixedMinN = 256;  // smaller direct solves stay in FP64
// the FP32 triangular solves keep the solution vector in LDS beside 4480
// floats of working space and index rows with 32-bit offsets
// (chol32_apply_inverse)
constexpr int64_t kMixedMaxN = 4096;
// ranks whose partial residuals the rank rule forms per operator application
constexpr int kRankChunkMax = 16;
static_assert(kMixedMaxN + 4480 <= (int64_t)(kEngLds / 4), "LDS of the FP32 triangular solves");
static_assert(kCholLds * 8 <= (int)kEngLds, "LDS of the Cholesky solve");
static_assert(kChol32Lds * 4 <= (int)kEngLds, "LDS of the FP32 Cholesky factor");

__device__ __forceinline__ double* xcore(const ttx_slot* s, int64_t k) { return s->x + k * s->cap_core; }
__device__ __forceinline__ double* zcore(const ttx_slot* s, int64_t k) { return s->z + k * s->cap_zcore; }
__device__ __forceinline__ double* envp(const ttx_slot* s, int kind, int side, int64_t k) {
  return s->env + s->env_off[kind * 2 + side] + k * s->cap_env[kind];
}
__device__ __forceinline__ double* scr(const ttx_slot* s, int id) { return s->scr + s->scr_off[id]; }

__device__ __forceinline__ void bfill(double* p, int64_t n, double v) {
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}
__device__ __forceinline__ void bcopy(double* dst, const double* src, int64_t n) {
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}
__device__ __forceinline__ double bnorm2(const double* a, const double* b, int64_t n, double* red) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = b ? a[i] - b[i] : a[i];
    s += v * v;
  }
  return sqrt(block_sum(s, red));
}

// ---------------------------------------------------------------- contractions
// left environment step (common.py:170-188): out from E with frame core F
// (rf, nr, rf2), target core T (rt, nc, rt2), operator core A (Ra, nr, nc, Ra2)
__device__ void env_left_blk(const double* E, const double* F, int64_t rf, int64_t nr, int64_t rf2,
                             const double* T, int64_t rt, int64_t nc, int64_t rt2, const double* A,
                             int64_t Ra, int64_t Ra2, double* out, double* s1, double* s2) {
  if (!A) {
    bgemm(1, rt, nr * rf2, rf, 1.0, E, rf, 1, 0, F, nr * rf2, 1, 0, 0.0, s1, nr * rf2, 1, 0);
    __syncthreads();
    bgemm(1, rt2, rf2, rt * nc, 1.0, T, 1, rt2, 0, s1, rf2, 1, 0, 0.0, out, rf2, 1, 0);
    return;
  }
  bgemm(1, rt * Ra, nr * rf2, rf, 1.0, E, rf, 1, 0, F, nr * rf2, 1, 0, 0.0, s1, nr * rf2, 1, 0);
  __syncthreads();
  bgemm(rt, rf2, nc * Ra2, Ra * nr, 1.0, s1, 1, rf2, Ra * nr * rf2, A, nc * Ra2, 1, 0, 0.0, s2, 1,
        rf2, nc * Ra2 * rf2);
  __syncthreads();
  bgemm(1, rt2, Ra2 * rf2, rt * nc, 1.0, T, 1, rt2, 0, s2, Ra2 * rf2, 1, 0, 0.0, out, Ra2 * rf2, 1, 0);
}

// right environment step (common.py:203-217)
__device__ void env_right_blk(const double* Phi, const double* F, int64_t rf, int64_t nr, int64_t rf2,
                              const double* T, int64_t rt, int64_t nc, int64_t rt2, const double* A,
                              int64_t Ra, int64_t Ra2, double* out, double* s1, double* s2) {
  if (!A) {
    bgemm(1, rt * nc, rf2, rt2, 1.0, T, rt2, 1, 0, Phi, rf2, 1, 0, 0.0, s1, rf2, 1, 0);
    __syncthreads();
    bgemm(1, rt, rf, nc * rf2, 1.0, s1, nc * rf2, 1, 0, F, 1, nr * rf2, 0, 0.0, out, rf, 1, 0);
    return;
  }
  bgemm(1, rt * nc, Ra2 * rf2, rt2, 1.0, T, rt2, 1, 0, Phi, Ra2 * rf2, 1, 0, 0.0, s1, Ra2 * rf2, 1, 0);
  __syncthreads();
  bgemm(rt, rf2, Ra * nr, nc * Ra2, 1.0, s1, 1, rf2, nc * Ra2 * rf2, A, 1, nc * Ra2, 0, 0.0, s2, 1,
        rf2, Ra * nr * rf2);
  __syncthreads();
  bgemm(1, rt * Ra, rf, nr * rf2, 1.0, s2, nr * rf2, 1, 0, F, 1, nr * rf2, 0, 0.0, out, rf, 1, 0);
}

// b_loc[f_l, i, f_r] = sum Psi[t, f_l] b[t, i, t'] Phi[t', f_r] (one_site.py:27-46)
__device__ void local_rhs_blk(const double* psi, int64_t rt, int64_t fl, const double* bc, int64_t n,
                              int64_t rt2, const double* phi, int64_t fr, double* out, double* s1) {
  bgemm(1, fl, n * rt2, rt, 1.0, psi, 1, fl, 0, bc, n * rt2, 1, 0, 0.0, s1, n * rt2, 1, 0);
  __syncthreads();
  bgemm(1, fl * n, fr, rt2, 1.0, s1, rt2, 1, 0, phi, fr, 1, 0, 0.0, out, fr, 1, 0);
}

// one-site operator Psi . A_k . Phi, matrix-free (one_site.py:78-115);
// rows (fl, i, fr), columns (rl, j, rr)
struct LocOp {
  const double* psi;
  const double* A;
  const double* phi;
  int64_t rl, Ra, fl, nr, nc, Rb, rr, fr;
  double* s1;
  double* s2;
  __device__ int64_t rows() const { return fl * nr * fr; }
  __device__ int64_t cols() const { return rl * nc * rr; }
  // out = alpha * op(y) + beta * out for B columns (column-major); ends synced
  __device__ void apply(const double* y, int64_t B, double* out, double alpha, double beta) const {
    bgemm(1, B * rl * nc, Rb * fr, rr, 1.0, y, rr, 1, 0, phi, Rb * fr, 1, 0, 0.0, s1, Rb * fr, 1, 0);
    __syncthreads();
    if (Ra * nr <= 16)
      bgemm_shared_a_m16(B * rl, Ra * nr, fr, nc * Rb, 1.0, A, nc * Rb, 1, s1, fr, 1, nc * Rb * fr,
                         0.0, s2, fr, 1, Ra * nr * fr);
    else
      bgemm(B * rl, Ra * nr, fr, nc * Rb, 1.0, A, nc * Rb, 1, 0, s1, fr, 1, nc * Rb * fr, 0.0, s2,
            fr, 1, Ra * nr * fr);
    __syncthreads();
    bgemm(B, fl, nr * fr, rl * Ra, alpha, psi, 1, fl, 0, s2, nr * fr, 1, rl * Ra * nr * fr, beta,
          out, nr * fr, 1, fl * nr * fr);
    __syncthreads();
  }
};

// The one-site operator in FP32 for the inner solves of mixed requests:
// FP32 copies of psi, A_k and phi, the three products of LocOp::apply on
// FP32 MFMA (bgemm32, FP32 accumulation) through FP32 scratch; with bf16 the
// operands of each product are rounded to BF16 as they are loaded
// (bgemm_bf16, FP32 accumulation).
struct LocOp32 {
  const float* psi;
  const float* A;
  const float* phi;
  int64_t rl, Ra, fl, nr, nc, Rb, rr, fr;
  float* s1;
  float* s2;
  bool bf16;
  __device__ void apply(const float* y, float* out) const {
    auto gemm = [&](int64_t batch, int64_t M, int64_t N, int64_t K, const float* A_, int64_t a_ms,
                    int64_t a_ks, int64_t a_bs, const float* B_, int64_t b_ks, int64_t b_ns,
                    int64_t b_bs, float* C_, int64_t c_ms, int64_t c_ns, int64_t c_bs) {
      if (bf16)
        bgemm_bf16(batch, M, N, K, 1.0f, A_, a_ms, a_ks, a_bs, B_, b_ks, b_ns, b_bs, 0.0f, C_, c_ms,
                   c_ns, c_bs);
      else
        bgemm32(batch, M, N, K, 1.0f, A_, a_ms, a_ks, a_bs, B_, b_ks, b_ns, b_bs, 0.0f, C_, c_ms, c_ns,
                c_bs);
      __syncthreads();
    };
    gemm(1, rl * nc, Rb * fr, rr, y, rr, 1, 0, phi, Rb * fr, 1, 0, s1, Rb * fr, 1, 0);
    gemm(rl, Ra * nr, fr, nc * Rb, A, nc * Rb, 1, 0, s1, fr, 1, nc * Rb * fr, s2, fr, 1, Ra * nr * fr);
    gemm(1, fl, nr * fr, rl * Ra, psi, 1, fl, 0, s2, nr * fr, 1, 0, out, nr * fr, 1, 0);
  }
};

// MINRES's operator for a mixed request: FP64 vectors in and out, the
// product on LocOp32 (one column).  Ends synced.
struct MixedOp {
  LocOp32 op;
  float* y32;
  float* o32;
  int64_t n;
  __device__ void apply(const double* y, int64_t B, double* out, double alpha, double beta) const {
    (void)B;  // MINRES applies one column at a time
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) y32[i] = (float)y[i];
    __syncthreads();
    op.apply(y32, o32);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double v = alpha * (double)o32[i];
      out[i] = beta == 0.0 ? v : v + beta * out[i];
    }
    __syncthreads();
  }
};

// explicit matrix into out (row-major, row stride ld):
//   M[(f_l,i,f_r),(t_l,j,t_r)] = sum_a psi[t_l,a,f_l] * T_a,  T_a = sum_b A[a,i,j,b] phi[t_r,b,f_r]
// Each thread owns a column (t_l, j, t_r) and a share of the (i, f_r) pairs;
// T_a does not depend on f_l, so it is formed once per pair and reused down
// the fl rows that share it.  Stores of a row stay coalesced across threads.
// psi, A_k and phi are staged in LDS when they fit in lds_cap doubles (phi as
// [b][f_r][t_r], so that consecutive threads, consecutive t_r, read
// consecutive words).  With lower, only entries on or below the diagonal are
// written.  Matches local_matrix_entry's sums term for term (a ascending,
// zero psi skipped, b ascending).  Ends synced.
//
// The column loop takes its three operands as separate pointers with their
// strides, so that on the staged path every operand pointer derives from the
// LDS array alone and the compiler reads them with LDS instructions.  It is
// instantiated for operator bonds up to KR: every load of an entry's terms
// (KR values, indices clamped to the bond so they stay in bounds) is issued
// before the sums that use them, so a thread waits once per entry rather than
// once per term; the sums skip the terms past the bonds.
template <typename OutT, int KR>
__device__ __forceinline__ void local_matrix_cols(const LocOp& op, OutT* out, int64_t ld, bool lower,
                                                  const double* psi, const double* phi,
                                                  const double* A, int pt, int pb, int pf);

// operator bonds up to 16: the terms are read where the sums use them (KR
// clamped loads per term would not fit in registers)
template <typename OutT>
__device__ __forceinline__ void local_matrix_cols_wide(const LocOp& op, OutT* out, int64_t ld,
                                                       bool lower, const double* psi,
                                                       const double* phi, const double* A, int pt,
                                                       int pb, int pf) {
  constexpr int kMaxRa = 16;
  const int cols = (int)op.cols();
  const int Ra = (int)op.Ra, Rb = (int)op.Rb, fl = (int)op.fl, fr = (int)op.fr, nr = (int)op.nr,
            nc = (int)op.nc, rr = (int)op.rr;
  const int npairs = nr * fr;
  const int S = cols >= (int)blockDim.x ? 1 : (int)blockDim.x / cols;
  const int a_stride = nr * nc * Rb;
  for (int it = threadIdx.x; it < cols * S; it += blockDim.x) {
    const int sl = it / cols, c = it - sl * cols;
    const int cq = c / rr, t_r = c - cq * rr, t_l = cq / nc, j = cq - t_l * nc;
    const double* ps = psi + t_l * Ra * fl;
    const double* ph = phi + t_r * pt;
    const int pbeg = npairs * sl / S, pend = npairs * (sl + 1) / S;
    for (int pq = pbeg; pq < pend; ++pq) {
      const int i = pq / fr, f_r = pq - i * fr;
      double T[kMaxRa];
#pragma unroll
      for (int a = 0; a < kMaxRa; ++a) {
        if (a < Ra) {
          const double* arow = A + a * a_stride + (i * nc + j) * Rb;
          double t = 0.0;
          for (int b = 0; b < Rb; ++b) t += arow[b] * ph[b * pb + f_r * pf];
          T[a] = t;
        }
      }
      OutT* o = out + (int64_t)(i * fr + f_r) * ld + c;
      const int64_t ostep = (int64_t)nr * fr * ld;
      int f0 = 0;
      if (lower) {
        const int need = c - (i * fr + f_r);
        f0 = need > 0 ? (need + nr * fr - 1) / (nr * fr) : 0;
      }
      for (int f_l = f0; f_l < fl; ++f_l) {
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a < kMaxRa; ++a) {
          if (a < Ra) {
            const double p = ps[a * fl + f_l];
            if (p != 0.0) sum += p * T[a];
          }
        }
        o[f_l * ostep] = (OutT)sum;
      }
    }
  }
}

template <typename OutT, int KR>
__device__ __forceinline__ void local_matrix_cols(const LocOp& op, OutT* out, int64_t ld, bool lower,
                                                  const double* psi, const double* phi,
                                                  const double* A, int pt, int pb, int pf) {
  const int cols = (int)op.cols();
  const int Ra = (int)op.Ra, Rb = (int)op.Rb, fl = (int)op.fl, fr = (int)op.fr, nr = (int)op.nr,
            nc = (int)op.nc, rr = (int)op.rr;
  const int am1 = Ra - 1, bm1 = Rb - 1;
  const int npairs = nr * fr;
  const int S = cols >= (int)blockDim.x ? 1 : (int)blockDim.x / cols;
  const int a_stride = nr * nc * Rb;
  for (int it = threadIdx.x; it < cols * S; it += blockDim.x) {
    const int sl = it / cols, c = it - sl * cols;
    const int cq = c / rr, t_r = c - cq * rr, t_l = cq / nc, j = cq - t_l * nc;
    const double* ps = psi + t_l * Ra * fl;
    const double* ph = phi + t_r * pt;
    const int pbeg = npairs * sl / S, pend = npairs * (sl + 1) / S;
    for (int pq = pbeg; pq < pend; ++pq) {
      const int i = pq / fr, f_r = pq - i * fr;
      double phv[KR];  // phi[t_r, b, f_r], read once for every a
#pragma unroll
      for (int b = 0; b < KR; ++b) phv[b] = ph[(b < Rb ? b : bm1) * pb + f_r * pf];
      double T[KR];
#pragma unroll
      for (int a = 0; a < KR; ++a) {
        const double* arow = A + (a < Ra ? a : am1) * a_stride + (i * nc + j) * Rb;
        double av[KR];
#pragma unroll
        for (int b = 0; b < KR; ++b) av[b] = arow[b < Rb ? b : bm1];
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < KR; ++b)
          if (b < Rb) t += av[b] * phv[b];
        T[a] = t;
      }
      OutT* o = out + (int64_t)(i * fr + f_r) * ld + c;
      const int64_t ostep = (int64_t)nr * fr * ld;
      // with lower, only rows >= c: f_l >= ceil((c - i*fr - f_r) / (nr*fr))
      int f0 = 0;
      if (lower) {
        const int need = c - (i * fr + f_r);
        f0 = need > 0 ? (need + nr * fr - 1) / (nr * fr) : 0;
      }
      for (int f_l = f0; f_l < fl; ++f_l) {
        double pv[KR];
#pragma unroll
        for (int a = 0; a < KR; ++a) pv[a] = ps[(a < Ra ? a : am1) * fl + f_l];
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a < KR; ++a)
          if (a < Ra && pv[a] != 0.0) sum += pv[a] * T[a];
        o[f_l * ostep] = (OutT)sum;
      }
    }
  }
}

template <typename OutT>
__device__ __forceinline__ void local_matrix_blk(const LocOp& op, OutT* out, int64_t ld, double* lds,
                                 int64_t lds_cap, bool lower) {
  constexpr int kMaxRa = 16;
  const int rows = (int)op.rows(), cols = (int)op.cols();
  add_flops(2.0 * (double)rows * (double)cols * (double)op.Ra * (double)op.Rb);
  const int Ra = (int)op.Ra, Rb = (int)op.Rb, fl = (int)op.fl, fr = (int)op.fr, nr = (int)op.nr,
            nc = (int)op.nc, rr = (int)op.rr;
  const int npsi = (int)(op.rl * Ra * fl), nphi = (int)(rr * Rb * fr), nA = Ra * nr * nc * Rb;
  if (Ra > kMaxRa || Rb > kMaxRa) {  // wide operator bonds: entry by entry
    for (int it = threadIdx.x; it < cols; it += blockDim.x) {
      const int cq = it / rr, t_r = it - cq * rr, t_l = cq / nc, j = cq - t_l * nc;
      for (int row = 0; row < rows; ++row) {
        const int rq = row / fr, f_r = row - rq * fr, f_l = rq / nr, i = rq - f_l * nr;
        if (lower && row < it) continue;
        out[(int64_t)row * ld + it] = (OutT)local_matrix_entry(
            op.psi + (int64_t)t_l * Ra * fl, fl, f_l, op.A + (i * nc + j) * Rb,
            (int64_t)nr * nc * Rb, Ra, Rb, op.phi + (int64_t)t_r * Rb * fr, fr, f_r);
      }
    }
    __syncthreads();
    return;
  }
  if ((int64_t)(npsi + nphi + nA) <= lds_cap) {
    for (int e = threadIdx.x; e < npsi; e += blockDim.x) lds[e] = op.psi[e];
    for (int e = threadIdx.x; e < nphi; e += blockDim.x) {
      const int t_r = e / (Rb * fr), bf = e - t_r * (Rb * fr);
      lds[npsi + bf * rr + t_r] = op.phi[e];
    }
    for (int e = threadIdx.x; e < nA; e += blockDim.x) lds[npsi + nphi + e] = op.A[e];
    __syncthreads();
    const double* sp = lds;
    const double* sf = lds + npsi;
    const double* sa = lds + npsi + nphi;
    if (Ra <= 4 && Rb <= 4)
      local_matrix_cols<OutT, 4>(op, out, ld, lower, sp, sf, sa, 1, fr * rr, rr);
    else if (Ra <= 8 && Rb <= 8)
      local_matrix_cols<OutT, 8>(op, out, ld, lower, sp, sf, sa, 1, fr * rr, rr);
    else
      local_matrix_cols_wide<OutT>(op, out, ld, lower, sp, sf, sa, 1, fr * rr, rr);
  } else {
    local_matrix_cols_wide<OutT>(op, out, ld, lower, op.psi, op.phi, op.A, Rb * fr, fr, 1);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- QR helpers
// Economic QR of the strided m x n matrix A; Q (m x k) and R (k x n) written
// with their strides (Q may be null).  Input is copied first, so outputs may
// overlap A.  work: m*n + k + m*k doubles.  Ends synced.
// the dynamic LDS of the kernel that runs the step (post and check kernels are
// launched with kEngLds bytes of it); the QRs use it for their panels
extern __shared__ double eng_lds[];

__device__ void block_qr(int64_t m, int64_t n, const double* A, int64_t a_rs, int64_t a_cs, double* Q,
                         int64_t q_rs, int64_t q_cs, double* R, int64_t r_rs, int64_t r_cs,
                         double* work, double* red) {
  const int64_t k = m < n ? m : n;
  double* W = work;
  double* tau = W + m * n;
  double* Qw = tau + k;
  for (int64_t e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int64_t j = q32(e, m), i = e - j * m;
    W[e] = A[i * a_rs + j * a_cs];
  }
  __syncthreads();
  block_householder_qr(m, n, W, tau, Qw, red, Q != nullptr, eng_lds, (int64_t)(kEngLds / 8));
  __syncthreads();
  if (Q) {
    for (int64_t e = threadIdx.x; e < m * k; e += blockDim.x) {
      const int64_t c = q32(e, m), i = e - c * m;
      Q[i * q_rs + c * q_cs] = Qw[e];
    }
  }
  for (int64_t e = threadIdx.x; e < k * n; e += blockDim.x) {
    const int64_t c = q32(e, k), i = e - c * k;
    R[i * r_rs + c * r_cs] = (i <= c) ? W[i + c * m] : 0.0;
  }
  __syncthreads();
}

// train[k] <- Q of (rl n x rr), train[k+1] <- R train[k+1]; returns the new
// bond (one_site.py:181-193)
__device__ int64_t shift_right_blk(double* core_k, double* core_k1, const double* src, int64_t rl,
                                   int64_t n, int64_t rr, int64_t rest, double* qb, double* rb,
                                   double* tmp, double* work, double* red) {
  const int64_t m = rl * n;
  const int64_t kk = m < rr ? m : rr;
  block_qr(m, rr, src, rr, 1, qb, kk, 1, rb, rr, 1, work, red);
  bgemm(1, kk, rest, rr, 1.0, rb, rr, 1, 0, core_k1, rest, 1, 0, 0.0, tmp, rest, 1, 0);
  __syncthreads();
  bcopy(core_k, qb, m * kk);
  bcopy(core_k1, tmp, kk * rest);
  __syncthreads();
  return kk;
}

// train[k] <- Q^T of the transposed right unfolding, train[k-1] <- train[k-1] R^T
// (one_site.py:196-209)
__device__ int64_t shift_left_blk(double* core_k, double* core_km1, const double* src, int64_t rl,
                                  int64_t n, int64_t rr, int64_t rows_prev, double* qb, double* rb,
                                  double* tmp, double* work, double* red) {
  const int64_t w = n * rr;
  const int64_t kk = w < rl ? w : rl;
  block_qr(w, rl, src, 1, w, qb, 1, w, rb, rl, 1, work, red);  // qb holds Q^T (kk x w)
  bgemm(1, rows_prev, kk, rl, 1.0, core_km1, rl, 1, 0, rb, 1, rl, 0, 0.0, tmp, kk, 1, 0);
  __syncthreads();
  bcopy(core_k, qb, kk * w);
  bcopy(core_km1, tmp, rows_prev * kk);
  __syncthreads();
  return kk;
}

// ---------------------------------------------------------------- MINRES
// scipy.sparse.linalg.minres (no preconditioner, shift 0) by one block, the
// recurrences of krylov.hip; V holds 9 n doubles (x, v, y, r1, r2, w, w1, w2,
// yA).  x0 may alias V's x.  Returns with x in V[0..n).
// Resumable MINRES with SciPy's recurrences.  The iteration state lives in
// the slot: vectors in V (x, v, y, r1/r2, w/w2, A v at fixed offsets; r and w
// rotate by parity bits instead of copies) and the scalars in st[], so an
// attempt can span several ticks.  Element i of every vector pass is touched
// only by thread i mod blockDim; each pass loads kB elements per thread before
// it computes.
enum {
  kMrBeta = 0, kMrOldb, kMrDbar, kMrEpsln, kMrPhibar, kMrTnorm2, kMrGmax, kMrGmin, kMrCs, kMrSn,
  kMrBeta1, kMrItn, kMrIstop, kMrPr, kMrPw, kMrRtol, kMrMaxit, kMrAttempt, kMrTarget,
  kMrPhiStop,  // > 0: also stop once the residual estimate phibar <= st[kMrPhiStop] beta1
  kMrCount
};

struct MrVec {
  double *x, *v, *y, *r1, *r2, *w, *w2, *yA;
  __device__ MrVec(double* V, int64_t n, int pr, int pw)
      : x(V), v(V + n), y(V + 2 * n), r1(V + (3 + pr) * n), r2(V + (4 - pr) * n),
        w(V + (5 + 2 * pw) * n), w2(V + (7 - 2 * pw) * n), yA(V + 8 * n) {}
};

// begin an attempt from x0 (x0 may alias V's x); st[kMrRtol], st[kMrMaxit] set
template <class Op>
__device__ void minres_start(const Op& op, int64_t n, const double* b, const double* x0, double* V,
                             double* red, double* st) {
  MrVec m(V, n, 0, 0);
  const double eps = DBL_EPSILON;
  if (x0) op.apply(x0, 1, m.yA, 1.0, 0.0);
  double sb = 0.0, sr = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double xi = x0 ? x0[i] : 0.0;
    const double ri = x0 ? b[i] - m.yA[i] : b[i];
    m.x[i] = xi;
    m.r1[i] = ri;
    m.r2[i] = ri;
    m.w[i] = 0.0;
    m.w2[i] = 0.0;
    sb += b[i] * b[i];
    sr += ri * ri;
  }
  const double beta1sq = block_sum(sr, red);
  const double bn = sqrt(block_sum(sb, red));
  double istop = 0.0;
  if (beta1sq == 0.0) {
    istop = 100.0;  // x0 already solves it
  } else if (bn == 0.0) {
    bcopy(m.x, b, n);
    istop = 100.0;
  } else {
    const double beta1 = sqrt(beta1sq);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m.v[i] = m.r1[i] / beta1;
    if (threadIdx.x == 0) {
      st[kMrBeta] = beta1;
      st[kMrBeta1] = beta1;
      st[kMrPhibar] = beta1;
    }
  }
  if (threadIdx.x == 0) {
    st[kMrOldb] = 0.0;
    st[kMrDbar] = 0.0;
    st[kMrEpsln] = 0.0;
    st[kMrTnorm2] = 0.0;
    st[kMrGmax] = 0.0;
    st[kMrGmin] = DBL_MAX;
    st[kMrCs] = -1.0;
    st[kMrSn] = 0.0;
    st[kMrItn] = 0.0;
    st[kMrIstop] = istop;
    st[kMrPr] = 0.0;
    st[kMrPw] = 0.0;
  }
  (void)eps;
  __syncthreads();
}

// up to `budget` iterations of the current attempt; returns the number done
template <class Op>
__device__ int64_t minres_run(const Op& op, int64_t n, double* V, double* red, double* st,
                              int64_t budget, int64_t* iters, int64_t* clk = nullptr) {
  constexpr int kB = 4;
  const double eps = DBL_EPSILON;
  const int64_t T = blockDim.x;
  double beta = st[kMrBeta], oldb = st[kMrOldb], dbar = st[kMrDbar], epsln = st[kMrEpsln];
  double phibar = st[kMrPhibar], tnorm2 = st[kMrTnorm2], gmax = st[kMrGmax], gmin = st[kMrGmin];
  double cs = st[kMrCs], sn = st[kMrSn];
  const double beta1 = st[kMrBeta1], rtol = st[kMrRtol], phistop = st[kMrPhiStop];
  const int64_t maxiter = (int64_t)st[kMrMaxit];
  int64_t itn = (int64_t)st[kMrItn];
  int istop = (int)st[kMrIstop];
  int pr = (int)st[kMrPr], pw = (int)st[kMrPw];
  __syncthreads();  // every thread has read st before thread 0 rewrites it
  MrVec m(V, n, pr, pw);
  int64_t done = 0;
  int64_t tc = clk ? wall_clock64() : 0;
  while (istop == 0 && itn < maxiter && done < budget) {
    ++itn;
    ++done;
    op.apply(m.v, 1, m.yA, 1.0, 0.0);
    if (clk) {  // clk[0]: operator applications, clk[1]: the rest of the iterations
      const int64_t t = wall_clock64();
      if (threadIdx.x == 0) clk[0] += t - tc;
      tc = t;
    }
    const double f1 = itn >= 2 ? beta / oldb : 0.0;
    double s = 0.0;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += kB * T) {
      double ya[kB], rr[kB], vv[kB];
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        ya[q] = i < n ? m.yA[i] : 0.0;
        rr[q] = i < n ? m.r1[i] : 0.0;
        vv[q] = i < n ? m.v[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        if (i < n) {
          double yi = ya[q];
          if (itn >= 2) yi -= f1 * rr[q];
          m.y[i] = yi;
          s += vv[q] * yi;
        }
      }
    }
    const double alfa = block_sum(s, red);
    const double f2 = alfa / beta;
    s = 0.0;
    // r1 <- r2, r2 <- y - f2 r2 (the new r2 goes into the old r1's storage)
    double* nr2 = m.r1;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += kB * T) {
      double yy[kB], rr[kB];
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        yy[q] = i < n ? m.y[i] : 0.0;
        rr[q] = i < n ? m.r2[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        if (i < n) {
          const double yi = yy[q] - f2 * rr[q];
          nr2[i] = yi;
          s += yi * yi;
        }
      }
    }
    m.r1 = m.r2;
    m.r2 = nr2;
    pr ^= 1;
    const double nbeta = sqrt(block_sum(s, red));
    oldb = beta;
    tnorm2 += alfa * alfa + oldb * oldb + nbeta * nbeta;
    if (itn == 1 && nbeta / beta1 <= 10.0 * eps) istop = -1;
    const double oldeps = epsln;
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    epsln = sn * nbeta;
    dbar = -cs * nbeta;
    const double root = sqrt(gbar * gbar + dbar * dbar);
    double gamma = sqrt(gbar * gbar + nbeta * nbeta);
    gamma = gamma > eps ? gamma : eps;
    cs = gbar / gamma;
    sn = nbeta / gamma;
    const double phi = cs * phibar;
    phibar = sn * phibar;
    const double denom = 1.0 / gamma;
    s = 0.0;
    // w1 <- w2, w2 <- w, w <- (v - oldeps w1 - delta w2) / gamma (into the old
    // w2's storage), x += phi w, and the next Lanczos vector v <- r2 / beta
    double* nw_ = m.w2;
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += kB * T) {
      double o2[kB], o1[kB], vv[kB], xx[kB], rr[kB];
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        o2[q] = i < n ? m.w2[i] : 0.0;
        o1[q] = i < n ? m.w[i] : 0.0;
        vv[q] = i < n ? m.v[i] : 0.0;
        xx[q] = i < n ? m.x[i] : 0.0;
        rr[q] = i < n ? m.r2[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kB; ++q) {
        const int64_t i = i0 + q * T;
        if (i < n) {
          const double nw = (vv[q] - oldeps * o2[q] - delta * o1[q]) * denom;
          nw_[i] = nw;
          const double xi = xx[q] + phi * nw;
          m.x[i] = xi;
          s += xi * xi;
          m.v[i] = rr[q] / nbeta;
        }
      }
    }
    m.w2 = m.w;
    m.w = nw_;
    pw ^= 1;
    const double ynorm = sqrt(block_sum(s, red));
    beta = nbeta;
    gmax = fmax(gmax, gamma);
    gmin = fmin(gmin, gamma);
    const double Anorm = sqrt(tnorm2);
    const double epsx = Anorm * ynorm * eps;
    const double test1 = (ynorm == 0.0 || Anorm == 0.0) ? INFINITY : phibar / (Anorm * ynorm);
    const double test2 = (Anorm == 0.0) ? INFINITY : root / Anorm;
    const double Acond = gmax / gmin;
    if (istop == 0) {
      if (1.0 + test2 <= 1.0) istop = 2;
      if (1.0 + test1 <= 1.0) istop = 1;
      if (itn >= maxiter) istop = 6;
      if (Acond >= 0.1 / eps) istop = 4;
      if (epsx >= beta1) istop = 3;
      if (test2 <= rtol) istop = 2;
      if (test1 <= rtol) istop = 1;
      if (phistop > 0.0 && phibar <= phistop * beta1) istop = 1;
    }
    add_flops(18.0 * (double)n);
    if (clk) {
      const int64_t t = wall_clock64();
      if (threadIdx.x == 0) clk[1] += t - tc;
      tc = t;
    }
  }
  if (istop == 0 && itn >= maxiter) istop = 6;
  if (threadIdx.x == 0) {
    st[kMrBeta] = beta;
    st[kMrOldb] = oldb;
    st[kMrDbar] = dbar;
    st[kMrEpsln] = epsln;
    st[kMrPhibar] = phibar;
    st[kMrTnorm2] = tnorm2;
    st[kMrGmax] = gmax;
    st[kMrGmin] = gmin;
    st[kMrCs] = cs;
    st[kMrSn] = sn;
    st[kMrItn] = (double)itn;
    st[kMrIstop] = (double)istop;
    st[kMrPr] = (double)pr;
    st[kMrPw] = (double)pw;
    if (iters) *iters += done;
  }
  __syncthreads();
  return done;
}

// ---------------------------------------------------------------- slot phases
struct Env {
  const double* ptr;
  int64_t t, a, f;  // target, operator (0: none), frame bonds
};

__device__ LocOp make_mfree(const ttx_slot* s, int64_t k, const double* psi, int64_t rl, int64_t Ra,
                            int64_t fl, const double* phi, int64_t rr, int64_t Rb, int64_t fr) {
  LocOp op;
  op.psi = psi;
  op.A = s->A[k];
  op.phi = phi;
  op.rl = rl;
  op.Ra = Ra;
  op.fl = fl;
  op.nr = s->n[k];
  op.nc = s->n[k];
  op.Rb = Rb;
  op.rr = rr;
  op.fr = fr;
  op.s1 = scr(s, TTX_SCR_S1);
  op.s2 = scr(s, TTX_SCR_S2);
  return op;
}

// the x/x/A local operator at k, matrix-free (the direct solve's residual and
// rank tests apply it too: the same operator as the explicit matrix, at the
// cost of three small GEMMs instead of a pass over N^2 entries)
__device__ LocOp system_op(const ttx_slot* s, int64_t k) {
  const int64_t rl = s->rx[k], rr = s->rx[k + 1];
  return make_mfree(s, k, envp(s, 0, 0, k), rl, s->ra[k], rl, envp(s, 0, 1, k), rr, s->ra[k + 1], rr);
}

// z_k = b_part - Psi_zx . A_k . Phi_zx x_core (amen.py:109-159), into ZNEW
__device__ void residual_core_blk(ttx_slot* s, int64_t k, const double* x_core, double* red) {
  const int64_t zl = s->rz[k], zr = s->rz[k + 1], n = s->n[k];
  double* out = scr(s, TTX_SCR_ZNEW);
  local_rhs_blk(envp(s, 2, 0, k), s->rb[k], zl, s->b[k], n, s->rb[k + 1], envp(s, 2, 1, k), zr, out,
                scr(s, TTX_SCR_S1));
  __syncthreads();
  LocOp op = make_mfree(s, k, envp(s, 3, 0, k), s->rx[k], s->ra[k], zl, envp(s, 3, 1, k), s->rx[k + 1],
                        s->ra[k + 1], zr);
  op.apply(x_core, 1, out, -1.0, 1.0);
  (void)red;
}

// rank of the solved core (common.py:812-929); same value on every thread
__device__ int64_t choose_rank_blk(ttx_slot* s, const LocOp& op, const double* bloc, const double* U,
                                   const double* S, const double* Vt, int64_t p, int64_t q, int64_t K,
                                   double exact, int* capped, double* red) {
  const int64_t mr = s->max_rank;
  if (!s->residual_mode) {
    // sigma_i / sigma_1 > error_threshold, floor 1, cap (common.py:872-885)
    if (K == 0 || S[0] == 0.0) { *capped = 0; return 1; }
    int64_t c = 0;
    for (int64_t i = 0; i < K; ++i) c += (S[i] / S[0] > s->error_threshold) ? 1 : 0;
    int64_t r = c > 1 ? c : 1;
    *capped = 0;
    if (mr > 0 && r > mr) { r = mr; *capped = 1; }
    return r;
  }
  *capped = 0;
  if (S[0] == 0.0) return 1;
  int64_t limit = K;
  if (mr > 0 && mr < limit) limit = mr;
  const int64_t N = p * q;
  const double bn = bnorm2(bloc, nullptr, N, red);
  const double a = s->damp * exact, f = 0.5 * s->trunc_tol * bn;
  const double thr = a > f ? a : f;
  __shared__ double red_many[kRankChunkMax * (kEngThreads / 64)];
  double* partial = scr(s, TTX_SCR_TMP);
  double* terms = scr(s, TTX_SCR_CHK);
  double* contrib = terms + s->chunk * s->cap_loc;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) partial[i] = -bloc[i];
  __syncthreads();
  for (int64_t c0 = 0; c0 < limit; c0 += s->chunk) {
    const int64_t cb = (limit - c0) < s->chunk ? (limit - c0) : s->chunk;
    for (int64_t e = threadIdx.x; e < N * cb; e += blockDim.x) {
      const int64_t j = q32(e, N), row = e - j * N;
      const int64_t i = q32(row, q), jj = row - i * q, r = c0 + j;
      terms[e] = U[i * K + r] * S[r] * Vt[r * q + jj];
    }
    __syncthreads();
    op.apply(terms, cb, contrib, 1.0, 0.0);
    // the cb partial residuals' norms in one pass and one reduction; each
    // norm sums the same terms in the same order as a reduction per rank
    double t[kRankChunkMax];
#pragma unroll
    for (int j = 0; j < kRankChunkMax; ++j) t[j] = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
      double v = partial[i];
#pragma unroll
      for (int j = 0; j < kRankChunkMax; ++j)
        if (j < cb) {
          v += contrib[j * N + i];
          t[j] += v * v;
        }
      partial[i] = v;
    }
    block_sum_many<kRankChunkMax>(t, (int)cb, red_many);
    int hit = -1;  // the first rank whose residual meets thr
#pragma unroll
    for (int j = kRankChunkMax - 1; j >= 0; --j)
      if (j < cb && sqrt(t[j]) <= thr) hit = j;
    if (hit >= 0) return c0 + hit + 1;
  }
  *capped = (mr > 0 && mr < K) ? 1 : 0;
  return limit;
}

// charge the ticks since *t0 to stage i and restart the stopwatch
__device__ __forceinline__ void stamp(ttx_slot* s, int i, int64_t& t0) {
  const int64_t t = wall_clock64();
  if (threadIdx.x == 0) s->clk[i] += t - t0;
  t0 = t;
}

__device__ __forceinline__ void fail(ttx_slot* s, int64_t code) {
  if (threadIdx.x == 0) s->status = code;
}

// ---------------------------------------------------------------- kernels
__device__ void env_body(ttx_slot* s, int kind) {
  if (s->status != TTX_SLOT_RUNNING || s->sstage != kSolveNone) return;
  const int64_t ph = s->phase, d = s->d;
  int side;
  int64_t k = s->k;
  bool zonly = false;
  switch (ph) {
    case TTX_PH_FIT_INIT: side = 1; zonly = true; break;
    case TTX_PH_FIT_FWD: side = 0; zonly = true; break;
    case TTX_PH_FIT_LAST: side = 0; zonly = true; k = d - 1; break;
    case TTX_PH_FIT_BWD: side = 1; zonly = true; break;
    case TTX_PH_INIT_R: side = 1; break;
    case TTX_PH_FWD: side = 0; break;
    case TTX_PH_LAST: side = 0; k = d - 1; break;
    case TTX_PH_BWD: side = 1; break;
    default: return;
  }
  if (zonly && kind != 2 && kind != 3) return;
  double* out = envp(s, kind, side, k);
  if ((side == 0 && k == 0) || (side == 1 && k == d - 1)) {
    if (threadIdx.x == 0) out[0] = 1.0;
    return;
  }
  const int64_t c = side == 0 ? k - 1 : k + 1;
  // frame/target trains of the kind: 0 (x,x,A) 1 (x,b) 2 (z,b) 3 (z,x,A) 4 (x,z)
  const double* F;
  const int64_t* fb;
  const double* T;
  const int64_t* tb;
  const bool frame_z = (kind == 2 || kind == 3);
  F = frame_z ? zcore(s, c) : xcore(s, c);
  fb = frame_z ? s->rz : s->rx;
  if (kind == 0 || kind == 3) { T = xcore(s, c); tb = s->rx; }
  else if (kind == 4) { T = zcore(s, c); tb = s->rz; }
  else { T = s->b[c]; tb = s->rb; }
  const bool with_op = (kind == 0 || kind == 3);
  const double* A = with_op ? s->A[c] : nullptr;
  const int64_t n = s->n[c];
  double* s1 = scr(s, TTX_SCR_ENVS) + (int64_t)kind * 2 * s->cap_core * s->ramax;
  double* s2 = s1 + s->cap_core * s->ramax;
  if (side == 0)
    env_left_blk(envp(s, kind, 0, k - 1), F, fb[c], n, fb[c + 1], T, tb[c], n, tb[c + 1], A,
                 with_op ? s->ra[c] : 1, with_op ? s->ra[c + 1] : 1, out, s1, s2);
  else
    env_right_blk(envp(s, kind, 1, k + 1), F, fb[c], n, fb[c + 1], T, tb[c], n, tb[c + 1], A,
                  with_op ? s->ra[c] : 1, with_op ? s->ra[c + 1] : 1, out, s1, s2);
}

// ---- the stages of a tick.  Each runs one kernel's share of a tick for one
// slot and banks its flops in s->flops[stage]; the tick kernels run one stage
// over all slots, eng_run_kernel all seven in turn for its own slot.

// environments of one kind (0..4); `several`: other workgroups update the
// same slot's counters at the same time (the tick kernel's five kinds)
__device__ __forceinline__ void stage_env(ttx_slot* s, int kind, bool several) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  const int64_t t0 = wall_clock64();
  env_body(s, kind);
  __syncthreads();
  if (threadIdx.x == 0 && s->status == TTX_SLOT_RUNNING) {
    const unsigned long long dt = (unsigned long long)(wall_clock64() - t0);
    if (several)
      atomicAdd((unsigned long long*)&s->clk[7], dt);
    else
      s->clk[7] += (int64_t)dt;
  }
  if (threadIdx.x == 0 && g_flops != 0.0) {
    if (several)
      atomicAdd(&s->flops[0], g_flops);
    else
      s->flops[0] += g_flops;
  }
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_env_kernel(ttx_slot* slots) {
  stage_env(slots + blockIdx.x, blockIdx.y, true);
}

// The local solve of a sweep step (one_site.py:118-178, common.py:729-809)
// runs as up to three kernels of one tick, each with its own register
// budget: eng_solve_kernel forms b_loc and, for a large symmetric system,
// runs restarted MINRES; a direct solve is handed on through s->sstage to
// eng_chol_kernel (symmetric: Cholesky) and, when that finds the matrix not
// positive definite or the problem is not symmetric, to eng_lu_kernel (LU
// with partial pivoting).  scal[0] receives the exact local residual.

// ---- mixed iterative local solves
// The inner solve runs MINRES on the correction equation A_32 d = r with
// the operator applied in FP32 (MixedOp, FP32 copies of psi, A_k, phi; FP32
// accumulation) until its residual estimate has fallen by
// max(kMixedInnerRtol, target/2 ||r||) (SciPy's own test is relative to
// ||A|| ||d|| and stops far too early on ill-conditioned systems); the outer loop adds d to u in FP64 and recomputes r = b_loc - A u
// with the FP64 operator, until ||r|| meets the reference's acceptance test
// (||r|| <= local_rtol ||b_loc||, common.py:787) or kMixedOuterMax passes,
// the reference's budget of MINRES attempts before it accepts an inexact
// solve (MAX_LOCAL_RESTARTS, common.py:779-800).
// The environments stay FP64: the outer residual, the exact residual of the
// rank rule and everything after the solve see the FP64 operator.
constexpr int kMixedOuterMax = 4;
constexpr double kMixedInnerRtol = 1e-5;
// BF16 operands (slot.mixed = 2): the operator's rounding error, ~1e-3 of its
// norm, bounds what one correction can gain, so the inner solve stops at 1e-3
// and the outer loop gets more passes
constexpr int kBf16OuterMax = 16;
constexpr double kBf16InnerRtol = 1e-3;
enum { kMrOuter = kMrCount };

// the FP32 operator of site k in the slot's M32 region into *m (the factors
// are converted when fill is set); false when the region cannot hold it
// (with m null: only that test)
__device__ bool mixed_op(ttx_slot* s, int64_t k, MixedOp* mp, bool fill = false) {
  const int64_t rl = s->rx[k], n = s->n[k], rr = s->rx[k + 1];
  const int64_t Ra = s->ra[k], Rb = s->ra[k + 1];
  const int64_t N = rl * n * rr;
  const int64_t npsi = rl * Ra * rl, nphi = rr * Rb * rr, nA = Ra * n * n * Rb;
  const int64_t ns1 = rl * n * Rb * rr, ns2 = rl * Ra * n * rr;
  const int64_t have = 2 * (s->scr_size - s->scr_off[TTX_SCR_M32]);  // floats
  if (!s->mixed || npsi + nphi + nA + ns1 + ns2 + 2 * N > have) return false;
  if (mp == nullptr) return true;
  MixedOp& m = *mp;
  float* base = reinterpret_cast<float*>(scr(s, TTX_SCR_M32));
  float* psi = base;
  float* phi = psi + npsi;
  float* A = phi + nphi;
  m.op.psi = psi;
  m.op.phi = phi;
  m.op.A = A;
  m.op.rl = rl;
  m.op.Ra = Ra;
  m.op.fl = rl;
  m.op.nr = n;
  m.op.nc = n;
  m.op.Rb = Rb;
  m.op.rr = rr;
  m.op.fr = rr;
  m.op.s1 = A + nA;
  m.op.s2 = m.op.s1 + ns1;
  m.op.bf16 = s->mixed == 2;
  m.y32 = m.op.s2 + ns2;
  m.o32 = m.y32 + N;
  m.n = N;
  if (fill) {
    const double* psi64 = envp(s, 0, 0, k);
    const double* phi64 = envp(s, 0, 1, k);
    const double* A64 = s->A[k];
    for (int64_t i = threadIdx.x; i < npsi; i += blockDim.x) psi[i] = (float)psi64[i];
    for (int64_t i = threadIdx.x; i < nphi; i += blockDim.x) phi[i] = (float)phi64[i];
    for (int64_t i = threadIdx.x; i < nA; i += blockDim.x) A[i] = (float)A64[i];
    __syncthreads();
  }
  return true;
}

// r = b_loc - A_64 u into r; returns ||r||
__device__ double mixed_residual(const LocOp& op, const double* u, const double* bloc, double* r,
                                 int64_t N, double* red) {
  op.apply(u, 1, r, -1.0, 0.0);
  double sq = 0.0;
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
    const double ri = bloc[i] + r[i];
    r[i] = ri;
    sq += ri * ri;
  }
  return sqrt(block_sum(sq, red));
}

__device__ __forceinline__ bool solve_step_active(const ttx_slot* s) {
  return s->status == TTX_SLOT_RUNNING && (s->phase == TTX_PH_FWD || s->phase == TTX_PH_BWD);
}

// ||A_loc u - b_loc|| into scal[0] when the residual budget needs it
__device__ void exact_local_residual(ttx_slot* s, int64_t k, int64_t N, double* red, int64_t& t0) {
  double* scal = scr(s, TTX_SCR_SCAL);
  if (s->residual_mode) {
    LocOp op = system_op(s, k);
    double* tmp = scr(s, TTX_SCR_TMP);
    op.apply(scr(s, TTX_SCR_U), 1, tmp, 1.0, 0.0);
    const double ex = bnorm2(tmp, scr(s, TTX_SCR_BLOC), N, red);
    if (threadIdx.x == 0) scal[0] = ex;
    stamp(s, 4, t0);
  } else if (threadIdx.x == 0) {
    scal[0] = 0.0;
  }
}

__device__ void solve_body(ttx_slot* s) {
  __shared__ double red[16];
  if (!solve_step_active(s)) return;
  const int64_t k = s->k;
  const int64_t rl = s->rx[k], n = s->n[k], rr = s->rx[k + 1];
  const int64_t N = rl * n * rr;
  int64_t t0 = wall_clock64();
  double* u = scr(s, TTX_SCR_U);
  double* bloc = scr(s, TTX_SCR_BLOC);
  double* tmp = scr(s, TTX_SCR_TMP);
  double* scal = scr(s, TTX_SCR_SCAL);
  double* st = scal + 8;  // MINRES state (kMrCount doubles)
  double* V = scr(s, TTX_SCR_MR);
  const int64_t stage = s->sstage;
  if (stage != kSolveNone && stage != kSolveMinres) return;  // a direct or mixed solve is pending
  LocOp op = system_op(s, k);
  if (stage == kSolveNone) {
    local_rhs_blk(envp(s, 1, 0, k), s->rb[k], rl, s->b[k], n, s->rb[k + 1], envp(s, 1, 1, k), rr,
                  bloc, scr(s, TTX_SCR_S1));
    __syncthreads();
    stamp(s, 0, t0);
    const bool iter = s->iterative && N > s->direct_size;
    if (!iter) {
      if (N > s->direct_cap) { fail(s, TTX_SLOT_ERR_CAPACITY); return; }
      if (threadIdx.x == 0)
        s->sstage = !s->chol ? kSolveLu
                             : (s->mixed && N >= kMixedMinN && N <= kMixedMaxN ? kSolveChol32
                                                                                : kSolveChol);
      return;
    }
    // first attempt, warm-started from the current core (one_site.py:151-153)
    const double target = s->local_rtol * bnorm2(bloc, nullptr, N, red);
    // GMRES (common.py:786) runs in eng_lu_kernel, the mixed FP32 inner
    // solves in eng_chol32_kernel
    const int64_t hand = s->iterative == 2 ? kSolveGmres
                         : (s->mixed && mixed_op(s, k, nullptr)) ? kSolveMinres32 : kSolveNone;
    if (hand != kSolveNone) {
      if (threadIdx.x == 0) {
        st[kMrTarget] = target;
        st[kMrRtol] = s->local_rtol;
        st[kMrAttempt] = 0.0;
        st[kMrOuter] = -1.0;  // not started
        s->sstage = hand;
      }
      __syncthreads();
      return;
    } else if (threadIdx.x == 0) {
      st[kMrRtol] = s->local_rtol;
      st[kMrMaxit] = (double)(5 * N);
      st[kMrAttempt] = 0.0;
      st[kMrTarget] = target;
      st[kMrPhiStop] = 0.0;  // SciPy's stopping tests only
      s->sstage = kSolveMinres;
    }
    minres_start(op, N, bloc, xcore(s, k), V, red, st);
  }

  // restarted MINRES with the true-residual check (common.py:776-800), at most
  // slot.minres_slice iterations in this tick
  int64_t budget = s->minres_slice > 0 ? s->minres_slice : 128;
  while (true) {
    budget -= minres_run(op, N, V, red, st, budget, threadIdx.x == 0 ? &s->minres_iters : nullptr,
                         &s->clk[20]);
    if (st[kMrIstop] == 0.0) break;  // out of budget: resume next tick
    op.apply(V, 1, tmp, 1.0, 0.0);
    const double res = bnorm2(tmp, bloc, N, red);
    const double target = st[kMrTarget];
    const int attempt = (int)st[kMrAttempt];
    const bool ok = res <= target;
    if (!ok && attempt >= 3 && s->allow_fallback && N <= s->direct_cap) {
      // the explicit-matrix fallback (common.py:793-809): the direct kernels
      // of this tick solve the system on the explicit local matrix
      __syncthreads();
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        s->local_fallbacks += 1;
        s->sstage = s->chol ? kSolveChol : kSolveLu;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    if (ok || attempt >= 3) {
      bcopy(u, V, N);
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        if (!ok) s->local_inexact += 1;
        scal[1] = 0.0;
        s->sstage = kSolveNone;
      }
      __syncthreads();
      stamp(s, 3, t0);
      exact_local_residual(s, k, N, red, t0);
      return;
    }
    __syncthreads();  // everyone has read st before it is rewritten
    if (threadIdx.x == 0) {
      st[kMrRtol] = fmax(st[kMrRtol] * 1e-2, 1e-15);
      st[kMrAttempt] = (double)(attempt + 1);
    }
    minres_start(op, N, bloc, V, V, red, st);  // warm start from the last iterate
    if (budget <= 0) break;
  }
  stamp(s, 3, t0);
}

// mixed precision: FP32 Cholesky on the lower triangle (entries formed in
// FP64, rounded once) and FP64 iterative refinement, u_{i+1} = u_i +
// (L L^T)^{-1} (b - A_loc u_i) with the residual through the matrix-free FP64
// operator, until it stops shrinking; accepted once ||r|| <= kRefineAccept
// ||b_loc||, else the FP64 Cholesky takes the step (same tick).
// The refinement stops at ||r|| <= 1e-12 ||b_loc||: three orders below the
// local tolerance the reference accepts from its iterative local solves
// (local_rtol = 0.1 residual_tol = 1e-9, common.py:374-375), and far below
// the residual-budget threshold 0.5 truncation_tol ||b_loc|| (~1e-9), so the
// rank decisions it feeds are those of an exact solve.
constexpr int kRefineMax = 8;
constexpr double kRefineStop = 1e-12;
constexpr double kRefineAccept = 1e-11;
// the FP32 factor (N rows of stride ld32 = N rounded up to 4 floats, so that
// 16-byte loads of its rows are aligned), the right-hand side (N, padded to
// 64) and the diagonal-block inverses (4096 per 64 rows) share the LU scratch
// of 2 N (N+1) floats, which holds them for N >= 113
static_assert(kMixedMinN >= 113, "mixed direct solves need N >= 113 for their scratch");

// (a function of its own, called from two kernels: it names the dynamic LDS
// itself, so that its LDS operands are read with LDS instructions whatever
// pointer a caller holds)
__device__ void chol32_body(ttx_slot* s) {
  double* lds = eng_lds;
  __shared__ double red[16];
  __shared__ int32_t s_info;
  if (!solve_step_active(s) || s->sstage != kSolveChol32) return;
  const int64_t k = s->k;
  const int64_t N = s->rx[k] * s->n[k] * s->rx[k + 1];
  int64_t t0 = wall_clock64();
  const int64_t ld32 = (N + 3) & ~(int64_t)3;
  float* W = reinterpret_cast<float*>(scr(s, TTX_SCR_LU));  // N ld32 floats of 2 N (N+1)
  float* d32 = W + N * ld32;
  float* Dinv = d32 + ((N + 63) / 64) * 64;  // 64 x 64 diagonal-block inverses of L
  const double* bloc = scr(s, TTX_SCR_BLOC);
  double* u = scr(s, TTX_SCR_U);
  double* r = scr(s, TTX_SCR_TMP);
  const LocOp op = system_op(s, k);
  local_matrix_blk<float>(op, W, ld32, lds, kEngLds / 8, true);
  if (threadIdx.x == 0) s_info = TTX_OK;
  __syncthreads();
  stamp(s, 1, t0);
  block_chol32_factor<kEngThreads>(N, ld32, W, reinterpret_cast<float*>(lds), &s_info,
                                  &s->clk[22]);
  __syncthreads();
  if (s_info == TTX_OK) chol32_invert_diag<kEngThreads>(N, ld32, W, Dinv, reinterpret_cast<float*>(lds));
  stamp(s, 18, t0);
  bool ok = false;
  int it = 0;
  double rn = 0.0;
  if (s_info == TTX_OK) {
    const double bn = bnorm2(bloc, nullptr, N, red);
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
      u[i] = 0.0;
      r[i] = bloc[i];
    }
    __syncthreads();
    double prev = bn;
    for (it = 1; it <= kRefineMax; ++it) {
      for (int64_t i = threadIdx.x; i < N; i += blockDim.x) d32[i] = (float)r[i];
      __syncthreads();
      chol32_apply_inverse<kEngThreads>(N, ld32, W, Dinv, d32, reinterpret_cast<float*>(lds));
      for (int64_t i = threadIdx.x; i < N; i += blockDim.x) u[i] += (double)d32[i];
      __syncthreads();
      op.apply(u, 1, r, -1.0, 0.0);  // r = -A u
      double sq = 0.0;
      for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        const double ri = bloc[i] + r[i];
        r[i] = ri;
        sq += ri * ri;
      }
      rn = sqrt(block_sum(sq, red));
      add_flops(4.0 * (double)N * (double)N);
      if (rn <= kRefineStop * bn) { ok = true; break; }
      if (rn > 0.25 * prev) { ok = rn <= kRefineAccept * bn; break; }
      prev = rn;
    }
    if (it > kRefineMax) ok = rn <= kRefineAccept * bn;
  }
  stamp(s, 19, t0);
  if (!ok) {  // not SPD in FP32, or the refinement stalled: FP64 Cholesky this tick
    if (threadIdx.x == 0) {
      s->mixed_fallbacks += 1;
      s->sstage = kSolveChol;
    }
    return;
  }
  if (threadIdx.x == 0) {
    double* scal = scr(s, TTX_SCR_SCAL);
    scal[1] = 1.0;
    scal[0] = s->residual_mode ? rn : 0.0;  // the refinement's last residual is exact
    s->lu_solves += 1;
    s->mixed_solves += 1;
    s->mixed_refines += it;
    s->sstage = kSolveNone;
  }
  __syncthreads();
}

// The mixed iterative local solve (see kMixedOuterMax): started here on the
// tick the solve kernel hands it over (slot.sstage = kSolveMinres32 with
// st[kMrOuter] = -1), then resumed tick by tick.
__device__ void minres32_body(ttx_slot* s) {
  __shared__ double red[16];
  if (!solve_step_active(s) || s->sstage != kSolveMinres32) return;
  const int64_t k = s->k;
  const int64_t N = s->rx[k] * s->n[k] * s->rx[k + 1];
  int64_t t0 = wall_clock64();
  double* u = scr(s, TTX_SCR_U);
  const double* bloc = scr(s, TTX_SCR_BLOC);
  double* tmp = scr(s, TTX_SCR_TMP);
  double* scal = scr(s, TTX_SCR_SCAL);
  double* st = scal + 8;
  double* V = scr(s, TTX_SCR_MR);
  const LocOp op = system_op(s, k);
  const int outer_max = s->mixed == 2 ? kBf16OuterMax : kMixedOuterMax;
  const double inner_rtol = s->mixed == 2 ? kBf16InnerRtol : kMixedInnerRtol;
  MixedOp mop;
  const bool start = st[kMrOuter] < 0.0;
  mixed_op(s, k, &mop, start);  // its size was checked when the solve was handed over
  if (start) {
    // u = the warm start (the current core, one_site.py:151-153), r = b_loc - A_64 u
    bcopy(u, xcore(s, k), N);
    __syncthreads();
    const double res = mixed_residual(op, u, bloc, tmp, N, red);
    const double target = st[kMrTarget];
    __syncthreads();  // every thread has read st before it changes
    if (res <= target) {  // the warm start already meets the test
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        s->mixed_solves += 1;
        scal[1] = 0.0;
        scal[0] = s->residual_mode ? res : 0.0;
        s->sstage = kSolveNone;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    if (threadIdx.x == 0) {
      st[kMrOuter] = 0.0;
      st[kMrRtol] = 1e-15;  // the residual-estimate test below ends the inner solve
      st[kMrPhiStop] = fmax(inner_rtol, 0.5 * target / res);
      st[kMrMaxit] = (double)(5 * N);
    }
    minres_start(mop, N, tmp, nullptr, V, red, st);
  }

  // FP32 inner solves, FP64 corrections; at most slot.minres_slice inner
  // iterations in this tick
  int64_t budget = s->minres_slice > 0 ? s->minres_slice : 128;
  while (true) {
    budget -= minres_run(mop, N, V, red, st, budget, threadIdx.x == 0 ? &s->minres_iters : nullptr);
    if (st[kMrIstop] == 0.0) break;  // out of budget: resume next tick
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) u[i] += V[i];
    __syncthreads();
    const double res = mixed_residual(op, u, bloc, tmp, N, red);
    const double target = st[kMrTarget];
    const int outer = (int)st[kMrOuter] + 1;
    const bool ok = res <= target;
    __syncthreads();  // everyone has read st before it is rewritten
    if (!ok && outer >= outer_max && s->allow_fallback && N <= s->direct_cap) {
      // the explicit-matrix fallback (common.py:793-809) in FP64
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        s->local_fallbacks += 1;
        s->mixed_refines += outer;
        s->sstage = kSolveChol;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    if (ok || outer >= outer_max) {
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        if (!ok) s->local_inexact += 1;
        s->mixed_solves += 1;
        s->mixed_refines += outer;
        scal[1] = 0.0;
        scal[0] = s->residual_mode ? res : 0.0;  // ||A_64 u - b_loc||, the exact residual
        s->sstage = kSolveNone;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    if (threadIdx.x == 0) {
      st[kMrOuter] = (double)outer;
      st[kMrPhiStop] = fmax(inner_rtol, 0.5 * target / res);
    }
  minres_start(mop, N, tmp, nullptr, V, red, st);
    if (budget <= 0) break;
  }
  stamp(s, 3, t0);
}

__device__ __forceinline__ void stage_chol32(ttx_slot* s) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  chol32_body(s);
  __syncthreads();
  minres32_body(s);
  __syncthreads();
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[2] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_chol32_kernel(ttx_slot* slots) {
  stage_chol32(slots + blockIdx.x);
}

__device__ __forceinline__ void chol_body(ttx_slot* s, double* lds) {
  __shared__ double red[16];
  __shared__ int32_t s_info;
  if (!solve_step_active(s) || s->sstage != kSolveChol) return;
  const int64_t k = s->k;
  const int64_t N = s->rx[k] * s->n[k] * s->rx[k + 1];
  int64_t t0 = wall_clock64();
  double* W = scr(s, TTX_SCR_LU);
  const double* bloc = scr(s, TTX_SCR_BLOC);
  local_matrix_blk<double>(system_op(s, k), W, N + 1, lds, kEngLds / 8, true);
  for (int64_t r = threadIdx.x; r < N; r += blockDim.x) W[r * (N + 1) + N] = bloc[r];
  if (threadIdx.x == 0) s_info = TTX_OK;
  __syncthreads();
  stamp(s, 1, t0);
  // the tile height below the diagonal blocks is the slot's (picked by the
  // host's tuner); both give the same values
  if (s->tiles & 1)
    block_chol_solve<kEngThreads, 1>(N, N + 1, W, scr(s, TTX_SCR_U), lds, &s_info, &s->clk[13]);
  else
    block_chol_solve<kEngThreads, 2>(N, N + 1, W, scr(s, TTX_SCR_U), lds, &s_info, &s->clk[13]);
  __syncthreads();
  stamp(s, 2, t0);
  if (s_info == TTX_ENOTSPD) {  // not positive definite: LU takes over this tick
    if (threadIdx.x == 0) {
      s->chol_fallbacks += 1;
      s->sstage = kSolveLu;
    }
    return;
  }
  if (threadIdx.x == 0) {
    scr(s, TTX_SCR_SCAL)[1] = 1.0;
    s->lu_solves += 1;
    s->sstage = kSolveNone;
  }
  __syncthreads();
  exact_local_residual(s, k, N, red, t0);
}

__device__ __forceinline__ void lu_body(ttx_slot* s, double* lds) {
  __shared__ double red[16];
  __shared__ int32_t s_info;
  if (!solve_step_active(s) || s->sstage != kSolveLu) return;
  const int64_t k = s->k;
  const int64_t N = s->rx[k] * s->n[k] * s->rx[k + 1];
  int64_t t0 = wall_clock64();
  double* W = scr(s, TTX_SCR_LU);
  const double* bloc = scr(s, TTX_SCR_BLOC);
  if ((int64_t)kEngLds < 4 * N + 8 * (N > 4224 ? N : 4224)) { fail(s, TTX_SLOT_ERR_CAPACITY); return; }
  local_matrix_blk<double>(system_op(s, k), W, N + 1, lds, kEngLds / 8, false);
  for (int64_t r = threadIdx.x; r < N; r += blockDim.x) W[r * (N + 1) + N] = bloc[r];
  if (threadIdx.x == 0) s_info = TTX_OK;
  __syncthreads();
  stamp(s, 1, t0);
  block_lu_solve<kEngThreads>(N, N + 1, W, scr(s, TTX_SCR_U), lds, kEngLds, &s_info, &s->clk[13]);
  __syncthreads();
  stamp(s, 2, t0);
  if (s_info != TTX_OK) { fail(s, TTX_SLOT_ERR_SINGULAR); return; }
  if (threadIdx.x == 0) {
    scr(s, TTX_SCR_SCAL)[1] = 1.0;
    s->lu_solves += 1;
    s->sstage = kSolveNone;
  }
  __syncthreads();
  exact_local_residual(s, k, N, red, t0);
}

__device__ __forceinline__ void stage_solve(ttx_slot* s) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  solve_body(s);
  __syncthreads();
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[1] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_solve_kernel(ttx_slot* slots) {
  stage_solve(slots + blockIdx.x);
}

__device__ __forceinline__ void stage_chol(ttx_slot* s, double* lds) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  chol_body(s, lds);
  __syncthreads();
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[3] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_chol_kernel(ttx_slot* slots) {
  extern __shared__ double lds[];
  stage_chol(slots + blockIdx.x, lds);
}

// GMRES local solves (symmetric=False, or local_solver="gmres") with the
// reference's restarts (common.py:776-809): up to four gmres calls, each from
// the last iterate with rtol shrunk 1e-2 (floor 1e-15), accepted once
// ||A u - b_loc|| <= local_rtol ||b_loc||; the routines of gmres_blk.h on the
// slot's GMRES work, whole restart cycles (m + 1 operator applies) while the
// tick's budget of minres_slice applies lasts.
constexpr int64_t kGmresRestart = 20;

__device__ void gmres_body(ttx_slot* s) {
  __shared__ double red[16];
  if (!solve_step_active(s) || s->sstage != kSolveGmres) return;
  const int64_t k = s->k;
  const int64_t N = s->rx[k] * s->n[k] * s->rx[k + 1];
  const int64_t m = N < kGmresRestart ? N : kGmresRestart;
  int64_t t0 = wall_clock64();
  double* u = scr(s, TTX_SCR_U);
  const double* bloc = scr(s, TTX_SCR_BLOC);
  double* scal = scr(s, TTX_SCR_SCAL);
  double* st = scal + 8;    // attempt bookkeeping (the MINRES slots)
  double* gst = scal + 40;  // GMRES state (TTX_GMRES_STATE doubles)
  double* work = scr(s, TTX_SCR_GMRES);
  GmresView g(work, N, m);
  const LocOp op = system_op(s, k);
  if (st[kMrOuter] < 0.0) {
    // first call, warm-started from the current core (one_site.py:151-153)
    op.apply(xcore(s, k), 1, g.ax, 1.0, 0.0);
    gmres_init_blk(N, m, bloc, xcore(s, k), g.ax, st[kMrRtol], 10 * N, work, gst, red);
    if (threadIdx.x == 0) st[kMrOuter] = 0.0;
    __syncthreads();
  }
  int64_t budget = s->minres_slice > 0 ? s->minres_slice : 128;
  while (true) {
    while (gst[G_DONE] == 0.0 && budget > 0) {
      for (int64_t col = 0; col < m; ++col) {
        if (gst[G_SKIP] != 0.0) break;  // the inner loop has stopped
        op.apply(g.V + col * N, 1, g.V + (col + 1) * N, 1.0, 0.0);
        gmres_arnoldi_blk(N, m, col, work, gst, red);
        --budget;
      }
      const int64_t steps = (int64_t)gst[G_COL] + 1;  // Arnoldi steps of this cycle
      gmres_finish_blk(N, m, work, gst);
      op.apply(g.x, 1, g.ax, 1.0, 0.0);
      gmres_check_blk(N, m, bloc, work, gst, red);
      --budget;
      if (threadIdx.x == 0) s->minres_iters += steps;
    }
    if (gst[G_DONE] == 0.0) break;  // out of budget: resume next tick
    // the attempt is over: true residual ||b - A x|| from its last check
    const double res = gst[G_RNORM];
    const double target = st[kMrTarget];
    const int attempt = (int)st[kMrAttempt];
    const bool ok = res <= target;
    __syncthreads();  // everyone has read st before it is rewritten
    if (!ok && attempt >= 3 && s->allow_fallback && N <= s->direct_cap) {
      // the explicit-matrix fallback (common.py:793-809): LU in this kernel
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        s->local_fallbacks += 1;
        s->sstage = s->chol ? kSolveChol : kSolveLu;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    if (ok || attempt >= 3) {
      bcopy(u, g.x, N);
      if (threadIdx.x == 0) {
        s->local_solves += 1;
        if (!ok) s->local_inexact += 1;
        scal[1] = 0.0;
        scal[0] = s->residual_mode ? res : 0.0;  // ||A u - b_loc||, the exact residual
        s->sstage = kSolveNone;
      }
      __syncthreads();
      stamp(s, 3, t0);
      return;
    }
    // next attempt from the last iterate (x and A x are the check's)
    const double rt = fmax(st[kMrRtol] * 1e-2, 1e-15);
    if (threadIdx.x == 0) {
      st[kMrRtol] = rt;
      st[kMrAttempt] = (double)(attempt + 1);
    }
    gmres_init_blk(N, m, bloc, g.x, g.ax, rt, 10 * N, work, gst, red);
    if (budget <= 0) break;
  }
  stamp(s, 3, t0);
}

__device__ __forceinline__ void stage_lu(ttx_slot* s, double* lds) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  gmres_body(s);
  __syncthreads();
  lu_body(s, lds);
  __syncthreads();
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[4] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_lu_kernel(ttx_slot* slots) {
  extern __shared__ double lds[];
  stage_lu(slots + blockIdx.x, lds);
}

// truncation, enrichment and shifts after a solve; QR sweeps and z fit
__device__ __forceinline__ void post_body(ttx_slot* s) {
  __shared__ double red[16];
  if (s->status != TTX_SLOT_RUNNING) return;
  const int64_t ph = s->phase, d = s->d, k = s->k;
  double* enr = scr(s, TTX_SCR_ENR);
  const int64_t wcap = (s->rmax + s->kick) * s->nmax * (s->rmax + s->kick);
  double* nb_buf = enr + wcap;         // neighbour update
  double* qb = nb_buf + wcap;          // Q
  double* rbuf = qb + wcap;            // R
  double* tmp = rbuf + wcap;           // GEMM output before copy-back
  double* qrw = tmp + wcap;            // QR work (3 wcap)
  const int64_t n = s->n[k];

  if (ph == TTX_PH_XORTH || ph == TTX_PH_ZORTH || ph == TTX_PH_RED_R) {
    // right-orthonormalising QR step at k (utils.py:115-136)
    const bool onz = ph == TTX_PH_ZORTH;
    double* ck = onz ? zcore(s, k) : xcore(s, k);
    double* cp = onz ? zcore(s, k - 1) : xcore(s, k - 1);
    int64_t* bond = onz ? s->rz : s->rx;
    const int64_t rl = bond[k], rr = bond[k + 1];
    const int64_t rows_prev = bond[k - 1] * s->n[k - 1];
    const int64_t kk = shift_left_blk(ck, cp, ck, rl, n, rr, rows_prev, qb, rbuf, tmp, qrw, red);
    if (threadIdx.x == 0) bond[k] = kk;
    return;
  }
  if (ph == TTX_PH_RED_L) {
    // left-orthonormalising QR step at k (utils.py:86-112)
    const int64_t rl = s->rx[k], rr = s->rx[k + 1];
    const int64_t rest = s->n[k + 1] * s->rx[k + 2];
    const int64_t kk = shift_right_blk(xcore(s, k), xcore(s, k + 1), xcore(s, k), rl, n, rr, rest, qb,
                                       rbuf, tmp, qrw, red);
    if (threadIdx.x == 0) s->rx[k + 1] = kk;
    return;
  }
  if (ph == TTX_PH_FIT_FWD || ph == TTX_PH_FIT_BWD) {
    // residual-train fit (amen.py:162-224)
    residual_core_blk(s, k, xcore(s, k), red);
    double* zn = scr(s, TTX_SCR_ZNEW);
    const int64_t zl = s->rz[k], zr = s->rz[k + 1];
    if (ph == TTX_PH_FIT_FWD) {
      const int64_t kk = shift_right_blk(zcore(s, k), zcore(s, k + 1), zn, zl, n, zr,
                                         s->n[k + 1] * s->rz[k + 2], qb, rbuf, tmp, qrw, red);
      if (threadIdx.x == 0) s->rz[k + 1] = kk;
    } else if (k > 0) {
      const int64_t kk = shift_left_blk(zcore(s, k), zcore(s, k - 1), zn, zl, n, zr,
                                        s->rz[k - 1] * s->n[k - 1], qb, rbuf, tmp, qrw, red);
      if (threadIdx.x == 0) s->rz[k] = kk;
    } else {
      bcopy(zcore(s, 0), zn, zl * n * zr);
    }
    return;
  }
  if (ph != TTX_PH_FWD && ph != TTX_PH_BWD) return;
  if (s->sstage != kSolveNone) return;  // the local solve is still running

  const int64_t rl = s->rx[k], rr = s->rx[k + 1];
  const int64_t N = rl * n * rr;
  const double* u = scr(s, TTX_SCR_U);
  const double* bloc = scr(s, TTX_SCR_BLOC);
  const double* scal = scr(s, TTX_SCR_SCAL);
  const double exact = scal[0];
  int64_t t0 = wall_clock64();
  residual_core_blk(s, k, u, red);  // z from the untruncated solution (amen.py:601-613)
  double* zn = scr(s, TTX_SCR_ZNEW);
  const int64_t zl = s->rz[k], zr = s->rz[k + 1];
  __syncthreads();
  stamp(s, 8, t0);

  if (ph == TTX_PH_BWD && k == 0) {
    // last core of the backward sweep is stored as solved (amen.py:654-677)
    bcopy(xcore(s, 0), u, N);
    bcopy(zcore(s, 0), zn, zl * n * zr);
    return;
  }
  const LocOp op = system_op(s, k);
  const bool fwd = ph == TTX_PH_FWD;
  // SVD of (rl n x rr) forward or (rl x n rr) backward (amen.py:254-260,301-307)
  const int64_t p = fwd ? rl * n : rl;
  const int64_t q = fwd ? rr : n * rr;
  const int64_t K = p < q ? p : q;
  double* U = scr(s, TTX_SCR_USVD);
  double* S = scr(s, TTX_SCR_SSVD);
  double* Vt = scr(s, TTX_SCR_VSVD);
  __shared__ int32_t s_svd;
  if (threadIdx.x == 0) s_svd = TTX_OK;
  __syncthreads();
  SvdArgs sa{p, q, u, q, 1, U, K, 1, S, Vt, q, 1, scr(s, TTX_SCR_SVDW), &s_svd};
  block_svd<false>(sa, scr(s, TTX_SCR_SVDW), red, eng_lds, (int64_t)(kEngLds / 8));
  __syncthreads();
  stamp(s, 9, t0);
  if (s_svd == TTX_ENONFINITE) { fail(s, TTX_SLOT_ERR_NONFINITE); return; }
  if (s_svd == TTX_ENOCONV && threadIdx.x == 0) s->svd_noconv += 1;
  int capped = 0;
  const int64_t rank = choose_rank_blk(s, op, bloc, U, S, Vt, p, q, K, exact, &capped, red);
  __syncthreads();
  stamp(s, 10, t0);
  const int64_t mr = s->max_rank;
  if (fwd) {
    // core = U_r, carry = S V_r, block = Psi_xz^T z (amen.py:614-627), enrich (amen.py:325-360)
    const int64_t W = rank + zr;
    const int64_t m = rl * n;
    for (int64_t e = threadIdx.x; e < m * rank; e += blockDim.x) {
      const int64_t row = q32(e, rank), c = e - row * rank;
      enr[row * W + c] = U[row * K + c];
    }
    bgemm(n, rl, zr, zl, 1.0, envp(s, 4, 0, k), 1, rl, 0, zn, n * zr, 1, zr, 0.0, enr + rank, n * W,
          1, W);
    double* carry = tmp;
    for (int64_t e = threadIdx.x; e < rank * rr; e += blockDim.x) {
      const int64_t c = q32(e, rr);
      carry[e] = S[c] * Vt[e];
    }
    __syncthreads();
    const int64_t rest = s->n[k + 1] * s->rx[k + 2];
    bgemm(1, rank, rest, rr, 1.0, carry, rr, 1, 0, xcore(s, k + 1), rest, 1, 0, 0.0, nb_buf, rest, 1, 0);
    bfill(nb_buf + rank * rest, zr * rest, 0.0);
    __syncthreads();
    const int64_t kk = m < W ? m : W;
    block_qr(m, W, enr, W, 1, qb, kk, 1, rbuf, W, 1, qrw, red);
    const int64_t kc = (mr > 0 && kk > mr) ? mr : kk;
    if (kc > s->rmax) { fail(s, TTX_SLOT_ERR_CAPACITY); return; }
    double* xk = xcore(s, k);
    for (int64_t e = threadIdx.x; e < m * kc; e += blockDim.x) {
      const int64_t row = q32(e, kc), c = e - row * kc;
      xk[e] = qb[row * kk + c];
    }
    bgemm(1, kc, rest, W, 1.0, rbuf, W, 1, 0, nb_buf, rest, 1, 0, 0.0, xcore(s, k + 1), rest, 1, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
      s->rx[k + 1] = kc;
      s->rank_decisions += 2;
      s->rank_capped += capped + (kc < kk ? 1 : 0);
    }
    stamp(s, 11, t0);
    // z centre moves right with the fresh core (amen.py:628)
    const int64_t kz = shift_right_blk(zcore(s, k), zcore(s, k + 1), zn, zl, n, zr,
                                       s->n[k + 1] * s->rz[k + 2], qb, rbuf, tmp, qrw, red);
    if (threadIdx.x == 0) s->rz[k + 1] = kz;
    stamp(s, 12, t0);
    return;
  }
  // backward: core = V_r, carry = U_r S, block = z Phi_xz (amen.py:679-703), enrich (amen.py:363-392)
  const int64_t W = rank + zl;
  const int64_t w = n * rr;
  bcopy(enr, Vt, rank * w);
  bgemm(1, zl * n, rr, zr, 1.0, zn, zr, 1, 0, envp(s, 4, 1, k), rr, 1, 0, 0.0, enr + rank * w, rr, 1, 0);
  double* carry = tmp;
  for (int64_t e = threadIdx.x; e < rl * rank; e += blockDim.x) {
    const int64_t i = q32(e, rank), c = e - i * rank;
    carry[e] = U[i * K + c] * S[c];
  }
  const int64_t rows_prev = s->rx[k - 1] * s->n[k - 1];
  bfill(nb_buf, rows_prev * W, 0.0);
  __syncthreads();
  bgemm(1, rows_prev, rank, rl, 1.0, xcore(s, k - 1), rl, 1, 0, carry, rank, 1, 0, 0.0, nb_buf, W, 1, 0);
  __syncthreads();
  const int64_t kk = w < W ? w : W;
  block_qr(w, W, enr, 1, w, qb, 1, w, rbuf, W, 1, qrw, red);  // qb = Q^T (kk x w)
  const int64_t kc = (mr > 0 && kk > mr) ? mr : kk;
  if (kc > s->rmax) { fail(s, TTX_SLOT_ERR_CAPACITY); return; }
  bcopy(xcore(s, k), qb, kc * w);
  bgemm(1, rows_prev, kc, W, 1.0, nb_buf, W, 1, 0, rbuf, 1, W, 0, 0.0, xcore(s, k - 1), kc, 1, 0);
  __syncthreads();
  if (threadIdx.x == 0) {
    s->rx[k] = kc;
    s->rank_decisions += 2;
    s->rank_capped += capped + (kc < kk ? 1 : 0);
  }
  stamp(s, 11, t0);
  const int64_t kz = shift_left_blk(zcore(s, k), zcore(s, k - 1), zn, zl, n, zr,
                                    s->rz[k - 1] * s->n[k - 1], qb, rbuf, tmp, qrw, red);
  if (threadIdx.x == 0) s->rz[k] = kz;
  stamp(s, 12, t0);
}

__device__ __forceinline__ void stage_post(ttx_slot* s) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  int64_t t0 = wall_clock64();
  post_body(s);
  __syncthreads();
  if (s->status == TTX_SLOT_RUNNING || s->status == TTX_SLOT_DONE) stamp(s, 5, t0);
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[5] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_post_kernel(ttx_slot* slots) {
  stage_post(slots + blockIdx.x);
}

// ||A x - b|| / ||b|| in TT arithmetic (common.py:118-146): the residual
// train is formed core by core and normed by a Q-less right QR sweep.  The
// sweep resumes across ticks: s->chk_core is the next core (d-1 .. 0) and the
// carried R^T factor (s->chk_cdim columns) stays in scratch; cores are taken
// until kCheckSliceWork (m L^2 of the QRs) is spent.  Returns true when the
// norm is complete (in *nrm).
constexpr double kCheckSliceWork = 1.5e8;

__device__ bool residual_norm_blk(ttx_slot* s, double* red, double* nrm) {
  const int64_t d = s->d;
  double* base = scr(s, TTX_SCR_CHK);
  const int64_t L2 = s->lmax * s->lmax;
  double* T = base;                       // L x n x R
  double* Wb = T + L2 * s->nmax;          // L x n x c
  double* C = Wb + L2 * s->nmax;          // R x c
  double* Rb = C + L2;                    // R factor
  double* qw = Rb + L2;                   // QR work
  int64_t i = s->chk_core, cdim = s->chk_cdim;
  if (i < 0) {
    i = d - 1;
    cdim = 1;
    if (threadIdx.x == 0) C[0] = 1.0;
  }
  __syncthreads();
  double work = 0.0;
  for (; i >= 0; --i) {
    const int64_t n = s->n[i];
    const int64_t Ra = s->ra[i], Rb2 = s->ra[i + 1], rx = s->rx[i], rx2 = s->rx[i + 1];
    const int64_t rbl = s->rb[i], rbr = s->rb[i + 1];
    const int64_t L = i == 0 ? 1 : Ra * rx + rbl;
    const int64_t R = i == d - 1 ? 1 : Rb2 * rx2 + rbr;
    if (work > 0.0 && work + (double)(n * cdim) * L * L > kCheckSliceWork) break;  // next tick
    work += (double)(n * cdim) * L * L;
    bfill(T, L * n * R, 0.0);
    __syncthreads();
    const double* A = s->A[i];
    const double* xc = xcore(s, i);
    const int64_t Lx = Ra * rx, Rx = Rb2 * rx2;
    for (int64_t e = threadIdx.x; e < Lx * n * Rx; e += blockDim.x) {
      const int64_t e_q = q32(e, Rx), r = e - e_q * Rx, l = q32(e_q, n), j = e_q - l * n;
      const int64_t a = q32(l, rx), m = l - a * rx, b = q32(r, rx2), nn = r - b * rx2;
      double t = 0.0;
      for (int64_t jj = 0; jj < n; ++jj) t += A[((a * n + j) * n + jj) * Rb2 + b] * xc[(m * n + jj) * rx2 + nn];
      T[(l * n + j) * R + r] = t;
    }
    const double sg = (i == s->bneg) ? -1.0 : 1.0;
    const double* bc = s->b[i];
    const int64_t offl = L - rbl, offr = R - rbr;
    for (int64_t e = threadIdx.x; e < rbl * n * rbr; e += blockDim.x) {
      const int64_t e_q = q32(e, rbr), r = e - e_q * rbr, l = q32(e_q, n), j = e_q - l * n;
      T[((offl + l) * n + j) * R + offr + r] = sg * bc[e];
    }
    __syncthreads();
    bgemm(1, L * n, cdim, R, 1.0, T, R, 1, 0, C, cdim, 1, 0, 0.0, Wb, cdim, 1, 0);
    __syncthreads();
    if (i == 0) {
      *nrm = bnorm2(Wb, nullptr, L * n * cdim, red);
      if (threadIdx.x == 0) s->chk_core = -1;
      return true;
    }
    // QR of the transposed unfolding ((n c) x L), keep R^T as the next carry
    const int64_t m = n * cdim;
    const int64_t kk = m < L ? m : L;
    block_qr(m, L, Wb, 1, m, nullptr, 0, 0, Rb, L, 1, qw, red);
    for (int64_t e = threadIdx.x; e < L * kk; e += blockDim.x) {
      const int64_t j = q32(e, kk), c = e - j * kk;
      C[e] = Rb[c * L + j];
    }
    cdim = kk;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    s->chk_core = i;
    s->chk_cdim = cdim;
  }
  return false;
}

__device__ void check_decide(ttx_slot* s, int64_t sw, double res);

__device__ void advance(ttx_slot* s) {
  const int64_t d = s->d;
  switch (s->phase) {
    case TTX_PH_XORTH:
      if (--s->k < 1) { s->phase = TTX_PH_ZORTH; s->k = d - 1; }
      break;
    case TTX_PH_ZORTH:
      if (--s->k < 1) { s->phase = TTX_PH_FIT_INIT; s->k = d - 1; }
      break;
    case TTX_PH_FIT_INIT:
      if (--s->k < 0) { s->phase = TTX_PH_FIT_FWD; s->k = 0; }
      break;
    case TTX_PH_FIT_FWD:
      if (++s->k >= d - 1) { s->phase = TTX_PH_FIT_LAST; s->k = d - 1; }
      break;
    case TTX_PH_FIT_LAST: s->phase = TTX_PH_FIT_BWD; s->k = d - 1; break;
    case TTX_PH_FIT_BWD:
      if (--s->k < 0) { s->phase = TTX_PH_INIT_R; s->k = d - 1; }
      break;
    case TTX_PH_INIT_R:
      if (--s->k < 0) { s->phase = TTX_PH_FWD; s->k = 0; s->sweep = 0; }
      break;
    case TTX_PH_FWD:
      if (++s->k >= d - 1) { s->phase = TTX_PH_LAST; s->k = d - 1; }
      break;
    case TTX_PH_LAST: s->phase = TTX_PH_BWD; s->k = d - 1; break;
    case TTX_PH_BWD:
      if (--s->k < 0) { s->phase = TTX_PH_CHECK; s->k = 0; }
      break;
    case TTX_PH_RED_L:
      if (++s->k >= s->red_last) { s->phase = TTX_PH_RED_R; s->k = s->red_last; }
      break;
    case TTX_PH_RED_R:
      if (--s->k < 1) { s->phase = TTX_PH_DONE; s->status = TTX_SLOT_DONE; }
      break;
    default: break;
  }
}

// per-sweep monitor (common.py:504-592) and the phase advance of every slot
__device__ void check_body(ttx_slot* s) {
  __shared__ double red[16];
  if (s->status != TTX_SLOT_RUNNING) return;
  if (s->phase != TTX_PH_CHECK) {
    if (s->sstage != kSolveNone) return;  // the local solve resumes next tick
    __syncthreads();
    if (threadIdx.x == 0) {
      s->ticks += 1;
      advance(s);
    }
    return;
  }
  const int64_t sw = s->sweep;
  double res = 0.0;
  if (s->track) {
    double rn = 0.0;
    if (!residual_norm_blk(s, red, &rn)) return;  // the sweep resumes next tick
    res = s->bnorm == 0.0 ? rn : rn / s->bnorm;
  }
  __syncthreads();
  if (threadIdx.x == 0) check_decide(s, sw, res);
}

__device__ __forceinline__ void stage_check(ttx_slot* s) {
  if (threadIdx.x == 0) g_flops = 0.0;
  __syncthreads();
  int64_t t0 = wall_clock64();
  check_body(s);
  __syncthreads();
  if (s->status == TTX_SLOT_RUNNING || s->status == TTX_SLOT_DONE) stamp(s, 6, t0);
  if (threadIdx.x == 0 && g_flops != 0.0) s->flops[6] += g_flops;
}

__global__ __launch_bounds__(kEngThreads, 2) void eng_check_kernel(ttx_slot* slots) {
  stage_check(slots + blockIdx.x);
}

// A pass: the seven stages of a tick for slot s, in the tick kernels' order
// (the five environment kinds one after another).  A slot's stages read and
// write only that slot, so a pass gives the bits a tick of the seven kernels
// gives.  The LDS operand of the stages is the kernel's dynamic LDS, named
// here (not passed in) so that a pass compiled as a function of its own
// still reads it with LDS instructions.
// Each stage is a function of its own, so that its registers are allocated
// as in its tick kernel (one function for all seven spills hundreds).
__device__ __noinline__ void call_env(ttx_slot* s, int kind) { stage_env(s, kind, false); }
__device__ __noinline__ void call_solve(ttx_slot* s) { stage_solve(s); }
__device__ __noinline__ void call_chol32(ttx_slot* s) { stage_chol32(s); }
__device__ __noinline__ void call_chol(ttx_slot* s) { stage_chol(s, eng_lds); }
__device__ __noinline__ void call_lu(ttx_slot* s) { stage_lu(s, eng_lds); }
__device__ __noinline__ void call_post(ttx_slot* s) { stage_post(s); }
__device__ __noinline__ void call_check(ttx_slot* s) { stage_check(s); }

__device__ __forceinline__ void run_pass(ttx_slot* s) {
  for (int kind = 0; kind < 5; ++kind) call_env(s, kind);
  call_solve(s);
  call_chol32(s);
  call_chol(s);
  call_lu(s);
  call_post(s);
  call_check(s);
}

// A whole tick in one launch: workgroup b runs a pass for slot b without
// waiting for the other slots between stages.
__global__ __launch_bounds__(kEngThreads, 2) void eng_pass_kernel(ttx_slot* slots) {
  ttx_slot* s = slots + blockIdx.x;
  if (s->status != TTX_SLOT_RUNNING) return;  // uniform across the workgroup
  run_pass(s);
}



// Many passes per launch, each slot at its own pace: workgroup b runs passes
// for slot b until the slot stops running, `passes` passes are done, or (with
// budget > 0) `budget` ticks of the 100 MHz wall clock have passed since it
// started (checked between passes; one pass at least).  Thread 0 decides and
// every thread reads the decision after a barrier.
__global__ __launch_bounds__(kEngThreads, 2) void eng_run_kernel(ttx_slot* slots, int64_t passes,
                                                                 int64_t budget) {
  __shared__ int s_go;
  ttx_slot* s = slots + blockIdx.x;
  const int64_t start = wall_clock64();
  for (int64_t p = 0; p < passes; ++p) {
    __syncthreads();  // every thread has read the last pass's s_go
    if (threadIdx.x == 0)
      s_go = s->status == TTX_SLOT_RUNNING &&
             (p == 0 || budget <= 0 || wall_clock64() - start < budget);
    __syncthreads();
    if (!s_go) break;
    run_pass(s);
  }
}

// the stopping test and the next phase after a completed sweep (thread 0)
__device__ void check_decide(ttx_slot* s, int64_t sw, double res) {
  s->ticks += 1;
  int64_t mb = 0;
  for (int64_t i = 0; i <= s->d; ++i) mb = s->rx[i] > mb ? s->rx[i] : mb;
  if (sw < TTX_MAXSWEEPS) {
    s->max_bonds[sw] = mb;
    s->residuals[sw] = res;
  }
  int64_t stop = 0;
  if (s->track) {
    if (s->residual_tol > 0.0 && res <= s->residual_tol) stop = 1;
    else if (s->stagnation > 0.0 && sw >= 1 && res > s->stagnation * s->residuals[sw - 1]) stop = 2;
  }
  if (!stop && sw + 1 >= s->max_sweeps) stop = 3;
  if (!stop) {
    s->sweep = sw + 1;
    s->phase = TTX_PH_FWD;
    s->k = 0;
    return;
  }
  s->stop_reason = stop;
  // bonds above the full-rank bound left by the last enrichment (amen.py:710-719)
  int64_t last = 0;
  double prodn = 1.0;
  for (int64_t i = 1; i < s->d; ++i) {
    prodn *= (double)s->n[i - 1];
    if ((double)s->rx[i] > prodn) last = i;
  }
  if (last > 0) {
    s->red_last = last;
    s->phase = TTX_PH_RED_L;
    s->k = 0;
  } else {
    s->phase = TTX_PH_DONE;
    s->status = TTX_SLOT_DONE;
  }
}

}  // namespace ttx

using namespace ttx;

extern "C" int64_t ttx_engine_slot_bytes(void) { return (int64_t)sizeof(ttx_slot); }

extern "C" int64_t ttx_engine_field_offset(int which) {
  switch (which) {
    case 0: return (int64_t)offsetof(ttx_slot, n);
    case 1: return (int64_t)offsetof(ttx_slot, residual_tol);
    case 2: return (int64_t)offsetof(ttx_slot, cap_core);
    case 3: return (int64_t)offsetof(ttx_slot, A);
    case 4: return (int64_t)offsetof(ttx_slot, rx);
    case 5: return (int64_t)offsetof(ttx_slot, residuals);
    case 6: return (int64_t)offsetof(ttx_slot, max_bonds);
    default: return -1;
  }
}

// Fill the layout fields of a host-side slot from its problem fields.
extern "C" int ttx_engine_plan(ttx_slot* s) {
  if (s->d < 2 || s->d > TTX_MAXD || s->rmax < 1 || s->kick < 1 || s->chunk < 1 ||
      s->chunk > kRankChunkMax)
    return TTX_EINVAL;
  int64_t nmax = 1, ramax = 1, rbmax = 1;
  for (int64_t i = 0; i < s->d; ++i) nmax = s->n[i] > nmax ? s->n[i] : nmax;
  for (int64_t i = 0; i <= s->d; ++i) {
    ramax = s->ra[i] > ramax ? s->ra[i] : ramax;
    rbmax = s->rb[i] > rbmax ? s->rb[i] : rbmax;
  }
  const int64_t rmax = s->rmax, kick = s->kick;
  s->nmax = nmax;
  s->ramax = ramax;
  s->rbmax = rbmax;
  s->cap_core = rmax * nmax * rmax;
  s->cap_zcore = kick * nmax * kick;
  s->cap_loc = s->cap_core;
  // the explicit-matrix fallback of an iterative solve (allow_fallback) may
  // factor a local system of any size the slot holds
  s->direct_cap = (s->allow_fallback || s->direct_size > s->cap_loc) ? s->cap_loc : s->direct_size;
  if (s->direct_cap < 1) s->direct_cap = 1;
  const int64_t zr = kick > rmax ? kick : rmax;
  s->cap_env[0] = rmax * ramax * rmax;
  s->cap_env[1] = rbmax * rmax;
  s->cap_env[2] = rbmax * kick;
  s->cap_env[3] = rmax * ramax * kick;
  s->cap_env[4] = kick * zr;
  int64_t off = 0;
  for (int kind = 0; kind < 5; ++kind)
    for (int side = 0; side < 2; ++side) {
      s->env_off[kind * 2 + side] = off;
      off += s->d * s->cap_env[kind];
    }
  s->env_size = off;
  s->lmax = ramax * rmax + rbmax;
  const int64_t L2 = s->lmax * s->lmax;
  const int64_t cl = s->cap_loc;
  const int64_t wcap = (rmax + kick) * nmax * (rmax + kick);
  int64_t sz[TTX_SCR_COUNT];
  sz[TTX_SCR_U] = cl;
  sz[TTX_SCR_BLOC] = cl;
  sz[TTX_SCR_TMP] = cl > wcap ? cl : wcap;
  sz[TTX_SCR_ZNEW] = kick * nmax * kick;
  sz[TTX_SCR_MAT] = 0;  // the explicit matrix is built straight into the LU buffer
  sz[TTX_SCR_LU] = s->direct_cap * (s->direct_cap + 1);
  sz[TTX_SCR_MR] = 9 * cl;
  const int64_t sapply = s->chunk * cl * ramax;
  sz[TTX_SCR_S1] = sapply > cl * ramax * 2 ? sapply : cl * ramax * 2;
  sz[TTX_SCR_S2] = sz[TTX_SCR_S1];
  sz[TTX_SCR_SVDW] = 2 * cl + 2 * rmax * rmax * nmax * nmax + 3 * rmax * nmax + 8;
  sz[TTX_SCR_USVD] = cl;
  sz[TTX_SCR_SSVD] = rmax * nmax;
  sz[TTX_SCR_VSVD] = cl;
  sz[TTX_SCR_ENR] = 8 * wcap;
  const int64_t chk_res = (4 * nmax + 3) * L2 + 2 * s->lmax + 8;
  const int64_t chk_rank = 2 * s->chunk * cl;
  sz[TTX_SCR_CHK] = chk_res > chk_rank ? chk_res : chk_rank;
  sz[TTX_SCR_ENVS] = 5 * 2 * s->cap_core * ramax;
  sz[TTX_SCR_SCAL] = 64;
  // mixed requests: FP32 psi, phi, A_k, the two products' scratch and two vectors
  sz[TTX_SCR_M32] = s->mixed ? (2 * s->cap_env[0] + ramax * nmax * nmax * ramax + 2 * cl * ramax +
                                2 * cl + 1) / 2
                             : 0;
  // GMRES slots: the Krylov basis and the rest of the GMRES work for the
  // largest local system (restart min(20, N))
  sz[TTX_SCR_GMRES] = s->iterative == 2 ? gmres_work_doubles(cl, cl < kGmresRestart ? cl : kGmresRestart)
                                        : 0;
  off = 0;
  for (int i = 0; i < TTX_SCR_COUNT; ++i) {
    s->scr_off[i] = off;
    off += (sz[i] + 31) & ~(int64_t)31;  // 256-byte aligned regions
  }
  s->scr_size = off;
  return TTX_OK;
}

static int engine_prepare() {
  static bool attr = false;
  if (!attr) {
    const void* fns[7] = {(const void*)eng_chol32_kernel, (const void*)eng_chol_kernel,
                          (const void*)eng_lu_kernel,     (const void*)eng_post_kernel,
                          (const void*)eng_check_kernel,  (const void*)eng_pass_kernel,
                          (const void*)eng_run_kernel};
    for (const void* f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kEngLds) !=
          hipSuccess)
        return TTX_EHIP;
    attr = true;
  }
  return TTX_OK;
}

// kernel i of a tick: env, solve (rhs + MINRES), FP32 Cholesky, Cholesky, LU, post, check
static void launch_stage(int i, hipStream_t st, ttx_slot* slots, int64_t nslots) {
  const dim3 g((unsigned)nslots), b(kEngThreads);
  switch (i) {
    case 0: hipLaunchKernelGGL(eng_env_kernel, dim3((unsigned)nslots, 5), b, 0, st, slots); break;
    case 1: hipLaunchKernelGGL(eng_solve_kernel, g, b, 0, st, slots); break;
    case 2: hipLaunchKernelGGL(eng_chol32_kernel, g, b, kEngLds, st, slots); break;
    case 3: hipLaunchKernelGGL(eng_chol_kernel, g, b, kEngLds, st, slots); break;
    case 4: hipLaunchKernelGGL(eng_lu_kernel, g, b, kEngLds, st, slots); break;
    case 5: hipLaunchKernelGGL(eng_post_kernel, g, b, kEngLds, st, slots); break;
    default: hipLaunchKernelGGL(eng_check_kernel, g, b, kEngLds, st, slots); break;
  }
}

// The kernels of a tick hand a local solve on through slot.sstage, and every
// kernel returns at once for a slot whose sstage names another kernel's solve
// (env, post and check wait for kSolveNone), so a slot waiting on a kernel
// that a tick leaves out is untouched by that tick.
// With TTX_ENGINE_FUSED the whole tick is one eng_pass_kernel launch.
static void launch_tick(hipStream_t st, ttx_slot* slots, int64_t nslots, uint32_t stages) {
  if (stages & TTX_ENGINE_FUSED) {
    hipLaunchKernelGGL(eng_pass_kernel, dim3((unsigned)nslots), dim3(kEngThreads), kEngLds, st, slots);
    return;
  }
  for (int i = 0; i < TTX_ENGINE_KERNELS; ++i)
    if ((stages >> i) & 1u) launch_stage(i, st, slots, nslots);
}

extern "C" int ttx_engine_tick(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t ticks,
                               uint32_t stages) {
  if (nslots < 1 || ticks < 0 || (stages & ~(uint32_t)(TTX_ENGINE_STAGES_ALL | TTX_ENGINE_FUSED)) != 0) return TTX_EINVAL;
  if (engine_prepare() != TTX_OK) return TTX_EHIP;
  hipStream_t st = (hipStream_t)stream;
  for (int64_t t = 0; t < ticks; ++t) launch_tick(st, slots, nslots, stages);
  return launch_status();
}

extern "C" int ttx_engine_run(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t passes,
                              int64_t budget) {
  if (nslots < 1 || passes < 0) return TTX_EINVAL;
  if (engine_prepare() != TTX_OK) return TTX_EHIP;
  if (passes == 0) return TTX_OK;
  hipLaunchKernelGGL(eng_run_kernel, dim3((unsigned)nslots), dim3(kEngThreads), kEngLds,
                     (hipStream_t)stream, slots, passes, budget);
  return launch_status();
}

// `ticks` ticks of the engine recorded once as a HIP graph.  The kernels'
// only argument is the device slot table, which outlives the graph, and every
// per-request change goes through that table, so one recording serves every
// later poll of the engine.
struct EngineGraph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int ttx_engine_graph_create(ttx_slot* slots, int64_t nslots, int64_t ticks,
                                       uint32_t stages, void** graph_out) {
  if (nslots < 1 || ticks < 1 || graph_out == nullptr ||
      (stages & ~(uint32_t)(TTX_ENGINE_STAGES_ALL | TTX_ENGINE_FUSED)) != 0)
    return TTX_EINVAL;
  if (engine_prepare() != TTX_OK) return TTX_EHIP;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return TTX_EHIP;
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = TTX_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = TTX_EHIP;
  } else {
    for (int64_t t = 0; t < ticks; ++t) launch_tick(cs, slots, nslots, stages);
    if (hipStreamEndCapture(cs, &g) != hipSuccess) rc = TTX_EHIP;
  }
  if (rc == TTX_OK && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess)
    rc = TTX_EHIP;
  hipStreamDestroy(cs);
  if (rc != TTX_OK) {
    if (g != nullptr) hipGraphDestroy(g);
    return rc;
  }
  *graph_out = new EngineGraph{g, exec};
  return TTX_OK;
}

extern "C" int ttx_engine_graph_launch(ttx_stream_t stream, void* graph) {
  if (graph == nullptr) return TTX_EINVAL;
  EngineGraph* eg = (EngineGraph*)graph;
  return hipGraphLaunch(eg->exec, (hipStream_t)stream) == hipSuccess ? TTX_OK : TTX_EHIP;
}

extern "C" int ttx_engine_graph_destroy(void* graph) {
  if (graph == nullptr) return TTX_OK;
  EngineGraph* eg = (EngineGraph*)graph;
  hipGraphExecDestroy(eg->exec);
  hipGraphDestroy(eg->graph);
  delete eg;
  return TTX_OK;
}

extern "C" int ttx_engine_profile(ttx_stream_t stream, ttx_slot* slots, int64_t nslots, int64_t ticks,
                                  double* ms_out) {
  if (nslots < 1 || ticks < 1) return TTX_EINVAL;
  if (engine_prepare() != TTX_OK) return TTX_EHIP;
  hipStream_t st = (hipStream_t)stream;
  constexpr int K = TTX_ENGINE_KERNELS;
  hipEvent_t ev[K + 1];
  for (int i = 0; i <= K; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) return TTX_EHIP;
  int rc = TTX_OK;
  for (int64_t t = 0; t < ticks && rc == TTX_OK; ++t) {
    hipEventRecord(ev[0], st);
    for (int i = 0; i < K; ++i) {
      launch_stage(i, st, slots, nslots);
      hipEventRecord(ev[i + 1], st);
    }
    if (hipEventSynchronize(ev[K]) != hipSuccess) { rc = TTX_EHIP; break; }
    for (int j = 0; j < K; ++j) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[j], ev[j + 1]) != hipSuccess) { rc = TTX_EHIP; break; }
      ms_out[j] += (double)ms;
    }
  }
  for (int i = 0; i <= K; ++i) hipEventDestroy(ev[i]);
  return rc == TTX_OK ? launch_status() : rc;
}
