// Block-level (one workgroup) dense FP64 building blocks shared by the
// single-matrix kernels and the batched solver engine.
#pragma once

#include <float.h>

#include "ttx_common.h"

namespace ttx {

constexpr double kEps = DBL_EPSILON;

// Algorithmic flop tally of the block-level routines of one workgroup
// (thread 0 adds; the engine kernels zero it on entry and bank it per slot).
static __shared__ double g_flops;
__device__ __forceinline__ void add_flops(double f) {
  if (threadIdx.x == 0) g_flops += f;
}

// ---------------------------------------------------------------------------
// Householder QR on a column-major m x n matrix W (leading dim m) held in
// global memory.  On return W holds R in its upper triangle and the reflector
// vectors below the diagonal; tau[j] the reflector scales.  Q (m x k,
// column-major, leading dim m) is formed in Qw.  Block-cooperative.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_householder_qr(int64_t m, int64_t n, double* W, double* tau, double* Qw,
                                     double* red, bool form_q = true) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const int64_t k = m < n ? m : n;
  __shared__ double s_tau;
  {
    const double dm = (double)m, dn = (double)n, dk = (double)k;
    add_flops(2.0 * dm * dn * dk - (dm + dn) * dk * dk + 2.0 * dk * dk * dk / 3.0);
    if (form_q) add_flops(4.0 * dm * dk * dk - 4.0 * dk * dk * dk / 3.0);
  }
  for (int64_t j = 0; j < k; ++j) {
    double* col = W + j * m;
    double s = 0.0;
    for (int64_t i = j + 1 + tid; i < m; i += blockDim.x) s += col[i] * col[i];
    s = block_sum(s, red);
    const double alpha = col[j];
    __syncthreads();
    if (s == 0.0) {
      if (tid == 0) { s_tau = 0.0; tau[j] = 0.0; }
    } else {
      const double beta = -copysign(sqrt(alpha * alpha + s), alpha);
      const double scal = 1.0 / (alpha - beta);
      for (int64_t i = j + 1 + tid; i < m; i += blockDim.x) col[i] *= scal;
      if (tid == 0) {
        s_tau = (beta - alpha) / beta;
        tau[j] = s_tau;
        col[j] = beta;
      }
    }
    __syncthreads();
    const double t = s_tau;
    if (t != 0.0) {
      for (int64_t c = j + 1 + wave; c < n; c += nwave) {
        double* cc = W + c * m;
        double d = 0.0;
        for (int64_t i = j + 1 + lane; i < m; i += 64) d += col[i] * cc[i];
        d = wave_sum(d) + cc[j];
        d *= t;
        if (lane == 0) cc[j] -= d;
        for (int64_t i = j + 1 + lane; i < m; i += 64) cc[i] -= d * col[i];
      }
    }
    __syncthreads();
  }
  if (!form_q) return;
  // Q = H_0 ... H_{k-1} [I_k; 0], accumulated backwards
  for (int64_t e = tid; e < m * k; e += blockDim.x) {
    const int64_t i = e % m, c = e / m;
    Qw[e] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int64_t j = k - 1; j >= 0; --j) {
    const double t = tau[j];
    const double* v = W + j * m;
    if (t != 0.0) {
      for (int64_t c = j + wave; c < k; c += nwave) {
        double* qc = Qw + c * m;
        double d = 0.0;
        for (int64_t i = j + 1 + lane; i < m; i += 64) d += v[i] * qc[i];
        d = wave_sum(d) + qc[j];
        d *= t;
        if (lane == 0) qc[j] -= d;
        for (int64_t i = j + 1 + lane; i < m; i += 64) qc[i] -= d * v[i];
      }
    }
    __syncthreads();
  }
}

struct SvdArgs {
  int64_t m, n;
  const double* A;
  int64_t a_rs, a_cs;
  double* U;
  int64_t u_rs, u_cs;
  double* S;
  double* Vt;
  int64_t v_rs, v_cs;
  double* work;
  int32_t* info;
};

__device__ __forceinline__ int64_t rr_player(int64_t t, int64_t round, int64_t qq) {
  return t == 0 ? 0 : 1 + (t - 1 + round) % (qq - 1);
}


// Thin SVD of one matrix by the whole block: QR of the tall side, then
// one-sided Jacobi on the square factor (round-robin pairs, one wave per
// pair), columns sorted by descending singular value.  `work` holds
// 2*P*Q + 2*Q*Q + 3*Q doubles (P >= Q the long and short sides); it may be LDS.
__device__ __forceinline__ void block_svd(const SvdArgs& p, double* work, double* red) {
  __shared__ int s_rot;
  __shared__ int s_bad;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const bool wide = p.m < p.n;
  // X = A (tall) or A^T (wide): P x Q with P >= Q
  const int64_t P = wide ? p.n : p.m;
  const int64_t Q = wide ? p.m : p.n;
  const int64_t x_rs = wide ? p.a_cs : p.a_rs;
  const int64_t x_cs = wide ? p.a_rs : p.a_cs;
  // Jacobi on the Q x Q factor and U = Q_x U_r (the QR of the tall side counts itself)
  add_flops(20.0 * (double)Q * (double)Q * (double)Q + 2.0 * (double)P * (double)Q * (double)Q);
  double* Wq = work;              // P x Q   (QR working matrix)
  double* tau = Wq + P * Q;       // Q
  double* Qx = tau + Q;           // P x Q   (orthonormal factor)
  double* G = Qx + P * Q;         // Q x Q   (Jacobi matrix)
  double* V = G + Q * Q;          // Q x Q   (accumulated rotations)
  double* sig = V + Q * Q;        // Q
  double* ordd = sig + Q;         // Q       (sort permutation, stored as doubles)

  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int64_t e = tid; e < P * Q; e += blockDim.x) {
    const int64_t i = e % P, j = e / P;
    const double v = p.A[i * x_rs + j * x_cs];
    if (!isfinite(v)) s_bad = 1;
    Wq[e] = v;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) p.info[0] = TTX_ENONFINITE;
    return;
  }
  const bool use_qr = P > Q;
  if (use_qr) {
    block_householder_qr(P, Q, Wq, tau, Qx, red);
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
      const int64_t i = e % Q, c = e / Q;
      G[e] = (i <= c) ? Wq[i + c * P] : 0.0;
    }
  } else {
    for (int64_t e = tid; e < Q * Q; e += blockDim.x) G[e] = Wq[e];
  }
  for (int64_t e = tid; e < Q * Q; e += blockDim.x) V[e] = ((e % Q) == (e / Q)) ? 1.0 : 0.0;
  __syncthreads();

  // one-sided Jacobi on the columns of G, round-robin pairs per round
  const int64_t QQ = Q + (Q & 1);
  const double tol = (double)Q * kEps;
  int converged = (Q < 2);
  for (int sweep = 0; sweep < 80 && !converged; ++sweep) {
    if (tid == 0) s_rot = 0;
    __syncthreads();
    for (int64_t round = 0; round < QQ - 1; ++round) {
      for (int64_t pk = wave; pk < QQ / 2; pk += nwave) {
        int64_t i = rr_player(pk, round, QQ);
        int64_t j = rr_player(QQ - 1 - pk, round, QQ);
        if (i >= Q || j >= Q) continue;
        if (i > j) { const int64_t t = i; i = j; j = t; }
        double* gi = G + i * Q;
        double* gj = G + j * Q;
        double a = 0.0, b = 0.0, c = 0.0;
        for (int64_t r = lane; r < Q; r += 64) {
          const double x = gi[r], y = gj[r];
          a += x * x;
          b += y * y;
          c += x * y;
        }
        a = wave_sum(a);
        b = wave_sum(b);
        c = wave_sum(c);
        if (c == 0.0 || a == 0.0 || b == 0.0) continue;
        if (fabs(c) <= tol * sqrt(a) * sqrt(b)) continue;
        const double zeta = (b - a) / (2.0 * c);
        double t;
        if (fabs(zeta) > 1e150)
          t = 0.5 / zeta;
        else
          t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t);
        const double sn = cs * t;
        for (int64_t r = lane; r < Q; r += 64) {
          const double x = gi[r], y = gj[r];
          gi[r] = cs * x - sn * y;
          gj[r] = sn * x + cs * y;
        }
        double* vi = V + i * Q;
        double* vj = V + j * Q;
        for (int64_t r = lane; r < Q; r += 64) {
          const double x = vi[r], y = vj[r];
          vi[r] = cs * x - sn * y;
          vj[r] = sn * x + cs * y;
        }
        if (lane == 0) s_rot = 1;
      }
      __syncthreads();
    }
    converged = (s_rot == 0);
    __syncthreads();
  }
  if (tid == 0) p.info[0] = converged ? TTX_OK : TTX_ENOCONV;
  __syncthreads();

  // singular values = column norms
  for (int64_t j = wave; j < Q; j += nwave) {
    const double* gj = G + j * Q;
    double s = 0.0;
    for (int64_t r = lane; r < Q; r += 64) s += gj[r] * gj[r];
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  __syncthreads();
  // descending order, stable on ties (selection by rank count)
  for (int64_t j = tid; j < Q; j += blockDim.x) {
    const double sj = sig[j];
    int64_t pos = 0;
    for (int64_t l = 0; l < Q; ++l) {
      const double sl = sig[l];
      if (sl > sj || (sl == sj && l < j)) ++pos;
    }
    ordd[pos] = (double)j;
  }
  __syncthreads();
  // normalise the columns of G in place: G[:, j] /= sig[j]
  for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
    const int64_t j = e / Q;
    const double s = sig[j];
    G[e] = s > 0.0 ? G[e] / s : 0.0;
  }
  __syncthreads();
  // Ux (P x Q) column o = (use_qr ? Qx : I) * G[:, ord[o]];  V column ord[o]
  // tall: U_A = Ux, Vt_A = V^T.  wide: U_A = V, Vt_A = Ux^T.
  for (int64_t e = tid; e < P * Q; e += blockDim.x) {
    const int64_t r = e % P, o = e / P;
    const int64_t src = (int64_t)ordd[o];
    double v;
    if (use_qr) {
      v = 0.0;
      const double* gcol = G + src * Q;
      for (int64_t l = 0; l < Q; ++l) v += Qx[r + l * P] * gcol[l];
    } else {
      v = G[r + src * Q];
    }
    if (!wide)
      p.U[r * p.u_rs + o * p.u_cs] = v;
    else
      p.Vt[o * p.v_rs + r * p.v_cs] = v;
  }
  for (int64_t e = tid; e < Q * Q; e += blockDim.x) {
    const int64_t r = e % Q, o = e / Q;
    const int64_t src = (int64_t)ordd[o];
    const double v = V[r + src * Q];
    if (!wide)
      p.Vt[o * p.v_rs + r * p.v_cs] = v;
    else
      p.U[r * p.u_rs + o * p.u_cs] = v;
  }
  for (int64_t o = tid; o < Q; o += blockDim.x) p.S[o] = sig[(int64_t)ordd[o]];
}

// ---------------------------------------------------------------------------
// Batched strided GEMM by one block: C[b] = alpha A[b] B[b] + beta C[b].
// Each wave owns 32x32 output tiles (2x2 MFMA 16x16x4 f64 blocks) and reads
// its operands straight from memory with the given strides.  No barrier
// inside: callers separate dependent calls with __syncthreads().
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bgemm(int64_t batch, int64_t M, int64_t N, int64_t K, double alpha,
                                      const double* A, int64_t a_ms, int64_t a_ks, int64_t a_bs,
                                      const double* B, int64_t b_ks, int64_t b_ns, int64_t b_bs,
                                      double beta, double* C, int64_t c_ms, int64_t c_ns,
                                      int64_t c_bs) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t tm = (M + 31) >> 5, tn = (N + 31) >> 5;
  const int64_t per = tm * tn;
  add_flops(2.0 * (double)batch * (double)M * (double)N * (double)K);
  for (int64_t t = wave; t < batch * per; t += nw) {
    const int64_t bt = t / per, tt = t % per;
    const int64_t m0 = (tt / tn) << 5, n0 = (tt % tn) << 5;
    const double* Ab = A + bt * a_bs;
    const double* Bb = B + bt * b_bs;
    double* Cb = C + bt * c_bs;
    const int64_t ma = m0 + r, mb = m0 + 16 + r, na = n0 + r, nb = n0 + 16 + r;
    const bool va = ma < M, vb = mb < M, wa = na < N, wb = nb < N;
    double4_t c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    for (int64_t k0 = 0; k0 < K; k0 += 4) {
      const int64_t kk = k0 + kq;
      const bool kin = kk < K;
      const double a0 = (kin && va) ? Ab[ma * a_ms + kk * a_ks] : 0.0;
      const double a1 = (kin && vb) ? Ab[mb * a_ms + kk * a_ks] : 0.0;
      const double b0 = (kin && wa) ? Bb[kk * b_ks + na * b_ns] : 0.0;
      const double b1 = (kin && wb) ? Bb[kk * b_ks + nb * b_ns] : 0.0;
      c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c00, 0, 0, 0);
      c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, c01, 0, 0, 0);
      c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, c10, 0, 0, 0);
      c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c11, 0, 0, 0);
    }
    // D slot (lane, e) of a 16x16 block is row 4e + lane/16, column lane%16
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row0 = m0 + 4 * e + kq, row1 = row0 + 16;
      if (row0 < M && wa) {
        double* d = Cb + row0 * c_ms + na * c_ns;
        *d = beta == 0.0 ? alpha * c00[e] : alpha * c00[e] + beta * (*d);
      }
      if (row0 < M && wb) {
        double* d = Cb + row0 * c_ms + nb * c_ns;
        *d = beta == 0.0 ? alpha * c01[e] : alpha * c01[e] + beta * (*d);
      }
      if (row1 < M && wa) {
        double* d = Cb + row1 * c_ms + na * c_ns;
        *d = beta == 0.0 ? alpha * c10[e] : alpha * c10[e] + beta * (*d);
      }
      if (row1 < M && wb) {
        double* d = Cb + row1 * c_ms + nb * c_ns;
        *d = beta == 0.0 ? alpha * c11[e] : alpha * c11[e] + beta * (*d);
      }
    }
  }
}

// y = alpha * M x + beta * y for a row-major M (row stride ld); one wave per
// row so the reads of a row coalesce.  No barrier inside.
__device__ __forceinline__ void bgemv(int64_t rows, int64_t cols, double alpha, const double* M,
                                      int64_t ld, const double* x, double beta, double* y) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  add_flops(2.0 * (double)rows * (double)cols);
  for (int64_t r = wave; r < rows; r += nw) {
    const double* mr = M + r * ld;
    double s = 0.0;
    for (int64_t c = lane; c < cols; c += 64) s += mr[c] * x[c];
    s = wave_sum(s);
    if (lane == 0) y[r] = beta == 0.0 ? alpha * s : alpha * s + beta * y[r];
  }
}

// ---------------------------------------------------------------------------
// LU with partial pivoting and the solve of [M | b] (row-major, row stride
// ld >= n + 1) by one block.  `lds` holds the panel (nb x n doubles) and is
// reused for the 64 x 64 diagonal blocks of the back substitution; `u`
// receives the solution.  Returns TTX_ESINGULAR through *info on a zero
// pivot (the factorization continues, as getrf does).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_lu_solve(int64_t n, int64_t ld, double* W, double* u,
                                               double* lds, int64_t nb0, double* red,
                                               int32_t* info) {
  __shared__ double rv[16];
  __shared__ int64_t ri[16];
  __shared__ int64_t s_piv[16];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwave = blockDim.x >> 6;
  const int64_t ncols = n + 1;
  {
    const double dn = (double)n;
    add_flops(2.0 * dn * dn * dn / 3.0 + 2.0 * dn * dn);
  }
  for (int64_t k = 0; k < n; k += nb0) {
    const int64_t nb = (n - k) < nb0 ? (n - k) : nb0;
    const int64_t mp = n - k;
    double* P = lds;
    for (int64_t e = tid; e < nb * mp; e += blockDim.x) {
      const int64_t c = e % nb, r = e / nb;
      P[c * mp + r] = W[(k + r) * ld + (k + c)];
    }
    __syncthreads();
    for (int64_t j = 0; j < nb; ++j) {
      double* pj = P + j * mp;
      double best = -1.0;
      int64_t bi = j;
      for (int64_t r = j + tid; r < mp; r += blockDim.x) {
        const double v = fabs(pj[r]);
        if (v > best) { best = v; bi = r; }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int64_t oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
      __syncthreads();
      if (tid == 0) {
        double b0 = rv[0];
        int64_t i0 = ri[0];
        for (int w = 1; w < nwave; ++w)
          if (rv[w] > b0 || (rv[w] == b0 && ri[w] < i0)) { b0 = rv[w]; i0 = ri[w]; }
        s_piv[j] = i0;
        if (b0 == 0.0) info[0] = TTX_ESINGULAR;
      }
      __syncthreads();
      const int64_t pr = s_piv[j];
      if (pr != j) {
        for (int64_t c = tid; c < nb; c += blockDim.x) {
          double* pc = P + c * mp;
          const double t = pc[j];
          pc[j] = pc[pr];
          pc[pr] = t;
        }
      }
      __syncthreads();
      const double dj = pj[j];
      if (dj != 0.0) {
        for (int64_t r = j + 1 + tid; r < mp; r += blockDim.x) pj[r] /= dj;
      }
      __syncthreads();
      for (int64_t c = j + 1; c < nb; ++c) {
        double* pc = P + c * mp;
        const double uu = pc[j];
        for (int64_t r = j + 1 + tid; r < mp; r += blockDim.x) pc[r] -= pj[r] * uu;
      }
      __syncthreads();
    }
    for (int64_t e = tid; e < nb * mp; e += blockDim.x) {
      const int64_t c = e % nb, r = e / nb;
      W[(k + r) * ld + (k + c)] = P[c * mp + r];
    }
    for (int64_t c = tid; c < ncols; c += blockDim.x) {
      if (c >= k && c < k + nb) continue;
      for (int64_t j = 0; j < nb; ++j) {
        const int64_t pr = k + s_piv[j];
        if (pr != k + j) {
          const double t = W[(k + j) * ld + c];
          W[(k + j) * ld + c] = W[pr * ld + c];
          W[pr * ld + c] = t;
        }
      }
      if (c >= k + nb) {
        for (int64_t i = 1; i < nb; ++i) {
          double s = W[(k + i) * ld + c];
          for (int64_t l = 0; l < i; ++l) s -= P[l * mp + i] * W[(k + l) * ld + c];
          W[(k + i) * ld + c] = s;
        }
      }
    }
    __syncthreads();
    const int64_t rows = n - k - nb, cols = ncols - k - nb;
    if (rows > 0 && cols > 0)
      bgemm(1, rows, cols, nb, -1.0, W + (k + nb) * ld + k, ld, 1, 0, W + k * ld + (k + nb), ld, 1, 0,
            1.0, W + (k + nb) * ld + (k + nb), ld, 1, 0);
    __syncthreads();
  }
  // back substitution U u = y, y = column n
  double* T = lds;          // 64 x 65
  double* xb = lds + 64 * 65;
  for (int64_t i = tid; i < n; i += blockDim.x) u[i] = W[i * ld + n];
  __syncthreads();
  for (int64_t j1 = n; j1 > 0; j1 -= 64) {
    const int64_t j0 = j1 - 64 > 0 ? j1 - 64 : 0;
    const int64_t nbk = j1 - j0;
    for (int64_t e = tid; e < nbk * nbk; e += blockDim.x) {
      const int64_t r = e / nbk, c = e % nbk;
      T[r * 65 + c] = W[(j0 + r) * ld + j0 + c];
    }
    __syncthreads();
    if (wave == 0) {
      const int64_t row = j0 + lane;
      double yv = row < j1 ? u[row] : 0.0;
      for (int64_t j = j1 - 1; j >= j0; --j) {
        const int jj = (int)(j - j0);
        const double xj = __shfl(yv, jj, 64) / T[jj * 65 + jj];
        if (row == j) yv = xj;
        if (row < j) yv -= xj * T[lane * 65 + jj];
      }
      if (row < j1) { u[row] = yv; xb[lane] = yv; }
    }
    __syncthreads();
    for (int64_t r = tid; r < j0; r += blockDim.x) {
      double s = u[r];
      const double* wr = W + r * ld + j0;
      for (int64_t c = 0; c < nbk; ++c) s -= wr[c] * xb[c];
      u[r] = s;
    }
    __syncthreads();
  }
  (void)red;
}

}  // namespace ttx
