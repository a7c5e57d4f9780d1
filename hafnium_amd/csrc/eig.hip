// Batched symmetric eigensolver for the lowest eigenpairs, one workgroup
// per matrix (the Fock diagonalisation of a batched SCF, where rocSOLVER's
// batched divide and conquer is about 1500 small launches per call):
//   1. Householder reduction to tridiagonal form, T = Q^T A Q (Golub and
//      Van Loan, "Matrix Computations", algorithm 8.3.1), the reflector
//      vectors kept in the rows of A above the diagonal;
//   2. the nvec lowest eigenvalues of T by multisection (bisection with
//      several points per round) on the Sturm count of the LDL^T recurrence
//      (ibid. section 8.4.1; the pivot guard of LAPACK's dstebz);
//   3. their eigenvectors by inverse iteration on T with a partially
//      pivoted tridiagonal LU (as LAPACK's dlagtf / dlagts), equal
//      eigenvalues' shifts separated by 10 eps |T|, then made orthonormal
//      (Cholesky QR twice, or Gram-Schmidt twice past 64 vectors);
//   4. back-transformation x = Q y by the stored reflectors.
// A is row-major and symmetric; on return its rows 0 .. nvec - 1 are the
// eigenvectors of the nvec lowest eigenvalues and w[0 .. nvec) holds those
// eigenvalues in ascending order (rows nvec .. n - 1 of A hold scratch).
#include <cfloat>
#include <cmath>
#include <string>

#include "hfjk_internal.h"

namespace {

typedef unsigned int u2 __attribute__((ext_vector_type(2)));

constexpr int kEigThreads = 512;
constexpr int kEigWaves = kEigThreads / 64;
constexpr int kEigMaxN = 512;
constexpr int kEigMaxCols = kEigMaxN / 64;  // 64-row chunks of a vector
constexpr int kEigRegVecs = 64;             // nvec up to this: vectors stay in registers, Cholesky QR
constexpr int kGs = 65;                     // row stride of the Gram matrix and its inverse factor in LDS

// Workspace doubles per matrix: d, e, beta (n each), the nvec eigenvectors [nvec][n], their LU
// factors (five [n][nvec] arrays) and the inverse-iteration vectors [n][nvec].
__host__ __device__ inline long long eig_work_per(int n, int nvec) { return 3LL * n + 7LL * nvec * n; }

// Sum over the workgroup (every thread passes its part; all get the total).
__device__ __forceinline__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();  // red is free (its previous readers are done)
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEigWaves; ++k) s += red[k];
  return s;
}

// A lane's partner value under a DPP row permutation (ctrl: 0x140 row mirror, 0x141 half-row
// mirror, quad permutations below 0x100); every lane of the row takes part.
template <int CTRL>
__device__ inline double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Eight sums over the wave at once: on return every lane holds the total of s[lane >> 3].
// Each halving exchange sends the half of the values the partner keeps (10 exchanges, not 48).
// The partners differ in lane bit 5, 4, 3, 2, 1, 0 in turn: lane ^ 32 and lane ^ 16 by
// shuffles, then within a row of 16 by DPP (row mirror 15 - l, half-row mirror 7 - l, quad
// permutations l ^ 2 and l ^ 1), which pair lanes the same way for these sums.
__device__ inline double wave_sums8(const double (&s)[8], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
  double t4[4], t2[2];
#pragma unroll
  for (int x = 0; x < 4; ++x) t4[x] = (h5 ? s[4 + x] : s[x]) + __shfl_xor(h5 ? s[x] : s[4 + x], 32);
#pragma unroll
  for (int x = 0; x < 2; ++x) t2[x] = (h4 ? t4[2 + x] : t4[x]) + __shfl_xor(h4 ? t4[x] : t4[2 + x], 16);
  double t1 = (h3 ? t2[1] : t2[0]) + dpp_d<0x140>(h3 ? t2[0] : t2[1]);
  t1 += dpp_d<0x141>(t1);
  t1 += dpp_d<0x4E>(t1);  // quad permutation [2, 3, 0, 1]
  t1 += dpp_d<0xB1>(t1);  // quad permutation [1, 0, 3, 2]
  return t1;
}

// Sum over the wave, every lane gets it (shuffles across rows, DPP within a row).
__device__ inline double wave_sum(double v) {
  v += __shfl_xor(v, 32);
  v += __shfl_xor(v, 16);
  v += dpp_d<0x140>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0xB1>(v);
  return v;
}

// Sum over the workgroup with one barrier: red must not be written again before a later barrier
// (callers alternate two buffers between barriers).
__device__ __forceinline__ double wg_sum1(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kEigWaves; ++k) s += red[k];
  return s;
}

__device__ inline double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1 / q from the hardware estimate and one Newton step (the Sturm count needs only the signs
// of a recurrence that is stable under small relative perturbations).
__device__ inline double recip(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return fma(r, fma(-q, r, 1.0), r);
}

// A deterministic start vector entry in (-1, 1).
__device__ double start_entry(int j, int i) {
  unsigned h = (unsigned)(j * 73856093) ^ (unsigned)(i * 19349663) ^ 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return (double)(h & 0xFFFFFFu) / (double)0x800000 - 1.0;
}

constexpr int kOutOfRange = 1 << 30;  // a byte offset beyond any matrix of at most 512 x 512 doubles

__device__ inline double buf_ld(__amdgpu_buffer_rsrc_t rsrc, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0));
}
__device__ inline void buf_st(__amdgpu_buffer_rsrc_t rsrc, double v, int off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rsrc, off, 0, 0);
}

// R rows j0, j0 + 8, ... of one pass, CB chunks each, without branches (so that the compiler
// keeps the per-row sums in place): rows past m load the last row and store out of range,
// lanes off the lower triangle enter the sums with a zero factor.
template <int NC, int CB, int R>
__device__ __forceinline__ void tri_batch(__amdgpu_buffer_rsrc_t rsrc, int boff, int n, int m, int j0, int lane,
                                          const double (&pvc)[NC], const double (&pwc)[NC], const double (&vc)[NC],
                                          double (&acc)[NC], const double* pv, const double* pw, const double* vv,
                                          double v0, double* nxt, double* prow, double* lsink, int lane_off,
                                          int last_off) {
  double val[R][CB];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int roff = boff + min(j0 + r * kEigWaves, m - 1) * n * 8;
#pragma unroll
    for (int c = 0; c < CB; ++c) val[r][c] = buf_ld(rsrc, roff + (c == NC - 1 ? last_off : lane_off) + 512 * c);
  }
  double rs[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) rs[r] = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r * kEigWaves;
    const bool live = j < m;
    const int jc = min(j, m - 1);
    const int jm = live ? j : -1;  // lanes i <= jm are in the lower triangle of a live row
    const double vj = live ? (jc == 0 ? v0 : vv[jc]) : 0.0;
    const double pvj = pv[1 + jc], pwj = pw[1 + jc];
    const int roff = live ? boff + j * n * 8 : kOutOfRange;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int i = lane + 64 * c;
      const double a = val[r][c] - pvj * pwc[c] - pwj * pvc[c];
      buf_st(rsrc, a, roff + (c == NC - 1 ? last_off : lane_off) + 512 * c);
      if (c + 2 < CB) {  // below the diagonal in every row of the batch (a batch spans < 64 rows)
        acc[c] = fma(a, vj, acc[c]);
        rs[r] = fma(a, vc[c], rs[r]);  // the sums of rows past m are not used
      } else {
        acc[c] = fma(a, i <= jm ? vj : 0.0, acc[c]);
        rs[r] = fma(a, i < jm ? vc[c] : 0.0, rs[r]);
      }
      if (c == 0) *(live && lane == 0 ? nxt + j : lsink + lane) = a;
    }
  }
  const double tot = wave_sums8(rs, lane);
  if ((lane & 7) == 0) {
    const int r = lane >> 3, j = j0 + r * kEigWaves;
    if (r < R && j < m) prow[j] = tot;
  }
}

template <int NC, int CB, int R>
__device__ __forceinline__ void tri_batch_cb(int cb, __amdgpu_buffer_rsrc_t rsrc, int boff, int n, int m, int j0,
                                             int lane, const double (&pvc)[NC], const double (&pwc)[NC],
                                             const double (&vc)[NC], double (&acc)[NC], const double* pv,
                                             const double* pw, const double* vv, double v0, double* nxt, double* prow,
                                             double* lsink, int lane_off, int last_off) {
  if constexpr (CB <= NC) {
    if (cb == CB)
      tri_batch<NC, CB, R>(rsrc, boff, n, m, j0, lane, pvc, pwc, vc, acc, pv, pw, vv, v0, nxt, prow, lsink, lane_off,
                           last_off);
    else
      tri_batch_cb<NC, CB + 1, R>(cb, rsrc, boff, n, m, j0, lane, pvc, pwc, vc, acc, pv, pw, vv, v0, nxt, prow, lsink,
                                  lane_off, last_off);
  }
}

// One pass over the lower triangle of the m x m trailing block (rows and columns c0 .. n - 1 of
// A, at byte offset boff of the buffer rsrc, which covers A): applies the previous step's rank-2
// update A -= v' w'^T + w' v'^T (pv, pw indexed from row c0 - 1) and forms the product of the
// updated block with v (v[0] is v0): the column part sum_{j >= i} a_ji v_j into partw[i], the
// row part sum_{i < j} a_ji v_i into prow[j]; column c0 of the result also goes to nxt. Wave
// `wave` takes rows wave, wave + 8, ..., R rows per batch, each batch as many 64-column chunks as
// its last row reaches. Buffer loads and stores keep the VALU work per chunk to the arithmetic:
// the chunk's offset is an immediate, and lanes of the last chunk past the row end get an
// offset outside the buffer (loads read zero, stores are dropped). Lanes above the diagonal load
// and store the upper triangle of the block, which no pass reads.
template <int NC>
__device__ __forceinline__ void tri_pass(__amdgpu_buffer_rsrc_t rsrc, int boff, int n, int m, int wave, int lane,
                                         const double* pv, const double* pw, const double* vv, double v0,
                                         double* nxt, double* partw, double* prow, double* lsink) {
  constexpr int R = NC <= 4 ? 8 : 4;
  double pvc[NC], pwc[NC], vc[NC], acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = lane + 64 * c;
    pvc[c] = i < m ? pv[1 + i] : 0.0;
    pwc[c] = i < m ? pw[1 + i] : 0.0;
    vc[c] = i == 0 ? v0 : i < m ? vv[i] : 0.0;
    acc[c] = 0.0;
  }
  const int lane_off = lane * 8;
  const int last_off = lane + 64 * (NC - 1) < m ? lane * 8 : kOutOfRange;  // last chunk, less 512 (NC - 1)
  for (int j0 = wave; j0 < m; j0 += kEigWaves * R) {
    const int cb = (min(j0 + kEigWaves * (R - 1), m - 1) >> 6) + 1;
    tri_batch_cb<NC, 1, R>(cb, rsrc, boff, n, m, j0, lane, pvc, pwc, vc, acc, pv, pw, vv, v0, nxt, prow, lsink,
                           lane_off, last_off);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = lane + 64 * c;
    if (i < m) partw[i] = acc[c];
  }
}

// y[q] = Y[j0 + q] for the wave's eight vectors (zero past nvec or n).
template <int NC>
__device__ inline void load_vecs(double (&y)[8][NC], const double* Y, int n, int nvec, int j0, int lane) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = min(j0 + q, nvec - 1);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int r = lane + 64 * c;
      const double v = Y[(long long)j * n + min(r, n - 1)];
      y[q][c] = (j0 + q < nvec && r < n) ? v : 0.0;
    }
  }
}

template <int NC>
__device__ inline void store_vecs(const double (&y)[8][NC], double* Y, int n, int nvec, int j0, int lane) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (j0 + q < nvec) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int r = lane + 64 * c;
        if (r < n) Y[(long long)(j0 + q) * n + r] = y[q][c];
      }
    }
  }
}

// y <- H_0 H_1 ... H_{n-3} y for the wave's eight vectors in registers; reflector k is
// I - beta_k u u^T with u in row k of A (columns k + 1 .. n - 1). The next reflector's loads are
// issued before the current one is applied.
template <int NC>
__device__ __forceinline__ void back_transform(double (&y)[8][NC], const double* A, const double* lbeta, int n, int lane) {
  if (n < 3) return;
  double un[NC];
  auto load_u = [&](int k) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int r = lane + 64 * c;
      const double v = A[(long long)k * n + min(max(r, k + 1), n - 1)];
      un[c] = (r > k && r < n) ? v : 0.0;
    }
  };
  load_u(n - 3);
  for (int k = n - 3; k >= 0; --k) {
    double u[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) u[c] = un[c];
    if (k > 0) load_u(k - 1);
    const double bk = lbeta[k];
    if (bk == 0.0) continue;
    double s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double a = 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) a += u[c] * y[q][c];
      s[q] = a;
    }
    const double tot = wave_sums8(s, lane);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double f = bk * readlane_d(tot, 8 * q);
#pragma unroll
      for (int c = 0; c < NC; ++c) y[q][c] -= f * u[c];
    }
  }
}

// Rows j0 .. j0 + 7 of the Gram matrix G = Y Y^T against columns 0 .. min(j0 + 7, nvec - 1)
// into gm (lower triangle), the wave's own vectors in y, the others read from Y.
template <int NC>
__device__ __forceinline__ void gram_rows(const double (&y)[8][NC], const double* Y, int n, int nvec, int j0, int lane,
                          double* gm) {
  constexpr int OB = NC <= 2 ? 8 : NC <= 4 ? 4 : 2;  // other vectors per batch of loads
  const int kend = min(j0 + 8, nvec);
  for (int kb = 0; kb < kend; kb += OB) {
    double ob[OB][NC];
#pragma unroll
    for (int o = 0; o < OB; ++o) {
      const int kk = min(kb + o, nvec - 1);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int r = lane + 64 * c;
        const double v = Y[(long long)kk * n + min(r, n - 1)];
        ob[o][c] = r < n ? v : 0.0;
      }
    }
#pragma unroll
    for (int g = 0; g < 8 * OB; g += 8) {
      double s[8];
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        const int q = (g + x) / OB, o = (g + x) % OB;
        double a = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) a += y[q][c] * ob[o][c];
        s[x] = a;
      }
      const double tot = wave_sums8(s, lane);
      if ((lane & 7) == 0) {
        const int x = lane >> 3;
        const int q = (g + x) / OB, o = (g + x) % OB;
        const int j = j0 + q, kk = kb + o;
        if (j < nvec && kk <= j) gm[j * kGs + kk] = tot;
      }
    }
  }
}

// y <- rows j0 .. j0 + 7 of Linv Y (Linv lower triangular in li), Y read from memory.
template <int NC>
__device__ __forceinline__ void tri_product(double (&y)[8][NC], const double* Y, int n, int nvec, int j0, int lane,
                            const double* li) {
  constexpr int OB = NC <= 2 ? 8 : NC <= 4 ? 4 : 2;
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int c = 0; c < NC; ++c) y[q][c] = 0.0;
  const int kend = min(j0 + 8, nvec);
  for (int kb = 0; kb < kend; kb += OB) {
    double ob[OB][NC];
#pragma unroll
    for (int o = 0; o < OB; ++o) {
      const int kk = min(kb + o, nvec - 1);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int r = lane + 64 * c;
        const double v = Y[(long long)kk * n + min(r, n - 1)];
        ob[o][c] = r < n ? v : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = j0 + q;
#pragma unroll
      for (int o = 0; o < OB; ++o) {
        const int kk = kb + o;
        const double coef = (j < nvec && kk <= j) ? li[j * kGs + kk] : 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) y[q][c] += coef * ob[o][c];
      }
    }
  }
}

// G = L L^T (G in the lower triangle of gm, nv <= 64) and li = L^{-1}, by wave 0 with every
// row in registers: lane i holds row i of G, each column of L is broadcast through LDS (the
// first row of li), and the update of the rows is unrolled over the columns; then L goes to gm
// and lane c forms column c of L^{-1} by forward substitution, reading L by broadcast. A
// non-positive pivot is raised to DBL_MIN. Not inlined: one copy serves every vector length.
__device__ __attribute__((noinline)) void cholesky_inverse(double* gm, double* li, int nv, int lane) {
  const int i = lane;
  double g[kEigRegVecs];
#pragma unroll
  for (int j = 0; j < kEigRegVecs; ++j) g[j] = (i < nv && j <= i) ? gm[i * kGs + j] : 0.0;
#pragma unroll
  for (int k = 0; k < kEigRegVecs; ++k) {
    if (k < nv) {
      const double lkk = sqrt(fmax(readlane_d(g[k], k), DBL_MIN));
      const double lik = i == k ? lkk : g[k] / lkk;  // L[i][k] for i >= k (lanes above: unused)
      g[k] = lik;
      li[i] = lik;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int j = k + 1; j < kEigRegVecs; ++j) g[j] = fma(-lik, li[j], g[j]);
      __builtin_amdgcn_wave_barrier();
    }
  }
#pragma unroll
  for (int j = 0; j < kEigRegVecs; ++j) gm[i * kGs + j] = g[j];
  __builtin_amdgcn_wave_barrier();
  const int c = lane;
  double x[kEigRegVecs];  // column c of L^{-1}; x[r] = 0 for r < c falls out of the recurrence
#pragma unroll
  for (int r = 0; r < kEigRegVecs; ++r) {
    if (r < nv) {
      double sacc = r == c ? 1.0 : 0.0;
#pragma unroll
      for (int p = 0; p < r; ++p) sacc = fma(-gm[r * kGs + p], x[p], sacc);
      x[r] = sacc / gm[r * kGs + r];
    } else {
      x[r] = 0.0;
    }
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int r = 0; r < kEigRegVecs; ++r)
    if (r < nv && c < nv) li[r * kGs + c] = x[r];
}

// Orthonormalisation (Cholesky QR twice) and back-transformation for nvec <= 64, each wave's
// eight vectors held in registers; the results go to rows 0 .. nvec - 1 of A.
template <int NC>
__device__ __forceinline__ void finish_in_registers(double* A, double* Y, const double* lbeta, double* gm, double* li, int n,
                                    int nvec, int wave, int lane) {
  const int j0 = 8 * wave;
  double y[8][NC];
  load_vecs<NC>(y, Y, n, nvec, j0, lane);
  for (int pass = 0; pass < 2; ++pass) {
    if (j0 < nvec) gram_rows<NC>(y, Y, n, nvec, j0, lane, gm);
    __syncthreads();
    if (wave == 0) cholesky_inverse(gm, li, nvec, lane);
    __syncthreads();
    if (j0 < nvec) tri_product<NC>(y, Y, n, nvec, j0, lane, li);
    __syncthreads();  // every wave has read Y
    if (pass == 0) {
      store_vecs<NC>(y, Y, n, nvec, j0, lane);
      __syncthreads();
    }
  }
  if (j0 < nvec) back_transform<NC>(y, A, lbeta, n, lane);
  __syncthreads();  // every wave is done with the reflectors in A
  store_vecs<NC>(y, A, n, nvec, j0, lane);
}

// Back-transformation of Y in chunks of 64 vectors (nvec > 64), results in Y.
template <int NC>
__device__ __forceinline__ void back_transform_chunks(const double* A, double* Y, const double* lbeta, int n, int nvec, int wave,
                                      int lane) {
  for (int cb = 0; cb < nvec; cb += 64) {
    const int j0 = cb + 8 * wave;
    if (j0 >= nvec) continue;
    double y[8][NC];
    load_vecs<NC>(y, Y, n, nvec, j0, lane);
    back_transform<NC>(y, A, lbeta, n, lane);
    store_vecs<NC>(y, Y, n, nvec, j0, lane);
  }
}

__global__ __launch_bounds__(kEigThreads) void eig_lowest_kernel(int n, int nvec, double* __restrict__ a_all,
                                                                  double* __restrict__ w_all,
                                                                  double* __restrict__ work_all, int* __restrict__ info) {
  __shared__ double part[kEigWaves][kEigMaxN];
  __shared__ double lv[2][kEigMaxN], lw[2][kEigMaxN], nxt[kEigMaxN], prow[kEigMaxN], lbeta[kEigMaxN];
  __shared__ double gm[kEigRegVecs * kGs], li[kEigRegVecs * kGs];
  __shared__ double red[kEigWaves], red2[2][kEigWaves];
  __shared__ double bounds[4];
  __shared__ double lsink[64];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  if (nvec == 0) {
    if (t == 0) info[blockIdx.x] = 0;
    return;
  }
  double* A = a_all + (long long)blockIdx.x * n * n;
  double* w = w_all + (long long)blockIdx.x * n;
  double* wk = work_all + (long long)blockIdx.x * eig_work_per(n, nvec);
  double* d = wk;
  double* e = wk + n;
  double* beta = wk + 2 * n;
  double* Y = wk + 3 * n;                // [nvec][n]
  double* F = Y + (long long)nvec * n;   // [5][n][nvec]: 1 / pivot, U's two superdiagonals, multiplier, swap flag
  double* Yt = F + 5LL * nvec * n;       // [n][nvec]

  // 1. Householder tridiagonalisation on the lower triangle. Step k reflects column k below the
  // diagonal; its rank-2 update of the trailing block is applied in the next step's single pass
  // over the block, which also forms that step's product with the trailing block. nxt holds
  // column k of the matrix as the previous pass left it; lv / lw hold v and w of the current and
  // the previous step. The reflector vectors go to the upper triangle.
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(A, 0, n * n * 8, 0x00020000);
  for (int i = t; i < n; i += kEigThreads) {
    nxt[i] = A[(long long)i * n];
    lv[1][i] = 0.0;
    lw[1][i] = 0.0;
  }
  __syncthreads();
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    const int c0 = k + 1;
    const int cur = k & 1;
    double* vv = lv[cur];
    double* ww = lw[cur];
    const double* pv = lv[cur ^ 1];
    const double* pw = lw[cur ^ 1];
    const double pv0 = pv[0], pw0 = pw[0];
    double sq = 0.0;
    if (t < m) {
      const double x = nxt[1 + t] - pv[1 + t] * pw0 - pw[1 + t] * pv0;
      vv[t] = x;
      sq = x * x;
    }
    const double x0 = nxt[1] - pv[1] * pw0 - pw[1] * pv0;
    const double dk = nxt[0] - 2.0 * pv0 * pw0;
    const double ss = wg_sum1(sq, red2[0]);
    const double sigma = ss - x0 * x0;
    double bk = 0.0, ek = x0, v0 = x0;
    if (sigma > 0.0) {
      const double norm = sqrt(ss);
      const double alpha = x0 > 0.0 ? -norm : norm;
      v0 = x0 - alpha;
      bk = 2.0 / (sigma + v0 * v0);
      ek = alpha;
    }
    if (t == 0) {
      d[k] = dk;
      e[k] = ek;
      vv[0] = v0;
    }
    if (t == k) lbeta[k] = bk;
    const int boff = (c0 * n + c0) * 8;
    switch ((m + 63) >> 6) {
      case 1: tri_pass<1>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 2: tri_pass<2>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 3: tri_pass<3>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 4: tri_pass<4>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 5: tri_pass<5>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 6: tri_pass<6>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      case 7: tri_pass<7>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
      default: tri_pass<8>(rsrc, boff, n, m, wave, lane, pv, pw, vv, v0, nxt, part[wave], prow, lsink); break;
    }
    __syncthreads();
    double p = 0.0, pdv = 0.0, vi = 0.0;
    if (t < m) {
      double sacc = 0.0;
#pragma unroll
      for (int g = 0; g < kEigWaves; ++g) sacc += part[g][t];
      p = bk * (sacc + prow[t]);
      vi = t == 0 ? v0 : vv[t];
      pdv = p * vi;
      A[(long long)k * n + c0 + t] = vi;  // the reflector vector, row k above the diagonal
    }
    const double K = 0.5 * bk * wg_sum1(pdv, red2[1]);
    if (t < m) ww[t] = p - K * vi;
    __syncthreads();
  }
  if (t == 0) {
    if (n >= 3) {
      const int cur = (n - 3) & 1;
      const double v0 = lv[cur][0], v1 = lv[cur][1], w0 = lw[cur][0], w1 = lw[cur][1];
      d[n - 2] = nxt[0] - 2.0 * v0 * w0;
      e[n - 2] = nxt[1] - v1 * w0 - w1 * v0;
      d[n - 1] = A[(long long)(n - 1) * n + n - 1] - 2.0 * v1 * w1;
    } else if (n == 2) {
      d[0] = A[0];
      e[0] = A[2];
      d[1] = A[3];
    } else {
      d[0] = A[0];
    }
    if (n >= 2) lbeta[n - 2] = 0.0;
    e[n - 1] = 0.0;
    lbeta[n - 1] = 0.0;
    // Gershgorin interval and the pivot guard.
    double lo = d[0], hi = d[0], emax = 0.0;
    for (int i = 0; i < n; ++i) {
      const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
      lo = fmin(lo, d[i] - r);
      hi = fmax(hi, d[i] + r);
      if (i + 1 < n) emax = fmax(emax, e[i] * e[i]);
    }
    const double tnorm = fmax(fabs(lo), fabs(hi));
    bounds[0] = lo - 2.0 * DBL_EPSILON * tnorm - DBL_MIN;
    bounds[1] = hi + 2.0 * DBL_EPSILON * tnorm + DBL_MIN;
    bounds[2] = fmax(tnorm, DBL_MIN);
    bounds[3] = DBL_MIN * fmax(1.0, emax);  // pivmin
  }
  __syncthreads();
  const double lo = bounds[0], hi = bounds[1], tnorm = bounds[2], pivmin = bounds[3];
  for (int i = t; i < n; i += kEigThreads) beta[i] = lbeta[i];

  // 2. The nvec lowest eigenvalues by multisection: G threads per eigenvalue (consecutive lanes
  // of one wave) each count at one of G points of the interval, which shrinks G + 1 times per
  // round; d, e and e^2 from LDS.
  double* ld = part[0];
  double* le2 = part[1];
  double* le = part[2];
  double* lamv = part[3];
  for (int i = t; i < n; i += kEigThreads) {
    ld[i] = d[i];
    le[i] = e[i];
    le2[i] = i + 1 < n ? e[i] * e[i] : 0.0;
  }
  __syncthreads();
  {
    const int G = nvec <= 64 ? 8 : nvec <= 128 ? 4 : nvec <= 256 ? 2 : 1;
    const int groups = kEigThreads / G;
    const int s_in = lane % G;      // this thread's point in its group
    const int gbase = lane - s_in;  // first lane of the group
    const double tol = 0.5 * DBL_EPSILON * tnorm;
    for (int jb = 0; jb < nvec; jb += groups) {
      const int j = jb + t / G;
      double a = lo, b = hi;
      for (int it = 0; it < 200; ++it) {
        const bool done = b - a <= 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b)) + tol;
        const double x = a + (b - a) * (double)(s_in + 1) / (double)(G + 1);
        int cnt = 0;
        if (j < nvec && !done) {
          double q = ld[0] - x;
          if (fabs(q) < pivmin) q = -pivmin;
          cnt = q < 0.0;
          for (int i = 1; i < n; ++i) {
            q = (ld[i] - x) - le2[i - 1] * recip(q);
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
          }
        }
        // Every lane of the wave takes part in the ballots (a group decides as one).
        const unsigned long long above = __ballot(cnt > j);
        const unsigned long long mine = (above >> gbase) & ((1ull << G) - 1);
        if (__ballot(j < nvec && !done) == 0) break;  // the whole wave is done
        if (j < nvec && !done) {
          // The first point with count > j bounds the eigenvalue from above.
          const int sf = mine ? __builtin_ctzll(mine) : G;
          const double na = sf == 0 ? a : a + (b - a) * (double)sf / (double)(G + 1);
          const double nb = sf == G ? b : a + (b - a) * (double)(sf + 1) / (double)(G + 1);
          a = na;
          b = nb;
        }
      }
      if (j < nvec && s_in == 0) {
        w[j] = 0.5 * (a + b);
        lamv[j] = 0.5 * (a + b);
      }
    }
  }
  __syncthreads();

  // 3. Eigenvectors of T, one thread each: the shifts of equal eigenvalues are separated by
  // 10 eps |T| (so that their iterations differ), two inverse iterations from a seeded start
  // with a partially pivoted LU of T - lambda I. Factors and vectors in [n][nvec] order
  // (coalesced over the threads), loads in batches of eight.
  if (t == 0) {
    double prev = 0.0;
    for (int j = 0; j < nvec; ++j) {
      double lam = lamv[j];
      if (j > 0 && lam - prev < 10.0 * DBL_EPSILON * tnorm) lam = prev + 10.0 * DBL_EPSILON * tnorm;
      lamv[j] = lam;
      prev = lam;
    }
  }
  __syncthreads();
  if (t < nvec) {
    const int j = t;
    const double lam = lamv[j];
    const double tiny = DBL_EPSILON * tnorm;
    auto fi = [&](int f, int k) { return ((long long)f * n + k) * nvec + j; };
    double p0 = ld[0] - lam, p1 = n > 1 ? le[0] : 0.0;
    for (int k = 0; k + 1 < n; ++k) {
      const double c = le[k];                                                // below the diagonal
      const double a1 = ld[k + 1] - lam, b1 = k + 2 < n ? le[k + 1] : 0.0;  // the next row
      if (fabs(c) > fabs(p0)) {
        const double mult = p0 / c;
        F[fi(0, k)] = 1.0 / c;
        F[fi(1, k)] = a1;
        F[fi(2, k)] = b1;
        F[fi(3, k)] = mult;
        F[fi(4, k)] = 1.0;
        p0 = p1 - mult * a1;
        p1 = -mult * b1;
      } else {
        if (fabs(p0) < tiny) p0 = p0 < 0.0 ? -tiny : tiny;
        const double mult = c / p0;
        F[fi(0, k)] = 1.0 / p0;
        F[fi(1, k)] = p1;
        F[fi(2, k)] = 0.0;
        F[fi(3, k)] = mult;
        F[fi(4, k)] = 0.0;
        p0 = a1 - mult * p1;
        p1 = b1;
      }
    }
    if (fabs(p0) < tiny) p0 = p0 < 0.0 ? -tiny : tiny;
    F[fi(0, n - 1)] = 1.0 / p0;
    F[fi(1, n - 1)] = 0.0;
    F[fi(2, n - 1)] = 0.0;
    double scale = 1.0;
#pragma unroll
    for (int iter = 0; iter < 2; ++iter) {
      // Forward elimination of the right-hand side (the seeded start, or the previous iterate
      // times scale), then back substitution; scale makes the result unit length.
      double carry = iter == 0 ? start_entry(j, 0) : Yt[j] * scale;
      for (int k0 = 0; k0 + 1 < n; k0 += 8) {
        double lmv[8], pvv[8], yn[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int k = min(k0 + s, n - 2);
          lmv[s] = F[fi(3, k)];
          pvv[s] = F[fi(4, k)];
          yn[s] = iter == 0 ? start_entry(j, k + 1) : Yt[(long long)(k + 1) * nvec + j] * scale;
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int k = k0 + s;
          if (k + 1 < n) {
            double yk = carry, yk1 = yn[s];
            if (pvv[s] != 0.0) {
              yk = yn[s];
              yk1 = carry;
            }
            Yt[(long long)k * nvec + j] = yk;
            carry = yk1 - lmv[s] * yk;
          }
        }
      }
      Yt[(long long)(n - 1) * nvec + j] = carry;
      double nrm = 0.0, x1 = 0.0, x2 = 0.0;  // x_{i+1}, x_{i+2}
      for (int i0 = n - 1; i0 >= 0; i0 -= 8) {
        double yv[8], a1[8], a2[8], rr[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int i = max(i0 - s, 0);
          yv[s] = Yt[(long long)i * nvec + j];
          rr[s] = F[fi(0, i)];
          a1[s] = F[fi(1, i)];
          a2[s] = F[fi(2, i)];
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int i = i0 - s;
          if (i >= 0) {
            const double xi = (yv[s] - a1[s] * x1 - a2[s] * x2) * rr[s];
            Yt[(long long)i * nvec + j] = xi;
            nrm += xi * xi;
            x2 = x1;
            x1 = xi;
          }
        }
      }
      scale = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    }
    for (int i0 = 0; i0 < n; i0 += 8) {
      double yv[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) yv[s] = Yt[(long long)min(i0 + s, n - 1) * nvec + j];
#pragma unroll
      for (int s = 0; s < 8; ++s)
        if (i0 + s < n) Y[(long long)j * n + i0 + s] = yv[s] * scale;
    }
  }
  __syncthreads();

  // 4. Orthonormalisation and back-transformation x = Q y = H_0 (H_1 (... H_{n-3} y)).
  const int nc = (n + 63) >> 6;
  if (nvec <= kEigRegVecs) {
    switch (nc) {
      case 1: finish_in_registers<1>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 2: finish_in_registers<2>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 3: finish_in_registers<3>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 4: finish_in_registers<4>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 5: finish_in_registers<5>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 6: finish_in_registers<6>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      case 7: finish_in_registers<7>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
      default: finish_in_registers<8>(A, Y, lbeta, gm, li, n, nvec, wave, lane); break;
    }
  } else {
    // Gram-Schmidt in order: vector j minus its components along vectors 0 .. j - 1 (dots by
    // waves, one earlier vector each), twice, then unit length.
    for (int j = 1; j < nvec; ++j) {
      double* y = Y + (long long)j * n;
      for (int pass = 0; pass < 2; ++pass) {
        for (int k = wave; k < j; k += kEigWaves) {
          const double* z = Y + (long long)k * n;
          double sdot = 0.0;
          for (int i = lane; i < n; i += 64) sdot += z[i] * y[i];
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1) sdot += __shfl_xor(sdot, o);
          if (lane == 0) part[0][k] = sdot;
        }
        __syncthreads();
        for (int i = t; i < n; i += kEigThreads) {
          double v = y[i];
          for (int k = 0; k < j; ++k) v -= part[0][k] * Y[(long long)k * n + i];
          y[i] = v;
        }
        __syncthreads();
      }
      double sq = 0.0;
      for (int i = t; i < n; i += kEigThreads) sq += y[i] * y[i];
      const double nrm = wg_sum(sq, red);
      const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
      for (int i = t; i < n; i += kEigThreads) y[i] *= inv;
      __syncthreads();
    }
    switch (nc) {
      case 1: back_transform_chunks<1>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 2: back_transform_chunks<2>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 3: back_transform_chunks<3>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 4: back_transform_chunks<4>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 5: back_transform_chunks<5>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 6: back_transform_chunks<6>(A, Y, lbeta, n, nvec, wave, lane); break;
      case 7: back_transform_chunks<7>(A, Y, lbeta, n, nvec, wave, lane); break;
      default: back_transform_chunks<8>(A, Y, lbeta, n, nvec, wave, lane); break;
    }
    __syncthreads();
    for (long long f = t; f < (long long)nvec * n; f += kEigThreads) A[f] = Y[f];
  }
  if (t == 0) info[blockIdx.x] = 0;
}

}  // namespace

extern "C" {

long long hfjk_eigh_lowest_batched_work(int n, int count, int nvec) {
  if (n < 1 || count < 0 || nvec < 0) return 0;
  return (long long)count * eig_work_per(n, nvec);
}

int hfjk_eigh_lowest_batched(hfjk_ctx* ctx, int n, int count, int nvec, double* a, double* w, double* work, int* info,
                             void* stream) {
  if (ctx == nullptr) {
    hfjk_internal_set_error("null context");
    return HFJK_ERR_INPUT;
  }
  if (n < 1 || n > kEigMaxN || count < 0 || nvec < 0 || nvec > n) {
    hfjk_internal_set_error(("hfjk_eigh_lowest_batched needs 1 <= n <= " + std::to_string(kEigMaxN) +
                             ", count >= 0 and 0 <= nvec <= n")
                                .c_str());
    return HFJK_ERR_INPUT;
  }
  if (count == 0) return HFJK_OK;
  if (a == nullptr || w == nullptr || work == nullptr || info == nullptr) {
    hfjk_internal_set_error("hfjk_eigh_lowest_batched needs device buffers");
    return HFJK_ERR_INPUT;
  }
  hipError_t err = hipSetDevice(ctx->device);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(eig_lowest_kernel, dim3((unsigned)count), dim3(kEigThreads), 0, (hipStream_t)stream, n, nvec,
                       a, w, work, info);
    err = hipGetLastError();
  }
  if (err != hipSuccess) {
    hfjk_internal_set_error((std::string("eig_lowest_kernel launch: ") + hipGetErrorString(err)).c_str());
    return HFJK_ERR_HIP;
  }
  return HFJK_OK;
}

}  // extern "C"
