// DF-J/K kernels for gfx950 (CDNA4): the Coulomb charges and matrix as HBM
// reductions over the tiled packed tensor, and the exchange as a
// half-transform X[P] = B[P] C followed by K = sum_P X[P] X[P]^T, both on
// MFMA with 32-wide LDS-staged k tiles.
//
// MFMA lane layouts (measured on gfx950):
//  * v_mfma_f64_16x16x4f64 : A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
//    D register r of lane l holds D[l / 16 + 4 r][l % 16].
//  * v_mfma_f32_16x16x4f32 : same A/B, D register r of lane l holds
//    D[4 (l / 16) + r][l % 16].
//  * v_mfma_f32_16x16x32_bf16 : A[i][8 (l / 16) + e] in element e of lane i + ...
//    (l % 16 = i), B likewise with l % 16 = j; D as for f32 16x16x4.
#include "hfjk_internal.h"

namespace hfjk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ __forceinline__ void tile_rc(int tile, int& r, int& c) {
  int rr = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((rr + 1) * (rr + 2) / 2 <= tile) ++rr;
  while (rr * (rr + 1) / 2 > tile) --rr;
  r = rr;
  c = tile - rr * (rr + 1) / 2;
}

__device__ __forceinline__ long long tile_index(int r, int c) { return (long long)r * (r + 1) / 2 + c; }

__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Four consecutive packed elements of B[q] starting at e (e % 4 == 0), as
// float64. PREC selects the stored copy.
template <int PREC>
__device__ __forceinline__ void load4(const void* b, long long p, long long naux, int q, long long e,
                                      double out[4]) {
  if (PREC == HFJK_FP64) {
    const double2* src = (const double2*)((const double*)b + (long long)q * p + e);
    const double2 v0 = src[0], v1 = src[1];
    out[0] = v0.x;
    out[1] = v0.y;
    out[2] = v1.x;
    out[3] = v1.y;
  } else if (PREC == HFJK_FP32) {
    const float4 v = *(const float4*)((const float*)b + (long long)q * p + e);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    out[3] = v.w;
  } else {
    const uint16_t* hi = (const uint16_t*)b + (long long)q * p + e;
    const uint16_t* lo = hi + naux * p;
    const ushort4 h = *(const ushort4*)hi;
    const ushort4 l = *(const ushort4*)lo;
    out[0] = (double)(bf16f(h.x) + bf16f(l.x));
    out[1] = (double)(bf16f(h.y) + bf16f(l.y));
    out[2] = (double)(bf16f(h.z) + bf16f(l.z));
    out[3] = (double)(bf16f(h.w) + bf16f(l.w));
  }
}

// ---------------------------------------------------------------- Coulomb

// W = packed weights of D: D on diagonal tiles, D + D^T elsewhere, so that
// <B[P], D> = <packed B[P], W> for symmetric B.
__global__ __launch_bounds__(kThreads) void weights_kernel(const double* __restrict__ d, int nao,
                                                            double* __restrict__ w) {
  const int tile = blockIdx.x;
  int r, c;
  tile_rc(tile, r, c);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int local = threadIdx.x * 4 + v;
    const int row = r * kTile + local / kTile;
    const int col = c * kTile + local % kTile;
    double value = 0.0;
    if (row < nao && col < nao) {
      value = d[(long long)row * nao + col];
      if (r != c) value += d[(long long)col * nao + row];
    }
    w[(long long)tile * kTileElems + local] = value;
  }
}

constexpr int kGammaQ = 64;  // aux functions per gamma block

// partial[tile][q] = <B[q] restricted to tile, W restricted to tile>.
template <int PREC>
__global__ __launch_bounds__(kThreads) void gamma_kernel(const void* __restrict__ b, long long p, int naux,
                                                          const double* __restrict__ w,
                                                          double* __restrict__ partial) {
  __shared__ double red[kThreads / 64][kGammaQ];
  const int tile = blockIdx.x;
  const int q0 = blockIdx.y * kGammaQ;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long e = (long long)tile * kTileElems + threadIdx.x * 4;
  double wv[4];
  {
    const double2* src = (const double2*)(w + e);
    const double2 a = src[0], bb = src[1];
    wv[0] = a.x;
    wv[1] = a.y;
    wv[2] = bb.x;
    wv[3] = bb.y;
  }
  const int qn = (naux - q0) < kGammaQ ? (naux - q0) : kGammaQ;
  for (int qi = 0; qi < qn; ++qi) {
    double bv[4];
    load4<PREC>(b, p, naux, q0 + qi, e, bv);
    double s = bv[0] * wv[0] + bv[1] * wv[1] + bv[2] * wv[2] + bv[3] * wv[3];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) red[wave][qi] = s;
  }
  __syncthreads();
  if (threadIdx.x < qn) {
    const int qi = threadIdx.x;
    const double s = red[0][qi] + red[1][qi] + red[2][qi] + red[3][qi];
    partial[(long long)tile * naux + q0 + qi] = s;
  }
}

__global__ __launch_bounds__(kThreads) void gamma_reduce_kernel(const double* __restrict__ partial, int ntiles,
                                                                 int naux, double* __restrict__ gamma) {
  const int q = blockIdx.x * kThreads + threadIdx.x;
  if (q >= naux) return;
  double s = 0.0;
  for (int t = 0; t < ntiles; ++t) s += partial[(long long)t * naux + q];
  gamma[q] = s;
}

// partial[s][e] = sum over the s-th aux slice of gamma[q] B[q][e], one
// (tile, slice) item per workgroup iteration (the grid may be capped).
template <int PREC>
__global__ __launch_bounds__(kThreads) void coulomb_accumulate_kernel(const void* __restrict__ b, long long p,
                                                                       int naux, int ntiles, int nsplit,
                                                                       const double* __restrict__ gamma,
                                                                       double* __restrict__ partial) {
  const long long nitems = (long long)ntiles * nsplit;
  for (long long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int tile = (int)(item % ntiles);
    const int s = (int)(item / ntiles);
    const int qb = (int)((long long)s * naux / nsplit);
    const int qe = (int)((long long)(s + 1) * naux / nsplit);
    const long long e = (long long)tile * kTileElems + threadIdx.x * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = qb; q < qe; ++q) {
      double bv[4];
      load4<PREC>(b, p, naux, q, e, bv);
      const double g = gamma[q];
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += g * bv[v];
    }
    double2* dst = (double2*)(partial + (long long)s * p + e);
    dst[0] = make_double2(acc[0], acc[1]);
    dst[1] = make_double2(acc[2], acc[3]);
  }
}

__global__ __launch_bounds__(kThreads) void coulomb_unpack_kernel(const double* __restrict__ partial, int nsplit,
                                                                   long long p, int nao,
                                                                   double* __restrict__ out) {
  const long long total = (long long)nao * nao;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int row = (int)(idx / nao);
    const int col = (int)(idx % nao);
    const int r = row / kTile, c = col / kTile;
    long long src;
    if (r >= c) {
      src = tile_index(r, c) * kTileElems + (row % kTile) * kTile + col % kTile;
    } else {
      src = tile_index(c, r) * kTileElems + (col % kTile) * kTile + row % kTile;
    }
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += partial[(long long)k * p + src];
    out[idx] = s;
  }
}

// ---------------------------------------------------------------- Exchange

// Orbitals [nao][nocc] float64 -> padded [npad][ldo] in T (npad = nt * 32).
template <typename T>
__global__ __launch_bounds__(kThreads) void orbitals_kernel(const double* __restrict__ c, int nao, int nocc,
                                                             int npad, int ldo, T* __restrict__ out) {
  const long long total = (long long)npad * ldo;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int k = (int)(idx / ldo);
    const int i = (int)(idx % ldo);
    const double v = (k < nao && i < nocc) ? c[(long long)k * nocc + i] : 0.0;
    out[idx] = (T)v;
  }
}

__global__ __launch_bounds__(kThreads) void occ_signs_kernel(const double* __restrict__ in, int nocc, int ldo,
                                                              double* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= ldo) return;
  out[i] = (i < nocc) ? (in != nullptr ? in[i] : 1.0) : 0.0;
}

// C[k][j'] = V[k][j] sqrt(|l_j|), signs[j'] = sign(l_j), for the retained
// columns j = j' (j' < first) and j = (nao - count) + j' - first otherwise.
// V is column-major (rocSOLVER): V[k][j] = vecs[j * nao + k].
__global__ __launch_bounds__(kThreads) void split_factor_kernel(const double* __restrict__ vecs,
                                                                 const double* __restrict__ vals, int nao,
                                                                 int nneg, int npos, double* __restrict__ c,
                                                                 double* __restrict__ signs) {
  const int r = nneg + npos;
  const long long total = (long long)nao * r;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int k = (int)(idx / r);
    const int jj = (int)(idx % r);
    const int j = jj < nneg ? jj : (nao - npos) + (jj - nneg);
    const double l = vals[j];
    c[idx] = vecs[(long long)j * nao + k] * sqrt(fabs(l));
    if (k == 0) signs[jj] = l < 0.0 ? -1.0 : 1.0;
  }
}

// Pivoted Cholesky of a symmetric positive semidefinite D [n][n], one column
// per step: D = L L^T with L stored by columns (l[k][i] = L[i][k], column
// stride n), column k pivoted on the largest remaining diagonal r_p,
// L[i][k] = (D[p][i] - sum_{j<k} L[i][j] L[p][j]) / sqrt(r_p), r -= L[:,k]^2,
// stopping once that diagonal is <= tol * (largest initial diagonal). Each
// step is two launches (the pivot in one workgroup, the column across the
// GPU), and the state lives on the device so the host checks it only
// every few steps.
constexpr int kPcholThreads = 1024;

// diag = diag(D); st: cut = tol * max diag, no stop, rank 0. One workgroup.
__global__ __launch_bounds__(kPcholThreads) void pchol_init_kernel(const double* __restrict__ d, int n, double tol,
                                                                  double* __restrict__ diag,
                                                                  PcholState* __restrict__ st) {
  __shared__ double wmax[kPcholThreads / 64];
  double top = 0.0;
  for (int i = threadIdx.x; i < n; i += kPcholThreads) {
    const double v = d[(long long)i * n + i];
    diag[i] = v;
    top = fmax(top, v);
  }
  for (int off = 32; off > 0; off >>= 1) top = fmax(top, __shfl_xor(top, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = top;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPcholThreads / 64; ++w) top = fmax(top, wmax[w]);
    st->cut = tol * top;
    st->stop = 0;
    st->rank = 0;
    st->p = -1;
    st->piv = 0.0;
  }
}

// Pivot of column k: the largest remaining diagonal (ties: lowest index);
// stop with rank k when it is not above the cut, or at k = n. One workgroup.
__global__ __launch_bounds__(kPcholThreads) void pchol_pivot_kernel(const double* __restrict__ diag, int n, int k,
                                                                   PcholState* __restrict__ st) {
  __shared__ double wval[kPcholThreads / 64];
  __shared__ int widx[kPcholThreads / 64];
  if (st->stop) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double best = -1.0;
  int at = -1;
  for (int i = t; i < n; i += kPcholThreads) {
    const double v = diag[i];
    if (v > best) {
      best = v;
      at = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off);
    const int oi = __shfl_xor(at, off);
    if (ov > best || (ov == best && oi >= 0 && (at < 0 || oi < at))) {
      best = ov;
      at = oi;
    }
  }
  if (lane == 0) {
    wval[wave] = best;
    widx[wave] = at;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kPcholThreads / 64; ++w)
      if (wval[w] > best || (wval[w] == best && widx[w] >= 0 && (at < 0 || widx[w] < at))) {
        best = wval[w];
        at = widx[w];
      }
    if (k >= n || at < 0 || !(best > st->cut)) {
      st->stop = 1;
      st->rank = k;
    } else {
      st->p = at;
      st->piv = best;
    }
  }
}

// Column k of L: four lanes per row split the sum over the earlier columns.
__global__ __launch_bounds__(kThreads) void pchol_column_kernel(const double* __restrict__ d, int n, int k,
                                                                double* __restrict__ l, double* __restrict__ diag,
                                                                const PcholState* __restrict__ st) {
  if (st->stop) return;
  const int p = st->p;
  const double inv = 1.0 / sqrt(st->piv);
  const long long g = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int row = (int)(g >> 2), part = (int)(g & 3);
  double s = 0.0;
  if (row < n)
    for (int j = part; j < k; j += 4) s += l[(long long)j * n + row] * l[(long long)j * n + p];
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  if (row < n && part == 0) {
    const double v = (d[(long long)p * n + row] - s) * inv;
    l[(long long)k * n + row] = v;
    diag[row] -= v * v;
  }
}

// max |D[i][j] - sum_{k < rank} L[i][k] L[j][k]| into *out (bits of a
// non-negative double, atomic max; L by columns as above); rank read from
// the device.
__global__ __launch_bounds__(kThreads) void factor_residual_kernel(const double* __restrict__ d, int n,
                                                                   const double* __restrict__ l,
                                                                   const int* __restrict__ rank,
                                                                   unsigned long long* __restrict__ out) {
  const int r = *rank;
  double worst = 0.0;
  const long long total = (long long)n * n;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / n), j = (int)(idx % n);
    double v = d[idx];
    for (int k = 0; k < r; ++k) v -= l[(long long)k * n + i] * l[(long long)k * n + j];
    worst = fmax(worst, fabs(v));
  }
  for (int off = 32; off > 0; off >>= 1) worst = fmax(worst, __shfl_xor(worst, off));
  if ((threadIdx.x & 63) == 0 && worst > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(worst));
}

// c [nao][r] = the first r columns of L (stored by columns, stride nao);
// signs = +1.
__global__ __launch_bounds__(kThreads) void compact_factor_kernel(const double* __restrict__ l, int nao, int r,
                                                                  double* __restrict__ c,
                                                                  double* __restrict__ signs) {
  const long long total = (long long)nao * r;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / r), k = (int)(idx % r);
    c[idx] = l[(long long)k * nao + i];
    if (i == 0) signs[k] = 1.0;
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers

hipError_t launch_pchol_init(const double* density, int nao, double tol, double* diag, PcholState* st,
                             hipStream_t s) {
  hipLaunchKernelGGL(pchol_init_kernel, dim3(1), dim3(kPcholThreads), 0, s, density, nao, tol, diag, st);
  return hipGetLastError();
}

hipError_t launch_pchol_step(const double* density, int nao, int k, double* l, double* diag, PcholState* st,
                             hipStream_t s) {
  hipLaunchKernelGGL(pchol_pivot_kernel, dim3(1), dim3(kPcholThreads), 0, s, diag, nao, k, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || k >= nao) return e;
  const long long threads = 4LL * nao;
  hipLaunchKernelGGL(pchol_column_kernel, dim3((unsigned)((threads + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     s, density, nao, k, l, diag, st);
  return hipGetLastError();
}

hipError_t launch_factor_residual(const double* density, int nao, const double* l, const PcholState* st,
                                  unsigned long long* residual, hipStream_t s) {
  const long long nn = (long long)nao * nao;
  hipLaunchKernelGGL(factor_residual_kernel, dim3(grid_for(nn)), dim3(kThreads), 0, s, density, nao, l, &st->rank,
                     residual);
  return hipGetLastError();
}

hipError_t launch_compact_factor(const double* l, int nao, int r, double* c, double* signs, hipStream_t s) {
  const long long n = (long long)nao * r;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(compact_factor_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, l, nao, r, c, signs);
  return hipGetLastError();
}

hipError_t launch_weights(const double* density, int nao, double* w, hipStream_t s) {
  hipLaunchKernelGGL(weights_kernel, dim3((unsigned)n_tile_pairs(nao)), dim3(kThreads), 0, s, density, nao, w);
  return hipGetLastError();
}

hipError_t launch_gamma(const void* b, int precision, long long p, int naux, int ntiles, const double* w,
                        double* partial, hipStream_t s) {
  const dim3 grid((unsigned)ntiles, (unsigned)((naux + kGammaQ - 1) / kGammaQ));
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(gamma_kernel<HFJK_FP64>, grid, dim3(kThreads), 0, s, b, p, naux, w, partial);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(gamma_kernel<HFJK_FP32>, grid, dim3(kThreads), 0, s, b, p, naux, w, partial);
  else
    hipLaunchKernelGGL(gamma_kernel<HFJK_BF16X3>, grid, dim3(kThreads), 0, s, b, p, naux, w, partial);
  return hipGetLastError();
}

hipError_t launch_gamma_reduce(const double* partial, int ntiles, int naux, double* gamma, hipStream_t s) {
  hipLaunchKernelGGL(gamma_reduce_kernel, dim3((unsigned)((naux + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     partial, ntiles, naux, gamma);
  return hipGetLastError();
}

hipError_t launch_coulomb_accumulate(const void* b, int precision, long long p, int naux, int ntiles, int nsplit,
                                     const double* gamma, double* partial, int max_groups, hipStream_t s) {
  long long groups = (long long)ntiles * nsplit;
  if (max_groups > 0) groups = std::min<long long>(groups, max_groups);
  const dim3 grid((unsigned)groups);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_FP64>, grid, dim3(kThreads), 0, s, b, p, naux, ntiles, nsplit,
                       gamma, partial);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_FP32>, grid, dim3(kThreads), 0, s, b, p, naux, ntiles, nsplit,
                       gamma, partial);
  else
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_BF16X3>, grid, dim3(kThreads), 0, s, b, p, naux, ntiles,
                       nsplit, gamma, partial);
  return hipGetLastError();
}

hipError_t launch_coulomb_unpack(const double* partial, int nsplit, long long p, int nao, double* out,
                                 hipStream_t s) {
  const long long n = (long long)nao * nao;
  hipLaunchKernelGGL(coulomb_unpack_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s,
                     partial, nsplit, p, nao, out);
  return hipGetLastError();
}

hipError_t launch_orbitals(const double* c, int nao, int nocc, int precision, void* out, hipStream_t s) {
  const int npad = n_tiles(nao) * kTile;
  const int ldo = occ_ld(nocc);
  const long long n = (long long)npad * ldo;
  const unsigned blocks = grid_for(n);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(orbitals_kernel<double>, dim3(blocks), dim3(kThreads), 0, s, c, nao, nocc, npad, ldo,
                       (double*)out);
  else
    hipLaunchKernelGGL(orbitals_kernel<float>, dim3(blocks), dim3(kThreads), 0, s, c, nao, nocc, npad, ldo,
                       (float*)out);
  return hipGetLastError();
}

hipError_t launch_occ_signs(const double* signs_in, int nocc, int ldo, double* signs_out, hipStream_t s) {
  hipLaunchKernelGGL(occ_signs_kernel, dim3((unsigned)((ldo + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     signs_in, nocc, ldo, signs_out);
  return hipGetLastError();
}

hipError_t launch_split_factor(const double* vecs, const double* vals, int nao, int nneg, int npos, double* c,
                               double* signs, hipStream_t s) {
  const long long n = (long long)nao * (nneg + npos);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(split_factor_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, vecs,
                     vals, nao, nneg, npos, c, signs);
  return hipGetLastError();
}

}  // namespace hfjk
