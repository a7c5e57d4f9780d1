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

// partial[s][e] = sum over the s-th aux slice of gamma[q] B[q][e].
template <int PREC>
__global__ __launch_bounds__(kThreads) void coulomb_accumulate_kernel(const void* __restrict__ b, long long p,
                                                                       int naux, int nsplit,
                                                                       const double* __restrict__ gamma,
                                                                       double* __restrict__ partial) {
  const int tile = blockIdx.x;
  const int s = blockIdx.y;
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

// X[q][m][i] = sum_k B[q][m][k] C[k][i]. Workgroup tile 64 rows x BN
// columns; waves 2 x 2, each 32 x BN/2 as (2 x BN/32) MFMA 16x16 blocks.
// The k loop walks the 32-wide tile columns of the packed tensor: the two
// 32-row tiles of the block are read directly (r >= kt) or transposed
// (r < kt, stored as tile (kt, r)) into the k-major LDS tile As[k][m].
template <typename T, int BN>
__global__ __launch_bounds__(kThreads) void half_transform_kernel(const T* __restrict__ b, long long p, int nt,
                                                                   const T* __restrict__ c, int ldo,
                                                                   T* __restrict__ x, int xrows, int qbase) {
  constexpr int NI = BN / 32;
  __shared__ T As[kTile][kBlock + 2];
  __shared__ T Bs[kTile][BN + 2];
  const int nb = blockIdx.x;
  const int mb = blockIdx.y;
  const int q = qbase + blockIdx.z;
  const int n0 = nb * BN;
  const T* bq = b + (long long)q * p;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  typename Mfma<T>::acc_t acc[2][NI];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  for (int kt = 0; kt < nt; ++kt) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * mb + h;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int e = t * 4 + v;
        const int a = e / kTile;
        const int bc = e % kTile;
        if (r >= nt) {
          As[bc][h * kTile + a] = (T)0;
        } else if (r >= kt) {
          // element (a, bc) of tile (r, kt) is B[m = r*32 + a][k = kt*32 + bc]
          As[bc][h * kTile + a] = bq[tile_index(r, kt) * kTileElems + e];
        } else {
          // element (a, bc) of tile (kt, r) is B[k = kt*32 + a][m = r*32 + bc]
          As[a][h * kTile + bc] = bq[tile_index(kt, r) * kTileElems + e];
        }
      }
    }
    for (int idx = t; idx < kTile * BN; idx += kThreads) {
      const int kk = idx / BN;
      const int jj = idx % BN;
      const int col = n0 + jj;
      Bs[kk][jj] = (col < ldo) ? c[(long long)(kt * kTile + kk) * ldo + col] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
      T af[2], bf[NI];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) af[mi] = As[kk + (lane >> 4)][wm * 32 + mi * 16 + (lane & 15)];
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) bf[ni] = Bs[kk + (lane >> 4)][wn * (BN / 2) + ni * 16 + (lane & 15)];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = Mfma<T>::mma(af[mi], bf[ni], acc[mi][ni]);
    }
    __syncthreads();
  }
  T* xq = x + (long long)q * xrows * ldo;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int col = n0 + wn * (BN / 2) + ni * 16 + (lane & 15);
      if (col < ldo) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mb * kBlock + wm * 32 + mi * 16 + Mfma<T>::row(lane, r);
          xq[(long long)row * ldo + col] = acc[mi][ni][r];
        }
      }
    }
}

// partial[s][pair][64][64] = sum over the s-th aux slice and all columns i of
// X[q][mb*64 + .][i] sign[i] X[q][nb*64 + .][i], pair = mb (mb + 1) / 2 + nb.
template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_kernel(const T* __restrict__ x, int naux, int xrows, int ldo,
                                                         const double* __restrict__ signs, int nsplit,
                                                         int npairs, T* __restrict__ partial) {
  __shared__ T As[kTile][kBlock + 2];
  __shared__ T Bs[kTile][kBlock + 2];
  const int pair = blockIdx.x;
  const int s = blockIdx.y;
  int mb, nb;
  tile_rc(pair, mb, nb);
  const int qb = (int)((long long)s * naux / nsplit);
  const int qe = (int)((long long)(s + 1) * naux / nsplit);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  typename Mfma<T>::acc_t acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  for (int q = qb; q < qe; ++q) {
    const T* xm = x + ((long long)q * xrows + mb * kBlock) * ldo;
    const T* xn = x + ((long long)q * xrows + nb * kBlock) * ldo;
    for (int i0 = 0; i0 < ldo; i0 += kTile) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int idx = t + kThreads * v;
        const int row = idx / kTile;
        const int col = idx % kTile;
        As[col][row] = xm[(long long)row * ldo + i0 + col];
        Bs[col][row] = xn[(long long)row * ldo + i0 + col] * (T)signs[i0 + col];
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kTile; kk += 4) {
        T af[2], bf[2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) af[mi] = As[kk + (lane >> 4)][wm * 32 + mi * 16 + (lane & 15)];
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) bf[ni] = Bs[kk + (lane >> 4)][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = Mfma<T>::mma(af[mi], bf[ni], acc[mi][ni]);
      }
      __syncthreads();
    }
  }
  T* dst = partial + ((long long)s * npairs + pair) * (kBlock * kBlock);
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      const int col = wn * 32 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * 32 + mi * 16 + Mfma<T>::row(lane, r);
        dst[row * kBlock + col] = acc[mi][ni][r];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_reduce_kernel(const T* __restrict__ partial, int nsplit,
                                                                int npairs, int nao, double* __restrict__ out) {
  const long long total = (long long)nao * nao;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int row = (int)(idx / nao);
    const int col = (int)(idx % nao);
    const int bm = row / kBlock, bn = col / kBlock;
    long long pair;
    int local;
    if (bm >= bn) {
      pair = tile_index(bm, bn);
      local = (row % kBlock) * kBlock + col % kBlock;
    } else {
      pair = tile_index(bn, bm);
      local = (col % kBlock) * kBlock + row % kBlock;
    }
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k)
      s += (double)partial[((long long)k * npairs + pair) * (kBlock * kBlock) + local];
    out[idx] = s;
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

}  // namespace

// ---------------------------------------------------------------- launchers

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
                                     const double* gamma, double* partial, hipStream_t s) {
  const dim3 grid((unsigned)ntiles, (unsigned)nsplit);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_FP64>, grid, dim3(kThreads), 0, s, b, p, naux, nsplit, gamma,
                       partial);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_FP32>, grid, dim3(kThreads), 0, s, b, p, naux, nsplit, gamma,
                       partial);
  else
    hipLaunchKernelGGL(coulomb_accumulate_kernel<HFJK_BF16X3>, grid, dim3(kThreads), 0, s, b, p, naux, nsplit,
                       gamma, partial);
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

template <typename T, int BN>
static hipError_t launch_half_transform_t(const T* b, long long p, int naux, int nao, const T* c, int ldo, T* x,
                                          hipStream_t s) {
  const int nt = n_tiles(nao);
  const int xrows = x_rows(nao);
  for (int base = 0; base < naux; base += 65535) {
    const int count = naux - base < 65535 ? naux - base : 65535;
    const dim3 grid((unsigned)((ldo + BN - 1) / BN), (unsigned)(xrows / kBlock), (unsigned)count);
    hipLaunchKernelGGL((half_transform_kernel<T, BN>), grid, dim3(kThreads), 0, s, b, p, nt, c, ldo, x, xrows, base);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

template <typename T>
static hipError_t launch_half_transform_bn(const T* b, long long p, int naux, int nao, const T* c, int ldo, T* x,
                                           hipStream_t s) {
  if (ldo <= 32) return launch_half_transform_t<T, 32>(b, p, naux, nao, c, ldo, x, s);
  if (ldo <= 64) return launch_half_transform_t<T, 64>(b, p, naux, nao, c, ldo, x, s);
  return launch_half_transform_t<T, 128>(b, p, naux, nao, c, ldo, x, s);
}

hipError_t launch_half_transform(const void* b, int precision, long long p, int naux, int nao, const void* c, int ldo,
                                 void* x, hipStream_t s) {
  if (precision == HFJK_FP64)
    return launch_half_transform_bn<double>((const double*)b, p, naux, nao, (const double*)c, ldo, (double*)x, s);
  return launch_half_transform_bn<float>((const float*)b, p, naux, nao, (const float*)c, ldo, (float*)x, s);
}

hipError_t launch_syrk(const void* x, int precision, int naux, int nao, int ldo, const double* signs, int nsplit,
                       void* partial, hipStream_t s) {
  const int xrows = x_rows(nao);
  const int nb = xrows / kBlock;
  const int npairs = nb * (nb + 1) / 2;
  const dim3 grid((unsigned)npairs, (unsigned)nsplit);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(syrk_kernel<double>, grid, dim3(kThreads), 0, s, (const double*)x, naux, xrows, ldo, signs,
                       nsplit, npairs, (double*)partial);
  else
    hipLaunchKernelGGL(syrk_kernel<float>, grid, dim3(kThreads), 0, s, (const float*)x, naux, xrows, ldo, signs,
                       nsplit, npairs, (float*)partial);
  return hipGetLastError();
}

hipError_t launch_syrk_reduce(const void* partial, int precision, int nsplit, int nao, double* out, hipStream_t s) {
  const int nb = x_rows(nao) / kBlock;
  const int npairs = nb * (nb + 1) / 2;
  const long long n = (long long)nao * nao;
  const unsigned blocks = grid_for(n);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(syrk_reduce_kernel<double>, dim3(blocks), dim3(kThreads), 0, s, (const double*)partial, nsplit,
                       npairs, nao, out);
  else
    hipLaunchKernelGGL(syrk_reduce_kernel<float>, dim3(blocks), dim3(kThreads), 0, s, (const float*)partial, nsplit,
                       npairs, nao, out);
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
