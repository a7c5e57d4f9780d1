// Exchange kernels for gfx950 (CDNA4): the half-transform X[P] = B[P] C and
// the symmetric rank-k update K = sum_P X[P] S X[P]^T, on MFMA, in three
// arithmetics (FP64, FP32, BF16x3).
//
// Structure shared by all of them:
//  * one 256-thread workgroup (4 waves of 64) per output tile;
//  * global -> register prefetch of the next k step while the MFMAs of the
//    current one run from LDS (one LDS buffer, two barriers per step);
//  * the SYRK grid is XCD-aware: workgroups are dispatched round-robin over
//    the 8 XCDs, so linear id b runs on XCD b % 8, and every tile of one
//    aux-slice split s is mapped to the same XCD. The slice of X that those
//    tiles stream is then fetched from HBM into one L2 instead of eight.
//
// MFMA lane layouts (measured on gfx950):
//  * v_mfma_f64_16x16x4f64 : A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
//    D register r of lane l = D[l / 16 + 4 r][l % 16];
//  * v_mfma_f32_16x16x4f32 : same A/B, D register r of lane l = D[4 (l / 16) + r][l % 16];
//  * v_mfma_f32_16x16x32_bf16 : A[i][8 (l / 16) + e] in element e of lane l with
//    i = l % 16, B[8 (l / 16) + e][j] likewise with j = l % 16; D as f32 16x16x4.
//
// FP64/FP32 operands are single elements per lane, so the packed B tiles are
// kept in LDS in their stored (row-major) order and a transposed tile is read
// with swapped addressing. BF16x3 operands are 8 consecutive k per lane, so
// tiles stored transposed are transposed in registers (4 x 4 blocks) on their
// way into a [row][k] LDS tile.
#include <algorithm>
#include <type_traits>

#include "hfjk_internal.h"

namespace hfjk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ long long tidx(int r, int c) { return (long long)r * (r + 1) / 2 + c; }

__device__ __forceinline__ void tile_of(int tile, int& r, int& c) {
  int rr = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((rr + 1) * (rr + 2) / 2 <= tile) ++rr;
  while (rr * (rr + 1) / 2 > tile) --rr;
  r = rr;
  c = tile - rr * (rr + 1) / 2;
}

__device__ __forceinline__ uint16_t bf16_of(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float f32_of(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Logical index of linear workgroup b of a W-workgroup 1-D grid: XCD x
// (b % 8) receives the consecutive logical range [x W/8, (x+1) W/8), so
// neighbouring tiles (which share operands) share one L2.
__device__ __forceinline__ int xcd_remap(int b, int w) {
  const int full = w & ~7;
  if (b >= full) return b;
  return (b & 7) * (full >> 3) + (b >> 3);
}

template <typename T>
struct Op;

template <>
struct Op<double> {
  typedef d4 acc_t;
  typedef double2 vec2;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Op<float> {
  typedef f4 acc_t;
  typedef float2 vec2;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Zero a vector unless `keep`, component by component (a select of whole
// vectors is lowered through a private-memory copy).
__device__ __forceinline__ uint2 mask2(bool keep, uint2 v) {
  const uint32_t m = keep ? 0xFFFFFFFFu : 0u;
  return make_uint2(v.x & m, v.y & m);
}

__device__ __forceinline__ uint4 mask4(bool keep, uint4 v) {
  const uint32_t m = keep ? 0xFFFFFFFFu : 0u;
  return make_uint4(v.x & m, v.y & m, v.z & m, v.w & m);
}

// 8 bf16 from LDS as one 16-byte load, reinterpreted as the MFMA operand.
__device__ __forceinline__ bf8 lds_bf8(const uint16_t* p) {
  const uint4 v = *(const uint4*)p;
  return __builtin_bit_cast(bf8, v);
}

__device__ __forceinline__ f4 mma3(bf8 ah, bf8 al, bf8 bh, bf8 bl, f4 acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// Wave arrangement of a 64 x BN half-transform tile: WN waves across the
// columns; each wave covers (64 / WM) rows and (BN / WN) columns.
template <int BN>
struct XShape {
  static constexpr int WN = BN >= 256 ? 4 : 2;
  static constexpr int WM = 4 / WN;
  static constexpr int MI = 64 / WM / 16;   // 16-row blocks per wave
  static constexpr int NI = BN / WN / 16;   // 16-column blocks per wave
};

// ------------------------------------------------------------------ FP64 / FP32 half-transform

// X[q][m][i] = sum_k B[q][m][k] C[k][i]; workgroup tile BM rows (BM / 32
// layout tile rows r = (BM / 32) mb + h) x BN columns; one layout tile
// column kt per k step. BM = 128 doubles the MFMAs per LDS read and halves
// the staged bytes per flop against BM = 64, at two workgroups per CU.
template <typename T, int BN, int BM = 64>
struct XformSmem {
  T At[BM / kTile][kTile][kTile + 1];
  T Cs[kTile][BN + 1];
};

// One BM x BN tile (row block mb, column block nb) of X_q = B_q C for one
// aux function: bq = its packed B, xq = its X rows (xrows of them). The
// whole workgroup calls it; LDS is free again when it returns. With SKIP the
// MFMAs of 16-row blocks at or past nrow, of 16-column blocks at or past ncol
// and of k steps at or past kmax are left out (B and C are zero there, so X
// is the same); each wave decides for its own blocks.
template <typename T, int BN, int BM = 64, bool SKIP = false>
__device__ __forceinline__ void xform_tile(const T* __restrict__ bq, int nt, const T* __restrict__ c, int ldo,
                                           T* __restrict__ xq, int xrows, int mb, int nb,
                                           XformSmem<T, BN, BM>& sm, int nrow = 0, int ncol = 0, int kmax = 0) {
  typedef XShape<BN> S;
  constexpr int NTL = BM / kTile;                  // layout tiles per row block
  constexpr int MI = BM / S::WM / 16;              // 16-row blocks per wave
  constexpr int CCH = kTile * BN / 2 / kThreads;  // 2-element chunks of the C tile per thread
  auto& At = sm.At;
  auto& Cs = sm.Cs;
  const int n0 = nb * BN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave % S::WM;
  const int wn = wave / S::WM;

  typename Op<T>::acc_t acc[MI][S::NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < S::NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  // Prefetch registers: 4 consecutive elements of each of the NTL tiles,
  // and CCH 2-element chunks of the C tile.
  typename Op<T>::vec2 ra[NTL][2];
  typename Op<T>::vec2 rc[CCH];

  auto load = [&](int kt) {
#pragma unroll
    for (int h = 0; h < NTL; ++h) {
      const int r = NTL * mb + h;
      if (r < nt) {
        const long long off = (r >= kt ? tidx(r, kt) : tidx(kt, r)) * kTileElems + t * 4;
        const typename Op<T>::vec2* src = (const typename Op<T>::vec2*)(bq + off);
        ra[h][0] = src[0];
        ra[h][1] = src[1];
      } else {
        ra[h][0].x = ra[h][0].y = ra[h][1].x = ra[h][1].y = (T)0;
      }
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const int kk = chunk / (BN / 2);
      const int col = n0 + (chunk % (BN / 2)) * 2;
      if (col < ldo) {
        rc[v] = *(const typename Op<T>::vec2*)(c + (long long)(kt * kTile + kk) * ldo + col);
      } else {
        rc[v].x = rc[v].y = (T)0;
      }
    }
  };
  auto store = [&]() {
    const int a = (t * 4) / kTile;
    const int bc = (t * 4) % kTile;
#pragma unroll
    for (int h = 0; h < NTL; ++h) {
      At[h][a][bc + 0] = ra[h][0].x;
      At[h][a][bc + 1] = ra[h][0].y;
      At[h][a][bc + 2] = ra[h][1].x;
      At[h][a][bc + 3] = ra[h][1].y;
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const int kk = chunk / (BN / 2);
      const int jj = (chunk % (BN / 2)) * 2;
      Cs[kk][jj] = rc[v].x;
      Cs[kk][jj + 1] = rc[v].y;
    }
  };

  // Wave-uniform block masks of SKIP.
  bool rv[MI], cv[S::NI];
  if constexpr (SKIP) {
    const int wu = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) rv[mi] = mb * BM + (wu % S::WM) * (BM / S::WM) + mi * 16 < nrow;
#pragma unroll
    for (int ni = 0; ni < S::NI; ++ni) cv[ni] = n0 + (wu / S::WM) * (BN / S::WN) + ni * 16 < ncol;
  }

  load(0);
  store();
  __syncthreads();
  for (int kt = 0; kt < nt; ++kt) {
    if (kt + 1 < nt) load(kt + 1);
    // Tile h is stored as (r, kt) when r >= kt (read A[i][k] = At[h][i][k]) and
    // as (kt, r) otherwise (read A[i][k] = At[h][k][i]).
    bool direct[NTL];
#pragma unroll
    for (int h = 0; h < NTL; ++h) direct[h] = (NTL * mb + h) >= kt;
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
      if constexpr (SKIP) {
        if (kt * kTile + kk >= kmax) break;
      }
      const int k = kk + (lane >> 4);
      T af[MI], bf[S::NI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) {
        const int row = wm * (BM / S::WM) + mi * 16 + (lane & 15);
        const int h = row >> 5;
        const int i = row & 31;
        af[mi] = direct[h] ? At[h][i][k] : At[h][k][i];
      }
#pragma unroll
      for (int ni = 0; ni < S::NI; ++ni) bf[ni] = Cs[k][wn * (BN / S::WN) + ni * 16 + (lane & 15)];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < S::NI; ++ni) {
          if constexpr (SKIP) {
            if (!(rv[mi] && cv[ni])) continue;
          }
          acc[mi][ni] = Op<T>::mma(af[mi], bf[ni], acc[mi][ni]);
        }
    }
    __syncthreads();
    if (kt + 1 < nt) {
      store();
      __syncthreads();
    }
  }
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < S::NI; ++ni) {
      const int col = n0 + wn * (BN / S::WN) + ni * 16 + (lane & 15);
      if (col < ldo) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mb * BM + wm * (BM / S::WM) + mi * 16 + Op<T>::row(lane, r);
          if (row < xrows) xq[(long long)row * ldo + col] = acc[mi][ni][r];
        }
      }
    }
}

// The half-transform tile of xform_tile for each aux function q in
// [q0, q1) (the batched kernels' items): b = the packed B of aux function 0 (q's at
// stride p), x = X of aux function 0 (q's at stride xrows * ldo). The
// (q, k tile) steps run as one pipeline: the next step's loads are issued
// before the current step's MFMAs, across aux functions too, and each q's
// tile is written out after its last k tile. The whole workgroup calls it;
// LDS is free again when it returns. With SKIP the MFMAs of 16-row blocks
// at or past nrow, of 16-column blocks at or past ncol and of k steps at or
// past kmax are left out (B and C are zero there, so X is the same); each
// wave decides for its own blocks.
template <typename T, int BN, int BM = 64, bool SKIP = false>
__device__ __forceinline__ void xform_tiles(const T* __restrict__ b, long long p, int nt, const T* __restrict__ c,
                                            int ldo, T* __restrict__ x, int xrows, int mb, int nb, int q0, int q1,
                                            XformSmem<T, BN, BM>& sm, int nrow = 0, int ncol = 0, int kmax = 0) {
  typedef XShape<BN> S;
  constexpr int NTL = BM / kTile;                  // layout tiles per row block
  constexpr int MI = BM / S::WM / 16;              // 16-row blocks per wave
  constexpr int CCH = kTile * BN / 2 / kThreads;  // 2-element chunks of the C tile per thread
  auto& At = sm.At;
  auto& Cs = sm.Cs;
  const int n0 = nb * BN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave % S::WM;
  const int wn = wave / S::WM;

  typename Op<T>::acc_t acc[MI][S::NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < S::NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  // Prefetch registers: 4 consecutive elements of each of the NTL tiles,
  // and CCH 2-element chunks of the C tile.
  typename Op<T>::vec2 ra[NTL][2];
  typename Op<T>::vec2 rc[CCH];

  auto load = [&](int q, int kt) {
    const T* bq = b + (long long)q * p;
#pragma unroll
    for (int h = 0; h < NTL; ++h) {
      const int r = NTL * mb + h;
      if (r < nt) {
        const long long off = (r >= kt ? tidx(r, kt) : tidx(kt, r)) * kTileElems + t * 4;
        const typename Op<T>::vec2* src = (const typename Op<T>::vec2*)(bq + off);
        ra[h][0] = src[0];
        ra[h][1] = src[1];
      } else {
        ra[h][0].x = ra[h][0].y = ra[h][1].x = ra[h][1].y = (T)0;
      }
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const int kk = chunk / (BN / 2);
      const int col = n0 + (chunk % (BN / 2)) * 2;
      if (col < ldo) {
        rc[v] = *(const typename Op<T>::vec2*)(c + (long long)(kt * kTile + kk) * ldo + col);
      } else {
        rc[v].x = rc[v].y = (T)0;
      }
    }
  };
  auto store = [&]() {
    const int a = (t * 4) / kTile;
    const int bc = (t * 4) % kTile;
#pragma unroll
    for (int h = 0; h < NTL; ++h) {
      At[h][a][bc + 0] = ra[h][0].x;
      At[h][a][bc + 1] = ra[h][0].y;
      At[h][a][bc + 2] = ra[h][1].x;
      At[h][a][bc + 3] = ra[h][1].y;
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const int kk = chunk / (BN / 2);
      const int jj = (chunk % (BN / 2)) * 2;
      Cs[kk][jj] = rc[v].x;
      Cs[kk][jj + 1] = rc[v].y;
    }
  };

  // Wave-uniform block masks of SKIP.
  bool rv[MI], cv[S::NI];
  if constexpr (SKIP) {
    const int wu = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) rv[mi] = mb * BM + (wu % S::WM) * (BM / S::WM) + mi * 16 < nrow;
#pragma unroll
    for (int ni = 0; ni < S::NI; ++ni) cv[ni] = n0 + (wu / S::WM) * (BN / S::WN) + ni * 16 < ncol;
  }

  const int steps = (q1 - q0) * nt;
  if (steps <= 0) return;
  load(q0, 0);
  store();
  __syncthreads();
  int q = q0, kt = 0;
  for (int step = 0; step < steps; ++step) {
    const int nq = kt + 1 < nt ? q : q + 1;
    const int nkt = kt + 1 < nt ? kt + 1 : 0;
    if (step + 1 < steps) load(nq, nkt);
    // Tile h is stored as (r, kt) when r >= kt (read A[i][k] = At[h][i][k]) and
    // as (kt, r) otherwise (read A[i][k] = At[h][k][i]).
    bool direct[NTL];
#pragma unroll
    for (int h = 0; h < NTL; ++h) direct[h] = (NTL * mb + h) >= kt;
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
      if constexpr (SKIP) {
        if (kt * kTile + kk >= kmax) break;
      }
      const int k = kk + (lane >> 4);
      T af[MI], bf[S::NI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) {
        const int row = wm * (BM / S::WM) + mi * 16 + (lane & 15);
        const int h = row >> 5;
        const int i = row & 31;
        af[mi] = direct[h] ? At[h][i][k] : At[h][k][i];
      }
#pragma unroll
      for (int ni = 0; ni < S::NI; ++ni) bf[ni] = Cs[k][wn * (BN / S::WN) + ni * 16 + (lane & 15)];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < S::NI; ++ni) {
          if constexpr (SKIP) {
            if (!(rv[mi] && cv[ni])) continue;
          }
          acc[mi][ni] = Op<T>::mma(af[mi], bf[ni], acc[mi][ni]);
        }
    }
    __syncthreads();
    if (kt == nt - 1) {
      // This aux function's tile is complete: write it and start the next one from zero.
      const int qw = q;
      T* xq = x + (long long)qw * xrows * ldo;
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < S::NI; ++ni) {
          const int col = n0 + wn * (BN / S::WN) + ni * 16 + (lane & 15);
          if (col < ldo) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = mb * BM + wm * (BM / S::WM) + mi * 16 + Op<T>::row(lane, r);
              if (row < xrows) xq[(long long)row * ldo + col] = acc[mi][ni][r];
            }
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;
        }
    }
    if (step + 1 < steps) {
      store();
      __syncthreads();
    }
    q = nq;
    kt = nkt;
  }
}

template <typename T, int BN, int BM>
__global__ __launch_bounds__(kThreads) void xform_kernel(const T* __restrict__ b, long long p, int nt,
                                                          const T* __restrict__ c, int ldo, T* __restrict__ x,
                                                          int xrows, int qbase, int nblocks_n, int nblocks_m,
                                                          int nwg) {
  __shared__ XformSmem<T, BN, BM> sm;
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  xform_tile<T, BN, BM>(b + (long long)q * p, nt, c, ldo, x + (long long)q * xrows * ldo, xrows, mb, nb, sm);
}

// The same kernel held to 256 registers per lane (two waves per SIMD): the
// 128 x 128 FP64 tile needs 258 otherwise.
template <typename T, int BN, int BM>
__global__ __launch_bounds__(kThreads, 2) void xform_kernel_w2(const T* __restrict__ b, long long p, int nt,
                                                                const T* __restrict__ c, int ldo, T* __restrict__ x,
                                                                int xrows, int qbase, int nblocks_n, int nblocks_m,
                                                                int nwg) {
  __shared__ XformSmem<T, BN, BM> sm;
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  xform_tile<T, BN, BM>(b + (long long)q * p, nt, c, ldo, x + (long long)q * xrows * ldo, xrows, mb, nb, sm);
}

// ------------------------------------------------------------------ BF16x3 half-transform

constexpr int kRowK = 40;  // bf16 LDS row: 32 k + 8 pad (80 bytes, 16-byte aligned rows)

// X in BF16x3: workgroup tile 128 rows (four layout tile rows) x BN columns,
// waves 2 x 2 each (64 x BN/2). A is staged [m][k]; C^T (stored [i][k])
// is staged [i][k].
template <int BN>
struct Xform16Smem {
  uint16_t Ah[128][kRowK];
  uint16_t Al[128][kRowK];
  uint16_t Bh[BN][kRowK];
  uint16_t Bl[BN][kRowK];
};

// One 128 x BN tile of X_q in BF16x3: hq/lq = the packed hi/lo planes of
// B_q, xhi/xlo = X_q's hi/lo rows (xrows rows of ldo).
template <int BN>
__device__ __forceinline__ void xform16_tile(const uint16_t* __restrict__ hq, const uint16_t* __restrict__ lq, int nt,
                                             const uint16_t* __restrict__ cthi, const uint16_t* __restrict__ ctlo,
                                             int npad, int ldo, uint16_t* __restrict__ xhi,
                                             uint16_t* __restrict__ xlo, int xrows, int mb, int nb,
                                             Xform16Smem<BN>& sm) {
  constexpr int MI = 4;            // 64 rows per wave
  constexpr int NI = BN / 2 / 16;  // BN/2 columns per wave
  constexpr int CCH = BN * 4 / kThreads > 0 ? BN * 4 / kThreads : 1;  // 16-byte chunks of C^T per plane
  auto& Ah = sm.Ah;
  auto& Al = sm.Al;
  auto& Bh = sm.Bh;
  auto& Bl = sm.Bl;
  const int n0 = nb * BN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;

  f4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  // Each thread owns one 4 x 4 block of one of the four layout tiles.
  const int th = t >> 6;            // tile row within the block: r = 4 mb + th
  const int a0 = ((t & 63) >> 3) * 4;
  const int b0 = (t & 7) * 4;
  const int r_own = 4 * mb + th;
  uint2 rah[4], ral[4];
  uint4 rch[CCH], rcl[CCH];

  // Loads are unconditional (from a valid clamped address); the masks (bitwise,
  // which keeps the prefetch arrays in registers) are applied in store(), so
  // the wait for the loads falls at the end of the k step, not at its start.
  const bool own = r_own < nt;
  auto load = [&](int kt) {
    const int rc = own ? r_own : 0;
    const long long base = (rc >= kt ? tidx(rc, kt) : tidx(kt, rc)) * kTileElems + a0 * kTile + b0;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      rah[rr] = *(const uint2*)(hq + base + rr * kTile);
      ral[rr] = *(const uint2*)(lq + base + rr * kTile);
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;  // BN rows x 4 chunks of 8 bf16
      const int jj = chunk >> 2;
      const int k8 = (chunk & 3) * 8;
      const int col = n0 + jj;
      const bool ok = chunk < BN * 4 && col < ldo;
      const long long off = ok ? (long long)col * npad + kt * kTile + k8 : 0;
      rch[v] = *(const uint4*)(cthi + off);
      rcl[v] = *(const uint4*)(ctlo + off);
    }
  };
  auto store = [&](int kt) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      rah[rr] = mask2(own, rah[rr]);
      ral[rr] = mask2(own, ral[rr]);
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const bool ok = chunk < BN * 4 && n0 + (chunk >> 2) < ldo;
      rch[v] = mask4(ok, rch[v]);
      rcl[v] = mask4(ok, rcl[v]);
    }
    const bool direct = r_own >= kt;  // also true for r_own >= nt (zeros)
    if (direct) {
      // element (a, b) is B[m = r*32 + a][k = kt*32 + b]
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        *(uint2*)&Ah[th * kTile + a0 + rr][b0] = rah[rr];
        *(uint2*)&Al[th * kTile + a0 + rr][b0] = ral[rr];
      }
    } else {
      // element (a, b) is B[k = kt*32 + a][m = r*32 + b]: transpose the 4 x 4 block
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        uint2 h, l;
        const int sh = (cc & 1) * 16;
        auto pick = [&](const uint2& v) -> uint32_t { return ((cc < 2 ? v.x : v.y) >> sh) & 0xFFFFu; };
        h.x = pick(rah[0]) | (pick(rah[1]) << 16);
        h.y = pick(rah[2]) | (pick(rah[3]) << 16);
        l.x = pick(ral[0]) | (pick(ral[1]) << 16);
        l.y = pick(ral[2]) | (pick(ral[3]) << 16);
        *(uint2*)&Ah[th * kTile + b0 + cc][a0] = h;
        *(uint2*)&Al[th * kTile + b0 + cc][a0] = l;
      }
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      if (chunk < BN * 4) {
        const int jj = chunk >> 2;
        const int k8 = (chunk & 3) * 8;
        *(uint4*)&Bh[jj][k8] = rch[v];
        *(uint4*)&Bl[jj][k8] = rcl[v];
      }
    }
  };

  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nt; ++kt) {
    if (kt + 1 < nt) load(kt + 1);
    const int kq = 8 * (lane >> 4);
    bf8 ah[MI], al[MI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
      ah[mi] = lds_bf8(&Ah[wm * 64 + mi * 16 + (lane & 15)][kq]);
      al[mi] = lds_bf8(&Al[wm * 64 + mi * 16 + (lane & 15)][kq]);
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const bf8 bh = lds_bf8(&Bh[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
      const bf8 bl = lds_bf8(&Bl[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) acc[mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mi][ni]);
    }
    __syncthreads();
    if (kt + 1 < nt) {
      store(kt + 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int col = n0 + wn * (BN / 2) + ni * 16 + (lane & 15);
      if (col < ldo) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mb * 128 + wm * 64 + mi * 16 + 4 * (lane >> 4) + r;
          if (row < xrows) {
            const float v = acc[mi][ni][r];
            const uint16_t h = bf16_of(v);
            const long long off = (long long)row * ldo + col;
            xhi[off] = h;
            xlo[off] = bf16_of(v - f32_of(h));
          }
        }
      }
    }
}

// xform16_tile for each aux function q in [q0, q1) (the batched kernel's
// items): hq/lq = the packed hi/lo planes of aux function 0 (q's at stride
// p), xhi/xlo = X of aux function 0 (q's at stride xrows * ldo). The (q, k
// tile) steps run as one pipeline, the next step's loads issued before the
// current step's MFMAs across aux functions too, and each q's tile is
// written after its last k tile. With SKIP the MFMAs of 16-row blocks at or
// past nrow and of 16-column blocks at or past ncol are left out (B and C
// are zero there, so X is the same); each wave decides for its own blocks.
template <int BN, bool SKIP>
__device__ __forceinline__ void xform16_tiles(const uint16_t* __restrict__ hq, const uint16_t* __restrict__ lq,
                                              long long p, int nt, const uint16_t* __restrict__ cthi,
                                              const uint16_t* __restrict__ ctlo, int npad, int ldo,
                                              uint16_t* __restrict__ xhi, uint16_t* __restrict__ xlo, int xrows,
                                              int mb, int nb, int q0, int q1, Xform16Smem<BN>& sm, int nrow,
                                              int ncol) {
  constexpr int MI = 4;            // 64 rows per wave
  constexpr int NI = BN / 2 / 16;  // BN/2 columns per wave
  constexpr int CCH = BN * 4 / kThreads > 0 ? BN * 4 / kThreads : 1;  // 16-byte chunks of C^T per plane
  auto& Ah = sm.Ah;
  auto& Al = sm.Al;
  auto& Bh = sm.Bh;
  auto& Bl = sm.Bl;
  const int n0 = nb * BN;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;

  f4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  // Each thread owns one 4 x 4 block of one of the four layout tiles.
  const int th = t >> 6;
  const int a0 = ((t & 63) >> 3) * 4;
  const int b0 = (t & 7) * 4;
  const int r_own = 4 * mb + th;
  uint2 rah[4], ral[4];
  uint4 rch[CCH], rcl[CCH];
  const bool own = r_own < nt;
  auto load = [&](int q, int kt) {
    const int rc = own ? r_own : 0;
    const long long base = (long long)q * p + (rc >= kt ? tidx(rc, kt) : tidx(kt, rc)) * kTileElems + a0 * kTile + b0;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      rah[rr] = *(const uint2*)(hq + base + rr * kTile);
      ral[rr] = *(const uint2*)(lq + base + rr * kTile);
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;  // BN rows x 4 chunks of 8 bf16
      const int jj = chunk >> 2;
      const int k8 = (chunk & 3) * 8;
      const int col = n0 + jj;
      const bool ok = chunk < BN * 4 && col < ldo;
      const long long off = ok ? (long long)col * npad + kt * kTile + k8 : 0;
      rch[v] = *(const uint4*)(cthi + off);
      rcl[v] = *(const uint4*)(ctlo + off);
    }
  };
  auto store = [&](int kt) {
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      rah[rr] = mask2(own, rah[rr]);
      ral[rr] = mask2(own, ral[rr]);
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      const bool ok = chunk < BN * 4 && n0 + (chunk >> 2) < ldo;
      rch[v] = mask4(ok, rch[v]);
      rcl[v] = mask4(ok, rcl[v]);
    }
    const bool direct = r_own >= kt;  // also true for r_own >= nt (zeros)
    if (direct) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        *(uint2*)&Ah[th * kTile + a0 + rr][b0] = rah[rr];
        *(uint2*)&Al[th * kTile + a0 + rr][b0] = ral[rr];
      }
    } else {
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        uint2 h, l;
        const int sh = (cc & 1) * 16;
        auto pick = [&](const uint2& v) -> uint32_t { return ((cc < 2 ? v.x : v.y) >> sh) & 0xFFFFu; };
        h.x = pick(rah[0]) | (pick(rah[1]) << 16);
        h.y = pick(rah[2]) | (pick(rah[3]) << 16);
        l.x = pick(ral[0]) | (pick(ral[1]) << 16);
        l.y = pick(ral[2]) | (pick(ral[3]) << 16);
        *(uint2*)&Ah[th * kTile + b0 + cc][a0] = h;
        *(uint2*)&Al[th * kTile + b0 + cc][a0] = l;
      }
    }
#pragma unroll
    for (int v = 0; v < CCH; ++v) {
      const int chunk = t + kThreads * v;
      if (chunk < BN * 4) {
        const int jj = chunk >> 2;
        const int k8 = (chunk & 3) * 8;
        *(uint4*)&Bh[jj][k8] = rch[v];
        *(uint4*)&Bl[jj][k8] = rcl[v];
      }
    }
  };

  // Wave-uniform block masks of SKIP.
  bool rv[MI], cv[NI];
  {
    const int wu = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) rv[mi] = !SKIP || mb * 128 + (wu & 1) * 64 + mi * 16 < nrow;
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) cv[ni] = !SKIP || n0 + (wu >> 1) * (BN / 2) + ni * 16 < ncol;
  }

  const int steps = (q1 - q0) * nt;
  if (steps <= 0) return;
  load(q0, 0);
  store(0);
  __syncthreads();
  int q = q0, kt = 0;
  for (int step = 0; step < steps; ++step) {
    const int nq = kt + 1 < nt ? q : q + 1;
    const int nkt = kt + 1 < nt ? kt + 1 : 0;
    if (step + 1 < steps) load(nq, nkt);
    const int kq = 8 * (lane >> 4);
    bf8 ah[MI], al[MI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
      ah[mi] = lds_bf8(&Ah[wm * 64 + mi * 16 + (lane & 15)][kq]);
      al[mi] = lds_bf8(&Al[wm * 64 + mi * 16 + (lane & 15)][kq]);
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      if (!cv[ni]) continue;
      const bf8 bh = lds_bf8(&Bh[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
      const bf8 bl = lds_bf8(&Bl[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
        if (rv[mi]) acc[mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mi][ni]);
    }
    __syncthreads();
    if (kt == nt - 1) {
      // This aux function's tile is complete: write it and start the next one from zero.
      const long long xq = (long long)q * xrows * ldo;
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
          const int col = n0 + wn * (BN / 2) + ni * 16 + (lane & 15);
          if (col < ldo) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = mb * 128 + wm * 64 + mi * 16 + 4 * (lane >> 4) + r;
              if (row < xrows) {
                const float v = acc[mi][ni][r];
                const uint16_t h = bf16_of(v);
                const long long off = xq + (long long)row * ldo + col;
                xhi[off] = h;
                xlo[off] = bf16_of(v - f32_of(h));
              }
            }
          }
          acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (step + 1 < steps) {
      store(nkt);
      __syncthreads();
    }
    q = nq;
    kt = nkt;
  }
}

template <int BN>
__global__ __launch_bounds__(kThreads) void xform_bf16x3_kernel(
    const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo, long long p, int nt,
    const uint16_t* __restrict__ cthi, const uint16_t* __restrict__ ctlo, int npad, int ldo,
    uint16_t* __restrict__ xhi, uint16_t* __restrict__ xlo, int xrows, int qbase, int nblocks_n, int nblocks_m,
    int nwg) {
  __shared__ __attribute__((aligned(16))) Xform16Smem<BN> sm;
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;  // 128-row block
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  const long long xbase = (long long)q * xrows * ldo;
  xform16_tile<BN>(bhi + (long long)q * p, blo + (long long)q * p, nt, cthi, ctlo, npad, ldo, xhi + xbase,
                   xlo + xbase, xrows, mb, nb, sm);
}

// ------------------------------------------------------------------ Coulomb charges from X
//
// gamma_q = <B_q, C C^T> = sum_{m, i} X_q[m][i] C[m][i]: once X is
// half-transformed, the charges of D = C C^T cost one pass over X instead of
// one over B. C is the zero-padded orbital copy in X's row layout
// ([npad][ldo]; FP64 for X in FP64, FP32 otherwise). Workgroup (q, g) sums
// rows [kChargeRows g, kChargeRows (g + 1)) of X_q into parts[q][g];
// charge_reduce_kernel adds the parts in a fixed order.
// F: 0 = X in FP64, 1 = FP32, 2 = BF16x3 (hi plane, then lo plane `plane`
// elements later), 3 = BF16x3 interleaved (each row as ldo / 32 blocks of 32
// hi then 32 lo values).
constexpr int kChargeRows = 128;

// V consecutive X values from element idx (16 bytes of the plane, or of each
// plane for BF16x3), as doubles.
template <int F>
struct XVec {
  static constexpr int V = F == 0 ? 2 : F == 1 ? 4 : 8;
  // Element (row m, column i) of X_q: F 0-2 row-major [xrows][ldo] per plane; F 3 (BF16x3
  // interleaved) [xrows][ldo / 32][64], hi then lo per 32-column block.
  static __device__ __forceinline__ long long index(long long xq, int m, int i, int ldo) {
    if constexpr (F == 3)
      return 2 * xq + (long long)m * 2 * ldo + (i >> 5) * 64 + (i & 31);
    else
      return xq + (long long)m * ldo + i;
  }
  static __device__ __forceinline__ void load(const void* __restrict__ x, long long plane, long long idx,
                                              double (&out)[V]) {
    if constexpr (F == 0) {
      const double2 v = *(const double2*)((const double*)x + idx);
      out[0] = v.x;
      out[1] = v.y;
    } else if constexpr (F == 1) {
      const float4 v = *(const float4*)((const float*)x + idx);
      out[0] = v.x;
      out[1] = v.y;
      out[2] = v.z;
      out[3] = v.w;
    } else {
      const uint16_t* h = (const uint16_t*)x;
      const uint4 hv = *(const uint4*)(h + idx);
      const uint4 lv = *(const uint4*)(h + idx + (F == 3 ? 32 : plane));
      const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
      const uint32_t lw[4] = {lv.x, lv.y, lv.z, lv.w};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint16_t hb = (uint16_t)(hw[e >> 1] >> ((e & 1) * 16));
        const uint16_t lb = (uint16_t)(lw[e >> 1] >> ((e & 1) * 16));
        out[e] = (double)f32_of(hb) + (double)f32_of(lb);
      }
    }
  }
};

// V consecutive padded orbital values from element idx, as doubles.
template <typename TC, int V>
__device__ __forceinline__ void load_orbitals(const TC* __restrict__ c, long long idx, double (&out)[V]) {
  if constexpr (sizeof(TC) == 8) {
#pragma unroll
    for (int e = 0; e < V; e += 2) {
      const double2 v = *(const double2*)(c + idx + e);
      out[e] = v.x;
      out[e + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < V; e += 4) {
      const float4 v = *(const float4*)(c + idx + e);
      out[e] = v.x;
      out[e + 1] = v.y;
      out[e + 2] = v.z;
      out[e + 3] = v.w;
    }
  }
}

// One (aux function, row group) item per workgroup iteration; the grid may
// be smaller than the item count (a capped grid leaves room on every CU for
// the SYRK it runs beside).
template <int F>
__global__ __launch_bounds__(kThreads) void charges_from_x_kernel(const void* __restrict__ x, long long plane, int nao,
                                                                   int xrows, int ldo, const void* __restrict__ cpad,
                                                                   int ngroups, long long nitems,
                                                                   double* __restrict__ parts) {
  typedef typename std::conditional<F == 0, double, float>::type TC;
  const TC* __restrict__ c = (const TC*)cpad;
  __shared__ double red[kThreads / 64];
  for (long long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int q = (int)(item / ngroups);
    const int g = (int)(item - (long long)q * ngroups);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long xq = (long long)q * xrows * ldo;
    const int m_end = min(nao, (g + 1) * kChargeRows);
    constexpr int V = XVec<F>::V;
    double part = 0.0;
    // Rows [m_begin, m_end) of X_q are one contiguous range of
    // (m_end - m_begin) * ldo elements (F 3: of both parts): the workgroup
    // streams it in 16-byte groups (ldo is a multiple of 32, so no group
    // straddles two rows), group f of thread t being f = t + kThreads j. The
    // padding is zero, so the loop body has no branch and the loads of
    // successive groups overlap.
    const int m_begin = g * kChargeRows;
    const int per_row = ldo / V;
    const int ngroups_blk = (m_end - m_begin) * per_row;
    const int step_m = kThreads / per_row;
    const int step_g = kThreads - step_m * per_row;
    int m = m_begin + (int)threadIdx.x / per_row;
    int gi = (int)threadIdx.x % per_row;  // group within the row
#pragma unroll 4
    for (int f = threadIdx.x; f < ngroups_blk; f += kThreads) {
      const int i0 = gi * V;
      double xv[V], cv[V];
      XVec<F>::load(x, plane, XVec<F>::index(xq, m, i0, ldo), xv);
      load_orbitals<TC, V>(c, (long long)m * ldo + i0, cv);
#pragma unroll
      for (int e = 0; e < V; ++e) part += xv[e] * cv[e];
      m += step_m;
      gi += step_g;
      if (gi >= per_row) {
        gi -= per_row;
        ++m;
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) sum += red[w];
      parts[item] = sum;
    }
    __syncthreads();
  }
}

// gamma[q] = sum_g parts[q][g], in a fixed order.
__global__ __launch_bounds__(kThreads) void charge_reduce_kernel(const double* __restrict__ parts, int naux,
                                                                  int ngroups, double* __restrict__ gamma) {
  const int q = blockIdx.x * kThreads + threadIdx.x;
  if (q >= naux) return;
  double sum = 0.0;
  for (int g = 0; g < ngroups; ++g) sum += parts[(long long)q * ngroups + g];
  gamma[q] = sum;
}

// Orbitals [nao][nocc] float64 -> transposed split Ct[ldo][npad], zero-padded:
// ILV = false: the hi plane, then the lo plane; ILV = true: each row as
// npad / 32 blocks of 64 (32 hi, then the same 32 lo), the layout of the
// bf16 full copy of B.
template <bool ILV>
__global__ __launch_bounds__(kThreads) void orbitals_bf16x3_kernel(const double* __restrict__ c, int nao, int nocc,
                                                                    int npad, int ldo, uint16_t* __restrict__ hi,
                                                                    uint16_t* __restrict__ lo) {
  const long long total = (long long)npad * ldo;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / npad);
    const int k = (int)(idx % npad);
    const float v = (k < nao && i < nocc) ? (float)c[(long long)k * nocc + i] : 0.0f;
    const uint16_t h = bf16_of(v);
    if constexpr (ILV) {
      const long long o = (long long)i * 2 * npad + (k >> 5) * 64 + (k & 31);
      hi[o] = h;
      hi[o + 32] = bf16_of(v - f32_of(h));
    } else {
      hi[idx] = h;
      lo[idx] = bf16_of(v - f32_of(h));
    }
  }
}

// ------------------------------------------------------------------ SYRK

constexpr int kSyrkTile = 128;

// Split s and tile pair of linear workgroup b (XCD-aware, nsplit % 8 == 0).
__device__ __forceinline__ void syrk_coords(int b, int npairs, int& s, int& pair) {
  const int xcd = b & 7;
  const int slot = b >> 3;
  s = xcd + 8 * (slot / npairs);
  pair = slot % npairs;
}

// Partial slots of split s and the aux functions of slot c. run = 0 (FP64):
// one slot per split, [s naux / nsplit, (s + 1) naux / nsplit). run > 0 (the
// FP32-accumulating modes): slot c holds aux functions [c run, (c + 1) run),
// and split s takes slots [s nslot / nsplit, (s + 1) nslot / nsplit), so an
// FP32 accumulator never sums more than run aux functions whatever the
// split count.
__device__ __forceinline__ void slot_range(int s, int nsplit, int naux, int run, int& cb, int& ce) {
  if (run == 0) {
    cb = s;
    ce = s + 1;
    return;
  }
  const int nslot = (naux + run - 1) / run;
  cb = (int)((long long)s * nslot / nsplit);
  ce = (int)((long long)(s + 1) * nslot / nsplit);
}

__device__ __forceinline__ void slot_aux(int c, int nsplit, int naux, int run, int& qb, int& qe) {
  if (run == 0) {
    qb = (int)((long long)c * naux / nsplit);
    qe = (int)((long long)(c + 1) * naux / nsplit);
    return;
  }
  qb = c * run;
  qe = min(naux, qb + run);
}

template <typename T>
struct SyrkSmem {
  T As[kSyrkTile][kTile + 1];
  T Bs[kSyrkTile][kTile + 1];
};

// FP64 / FP32: one TS x TS tile (row block mb, column block nb; TS = 128, or
// 64 for the small instances of a batch) of sum_{q in [qb, qe)} X_q S X_q^T
// into dst[TS][TS]; x = the instance's X.
template <typename T, int TS = kSyrkTile>
__device__ __forceinline__ void syrk_tile(const T* __restrict__ x, int qb, int qe, int xrows, int ldo,
                                          const double* __restrict__ signs, int mb, int nb, T* __restrict__ dst,
                                          SyrkSmem<T>& sm) {
  static_assert(TS == 64 || TS == 128, "SYRK tiles are 64 or 128 rows");
  constexpr int MI = TS / 32, NI = TS / 32;  // wave tile TS/2 x TS/2
  constexpr int NV = TS / 16;                // 2-element chunks per thread and operand
  auto& As = sm.As;
  auto& Bs = sm.Bs;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  typename Op<T>::acc_t acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  // TS rows x 32 columns per operand = 16 TS 2-element chunks: TS / 16 per
  // thread. The prefetch only copies loaded values into registers; the row
  // masks and the column signs are applied in store(), so the wait for the
  // loads falls at the end of the k step, behind its MFMAs, not at its start.
  typename Op<T>::vec2 ra[NV], rb[NV];
  T sg[2];
  const int nchunk = ldo / kTile;
  const int mrows = min(TS, xrows - mb * TS);
  const int nrows = min(TS, xrows - nb * TS);
  auto load = [&](int q, int i0) {
    const T* xm = x + ((long long)q * xrows + mb * TS) * ldo + i0;
    const T* xn = x + ((long long)q * xrows + nb * TS) * ldo + i0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int chunk = t + kThreads * v;
      const int row = chunk >> 4;
      const int c2 = (chunk & 15) * 2;
      ra[v] = *(const typename Op<T>::vec2*)(xm + (long long)(row < mrows ? row : 0) * ldo + c2);
      rb[v] = *(const typename Op<T>::vec2*)(xn + (long long)(row < nrows ? row : 0) * ldo + c2);
    }
    // Every chunk of a thread has the same column pair c2 = (t & 15) * 2.
    sg[0] = (T)signs[i0 + (t & 15) * 2];
    sg[1] = (T)signs[i0 + (t & 15) * 2 + 1];
  };
  auto store = [&]() {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int chunk = t + kThreads * v;
      const int row = chunk >> 4;
      const int c2 = (chunk & 15) * 2;
      const T sm = row < mrows ? (T)1 : (T)0;
      const T sn = row < nrows ? (T)1 : (T)0;
      As[row][c2] = ra[v].x * sm;
      As[row][c2 + 1] = ra[v].y * sm;
      Bs[row][c2] = rb[v].x * (sn * sg[0]);
      Bs[row][c2 + 1] = rb[v].y * (sn * sg[1]);
    }
  };

  const int steps = (qe - qb) * nchunk;
  if (steps > 0) {
    load(qb, 0);
    store();
  }
  __syncthreads();
  for (int step = 0; step < steps; ++step) {
    const int nxt = step + 1;
    if (nxt < steps) load(qb + nxt / nchunk, (nxt % nchunk) * kTile);
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
      const int k = kk + (lane >> 4);
      T af[MI], bf[NI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) af[mi] = As[wm * (TS / 2) + mi * 16 + (lane & 15)][k];
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) bf[ni] = Bs[wn * (TS / 2) + ni * 16 + (lane & 15)][k];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = Op<T>::mma(af[mi], bf[ni], acc[mi][ni]);
    }
    __syncthreads();
    if (nxt < steps) {
      store();
      __syncthreads();
    }
  }
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int col = wn * (TS / 2) + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * (TS / 2) + mi * 16 + Op<T>::row(lane, r);
        dst[row * TS + col] = acc[mi][ni][r];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_kernel(const T* __restrict__ x, int naux, int xrows, int ldo,
                                                         const double* __restrict__ signs, int nsplit, int npairs,
                                                         int run, T* __restrict__ partial) {
  __shared__ SyrkSmem<T> sm;
  int s, pair;
  syrk_coords(blockIdx.x, npairs, s, pair);
  int mb, nb;
  tile_of(pair, mb, nb);
  int cb, ce;
  slot_range(s, nsplit, naux, run, cb, ce);
  for (int c = cb; c < ce; ++c) {
    int qb, qe;
    slot_aux(c, nsplit, naux, run, qb, qe);
    syrk_tile<T>(x, qb, qe, xrows, ldo, signs, mb, nb,
                 partial + ((long long)c * npairs + pair) * (kSyrkTile * kSyrkTile), sm);
  }
}

// ---------------------------------------------- SYRK over direct LDS loads
//
// FP64 / FP32 SYRK whose operand rows go from global memory straight into
// LDS (global_load_lds_dwordx4, gfx950), through a ring of kDmaStages stages:
// the loads of the next two k steps are in flight while one step computes
// and the first operands of the following step are read, no registers
// stage them, and each k step has one barrier.
//
// A stage holds CH chunks of k of the 128 rows of both operands as
// 1 KB chunks (one load instruction of one wave; lane l fills bytes
// [16 l, 16 l + 16)): chunk = 16 rows x 4 groups of 16 bytes along k, lane l
// carrying row l % 16 and group l / 16. The MFMA side reads a chunk back
// with one 16-byte LDS read per lane, lane l again getting row l % 16, group
// l / 16 (a contiguous, conflict-free 1 KB read). Element e of the group
// feeds MFMA e: that MFMA sums over k = E (l / 16) + e, a permutation of the
// chunk's k that both operands share, so the E MFMAs together sum all of it.
//
// CH = 2 (32 KB stages, 128 KB ring): one workgroup per CU, so one wave per
// SIMD, which measured at most 86 % of the FP64 MFMA rate even without any
// memory traffic. CH = 1 (16 KB stages, 64 KB ring, no sign table): two
// workgroups per CU, two waves per SIMD that fill each other's barrier and
// LDS-read gaps (99 % in the same measurement).
constexpr int kDmaStages = 4;
constexpr int kDmaMaxLdo = 2048;

typedef __attribute__((address_space(3))) void lds_void_t;

template <typename T, int CH>
struct SyrkDma {
  static constexpr int E = 16 / (int)sizeof(T);       // elements per 16-byte group
  static constexpr int CHUNK = 1024 / (int)sizeof(T);  // elements per 1 KB chunk
  static constexpr int KC = 4 * E;                      // k per chunk
  static constexpr int KS = CH * KC;                    // k per stage (CH chunks per 16-row block)
  static constexpr int STAGE = 16 * CH * CHUNK;         // 2 operands x 8 row blocks x CH chunks
  static constexpr int LOADS = 4 * CH;                  // load instructions per wave per stage
  typedef T unit __attribute__((ext_vector_type(16 / sizeof(T))));
};

// s_waitcnt vmcnt(VM) lgkmcnt(0) (gfx9 encoding: vmcnt in bits 3:0 and 15:14).
template <int VM>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(VM >= 0 && VM < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((VM & 15) | ((VM >> 4) << 14) | (0x7 << 4));
}

// Workgroup barrier without a memory fence (a fence would wait for every
// load in flight, the stages ahead included). The clobber keeps the compiler
// from moving memory accesses across it.
__device__ __forceinline__ void barrier_nofence() { asm volatile("s_barrier" ::: "memory"); }

// The ring: acc += sum over `steps` k steps of (A rows) x (B rows)^T, the
// 128 rows of each operand read at step k from src(k, operand) plus this
// lane's byte offsets off[] (operand A = chunks of waves 0-1, B = waves 2-3).
// With SIGNED, the B operand's k values are scaled by sgn[sign_at(k) + k']
// times scale_at(k) (one factor per step). With MASKED, the MFMAs of this
// wave's 16 x 16 block (mi, ni) run only where bit mi of rbits and bit ni of
// cbits are set (wave-uniform masks; the skipped accumulators stay as they
// were).
template <typename T, bool SIGNED, int CH, bool MASKED = false, typename Src, typename SignAt, typename ScaleAt>
__device__ __forceinline__ void dma_ring(const Src& src, const int (&off)[4 * CH], int steps,
                                         const T* __restrict__ sgn, const SignAt& sign_at, const ScaleAt& scale_at,
                                         typename Op<T>::acc_t (&acc)[4][4], int rbits = 15, int cbits = 15) {
  typedef SyrkDma<T, CH> D;
  typedef typename D::unit unit;
  constexpr int E = D::E;
  constexpr int LOADS = D::LOADS;
  __shared__ __attribute__((aligned(16))) T st0[D::STAGE];
  __shared__ __attribute__((aligned(16))) T st1[D::STAGE];
  __shared__ __attribute__((aligned(16))) T st2[D::STAGE];
  __shared__ __attribute__((aligned(16))) T st3[D::STAGE];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int operand = wave >> 1;  // the chunks of waves 0-1 are A's, of waves 2-3 B's

  auto issue = [&](int step, T* stage) {
    const T* base = src(step, operand);
#pragma unroll
    for (int v = 0; v < LOADS; ++v)
      __builtin_amdgcn_global_load_lds((const void*)((const char*)base + (unsigned)off[v]),
                                       (lds_void_t*)(stage + (wave * LOADS + v) * D::CHUNK), 16, 0, 0);
  };

  // Operands of one chunk (16 rows x KC k per 16-row block) of a stage, with
  // the column signs applied to the B side.
  auto read_chunk = [&](const T* stage, int step, int c, unit (&a)[4], unit (&b)[4]) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) a[mi] = *(const unit*)(stage + ((((wm * 4 + mi) * CH + c) * 64 + lane) * E));
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
      b[ni] = *(const unit*)(stage + D::STAGE / 2 + ((((wn * 4 + ni) * CH + c) * 64 + lane) * E));
    if constexpr (SIGNED) {
      const unit sg = *(const unit*)(sgn + sign_at(step) + c * D::KC + (lane >> 4) * E) * scale_at(step);
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) b[ni] *= sg;
    }
  };
  auto mma_chunk = [&](const unit (&a)[4], const unit (&b)[4]) {
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
          if constexpr (MASKED) {
            if (!((rbits >> mi) & (cbits >> ni) & 1)) continue;
          }
          acc[mi][ni] = Op<T>::mma(a[mi][e], b[ni][e], acc[mi][ni]);
        }
  };

  // Step k: wait until this wave's loads of stage k + 1 have landed (those
  // of stage k + 2 stay in flight) and every LDS read issued so far has
  // returned; after the barrier every wave's loads of stages k and k + 1 are
  // in LDS and nobody reads stage k - 1 any more, so its buffer takes the
  // loads of step k + 3. The first chunk of stage k is already in registers
  // (read during step k - 1). With CH = 2 the step reads its second chunk,
  // runs the first chunk's MFMAs, reads the first chunk of stage k + 1, then
  // runs the second chunk's MFMAs; with CH = 1 it reads stage k + 1's chunk
  // into the other register set while the MFMAs of stage k run. Either way
  // no MFMA waits for an LDS read after the barrier.
  // Past the last step the loads repeat the last stage into buffers nobody
  // reads and the MFMAs are skipped, so there are only the seven load sites
  // below, each naming its stage buffer statically: the compiler then tells
  // the stages being read from the ones being loaded and adds no wait of its
  // own.
  unit a0[4], b0[4], a1[4], b1[4];
  // odd: which register set holds stage k's first chunk (CH = 1 alternates).
  auto step_body = [&](int k, const T* cur, const T* next, T* free, bool odd) {
    // Keep the previous step's MFMAs ahead of the wait: issued before it,
    // they keep the matrix pipe busy while the wave waits at the barrier.
    __builtin_amdgcn_sched_barrier(0);
    wait_vm_lgkm0<LOADS>();
    barrier_nofence();
    issue(min(k + 3, steps - 1), free);
    if (k < steps) {
      if constexpr (CH == 2) {
        read_chunk(cur, k, 1, a1, b1);
        mma_chunk(a0, b0);
        read_chunk(next, k + 1, 0, a0, b0);
        mma_chunk(a1, b1);
      } else if (!odd) {
        read_chunk(next, k + 1, 0, a1, b1);
        mma_chunk(a0, b0);
      } else {
        read_chunk(next, k + 1, 0, a0, b0);
        mma_chunk(a1, b1);
      }
    }
  };

  if (steps > 0) {
    issue(0, st0);
    issue(min(1, steps - 1), st1);
    issue(min(2, steps - 1), st2);
    wait_vm_lgkm0<2 * LOADS>();
    barrier_nofence();
    read_chunk(st0, 0, 0, a0, b0);
    for (int k = 0; k < steps; k += kDmaStages) {
      step_body(k, st0, st1, st3, false);
      step_body(k + 1, st1, st2, st0, true);
      step_body(k + 2, st2, st3, st1, false);
      step_body(k + 3, st3, st0, st2, true);
    }
    // No load may still be writing this workgroup's LDS when it returns.
    wait_vm_lgkm0<0>();
  }
}

// SIGNED = false: every sign is +1 on the occupied columns (the padded
// columns of X are zero), so the sign table is neither built nor applied.
template <typename T, bool SIGNED, int CH>
__global__ __launch_bounds__(kThreads, CH == 1 ? 2 : 1) void syrk_dma_kernel(
    const T* __restrict__ x, int naux, int xrows, int ldo, const double* __restrict__ signs,
    const double* __restrict__ auxsign, int shift, int nsplit, int npairs, int run, T* __restrict__ partial) {
  typedef SyrkDma<T, CH> D;
  constexpr int E = D::E;
  constexpr int LOADS = D::LOADS;
  __shared__ __attribute__((aligned(16))) T sgn[SIGNED ? kDmaMaxLdo : 1];

  int s, pair;
  syrk_coords(blockIdx.x, npairs, s, pair);
  int mb, nb;
  tile_of(pair, mb, nb);
  int cb, ce;
  slot_range(s, nsplit, naux, run, cb, ce);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;

  if (SIGNED)
    for (int i = t; i < ldo; i += kThreads) sgn[i] = (T)signs[i];

  // Byte offsets (within one aux function's X, from k = i0) of this lane's
  // group in the wave's LOADS chunks: chunks LOADS wave .. LOADS wave +
  // LOADS - 1 of the 16 CH of a stage (the first 8 CH operand A = row block
  // mb, the rest operand B = row block nb; chunk = CH * (16-row block) + k
  // part). Rows past xrows (the last 128-row block may hold only 64) read
  // row 0 instead: their products land in K rows and columns past nao, which
  // are never read.
  int off[LOADS];
#pragma unroll
  for (int v = 0; v < LOADS; ++v) {
    const int ch = wave * LOADS + v;
    const int blk = ch >= 8 * CH ? nb : mb;
    int row = blk * kSyrkTile + ((ch / CH) % 8) * 16 + (lane & 15);
    row = row < xrows ? row : 0;
    off[v] = (row * ldo + (ch % CH) * D::KC + (lane >> 4) * E) * (int)sizeof(T);
  }
  const int nch = ldo / D::KS;
  auto sign_at = [&](int step) { return (step % nch) * D::KS; };

  for (int c = cb; c < ce; ++c) {
    int qb, qe;
    slot_aux(c, nsplit, naux, run, qb, qe);
    const int steps = (qe - qb) * nch;
    // Step k covers k values i0 = (k % nch) KS of aux function qb + k / nch;
    // the second operand reads columns i0 + shift (mod ldo).
    auto src = [&](int step, int operand) -> const T* {
      const int qq = step / nch;
      int i0 = (step - qq * nch) * D::KS;
      if (operand) i0 = (i0 + shift) % ldo;
      return x + (long long)(qb + qq) * xrows * ldo + i0;
    };
    // Per-aux signs (the signature of an indefinite metric), when given.
    auto scale_at = [&](int step) -> T { return auxsign != nullptr ? (T)auxsign[qb + step / nch] : (T)1; };

    typename Op<T>::acc_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

    // The sign table is written, and every wave's LDS reads of the previous
    // slot's ring have returned (each ring ends waiting for them), so the
    // stage buffers may take this slot's loads.
    __syncthreads();
    dma_ring<T, SIGNED, CH>(src, off, steps, sgn, sign_at, scale_at, acc);

    T* dst = partial + ((long long)c * npairs + pair) * (kSyrkTile * kSyrkTile);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const int col = wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm * 64 + mi * 16 + Op<T>::row(lane, r);
          dst[row * kSyrkTile + col] = acc[mi][ni][r];
        }
      }
  }
}

// One 128 x 128 tile (mb, nb) of sum_{q in [qb, qe)} X_q X_q^T on the ring
// (all column signs +1: X's padded columns are zero), stored row-major into
// dst; the MFMAs of 16-row blocks at or past nrow (X's rows past nao) are
// skipped, and those rows and columns of dst hold zeros. The whole
// workgroup calls it; its LDS reads of an earlier call must have returned (a
// workgroup barrier between calls).
template <typename T, int CH>
__device__ __forceinline__ void syrk_ring_tile(const T* __restrict__ x, int xrows, int ldo, int mb, int nb, int qb,
                                               int qe, T* __restrict__ dst, int nrow) {
  typedef SyrkDma<T, CH> D;
  constexpr int E = D::E;
  constexpr int LOADS = D::LOADS;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  // Lane byte offsets within one aux function's X, as in syrk_dma_kernel;
  // rows past xrows read row 0 (their products land in rows never read).
  int off[LOADS];
#pragma unroll
  for (int v = 0; v < LOADS; ++v) {
    const int ch = wave * LOADS + v;
    const int blk = ch >= 8 * CH ? nb : mb;
    int row = blk * kSyrkTile + ((ch / CH) % 8) * 16 + (lane & 15);
    row = row < xrows ? row : 0;
    off[v] = (row * ldo + (ch % CH) * D::KC + (lane >> 4) * E) * (int)sizeof(T);
  }
  const int nch = ldo / D::KS;
  auto src = [&](int step, int operand) -> const T* {
    (void)operand;
    const int qq = step / nch;
    return x + (long long)(qb + qq) * xrows * ldo + (step - qq * nch) * D::KS;
  };
  auto sign_at = [&](int) { return 0; };
  auto scale_at = [&](int) -> T { return (T)1; };
  typename Op<T>::acc_t acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;
  // The 16-row blocks of each operand at or past nrow hold padding: their MFMAs are skipped.
  int rbits = 0, cbits = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    rbits |= (mb * kSyrkTile + wm * 64 + b * 16 < nrow) << b;
    cbits |= (nb * kSyrkTile + wn * 64 + b * 16 < nrow) << b;
  }
  dma_ring<T, false, CH, true>(src, off, (qe - qb) * nch, (const T*)nullptr, sign_at, scale_at, acc, rbits, cbits);
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * 64 + mi * 16 + Op<T>::row(lane, r);
        dst[row * kSyrkTile + col] = acc[mi][ni][r];
      }
    }
}

// FP64 / FP32 half-transform on the ring, from a full row-major copy of B
// (bfull[q][npad][npad], both triangles) and C^T (ct[ldo][npad], zero past
// nocc and past nao): X_q[m][i] = sum_k B_q[m][k] C^T[i][k], one 128 x 128
// tile (row block mb, column block nb) of one aux function per workgroup.
// Both operands are rows along k, as in the SYRK, so the ring is the same.
// X rows past npad (xrows may exceed it by up to 32) are written as zeros.
template <typename T, int CH>
__global__ __launch_bounds__(kThreads, CH == 1 ? 2 : 1) void xform_dma_kernel(
    const T* __restrict__ bfull, int npad, const T* __restrict__ ct, int ldo, T* __restrict__ x, int xrows,
    int qbase, int nblocks_n, int nblocks_m, int nwg) {
  typedef SyrkDma<T, CH> D;
  constexpr int E = D::E;
  constexpr int LOADS = D::LOADS;
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;

  // Lane offsets: operand A rows m of B_q (clamped below npad), operand B
  // rows i of C^T (clamped below ldo); both have row stride npad.
  int off[LOADS];
#pragma unroll
  for (int v = 0; v < LOADS; ++v) {
    const int ch = wave * LOADS + v;
    const bool is_b = ch >= 8 * CH;
    int row = (is_b ? nb : mb) * kSyrkTile + ((ch / CH) % 8) * 16 + (lane & 15);
    row = row < (is_b ? ldo : npad) ? row : 0;
    off[v] = (row * npad + (ch % CH) * D::KC + (lane >> 4) * E) * (int)sizeof(T);
  }
  const T* bq = bfull + (long long)q * npad * npad;
  auto src = [&](int step, int operand) -> const T* { return (operand ? ct : bq) + step * D::KS; };
  auto no_sign = [](int) { return 0; };
  auto no_scale = [](int) { return (T)1; };

  typename Op<T>::acc_t acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mi][ni][r] = (T)0;

  dma_ring<T, false, CH>(src, off, npad / D::KS, (const T*)nullptr, no_sign, no_scale, acc);

  T* xq = x + (long long)q * xrows * ldo;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = nb * kSyrkTile + wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mb * kSyrkTile + wm * 64 + mi * 16 + Op<T>::row(lane, r);
        if (row < xrows && col < ldo) xq[(long long)row * ldo + col] = row < npad ? acc[mi][ni][r] : (T)0;
      }
    }
}

// Full row-major copy of tiled packed B in T: out[q][r][c] = B_q[r][c] for
// r, c < npad (both triangles; zero past nao, as the packed padding). One
// workgroup per stored 32 x 32 tile (tr >= tc) of one aux function, in a
// grid-stride loop: the tile is read once, contiguously, written to block
// (tr, tc) and, through LDS, transposed into block (tc, tr), so every global
// access is a run of 32 consecutive elements.
template <typename T>
__global__ __launch_bounds__(kThreads) void expand_full_kernel(const double* __restrict__ packed, long long p,
                                                                int naux, int ntp, int npad, T* __restrict__ out) {
  __shared__ T tile[kTile][kTile + 1];
  const long long jobs = (long long)naux * ntp;
  for (long long job = blockIdx.x; job < jobs; job += gridDim.x) {
    const int q = (int)(job / ntp);
    const int t = (int)(job - (long long)q * ntp);
    int tr = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
    while (tr * (tr + 1) / 2 > t) --tr;
    const int tc = t - tr * (tr + 1) / 2;
    const double* src = packed + q * p + (long long)t * kTileElems;
    T* oq = out + (long long)q * npad * npad;
    for (int e = threadIdx.x; e < kTileElems; e += kThreads) {
      const int i = e / kTile, j = e % kTile;
      const T v = (T)src[e];
      tile[i][j] = v;
      oq[(long long)(tr * kTile + i) * npad + tc * kTile + j] = v;
    }
    if (tr != tc) {
      __syncthreads();
      for (int e = threadIdx.x; e < kTileElems; e += kThreads) {
        const int i = e / kTile, j = e % kTile;
        oq[(long long)(tc * kTile + i) * npad + tr * kTile + j] = tile[j][i];
      }
    }
    __syncthreads();
  }
}

// C^T in T: ct[i][k] = C[k][i] for k < nao, i < nocc, zero elsewhere
// ([ldo][npad]).
template <typename T>
__global__ __launch_bounds__(kThreads) void orbitals_t_kernel(const double* __restrict__ c, int nao, int nocc,
                                                               int npad, int ldo, T* __restrict__ ct) {
  const long long total = (long long)npad * ldo;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / npad);
    const int k = (int)(idx % npad);
    ct[idx] = (T)((k < nao && i < nocc) ? c[(long long)k * nocc + i] : 0.0);
  }
}

struct Syrk16Smem {
  uint16_t Ah[kSyrkTile][kRowK];
  uint16_t Al[kSyrkTile][kRowK];
  uint16_t Bh[kSyrkTile][kRowK];
  uint16_t Bl[kSyrkTile][kRowK];
};

// BF16x3: one 128 x 128 FP32 tile of sum_{q in [qb, qe)} X_q S X_q^T.
__device__ __forceinline__ void syrk16_tile(const uint16_t* __restrict__ xhi, const uint16_t* __restrict__ xlo,
                                            int qb, int qe, int xrows, int ldo, const double* __restrict__ signs,
                                            int mb, int nb, float* __restrict__ dst, Syrk16Smem& sm) {
  constexpr int MI = 4, NI = 4;
  auto& Ah = sm.Ah;
  auto& Al = sm.Al;
  auto& Bh = sm.Bh;
  auto& Bl = sm.Bl;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  f4 acc[MI][NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  // 128 rows x 32 bf16 per plane = 512 16-byte chunks: 2 per thread per plane.
  // As in the FP64/FP32 SYRK, the prefetch only copies: row masks and sign
  // flips are applied in store(), at the end of the k step.
  uint4 rah[2], ral[2], rbh[2], rbl[2];
  double rsg[8];
  const int nchunk = ldo / kTile;
  const int mrows = min(kSyrkTile, xrows - mb * kSyrkTile);
  const int nrows = min(kSyrkTile, xrows - nb * kSyrkTile);
  // Every chunk of a thread has the same column block c8 = (t & 3) * 8.
  const int c8 = (t & 3) * 8;
  auto load = [&](int q, int i0) {
    const long long om = ((long long)q * xrows + mb * kSyrkTile) * ldo + i0;
    const long long on = ((long long)q * xrows + nb * kSyrkTile) * ldo + i0;
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int row = (t + kThreads * v) >> 2;
      const long long am = om + (long long)(row < mrows ? row : 0) * ldo + c8;
      const long long an = on + (long long)(row < nrows ? row : 0) * ldo + c8;
      rah[v] = *(const uint4*)(xhi + am);
      ral[v] = *(const uint4*)(xlo + am);
      rbh[v] = *(const uint4*)(xhi + an);
      rbl[v] = *(const uint4*)(xlo + an);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) rsg[e] = signs[i0 + c8 + e];
  };
  auto store = [&]() {
    // A negative sign flips the sign bit of the element in both planes.
    const uint32_t f0 = (rsg[0] < 0.0 ? 0x8000u : 0u) | (rsg[1] < 0.0 ? 0x80000000u : 0u);
    const uint32_t f1 = (rsg[2] < 0.0 ? 0x8000u : 0u) | (rsg[3] < 0.0 ? 0x80000000u : 0u);
    const uint32_t f2 = (rsg[4] < 0.0 ? 0x8000u : 0u) | (rsg[5] < 0.0 ? 0x80000000u : 0u);
    const uint32_t f3 = (rsg[6] < 0.0 ? 0x8000u : 0u) | (rsg[7] < 0.0 ? 0x80000000u : 0u);
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int row = (t + kThreads * v) >> 2;
      const bool okm = row < mrows;
      const bool okn = row < nrows;
      uint4 h = rbh[v], l = rbl[v];
      h.x ^= f0;
      h.y ^= f1;
      h.z ^= f2;
      h.w ^= f3;
      l.x ^= f0;
      l.y ^= f1;
      l.z ^= f2;
      l.w ^= f3;
      *(uint4*)&Ah[row][c8] = mask4(okm, rah[v]);
      *(uint4*)&Al[row][c8] = mask4(okm, ral[v]);
      *(uint4*)&Bh[row][c8] = mask4(okn, h);
      *(uint4*)&Bl[row][c8] = mask4(okn, l);
    }
  };

  const int steps = (qe - qb) * nchunk;
  if (steps > 0) {
    load(qb, 0);
    store();
  }
  __syncthreads();
  for (int step = 0; step < steps; ++step) {
    const int nxt = step + 1;
    if (nxt < steps) load(qb + nxt / nchunk, (nxt % nchunk) * kTile);
    const int kq = 8 * (lane >> 4);
    bf8 ah[MI], al[MI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
      ah[mi] = lds_bf8(&Ah[wm * 64 + mi * 16 + (lane & 15)][kq]);
      al[mi] = lds_bf8(&Al[wm * 64 + mi * 16 + (lane & 15)][kq]);
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const bf8 bh = lds_bf8(&Bh[wn * 64 + ni * 16 + (lane & 15)][kq]);
      const bf8 bl = lds_bf8(&Bl[wn * 64 + ni * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) acc[mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mi][ni]);
    }
    __syncthreads();
    if (nxt < steps) {
      store();
      __syncthreads();
    }
  }
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int col = wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * 64 + mi * 16 + 4 * (lane >> 4) + r;
        dst[row * kSyrkTile + col] = acc[mi][ni][r];
      }
    }
}

__global__ __launch_bounds__(kThreads) void syrk_bf16x3_kernel(const uint16_t* __restrict__ xhi,
                                                                const uint16_t* __restrict__ xlo, int naux,
                                                                int xrows, int ldo, const double* __restrict__ signs,
                                                                int nsplit, int npairs, int run,
                                                                float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) Syrk16Smem sm;
  int s, pair;
  syrk_coords(blockIdx.x, npairs, s, pair);
  int mb, nb;
  tile_of(pair, mb, nb);
  int cb, ce;
  slot_range(s, nsplit, naux, run, cb, ce);
  for (int c = cb; c < ce; ++c) {
    int qb, qe;
    slot_aux(c, nsplit, naux, run, qb, qe);
    syrk16_tile(xhi, xlo, qb, qe, xrows, ldo, signs, mb, nb,
                partial + ((long long)c * npairs + pair) * (kSyrkTile * kSyrkTile), sm);
  }
}

// ------------------------------------------------------------------ BF16x3 on the direct-LDS ring
//
// The bf16 form of the ring: a 1 KB chunk (16 rows x 32 k of bf16, lane l
// carrying row l % 16 and k 8 (l / 16) .. + 8) is exactly one operand of
// v_mfma_f32_16x16x32_bf16, read back by lane l as one 16-byte LDS read. A
// stage holds 32 chunks: for each operand (A = the 128 rows of the first
// block, B = those of the second) its 8 row blocks in the hi and the lo
// plane, 32 KB; four stages fill 128 KB, one workgroup per CU. Wave w loads
// chunks 8 w .. 8 w + 7 (chunk = 16 operand + 2 block + plane), so waves
// 0-1 load A and waves 2-3 B. Each k step of 32 runs 48 MFMAs per wave
// (hi.hi + hi.lo + lo.hi on 4 x 4 output blocks) and reads the next stage's
// operands into the second register set meanwhile; the barrier and wait
// structure is the FP64/FP32 ring's (one barrier per step, the loads of
// two steps ahead in flight).
constexpr int kRing16Loads = 8;              // load instructions per wave per stage
constexpr int kRing16Stage = 32 * 512;       // uint16 elements per stage (32 KB)

struct Ops16 {
  bf8 ah[4], al[4], bh[4], bl[4];
};

// src(step, operand, plane): the uint16 base of operand `operand`'s `plane`
// at step `step`, to which this lane's byte offsets off[v] (row and k part)
// are added. With SIGNED, the B operand's elements are xored with the
// 16-bit masks sgn[sign_at(step) + k'] (0x8000 flips the sign).
template <bool SIGNED, typename Src, typename SignAt>
__device__ __forceinline__ void dma_ring16(const Src& src, const int (&off)[kRing16Loads], int steps,
                                           const uint16_t* __restrict__ sgn, const SignAt& sign_at,
                                           f4 (&acc)[4][4]) {
  __shared__ __attribute__((aligned(16))) uint16_t st0[kRing16Stage];
  __shared__ __attribute__((aligned(16))) uint16_t st1[kRing16Stage];
  __shared__ __attribute__((aligned(16))) uint16_t st2[kRing16Stage];
  __shared__ __attribute__((aligned(16))) uint16_t st3[kRing16Stage];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int operand = wave >> 1;

  auto issue = [&](int step, uint16_t* stage) {
    const uint16_t* hi = src(step, operand, 0);
    const uint16_t* lo = src(step, operand, 1);
#pragma unroll
    for (int v = 0; v < kRing16Loads; ++v) {
      const uint16_t* base = (v & 1) ? lo : hi;
      __builtin_amdgcn_global_load_lds((const void*)((const char*)base + (unsigned)off[v]),
                                       (lds_void_t*)(stage + (wave * kRing16Loads + v) * 512), 16, 0, 0);
    }
  };
  auto chunk = [&](const uint16_t* stage, int op, int blk, int plane) -> bf8 {
    return lds_bf8(stage + ((op * 16 + blk * 2 + plane) * 512 + lane * 8));
  };
  auto read = [&](const uint16_t* stage, int step, Ops16& o) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      o.ah[mi] = chunk(stage, 0, wm * 4 + mi, 0);
      o.al[mi] = chunk(stage, 0, wm * 4 + mi, 1);
    }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      o.bh[ni] = chunk(stage, 1, wn * 4 + ni, 0);
      o.bl[ni] = chunk(stage, 1, wn * 4 + ni, 1);
    }
    if constexpr (SIGNED) {
      const uint4 m = *(const uint4*)(sgn + sign_at(step) + (lane >> 4) * 8);
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        uint4 h = __builtin_bit_cast(uint4, o.bh[ni]);
        uint4 l = __builtin_bit_cast(uint4, o.bl[ni]);
        h.x ^= m.x; h.y ^= m.y; h.z ^= m.z; h.w ^= m.w;
        l.x ^= m.x; l.y ^= m.y; l.z ^= m.z; l.w ^= m.w;
        o.bh[ni] = __builtin_bit_cast(bf8, h);
        o.bl[ni] = __builtin_bit_cast(bf8, l);
      }
    }
  };
  auto mma = [&](const Ops16& o) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = mma3(o.ah[mi], o.al[mi], o.bh[ni], o.bl[ni], acc[mi][ni]);
  };

  Ops16 r0, r1;
  auto step_body = [&](int k, const uint16_t* next, uint16_t* free, bool odd) {
    __builtin_amdgcn_sched_barrier(0);
    wait_vm_lgkm0<kRing16Loads>();
    barrier_nofence();
    issue(min(k + 3, steps - 1), free);
    if (k < steps) {
      if (!odd) {
        read(next, k + 1, r1);
        mma(r0);
      } else {
        read(next, k + 1, r0);
        mma(r1);
      }
    }
  };
  if (steps > 0) {
    issue(0, st0);
    issue(min(1, steps - 1), st1);
    issue(min(2, steps - 1), st2);
    wait_vm_lgkm0<2 * kRing16Loads>();
    barrier_nofence();
    read(st0, 0, r0);
    for (int k = 0; k < steps; k += 4) {
      step_body(k, st1, st3, false);
      step_body(k + 1, st2, st0, true);
      step_body(k + 2, st3, st1, false);
      step_body(k + 3, st0, st2, true);
    }
    wait_vm_lgkm0<0>();
  }
}

// BF16x3 SYRK over the ring: the tile (mb, nb) of sum_q X_q S X_q^T from the
// hi/lo planes of X ([naux][xrows][ldo] each), partial slots as in the FP32
// SYRK (one per run of `run` aux functions). With SIGNED the sign masks of
// the columns sit in LDS. `shift` rotates the second operand's columns by
// that many (a multiple of 32): sum_k X[m][k] S_k X[n][(k + shift) mod ldo].
template <bool SIGNED>
__global__ __launch_bounds__(kThreads, 1) void syrk16_dma_kernel(const uint16_t* __restrict__ xhi,
                                                                 const uint16_t* __restrict__ xlo, int naux, int xrows,
                                                                 int ldo, const double* __restrict__ signs, int shift,
                                                                 int nsplit, int npairs, int run,
                                                                 float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) uint16_t sgn[SIGNED ? kDmaMaxLdo : 8];
  int s, pair;
  syrk_coords(blockIdx.x, npairs, s, pair);
  int mb, nb;
  tile_of(pair, mb, nb);
  int cb, ce;
  slot_range(s, nsplit, naux, run, cb, ce);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  if (SIGNED)
    for (int i = t; i < ldo; i += kThreads) sgn[i] = signs[i] < 0.0 ? (uint16_t)0x8000 : (uint16_t)0;

  // Lane byte offsets within one aux function's plane: load v of this wave is
  // chunk 8 wave + v = 16 operand + 2 block + plane.
  int off[kRing16Loads];
#pragma unroll
  for (int v = 0; v < kRing16Loads; ++v) {
    const int ch = wave * kRing16Loads + v;
    const int blk = (ch >> 1) & 7;
    int row = ((ch >> 4) ? nb : mb) * kSyrkTile + blk * 16 + (lane & 15);
    row = row < xrows ? row : 0;
    off[v] = (row * ldo + (lane >> 4) * 8) * 2;
  }
  const int nch = ldo / 32;
  auto sign_at = [&](int step) { return (step % nch) * 32; };
  for (int c = cb; c < ce; ++c) {
    int qb, qe;
    slot_aux(c, nsplit, naux, run, qb, qe);
    auto src = [&](int step, int operand, int plane) -> const uint16_t* {
      const int qq = step / nch;
      int i0 = (step - qq * nch) * 32;
      if (operand) i0 = (i0 + shift) % ldo;
      return (plane ? xlo : xhi) + (long long)(qb + qq) * xrows * ldo + i0;
    };
    f4 acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
    __syncthreads();  // sign masks written; the previous slot's LDS reads have returned
    dma_ring16<SIGNED>(src, off, (qe - qb) * nch, sgn, sign_at, acc);
    float* dst = partial + ((long long)c * npairs + pair) * (kSyrkTile * kSyrkTile);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const int col = wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[(wm * 64 + mi * 16 + 4 * (lane >> 4) + r) * kSyrkTile + col] = acc[mi][ni][r];
      }
  }
}

// One 128 x 128 tile (mb, nb) of sum_{q in [qb, qe)} X_q X_q^T in BF16x3 on
// the ring (all column signs +1) from X's hi/lo planes, stored row-major
// into dst. Called by the whole workgroup, a barrier between calls.
__device__ __forceinline__ void syrk16_ring_tile(const uint16_t* __restrict__ xhi, const uint16_t* __restrict__ xlo,
                                                 int xrows, int ldo, int mb, int nb, int qb, int qe,
                                                 float* __restrict__ dst) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  int off[kRing16Loads];
#pragma unroll
  for (int v = 0; v < kRing16Loads; ++v) {
    const int ch = wave * kRing16Loads + v;
    const int blk = (ch >> 1) & 7;
    int row = ((ch >> 4) ? nb : mb) * kSyrkTile + blk * 16 + (lane & 15);
    row = row < xrows ? row : 0;
    off[v] = (row * ldo + (lane >> 4) * 8) * 2;
  }
  const int nch = ldo / 32;
  auto src = [&](int step, int operand, int plane) -> const uint16_t* {
    (void)operand;
    const int qq = step / nch;
    return (plane ? xlo : xhi) + (long long)(qb + qq) * xrows * ldo + (step - qq * nch) * 32;
  };
  auto sign_at = [&](int) { return 0; };
  f4 acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
  dma_ring16<false>(src, off, (qe - qb) * nch, (const uint16_t*)nullptr, sign_at, acc);
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(wm * 64 + mi * 16 + 4 * (lane >> 4) + r) * kSyrkTile + col] = acc[mi][ni][r];
    }
}

// The MFMA core of the BF16x3 kernels with 256 x 256 tiles (eight waves):
// acc += A B^T over `steps` steps of 32 k, A and B each 256 rows of bf16 hi
// and lo planes. Wave w computes rows 128 (w & 1) .. + 128 of A against rows
// 64 (w >> 1) .. + 64 of B (8 x 4 blocks of 16 x 16, 128 FP32 accumulators
// per lane). A stage holds one step of both operands in 1 KB chunks of 16
// rows x 32 k (chunk 32 operand + 2 block + plane, 64 KB); two stages
// alternate, the loads of step k + 1 in flight during step k's MFMAs, so a
// step moves half the bytes per MFMA of the 128 x 128 ring. Waves 0-3 load
// A's chunks, waves 4-7 B's: addr(step, v) is this lane's 16-byte source of
// load v of its wave (chunk 8 wave + v) at that step. The stages are two
// separate LDS arrays, chosen statically (the loop is unrolled by two), which
// lets the compiler see that a step's LDS reads do not touch the stage the
// loads in flight write, so it does not wait for those loads before the
// reads. Called by the whole workgroup; it returns with every load and LDS
// read complete, so the stages may take new loads after the next barrier.
constexpr int kR256Threads = 512;
constexpr int kR256Loads = 8;          // load instructions per wave per stage
constexpr int kR256Stage = 64 * 512;   // uint16 elements per stage (64 KB)

template <bool ILV = false, class Addr>
__device__ __forceinline__ void ring256(const Addr& addr, int steps, f4 (&acc)[8][4]) {
  __shared__ __attribute__((aligned(16))) uint16_t st0[kR256Stage];
  __shared__ __attribute__((aligned(16))) uint16_t st1[kR256Stage];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;   // row half of A
  const int wn = wave >> 1;  // row quarter of B
  // ILV: chunk 2 block + h of an operand holds rows 8 h .. + 8 of the 16-row block, each row as
  // 8 segments of 16 bytes (logical segment 4 plane + k group, stored at segment logical ^ row),
  // so that the MFMA operand reads of 8 consecutive rows fall in distinct banks.
  const int r8 = lane & 7;
  const int ilv_off = ((lane >> 3) & 1) * 512 + r8 * 64;
  auto chunk = [&](const uint16_t* stage, int op, int blk, int plane) -> bf8 {
    if constexpr (ILV)
      return lds_bf8(stage + ((op * 32 + blk * 2) * 512 + ilv_off + ((plane * 4 + (lane >> 4)) ^ r8) * 8));
    else
      return lds_bf8(stage + ((op * 32 + blk * 2 + plane) * 512 + lane * 8));
  };
  // The wave's 8 row blocks of A in two halves of 4 (operand registers for 4 A blocks at a time).
  auto compute = [&](const uint16_t* stage) {
#pragma unroll
    for (int mh = 0; mh < 2; ++mh) {
      bf8 ah[4], al[4];
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        ah[mi] = chunk(stage, 0, wm * 8 + mh * 4 + mi, 0);
        al[mi] = chunk(stage, 0, wm * 8 + mh * 4 + mi, 1);
      }
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const bf8 bh = chunk(stage, 1, wn * 4 + ni, 0);
        const bf8 bl = chunk(stage, 1, wn * 4 + ni, 1);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) acc[mh * 4 + mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mh * 4 + mi][ni]);
      }
    }
  };
  auto issue = [&](int step, uint16_t* stage) {
#pragma unroll
    for (int v = 0; v < kR256Loads; ++v)
      __builtin_amdgcn_global_load_lds((const void*)addr(step, v),
                                       (lds_void_t*)(stage + (wave * kR256Loads + v) * 512), 16, 0, 0);
  };
  if (steps <= 0) return;
  issue(0, st0);
  wait_vm_lgkm0<0>();
  barrier_nofence();
  for (int k = 0; k < steps; k += 2) {
    // Past the last step the loads repeat it into the stage no later step reads.
    issue(min(k + 1, steps - 1), st1);
    __builtin_amdgcn_sched_barrier(0);  // the loads go out before the step's MFMAs
    compute(st0);
    __builtin_amdgcn_sched_barrier(0);
    wait_vm_lgkm0<0>();
    barrier_nofence();
    if (k + 1 < steps) {
      issue(min(k + 2, steps - 1), st0);
      __builtin_amdgcn_sched_barrier(0);
      compute(st1);
      __builtin_amdgcn_sched_barrier(0);
      wait_vm_lgkm0<0>();
      barrier_nofence();
    }
  }
}

// BF16x3 SYRK with 256 x 256 tiles: the four 128 x 128 quadrants of a
// 256-row tile pair in one workgroup (ring256). XILV: X interleaved (hi and
// lo per 32-column block, as the 256 x 256 half-transform writes it), loaded
// as whole 128-byte lines into the ring's interleaved chunks (7.5 ms against
// 8.7 from the hi/lo planes at nao 1500 / naux 5000 / nocc 250). Partial slots as in
// syrk16_dma_kernel: the tile's quadrants are written as the partial tiles
// of their 128-row tile pairs (the upper quadrant of a diagonal pair, and
// quadrants past the last 128-row block, are not written). Column signs +1
// (the occupied orbitals).
template <bool XILV>
__global__ __launch_bounds__(kR256Threads, 1) void syrk16_256_kernel(const uint16_t* __restrict__ xhi,
                                                                     const uint16_t* __restrict__ xlo, int naux,
                                                                     int xrows, int ldo, int nsplit, int npairs,
                                                                     int npairs128, int nb128, int run,
                                                                     float* __restrict__ partial) {
  int s, pair;
  syrk_coords(blockIdx.x, npairs, s, pair);
  int mb, nb;
  tile_of(pair, mb, nb);
  int cb, ce;
  slot_range(s, nsplit, naux, run, cb, ce);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  // Load v of this wave is chunk 8 wave + v = 32 operand + 2 block + plane (waves 0-3 operand A,
  // the rows of block mb; waves 4-7 operand B, block nb). Rows past xrows read row 0: their
  // products land in rows and columns of K past nao, never read.
  // XILV (X interleaved, hi and lo per 32-column block): load v fills the ring's interleaved
  // chunk 8 wave + v, rows 8 (chunk & 31) .. + 8, lane l the row's segment (l & 7) ^ row.
  int off[kR256Loads];
#pragma unroll
  for (int v = 0; v < kR256Loads; ++v) {
    const int ch = wave * kR256Loads + v;
    if constexpr (XILV) {
      const int r = lane >> 3;
      int row = ((ch >> 5) ? nb : mb) * 256 + (ch & 31) * 8 + r;
      row = row < xrows ? row : 0;
      off[v] = (row * 2 * ldo + ((lane & 7) ^ r) * 8) * 2;
    } else {
      const int blk = (ch >> 1) & 15;
      int row = ((ch >> 5) ? nb : mb) * 256 + blk * 16 + (lane & 15);
      row = row < xrows ? row : 0;
      off[v] = (row * ldo + (lane >> 4) * 8) * 2;
    }
  }
  const int nch = ldo / 32;
  // The partial tile this wave writes into: 128-row tile pair (2 mb + wm, 2 nb + wn / 2), columns
  // 64 (wn & 1) .. + 64 of it.
  const int qr = 2 * mb + wm, qc = 2 * nb + (wn >> 1);
  const bool writes = qr < nb128 && qc < nb128 && qr >= qc;
  for (int c = cb; c < ce; ++c) {
    int qb, qe;
    slot_aux(c, nsplit, naux, run, qb, qe);
    auto addr = [&](int step, int v) -> const char* {
      const int qq = step / nch;
      if constexpr (XILV) {
        const long long base = (long long)(qb + qq) * xrows * 2 * ldo + (step - qq * nch) * 64;
        return (const char*)(xhi + base) + (unsigned)off[v];
      } else {
        const long long base = (long long)(qb + qq) * xrows * ldo + (step - qq * nch) * 32;
        return (const char*)(((v & 1) ? xlo : xhi) + base) + (unsigned)off[v];
      }
    };
    f4 acc[8][4];
#pragma unroll
    for (int mi = 0; mi < 8; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
    // The previous slot's LDS reads and loads have all returned (ring256 ends on a wait for
    // both), so the stages may take this slot's loads once every wave is here.
    __syncthreads();
    ring256<XILV>(addr, (qe - qb) * nch, acc);
    if (writes) {
      // The lane's offset is opaque here, so its 32 store addresses are formed per slot rather than
      // hoisted out of the slot loop (where they would outlive the accumulators' registers).
      int at = (4 * (lane >> 4)) * kSyrkTile + (lane & 15);
      asm volatile("" : "+v"(at));
      float* dst = partial + ((long long)c * npairs128 + tidx(qr, qc)) * (kSyrkTile * kSyrkTile) + (wn & 1) * 64 + at;
#pragma unroll
      for (int mi = 0; mi < 8; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) dst[(mi * 16 + r) * kSyrkTile + ni * 16] = acc[mi][ni][r];
    }
  }
}

// BF16x3 half-transform with 256 x 256 tiles (ring256 with its interleaved
// chunk layout) from the bf16 full copy of B and C^T, both in the
// interleaved layout of xform16_dma_kernel: tile (mb, nb) of X_q into X's
// hi/lo planes, one tile per workgroup. Each load reads whole 128-byte lines
// (8 rows x 32 k, hi and lo): 15.9 ms at nao 1500 / naux 5000 / nocc 250,
// against 17.9 with separate planes (every load then used half of each line
// it touched) and 17.6 with interleaved planes loaded as 16-row half-line
// chunks. Rows of B past npad and of C^T past ldo read row 0; X rows past
// npad (up to xrows) are written as zeros.
template <bool XILV>
__global__ __launch_bounds__(kR256Threads, 1) void xform16_256_kernel(
    const uint16_t* __restrict__ b, int npad, const uint16_t* __restrict__ ct, int ldo, uint16_t* __restrict__ xhi,
    uint16_t* __restrict__ xlo, int xrows, int qbase, int nblocks_n, int nblocks_m, int nwg) {
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  // Waves 0-3 load B_q's rows of block mb, waves 4-7 C^T's rows of block nb.
  const bool is_c = wave >= 4;
  // Load v of this wave fills chunk 8 wave + v: rows 8 (chunk & 31) .. + 8 of its operand, lane l
  // the row's 16-byte segment l & 7 (the ring's interleaved chunk layout), i.e. logical segment
  // (l & 7) ^ row of the row's 128-byte line of the step (32 hi, then 32 lo).
  int off[kR256Loads];
#pragma unroll
  for (int v = 0; v < kR256Loads; ++v) {
    const int c32 = (wave * kR256Loads + v) & 31;
    const int r = lane >> 3;
    int row = (is_c ? nb : mb) * 256 + c32 * 8 + r;
    row = row < (is_c ? ldo : npad) ? row : 0;
    off[v] = (row * 2 * npad + ((lane & 7) ^ r) * 8) * 2;
  }
  const uint16_t* src = is_c ? ct : b + (long long)q * 2 * npad * npad;
  auto addr = [&](int step, int v) -> const char* { return (const char*)(src + step * 64) + (unsigned)off[v]; };
  f4 acc[8][4];
#pragma unroll
  for (int mi = 0; mi < 8; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
  ring256<true>(addr, npad / 32, acc);
  // XILV: X_q as [xrows][ldo / 32][64] (hi then lo per 32-column block) from xhi; else the hi
  // plane at xhi, the lo plane at xlo.
  const int ldx = XILV ? 2 * ldo : ldo;                                    // row stride
  auto col_at = [&](int col) { return XILV ? (col >> 5) * 64 + (col & 31) : col; };
  const long long lo_at = XILV ? 32 : (long long)(xlo - xhi);                // lo part from hi
  const long long xbase = (long long)q * xrows * ldx;
  if ((mb + 1) * 256 <= npad && (nb + 1) * 256 <= ldo) {
    // A tile inside X (and inside B's rows): no bounds to check. Column wn 64 + ni 16 + lane & 15
    // stays within one 32-column block for each ni.
    uint16_t* th = xhi + xbase + (long long)(mb * 256 + wm * 128 + 4 * (lane >> 4)) * ldx +
                   col_at(nb * 256 + wn * 64 + (lane & 15));
#pragma unroll
    for (int mi = 0; mi < 8; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float v = acc[mi][ni][r];
          const uint16_t h = bf16_of(v);
          const long long o = (mi * 16 + r) * ldx + (XILV ? (ni >> 1) * 64 + (ni & 1) * 16 : ni * 16);
          th[o] = h;
          th[o + lo_at] = bf16_of(v - f32_of(h));
        }
  } else {
#pragma unroll
    for (int mi = 0; mi < 8; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const int col = nb * 256 + wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mb * 256 + wm * 128 + mi * 16 + 4 * (lane >> 4) + r;
          if (row < xrows && col < ldo) {
            const float v = row < npad ? acc[mi][ni][r] : 0.f;
            const uint16_t h = bf16_of(v);
            const long long o = xbase + (long long)row * ldx + col_at(col);
            xhi[o] = h;
            xhi[o + lo_at] = bf16_of(v - f32_of(h));
          }
        }
      }
  }
}

// BF16x3 half-transform over the ring from the bf16 full copy of B (b:
// [naux][npad][npad / 32][64], each 32-column block as 32 hi then the same
// 32 lo) and C^T in the same interleaved layout (ct: [ldo][npad / 32][64]):
// the 128 x 128 tile (mb, nb) of X_q into X's hi/lo planes. A stage's hi and
// lo chunks of a 16-row block come from the two halves of the same 128-byte
// lines.
__global__ __launch_bounds__(kThreads, 1) void xform16_dma_kernel(const uint16_t* __restrict__ b, int npad,
                                                                  const uint16_t* __restrict__ ct, int ldo,
                                                                  uint16_t* __restrict__ xhi, uint16_t* __restrict__ xlo,
                                                                  int xrows, int qbase, int nblocks_n, int nblocks_m,
                                                                  int nwg) {
  const int logical = xcd_remap(blockIdx.x, nwg);
  const int nb = logical % nblocks_n;
  const int mb = (logical / nblocks_n) % nblocks_m;
  const int q = qbase + logical / (nblocks_n * nblocks_m);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1;
  const int wn = wave >> 1;
  int off[kRing16Loads];
#pragma unroll
  for (int v = 0; v < kRing16Loads; ++v) {
    const int ch = wave * kRing16Loads + v;
    const bool is_b = (ch >> 4) != 0;
    const int blk = (ch >> 1) & 7;
    int row = (is_b ? nb : mb) * kSyrkTile + blk * 16 + (lane & 15);
    row = row < (is_b ? ldo : npad) ? row : 0;
    off[v] = (row * 2 * npad + (lane >> 4) * 8) * 2;
  }
  const long long plane = (long long)q * 2 * npad * npad;
  auto src = [&](int step, int operand, int pl) -> const uint16_t* {
    return (operand ? ct : b + plane) + step * 64 + pl * 32;
  };
  auto no_sign = [](int) { return 0; };
  f4 acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};
  dma_ring16<false>(src, off, npad / 32, (const uint16_t*)nullptr, no_sign, acc);
  const long long xbase = (long long)q * xrows * ldo;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int col = nb * kSyrkTile + wn * 64 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mb * kSyrkTile + wm * 64 + mi * 16 + 4 * (lane >> 4) + r;
        if (row < xrows && col < ldo) {
          const float v = row < npad ? acc[mi][ni][r] : 0.f;
          const uint16_t h = bf16_of(v);
          xhi[xbase + (long long)row * ldo + col] = h;
          xlo[xbase + (long long)row * ldo + col] = bf16_of(v - f32_of(h));
        }
      }
    }
}

// The full bf16 hi/lo copy of the packed FP64 tensor (hi plane [naux][npad][npad], then the lo
// plane): the element split of launch_split (hi = bf16(fp32(b)), lo = bf16(fp32(b) - hi)), tiles
// transposed through LDS as in expand_full_kernel.
__global__ __launch_bounds__(kThreads) void expand_full16_kernel(const double* __restrict__ packed, long long p,
                                                                 int naux, int ntp, int npad,
                                                                 uint16_t* __restrict__ out) {
  __shared__ uint32_t tile[kTile][kTile + 1];
  const long long jobs = (long long)naux * ntp;
  for (long long job = blockIdx.x; job < jobs; job += gridDim.x) {
    const int q = (int)(job / ntp);
    const int t = (int)(job - (long long)q * ntp);
    int tr = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
    while (tr * (tr + 1) / 2 > t) --tr;
    const int tc = t - tr * (tr + 1) / 2;
    const double* src = packed + q * p + (long long)t * kTileElems;
    const long long oq = (long long)q * 2 * npad * npad;
    for (int e = threadIdx.x; e < kTileElems; e += kThreads) {
      const int i = e / kTile, j = e % kTile;
      const float x = (float)src[e];
      const uint16_t h = bf16_of(x);
      const uint16_t l = bf16_of(x - f32_of(h));
      tile[i][j] = (uint32_t)h | ((uint32_t)l << 16);
      // Row tr 32 + i, column block tc: hi at 64 tc + j, lo 32 further.
      const long long o = oq + (long long)(tr * kTile + i) * 2 * npad + tc * 2 * kTile + j;
      out[o] = h;
      out[o + kTile] = l;
    }
    if (tr != tc) {
      __syncthreads();
      for (int e = threadIdx.x; e < kTileElems; e += kThreads) {
        const int i = e / kTile, j = e % kTile;
        const uint32_t v = tile[j][i];
        const long long o = oq + (long long)(tc * kTile + i) * 2 * npad + tr * 2 * kTile + j;
        out[o] = (uint16_t)(v & 0xFFFFu);
        out[o + kTile] = (uint16_t)(v >> 16);
      }
    }
    __syncthreads();
  }
}

// K from the partial tiles: one workgroup per (lower tile pair, 8-row strip
// of the tile); each thread sums 4 elements of the strip over the slots in
// a fixed order (reads coalesced along the tile's rows), writes them to
// K[bm rows][bn columns], and for an off-diagonal pair the mirror
// K[bn rows][bm columns] through an LDS transpose (writes coalesced too).
template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_reduce_kernel(const T* __restrict__ partial, int nslot,
                                                                int npairs, int nao, double* __restrict__ out) {
  constexpr int kStrip = 8;
  __shared__ double mirror[kStrip][kSyrkTile + 1];
  const int pair = blockIdx.x / (kSyrkTile / kStrip);
  const int strip = blockIdx.x % (kSyrkTile / kStrip);
  int bm, bn;
  tile_of(pair, bm, bn);
  const int t = threadIdx.x;
  const int col = t & (kSyrkTile - 1);  // 0..127
  const int r0 = t >> 7;                 // 0..1
#pragma unroll
  for (int rr = 0; rr < kStrip / 2; ++rr) {
    const int lr = r0 + 2 * rr;              // row within the strip
    const int trow = strip * kStrip + lr;    // row within the tile
    const T* src = partial + (long long)pair * (kSyrkTile * kSyrkTile) + trow * kSyrkTile + col;
    double sum = 0.0;
    for (int k = 0; k < nslot; ++k) sum += (double)src[(long long)k * npairs * (kSyrkTile * kSyrkTile)];
    const int row = bm * kSyrkTile + trow, c = bn * kSyrkTile + col;
    if (row < nao && c < nao) out[(long long)row * nao + c] = sum;
    mirror[lr][col] = sum;
  }
  if (bm == bn) return;
  __syncthreads();
  // Mirror: K[bn * 128 + col][bm * 128 + strip rows] = strip[row][col].
  for (int e = t; e < kStrip * kSyrkTile; e += kThreads) {
    const int mc = e / kStrip;   // column of the tile -> row of the mirror
    const int mr = e % kStrip;   // row of the strip -> column of the mirror
    const int row = bn * kSyrkTile + mc, c = bm * kSyrkTile + strip * kStrip + mr;
    if (row < nao && c < nao) out[(long long)row * nao + c] = mirror[mr][mc];
  }
}

// The symmetric-plus-antisymmetric reduce of the non-symmetric exchange:
// per 8-row strip of a lower tile pair, S = sum of `sym`'s slots and A = sum
// of `anti`'s; K[m][n] (+)= S + A on the tile, K[n][m] (+)= S - A on its
// mirror (skipped for diagonal tiles, which hold both triangles).
template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_reduce2_kernel(const T* __restrict__ sym, const T* __restrict__ anti,
                                                                 int nslot, int npairs, int nao, int accumulate,
                                                                 double* __restrict__ out) {
  constexpr int kStrip = 8;
  __shared__ double mirror[kStrip][kSyrkTile + 1];
  const int pair = blockIdx.x / (kSyrkTile / kStrip);
  const int strip = blockIdx.x % (kSyrkTile / kStrip);
  int bm, bn;
  tile_of(pair, bm, bn);
  const int t = threadIdx.x;
  const int col = t & (kSyrkTile - 1);
  const int r0 = t >> 7;
#pragma unroll
  for (int rr = 0; rr < kStrip / 2; ++rr) {
    const int lr = r0 + 2 * rr;
    const int trow = strip * kStrip + lr;
    const long long base = (long long)pair * (kSyrkTile * kSyrkTile) + trow * kSyrkTile + col;
    double ss = 0.0, aa = 0.0;
    for (int k = 0; k < nslot; ++k) {
      const long long o = base + (long long)k * npairs * (kSyrkTile * kSyrkTile);
      ss += (double)sym[o];
      aa += (double)anti[o];
    }
    const int row = bm * kSyrkTile + trow, c = bn * kSyrkTile + col;
    if (row < nao && c < nao) {
      const long long o = (long long)row * nao + c;
      out[o] = (accumulate ? out[o] : 0.0) + ss + aa;
    }
    mirror[lr][col] = ss - aa;
  }
  if (bm == bn) return;
  __syncthreads();
  for (int e = t; e < kStrip * kSyrkTile; e += kThreads) {
    const int mc = e / kStrip;
    const int mr = e % kStrip;
    const int row = bn * kSyrkTile + mc, c = bm * kSyrkTile + strip * kStrip + mr;
    if (row < nao && c < nao) {
      const long long o = (long long)row * nao + c;
      out[o] = (accumulate ? out[o] : 0.0) + mirror[mr][mc];
    }
  }
}

template <typename T, int BN, int BM, bool W2 = false>
hipError_t launch_x_t(const T* b, long long p, int naux, int nao, const T* c, int ldo, T* x, hipStream_t s) {
  const int nt = n_tiles(nao);
  const int xrows = x_rows(nao);
  const int nbn = (ldo + BN - 1) / BN;
  const int nbm = (xrows + BM - 1) / BM;
  const int per = std::max(1, (1 << 24) / (nbn * nbm));  // keep a launch below 2^32 work-items
  for (int base = 0; base < naux; base += per) {
    const int count = std::min(per, naux - base);
    const int nwg = nbn * nbm * count;
    if constexpr (W2)
      hipLaunchKernelGGL((xform_kernel_w2<T, BN, BM>), dim3((unsigned)nwg), dim3(kThreads), 0, s, b, p, nt, c, ldo,
                         x, xrows, base, nbn, nbm, nwg);
    else
      hipLaunchKernelGGL((xform_kernel<T, BN, BM>), dim3((unsigned)nwg), dim3(kThreads), 0, s, b, p, nt, c, ldo, x,
                         xrows, base, nbn, nbm, nwg);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

template <int BN>
hipError_t launch_x16_t(const uint16_t* bhi, const uint16_t* blo, long long p, int naux, int nao,
                        const uint16_t* cthi, const uint16_t* ctlo, int ldo, uint16_t* xhi, uint16_t* xlo,
                        hipStream_t s) {
  const int nt = n_tiles(nao);
  const int npad = nt * kTile;
  const int xrows = x_rows(nao);
  const int nbm = (xrows + 127) / 128;
  const int nbn = (ldo + BN - 1) / BN;
  const int per = std::max(1, (1 << 24) / (nbn * nbm));
  for (int base = 0; base < naux; base += per) {
    const int count = std::min(per, naux - base);
    const int nwg = nbn * nbm * count;
    hipLaunchKernelGGL((xform_bf16x3_kernel<BN>), dim3((unsigned)nwg), dim3(kThreads), 0, s, bhi, blo, p, nt, cthi,
                       ctlo, npad, ldo, xhi, xlo, xrows, base, nbn, nbm, nwg);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

}  // namespace

hipError_t launch_expand_full(const double* packed, long long p, int naux, int nao, int precision, void* out,
                              hipStream_t s) {
  const int npad = n_tiles(nao) * kTile;
  const int ntp = (int)n_tile_pairs(nao);
  const dim3 grid((unsigned)std::min<long long>((long long)naux * ntp, 1 << 20));
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(expand_full_kernel<double>, grid, dim3(kThreads), 0, s, packed, p, naux, ntp, npad,
                       (double*)out);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(expand_full_kernel<float>, grid, dim3(kThreads), 0, s, packed, p, naux, ntp, npad,
                       (float*)out);
  else
    hipLaunchKernelGGL(expand_full16_kernel, grid, dim3(kThreads), 0, s, packed, p, naux, ntp, npad,
                       (uint16_t*)out);
  return hipGetLastError();
}

hipError_t launch_orbitals_t(const double* c, int nao, int nocc, int precision, void* ct, hipStream_t s) {
  const int npad = n_tiles(nao) * kTile;
  const int ldo = occ_ld(nocc);
  const dim3 grid(grid_for((long long)npad * ldo));
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(orbitals_t_kernel<double>, grid, dim3(kThreads), 0, s, c, nao, nocc, npad, ldo, (double*)ct);
  else
    hipLaunchKernelGGL(orbitals_t_kernel<float>, grid, dim3(kThreads), 0, s, c, nao, nocc, npad, ldo, (float*)ct);
  return hipGetLastError();
}

bool x16_tiles256(int ldo, int bn16) { return ldo > kSyrkTile && (bn16 == 0 || bn16 == 256); }

hipError_t launch_half_transform_full(const void* bfull, int precision, int naux, int nao, const void* ct, int ldo,
                                      void* x, hipStream_t s, int bn16, bool x_ilv) {
  const int npad = n_tiles(nao) * kTile;
  const int xrows = x_rows(nao);
  if (x_ilv && !(precision == HFJK_BF16X3 && x16_tiles256(ldo, bn16))) return hipErrorInvalidValue;
  if (precision == HFJK_BF16X3 && x16_tiles256(ldo, bn16)) {
    // BF16x3 with more than 128 occupied columns: the 256 x 256 tiles.
    const uint16_t* bh = (const uint16_t*)bfull;
    const uint16_t* ch = (const uint16_t*)ct;
    uint16_t* xh = (uint16_t*)x;
    const long long nx16 = (long long)naux * xrows * ldo;
    const int nbn = (ldo + 255) / 256;
    const int nbm = (xrows + 255) / 256;
    const int per = std::max(1, (1 << 22) / (nbn * nbm));  // keep a launch below 2^32 work-items
    for (int base = 0; base < naux; base += per) {
      const int count = std::min(per, naux - base);
      const int nwg = nbn * nbm * count;
      if (x_ilv)
        hipLaunchKernelGGL(xform16_256_kernel<true>, dim3((unsigned)nwg), dim3(kR256Threads), 0, s, bh, npad, ch, ldo,
                           xh, xh + nx16, xrows, base, nbn, nbm, nwg);
      else
        hipLaunchKernelGGL(xform16_256_kernel<false>, dim3((unsigned)nwg), dim3(kR256Threads), 0, s, bh, npad, ch,
                           ldo, xh, xh + nx16, xrows, base, nbn, nbm, nwg);
      hipError_t err = hipGetLastError();
      if (err != hipSuccess) return err;
    }
    return hipSuccess;
  }
  const int nbn = (ldo + kSyrkTile - 1) / kSyrkTile;
  const int nbm = (xrows + kSyrkTile - 1) / kSyrkTile;
  const int per = std::max(1, (1 << 24) / (nbn * nbm));  // keep a launch below 2^32 work-items
  for (int base = 0; base < naux; base += per) {
    const int count = std::min(per, naux - base);
    const int nwg = nbn * nbm * count;
    if (precision == HFJK_FP64) {
      hipLaunchKernelGGL((xform_dma_kernel<double, 1>), dim3((unsigned)nwg), dim3(kThreads), 0, s,
                         (const double*)bfull, npad, (const double*)ct, ldo, (double*)x, xrows, base, nbn, nbm, nwg);
    } else if (precision == HFJK_FP32) {
      hipLaunchKernelGGL((xform_dma_kernel<float, 1>), dim3((unsigned)nwg), dim3(kThreads), 0, s,
                         (const float*)bfull, npad, (const float*)ct, ldo, (float*)x, xrows, base, nbn, nbm, nwg);
    } else {
      // BF16x3: B and C^T interleaved (hi and lo per 32-column block), X as hi then lo plane.
      uint16_t* xh = (uint16_t*)x;
      const long long nx16 = (long long)naux * xrows * ldo;
      hipLaunchKernelGGL(xform16_dma_kernel, dim3((unsigned)nwg), dim3(kThreads), 0, s, (const uint16_t*)bfull, npad,
                         (const uint16_t*)ct, ldo, xh, xh + nx16, xrows, base, nbn, nbm, nwg);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

hipError_t launch_orbitals_bf16x3(const double* c, int nao, int nocc, void* out, hipStream_t s, bool interleaved) {
  const int npad = n_tiles(nao) * kTile;
  const int ldo = occ_ld(nocc);
  const long long n = (long long)npad * ldo;
  uint16_t* hi = (uint16_t*)out;
  if (interleaved)
    hipLaunchKernelGGL(orbitals_bf16x3_kernel<true>, dim3(grid_for(n)), dim3(kThreads), 0, s, c, nao, nocc, npad, ldo,
                       hi, hi + n);
  else
    hipLaunchKernelGGL(orbitals_bf16x3_kernel<false>, dim3(grid_for(n)), dim3(kThreads), 0, s, c, nao, nocc, npad,
                       ldo, hi, hi + n);
  return hipGetLastError();
}

int charge_groups(int nao) { return (nao + kChargeRows - 1) / kChargeRows; }

hipError_t launch_charges_from_x(const void* x, int precision, int naux, int nao, int ldo, const void* cpad,
                                 double* parts, double* gamma, int max_groups, hipStream_t s, bool x_ilv) {
  const int xrows = x_rows(nao);
  const int ng = charge_groups(nao);
  const long long plane = (long long)naux * xrows * ldo;
  const long long nitems = (long long)naux * ng;
  // At most 2^30 workgroups (each loops over its items), or max_groups when set.
  long long groups = std::min<long long>(nitems, 1LL << 30);
  if (max_groups > 0) groups = std::min<long long>(groups, max_groups);
  const dim3 grid((unsigned)groups);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(charges_from_x_kernel<0>, grid, dim3(kThreads), 0, s, x, plane, nao, xrows, ldo, cpad, ng,
                       nitems, parts);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(charges_from_x_kernel<1>, grid, dim3(kThreads), 0, s, x, plane, nao, xrows, ldo, cpad, ng,
                       nitems, parts);
  else if (!x_ilv)
    hipLaunchKernelGGL(charges_from_x_kernel<2>, grid, dim3(kThreads), 0, s, x, plane, nao, xrows, ldo, cpad, ng,
                       nitems, parts);
  else
    hipLaunchKernelGGL(charges_from_x_kernel<3>, grid, dim3(kThreads), 0, s, x, plane, nao, xrows, ldo, cpad, ng,
                       nitems, parts);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(charge_reduce_kernel, dim3((unsigned)((naux + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     parts, naux, ng, gamma);
  return hipGetLastError();
}

int syrk_pairs(int nao) {
  const int nb = (x_rows(nao) + kSyrkTile - 1) / kSyrkTile;
  return nb * (nb + 1) / 2;
}

int syrk_tile() { return kSyrkTile; }

// Column tile: the tuned value when one is set, else the smallest of
// 32/64/128/256 that covers ldo.
static int pick_bn(int bn, int ldo) {
  if (bn == 32 || bn == 64 || bn == 128 || bn == 256) return bn;
  return ldo <= 32 ? 32 : ldo <= 64 ? 64 : ldo <= 128 ? 128 : 256;
}

template <typename T>
hipError_t launch_x_typed(const T* b, long long p, int naux, int nao, const T* c, int ldo, T* x, int bn, int bm,
                          hipStream_t s) {
  // 128-row tiles for column tiles up to 128 (a 128 x 256 tile would not fit the register file).
  if (bm == 128 && bn <= 128) {
    if (bn == 32) return launch_x_t<T, 32, 128>(b, p, naux, nao, c, ldo, x, s);
    if (bn == 64) return launch_x_t<T, 64, 128>(b, p, naux, nao, c, ldo, x, s);
    // FP64 128 x 128 is held to two waves per SIMD (see xform_kernel_w2).
    return launch_x_t<T, 128, 128, sizeof(T) == 8>(b, p, naux, nao, c, ldo, x, s);
  }
  if (bn == 32) return launch_x_t<T, 32, 64>(b, p, naux, nao, c, ldo, x, s);
  if (bn == 64) return launch_x_t<T, 64, 64>(b, p, naux, nao, c, ldo, x, s);
  if (bn == 128) return launch_x_t<T, 128, 64>(b, p, naux, nao, c, ldo, x, s);
  return launch_x_t<T, 256, 64>(b, p, naux, nao, c, ldo, x, s);
}

hipError_t launch_half_transform(const void* b, int precision, long long p, int naux, int nao, const void* c, int ldo,
                                 void* x, int bn, int bm, hipStream_t s) {
  bn = pick_bn(bn, ldo);
  if (precision == HFJK_FP64)
    return launch_x_typed<double>((const double*)b, p, naux, nao, (const double*)c, ldo, (double*)x, bn, bm, s);
  return launch_x_typed<float>((const float*)b, p, naux, nao, (const float*)c, ldo, (float*)x, bn, bm, s);
}

hipError_t launch_half_transform_bf16x3(const uint16_t* b, long long p, int naux, int nao, const void* ct, int ldo,
                                        void* x, int bn, hipStream_t s, const uint16_t* b_lo) {
  bn = pick_bn(bn, ldo);
  const uint16_t* bhi = b;
  const uint16_t* blo = b_lo != nullptr ? b_lo : b + (long long)naux * p;
  const long long nct = (long long)n_tiles(nao) * kTile * ldo;
  const uint16_t* cthi = (const uint16_t*)ct;
  const uint16_t* ctlo = cthi + nct;
  const long long nx = (long long)naux * x_rows(nao) * ldo;
  uint16_t* xhi = (uint16_t*)x;
  uint16_t* xlo = xhi + nx;
  if (bn == 32) return launch_x16_t<32>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
  if (bn == 64) return launch_x16_t<64>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
  if (bn == 128) return launch_x16_t<128>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
  return launch_x16_t<256>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
}

int syrk_run_aux(int precision, int ldo) {
  if (precision == HFJK_FP64) return 0;
  return std::max(1, kSyrkRunProducts / std::max(ldo, 1));
}

int syrk_slots(int precision, int naux, int ldo, int nsplit) {
  const int run = syrk_run_aux(precision, ldo);
  return run == 0 ? nsplit : (naux + run - 1) / run;
}

bool syrk16_tiles256(int nao, int ldo, bool is_signed, int kind, int shift) {
  return !is_signed && shift == 0 && kind != 1 && kind != 2 && ldo <= kDmaMaxLdo &&
         (long long)x_rows(nao) * ldo * 2 < (1LL << 31) && syrk_run_aux(HFJK_BF16X3, ldo) > 0;
}

hipError_t launch_syrk(const void* x, int precision, int naux, int nao, int ldo, const double* signs,
                       bool is_signed, int nsplit, void* partial, int kind, hipStream_t s, const double* auxsign,
                       int shift, bool x_ilv) {
  const int xrows = x_rows(nao);
  const int npairs = syrk_pairs(nao);
  const int run = syrk_run_aux(precision, ldo);
  // Per-aux signs are applied by the signed FP64/FP32 direct-LDS kernels only; a column shift by
  // the signed direct-LDS kernels of every precision.
  const bool ring_ok = is_signed && kind != 1 && ldo <= kDmaMaxLdo && (long long)xrows * ldo < (1LL << 30);
  if (auxsign != nullptr && (!ring_ok || precision == HFJK_BF16X3)) return hipErrorInvalidValue;
  if (shift != 0 && (!ring_ok || shift % 32 != 0 || shift < 0 || shift >= ldo)) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)(npairs * nsplit);
  // The direct-load kernel keeps per-lane element offsets within one aux
  // function's X in 32 bits and the sign table in LDS.
  const bool dma = kind != 1 && precision != HFJK_BF16X3 && ldo <= kDmaMaxLdo &&
                   (long long)xrows * ldo < (1LL << 31);
  // Signed columns need the LDS sign table: the one-workgroup-per-CU ring.
  const bool two_wg = kind != 3 && !is_signed;
  if (dma && precision == HFJK_FP64 && is_signed)
    hipLaunchKernelGGL((syrk_dma_kernel<double, true, 2>), dim3(grid), dim3(kThreads), 0, s, (const double*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (double*)partial);
  else if (dma && precision == HFJK_FP64 && two_wg)
    hipLaunchKernelGGL((syrk_dma_kernel<double, false, 1>), dim3(grid), dim3(kThreads), 0, s, (const double*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (double*)partial);
  else if (dma && precision == HFJK_FP64)
    hipLaunchKernelGGL((syrk_dma_kernel<double, false, 2>), dim3(grid), dim3(kThreads), 0, s, (const double*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (double*)partial);
  else if (dma && is_signed)
    hipLaunchKernelGGL((syrk_dma_kernel<float, true, 2>), dim3(grid), dim3(kThreads), 0, s, (const float*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (float*)partial);
  else if (dma && two_wg)
    hipLaunchKernelGGL((syrk_dma_kernel<float, false, 1>), dim3(grid), dim3(kThreads), 0, s, (const float*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (float*)partial);
  else if (dma)
    hipLaunchKernelGGL((syrk_dma_kernel<float, false, 2>), dim3(grid), dim3(kThreads), 0, s, (const float*)x, naux,
                       xrows, ldo, signs, auxsign, shift, nsplit, npairs, run, (float*)partial);
  else if (precision == HFJK_FP64)
    hipLaunchKernelGGL(syrk_kernel<double>, dim3(grid), dim3(kThreads), 0, s, (const double*)x, naux, xrows, ldo,
                       signs, nsplit, npairs, run, (double*)partial);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(syrk_kernel<float>, dim3(grid), dim3(kThreads), 0, s, (const float*)x, naux, xrows, ldo, signs,
                       nsplit, npairs, run, (float*)partial);
  else {
    const long long nx = (long long)naux * xrows * ldo;
    const uint16_t* xhi = (const uint16_t*)x;
    const bool ring16 = kind != 1 && ldo <= kDmaMaxLdo && (long long)xrows * ldo * 2 < (1LL << 31);
    // Unsigned (occupied orbitals, no column shift): the 256 x 256 tiles unless the 128-row ring is
    // asked for (kind 2); their grid holds about four workgroups per CU of a 256-CU device.
    if (x_ilv && !syrk16_tiles256(nao, ldo, is_signed, kind, shift)) return hipErrorInvalidValue;
    if (ring16 && !is_signed && shift == 0 && kind != 2 && run > 0) {
      const int nb128 = (xrows + kSyrkTile - 1) / kSyrkTile;
      const int nb256 = (nb128 + 1) / 2;
      const int np256 = nb256 * (nb256 + 1) / 2;
      const int nslot = (naux + run - 1) / run;
      const int ns256 = std::max(8, std::min((nslot + 7) / 8 * 8, 1024 / np256 / 8 * 8));
      if (x_ilv)
        hipLaunchKernelGGL(syrk16_256_kernel<true>, dim3((unsigned)(np256 * ns256)), dim3(kR256Threads), 0, s, xhi,
                           xhi + nx, naux, xrows, ldo, ns256, np256, npairs, nb128, run, (float*)partial);
      else
        hipLaunchKernelGGL(syrk16_256_kernel<false>, dim3((unsigned)(np256 * ns256)), dim3(kR256Threads), 0, s, xhi,
                           xhi + nx, naux, xrows, ldo, ns256, np256, npairs, nb128, run, (float*)partial);
    } else if (ring16 && is_signed)
      hipLaunchKernelGGL(syrk16_dma_kernel<true>, dim3(grid), dim3(kThreads), 0, s, xhi, xhi + nx, naux, xrows, ldo,
                         signs, shift, nsplit, npairs, run, (float*)partial);
    else if (ring16)
      hipLaunchKernelGGL(syrk16_dma_kernel<false>, dim3(grid), dim3(kThreads), 0, s, xhi, xhi + nx, naux, xrows,
                         ldo, signs, 0, nsplit, npairs, run, (float*)partial);
    else
      hipLaunchKernelGGL(syrk_bf16x3_kernel, dim3(grid), dim3(kThreads), 0, s, xhi, xhi + nx, naux, xrows, ldo,
                         signs, nsplit, npairs, run, (float*)partial);
  }
  return hipGetLastError();
}

hipError_t launch_syrk_reduce2(const void* sym, const void* anti, int precision, int nslot, int nao,
                               bool accumulate, double* out, hipStream_t s) {
  const int npairs = syrk_pairs(nao);
  const unsigned grid = (unsigned)(npairs * (kSyrkTile / 8));
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(syrk_reduce2_kernel<double>, dim3(grid), dim3(kThreads), 0, s, (const double*)sym,
                       (const double*)anti, nslot, npairs, nao, accumulate ? 1 : 0, out);
  else
    hipLaunchKernelGGL(syrk_reduce2_kernel<float>, dim3(grid), dim3(kThreads), 0, s, (const float*)sym,
                       (const float*)anti, nslot, npairs, nao, accumulate ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t launch_syrk_reduce(const void* partial, int precision, int nslot, int nao, double* out, hipStream_t s) {
  const int npairs = syrk_pairs(nao);
  // One workgroup per 8-row strip of every lower tile pair.
  const unsigned grid = (unsigned)(npairs * (kSyrkTile / 8));
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(syrk_reduce_kernel<double>, dim3(grid), dim3(kThreads), 0, s, (const double*)partial, nslot,
                       npairs, nao, out);
  else
    hipLaunchKernelGGL(syrk_reduce_kernel<float>, dim3(grid), dim3(kThreads), 0, s, (const float*)partial, nslot,
                       npairs, nao, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ batched (persistent)
//
// Every instance of a batch contributes items to each phase; a fixed grid of
// workgroups pulls items from per-XCD counters (batch_next) until its
// share is exhausted, so instances of very different sizes share one launch
// and balance dynamically.

namespace {

template <typename T, int BN>
__global__ __launch_bounds__(kThreads) void batch_xform_kernel(BatchView v, const T* __restrict__ b,
                                                                const T* __restrict__ cbase, T* __restrict__ xbase) {
  __shared__ XformSmem<T, BN> sm;
  const long long* pre = v.prefix + (long long)kPhXform * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhXform, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int nb = (int)(local % in.nbn);
    const int mb = (int)((local / in.nbn) % in.nbm);
    const int q = (int)(local / ((long long)in.nbn * in.nbm));
    xform_tile<T, BN>(b + in.b_off + (long long)q * in.p, in.nt, cbase + in.cp_off, in.ldo,
                      xbase + in.x_off + (long long)q * in.xrows * in.ldo, in.xrows, mb, nb, sm);
  }
}

template <int BN>
__global__ __launch_bounds__(kThreads) void batch_xform16_kernel(BatchView v, const uint16_t* __restrict__ b,
                                                                  long long b_plane, const uint16_t* __restrict__ cbase,
                                                                  long long c_plane, uint16_t* __restrict__ xbase,
                                                                  long long x_plane) {
  __shared__ __attribute__((aligned(16))) Xform16Smem<BN> sm;
  const long long* pre = v.prefix + (long long)kPhXform * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhXform, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int nb = (int)(local % in.nbn);
    const int mb = (int)((local / in.nbn) % in.nbm);
    const int q = (int)(local / ((long long)in.nbn * in.nbm));
    const uint16_t* hq = b + in.b_off + (long long)q * in.p;
    const uint16_t* ct = cbase + in.cp_off;
    uint16_t* xh = xbase + in.x_off + (long long)q * in.xrows * in.ldo;
    xform16_tile<BN>(hq, hq + b_plane, in.nt, ct, ct + c_plane, in.nt * kTile, in.ldo, xh, xh + x_plane, in.xrows,
                     mb, nb, sm);
  }
}

// The half-transform with each instance's own tile (in.bm x in.bn: 64 x 32,
// 64 x 64, 64 x 128 or 128 x 64): the tile shapes share one LDS buffer.
template <typename T>
union XformSmemAny {
  XformSmem<T, 32> s32;
  XformSmem<T, 64> s64;
  XformSmem<T, 128> s128;
  XformSmem<T, 64, 128> t64;
};

template <typename T>
__global__ __launch_bounds__(kThreads, 2) void batch_xform_any_kernel(BatchView v, const T* __restrict__ b,
                                                                    const T* __restrict__ cbase,
                                                                    T* __restrict__ xbase) {
  __shared__ XformSmemAny<T> sm;
  const long long* pre = v.prefix + (long long)kPhXform * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhXform, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int nb = (int)(local % in.nbn);
    const int mb = (int)((local / in.nbn) % in.nbm);
    const int q0 = (int)(local / ((long long)in.nbn * in.nbm)) * in.qg;
    const int q1 = min(in.naux, q0 + in.qg);
    const T* bb = b + in.b_off;
    const T* cp = cbase + in.cp_off;
    T* xb = xbase + in.x_off;
    if (in.bm == 128)
      xform_tiles<T, 64, 128, true>(bb, in.p, in.nt, cp, in.ldo, xb, in.xrows, mb, nb, q0, q1, sm.t64, in.nao,
                                    in.nocc, in.nao);
    else if (in.bn == 32)
      xform_tiles<T, 32, 64, true>(bb, in.p, in.nt, cp, in.ldo, xb, in.xrows, mb, nb, q0, q1, sm.s32, in.nao,
                                   in.nocc, in.nao);
    else if (in.bn == 64)
      xform_tiles<T, 64, 64, true>(bb, in.p, in.nt, cp, in.ldo, xb, in.xrows, mb, nb, q0, q1, sm.s64, in.nao,
                                   in.nocc, in.nao);
    else
      xform_tiles<T, 128, 64, true>(bb, in.p, in.nt, cp, in.ldo, xb, in.xrows, mb, nb, q0, q1, sm.s128, in.nao,
                                    in.nocc, in.nao);
  }
}

union Xform16SmemAny {
  Xform16Smem<32> s32;
  Xform16Smem<64> s64;
  Xform16Smem<128> s128;
};

__global__ __launch_bounds__(kThreads, 2) void batch_xform16_any_kernel(BatchView v, const uint16_t* __restrict__ b,
                                                                      long long b_plane,
                                                                      const uint16_t* __restrict__ cbase,
                                                                      long long c_plane, uint16_t* __restrict__ xbase,
                                                                      long long x_plane) {
  __shared__ __attribute__((aligned(16))) Xform16SmemAny sm;
  const long long* pre = v.prefix + (long long)kPhXform * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhXform, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int nb = (int)(local % in.nbn);
    const int mb = (int)((local / in.nbn) % in.nbm);
    const int q0 = (int)(local / ((long long)in.nbn * in.nbm)) * in.qg;
    const int q1 = min(in.naux, q0 + in.qg);
    const uint16_t* hq = b + in.b_off;
    const uint16_t* ct = cbase + in.cp_off;
    uint16_t* xh = xbase + in.x_off;
    if (in.bn == 32)
      xform16_tiles<32, true>(hq, hq + b_plane, in.p, in.nt, ct, ct + c_plane, in.nt * kTile, in.ldo, xh,
                              xh + x_plane, in.xrows, mb, nb, q0, q1, sm.s32, in.nao, in.nocc);
    else if (in.bn == 64)
      xform16_tiles<64, true>(hq, hq + b_plane, in.p, in.nt, ct, ct + c_plane, in.nt * kTile, in.ldo, xh,
                              xh + x_plane, in.xrows, mb, nb, q0, q1, sm.s64, in.nao, in.nocc);
    else
      xform16_tiles<128, true>(hq, hq + b_plane, in.p, in.nt, ct, ct + c_plane, in.nt * kTile, in.ldo, xh,
                               xh + x_plane, in.xrows, mb, nb, q0, q1, sm.s128, in.nao, in.nocc);
  }
}

// SYRK items of the instances with 128-row tiles whose X offsets fit the
// direct-LDS ring (32-bit byte offsets within one aux function's X): one
// (tile pair, split) per item, the partial layout of batch_syrk_kernel.
template <typename T>
__global__ __launch_bounds__(kThreads, 2) void batch_syrk_ring_kernel(BatchView v, const T* __restrict__ xbase,
                                                                       T* __restrict__ kbase) {
  const long long* pre = v.prefix + (long long)kPhSyrkRing * (v.ninst + 1);
  for (;;) {
    long long end;
    // batch_next's barriers also order this item's LDS loads after the previous item's reads.
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhSyrkRing, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int pair = (int)(local % in.npairs);
    const int s = (int)(local / in.npairs);
    int mb, nb;
    tile_of(pair, mb, nb);
    const int qb = (int)((long long)s * in.naux / in.nsplit);
    const int qe = (int)((long long)(s + 1) * in.naux / in.nsplit);
    syrk_ring_tile<T, 1>(xbase + in.x_off, in.xrows, in.ldo, mb, nb, qb, qe,
                         kbase + in.kp_off + ((long long)s * in.npairs + pair) * (kSyrkTile * kSyrkTile), in.nao);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void batch_syrk_kernel(BatchView v, const T* __restrict__ xbase,
                                                               const double* __restrict__ ws64, T* __restrict__ kbase) {
  __shared__ SyrkSmem<T> sm;
  const long long* pre = v.prefix + (long long)kPhSyrk * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhSyrk, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int pair = (int)(local % in.npairs);
    const int s = (int)(local / in.npairs);
    int mb, nb;
    tile_of(pair, mb, nb);
    const int qb = (int)((long long)s * in.naux / in.nsplit);
    const int qe = (int)((long long)(s + 1) * in.naux / in.nsplit);
    T* dst = kbase + in.kp_off + ((long long)s * in.npairs + pair) * ((long long)in.st * in.st);
    if (in.st == 64)
      syrk_tile<T, 64>(xbase + in.x_off, qb, qe, in.xrows, in.ldo, ws64 + in.sg_off, mb, nb, dst, sm);
    else
      syrk_tile<T, kSyrkTile>(xbase + in.x_off, qb, qe, in.xrows, in.ldo, ws64 + in.sg_off, mb, nb, dst, sm);
  }
}

__global__ __launch_bounds__(kThreads) void batch_syrk16_kernel(BatchView v, const uint16_t* __restrict__ xbase,
                                                                 long long x_plane, const double* __restrict__ ws64,
                                                                 float* __restrict__ kbase) {
  __shared__ __attribute__((aligned(16))) Syrk16Smem sm;
  const long long* pre = v.prefix + (long long)kPhSyrk * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhSyrk, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int pair = (int)(local % in.npairs);
    const int s = (int)(local / in.npairs);
    int mb, nb;
    tile_of(pair, mb, nb);
    const int qb = (int)((long long)s * in.naux / in.nsplit);
    const int qe = (int)((long long)(s + 1) * in.naux / in.nsplit);
    const uint16_t* xh = xbase + in.x_off;
    syrk16_tile(xh, xh + x_plane, qb, qe, in.xrows, in.ldo, ws64 + in.sg_off, mb, nb,
                kbase + in.kp_off + ((long long)s * in.npairs + pair) * (kSyrkTile * kSyrkTile), sm);
  }
}

// The BF16x3 SYRK items on the direct-LDS ring (every instance whose X offsets fit it).
__global__ __launch_bounds__(kThreads, 1) void batch_syrk16_ring_kernel(BatchView v,
                                                                         const uint16_t* __restrict__ xbase,
                                                                         long long x_plane,
                                                                         float* __restrict__ kbase) {
  const long long* pre = v.prefix + (long long)kPhSyrkRing * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhSyrkRing, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long local = item - pre[i];
    const int pair = (int)(local % in.npairs);
    const int s = (int)(local / in.npairs);
    int mb, nb;
    tile_of(pair, mb, nb);
    const int qb = (int)((long long)s * in.naux / in.nsplit);
    const int qe = (int)((long long)(s + 1) * in.naux / in.nsplit);
    const uint16_t* xh = xbase + in.x_off;
    syrk16_ring_tile(xh, xh + x_plane, in.xrows, in.ldo, mb, nb, qb, qe,
                     kbase + in.kp_off + ((long long)s * in.npairs + pair) * (kSyrkTile * kSyrkTile));
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void batch_syrk_reduce_kernel(BatchView v, const T* __restrict__ kbase,
                                                                      double* __restrict__ kout) {
  const long long* pre = v.prefix + (long long)kPhSyrkReduce * (v.ninst + 1);
  for (;;) {
    long long end;
    const long long item = batch_next(v.prefix, v.counters, v.ninst, kPhSyrkReduce, end);
    if (item >= end) break;
    const int i = batch_find(pre, v.ninst, item);
    const BatchInst in = v.insts[i];
    const long long idx = (item - pre[i]) * kThreads + threadIdx.x;
    if (idx < (long long)in.nao * in.nao) {
      const int row = (int)(idx / in.nao);
      const int col = (int)(idx % in.nao);
      const int st = in.st;
      const int bm = row / st, bn = col / st;
      long long pair;
      int local;
      if (bm >= bn) {
        pair = tidx(bm, bn);
        local = (row % st) * st + col % st;
      } else {
        pair = tidx(bn, bm);
        local = (col % st) * st + row % st;
      }
      const T* part = kbase + in.kp_off;
      double sum = 0.0;
      for (int s = 0; s < in.nsplit; ++s)
        sum += (double)part[((long long)s * in.npairs + pair) * ((long long)st * st) + local];
      kout[in.d_off + idx] = sum;
    }
  }
}

}  // namespace

int batch_tile_rows(int precision, int xrows, int bn) {
  if (precision == HFJK_BF16X3) return 128;
  return xrows >= 128 && bn == 64 ? 128 : kBlock;
}

hipError_t launch_batch_xform(BatchView v, int precision, int bn, const void* b, long long b_plane, const void* cbase,
                              long long c_plane, void* xbase, long long x_plane, int grid, hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  // bn = 256: one fixed tile for every instance; otherwise each instance's own tile.
  if (precision == HFJK_FP64) {
    const double* bb = (const double*)b;
    const double* cb = (const double*)cbase;
    double* xb = (double*)xbase;
    if (bn == 256) hipLaunchKernelGGL((batch_xform_kernel<double, 256>), g, blk, 0, s, v, bb, cb, xb);
    else hipLaunchKernelGGL(batch_xform_any_kernel<double>, g, blk, 0, s, v, bb, cb, xb);
  } else if (precision == HFJK_FP32) {
    const float* bb = (const float*)b;
    const float* cb = (const float*)cbase;
    float* xb = (float*)xbase;
    if (bn == 256) hipLaunchKernelGGL((batch_xform_kernel<float, 256>), g, blk, 0, s, v, bb, cb, xb);
    else hipLaunchKernelGGL(batch_xform_any_kernel<float>, g, blk, 0, s, v, bb, cb, xb);
  } else {
    const uint16_t* bb = (const uint16_t*)b;
    const uint16_t* cb = (const uint16_t*)cbase;
    uint16_t* xb = (uint16_t*)xbase;
    if (bn == 256)
      hipLaunchKernelGGL((batch_xform16_kernel<256>), g, blk, 0, s, v, bb, b_plane, cb, c_plane, xb, x_plane);
    else
      hipLaunchKernelGGL(batch_xform16_any_kernel, g, blk, 0, s, v, bb, b_plane, cb, c_plane, xb, x_plane);
  }
  return hipGetLastError();
}

hipError_t launch_batch_syrk(BatchView v, int precision, const void* xbase, long long x_plane, const double* ws64,
                             void* kbase, int grid, hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  // grid = two workgroups per CU, the FP64 / FP32 ring kernel's occupancy; workgroups that find
  // no CU free start late, find their share taken and leave.
  if (precision == HFJK_FP64) {
    hipLaunchKernelGGL(batch_syrk_kernel<double>, g, blk, 0, s, v, (const double*)xbase, ws64, (double*)kbase);
    hipLaunchKernelGGL(batch_syrk_ring_kernel<double>, g, blk, 0, s, v, (const double*)xbase, (double*)kbase);
  } else if (precision == HFJK_FP32) {
    hipLaunchKernelGGL(batch_syrk_kernel<float>, g, blk, 0, s, v, (const float*)xbase, ws64, (float*)kbase);
    hipLaunchKernelGGL(batch_syrk_ring_kernel<float>, g, blk, 0, s, v, (const float*)xbase, (float*)kbase);
  } else {
    hipLaunchKernelGGL(batch_syrk16_kernel, g, blk, 0, s, v, (const uint16_t*)xbase, x_plane, ws64, (float*)kbase);
    hipLaunchKernelGGL(batch_syrk16_ring_kernel, g, blk, 0, s, v, (const uint16_t*)xbase, x_plane, (float*)kbase);
  }
  return hipGetLastError();
}

hipError_t launch_batch_syrk_reduce(BatchView v, int precision, const void* kbase, double* k, int grid,
                                    hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(batch_syrk_reduce_kernel<double>, g, blk, 0, s, v, (const double*)kbase, k);
  else
    hipLaunchKernelGGL(batch_syrk_reduce_kernel<float>, g, blk, 0, s, v, (const float*)kbase, k);
  return hipGetLastError();
}

}  // namespace hfjk
