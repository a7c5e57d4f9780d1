// Batched DF-J/K: many instances of uneven size in one persistent launch
// per phase (charges, Coulomb, orbitals, half-transform, SYRK, reductions).
//
// A plan (hfjk_batch_create) derives every instance's shapes, splits and
// buffer offsets on the host; hfjk_batch_bind uploads it once into the
// caller's workspace; hfjk_batch_run only enqueues kernels (a counter reset
// kernel and one persistent kernel per phase), so a run can be captured in a
// hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "hfjk_internal.h"

struct hfjk_batch {
  hfjk_ctx* ctx = nullptr;
  int precision = HFJK_FP64;
  int tile = 0;  // half-transform column tile
  int ninst = 0;
  std::vector<hfjk::BatchInst> insts;
  std::vector<long long> prefix;  // [kPhCount][ninst + 1]
  long long b_total = 0;          // elements of one plane of the batch tensor
  long long c_total = 0, d_total = 0;
  long long ws64_total = 0;       // fp64 workspace elements
  long long cp_total = 0, x_total = 0, kp_total = 0;  // mode elements (per plane for bf16)
  size_t off_insts = 0, off_prefix = 0, off_counters = 0, off_ws64 = 0, off_cp = 0, off_x = 0, off_kp = 0;
  size_t bytes = 0;
  void* ws = nullptr;
  int grid_mfma = 0, grid_elem = 0;
};

namespace hfjk {
namespace {

constexpr int kGammaQb = 64;

template <int PREC>
__device__ __forceinline__ void bload4(const void* b, long long plane, long long off, double out[4]) {
  if (PREC == HFJK_FP64) {
    const double2* src = (const double2*)((const double*)b + off);
    const double2 v0 = src[0], v1 = src[1];
    out[0] = v0.x;
    out[1] = v0.y;
    out[2] = v1.x;
    out[3] = v1.y;
  } else if (PREC == HFJK_FP32) {
    const float4 v = *(const float4*)((const float*)b + off);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    out[3] = v.w;
  } else {
    const uint16_t* hi = (const uint16_t*)b + off;
    const ushort4 h = *(const ushort4*)hi;
    const ushort4 l = *(const ushort4*)(hi + plane);
    auto f = [](uint16_t x) { return __uint_as_float(((uint32_t)x) << 16); };
    out[0] = (double)(f(h.x) + f(l.x));
    out[1] = (double)(f(h.y) + f(l.y));
    out[2] = (double)(f(h.z) + f(l.z));
    out[3] = (double)(f(h.w) + f(l.w));
  }
}

__device__ __forceinline__ void rc_of(int tile, int& r, int& c) {
  int rr = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((rr + 1) * (rr + 2) / 2 <= tile) ++rr;
  while (rr * (rr + 1) / 2 > tile) --rr;
  r = rr;
  c = tile - rr * (rr + 1) / 2;
}

__device__ __forceinline__ long long tix(int r, int c) { return (long long)r * (r + 1) / 2 + c; }

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(kThreads) void reset_kernel(unsigned long long* counters, int n) {
  for (int i = threadIdx.x; i < n; i += kThreads) counters[i] = 0ull;
}

#define BATCH_LOOP(PHASE)                                                           \
  const long long* pre = v.prefix + (long long)(PHASE) * (v.ninst + 1);             \
  for (;;) {                                                                        \
    long long end;                                                                  \
    const long long item = batch_next(v.prefix, v.counters, v.ninst, (PHASE), end); \
    if (item >= end) break;                                                         \
    const int inst = batch_find(pre, v.ninst, item);                                \
    const BatchInst in = v.insts[inst];                                             \
    const long long local = item - pre[inst];

#define BATCH_END }

// Packed weights of D (D on diagonal tiles, D + D^T elsewhere), one tile per item.
__global__ __launch_bounds__(kThreads) void b_weights_kernel(BatchView v, const double* __restrict__ d,
                                                             double* __restrict__ ws64) {
  BATCH_LOOP(kPhWeights)
  int r, c;
  rc_of((int)local, r, c);
  const double* di = d + in.d_off;
  double* w = ws64 + in.w_off + local * kTileElems;
  for (int v4 = 0; v4 < 4; ++v4) {
    const int e = threadIdx.x * 4 + v4;
    const int row = r * kTile + e / kTile;
    const int col = c * kTile + e % kTile;
    double value = 0.0;
    if (row < in.nao && col < in.nao) {
      value = di[(long long)row * in.nao + col];
      if (r != c) value += di[(long long)col * in.nao + row];
    }
    w[e] = value;
  }
  BATCH_END
}

// Charge partials: gp[tile][q] for a 64-q block, one (tile, q block) per item.
template <int PREC>
__global__ __launch_bounds__(kThreads) void b_gamma_kernel(BatchView v, const void* __restrict__ b, long long plane,
                                                           double* __restrict__ ws64) {
  __shared__ double red[kThreads / 64][kGammaQb];
  BATCH_LOOP(kPhGamma)
  const long long ntiles = in.p / kTileElems;
  const long long tile = local % ntiles;
  const int q0 = (int)(local / ntiles) * kGammaQb;
  const long long e = tile * kTileElems + threadIdx.x * 4;
  const double* w = ws64 + in.w_off + e;
  const double w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
  const int qn = min(kGammaQb, in.naux - q0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int qi = 0; qi < qn; ++qi) {
    double bv[4];
    bload4<PREC>(b, plane, in.b_off + (long long)(q0 + qi) * in.p + e, bv);
    double s = bv[0] * w0 + bv[1] * w1 + bv[2] * w2 + bv[3] * w3;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) red[wave][qi] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < qn)
    ws64[in.gp_off + tile * in.naux + q0 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  __syncthreads();
  BATCH_END
}

// gamma[q] = sum over tiles, 256 aux functions per item.
__global__ __launch_bounds__(kThreads) void b_gamma_reduce_kernel(BatchView v, double* __restrict__ ws64) {
  BATCH_LOOP(kPhGammaReduce)
  const int q = (int)local * kThreads + threadIdx.x;
  if (q < in.naux) {
    const long long ntiles = in.p / kTileElems;
    double s = 0.0;
    for (long long t = 0; t < ntiles; ++t) s += ws64[in.gp_off + t * in.naux + q];
    ws64[in.g_off + q] = s;
  }
  BATCH_END
}

// gamma[q] from a caller's flat charge vector (instance i at q_off), 256 per item.
__global__ __launch_bounds__(kThreads) void b_charges_load_kernel(BatchView v, const double* __restrict__ charges,
                                                                  double* __restrict__ ws64) {
  BATCH_LOOP(kPhGammaReduce)
  const int q = (int)local * kThreads + threadIdx.x;
  if (q < in.naux) ws64[in.g_off + q] = charges[in.q_off + q];
  BATCH_END
}

// gamma[q] = <B[q], C C^T> = sum_{mu < nao, i < nocc} C[mu][i] X[q][mu][i] from the
// half-transformed X of the exchange (mode layout; BF16X3: hi + lo planes)
// and the caller's FP64 orbitals, one aux function per item. The rows
// [0, nao) of X[q] are one contiguous range of nao * ldo elements, streamed
// in 16-byte groups (ldo is a multiple of 32, so no group straddles two
// rows); columns past nocc are skipped (X is zero there).
template <int PREC>
__global__ __launch_bounds__(kThreads) void b_charges_x_kernel(BatchView v, const double* __restrict__ c,
                                                               const void* __restrict__ xbase, long long x_plane,
                                                               double* __restrict__ ws64) {
  constexpr int V = PREC == HFJK_FP64 ? 2 : PREC == HFJK_FP32 ? 4 : 8;  // elements per 16-byte group
  __shared__ double red[kThreads / 64];
  BATCH_LOOP(kPhCharges)
  const double* ci = c + in.c_off;
  const long long xq = in.x_off + local * in.xrows * in.ldo;
  double s = 0.0;
  const int per_row = in.ldo / V;
  const int groups = in.nao * per_row;
  // Group g = mu * per_row + gi, stepped by the block size without a division.
  const int dm = kThreads / per_row, dg = kThreads % per_row;
  int mu = (int)threadIdx.x / per_row, gi = (int)threadIdx.x % per_row;
  for (int g = threadIdx.x; g < groups; g += kThreads) {
    const int i0 = gi * V;
    if (i0 < in.nocc) {
      const long long at = xq + (long long)mu * in.ldo + i0;
      float xf[V];
      double xd[V];
      if constexpr (PREC == HFJK_FP64) {
        const double2 w = *(const double2*)((const double*)xbase + at);
        xd[0] = w.x;
        xd[1] = w.y;
      } else if constexpr (PREC == HFJK_FP32) {
        const float4 w = *(const float4*)((const float*)xbase + at);
        xf[0] = w.x;
        xf[1] = w.y;
        xf[2] = w.z;
        xf[3] = w.w;
      } else {
        const uint16_t* xh = (const uint16_t*)xbase;
        const uint4 h = *(const uint4*)(xh + at);
        const uint4 l = *(const uint4*)(xh + at + x_plane);
        const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint32_t sh = (e & 1) * 16;
          xf[e] = __uint_as_float(((hw[e >> 1] >> sh) & 0xFFFFu) << 16) +
                  __uint_as_float(((lw[e >> 1] >> sh) & 0xFFFFu) << 16);
        }
      }
      const double* crow = ci + (long long)mu * in.nocc;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const double x = PREC == HFJK_FP64 ? xd[e] : (double)xf[e];
        if (i0 + e < in.nocc) s += crow[i0 + e] * x;
      }
    }
    mu += dm;
    gi += dg;
    if (gi >= per_row) {
      gi -= per_row;
      ++mu;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) ws64[in.g_off + local] = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  BATCH_END
}

// Packed J of one tile: sum_q gamma[q] B[q][tile], one tile per item.
template <int PREC>
__global__ __launch_bounds__(kThreads) void b_coulomb_kernel(BatchView v, const void* __restrict__ b, long long plane,
                                                             double* __restrict__ ws64) {
  BATCH_LOOP(kPhCoulomb)
  const long long e = local * kTileElems + threadIdx.x * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* g = ws64 + in.g_off;
  for (int q = 0; q < in.naux; ++q) {
    double bv[4];
    bload4<PREC>(b, plane, in.b_off + (long long)q * in.p + e, bv);
    const double gq = g[q];
    for (int k = 0; k < 4; ++k) acc[k] += gq * bv[k];
  }
  double* jp = ws64 + in.jp_off + e;
  for (int k = 0; k < 4; ++k) jp[k] = acc[k];
  BATCH_END
}

// J [nao][nao] from the packed J, 256 elements per item.
__global__ __launch_bounds__(kThreads) void b_coulomb_unpack_kernel(BatchView v, const double* __restrict__ ws64,
                                                                    double* __restrict__ j) {
  BATCH_LOOP(kPhCoulombUnpack)
  const long long idx = local * kThreads + threadIdx.x;
  if (idx < (long long)in.nao * in.nao) {
    const int row = (int)(idx / in.nao), col = (int)(idx % in.nao);
    const int r = row / kTile, c = col / kTile;
    const long long src = r >= c ? tix(r, c) * kTileElems + (row % kTile) * kTile + col % kTile
                                 : tix(c, r) * kTileElems + (col % kTile) * kTile + row % kTile;
    j[in.d_off + idx] = ws64[in.jp_off + src];
  }
  BATCH_END
}

// Padded orbitals in the mode's layout and the (all +1) column signs.
template <int PREC>
__global__ __launch_bounds__(kThreads) void b_orbitals_kernel(BatchView v, const double* __restrict__ c,
                                                              double* __restrict__ ws64, void* __restrict__ cbase,
                                                              long long c_plane) {
  BATCH_LOOP(kPhOrbitals)
  const int npad = in.nt * kTile;
  const long long idx = local * kThreads + threadIdx.x;
  if (local == 0)
    for (int i = threadIdx.x; i < in.ldo; i += kThreads) ws64[in.sg_off + i] = i < in.nocc ? 1.0 : 0.0;
  if (idx < (long long)npad * in.ldo) {
    const double* ci = c + in.c_off;
    if (PREC == HFJK_BF16X3) {
      const int i = (int)(idx / npad), k = (int)(idx % npad);  // Ct[i][k]
      const float x = (k < in.nao && i < in.nocc) ? (float)ci[(long long)k * in.nocc + i] : 0.0f;
      uint16_t* hi = (uint16_t*)cbase + in.cp_off;
      const uint16_t h = to_bf16(x);
      hi[idx] = h;
      hi[idx + c_plane] = to_bf16(x - __uint_as_float(((uint32_t)h) << 16));
    } else {
      const int k = (int)(idx / in.ldo), i = (int)(idx % in.ldo);  // C[k][i]
      const double x = (k < in.nao && i < in.nocc) ? ci[(long long)k * in.nocc + i] : 0.0;
      if (PREC == HFJK_FP64)
        ((double*)cbase)[in.cp_off + idx] = x;
      else
        ((float*)cbase)[in.cp_off + idx] = (float)x;
    }
  }
  BATCH_END
}

#undef BATCH_LOOP
#undef BATCH_END

}  // namespace

hipError_t launch_batch_reset(BatchView v, hipStream_t s) {
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(kThreads), 0, s, v.counters, kPhCount * 8);
  return hipGetLastError();
}

hipError_t launch_batch_coulomb(BatchView v, int precision, const void* b, long long b_plane, const double* d,
                                double* ws64, double* j, int grid, hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  hipLaunchKernelGGL(b_weights_kernel, g, blk, 0, s, v, d, ws64);
  if (precision == HFJK_FP64) {
    hipLaunchKernelGGL(b_gamma_kernel<HFJK_FP64>, g, blk, 0, s, v, b, b_plane, ws64);
    hipLaunchKernelGGL(b_gamma_reduce_kernel, g, blk, 0, s, v, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP64>, g, blk, 0, s, v, b, b_plane, ws64);
  } else if (precision == HFJK_FP32) {
    hipLaunchKernelGGL(b_gamma_kernel<HFJK_FP32>, g, blk, 0, s, v, b, b_plane, ws64);
    hipLaunchKernelGGL(b_gamma_reduce_kernel, g, blk, 0, s, v, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP32>, g, blk, 0, s, v, b, b_plane, ws64);
  } else {
    hipLaunchKernelGGL(b_gamma_kernel<HFJK_BF16X3>, g, blk, 0, s, v, b, b_plane, ws64);
    hipLaunchKernelGGL(b_gamma_reduce_kernel, g, blk, 0, s, v, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_BF16X3>, g, blk, 0, s, v, b, b_plane, ws64);
  }
  hipLaunchKernelGGL(b_coulomb_unpack_kernel, g, blk, 0, s, v, ws64, j);
  return hipGetLastError();
}

hipError_t launch_batch_coulomb_occ(BatchView v, int precision, const void* b, long long b_plane, const double* c,
                                    const void* xbase, long long x_plane, double* ws64, double* j, int grid,
                                    hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  if (precision == HFJK_FP64) {
    hipLaunchKernelGGL(b_charges_x_kernel<HFJK_FP64>, g, blk, 0, s, v, c, xbase, x_plane, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP64>, g, blk, 0, s, v, b, b_plane, ws64);
  } else if (precision == HFJK_FP32) {
    hipLaunchKernelGGL(b_charges_x_kernel<HFJK_FP32>, g, blk, 0, s, v, c, xbase, x_plane, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP32>, g, blk, 0, s, v, b, b_plane, ws64);
  } else {
    hipLaunchKernelGGL(b_charges_x_kernel<HFJK_BF16X3>, g, blk, 0, s, v, c, xbase, x_plane, ws64);
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_BF16X3>, g, blk, 0, s, v, b, b_plane, ws64);
  }
  hipLaunchKernelGGL(b_coulomb_unpack_kernel, g, blk, 0, s, v, ws64, j);
  return hipGetLastError();
}

hipError_t launch_batch_coulomb_charges(BatchView v, int precision, const void* b, long long b_plane,
                                        const double* charges, double* ws64, double* j, int grid, hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  hipLaunchKernelGGL(b_charges_load_kernel, g, blk, 0, s, v, charges, ws64);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP64>, g, blk, 0, s, v, b, b_plane, ws64);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_FP32>, g, blk, 0, s, v, b, b_plane, ws64);
  else
    hipLaunchKernelGGL(b_coulomb_kernel<HFJK_BF16X3>, g, blk, 0, s, v, b, b_plane, ws64);
  hipLaunchKernelGGL(b_coulomb_unpack_kernel, g, blk, 0, s, v, ws64, j);
  return hipGetLastError();
}

hipError_t launch_batch_orbitals(BatchView v, int precision, const double* c, double* ws64, void* cbase,
                                 long long c_plane, int grid, hipStream_t s) {
  const dim3 g((unsigned)grid), blk(kThreads);
  if (precision == HFJK_FP64)
    hipLaunchKernelGGL(b_orbitals_kernel<HFJK_FP64>, g, blk, 0, s, v, c, ws64, cbase, c_plane);
  else if (precision == HFJK_FP32)
    hipLaunchKernelGGL(b_orbitals_kernel<HFJK_FP32>, g, blk, 0, s, v, c, ws64, cbase, c_plane);
  else
    hipLaunchKernelGGL(b_orbitals_kernel<HFJK_BF16X3>, g, blk, 0, s, v, c, ws64, cbase, c_plane);
  return hipGetLastError();
}

}  // namespace hfjk

// ---------------------------------------------------------------- C ABI

namespace {

size_t al256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int hfjk_batch_create(hfjk_ctx* ctx, const hfjk_batch_instance* instances, int ninst, int precision, int tile,
                      hfjk_batch** out) {
  if (out != nullptr) *out = nullptr;
  if (ctx == nullptr || out == nullptr || (ninst > 0 && instances == nullptr) || ninst < 1) {
    hfjk_internal_set_error("hfjk_batch_create needs a context, an output pointer and at least one instance");
    return HFJK_ERR_INPUT;
  }
  if (precision != HFJK_FP64 && precision != HFJK_FP32 && precision != HFJK_BF16X3) {
    hfjk_internal_set_error("precision must be HFJK_FP64, HFJK_FP32 or HFJK_BF16X3");
    return HFJK_ERR_INPUT;
  }
  if (tile != 0 && tile != 32 && tile != 64 && tile != 128 && tile != 256) {
    hfjk_internal_set_error("the half-transform column tile is 0, 32, 64, 128 or 256");
    return HFJK_ERR_INPUT;
  }
  hfjk_batch* bt = new (std::nothrow) hfjk_batch();
  if (bt == nullptr) {
    hfjk_internal_set_error("host allocation failed");
    return HFJK_ERR_HIP;
  }
  bt->ctx = ctx;
  bt->precision = precision;
  bt->ninst = ninst;
  bt->insts.resize(ninst);
  int max_ldo = 32;
  for (int i = 0; i < ninst; ++i) max_ldo = std::max(max_ldo, hfjk::occ_ld(std::max(instances[i].nocc, 1)));
  bt->tile = tile != 0 ? tile : (max_ldo <= 32 ? 32 : max_ldo <= 64 ? 64 : 128);
  const int es = precision == HFJK_FP64 ? 8 : precision == HFJK_FP32 ? 4 : 2;  // bf16: per plane
  long long w64 = 0, cp = 0, xx = 0, kp = 0, qq = 0;
  for (int i = 0; i < ninst; ++i) {
    const hfjk_batch_instance& src = instances[i];
    if (src.nao < 1 || src.naux < 1 || src.nocc < 1 || src.nocc > src.nao || src.nao > 46000 ||
        src.b_offset < 0 || src.b_offset % hfjk::kTileElems != 0 || src.c_offset < 0 || src.d_offset < 0) {
      delete bt;
      hfjk_internal_set_error(("instance " + std::to_string(i) +
                               " needs 1 <= nocc <= nao <= 46000, naux >= 1, a tensor offset that is a multiple"
                               " of 1024 and non-negative offsets").c_str());
      return HFJK_ERR_INPUT;
    }
    hfjk::BatchInst& in = bt->insts[i];
    in.nao = src.nao;
    in.naux = src.naux;
    in.nocc = src.nocc;
    in.ldo = hfjk::occ_ld(src.nocc);
    in.nt = hfjk::n_tiles(src.nao);
    in.xrows = hfjk::x_rows(src.nao);
    // SYRK tile: 128 rows, on the direct-LDS ring where the instance's X offsets fit it. (Small
    // instances compute mostly padding in 128-row tiles, but the ring runs the 128-row tile near
    // the MFMA peak, about twice the rate of the 64-row register-staged tile, so it wins on the
    // whole batch; the 64-row tile serves only instances too large for the ring.)
    {
      const bool fits_ring = in.ldo <= 2048 && (long long)in.xrows * in.ldo * es < (1LL << 31);
      in.st = fits_ring || precision == HFJK_BF16X3 ? hfjk::syrk_tile() : 64;
      const long long nb = (in.xrows + in.st - 1) / in.st;
      in.npairs = (int)(nb * (nb + 1) / 2);
    }
    // Half-transform column tile: the smallest of 32 / 64 / 128 that covers the instance's
    // columns, at most the batch's tile (whose multiples then cover them); a batch tile of 256
    // applies to every instance.
    if (bt->tile == 256) {
      in.bn = 256;
    } else {
      in.bn = in.ldo <= 32 ? 32 : in.ldo <= 64 ? 64 : 128;
      in.bn = std::min(in.bn, bt->tile);
    }
    in.nsplit = (int)std::max(1LL, std::min(64LL, ((long long)src.naux * in.ldo + 8191) / 8192));
    in.bm = bt->tile == 256 ? hfjk::batch_tile_rows(precision, 0, 256) : hfjk::batch_tile_rows(precision, in.xrows, in.bn);
    // Items cover enough aux functions for about 32 pipelined k tiles (small instances have
    // few k tiles per aux function).
    in.qg = bt->tile == 256 ? 1 : std::max(1, std::min(src.naux, 32 / in.nt));
    in.nbm = (in.xrows + in.bm - 1) / in.bm;
    in.nbn = (in.ldo + in.bn - 1) / in.bn;
    in.p = hfjk::packed_elems(src.nao);
    in.b_off = src.b_offset;
    in.c_off = src.c_offset;
    in.d_off = src.d_offset;
    in.q_off = qq;
    qq += src.naux;
    bt->b_total = std::max(bt->b_total, src.b_offset + (long long)src.naux * in.p);
    bt->c_total = std::max(bt->c_total, src.c_offset + (long long)src.nao * src.nocc);
    bt->d_total = std::max(bt->d_total, src.d_offset + (long long)src.nao * src.nao);
    const long long ntiles = in.p / hfjk::kTileElems;
    auto take64 = [&](long long n) { const long long at = w64; w64 += (n + 31) / 32 * 32; return at; };
    in.w_off = take64(in.p);
    in.jp_off = take64(in.p);
    in.gp_off = take64(ntiles * src.naux);
    in.g_off = take64(src.naux);
    in.sg_off = take64(in.ldo);
    in.cp_off = cp;
    cp += (long long)in.nt * hfjk::kTile * in.ldo + 64;
    in.x_off = xx;
    xx += (long long)src.naux * in.xrows * in.ldo + 64;
    in.kp_off = kp;
    kp += (long long)in.nsplit * in.npairs * in.st * in.st;
  }
  bt->ws64_total = w64;
  bt->cp_total = cp;
  bt->x_total = xx;
  bt->kp_total = kp;
  // Items per instance and phase.
  bt->prefix.assign((size_t)hfjk::kPhCount * (ninst + 1), 0);
  for (int i = 0; i < ninst; ++i) {
    const hfjk::BatchInst& in = bt->insts[i];
    const long long ntiles = in.p / hfjk::kTileElems;
    const long long nn = (long long)in.nao * in.nao;
    long long items[hfjk::kPhCount];
    items[hfjk::kPhWeights] = ntiles;
    items[hfjk::kPhGamma] = ntiles * ((in.naux + 63) / 64);
    items[hfjk::kPhGammaReduce] = (in.naux + hfjk::kThreads - 1) / hfjk::kThreads;
    items[hfjk::kPhCoulomb] = ntiles;
    items[hfjk::kPhCoulombUnpack] = (nn + hfjk::kThreads - 1) / hfjk::kThreads;
    items[hfjk::kPhOrbitals] = ((long long)in.nt * hfjk::kTile * in.ldo + hfjk::kThreads - 1) / hfjk::kThreads;
    items[hfjk::kPhXform] = (long long)((in.naux + in.qg - 1) / in.qg) * in.nbm * in.nbn;
    // Instances whose X offsets fit the ring's 32-bit lane offsets run on it, in a phase of
    // their own.
    const bool ring = in.ldo <= 2048 && (long long)in.xrows * in.ldo * es < (1LL << 31);
    items[hfjk::kPhSyrk] = ring ? 0 : (long long)in.npairs * in.nsplit;
    items[hfjk::kPhSyrkRing] = ring ? (long long)in.npairs * in.nsplit : 0;
    items[hfjk::kPhSyrkReduce] = (nn + hfjk::kThreads - 1) / hfjk::kThreads;
    items[hfjk::kPhCharges] = in.naux;
    for (int ph = 0; ph < hfjk::kPhCount; ++ph) {
      const size_t base = (size_t)ph * (ninst + 1);
      bt->prefix[base + i + 1] = bt->prefix[base + i] + items[ph];
    }
  }
  // Workspace layout.
  size_t off = 0;
  bt->off_insts = off;
  off = al256(off + sizeof(hfjk::BatchInst) * ninst);
  bt->off_prefix = off;
  off = al256(off + sizeof(long long) * bt->prefix.size());
  bt->off_counters = off;
  off = al256(off + sizeof(unsigned long long) * hfjk::kPhCount * 8);
  bt->off_ws64 = off;
  off = al256(off + sizeof(double) * (size_t)w64);
  const int planes = precision == HFJK_BF16X3 ? 2 : 1;
  bt->off_cp = off;
  off = al256(off + (size_t)es * planes * cp);
  bt->off_x = off;
  off = al256(off + (size_t)es * planes * xx);
  bt->off_kp = off;
  off = al256(off + (size_t)(precision == HFJK_FP64 ? 8 : 4) * kp);
  bt->bytes = off;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) prop.multiProcessorCount = 256;
  const int cus = std::max(8, prop.multiProcessorCount);
  bt->grid_mfma = ((2 * cus + 7) / 8) * 8;
  bt->grid_elem = ((4 * cus + 7) / 8) * 8;
  *out = bt;
  return HFJK_OK;
}

int hfjk_batch_workspace_bytes(const hfjk_batch* batch, size_t* bytes) {
  if (batch == nullptr || bytes == nullptr) {
    hfjk_internal_set_error("null batch or output pointer");
    return HFJK_ERR_INPUT;
  }
  *bytes = batch->bytes;
  return HFJK_OK;
}

int hfjk_batch_sizes(const hfjk_batch* batch, long long* tensor_elems, long long* orbital_elems,
                     long long* matrix_elems) {
  if (batch == nullptr) {
    hfjk_internal_set_error("null batch");
    return HFJK_ERR_INPUT;
  }
  if (tensor_elems) *tensor_elems = batch->b_total;
  if (orbital_elems) *orbital_elems = batch->c_total;
  if (matrix_elems) *matrix_elems = batch->d_total;
  return HFJK_OK;
}

int hfjk_batch_bind(hfjk_batch* batch, void* workspace, size_t bytes, void* stream) {
  if (batch == nullptr || workspace == nullptr) {
    hfjk_internal_set_error("null batch or workspace");
    return HFJK_ERR_INPUT;
  }
  if (bytes < batch->bytes || ((uintptr_t)workspace & 255) != 0) {
    hfjk_internal_set_error(("the batch needs a 256-byte aligned workspace of " + std::to_string(batch->bytes) +
                             " bytes").c_str());
    return HFJK_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  hipError_t e = hipSetDevice(batch->ctx->device);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ws + batch->off_insts, batch->insts.data(), sizeof(hfjk::BatchInst) * batch->ninst,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ws + batch->off_prefix, batch->prefix.data(), sizeof(long long) * batch->prefix.size(),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    hfjk_internal_set_error((std::string("uploading the batch plan: ") + hipGetErrorString(e)).c_str());
    return HFJK_ERR_HIP;
  }
  batch->ws = workspace;
  return HFJK_OK;
}

int hfjk_batch_run(hfjk_batch* batch, const void* tensor, const double* orbitals, const double* density,
                   double* coulomb, double* exchange, void* stream) {
  if (batch == nullptr || batch->ws == nullptr) {
    hfjk_internal_set_error("the batch is not bound to a workspace (hfjk_batch_bind)");
    return HFJK_ERR_INPUT;
  }
  if (tensor == nullptr || orbitals == nullptr || coulomb == nullptr || exchange == nullptr) {
    hfjk_internal_set_error("hfjk_batch_run needs the tensor, orbitals and both outputs");
    return HFJK_ERR_INPUT;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)batch->ws;
  hfjk::BatchView v;
  v.insts = (const hfjk::BatchInst*)(ws + batch->off_insts);
  v.prefix = (const long long*)(ws + batch->off_prefix);
  v.counters = (unsigned long long*)(ws + batch->off_counters);
  v.ninst = batch->ninst;
  double* ws64 = (double*)(ws + batch->off_ws64);
  void* cbase = ws + batch->off_cp;
  void* xbase = ws + batch->off_x;
  void* kbase = ws + batch->off_kp;
  const int prec = batch->precision;
  hipError_t e = hipSetDevice(batch->ctx->device);
  if (e == hipSuccess) e = hfjk::launch_batch_reset(v, s);
  if (e == hipSuccess && density != nullptr)
    e = hfjk::launch_batch_coulomb(v, prec, tensor, batch->b_total, density, ws64, coulomb, batch->grid_elem, s);
  if (e == hipSuccess)
    e = hfjk::launch_batch_orbitals(v, prec, orbitals, ws64, cbase, batch->cp_total, batch->grid_elem, s);
  if (e == hipSuccess)
    e = hfjk::launch_batch_xform(v, prec, batch->tile, tensor, batch->b_total, cbase, batch->cp_total, xbase,
                                 batch->x_total, batch->grid_mfma, s);
  // Without densities D = C C^T: the charges come from X (one pass over X instead of one over B).
  if (e == hipSuccess && density == nullptr)
    e = hfjk::launch_batch_coulomb_occ(v, prec, tensor, batch->b_total, orbitals, xbase, batch->x_total, ws64,
                                       coulomb, batch->grid_elem, s);
  if (e == hipSuccess)
    e = hfjk::launch_batch_syrk(v, prec, xbase, batch->x_total, ws64, kbase, batch->grid_mfma, s);
  if (e == hipSuccess) e = hfjk::launch_batch_syrk_reduce(v, prec, kbase, exchange, batch->grid_elem, s);
  if (e != hipSuccess) {
    hfjk_internal_set_error((std::string("batched launch: ") + hipGetErrorString(e)).c_str());
    return HFJK_ERR_HIP;
  }
  return HFJK_OK;
}

int hfjk_batch_coulomb_charges(hfjk_batch* batch, const void* tensor, const double* charges, double* coulomb,
                               void* stream) {
  if (batch == nullptr || batch->ws == nullptr) {
    hfjk_internal_set_error("the batch is not bound to a workspace (hfjk_batch_bind)");
    return HFJK_ERR_INPUT;
  }
  if (tensor == nullptr || charges == nullptr || coulomb == nullptr) {
    hfjk_internal_set_error("hfjk_batch_coulomb_charges needs the tensor, the charges and the output");
    return HFJK_ERR_INPUT;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)batch->ws;
  hfjk::BatchView v;
  v.insts = (const hfjk::BatchInst*)(ws + batch->off_insts);
  v.prefix = (const long long*)(ws + batch->off_prefix);
  v.counters = (unsigned long long*)(ws + batch->off_counters);
  v.ninst = batch->ninst;
  double* ws64 = (double*)(ws + batch->off_ws64);
  hipError_t e = hipSetDevice(batch->ctx->device);
  if (e == hipSuccess) e = hfjk::launch_batch_reset(v, s);
  if (e == hipSuccess)
    e = hfjk::launch_batch_coulomb_charges(v, batch->precision, tensor, batch->b_total, charges, ws64, coulomb,
                                           batch->grid_elem, s);
  if (e != hipSuccess) {
    hfjk_internal_set_error((std::string("batched launch: ") + hipGetErrorString(e)).c_str());
    return HFJK_ERR_HIP;
  }
  return HFJK_OK;
}

int hfjk_batch_destroy(hfjk_batch* batch) {
  delete batch;
  return HFJK_OK;
}

}  // extern "C"
