// Producers of the tiled packed fitted tensor: the seeded synthetic
// generator, packing of a full tensor, and the FP32 / BF16 hi-lo copies.
//
// The generator must reproduce hafnium_amd/synthetic.py bit for bit, so its
// arithmetic is a fixed sequence of IEEE additions, multiplications,
// divisions and square roots with FMA contraction switched off; the Philox
// rounds are integer-exact.
#include "hfjk_internal.h"

namespace hfjk {
namespace {

constexpr uint32_t kM0 = 0xD2511F53u;
constexpr uint32_t kM1 = 0xCD9E8D57u;
constexpr uint32_t kW0 = 0x9E3779B9u;
constexpr uint32_t kW1 = 0xBB67AE85u;
constexpr uint32_t kTagTensor = 0x4A4B0001u;

__device__ inline void philox(uint32_t c[4], uint64_t seed) {
  uint32_t k0 = (uint32_t)(seed & 0xFFFFFFFFull);
  uint32_t k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)kM0 * c[0];
    uint64_t p1 = (uint64_t)kM1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += kW0;
    k1 += kW1;
  }
}

// ln(u), u in (0, 1]: u = m 2^e, m in [1/sqrt2, sqrt2), ln m = 2 atanh(s).
__device__ inline double det_log(double u) {
#pragma clang fp contract(off)
  int e;
  double m = frexp(u, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m * 2.0;
    e -= 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 0x1.642c8590b2164p-5;  // 1/23
  p = p * s2 + 0x1.8618618618618p-5;  // 1/21
  p = p * s2 + 0x1.af286bca1af28p-5;  // 1/19
  p = p * s2 + 0x1.e1e1e1e1e1e1ep-5;  // 1/17
  p = p * s2 + 0x1.1111111111111p-4;  // 1/15
  p = p * s2 + 0x1.3b13b13b13b14p-4;  // 1/13
  p = p * s2 + 0x1.745d1745d1746p-4;  // 1/11
  p = p * s2 + 0x1.c71c71c71c71cp-4;  // 1/9
  p = p * s2 + 0x1.2492492492492p-3;  // 1/7
  p = p * s2 + 0x1.999999999999ap-3;  // 1/5
  p = p * s2 + 0x1.5555555555555p-2;  // 1/3
  p = p * s2 + 1.0;
  double ed = (double)e;
  return ed * 0x1.62e42fefa39efp-1 + (s + s) * p;
}

// sin(x), |x| <= pi/2, Taylor series to x^23.
__device__ inline double det_sin(double x) {
#pragma clang fp contract(off)
  double x2 = x * x;
  double p = -0x1.761b413163819p-75;
  p = p * x2 + 0x1.71b8ef6dcf572p-66;
  p = p * x2 + -0x1.2f49b46814157p-57;
  p = p * x2 + 0x1.952c77030ad4ap-49;
  p = p * x2 + -0x1.ae7f3e733b81fp-41;
  p = p * x2 + 0x1.6124613a86d09p-33;
  p = p * x2 + -0x1.ae64567f544e4p-26;
  p = p * x2 + 0x1.71de3a556c734p-19;
  p = p * x2 + -0x1.a01a01a01a01ap-13;
  p = p * x2 + 0x1.1111111111111p-7;
  p = p * x2 + -0x1.5555555555555p-3;
  p = p * x2 + 1.0;
  return x * p;
}

__device__ inline double det_cos_two_pi(double u) {
#pragma clang fp contract(off)
  double v = (u < 0.5) ? u : 1.0 - u;
  double w = v - 0.25;
  return -det_sin(0x1.921fb54442d18p+2 * w);
}

// exp(x), x <= 0: 2^-k exp(r), |r| <= ln2/2, Taylor series to r^17.
__device__ inline double det_exp_neg(double x) {
#pragma clang fp contract(off)
  double k = floor(-x / 0x1.62e42fefa39efp-1 + 0.5);
  double r = x + k * 0x1.62e42fefa39efp-1;
  double p = 0x1.952c77030ad4ap-49;
  p = p * r + 0x1.ae7f3e733b81fp-45;
  p = p * r + 0x1.ae7f3e733b81fp-41;
  p = p * r + 0x1.93974a8c07c9dp-37;
  p = p * r + 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, -(int)k);
}

__device__ inline double det_normal(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed) {
#pragma clang fp contract(off)
  uint32_t c[4] = {c0, c1, c2, c3};
  philox(c, seed);
  uint64_t a = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
  uint64_t b = ((uint64_t)(c[2] >> 5) << 26) | (uint64_t)(c[3] >> 6);
  double u1 = 1.0 - (double)a * 0x1.0p-53;
  double u2 = (double)b * 0x1.0p-53;
  double l = det_log(u1);
  return sqrt(-2.0 * l) * det_cos_two_pi(u2);
}

// One thread per packed element; grid (tile pairs * 4, naux): each block of
// 256 threads covers a quarter tile.
__global__ __launch_bounds__(kThreads) void generate_kernel(uint64_t seed, int q0, int nao, int nt,
                                                             long long p, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int q = blockIdx.y;
  const long long e = (long long)blockIdx.x * kThreads + threadIdx.x;  // element within aux function
  if (e >= p) return;
  const int tile = (int)(e / kTileElems);
  const int local = (int)(e % kTileElems);
  // tile = r (r + 1) / 2 + c with c <= r
  int r = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= tile) ++r;
  while (r * (r + 1) / 2 > tile) --r;
  const int c = tile - r * (r + 1) / 2;
  const int row = r * kTile + local / kTile;
  const int col = c * kTile + local % kTile;
  double value = 0.0;
  if (row < nao && col < nao) {
    const int lo = row < col ? row : col;
    const int hi = row < col ? col : row;
    const double z = det_normal((uint32_t)lo, (uint32_t)hi, (uint32_t)(q0 + q), kTagTensor, seed);
    const double env = det_exp_neg(-((double)(hi - lo) / 20.0));
    value = z * env;
  }
  (void)nt;
  out[(long long)q * p + e] = value;
}

__global__ __launch_bounds__(kThreads) void normal_matrix_kernel(uint64_t seed, int rows, int cols,
                                                                  uint32_t tag, double* __restrict__ out) {
  const long long total = (long long)rows * cols;
  for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < total; e += (long long)gridDim.x * kThreads) {
    const int r = (int)(e / cols);
    const int c = (int)(e % cols);
    out[e] = det_normal((uint32_t)r, (uint32_t)c, 0xFFFFFFFFu, tag, seed);
  }
}

// Pack: one block per (tile, q); 256 threads x 4 elements.
__global__ __launch_bounds__(kThreads) void pack_kernel(const double* __restrict__ full, int nao, long long p,
                                                         double* __restrict__ out,
                                                         unsigned long long* __restrict__ asym_bits) {
  // asym_bits[0] <- max |B - B^T|, asym_bits[1] <- max |B| (as ordered bit patterns)
  const int q = blockIdx.y;
  const int tile = blockIdx.x;
  int r = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= tile) ++r;
  while (r * (r + 1) / 2 > tile) --r;
  const int c = tile - r * (r + 1) / 2;
  const double* src = full + (long long)q * nao * nao;
  double worst = 0.0;
  double mag = 0.0;
  for (int v = 0; v < 4; ++v) {
    const int local = threadIdx.x * 4 + v;
    const int row = r * kTile + local / kTile;
    const int col = c * kTile + local % kTile;
    double value = 0.0;
    if (row < nao && col < nao) {
      value = src[(long long)row * nao + col];
      if (asym_bits != nullptr) {
        const double other = src[(long long)col * nao + row];
        const double d = fabs(value - other);
        worst = d > worst ? d : worst;
        mag = fabs(value) > mag ? fabs(value) : mag;
      }
    }
    out[(long long)q * p + (long long)tile * kTileElems + local] = value;
  }
  if (asym_bits != nullptr) {
    // Non-negative doubles order like their bit patterns.
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(worst, off);
      const double m = __shfl_xor(mag, off);
      worst = o > worst ? o : worst;
      mag = m > mag ? m : mag;
    }
    if ((threadIdx.x & 63) == 0) {
      if (worst > 0.0) atomicMax(&asym_bits[0], (unsigned long long)__double_as_longlong(worst));
      if (mag > 0.0) atomicMax(&asym_bits[1], (unsigned long long)__double_as_longlong(mag));
    }
  }
}

__device__ inline uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_float(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__global__ __launch_bounds__(kThreads) void split32_kernel(const double* __restrict__ in, long long count,
                                                            float* __restrict__ out) {
  for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < count; e += (long long)gridDim.x * kThreads)
    out[e] = (float)in[e];
}

__global__ __launch_bounds__(kThreads) void split16_kernel(const double* __restrict__ in, long long count,
                                                            uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
  for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < count; e += (long long)gridDim.x * kThreads) {
    const float x = (float)in[e];
    const uint16_t h = bf16_rne(x);
    hi[e] = h;
    lo[e] = bf16_rne(x - bf16_to_float(h));
  }
}

}  // namespace

hipError_t launch_generate(uint64_t seed, int q0, int naux, int nao, double* out, hipStream_t s) {
  const long long p = packed_elems(nao);
  const long long blocks = (p + kThreads - 1) / kThreads;
  if (naux <= 0) return hipSuccess;
  // gridDim.y is limited to 65535 aux functions per launch.
  for (int base = 0; base < naux; base += 65535) {
    const int count = naux - base < 65535 ? naux - base : 65535;
    hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks, (unsigned)count), dim3(kThreads), 0, s, seed,
                       q0 + base, nao, n_tiles(nao), p, out + (long long)base * p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

hipError_t launch_normal_matrix(uint64_t seed, int rows, int cols, uint32_t tag, double* out, hipStream_t s) {
  const long long n = (long long)rows * cols;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(normal_matrix_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s,
                     seed, rows, cols, tag, out);
  return hipGetLastError();
}

hipError_t launch_pack(const double* full, int naux, int nao, double* out, unsigned long long* asym_bits,
                       hipStream_t s) {
  const long long p = packed_elems(nao);
  const long long tiles = n_tile_pairs(nao);
  for (int base = 0; base < naux; base += 65535) {
    const int count = naux - base < 65535 ? naux - base : 65535;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)tiles, (unsigned)count), dim3(kThreads), 0, s,
                       full + (long long)base * nao * nao, nao, p, out + (long long)base * p, asym_bits);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

hipError_t launch_split(const double* in, long long count, int precision, void* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const unsigned blocks = grid_for(count);
  if (precision == HFJK_FP32) {
    hipLaunchKernelGGL(split32_kernel, dim3(blocks), dim3(kThreads), 0, s, in, count, (float*)out);
  } else {
    uint16_t* hi = (uint16_t*)out;
    hipLaunchKernelGGL(split16_kernel, dim3(blocks), dim3(kThreads), 0, s, in, count, hi, hi + count);
  }
  return hipGetLastError();
}

}  // namespace hfjk
