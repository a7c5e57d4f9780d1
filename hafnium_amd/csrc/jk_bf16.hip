// BF16x3 exchange kernels for gfx950: every FP32 operand is carried as a
// bf16 pair (hi = bf16(x), lo = bf16(x - hi)) and each product is
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with FP32 accumulation
// (the lo*lo term, 2^-16 relative, is dropped).
//
// v_mfma_f32_16x16x32_bf16 lane layout (measured on gfx950): A[i][k] with
// i = lane % 16 and k = 8 (lane / 16) + e in element e of the lane's 8-vector;
// B[k][j] with j = lane % 16 and the same k; D register r of lane l holds
// D[4 (l / 16) + r][l % 16]. Both operands are therefore staged in LDS with
// k contiguous ([row][k]), so each lane reads its 8 values as one 16-byte word.
#include "hfjk_internal.h"

namespace hfjk {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kPadK = 40;  // LDS row length in bf16 (32 + 8): 80-byte rows keep 16-byte alignment

__device__ __forceinline__ long long tile_index16(int r, int c) { return (long long)r * (r + 1) / 2 + c; }

__device__ __forceinline__ void tile_rc16(int tile, int& r, int& c) {
  int rr = (int)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
  while ((rr + 1) * (rr + 2) / 2 <= tile) ++rr;
  while (rr * (rr + 1) / 2 > tile) --rr;
  r = rr;
  c = tile - rr * (rr + 1) / 2;
}

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ bf8 lds8(const uint16_t* p) { return *(const bf8*)p; }

__device__ __forceinline__ f4 mma3(bf8 ah, bf8 al, bf8 bh, bf8 bl, f4 acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// Orbitals [nao][nocc] float64 -> transposed split planes Ct[ldo][npad]
// (hi, then lo), zero-padded.
__global__ __launch_bounds__(kThreads) void orbitals_bf16x3_kernel(const double* __restrict__ c, int nao, int nocc,
                                                                    int npad, int ldo, uint16_t* __restrict__ hi,
                                                                    uint16_t* __restrict__ lo) {
  const long long total = (long long)npad * ldo;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / npad);
    const int k = (int)(idx % npad);
    const float v = (k < nao && i < nocc) ? (float)c[(long long)k * nocc + i] : 0.0f;
    const uint16_t h = to_bf16(v);
    hi[idx] = h;
    lo[idx] = to_bf16(v - from_bf16(h));
  }
}

// X[q][m][i] = sum_k B[q][m][k] C[k][i] in BF16x3, written as bf16 hi/lo
// planes. Workgroup tile 64 rows x BN columns, waves 2 x 2; one layout tile
// column (32) is exactly one MFMA k step.
template <int BN>
__global__ __launch_bounds__(kThreads) void half_transform_bf16x3_kernel(
    const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo, long long p, int nt,
    const uint16_t* __restrict__ cthi, const uint16_t* __restrict__ ctlo, int npad, int ldo,
    uint16_t* __restrict__ xhi, uint16_t* __restrict__ xlo, int xrows, int qbase) {
  constexpr int NI = BN / 32;
  __shared__ __attribute__((aligned(16))) uint16_t Ah[kBlock][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Al[kBlock][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Bh[BN][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Bl[BN][kPadK];
  const int nb = blockIdx.x;
  const int mb = blockIdx.y;
  const int q = qbase + blockIdx.z;
  const int n0 = nb * BN;
  const uint16_t* hq = bhi + (long long)q * p;
  const uint16_t* lq = blo + (long long)q * p;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  f4 acc[2][NI];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  for (int kt = 0; kt < nt; ++kt) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 2 * mb + h;
      const int e = t * 4;  // 4 consecutive elements of the 32 x 32 tile
      const int a = e / kTile;
      const int bc = e % kTile;
      if (r >= nt) {
        *(uint2*)&Ah[h * kTile + a][bc] = make_uint2(0u, 0u);
        *(uint2*)&Al[h * kTile + a][bc] = make_uint2(0u, 0u);
      } else if (r >= kt) {
        // tile (r, kt): element (a, bc) is B[m = r*32 + a][k = kt*32 + bc]
        const long long off = tile_index16(r, kt) * kTileElems + e;
        *(uint2*)&Ah[h * kTile + a][bc] = *(const uint2*)(hq + off);
        *(uint2*)&Al[h * kTile + a][bc] = *(const uint2*)(lq + off);
      } else {
        // tile (kt, r): element (a, bc) is B[k = kt*32 + a][m = r*32 + bc]
        const long long off = tile_index16(kt, r) * kTileElems + e;
        const ushort4 vh = *(const ushort4*)(hq + off);
        const ushort4 vl = *(const ushort4*)(lq + off);
        Ah[h * kTile + bc + 0][a] = vh.x;
        Ah[h * kTile + bc + 1][a] = vh.y;
        Ah[h * kTile + bc + 2][a] = vh.z;
        Ah[h * kTile + bc + 3][a] = vh.w;
        Al[h * kTile + bc + 0][a] = vl.x;
        Al[h * kTile + bc + 1][a] = vl.y;
        Al[h * kTile + bc + 2][a] = vl.z;
        Al[h * kTile + bc + 3][a] = vl.w;
      }
    }
    // C^T rows n0 .. n0+BN-1, k columns kt*32 .. +31, in 4-element chunks.
    for (int idx = t; idx < BN * (kTile / 4); idx += kThreads) {
      const int jj = idx / (kTile / 4);
      const int k4 = (idx % (kTile / 4)) * 4;
      const int col = n0 + jj;
      uint2 vh = make_uint2(0u, 0u), vl = make_uint2(0u, 0u);
      if (col < ldo) {
        const long long off = (long long)col * npad + kt * kTile + k4;
        vh = *(const uint2*)(cthi + off);
        vl = *(const uint2*)(ctlo + off);
      }
      *(uint2*)&Bh[jj][k4] = vh;
      *(uint2*)&Bl[jj][k4] = vl;
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    bf8 ah[2], al[2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
      ah[mi] = lds8(&Ah[wm * 32 + mi * 16 + (lane & 15)][kq]);
      al[mi] = lds8(&Al[wm * 32 + mi * 16 + (lane & 15)][kq]);
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const bf8 bh = lds8(&Bh[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
      const bf8 bl = lds8(&Bl[wn * (BN / 2) + ni * 16 + (lane & 15)][kq]);
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) acc[mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mi][ni]);
    }
    __syncthreads();
  }
  const long long xbase = (long long)q * xrows * ldo;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      const int col = n0 + wn * (BN / 2) + ni * 16 + (lane & 15);
      if (col < ldo) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mb * kBlock + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
          const float x = acc[mi][ni][r];
          const uint16_t h = to_bf16(x);
          const long long off = xbase + (long long)row * ldo + col;
          xhi[off] = h;
          xlo[off] = to_bf16(x - from_bf16(h));
        }
      }
    }
}

// partial[s][pair][64][64] = sum over the s-th aux slice and all i of
// X[q][mb*64 + .][i] sign[i] X[q][nb*64 + .][i] in BF16x3.
__global__ __launch_bounds__(kThreads) void syrk_bf16x3_kernel(const uint16_t* __restrict__ xhi,
                                                                const uint16_t* __restrict__ xlo, int naux,
                                                                int xrows, int ldo, const double* __restrict__ signs,
                                                                int nsplit, int npairs, float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) uint16_t Ah[kBlock][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Al[kBlock][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Bh[kBlock][kPadK];
  __shared__ __attribute__((aligned(16))) uint16_t Bl[kBlock][kPadK];
  const int pair = blockIdx.x;
  const int s = blockIdx.y;
  int mb, nb;
  tile_rc16(pair, mb, nb);
  const int qb = (int)((long long)s * naux / nsplit);
  const int qe = (int)((long long)(s + 1) * naux / nsplit);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  f4 acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  for (int q = qb; q < qe; ++q) {
    const long long rm = ((long long)q * xrows + mb * kBlock) * ldo;
    const long long rn = ((long long)q * xrows + nb * kBlock) * ldo;
    for (int i0 = 0; i0 < ldo; i0 += kTile) {
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int chunk = t + kThreads * v;  // 512 chunks of 4 elements: 64 rows x 8
        const int row = chunk / (kTile / 4);
        const int c4 = (chunk % (kTile / 4)) * 4;
        const long long om = rm + (long long)row * ldo + i0 + c4;
        const long long on = rn + (long long)row * ldo + i0 + c4;
        *(uint2*)&Ah[row][c4] = *(const uint2*)(xhi + om);
        *(uint2*)&Al[row][c4] = *(const uint2*)(xlo + om);
        uint2 nh = *(const uint2*)(xhi + on);
        uint2 nl = *(const uint2*)(xlo + on);
        // Negative signs flip the sign bit of each bf16 of both planes.
        const uint32_t f0 = (signs[i0 + c4 + 0] < 0.0 ? 0x8000u : 0u) | (signs[i0 + c4 + 1] < 0.0 ? 0x80000000u : 0u);
        const uint32_t f1 = (signs[i0 + c4 + 2] < 0.0 ? 0x8000u : 0u) | (signs[i0 + c4 + 3] < 0.0 ? 0x80000000u : 0u);
        nh.x ^= f0;
        nh.y ^= f1;
        nl.x ^= f0;
        nl.y ^= f1;
        *(uint2*)&Bh[row][c4] = nh;
        *(uint2*)&Bl[row][c4] = nl;
      }
      __syncthreads();
      const int kq = 8 * (lane >> 4);
      bf8 ah[2], al[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) {
        ah[mi] = lds8(&Ah[wm * 32 + mi * 16 + (lane & 15)][kq]);
        al[mi] = lds8(&Al[wm * 32 + mi * 16 + (lane & 15)][kq]);
      }
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const bf8 bh = lds8(&Bh[wn * 32 + ni * 16 + (lane & 15)][kq]);
        const bf8 bl = lds8(&Bl[wn * 32 + ni * 16 + (lane & 15)][kq]);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) acc[mi][ni] = mma3(ah[mi], al[mi], bh, bl, acc[mi][ni]);
      }
      __syncthreads();
    }
  }
  float* dst = partial + ((long long)s * npairs + pair) * (kBlock * kBlock);
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      const int col = wn * 32 + ni * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
        dst[row * kBlock + col] = acc[mi][ni][r];
      }
    }
}

template <int BN>
hipError_t launch_x16(const uint16_t* bhi, const uint16_t* blo, long long p, int naux, int nao, const uint16_t* cthi,
                      const uint16_t* ctlo, int ldo, uint16_t* xhi, uint16_t* xlo, hipStream_t s) {
  const int nt = n_tiles(nao);
  const int npad = nt * kTile;
  const int xrows = x_rows(nao);
  for (int base = 0; base < naux; base += 65535) {
    const int count = naux - base < 65535 ? naux - base : 65535;
    const dim3 grid((unsigned)((ldo + BN - 1) / BN), (unsigned)(xrows / kBlock), (unsigned)count);
    hipLaunchKernelGGL((half_transform_bf16x3_kernel<BN>), grid, dim3(kThreads), 0, s, bhi, blo, p, nt, cthi, ctlo,
                       npad, ldo, xhi, xlo, xrows, base);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

}  // namespace

hipError_t launch_orbitals_bf16x3(const double* c, int nao, int nocc, void* out, hipStream_t s) {
  const int npad = n_tiles(nao) * kTile;
  const int ldo = occ_ld(nocc);
  const long long n = (long long)npad * ldo;
  uint16_t* hi = (uint16_t*)out;
  hipLaunchKernelGGL(orbitals_bf16x3_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, c, nao, nocc, npad, ldo, hi,
                     hi + n);
  return hipGetLastError();
}

hipError_t launch_half_transform_bf16x3(const uint16_t* b, long long p, int naux, int nao, const void* ct, int ldo,
                                        void* x, hipStream_t s) {
  const uint16_t* bhi = b;
  const uint16_t* blo = b + (long long)naux * p;
  const long long nct = (long long)n_tiles(nao) * kTile * ldo;
  const uint16_t* cthi = (const uint16_t*)ct;
  const uint16_t* ctlo = cthi + nct;
  const long long nx = (long long)naux * x_rows(nao) * ldo;
  uint16_t* xhi = (uint16_t*)x;
  uint16_t* xlo = xhi + nx;
  if (ldo <= 32) return launch_x16<32>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
  if (ldo <= 64) return launch_x16<64>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
  return launch_x16<128>(bhi, blo, p, naux, nao, cthi, ctlo, ldo, xhi, xlo, s);
}

hipError_t launch_syrk_bf16x3(const void* x, int naux, int nao, int ldo, const double* signs, int nsplit,
                              void* partial, hipStream_t s) {
  const int xrows = x_rows(nao);
  const int nb = xrows / kBlock;
  const int npairs = nb * (nb + 1) / 2;
  const long long nx = (long long)naux * xrows * ldo;
  const uint16_t* xhi = (const uint16_t*)x;
  hipLaunchKernelGGL(syrk_bf16x3_kernel, dim3((unsigned)npairs, (unsigned)nsplit), dim3(kThreads), 0, s, xhi,
                     xhi + nx, naux, xrows, ldo, signs, nsplit, npairs, (float*)partial);
  return hipGetLastError();
}

}  // namespace hfjk
