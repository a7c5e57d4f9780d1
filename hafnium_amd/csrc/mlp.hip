// Inference and training kernels of the learned SCF guess (a small MLP on
// gfx950): a BF16 MFMA GEMM with FP32 accumulation and a fused bias / SiLU
// epilogue, the SiLU backward, column sums (bias gradients), a weighted
// squared-error loss with its gradient, and the Adam update that keeps FP32
// master weights beside their BF16 copy.
//
// Arithmetic contract (mirrored by the CPU restatement the tests check it
// against): every GEMM operand is rounded to bf16 (round to nearest even)
// when it is loaded, products are accumulated in FP32, and everything
// outside the GEMMs runs in FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "hfjk_internal.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kT = 256;
constexpr int kTileM = 64;
constexpr int kTileN = 64;
constexpr int kStepK = 32;
constexpr int kPitch = kStepK + 8;  // LDS row pitch (bf16): 80 bytes keeps 16-byte rows aligned

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint16_t load_bf16(const float* p) { return to_bf16(*p); }
__device__ __forceinline__ uint16_t load_bf16(const uint16_t* p) { return *p; }

__device__ __forceinline__ bf8 lds_bf8(const uint16_t* p) {
  const uint4 v = *(const uint4*)p;
  return __builtin_bit_cast(bf8, v);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

struct GemmArgs {
  int m, n, k;
  const void* a;
  int lda, ta;
  const int* a_rows;
  const void* b;
  int ldb, tb;
  const float* bias;
  float* c;
  int ldc;
  int accumulate;
  float* h;
};

// C[m][n] = op(A) op(B) (+ C) (+ bias); h = silu(C) when h is given.
// op(A) is m x k: A is stored m x k (ta = 0) or k x m (ta = 1), row-major
// with leading dimension lda; storage row r of A is read from row a_rows[r]
// when a_rows is given (a gather of minibatch rows). op(B) is k x n: B is
// stored k x n (tb = 0) or n x k (tb = 1).
// One 256-thread workgroup per 64 x 64 tile of C; wave w owns the 32 x 32
// quarter (w / 2, w % 2) as 2 x 2 v_mfma_f32_16x16x32_bf16 tiles
// (A[i][8 (l / 16) + e] in element e of lane l, i = l % 16; B likewise with
// j = l % 16; D register r of lane l = D[4 (l / 16) + r][l % 16]).
template <typename TA, typename TB>
__global__ __launch_bounds__(kT) void mlp_gemm_kernel(GemmArgs g) {
  __shared__ __align__(16) uint16_t as[kTileM][kPitch];
  __shared__ __align__(16) uint16_t bs[kTileN][kPitch];
  const TA* __restrict__ a = (const TA*)g.a;
  const TB* __restrict__ b = (const TB*)g.b;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = blockIdx.y * kTileM;
  const int col0 = blockIdx.x * kTileN;
  f4 acc[2][2];
  for (int mi = 0; mi < 2; ++mi)
    for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = f4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.k; k0 += kStepK) {
    // A tile -> as[i][kk] (bf16).
    if (!g.ta) {
      const int i = t >> 2, kk = (t & 3) * 8;
      const int gi = row0 + i;
      const long long sr = gi < g.m ? (g.a_rows ? g.a_rows[gi] : gi) : 0;
      for (int e = 0; e < 8; ++e) {
        const int gk = k0 + kk + e;
        as[i][kk + e] = (gi < g.m && gk < g.k) ? load_bf16(a + sr * g.lda + gk) : (uint16_t)0;
      }
    } else {
      const int kk = t >> 3, i = (t & 7) * 8;
      const int gk = k0 + kk;
      const long long sr = gk < g.k ? (g.a_rows ? g.a_rows[gk] : gk) : 0;
      for (int e = 0; e < 8; ++e) {
        const int gi = row0 + i + e;
        as[i + e][kk] = (gi < g.m && gk < g.k) ? load_bf16(a + sr * g.lda + gi) : (uint16_t)0;
      }
    }
    // B tile -> bs[j][kk] (bf16, transposed so that a lane reads 8 consecutive k).
    if (!g.tb) {
      const int kk = t >> 3, j = (t & 7) * 8;
      const int gk = k0 + kk;
      for (int e = 0; e < 8; ++e) {
        const int gj = col0 + j + e;
        bs[j + e][kk] = (gk < g.k && gj < g.n) ? load_bf16(b + (long long)gk * g.ldb + gj) : (uint16_t)0;
      }
    } else {
      const int j = t >> 2, kk = (t & 3) * 8;
      const int gj = col0 + j;
      for (int e = 0; e < 8; ++e) {
        const int gk = k0 + kk + e;
        bs[j][kk + e] = (gk < g.k && gj < g.n) ? load_bf16(b + (long long)gj * g.ldb + gk) : (uint16_t)0;
      }
    }
    __syncthreads();
    bf8 fa[2], fb[2];
    for (int mi = 0; mi < 2; ++mi) fa[mi] = lds_bf8(&as[wr * 32 + mi * 16 + (lane & 15)][8 * (lane >> 4)]);
    for (int ni = 0; ni < 2; ++ni) fb[ni] = lds_bf8(&bs[wc * 32 + ni * 16 + (lane & 15)][8 * (lane >> 4)]);
    for (int mi = 0; mi < 2; ++mi)
      for (int ni = 0; ni < 2; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
    __syncthreads();
  }

  for (int mi = 0; mi < 2; ++mi)
    for (int ni = 0; ni < 2; ++ni)
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wr * 32 + mi * 16 + 4 * (lane >> 4) + r;
        const int j = col0 + wc * 32 + ni * 16 + (lane & 15);
        if (i < g.m && j < g.n) {
          float v = acc[mi][ni][r];
          const long long at = (long long)i * g.ldc + j;
          if (g.accumulate) v += g.c[at];
          if (g.bias) v += g.bias[j];
          g.c[at] = v;
          if (g.h) g.h[at] = silu(v);
        }
      }
}

// grad_z = grad_h * silu'(z), silu'(z) = s (1 + z (1 - s)), s = 1 / (1 + e^-z).
__global__ __launch_bounds__(kT) void silu_backward_kernel(long long count, const float* __restrict__ grad_h,
                                                           const float* __restrict__ z, float* __restrict__ grad_z) {
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < count; e += (long long)gridDim.x * kT) {
    const float x = z[e];
    const float s = 1.0f / (1.0f + expf(-x));
    grad_z[e] = grad_h[e] * (s * (1.0f + x * (1.0f - s)));
  }
}

// out[j] = sum_i x[i][j]: 64 columns per workgroup, 4 row phases, fixed order.
__global__ __launch_bounds__(kT) void column_sum_kernel(int m, int n, const float* __restrict__ x, int ldx,
                                                        float* __restrict__ out) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63;
  const int phase = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  float s = 0.f;
  if (j < n)
    for (int i = phase; i < m; i += 4) s += x[(long long)i * ldx + j];
  part[phase][c] = s;
  __syncthreads();
  if (phase == 0 && j < n) out[j] = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
}

// Weighted squared error over y [m][n]: loss[0] = scale sum_ij w_j (y_ij - t_ij)^2,
// grad[i][j] = 2 scale w_j (y_ij - t_ij); target row i is t_rows[i] when
// given. One workgroup, fixed summation order.
__global__ __launch_bounds__(kT) void mse_kernel(int m, int n, const float* __restrict__ y, int ldy,
                                                 const float* __restrict__ tgt, int ldt, const int* __restrict__ t_rows,
                                                 const float* __restrict__ w, float scale, float* __restrict__ grad,
                                                 float* __restrict__ loss) {
  __shared__ float part[kT];
  float s = 0.f;
  const long long total = (long long)m * n;
  for (long long e = threadIdx.x; e < total; e += kT) {
    const int i = (int)(e / n), j = (int)(e % n);
    const long long tr = t_rows ? t_rows[i] : i;
    const float d = y[(long long)i * ldy + j] - tgt[tr * ldt + j];
    const float wj = w ? w[j] : 1.0f;
    s += wj * d * d;
    if (grad) grad[(long long)i * ldy + j] = 2.0f * scale * wj * d;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int half = kT / 2; half > 0; half >>= 1) {
    if (threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = scale * part[0];
}

// Adam (decoupled weight decay): m, v moments; p -= lr_t m / (sqrt(v) + eps) + lr wd p;
// the bf16 copy of p is refreshed in the same pass.
// The host forms 1 - beta in double (1 - 0.999f in float loses 4 digits).
__global__ __launch_bounds__(kT) void adam_kernel(long long count, float* __restrict__ p, const float* __restrict__ g,
                                                  float* __restrict__ m1, float* __restrict__ m2,
                                                  uint16_t* __restrict__ p16, float lr_t, float beta1,
                                                  float one_minus_beta1, float beta2, float one_minus_beta2,
                                                  float eps, float decay) {
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < count; e += (long long)gridDim.x * kT) {
    const float gi = g[e];
    const float a = beta1 * m1[e] + one_minus_beta1 * gi;
    const float b = beta2 * m2[e] + one_minus_beta2 * gi * gi;
    m1[e] = a;
    m2[e] = b;
    float x = p[e];
    x -= lr_t * a / (sqrtf(b) + eps) + decay * x;
    p[e] = x;
    if (p16) p16[e] = to_bf16(x);
  }
}

__global__ __launch_bounds__(kT) void to_bf16_kernel(long long count, const float* __restrict__ src,
                                                     uint16_t* __restrict__ dst) {
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < count; e += (long long)gridDim.x * kT)
    dst[e] = to_bf16(src[e]);
}

struct MlpFailure {
  int code;
  std::string message;
};

[[noreturn]] void mfail(int code, const std::string& m) { throw MlpFailure{code, m}; }

void mcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) mfail(HFJK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int mguard(hfjk_ctx* ctx, F&& body) {
  try {
    if (ctx == nullptr) mfail(HFJK_ERR_INPUT, "null context");
    mcheck(hipSetDevice(ctx->device), "hipSetDevice");
    body();
    return HFJK_OK;
  } catch (const MlpFailure& f) {
    hfjk_internal_set_error(f.message.c_str());
    return f.code;
  } catch (...) {
    hfjk_internal_set_error("unknown failure in an MLP routine");
    return HFJK_ERR_HIP;
  }
}

}  // namespace

extern "C" {

int hfjk_mlp_gemm(hfjk_ctx* ctx, int m, int n, int k, const void* a, int a_bf16, int lda, int trans_a,
                  const int* a_rows, const void* b, int b_bf16, int ldb, int trans_b, const float* bias, float* c,
                  int ldc, int accumulate, float* h_silu, void* stream) {
  return mguard(ctx, [&] {
    if (m < 0 || n < 0 || k < 0) mfail(HFJK_ERR_INPUT, "hfjk_mlp_gemm: negative size");
    if (m == 0 || n == 0) return;
    if (a == nullptr || b == nullptr || c == nullptr) mfail(HFJK_ERR_INPUT, "hfjk_mlp_gemm: null operand");
    // Leading dimensions cover the stored rows (gathered row indices are the caller's contract).
    if (lda < (trans_a ? m : k) || ldb < (trans_b ? k : n) || ldc < n)
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_gemm: a leading dimension is smaller than its row");
    if (h_silu == c) mfail(HFJK_ERR_INPUT, "hfjk_mlp_gemm: the activation output must not alias C");
    GemmArgs g{m, n, k, a, lda, trans_a ? 1 : 0, a_rows, b, ldb, trans_b ? 1 : 0, bias, c, ldc, accumulate ? 1 : 0,
               h_silu};
    dim3 grid((n + kTileN - 1) / kTileN, (m + kTileM - 1) / kTileM);
    if (grid.y > 65535u) mfail(HFJK_ERR_INPUT, "hfjk_mlp_gemm: more than 4194240 rows");
    hipStream_t s = (hipStream_t)stream;
    if (a_bf16 && b_bf16)
      hipLaunchKernelGGL((mlp_gemm_kernel<uint16_t, uint16_t>), grid, dim3(kT), 0, s, g);
    else if (a_bf16)
      hipLaunchKernelGGL((mlp_gemm_kernel<uint16_t, float>), grid, dim3(kT), 0, s, g);
    else if (b_bf16)
      hipLaunchKernelGGL((mlp_gemm_kernel<float, uint16_t>), grid, dim3(kT), 0, s, g);
    else
      hipLaunchKernelGGL((mlp_gemm_kernel<float, float>), grid, dim3(kT), 0, s, g);
    mcheck(hipGetLastError(), "mlp_gemm launch");
  });
}

int hfjk_mlp_silu_backward(hfjk_ctx* ctx, long long count, const float* grad_h, const float* z, float* grad_z,
                           void* stream) {
  return mguard(ctx, [&] {
    if (count < 0 || (count > 0 && (grad_h == nullptr || z == nullptr || grad_z == nullptr)))
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_silu_backward: null buffer");
    if (count == 0) return;
    hipLaunchKernelGGL(silu_backward_kernel, dim3(hfjk::grid_for(count)), dim3(kT), 0, (hipStream_t)stream, count,
                       grad_h, z, grad_z);
    mcheck(hipGetLastError(), "silu_backward launch");
  });
}

int hfjk_mlp_column_sum(hfjk_ctx* ctx, int m, int n, const float* x, int ldx, float* out, void* stream) {
  return mguard(ctx, [&] {
    if (m < 0 || n < 0 || ldx < n) mfail(HFJK_ERR_INPUT, "hfjk_mlp_column_sum: bad shape");
    if (n == 0) return;
    if (out == nullptr || (m > 0 && x == nullptr)) mfail(HFJK_ERR_INPUT, "hfjk_mlp_column_sum: null buffer");
    hipLaunchKernelGGL(column_sum_kernel, dim3((n + 63) / 64), dim3(kT), 0, (hipStream_t)stream, m, n, x, ldx, out);
    mcheck(hipGetLastError(), "column_sum launch");
  });
}

int hfjk_mlp_mse(hfjk_ctx* ctx, int m, int n, const float* y, int ldy, const float* target, int ldt,
                 const int* target_rows, const float* col_weights, float scale, float* grad, float* loss,
                 void* stream) {
  return mguard(ctx, [&] {
    if (m < 0 || n < 1 || ldy < n || ldt < n) mfail(HFJK_ERR_INPUT, "hfjk_mlp_mse: bad shape");
    if (loss == nullptr || (m > 0 && (y == nullptr || target == nullptr)))
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_mse: null buffer");
    hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(kT), 0, (hipStream_t)stream, m, n, y, ldy, target, ldt, target_rows,
                       col_weights, scale, grad, loss);
    mcheck(hipGetLastError(), "mse launch");
  });
}

int hfjk_mlp_adam(hfjk_ctx* ctx, long long count, float* params, const float* grads, float* moment1, float* moment2,
                  uint16_t* params_bf16, double lr, double lr_t, double beta1, double beta2, double eps,
                  double weight_decay, void* stream) {
  return mguard(ctx, [&] {
    if (count < 0 || (count > 0 && (params == nullptr || grads == nullptr || moment1 == nullptr ||
                                    moment2 == nullptr)))
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_adam: null buffer");
    if (!(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps >= 0.0))
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_adam: betas in [0, 1) and eps >= 0 are required");
    if (count == 0) return;
    hipLaunchKernelGGL(adam_kernel, dim3(hfjk::grid_for(count)), dim3(kT), 0, (hipStream_t)stream, count, params,
                       grads, moment1, moment2, params_bf16, (float)lr_t, (float)beta1, (float)(1.0 - beta1),
                       (float)beta2, (float)(1.0 - beta2), (float)eps, (float)(lr * weight_decay));
    mcheck(hipGetLastError(), "adam launch");
  });
}

int hfjk_mlp_to_bf16(hfjk_ctx* ctx, long long count, const float* src, uint16_t* dst, void* stream) {
  return mguard(ctx, [&] {
    if (count < 0 || (count > 0 && (src == nullptr || dst == nullptr)))
      mfail(HFJK_ERR_INPUT, "hfjk_mlp_to_bf16: null buffer");
    if (count == 0) return;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(hfjk::grid_for(count)), dim3(kT), 0, (hipStream_t)stream, count, src,
                       dst);
    mcheck(hipGetLastError(), "to_bf16 launch");
  });
}

}  // extern "C"
