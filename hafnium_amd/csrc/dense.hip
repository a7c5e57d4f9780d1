// FP64 dense steps of the SCF around the J/K build: row-major GEMM and
// triangular solve through rocBLAS, symmetric eigendecomposition and
// Cholesky through rocSOLVER, and this library's own elementwise and
// fixed-order reduction kernels (deterministic results run to run).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <string>

#include "hfjk_internal.h"

namespace {

thread_local std::string g_dense_error;

struct DenseFailure {
  int code;
  std::string message;
};

[[noreturn]] void dfail(int code, const std::string& m) { throw DenseFailure{code, m}; }

void dcheck_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) dfail(HFJK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void dcheck_blas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success) dfail(HFJK_ERR_SOLVER, std::string(what) + " failed (rocblas status " + std::to_string((int)st) + ")");
}

template <typename F>
int dguard(F&& body) {
  try {
    body();
    return HFJK_OK;
  } catch (const DenseFailure& f) {
    hfjk_internal_set_error(f.message.c_str());
    return f.code;
  } catch (...) {
    hfjk_internal_set_error("unknown failure in a dense FP64 routine");
    return HFJK_ERR_HIP;
  }
}

rocblas_handle handle_for(hfjk_ctx* ctx, hipStream_t s) {
  if (ctx->blas == nullptr) {
    rocblas_handle h = nullptr;
    dcheck_blas(rocblas_create_handle(&h), "rocblas_create_handle");
    ctx->blas = h;
  }
  rocblas_handle h = (rocblas_handle)ctx->blas;
  dcheck_blas(rocblas_set_stream(h, s), "rocblas_set_stream");
  return h;
}

constexpr int kT = 256;

__global__ __launch_bounds__(kT) void combine_kernel(long long n, double a, const double* __restrict__ x, double b,
                                                     const double* __restrict__ y, double c,
                                                     const double* __restrict__ z, double* __restrict__ out) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) {
    double v = a * x[i];
    if (y != nullptr) v += b * y[i];
    if (z != nullptr) v += c * z[i];
    out[i] = v;
  }
}

// Antisymmetric part: out = m - m^T ([n][n]).
__global__ __launch_bounds__(kT) void antisym_kernel(int n, const double* __restrict__ m, double* __restrict__ out) {
  const long long total = (long long)n * n;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < total; i += (long long)gridDim.x * kT) {
    const int r = (int)(i / n), c = (int)(i % n);
    out[i] = m[i] - m[(long long)c * n + r];
  }
}

// Fixed-order dot products: block partials, then one block sums them.
constexpr int kDotBlocks = 512;

__global__ __launch_bounds__(kT) void dot_partial_kernel(long long n, const double* __restrict__ x,
                                                         const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double red[kT];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n; i += (long long)gridDim.x * kT) s += x[i] * y[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  __shared__ double red[kT];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kT) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

extern "C" {

int hfjk_dgemm(hfjk_ctx* ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* a, int lda,
               const double* b, int ldb, double beta, double* c, int ldc, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (m < 0 || n < 0 || k < 0) dfail(HFJK_ERR_INPUT, "negative GEMM dimension");
    if ((m > 0 && n > 0) && (c == nullptr || (k > 0 && (a == nullptr || b == nullptr))))
      dfail(HFJK_ERR_INPUT, "null matrix pointer");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    rocblas_handle h = handle_for(ctx, (hipStream_t)stream);
    // Row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T.
    const rocblas_operation oa = trans_a ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation ob = trans_b ? rocblas_operation_transpose : rocblas_operation_none;
    dcheck_blas(rocblas_dgemm(h, ob, oa, n, m, k, &alpha, b, ldb, a, lda, &beta, c, ldc), "rocblas_dgemm");
  });
}

int hfjk_dsyevd(hfjk_ctx* ctx, int n, double* a, double* w, double* work, int* info, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (n < 1 || a == nullptr || w == nullptr || work == nullptr || info == nullptr)
      dfail(HFJK_ERR_INPUT, "hfjk_dsyevd needs n >= 1 and device buffers a[n*n], w[n], work[n], info[1]");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    rocblas_handle h = handle_for(ctx, (hipStream_t)stream);
    dcheck_blas(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, work, info),
                "rocsolver_dsyevd");
  });
}

int hfjk_dpotrf(hfjk_ctx* ctx, int n, double* a, int* info, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (n < 1 || a == nullptr || info == nullptr) dfail(HFJK_ERR_INPUT, "hfjk_dpotrf needs n >= 1 and device buffers");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    rocblas_handle h = handle_for(ctx, (hipStream_t)stream);
    // Row-major lower L (M = L L^T) is column-major upper U = L^T (M = U^T U).
    dcheck_blas(rocsolver_dpotrf(h, rocblas_fill_upper, n, a, n, info), "rocsolver_dpotrf");
  });
}

int hfjk_dtrsm_lower(hfjk_ctx* ctx, int m, int n, const double* l, double* b, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (m < 1 || n < 1 || l == nullptr || b == nullptr) dfail(HFJK_ERR_INPUT, "hfjk_dtrsm_lower needs m, n >= 1");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    rocblas_handle h = handle_for(ctx, (hipStream_t)stream);
    // Row-major B [m][n] <- L^-1 B with L [m][m] lower (row-major). In
    // column-major terms B^T [n][m] <- B^T L^-T, L^T stored as upper:
    // solve X op(A) = alpha B^T from the right with A = upper (the row-major
    // lower L read column-major), op = none: X U = B^T, U = L^T.
    const double one = 1.0;
    dcheck_blas(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n, m, &one, l, m, b, n),
                "rocblas_dtrsm");
  });
}

int hfjk_combine(hfjk_ctx* ctx, long long n, double a, const double* x, double b, const double* y, double c,
                 const double* z, double* out, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (n < 0 || (n > 0 && (x == nullptr || out == nullptr))) dfail(HFJK_ERR_INPUT, "hfjk_combine needs x and out");
    if (n == 0) return;
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    hipLaunchKernelGGL(combine_kernel, dim3(hfjk::grid_for(n)), dim3(kT), 0, (hipStream_t)stream, n, a, x, b, y, c, z,
                       out);
    dcheck_hip(hipGetLastError(), "combine launch");
  });
}

int hfjk_antisymmetrize(hfjk_ctx* ctx, int n, const double* m, double* out, void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (n < 1 || m == nullptr || out == nullptr || m == out) dfail(HFJK_ERR_INPUT, "hfjk_antisymmetrize needs distinct buffers");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    hipLaunchKernelGGL(antisym_kernel, dim3(hfjk::grid_for((long long)n * n)), dim3(kT), 0, (hipStream_t)stream, n, m,
                       out);
    dcheck_hip(hipGetLastError(), "antisymmetrize launch");
  });
}

long long hfjk_dot_workspace(void) { return kDotBlocks; }

int hfjk_ddot(hfjk_ctx* ctx, long long n, const double* x, const double* y, double* partial, double* out,
              void* stream) {
  return dguard([&] {
    if (ctx == nullptr) dfail(HFJK_ERR_INPUT, "null context");
    if (n < 0 || x == nullptr || y == nullptr || partial == nullptr || out == nullptr)
      dfail(HFJK_ERR_INPUT, "hfjk_ddot needs x, y, partial[hfjk_dot_workspace()] and out");
    dcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    long long blocks = (n + kT - 1) / kT;
    if (blocks > kDotBlocks) blocks = kDotBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(dot_partial_kernel, dim3((unsigned)blocks), dim3(kT), 0, s, n, x, y, partial);
    dcheck_hip(hipGetLastError(), "dot launch");
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kT), 0, s, (int)blocks, partial, out);
    dcheck_hip(hipGetLastError(), "dot launch");
  });
}

}  // extern "C"
