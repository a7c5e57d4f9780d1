// Batched FP64 dense steps of a batched SCF: many instances of one padded
// size n (each instance's real size n_i <= n, zero-padded), laid out at a
// fixed stride. Strided-batched rocBLAS GEMM and rocSOLVER eigensolvers, and
// this library's own small kernels: per-instance dot products (one
// workgroup per instance, fixed summation order), antisymmetrisation, the
// canonical orthogonaliser from eigenpairs, a diagonal shift of the
// discarded (padded) directions, column masks and copies between padded and
// compact per-instance storage.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>

#include "hfjk_internal.h"

namespace {

constexpr int kT = 256;

struct BdFailure {
  int code;
  std::string message;
};

[[noreturn]] void bfail(int code, const std::string& m) { throw BdFailure{code, m}; }

void bcheck_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) bfail(HFJK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void bcheck_blas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    bfail(HFJK_ERR_SOLVER, std::string(what) + " failed (rocblas status " + std::to_string((int)st) + ")");
}

template <typename F>
int bguard(hfjk_ctx* ctx, F&& body) {
  try {
    if (ctx == nullptr) bfail(HFJK_ERR_INPUT, "null context");
    bcheck_hip(hipSetDevice(ctx->device), "hipSetDevice");
    body();
    return HFJK_OK;
  } catch (const BdFailure& f) {
    hfjk_internal_set_error(f.message.c_str());
    return f.code;
  } catch (...) {
    hfjk_internal_set_error("unknown failure in a batched dense routine");
    return HFJK_ERR_HIP;
  }
}

rocblas_handle blas_for(hfjk_ctx* ctx, hipStream_t s) {
  if (ctx->blas == nullptr) {
    rocblas_handle h = nullptr;
    bcheck_blas(rocblas_create_handle(&h), "rocblas_create_handle");
    ctx->blas = h;
  }
  rocblas_handle h = (rocblas_handle)ctx->blas;
  bcheck_blas(rocblas_set_stream(h, s), "rocblas_set_stream");
  return h;
}

// out[i] = sum_e x_i[e] y_i[e]; one workgroup per instance.
__global__ __launch_bounds__(kT) void batched_dot_kernel(long long len, const double* __restrict__ x, long long sx,
                                                         const double* __restrict__ y, long long sy,
                                                         double* __restrict__ out) {
  __shared__ double red[kT];
  const double* xi = x + (long long)blockIdx.x * sx;
  const double* yi = y + (long long)blockIdx.x * sy;
  double s = 0.0;
  for (long long e = threadIdx.x; e < len; e += kT) s += xi[e] * yi[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// out[i][a] = <hist[newest][i], hist[a][i]> over len elements, hist [nslots][count][len];
// one workgroup per (instance, slot), fixed summation order.
__global__ __launch_bounds__(kT) void history_dots_kernel(long long len, int count, const double* __restrict__ hist,
                                                          int newest, double* __restrict__ out) {
  __shared__ double red[kT];
  const int i = blockIdx.x, a = blockIdx.y;
  const double* x = hist + ((long long)newest * count + i) * len;
  const double* y = hist + ((long long)a * count + i) * len;
  double s = 0.0;
  for (long long e = threadIdx.x; e < len; e += kT) s += x[e] * y[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(long long)i * gridDim.y + a] = red[0];
}

__global__ __launch_bounds__(kT) void batched_antisym_kernel(int n, const double* __restrict__ m,
                                                             double* __restrict__ out) {
  const long long nn = (long long)n * n;
  const double* mi = m + (long long)blockIdx.y * nn;
  double* oi = out + (long long)blockIdx.y * nn;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < nn; e += (long long)gridDim.x * kT) {
    const int r = (int)(e / n), c = (int)(e % n);
    oi[e] = mi[e] - mi[(long long)c * n + r];
  }
}

// a_i <- (a_i + a_i^T) / 2 in place: one thread per pair (r < c) writes both entries.
__global__ __launch_bounds__(kT) void batched_sym_kernel(int n, double* __restrict__ a) {
  const long long nn = (long long)n * n;
  double* ai = a + (long long)blockIdx.y * nn;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < nn; e += (long long)gridDim.x * kT) {
    const int r = (int)(e / n), c = (int)(e % n);
    if (r < c) {
      const double v = 0.5 * (ai[e] + ai[(long long)c * n + r]);
      ai[e] = v;
      ai[(long long)c * n + r] = v;
    }
  }
}

// x_i[u][j] = vecs_i[j][u] / sqrt(vals_i[j]) for vals_i[j] > threshold, else 0;
// kept_i[j] = 1 / 0 accordingly (rows of vecs are eigenvectors).
__global__ __launch_bounds__(kT) void batched_orth_kernel(int n, const double* __restrict__ vecs,
                                                          const double* __restrict__ vals, double threshold,
                                                          double* __restrict__ x, int* __restrict__ kept) {
  const long long nn = (long long)n * n;
  const int i = blockIdx.y;
  const double* vi = vecs + (long long)i * nn;
  const double* wi = vals + (long long)i * n;
  double* xi = x + (long long)i * nn;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < nn; e += (long long)gridDim.x * kT) {
    const int u = (int)(e / n), j = (int)(e % n);
    const double w = wi[j];
    xi[e] = w > threshold ? vi[(long long)j * n + u] / sqrt(w) : 0.0;
    if (u == 0) kept[(long long)i * n + j] = w > threshold ? 1 : 0;
  }
}

__global__ __launch_bounds__(kT) void batched_shift_kernel(int n, int count, const int* __restrict__ kept,
                                                           double value, double* __restrict__ a) {
  const long long total = (long long)n * count;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < total; e += (long long)gridDim.x * kT) {
    const long long i = e / n;
    const int j = (int)(e % n);
    if (!kept[e]) a[i * ((long long)n * n) + (long long)j * n + j] += value;
  }
}

__global__ __launch_bounds__(kT) void batched_mask_kernel(int rows, int cols, const int* __restrict__ keep,
                                                          double* __restrict__ a) {
  const long long rc = (long long)rows * cols;
  const int i = blockIdx.y;
  const int k = keep[i];
  double* ai = a + (long long)i * rc;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < rc; e += (long long)gridDim.x * kT)
    if ((int)(e % cols) >= k) ai[e] = 0.0;
}

// Padded (r, c) at i * stride + r * ld + c <-> compact at offsets[i] + r * cols_i + c.
__global__ __launch_bounds__(kT) void batched_copy_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                                                          const long long* __restrict__ offsets,
                                                          double* __restrict__ padded, int ld, long long stride,
                                                          double* __restrict__ compact, int to_compact) {
  const int i = blockIdx.y;
  const int ci = cols[i];
  const long long total = (long long)rows[i] * ci;
  double* pi = padded + (long long)i * stride;
  double* qi = compact + offsets[i];
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < total; e += (long long)gridDim.x * kT) {
    const long long r = e / ci, c = e % ci;
    if (to_compact)
      qi[e] = pi[r * ld + c];
    else
      pi[r * ld + c] = qi[e];
  }
}

__global__ __launch_bounds__(kT) void batched_select_kernel(long long len, const int* __restrict__ flags,
                                                            const double* __restrict__ a, const double* __restrict__ b,
                                                            double* __restrict__ out) {
  const long long base = (long long)blockIdx.y * len;
  const double* src = flags[blockIdx.y] ? a + base : b + base;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < len; e += (long long)gridDim.x * kT)
    out[base + e] = src[e];
}

unsigned blocks_for(long long elems) {
  long long b = (elems + kT - 1) / kT;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (unsigned)b;
}

void check_count(int count) {
  if (count < 1 || count > 65535) bfail(HFJK_ERR_INPUT, "batch count must be in 1..65535");
}

// DIIS coefficients, one 64-lane workgroup per instance (the reference's
// solve_diis, hafnium/scf/diis.py:82-106): the newest overlaps (dots[i][a]
// = <e_new, e_a>) first go into row and column `newest` of ov[i] (as 2
// <e_a, e_b>); over the valid slots, the bordered system [[B / s, -1],
// [-1^T, 0]] (s = the largest diagonal of B) is diagonalised by cyclic
// Jacobi rotations in LDS (lane k updates row / column entry k of each
// rotation) and solved for the right side (0, ..., 0, -1) with eigenvalues
// |w| <= 1e-14 max |w| dropped. coef[i] holds the solution on the valid
// slots, zero on the others.
constexpr int kDiisMax = 16;

struct DiisSlots {
  int n;
  int slot[kDiisMax];
};

__global__ __launch_bounds__(64) void diis_kernel(int m, double* __restrict__ ov, const double* __restrict__ dots,
                                                  int newest, DiisSlots valid, double* __restrict__ coef) {
  constexpr int N = kDiisMax + 1;
  __shared__ double a[N][N + 1], v[N][N + 1];
  __shared__ double rot[2];
  __shared__ int done;
  const int i = blockIdx.x;
  const int k = threadIdx.x;
  double* o = ov + (long long)i * m * m;
  double* c = coef + (long long)i * m;
  if (k < m) {
    const double x = 2.0 * dots[(long long)i * m + k];
    o[newest * m + k] = x;
    o[k * m + newest] = x;
    c[k] = 0.0;
  }
  __syncthreads();
  const int nv = valid.n;
  if (nv == 1) {
    if (k == 0) c[valid.slot[0]] = 1.0;
    return;
  }
  const int n = nv + 1;
  double scale = 0.0;
  for (int p = 0; p < nv; ++p) scale = fmax(scale, o[valid.slot[p] * m + valid.slot[p]]);
  scale = scale > 0.0 ? scale : 1.0;
  if (k < n)
    for (int q = 0; q < n; ++q) {
      a[k][q] = k < nv && q < nv ? o[valid.slot[k] * m + valid.slot[q]] / scale : (k == nv && q == nv ? 0.0 : -1.0);
      v[k][q] = k == q ? 1.0 : 0.0;
    }
  __syncthreads();
  for (int sweep = 0; sweep < 64; ++sweep) {
    if (k == 0) {
      double off = 0.0, diag = 0.0;
      for (int p = 0; p < n; ++p) {
        diag += a[p][p] * a[p][p];
        for (int q = p + 1; q < n; ++q) off += a[p][q] * a[p][q];
      }
      done = off <= 1e-34 * diag;
    }
    __syncthreads();
    if (done) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (k == 0) {
          const double apq = a[p][q];
          double cs = 1.0, sn = 0.0;
          if (apq != 0.0) {
            const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
            cs = 1.0 / sqrt(1.0 + t * t);
            sn = t * cs;
          }
          rot[0] = cs;
          rot[1] = sn;
        }
        __syncthreads();
        const double cs = rot[0], sn = rot[1];
        if (k < n) {
          const double kp = a[k][p], kq = a[k][q];
          a[k][p] = cs * kp - sn * kq;
          a[k][q] = sn * kp + cs * kq;
          const double vp = v[k][p], vq = v[k][q];
          v[k][p] = cs * vp - sn * vq;
          v[k][q] = sn * vp + cs * vq;
        }
        __syncthreads();
        if (k < n) {
          const double pk = a[p][k], qk = a[q][k];
          a[p][k] = cs * pk - sn * qk;
          a[q][k] = sn * pk + cs * qk;
        }
        __syncthreads();
      }
  }
  // x_p = sum_j v[p][j] z_j, z_j = (V^T r)_j / w_j = -v[n-1][j] / w_j over the kept eigenvalues.
  if (k < nv) {
    double wmax = 0.0;
    for (int j = 0; j < n; ++j) wmax = fmax(wmax, fabs(a[j][j]));
    double x = 0.0;
    for (int j = 0; j < n; ++j) {
      const double w = a[j][j];
      if (fabs(w) > 1e-14 * wmax) x += v[k][j] * (-v[n - 1][j] / w);
    }
    c[valid.slot[k]] = x;
  }
}

// The DIIS extrapolation out[i] = sum_a coef[i][a] hist[a][i] of count
// instances (hist [nslots][count][len]); slots with a zero coefficient are
// not read.
__global__ __launch_bounds__(kT) void diis_combine_kernel(long long len, int count, int nslots,
                                                          const double* __restrict__ hist,
                                                          const double* __restrict__ coef, double* __restrict__ out) {
  const int i = blockIdx.y;
  const double* ci = coef + (long long)i * nslots;
  for (long long e = (long long)blockIdx.x * kT + threadIdx.x; e < len; e += (long long)gridDim.x * kT) {
    double acc = 0.0;
    for (int a = 0; a < nslots; ++a) {
      const double w = ci[a];
      if (w != 0.0) acc += w * hist[((long long)a * count + i) * len + e];
    }
    out[(long long)i * len + e] = acc;
  }
}

}  // namespace

extern "C" {

int hfjk_dgemm_batched(hfjk_ctx* ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* a,
                       int lda, long long stride_a, const double* b, int ldb, long long stride_b, double beta,
                       double* c, int ldc, long long stride_c, int count, void* stream) {
  return bguard(ctx, [&] {
    if (m < 0 || n < 0 || k < 0 || count < 0) bfail(HFJK_ERR_INPUT, "negative batched GEMM dimension");
    if (m == 0 || n == 0 || count == 0) return;
    if (c == nullptr || (k > 0 && (a == nullptr || b == nullptr))) bfail(HFJK_ERR_INPUT, "null matrix pointer");
    rocblas_handle h = blas_for(ctx, (hipStream_t)stream);
    const rocblas_operation oa = trans_a ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation ob = trans_b ? rocblas_operation_transpose : rocblas_operation_none;
    // Row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T.
    bcheck_blas(rocblas_dgemm_strided_batched(h, ob, oa, n, m, k, &alpha, b, ldb, stride_b, a, lda, stride_a, &beta, c,
                                              ldc, stride_c, count),
                "rocblas_dgemm_strided_batched");
  });
}

long long hfjk_dsyev_batched_work(int n, int count) { return (long long)count * n + 2LL * count; }

int hfjk_dsyev_batched(hfjk_ctx* ctx, int n, int count, double* a, double* w, double* work, int* info, int method,
                       void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (n < 1 || a == nullptr || w == nullptr || work == nullptr || info == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_dsyev_batched needs n >= 1 and device buffers");
    rocblas_handle h = blas_for(ctx, (hipStream_t)stream);
    const rocblas_stride sa = (rocblas_stride)n * n;
    if (method == 0) {
      bcheck_blas(rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, n, a, n, sa, w, n,
                                                   work, n, info, count),
                  "rocsolver_dsyevd_strided_batched");
    } else if (method == 1) {
      double* residual = work + (long long)count * n;
      int* sweeps = (int*)(residual + count);
      bcheck_blas(rocsolver_dsyevj_strided_batched(h, rocblas_esort_ascending, rocblas_evect_original,
                                                   rocblas_fill_upper, n, a, n, sa, 0.0, residual, 100, sweeps, w, n,
                                                   info, count),
                  "rocsolver_dsyevj_strided_batched");
    } else {
      bfail(HFJK_ERR_INPUT, "method is 0 (divide and conquer) or 1 (Jacobi)");
    }
  });
}

int hfjk_ddot_batched(hfjk_ctx* ctx, int count, long long len, const double* x, long long stride_x, const double* y,
                      long long stride_y, double* out, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (len < 0 || x == nullptr || y == nullptr || out == nullptr) bfail(HFJK_ERR_INPUT, "hfjk_ddot_batched: bad input");
    hipLaunchKernelGGL(batched_dot_kernel, dim3(count), dim3(kT), 0, (hipStream_t)stream, len, x, stride_x, y,
                       stride_y, out);
    bcheck_hip(hipGetLastError(), "batched_dot launch");
  });
}

int hfjk_history_dots_batched(hfjk_ctx* ctx, int count, long long len, int nslots, const double* hist, int newest,
                              double* out, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (len < 0 || nslots < 1 || nslots > 64 || newest < 0 || newest >= nslots || hist == nullptr || out == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_history_dots_batched: bad input");
    hipLaunchKernelGGL(history_dots_kernel, dim3(count, nslots), dim3(kT), 0, (hipStream_t)stream, len, count, hist,
                       newest, out);
    bcheck_hip(hipGetLastError(), "history_dots launch");
  });
}

int hfjk_diis_batched(hfjk_ctx* ctx, int count, int nslots, double* ov, const double* dots, int newest,
                      const int* valid, int nvalid, double* coef, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (nslots < 1 || nslots > kDiisMax || newest < 0 || newest >= nslots || nvalid < 1 || nvalid > nslots ||
        valid == nullptr || ov == nullptr || dots == nullptr || coef == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_diis_batched: 1 <= nvalid <= nslots <= 16, 0 <= newest < nslots and buffers needed");
    DiisSlots slots{};
    slots.n = nvalid;
    for (int a = 0; a < nvalid; ++a) {
      if (valid[a] < 0 || valid[a] >= nslots) bfail(HFJK_ERR_INPUT, "hfjk_diis_batched: a valid slot is out of range");
      for (int b = 0; b < a; ++b)
        if (valid[b] == valid[a]) bfail(HFJK_ERR_INPUT, "hfjk_diis_batched: a valid slot is repeated");
      slots.slot[a] = valid[a];
    }
    hipLaunchKernelGGL(diis_kernel, dim3((unsigned)count), dim3(64), 0, (hipStream_t)stream, nslots, ov, dots, newest,
                       slots, coef);
    bcheck_hip(hipGetLastError(), "diis launch");
  });
}

int hfjk_diis_combine_batched(hfjk_ctx* ctx, int count, long long len, int nslots, const double* hist,
                              const double* coef, double* out, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (len < 0 || nslots < 1 || hist == nullptr || coef == nullptr || out == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_diis_combine_batched: bad input");
    if (len == 0) return;
    const unsigned gx = (unsigned)std::min<long long>(1024, blocks_for(len));
    hipLaunchKernelGGL(diis_combine_kernel, dim3(gx, count), dim3(kT), 0, (hipStream_t)stream, len, count, nslots,
                       hist, coef, out);
    bcheck_hip(hipGetLastError(), "diis_combine launch");
  });
}

int hfjk_antisymmetrize_batched(hfjk_ctx* ctx, int n, int count, const double* m, double* out, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (n < 1 || m == nullptr || out == nullptr || m == out)
      bfail(HFJK_ERR_INPUT, "hfjk_antisymmetrize_batched needs distinct buffers");
    hipLaunchKernelGGL(batched_antisym_kernel, dim3(blocks_for((long long)n * n), count), dim3(kT), 0,
                       (hipStream_t)stream, n, m, out);
    bcheck_hip(hipGetLastError(), "batched_antisym launch");
  });
}

int hfjk_symmetrize_batched(hfjk_ctx* ctx, int n, long long count, double* a, void* stream) {
  return bguard(ctx, [&] {
    if (n < 1 || count < 0 || a == nullptr) bfail(HFJK_ERR_INPUT, "hfjk_symmetrize_batched: bad input");
    // Chunks of at most 65535 matrices per launch (grid.y limit).
    for (long long at = 0; at < count; at += 65535) {
      const unsigned chunk = (unsigned)std::min<long long>(65535, count - at);
      hipLaunchKernelGGL(batched_sym_kernel, dim3(blocks_for((long long)n * n), chunk), dim3(kT), 0,
                         (hipStream_t)stream, n, a + at * ((long long)n * n));
      bcheck_hip(hipGetLastError(), "batched_sym launch");
    }
  });
}

int hfjk_orthogonalizer_batched(hfjk_ctx* ctx, int n, int count, const double* vecs, const double* vals,
                                double threshold, double* x, int* kept, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (n < 1 || vecs == nullptr || vals == nullptr || x == nullptr || kept == nullptr || x == vecs)
      bfail(HFJK_ERR_INPUT, "hfjk_orthogonalizer_batched: bad input");
    hipLaunchKernelGGL(batched_orth_kernel, dim3(blocks_for((long long)n * n), count), dim3(kT), 0,
                       (hipStream_t)stream, n, vecs, vals, threshold, x, kept);
    bcheck_hip(hipGetLastError(), "batched_orth launch");
  });
}

int hfjk_shift_discarded_batched(hfjk_ctx* ctx, int n, int count, const int* kept, double value, double* a,
                                 void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (n < 1 || kept == nullptr || a == nullptr) bfail(HFJK_ERR_INPUT, "hfjk_shift_discarded_batched: bad input");
    hipLaunchKernelGGL(batched_shift_kernel, dim3(blocks_for((long long)n * count)), dim3(kT), 0,
                       (hipStream_t)stream, n, count, kept, value, a);
    bcheck_hip(hipGetLastError(), "batched_shift launch");
  });
}

int hfjk_mask_columns_batched(hfjk_ctx* ctx, int rows, int cols, int count, const int* keep, double* a,
                              void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (rows < 1 || cols < 1 || keep == nullptr || a == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_mask_columns_batched: bad input");
    hipLaunchKernelGGL(batched_mask_kernel, dim3(blocks_for((long long)rows * cols), count), dim3(kT), 0,
                       (hipStream_t)stream, rows, cols, keep, a);
    bcheck_hip(hipGetLastError(), "batched_mask launch");
  });
}

int hfjk_select_batched(hfjk_ctx* ctx, int count, long long len, const int* flags, const double* a, const double* b,
                        double* out, void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (len < 0 || flags == nullptr || a == nullptr || b == nullptr || out == nullptr)
      bfail(HFJK_ERR_INPUT, "hfjk_select_batched: bad input");
    if (len == 0) return;
    hipLaunchKernelGGL(batched_select_kernel, dim3(blocks_for(len), count), dim3(kT), 0, (hipStream_t)stream, len,
                       flags, a, b, out);
    bcheck_hip(hipGetLastError(), "batched_select launch");
  });
}

int hfjk_copy_batched(hfjk_ctx* ctx, int count, const int* rows, const int* cols, const long long* offsets,
                      long long max_elems, double* padded, int ld, long long stride, double* compact, int to_compact,
                      void* stream) {
  return bguard(ctx, [&] {
    check_count(count);
    if (rows == nullptr || cols == nullptr || offsets == nullptr || padded == nullptr || compact == nullptr ||
        ld < 1 || stride < 1 || max_elems < 0)
      bfail(HFJK_ERR_INPUT, "hfjk_copy_batched: bad input");
    if (max_elems == 0) return;
    hipLaunchKernelGGL(batched_copy_kernel, dim3(blocks_for(max_elems), count), dim3(kT), 0, (hipStream_t)stream,
                       rows, cols, offsets, padded, ld, stride, compact, to_compact ? 1 : 0);
    bcheck_hip(hipGetLastError(), "batched_copy launch");
  });
}

}  // extern "C"
