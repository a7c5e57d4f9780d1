// k-point DF-J/K kernels (FP64, complex values as interleaved double2).
//
// Coulomb: the charges c_P = sum_k w_k sum_xw conj(B_k[P][x][w]) D_k[x][w]
// over the zero-transfer blocks (one workgroup per (k, P), a fixed-order
// reduction), then J_k = sum_P s_k[P] c_P B_k[P] streamed once over the
// blocks (one thread per matrix element, coalesced over elements).
//
// Exchange of one transfer block B = B[k1, k2] (general complex [naux][n][n])
// against a Hermitian D(k2) = V diag(l) V^dagger: with C = V |l|^1/2 sqrt(w)
// ([n][r] complex) the realified operands
//   Bh[P] = [[Re B, -Im B], [Im B, Re B]]   ([2n][2n] real),
//   Ch    = [Re C; Im C]                    ([2n][r] real)
// give Xh[P] = Bh[P] Ch = [Re X; Im X] for X = B[P] C, with no redundant
// products. The existing FP64 half-transform (direct-LDS ring over a full
// row-major B) and the signed SYRK then form
//   Kh = sum_P s[P] Xh[P] diag(sign l) Xh[P]^T ([2n][2n] symmetric),
// and K += (Kh_TL + Kh_BR) + i (Kh_BL - Kh_TR) is the block's
// sum_P s[P] B[P] D B[P]^dagger. Per-aux signs s[P] scale the SYRK's second
// operand per aux function.
#include <algorithm>

#include "hfjk_internal.h"

namespace hfjk {
namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Fixed-order sum of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double out = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return out;
}

// inner[k][P] = w[k] sum_xw conj(B_k[P][xw]) D_k[xw]; one workgroup per (k, P).
__global__ __launch_bounds__(kThreads) void kp_inner_kernel(const double2* __restrict__ b, long long kstride,
                                                            int naux, long long nn, const double2* __restrict__ d,
                                                            const double* __restrict__ w, double2* __restrict__ inner) {
  __shared__ double red[4];
  const int k = blockIdx.x / naux;
  const int p = blockIdx.x % naux;
  const double2* bp = b + (long long)k * kstride + (long long)p * nn;
  const double2* dk = d + (long long)k * nn;
  double re = 0.0, im = 0.0;
  for (long long i = threadIdx.x; i < nn; i += kThreads) {
    const double2 x = bp[i], y = dk[i];
    re += x.x * y.x + x.y * y.y;
    im += x.x * y.y - x.y * y.x;
  }
  re = block_sum(re, red);
  im = block_sum(im, red);
  if (threadIdx.x == 0) inner[(long long)k * naux + p] = make_double2(w[k] * re, w[k] * im);
}

// charges[P] = sum_k inner[k][P], k in order.
__global__ __launch_bounds__(kThreads) void kp_charges_kernel(const double2* __restrict__ inner, int nk, int naux,
                                                              double2* __restrict__ charges) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= naux) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int k = 0; k < nk; ++k) {
    const double2 v = inner[(long long)k * naux + p];
    acc.x += v.x;
    acc.y += v.y;
  }
  charges[p] = acc;
}

// J_k[xw] = sum_P s_k[P] charges[P] B_k[P][xw]; grid (element blocks, nk).
__global__ __launch_bounds__(kThreads) void kp_coulomb_kernel(const double2* __restrict__ b, long long kstride,
                                                              int naux, long long nn, const double* __restrict__ sig,
                                                              long long sig_kstride, const double2* __restrict__ charges,
                                                              double2* __restrict__ j) {
  const int k = blockIdx.y;
  const double2* bk = b + (long long)k * kstride;
  const double* sk = sig + (long long)k * sig_kstride;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < nn; i += (long long)gridDim.x * kThreads) {
    double2 acc = make_double2(0.0, 0.0);
    for (int p = 0; p < naux; ++p) {
      const double2 c = charges[p];
      const double s = sk[p];
      const double2 v = cmul(make_double2(s * c.x, s * c.y), bk[(long long)p * nn + i]);
      acc.x += v.x;
      acc.y += v.y;
    }
    j[(long long)k * nn + i] = acc;
  }
}

// Bh[P][r][c] ([npad2][npad2], zero past 2n) of a complex [naux][n][n] block.
__global__ __launch_bounds__(kThreads) void kp_realify_kernel(const double2* __restrict__ b, int naux, int n,
                                                              int npad2, double* __restrict__ out) {
  const long long plane = (long long)npad2 * npad2;
  const long long total = (long long)naux * plane;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const long long p = idx / plane;
    const int r = (int)((idx % plane) / npad2);
    const int c = (int)(idx % npad2);
    double v = 0.0;
    if (r < 2 * n && c < 2 * n) {
      const int rr = r < n ? r : r - n, cc = c < n ? c : c - n;
      const double2 z = b[(p * n + rr) * n + cc];
      v = (r < n) == (c < n) ? z.x : (r < n ? -z.y : z.y);
    }
    out[idx] = v;
  }
}

// Ch^T [ldo][npad2] and the column signs [ldo] from the eigenpairs of the
// column-major view of a Hermitian row-major D (that view is conj(D), so its
// eigenvectors u give D's as conj(u)). Columns: the nneg most negative then
// the npos most positive eigenvalues (ascending order in w); scale
// sqrt(weight |l|); zero past r = nneg + npos and past 2n.
__global__ __launch_bounds__(kThreads) void kp_factor_kernel(const double2* __restrict__ u, const double* __restrict__ w,
                                                             int n, int nneg, int npos, double weight, int npad2,
                                                             int ldo, double* __restrict__ ct, double* __restrict__ sg) {
  const long long total = (long long)ldo * npad2;
  const int r = nneg + npos;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int i = (int)(idx / npad2);
    const int k = (int)(idx % npad2);
    double v = 0.0;
    if (i < r && k < 2 * n) {
      const int jj = i < nneg ? i : n - npos + (i - nneg);
      const double scale = sqrt(weight * fabs(w[jj]));
      const double2 z = u[(long long)jj * n + (k < n ? k : k - n)];
      v = k < n ? scale * z.x : -scale * z.y;
    }
    ct[idx] = v;
    if (k == 0) sg[i] = i < nneg ? -1.0 : 1.0;
  }
}

// K[u][v] (+)= (Kh[u][v] + Kh[n+u][n+v]) + i (Kh[n+u][v] - Kh[u][n+v]); Kh is [2n][2n].
__global__ __launch_bounds__(kThreads) void kp_fold_kernel(const double* __restrict__ kh, int n, int accumulate,
                                                           double2* __restrict__ k) {
  const long long nn = (long long)n * n;
  const long long m = 2LL * n;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < nn;
       idx += (long long)gridDim.x * kThreads) {
    const long long u = idx / n, v = idx % n;
    const double re = kh[u * m + v] + kh[(n + u) * m + n + v];
    const double im = kh[(n + u) * m + v] - kh[u * m + n + v];
    if (accumulate) {
      const double2 old = k[idx];
      k[idx] = make_double2(old.x + re, old.y + im);
    } else {
      k[idx] = make_double2(re, im);
    }
  }
}

// c (+)= alpha a b for complex row-major [n][n]; one thread per output element.
__global__ __launch_bounds__(kThreads) void kp_zgemm_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                                                            int n, double alpha, int accumulate,
                                                            double2* __restrict__ c) {
  const long long nn = (long long)n * n;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < nn;
       idx += (long long)gridDim.x * kThreads) {
    const long long r = idx / n, col = idx % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int q = 0; q < n; ++q) {
      const double2 v = cmul(a[r * n + q], b[(long long)q * n + col]);
      acc.x += v.x;
      acc.y += v.y;
    }
    acc.x *= alpha;
    acc.y *= alpha;
    if (accumulate) {
      const double2 old = c[idx];
      acc.x += old.x;
      acc.y += old.y;
    }
    c[idx] = acc;
  }
}

// stats[0] = max |D - D^dagger|, stats[1] = max |D| (bits of doubles, atomic max; clear first).
__global__ __launch_bounds__(kThreads) void kp_hermitian_stats_kernel(const double2* __restrict__ d, int n,
                                                                      unsigned long long* __restrict__ stats) {
  double asym = 0.0, mag = 0.0;
  const long long nn = (long long)n * n;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < nn;
       idx += (long long)gridDim.x * kThreads) {
    const long long r = idx / n, c = idx % n;
    const double2 a = d[idx], t = d[c * n + r];
    const double a1 = hypot(a.x - t.x, a.y + t.y);
    const double m = hypot(a.x, a.y);
    asym = a1 > asym ? a1 : asym;
    mag = m > mag ? m : mag;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double a = __shfl_xor(asym, off);
    const double m = __shfl_xor(mag, off);
    asym = a > asym ? a : asym;
    mag = m > mag ? m : mag;
  }
  if ((threadIdx.x & 63) == 0) {
    if (asym > 0.0) atomicMax(&stats[0], (unsigned long long)__double_as_longlong(asym));
    if (mag > 0.0) atomicMax(&stats[1], (unsigned long long)__double_as_longlong(mag));
  }
}

// H = S + i A of a real row-major D (S, A its symmetric and antisymmetric parts), complex row-major.
__global__ __launch_bounds__(kThreads) void hermitian_of_real_kernel(const double* __restrict__ d, int n,
                                                                     double2* __restrict__ h) {
  const long long nn = (long long)n * n;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < nn;
       idx += (long long)gridDim.x * kThreads) {
    const long long r = idx / n, c = idx % n;
    const double a = d[idx], b = d[c * n + r];
    h[idx] = make_double2(0.5 * (a + b), 0.5 * (a - b));
  }
}

// Columns [j0, j0 + rb) of the retained eigenpairs of H (the nneg lowest,
// then the npos highest; u column-major from zheevd of the view conj(H), so
// H's eigenvectors are conj(u)) as the real orbitals of the non-symmetric
// exchange: c [n][2 hb] = [Re V |l|^1/2 (rb columns), 0, Im V |l|^1/2, 0]
// with hb the half width, and the column signs of its two SYRKs:
// sym = (sign l, +1, sign l, +1), cross = (-sign l, +1, +sign l, +1).
__global__ __launch_bounds__(kThreads) void general_factor_kernel(const double2* __restrict__ u,
                                                                  const double* __restrict__ w, int n, int nneg,
                                                                  int npos, int j0, int rb, int hb,
                                                                  double* __restrict__ c, double* __restrict__ sym,
                                                                  double* __restrict__ cross) {
  const int ld = 2 * hb;
  const long long total = (long long)n * ld;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * kThreads) {
    const int m = (int)(idx / ld);
    const int col = (int)(idx % ld);
    const int half = col >= hb ? 1 : 0;
    const int i = col - half * hb;
    double v = 0.0;
    double sg = 1.0;
    if (i < rb) {
      const int k = j0 + i;  // index among the retained pairs
      const int jj = k < nneg ? k : n - npos + (k - nneg);
      const double l = w[jj];
      const double2 z = u[(long long)jj * n + m];
      const double scale = sqrt(fabs(l));
      v = half ? -scale * z.y : scale * z.x;
      sg = l < 0.0 ? -1.0 : 1.0;
    }
    c[idx] = v;
    if (m == 0) {
      sym[col] = sg;
      cross[col] = i < rb ? (half ? sg : -sg) : 1.0;
    }
  }
}

}  // namespace

hipError_t launch_kp_charges(const void* b, long long kstride, int nk, int naux, int nao, const void* density,
                             const double* weights, void* inner, void* charges, hipStream_t s) {
  const long long nn = (long long)nao * nao;
  hipLaunchKernelGGL(kp_inner_kernel, dim3((unsigned)((long long)nk * naux)), dim3(kThreads), 0, s,
                     (const double2*)b, kstride, naux, nn, (const double2*)density, weights, (double2*)inner);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(kp_charges_kernel, dim3((unsigned)((naux + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                     (const double2*)inner, nk, naux, (double2*)charges);
  return hipGetLastError();
}

hipError_t launch_kp_coulomb(const void* b, long long kstride, int nk, int naux, int nao, const double* sig,
                             long long sig_kstride, const void* charges, void* out, hipStream_t s) {
  const long long nn = (long long)nao * nao;
  const unsigned gx = (unsigned)std::min<long long>((nn + kThreads - 1) / kThreads, 1 << 16);
  hipLaunchKernelGGL(kp_coulomb_kernel, dim3(gx, (unsigned)nk), dim3(kThreads), 0, s, (const double2*)b, kstride,
                     naux, nn, sig, sig_kstride, (const double2*)charges, (double2*)out);
  return hipGetLastError();
}

hipError_t launch_kp_realify(const void* b, int naux, int nao, double* out, hipStream_t s) {
  const int npad2 = n_tiles(2 * nao) * kTile;
  hipLaunchKernelGGL(kp_realify_kernel, dim3(grid_for((long long)naux * npad2 * npad2)), dim3(kThreads), 0, s,
                     (const double2*)b, naux, nao, npad2, out);
  return hipGetLastError();
}

hipError_t launch_kp_factor(const void* eigvecs, const double* eigvals, int nao, int nneg, int npos, double weight,
                            int ldo, double* ct, double* signs, hipStream_t s) {
  const int npad2 = n_tiles(2 * nao) * kTile;
  hipLaunchKernelGGL(kp_factor_kernel, dim3(grid_for((long long)ldo * npad2)), dim3(kThreads), 0, s,
                     (const double2*)eigvecs, eigvals, nao, nneg, npos, weight, npad2, ldo, ct, signs);
  return hipGetLastError();
}

hipError_t launch_kp_fold(const double* kh, int nao, bool accumulate, void* k, hipStream_t s) {
  hipLaunchKernelGGL(kp_fold_kernel, dim3(grid_for((long long)nao * nao)), dim3(kThreads), 0, s, kh, nao,
                     accumulate ? 1 : 0, (double2*)k);
  return hipGetLastError();
}

hipError_t launch_kp_zgemm(const void* a, const void* b, int nao, double alpha, bool accumulate, void* c,
                           hipStream_t s) {
  hipLaunchKernelGGL(kp_zgemm_kernel, dim3(grid_for((long long)nao * nao)), dim3(kThreads), 0, s,
                     (const double2*)a, (const double2*)b, nao, alpha, accumulate ? 1 : 0, (double2*)c);
  return hipGetLastError();
}

hipError_t launch_kp_hermitian_stats(const void* d, int nao, unsigned long long* stats, hipStream_t s) {
  hipLaunchKernelGGL(kp_hermitian_stats_kernel, dim3(grid_for((long long)nao * nao)), dim3(kThreads), 0, s,
                     (const double2*)d, nao, stats);
  return hipGetLastError();
}

hipError_t launch_hermitian_of_real(const double* d, int nao, void* h, hipStream_t s) {
  hipLaunchKernelGGL(hermitian_of_real_kernel, dim3(grid_for((long long)nao * nao)), dim3(kThreads), 0, s, d, nao,
                     (double2*)h);
  return hipGetLastError();
}

hipError_t launch_general_factor(const void* eigvecs, const double* eigvals, int nao, int nneg, int npos, int j0,
                                 int rb, int hb, double* c, double* sym, double* cross, hipStream_t s) {
  hipLaunchKernelGGL(general_factor_kernel, dim3(grid_for((long long)nao * 2 * hb)), dim3(kThreads), 0, s,
                     (const double2*)eigvecs, eigvals, nao, nneg, npos, j0, rb, hb, c, sym, cross);
  return hipGetLastError();
}

}  // namespace hfjk
