// C ABI of the DF-J/K library: argument checks, workspace carving and the
// launch sequences of J and K. No exception crosses this boundary; every
// entry point returns a status and leaves a thread-local message.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>

#include "hfjk_internal.h"

// Scratch of the allocation-free exchange (hfjk_exchange_occ_noalloc): one
// copy per GPU the library runs on, held until the process ends. Host code
// never passes this name to a kernel; it uses the per-device address that
// hipGetSymbolAddress returns at context creation.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

thread_local std::string g_last_error;

struct Failure {
  int code;
  std::string message;
};

[[noreturn]] void fail(int code, const std::string& message) { throw Failure{code, message}; }

void check_hip(hipError_t err, const char* what) {
  if (err != hipSuccess) fail(HFJK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
}

template <typename F>
int guarded(F&& body) {
  try {
    body();
    g_last_error.clear();
    return HFJK_OK;
  } catch (const Failure& f) {
    g_last_error = f.message;
    return f.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return HFJK_ERR_HIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return HFJK_ERR_HIP;
  } catch (...) {
    g_last_error = "unknown failure";
    return HFJK_ERR_HIP;
  }
}

const char* precision_name(int precision) {
  switch (precision) {
    case HFJK_FP64: return "fp64";
    case HFJK_FP32: return "fp32";
    case HFJK_BF16X3: return "bf16x3";
    default: return "unknown";
  }
}

void check_precision(int precision) {
  if (precision != HFJK_FP64 && precision != HFJK_FP32 && precision != HFJK_BF16X3)
    fail(HFJK_ERR_INPUT, "precision must be HFJK_FP64, HFJK_FP32 or HFJK_BF16X3");
}

void check_ctx(const hfjk_ctx* ctx) {
  if (ctx == nullptr) fail(HFJK_ERR_INPUT, "null context");
}

void check_bound(const hfjk_ctx* ctx, int precision) {
  if (ctx->naux <= 0 || ctx->nao <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
  const void* b = precision == HFJK_FP64 ? (const void*)ctx->b64
                  : precision == HFJK_FP32 ? (const void*)ctx->b32
                                           : (const void*)ctx->b16;
  if (b == nullptr)
    fail(HFJK_ERR_BACKEND, std::string("no ") + precision_name(precision) + " copy of the fitted tensor is bound");
}

const void* bound_tensor(const hfjk_ctx* ctx, int precision) {
  return precision == HFJK_FP64 ? (const void*)ctx->b64
         : precision == HFJK_FP32 ? (const void*)ctx->b32
                                  : (const void*)ctx->b16;
}

size_t elem_size(int precision) { return precision == HFJK_FP64 ? 8 : 4; }

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

// Split counts (deterministic functions of the shapes).
int coulomb_splits(long long ntiles, int naux) {
  long long s = (2048 + ntiles - 1) / ntiles;
  long long cap = std::max(1, naux / 4);
  return (int)std::max(1LL, std::min(s, cap));
}

// Aux-slice splits of the SYRK: a multiple of 8 (one XCD per split group),
// about 2048 workgroups in total, never more slices than 8 * ceil(naux / 8).
int syrk_splits(long long npairs, int naux) {
  long long s = (2048 + npairs - 1) / npairs;
  s = std::min(s, (long long)naux);
  return (int)(((std::max(1LL, s) + 7) / 8) * 8);
}

// Exchange operands in BF16X3 are not implemented by a dedicated kernel yet;
// the exchange in that precision runs on the FP32 tensor copy through the
// FP32 MFMA path only when that copy is bound (see exchange_precision).

struct Layout {
  // density factorisation region
  size_t dens_a, dens_w, dens_e, dens_info, dens_stats, dens_c, dens_signs, dens_end;
  // coulomb region (shares the base of the exchange region)
  size_t cw, cgp, cg, cjp, c_end;
  // exchange region (xct: C^T in FP64 for the half-transform from the full copy)
  size_t xc, xs, xx, xp, xct, x_end;
  // the Coulomb half of hfjk_jk_occ, after both regions (it runs while the
  // SYRK still reads X): row-group parts of the charges [naux][charge_groups],
  // gamma [naux], an FP32 padded orbital copy for BF16X3 (whose exchange keeps
  // only the bf16 planes of C^T), and the J partials
  size_t qt, qg, qc, qj, q_end;
  size_t total;
  int nsplit_j, nsplit_k, nslot_k, ldo, xrows, npairs;
  long long p, ntiles;
};

Layout make_layout(const hfjk_ctx* ctx, int nocc, int precision, int nsplit_k = 0) {
  Layout L{};
  const int n = ctx->nao, naux = ctx->naux;
  const size_t es = elem_size(precision);
  L.p = hfjk::packed_elems(n);
  L.ntiles = hfjk::n_tile_pairs(n);
  L.ldo = hfjk::occ_ld(std::max(nocc, 1));
  L.xrows = hfjk::x_rows(n);
  const int nb = L.xrows / hfjk::kBlock;
  (void)nb;
  L.npairs = hfjk::syrk_pairs(n);
  L.nsplit_j = coulomb_splits(L.ntiles, naux);
  L.nsplit_k = nsplit_k > 0 ? nsplit_k
               : ctx->syrk_splits > 0 ? ((ctx->syrk_splits + 7) / 8) * 8
                                       : syrk_splits(L.npairs, naux);
  L.nslot_k = hfjk::syrk_slots(precision, naux, L.ldo, L.nsplit_k);
  size_t off = 0;
  const size_t nn = (size_t)n * n;
  L.dens_a = off; off = align_up(off + nn * 8);
  L.dens_w = off; off = align_up(off + (size_t)n * 8);
  L.dens_e = off; off = align_up(off + (size_t)n * 8);
  L.dens_info = off; off = align_up(off + 8);
  L.dens_stats = off; off = align_up(off + 2 * 8);
  L.dens_c = off; off = align_up(off + nn * 8);
  L.dens_signs = off; off = align_up(off + (size_t)n * 8);
  L.dens_end = off;
  const size_t base = off;
  // coulomb
  off = base;
  L.cw = off; off = align_up(off + (size_t)L.p * 8);
  L.cgp = off; off = align_up(off + (size_t)L.ntiles * naux * 8);
  L.cg = off; off = align_up(off + (size_t)naux * 8);
  L.cjp = off; off = align_up(off + (size_t)L.nsplit_j * L.p * 8);
  L.c_end = off;
  // exchange
  off = base;
  const size_t npad = (size_t)hfjk::n_tiles(n) * hfjk::kTile;
  L.xc = off; off = align_up(off + npad * L.ldo * es);
  L.xs = off; off = align_up(off + (size_t)L.ldo * 8);
  L.xx = off; off = align_up(off + (size_t)naux * L.xrows * L.ldo * es);
  const size_t tile2 = (size_t)hfjk::syrk_tile() * hfjk::syrk_tile();
  L.xp = off; off = align_up(off + (size_t)L.nslot_k * L.npairs * tile2 * es);
  L.xct = off; off = align_up(off + npad * L.ldo * 8);
  L.x_end = off;
  off = std::max(L.c_end, L.x_end);
  L.qt = off; off = align_up(off + (size_t)naux * hfjk::charge_groups(n) * 8);
  L.qg = off; off = align_up(off + (size_t)naux * 8);
  L.qc = off; off = align_up(off + npad * L.ldo * 4);
  L.qj = off; off = align_up(off + (size_t)L.nsplit_j * L.p * 8);
  L.q_end = off;
  L.total = L.q_end;
  return L;
}

void check_workspace_size(size_t need, void* ws, size_t bytes) {
  if (ws == nullptr && need > 0) fail(HFJK_ERR_WORKSPACE, "null workspace");
  if (bytes < need)
    fail(HFJK_ERR_WORKSPACE, "workspace of " + std::to_string(bytes) + " bytes is smaller than the " +
                                 std::to_string(need) + " bytes the workspace query reports");
  if (((uintptr_t)ws & 255) != 0) fail(HFJK_ERR_WORKSPACE, "workspace must be 256-byte aligned");
}

void check_workspace(const Layout& L, void* ws, size_t bytes) { check_workspace_size(L.total, ws, bytes); }

template <typename T>
T* at(void* ws, size_t off) {
  return (T*)((char*)ws + off);
}

constexpr size_t kMaxStages = 4096;

// RAII stage timer: records an event pair around the launches in its scope.
// Set for the duration of an allocation-free call: stage timers (which
// create events on first use) stay off.
thread_local bool g_no_alloc_call = false;

struct Stage {
  hfjk_ctx* ctx;
  hipStream_t s;
  hfjk::StageTimer* timer = nullptr;
  Stage(hfjk_ctx* c, hipStream_t stream, const char* name) : ctx(c), s(stream) {
    if (!ctx->profile || g_no_alloc_call || ctx->n_stages >= kMaxStages) return;
    if (ctx->n_stages == ctx->stages.size()) {
      hfjk::StageTimer t;
      check_hip(hipEventCreate(&t.start), "hipEventCreate");
      check_hip(hipEventCreate(&t.stop), "hipEventCreate");
      ctx->stages.push_back(t);
    }
    timer = &ctx->stages[ctx->n_stages++];
    timer->name = name;
    check_hip(hipEventRecord(timer->start, s), "hipEventRecord");
  }
  ~Stage() {
    // A failed record leaves the stage unreadable; hfjk_profile_read reports it.
    if (timer != nullptr) (void)hipEventRecord(timer->stop, s);
  }
};

void begin_call(hfjk_ctx* ctx) {
  check_hip(hipSetDevice(ctx->device), "hipSetDevice");
}

void run_coulomb(hfjk_ctx* ctx, const double* density, double* out, int precision, void* ws, const Layout& L,
                 hipStream_t s) {
  const void* b = bound_tensor(ctx, precision);
  double* w = at<double>(ws, L.cw);
  double* gp = at<double>(ws, L.cgp);
  double* g = at<double>(ws, L.cg);
  double* jp = at<double>(ws, L.cjp);
  {
    Stage st(ctx, s, "coulomb_weights");
    check_hip(hfjk::launch_weights(density, ctx->nao, w, s), "weights launch");
  }
  {
    Stage st(ctx, s, "coulomb_gamma");
    check_hip(hfjk::launch_gamma(b, precision, L.p, ctx->naux, (int)L.ntiles, w, gp, s), "gamma launch");
    check_hip(hfjk::launch_gamma_reduce(gp, (int)L.ntiles, ctx->naux, g, s), "gamma reduce launch");
  }
  {
    Stage st(ctx, s, "coulomb_accumulate");
    check_hip(hfjk::launch_coulomb_accumulate(b, precision, L.p, ctx->naux, (int)L.ntiles, L.nsplit_j, g, jp, 0, s),
              "coulomb accumulate launch");
  }
  {
    Stage st(ctx, s, "coulomb_unpack");
    check_hip(hfjk::launch_coulomb_unpack(jp, L.nsplit_j, L.p, ctx->nao, out, s), "coulomb unpack launch");
  }
}

// X of the occupied-form exchange in the interleaved BF16X3 layout (hi and
// lo per 32-column block): written by the 256 x 256-tile half-transform from
// the full copy and read by the 256 x 256-tile SYRK and the charges, when
// both kernels run for these arguments; the hi and lo planes otherwise.
bool x16_interleaved(const hfjk_ctx* ctx, int precision, bool is_signed, int ldo) {
  return precision == HFJK_BF16X3 && ctx->b16_full != nullptr && ctx->x_full == 0 &&
         hfjk::x16_tiles256(ldo, ctx->x16_bn) && hfjk::syrk16_tiles256(ctx->nao, ldo, is_signed, ctx->syrk_kind, 0);
}

// Orbitals (and signs) into the workspace and the half-transform X = B C.
void exchange_front(hfjk_ctx* ctx, const double* orbitals, int nocc, const double* signs, int precision, void* ws,
                    const Layout& L, hipStream_t s) {
  // FP64 and FP32 run the same MFMA kernels on their tensor copy; BF16X3
  // runs the split-product kernels on the bf16 hi/lo planes. The FP32 and
  // BF16X3 buffers have the same sizes (4 bytes per element).
  check_bound(ctx, precision);
  const void* b = bound_tensor(ctx, precision);
  void* cp = at<void>(ws, L.xc);
  double* sg = at<double>(ws, L.xs);
  void* x = at<void>(ws, L.xx);
  const bool split = precision == HFJK_BF16X3;
  // FP64 / FP32 with a bound full copy of B: the half-transform on the direct-LDS ring.
  const void* bfull = precision == HFJK_FP64 ? (const void*)ctx->b64_full
                      : precision == HFJK_FP32 ? (const void*)ctx->b32_full
                                               : (const void*)ctx->b16_full;
  const bool full = bfull != nullptr && ctx->x_full == 0;
  // BF16X3 takes C^T's bf16 hi/lo (cp; interleaved for the ring from the
  // full copy) as the second operand.
  void* ct = split ? cp : at<void>(ws, L.xct);
  {
    Stage st(ctx, s, "exchange_prepare");
    if (split)
      check_hip(hfjk::launch_orbitals_bf16x3(orbitals, ctx->nao, nocc, cp, s, full), "orbitals launch");
    else
      check_hip(hfjk::launch_orbitals(orbitals, ctx->nao, nocc, precision, cp, s), "orbitals launch");
    if (full && !split)
      check_hip(hfjk::launch_orbitals_t(orbitals, ctx->nao, nocc, precision, ct, s), "orbitals launch");
    check_hip(hfjk::launch_occ_signs(signs, nocc, L.ldo, sg, s), "signs launch");
  }
  {
    Stage st(ctx, s, "exchange_half_transform");
    if (full)
      check_hip(hfjk::launch_half_transform_full(bfull, precision, ctx->naux, ctx->nao, ct, L.ldo, x, s,
                                                 ctx->x16_bn, x16_interleaved(ctx, precision, signs != nullptr, L.ldo)),
                "half transform launch");
    else if (split)
      check_hip(hfjk::launch_half_transform_bf16x3((const uint16_t*)b, L.p, ctx->naux, ctx->nao, cp, L.ldo, x,
                                                   ctx->x16_bn, s),
                "half transform launch");
    else
      check_hip(hfjk::launch_half_transform(b, precision, L.p, ctx->naux, ctx->nao, cp, L.ldo, x, ctx->x_bn,
                                            ctx->x_bm == 0 ? 64 : ctx->x_bm, s),
                "half transform launch");
  }
}

// K = sum_P X_P S X_P^T from the X that exchange_front left in the workspace.
void exchange_back(hfjk_ctx* ctx, bool is_signed, double* out, int precision, void* ws, const Layout& L,
                   hipStream_t s) {
  double* sg = at<double>(ws, L.xs);
  void* x = at<void>(ws, L.xx);
  void* part = at<void>(ws, L.xp);
  {
    Stage st(ctx, s, "exchange_syrk");
    check_hip(hfjk::launch_syrk(x, precision, ctx->naux, ctx->nao, L.ldo, sg, is_signed, L.nsplit_k, part,
                                ctx->syrk_kind, s, nullptr, 0, x16_interleaved(ctx, precision, is_signed, L.ldo)),
              "syrk launch");
  }
  {
    Stage st(ctx, s, "exchange_reduce");
    const int rprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;  // BF16X3 partials are FP32
    check_hip(hfjk::launch_syrk_reduce(part, rprec, L.nslot_k, ctx->nao, out, s), "syrk reduce launch");
  }
}

void run_exchange(hfjk_ctx* ctx, const double* orbitals, int nocc, const double* signs, double* out, int precision,
                  void* ws, const Layout& L, hipStream_t s) {
  exchange_front(ctx, orbitals, nocc, signs, precision, ws, L, s);
  exchange_back(ctx, signs != nullptr, out, precision, ws, L, s);
}

// Grid cap of the Coulomb kernels that hfjk_jk_occ runs beside the SYRK. In
// FP64 two workgroups per CU by default (large configuration, one box:
// 140.3 ms per build against 144.5 at one per CU, 141.0 uncapped and 141.8
// with the Coulomb half after the SYRK). The FP32 and BF16x3 SYRKs are too
// short to hide a capped Coulomb half (it reads B in FP64 or its FP32 copy
// either way), so they run it uncapped. 0 = no cap.
int side_groups(const hfjk_ctx* ctx, int precision) {
  if (ctx->side_groups > 0) return ctx->side_groups;
  return precision == HFJK_FP64 ? 2 * ctx->cus : 0;
}

rocblas_handle blas_handle(hfjk_ctx* ctx) {
  if (ctx->blas == nullptr) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) fail(HFJK_ERR_SOLVER, "rocblas_create_handle failed");
    ctx->blas = h;
  }
  return (rocblas_handle)ctx->blas;
}

// Layout of the non-symmetric exchange from offset `base` of the workspace,
// for aux chunks of `pc` functions and column blocks of at most kGeneralCols
// retained eigenpairs.
constexpr int kGeneralCols = 1024;

struct GeneralLayout {
  size_t hz, w, e, info, c, sym, cross, cp, ct, x, p1, p2, total;
  int ld, npairs, nsplit, nslot, xrows;
};

GeneralLayout general_layout(const hfjk_ctx* ctx, int precision, int rank, int pc, size_t base) {
  GeneralLayout G{};
  const size_t n = (size_t)ctx->nao;
  const size_t npad = (size_t)hfjk::n_tiles(ctx->nao) * hfjk::kTile;
  const size_t es = precision == HFJK_FP64 ? 8 : 4;
  const int rb = std::max(1, std::min(rank, kGeneralCols));
  G.ld = 2 * hfjk::occ_ld(rb);
  G.xrows = hfjk::x_rows(ctx->nao);
  G.npairs = hfjk::syrk_pairs(ctx->nao);
  // Split counts capped so the two partial sets stay small for small matrices.
  G.nsplit = std::min(syrk_splits(G.npairs, pc), std::max(8, ((512 / G.npairs + 7) / 8) * 8));
  G.nslot = hfjk::syrk_slots(precision, pc, G.ld, G.nsplit);
  const size_t tile2 = (size_t)hfjk::syrk_tile() * hfjk::syrk_tile();
  size_t off = base;
  G.hz = off; off = align_up(off + n * n * 16);
  G.w = off; off = align_up(off + n * 8);
  G.e = off; off = align_up(off + n * 8);
  G.info = off; off = align_up(off + 8);
  G.c = off; off = align_up(off + n * G.ld * 8);
  G.sym = off; off = align_up(off + (size_t)G.ld * 8);
  G.cross = off; off = align_up(off + (size_t)G.ld * 8);
  G.cp = off; off = align_up(off + npad * G.ld * es);
  G.ct = off; off = align_up(off + npad * G.ld * es);
  G.x = off; off = align_up(off + (size_t)pc * G.xrows * G.ld * es);
  G.p1 = off; off = align_up(off + (size_t)G.nslot * G.npairs * tile2 * es);
  G.p2 = off; off = align_up(off + (size_t)G.nslot * G.npairs * tile2 * es);
  G.total = off;
  return G;
}

__global__ void zero_fill_kernel(double* __restrict__ out, long long count) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
    out[i] = 0.0;
}

// K = sum_P B_P D B_P for a real non-symmetric D (the density form of
// coulomb_exchange.py:184-190) on the library's own kernels, in any
// precision. With S and A the symmetric and antisymmetric parts of D,
// H = S + i A is Hermitian, H = V diag(l) V^H (rocSOLVER zheevd, |l| <= 1e-12
// max |l| dropped), and for a symmetric B_P and X = B_P V |l|^1/2:
// B_P H B_P = X diag(sign l) X^H, whose real part is B_P S B_P and whose
// imaginary part is B_P A B_P. With X' = B [Re, 0, Im, 0] of those columns
// (the occupied-form half-transform on 2 hb columns):
//   B S B = sum_k X'[m][k] s_k X'[n][k]                  (the signed SYRK),
//   B A B = sum_k X'[m][k] t_k X'[n][(k + hb) mod 2 hb]  (its column-shifted form),
// s = (sign l, sign l), t = (-sign l, sign l) over the two halves, and the
// lower tiles give K[m][n] = S + A, their mirrors K[n][m] = S - A. Retained
// eigenpairs in blocks of at most kGeneralCols (2 hb <= 2048, the ring's
// limit), aux functions in chunks that fit the workspace beyond `base`;
// every block and chunk adds into K. 16 n^2 naux r flops for rank r.
void exchange_general(hfjk_ctx* ctx, const double* density, double* out, int precision, void* ws, size_t ws_bytes,
                      size_t base, hipStream_t s) {
  check_bound(ctx, precision);
  const int n = ctx->nao, naux = ctx->naux;
  const long long nn = (long long)n * n;
  const long long p = hfjk::packed_elems(n);
  const long long npad = (long long)hfjk::n_tiles(n) * hfjk::kTile;
  // Chunk of aux functions: the most that fit (for full-rank column blocks).
  GeneralLayout G = general_layout(ctx, precision, n, 1, base);
  if (G.total > ws_bytes)
    fail(HFJK_ERR_WORKSPACE, "workspace too small for the non-symmetric exchange (" + std::to_string(G.total) +
                                 " bytes for one aux function)");
  int pc = 1;
  for (int lo = 1, hi = naux; lo <= hi;) {
    const int mid = lo + (hi - lo) / 2;
    if (general_layout(ctx, precision, n, mid, base).total <= ws_bytes) {
      pc = mid;
      lo = mid + 1;
    } else {
      hi = mid - 1;
    }
  }
  if (ctx->general_chunk > 0) pc = std::min(pc, ctx->general_chunk);
  G = general_layout(ctx, precision, n, pc, base);
  auto* hz = at<rocblas_double_complex>(ws, G.hz);
  double* w = at<double>(ws, G.w);
  double* e = at<double>(ws, G.e);
  rocblas_int* info = at<rocblas_int>(ws, G.info);
  {
    Stage st(ctx, s, "general_eigh");
    check_hip(hfjk::launch_hermitian_of_real(density, n, hz, s), "hermitian launch");
    rocblas_handle h = blas_handle(ctx);
    if (rocblas_set_stream(h, s) != rocblas_status_success) fail(HFJK_ERR_SOLVER, "rocblas_set_stream failed");
    // The column-major view of the row-major Hermitian H is conj(H); its
    // eigenvectors are conjugated when the columns are formed.
    if (rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n, hz, n, w, e, info) !=
        rocblas_status_success)
      fail(HFJK_ERR_SOLVER, "rocsolver_zheevd failed to launch");
  }
  std::vector<double> vals(n);
  rocblas_int host_info = 0;
  check_hip(hipMemcpyAsync(vals.data(), w, sizeof(double) * n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipMemcpyAsync(&host_info, info, sizeof(host_info), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (host_info != 0) fail(HFJK_ERR_SOLVER, "rocsolver_zheevd did not converge (info " + std::to_string(host_info) + ")");
  double top = 0.0;
  for (double v : vals) top = std::max(top, std::fabs(v));
  const double cut = 1e-12 * top;
  int nneg = 0, npos = 0;
  for (int j = 0; j < n && vals[j] < -cut; ++j) ++nneg;
  for (int j = n - 1; j >= nneg && vals[j] > cut; --j) ++npos;
  const int r = nneg + npos;
  if (r == 0) {
    hipLaunchKernelGGL(zero_fill_kernel, dim3(hfjk::grid_for(nn)), dim3(256), 0, s, out, nn);
    check_hip(hipGetLastError(), "zero launch");
    return;
  }
  const bool split = precision == HFJK_BF16X3;
  const void* bfull = precision == HFJK_FP64 ? (const void*)ctx->b64_full
                      : precision == HFJK_FP32 ? (const void*)ctx->b32_full
                                               : (const void*)ctx->b16_full;
  const bool full = bfull != nullptr && ctx->x_full == 0;
  const void* b = bound_tensor(ctx, precision);
  const size_t es = precision == HFJK_FP64 ? 8 : 4;
  const int rprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
  double* c = at<double>(ws, G.c);
  double* sym = at<double>(ws, G.sym);
  double* cross = at<double>(ws, G.cross);
  void* cp = at<void>(ws, G.cp);
  void* ct = split ? cp : at<void>(ws, G.ct);
  void* x = at<void>(ws, G.x);
  void* p1 = at<void>(ws, G.p1);
  void* p2 = at<void>(ws, G.p2);
  bool accumulate = false;
  for (int j0 = 0; j0 < r; j0 += kGeneralCols) {
    const int rb = std::min(kGeneralCols, r - j0);
    const int hb = hfjk::occ_ld(rb);
    const int ld = 2 * hb;
    {
      Stage st(ctx, s, "general_prepare");
      check_hip(hfjk::launch_general_factor(hz, w, n, nneg, npos, j0, rb, hb, c, sym, cross, s),
                "general factor launch");
      if (split)
        check_hip(hfjk::launch_orbitals_bf16x3(c, n, ld, cp, s, full), "orbitals launch");
      else if (full)
        check_hip(hfjk::launch_orbitals_t(c, n, ld, precision, ct, s), "orbitals launch");
      else
        check_hip(hfjk::launch_orbitals(c, n, ld, precision, cp, s), "orbitals launch");
    }
    for (int q0 = 0; q0 < naux; q0 += pc) {
      const int qc = std::min(pc, naux - q0);
      const int nsplit = std::min(syrk_splits(G.npairs, qc), G.nsplit);
      const int nslot = hfjk::syrk_slots(precision, qc, ld, nsplit);
      {
        Stage st(ctx, s, "general_half_transform");
        if (full && split) {
          const uint16_t* bh = (const uint16_t*)bfull;
          check_hip(hfjk::launch_half_transform_full(bh + 2LL * q0 * npad * npad, precision, qc, n, ct, ld, x, s,
                                                     ctx->x16_bn),
                    "half transform launch");
        } else if (full) {
          check_hip(hfjk::launch_half_transform_full((const char*)bfull + q0 * npad * npad * (long long)es, precision,
                                                     qc, n, ct, ld, x, s),
                    "half transform launch");
        } else if (split) {
          const uint16_t* bh = (const uint16_t*)b;
          check_hip(hfjk::launch_half_transform_bf16x3(bh + q0 * p, p, qc, n, cp, ld, x, ctx->x16_bn, s,
                                                       bh + ((long long)naux + q0) * p),
                    "half transform launch");
        } else {
          check_hip(hfjk::launch_half_transform((const char*)b + q0 * p * (long long)es, precision, p, qc, n, cp, ld,
                                                x, ctx->x_bn, ctx->x_bm == 0 ? 64 : ctx->x_bm, s),
                    "half transform launch");
        }
      }
      {
        Stage st(ctx, s, "general_syrk");
        check_hip(hfjk::launch_syrk(x, precision, qc, n, ld, sym, true, nsplit, p1, 0, s), "syrk launch");
        check_hip(hfjk::launch_syrk(x, precision, qc, n, ld, cross, true, nsplit, p2, 0, s, nullptr, hb),
                  "syrk launch");
      }
      {
        Stage st(ctx, s, "general_reduce");
        check_hip(hfjk::launch_syrk_reduce2(p1, p2, rprec, nslot, n, accumulate, out, s), "reduce launch");
      }
      accumulate = true;
    }
  }
}

__global__ void density_stats_kernel(const double* __restrict__ d, int n, unsigned long long* __restrict__ stats) {
  double asym = 0.0, mag = 0.0;
  const long long total = (long long)n * n;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int r = (int)(idx / n), c = (int)(idx % n);
    const double a = fabs(d[idx] - d[(long long)c * n + r]);
    const double m = fabs(d[idx]);
    asym = a > asym ? a : asym;
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

}  // namespace

namespace {

// Device address of `scratch` on `device`, looked up once per device with
// that device current (each GPU holds its own copy of the array).
void* noalloc_scratch(int device) {
  static std::mutex lock;
  static void* addresses[64] = {};
  if (device < 0 || device >= 64) fail(HFJK_ERR_INPUT, "device index beyond 63");
  std::lock_guard<std::mutex> guard(lock);
  if (addresses[device] == nullptr) {
    check_hip(hipSetDevice(device), "hipSetDevice");
    void* ptr = nullptr;
    check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (ptr == nullptr) fail(HFJK_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    addresses[device] = ptr;
  }
  return addresses[device];
}

}  // namespace

void hfjk_internal_set_error(const char* message) { g_last_error = message != nullptr ? message : ""; }

extern "C" {

int hfjk_version(void) { return 100; }

const char* hfjk_last_error(void) { return g_last_error.c_str(); }

int hfjk_create(hfjk_ctx** out, int device) {
  return guarded([&] {
    if (out == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    *out = nullptr;
    int count = 0;
    check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
      fail(HFJK_ERR_INPUT, "device " + std::to_string(device) + " out of range (" + std::to_string(count) + ")");
    hipDeviceProp_t prop;
    check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(HFJK_ERR_BACKEND, std::string("the DF-J/K kernels are built for gfx950 (MI355X), not ") + prop.gcnArchName);
    void* scratch_ptr = noalloc_scratch(device);
    hfjk_ctx* ctx = new hfjk_ctx();
    ctx->device = device;
    ctx->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    ctx->scratch = scratch_ptr;
    *out = ctx;
  });
}

int hfjk_destroy(hfjk_ctx* ctx) {
  return guarded([&] {
    if (ctx == nullptr) return;
    (void)hipSetDevice(ctx->device);
    for (auto& t : ctx->stages) {
      if (t.start) (void)hipEventDestroy(t.start);
      if (t.stop) (void)hipEventDestroy(t.stop);
    }
    if (ctx->blas != nullptr) rocblas_destroy_handle((rocblas_handle)ctx->blas);
    if (ctx->fork != nullptr) (void)hipEventDestroy(ctx->fork);
    if (ctx->join != nullptr) (void)hipEventDestroy(ctx->join);
    if (ctx->side != nullptr) (void)hipStreamDestroy(ctx->side);
    delete ctx;
  });
}

long long hfjk_packed_elems(int nao) { return nao > 0 ? hfjk::packed_elems(nao) : 0; }

int hfjk_pack(hfjk_ctx* ctx, const double* full, int naux, int nao, double* packed, double* stats, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    if (naux <= 0 || nao <= 0) fail(HFJK_ERR_INPUT, "the fitted tensor has shape [aux, ao, ao] with aux, ao > 0");
    if (full == nullptr || packed == nullptr) fail(HFJK_ERR_INPUT, "null tensor pointer");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    Stage st(ctx, s, "pack");
    if (stats != nullptr) check_hip(hipMemsetAsync(stats, 0, 2 * sizeof(double), s), "hipMemsetAsync");
    check_hip(hfjk::launch_pack(full, naux, nao, packed, (unsigned long long*)stats, s), "pack launch");
  });
}

int hfjk_generate(hfjk_ctx* ctx, uint64_t seed, int q0, int naux, int nao, double* packed, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    if (naux <= 0 || nao <= 0 || q0 < 0) fail(HFJK_ERR_INPUT, "naux, nao > 0 and q0 >= 0 are required");
    if (packed == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    Stage st(ctx, s, "generate");
    check_hip(hfjk::launch_generate(seed, q0, naux, nao, packed, s), "generate launch");
  });
}

int hfjk_normal_matrix(hfjk_ctx* ctx, uint64_t seed, int rows, int cols, uint32_t tag, double* out, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    if (rows < 0 || cols < 0) fail(HFJK_ERR_INPUT, "negative shape");
    if (out == nullptr && rows * (long long)cols > 0) fail(HFJK_ERR_INPUT, "null output pointer");
    begin_call(ctx);
    check_hip(hfjk::launch_normal_matrix(seed, rows, cols, tag, out, (hipStream_t)stream), "normal launch");
  });
}

int hfjk_split(hfjk_ctx* ctx, const double* packed, long long count, int precision, void* out, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    if (precision != HFJK_FP32 && precision != HFJK_BF16X3)
      fail(HFJK_ERR_INPUT, "hfjk_split makes HFJK_FP32 or HFJK_BF16X3 copies");
    if (count < 0) fail(HFJK_ERR_INPUT, "negative count");
    if (count > 0 && (packed == nullptr || out == nullptr)) fail(HFJK_ERR_INPUT, "null pointer");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    Stage st(ctx, s, "split");
    check_hip(hfjk::launch_split(packed, count, precision, out, s), "split launch");
  });
}

int hfjk_bind(hfjk_ctx* ctx, int naux, int nao, const double* b64, const float* b32, const uint16_t* b16) {
  return guarded([&] {
    check_ctx(ctx);
    if (naux <= 0 || nao <= 0) fail(HFJK_ERR_INPUT, "the fitted tensor has shape [aux, ao, ao] with aux, ao > 0");
    if (b64 == nullptr && b32 == nullptr && b16 == nullptr) fail(HFJK_ERR_INPUT, "no tensor copy given");
    if (nao > 46000) fail(HFJK_ERR_INPUT, "nao beyond the supported 46000");
    ctx->naux = naux;
    ctx->nao = nao;
    ctx->b64 = b64;
    ctx->b32 = b32;
    ctx->b16 = b16;
    ctx->b64_full = nullptr;  // a full copy belongs to the tensor it was expanded from
    ctx->b32_full = nullptr;
    ctx->b16_full = nullptr;
  });
}

int hfjk_expand_full(hfjk_ctx* ctx, const double* packed, int naux, int nao, int precision, void* full,
                     void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (naux <= 0 || nao <= 0) fail(HFJK_ERR_INPUT, "the fitted tensor has shape [aux, ao, ao] with aux, ao > 0");
    if (packed == nullptr || full == nullptr) fail(HFJK_ERR_INPUT, "null pointer");
    const long long npad = (long long)hfjk::n_tiles(nao) * hfjk::kTile;
    if (npad * npad * 8 >= (1LL << 31)) fail(HFJK_ERR_INPUT, "the full copy needs nao below 16384");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    Stage st(ctx, s, "expand_full");
    check_hip(hfjk::launch_expand_full(packed, hfjk::packed_elems(nao), naux, nao, precision, full, s),
              "expand launch");
  });
}

int hfjk_bind_full(hfjk_ctx* ctx, int precision, const void* full) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (full != nullptr && ctx->naux <= 0) fail(HFJK_ERR_INPUT, "bind the packed tensor first (hfjk_bind)");
    if (full != nullptr && (long long)hfjk::n_tiles(ctx->nao) * hfjk::kTile * hfjk::n_tiles(ctx->nao) * hfjk::kTile * 8 >=
                               (1LL << 31))
      fail(HFJK_ERR_INPUT, "the full copy needs nao below 16384");
    if (precision == HFJK_FP64)
      ctx->b64_full = (const double*)full;
    else if (precision == HFJK_FP32)
      ctx->b32_full = (const float*)full;
    else
      ctx->b16_full = (const uint16_t*)full;
  });
}

int hfjk_workspace_bytes(hfjk_ctx* ctx, int nocc, int precision, size_t* bytes) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (bytes == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    if (nocc < 0) fail(HFJK_ERR_INPUT, "negative nocc");
    const int kprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
    *bytes = make_layout(ctx, std::max(nocc, 1), kprec).total;
  });
}

int hfjk_coulomb_charges(hfjk_ctx* ctx, const double* charges, double* coulomb, int precision, void* workspace,
                         size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    check_bound(ctx, precision);
    if (charges == nullptr || coulomb == nullptr) fail(HFJK_ERR_INPUT, "null pointer");
    const Layout L = make_layout(ctx, 1, precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32);
    if (workspace_bytes < L.c_end || workspace == nullptr)
      fail(HFJK_ERR_WORKSPACE, "workspace smaller than the coulomb needs (" + std::to_string(L.c_end) + " bytes)");
    if (((uintptr_t)workspace & 255) != 0) fail(HFJK_ERR_WORKSPACE, "workspace must be 256-byte aligned");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    const void* b = bound_tensor(ctx, precision);
    double* jp = at<double>(workspace, L.cjp);
    {
      Stage st(ctx, s, "coulomb_accumulate");
      check_hip(hfjk::launch_coulomb_accumulate(b, precision, L.p, ctx->naux, (int)L.ntiles, L.nsplit_j, charges, jp, 0,
                                                s),
                "coulomb accumulate launch");
    }
    {
      Stage st(ctx, s, "coulomb_unpack");
      check_hip(hfjk::launch_coulomb_unpack(jp, L.nsplit_j, L.p, ctx->nao, coulomb, s), "coulomb unpack launch");
    }
  });
}

int hfjk_coulomb(hfjk_ctx* ctx, const double* density, double* coulomb, int precision, void* workspace,
                 size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    check_bound(ctx, precision);
    if (density == nullptr || coulomb == nullptr) fail(HFJK_ERR_INPUT, "null matrix pointer");
    const Layout L = make_layout(ctx, 1, precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32);
    if (workspace_bytes < L.c_end || workspace == nullptr)
      fail(HFJK_ERR_WORKSPACE, "workspace smaller than the coulomb needs (" + std::to_string(L.c_end) + " bytes)");
    if (((uintptr_t)workspace & 255) != 0) fail(HFJK_ERR_WORKSPACE, "workspace must be 256-byte aligned");
    begin_call(ctx);
    run_coulomb(ctx, density, coulomb, precision, workspace, L, (hipStream_t)stream);
  });
}

int hfjk_exchange_occ(hfjk_ctx* ctx, const double* orbitals, int nocc, const double* signs, double* exchange,
                      int precision, void* workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    if (nocc < 0) fail(HFJK_ERR_INPUT, "negative nocc");
    if (exchange == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    if (nocc > 0 && orbitals == nullptr) fail(HFJK_ERR_INPUT, "null orbitals pointer");
    const int kprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
    const Layout L = make_layout(ctx, std::max(nocc, 1), kprec);
    if (workspace_bytes < L.x_end || workspace == nullptr)
      fail(HFJK_ERR_WORKSPACE, "workspace smaller than the exchange needs (" + std::to_string(L.x_end) + " bytes)");
    if (((uintptr_t)workspace & 255) != 0) fail(HFJK_ERR_WORKSPACE, "workspace must be 256-byte aligned");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    if (nocc == 0) {
      check_hip(hipMemsetAsync(exchange, 0, sizeof(double) * ctx->nao * ctx->nao, s), "hipMemsetAsync");
      return;
    }
    run_exchange(ctx, orbitals, nocc, signs, exchange, precision, workspace, L, s);
  });
}

int hfjk_exchange_occ_noalloc(hfjk_ctx* ctx, const double* orbitals, int nocc, const double* signs,
                              double* exchange, int precision, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    check_bound(ctx, precision);
    if (nocc < 0) fail(HFJK_ERR_INPUT, "negative nocc");
    if (exchange == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    if (nocc > 0 && orbitals == nullptr) fail(HFJK_ERR_INPUT, "null orbitals pointer");
    if (ctx->scratch == nullptr) fail(HFJK_ERR_BACKEND, "the context holds no scratch address");
    const int kprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
    // The workspace layout of hfjk_exchange_occ, with the fewest SYRK splits
    // (8) when the usual count does not fit the scratch.
    Layout L = make_layout(ctx, std::max(nocc, 1), kprec);
    if (L.x_end > kNoAllocScratchBytes) L = make_layout(ctx, std::max(nocc, 1), kprec, 8);
    if (L.x_end > kNoAllocScratchBytes)
      fail(HFJK_ERR_WORKSPACE, "this exchange needs " + std::to_string(L.x_end) +
                                   " bytes of scratch; the allocation-free entry point holds " +
                                   std::to_string(kNoAllocScratchBytes) + " (use hfjk_exchange_occ with a workspace)");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    g_no_alloc_call = true;
    struct Reset {
      ~Reset() { g_no_alloc_call = false; }
    } reset;
    if (nocc == 0) {
      const long long count = (long long)ctx->nao * ctx->nao;
      hipLaunchKernelGGL(zero_fill_kernel, dim3(hfjk::grid_for(count)), dim3(256), 0, s, exchange, count);
      check_hip(hipGetLastError(), "zero launch");
      return;
    }
    run_exchange(ctx, orbitals, nocc, signs, exchange, precision, ctx->scratch, L, s);
  });
}

int hfjk_jk_occ(hfjk_ctx* ctx, const double* orbitals, int nocc, double* coulomb, double* exchange, int precision,
                void* workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    check_bound(ctx, precision);
    if (nocc < 0) fail(HFJK_ERR_INPUT, "negative nocc");
    if (coulomb == nullptr || exchange == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    if (nocc > 0 && orbitals == nullptr) fail(HFJK_ERR_INPUT, "null orbitals pointer");
    const int kprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
    const Layout L = make_layout(ctx, std::max(nocc, 1), kprec);
    if (workspace_bytes < L.q_end || workspace == nullptr)
      fail(HFJK_ERR_WORKSPACE, "workspace smaller than the combined build needs (" + std::to_string(L.q_end) +
                                   " bytes)");
    if (((uintptr_t)workspace & 255) != 0) fail(HFJK_ERR_WORKSPACE, "workspace must be 256-byte aligned");
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = sizeof(double) * ctx->nao * ctx->nao;
    if (nocc == 0) {
      check_hip(hipMemsetAsync(coulomb, 0, bytes, s), "hipMemsetAsync");
      check_hip(hipMemsetAsync(exchange, 0, bytes, s), "hipMemsetAsync");
      return;
    }
    double* parts = at<double>(workspace, L.qt);
    double* gamma = at<double>(workspace, L.qg);
    double* jp = at<double>(workspace, L.qj);
    const bool serial = ctx->coulomb_serial != 0;
    if (ctx->side == nullptr && !serial) {
      check_hip(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
      check_hip(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming), "hipEventCreateWithFlags");
      check_hip(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming), "hipEventCreateWithFlags");
    }
    hipStream_t side = serial ? s : ctx->side;
    exchange_front(ctx, orbitals, nocc, nullptr, precision, workspace, L, s);
    // The Coulomb half needs only X and the padded orbitals: it runs on the
    // side stream (HBM-bound) while the SYRK (MFMA-bound) runs on s, or with
    // coulomb_serial on s after the SYRK. Its buffers (charges, J partials)
    // lie outside the regions the SYRK uses.
    if (serial) {
      exchange_back(ctx, false, exchange, precision, workspace, L, s);
    } else {
      check_hip(hipEventRecord(ctx->fork, s), "hipEventRecord");
      check_hip(hipStreamWaitEvent(side, ctx->fork, 0), "hipStreamWaitEvent");
    }
    {
      // gamma_P = sum C o X_P, one pass over X.
      Stage st(ctx, side, "coulomb_gamma");
      const void* cpad = at<void>(workspace, L.xc);
      if (precision == HFJK_BF16X3) {
        check_hip(hfjk::launch_orbitals(orbitals, ctx->nao, nocc, HFJK_FP32, at<void>(workspace, L.qc), side),
                  "orbitals launch");
        cpad = at<void>(workspace, L.qc);
      }
      check_hip(hfjk::launch_charges_from_x(at<void>(workspace, L.xx), precision, ctx->naux, ctx->nao, L.ldo, cpad,
                                            parts, gamma, serial ? 0 : side_groups(ctx, precision), side,
                                            x16_interleaved(ctx, precision, false, L.ldo)),
                "charges launch");
    }
    const void* b = bound_tensor(ctx, precision);
    {
      Stage st(ctx, side, "coulomb_accumulate");
      check_hip(hfjk::launch_coulomb_accumulate(b, precision, L.p, ctx->naux, (int)L.ntiles, L.nsplit_j, gamma, jp,
                                                serial ? 0 : side_groups(ctx, precision), side),
                "coulomb accumulate launch");
    }
    {
      Stage st(ctx, side, "coulomb_unpack");
      check_hip(hfjk::launch_coulomb_unpack(jp, L.nsplit_j, L.p, ctx->nao, coulomb, side), "coulomb unpack launch");
    }
    if (!serial) {
      check_hip(hipEventRecord(ctx->join, side), "hipEventRecord");
      exchange_back(ctx, false, exchange, precision, workspace, L, s);
      check_hip(hipStreamWaitEvent(s, ctx->join, 0), "hipStreamWaitEvent");
    }
  });
}

}  // extern "C"

namespace {

// For a symmetric density (max |D - D^T| <= 1e-12 max |D|) the signed factor
// D = C diag(signs) C^T into cf [nao][r] (row-major, the orbital layout of
// the occupied-form exchange) and the signs into sg [r]: by pivoted Cholesky
// (all signs +1) when D is positive semidefinite, else from rocSOLVER dsyevd
// with eigenpairs |l| <= 1e-12 max |l| dropped, C = V |l|^1/2. Returns r, or
// -1 for a non-symmetric density (nothing written). Uses the density region
// of the workspace layout; synchronises s.
int density_factor(hfjk_ctx* ctx, const double* density, void* workspace, const Layout& L, double* cf, double* sg,
                   hipStream_t s, double* asym_out, int* nneg_out) {
  *nneg_out = 0;
  const int n = ctx->nao;
  double* a = at<double>(workspace, L.dens_a);
  double* w = at<double>(workspace, L.dens_w);
  double* e = at<double>(workspace, L.dens_e);
  rocblas_int* info = at<rocblas_int>(workspace, L.dens_info);
  unsigned long long* stats = at<unsigned long long>(workspace, L.dens_stats);
  const long long nn = (long long)n * n;
  unsigned long long host_stats[2];
  {
    Stage st(ctx, s, "density_stats");
    check_hip(hipMemsetAsync(stats, 0, 2 * sizeof(unsigned long long), s), "hipMemsetAsync");
    hipLaunchKernelGGL(density_stats_kernel, dim3(hfjk::grid_for(nn)), dim3(256), 0, s, density, n, stats);
    check_hip(hipGetLastError(), "density stats launch");
  }
  check_hip(hipMemcpyAsync(host_stats, stats, sizeof(host_stats), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  double asym, mag;
  std::memcpy(&asym, &host_stats[0], sizeof(double));
  std::memcpy(&mag, &host_stats[1], sizeof(double));
  if (asym_out != nullptr) *asym_out = asym;
  if (!(asym <= 1e-12 * (mag > 0 ? mag : 1.0))) return -1;
  // A positive semidefinite density (every aufbau or averaged-occupation
  // density) factorises by pivoted Cholesky, D = L L^T, in milliseconds; its
  // residual max |D - L L^T| <= 1e-11 max |D| decides, else dsyevd below.
  hfjk::PcholState* pst = (hfjk::PcholState*)info;  // the 256-byte dens_info region
  static_assert(sizeof(hfjk::PcholState) <= 256, "the Cholesky state fits the density info region");
  hfjk::PcholState host_st{};
  {
    Stage st(ctx, s, "density_cholesky");
    check_hip(hfjk::launch_pchol_init(density, n, 1e-12, e, pst, s), "pivoted Cholesky launch");
    // Columns in batches of 32 steps; the state is read back after each batch.
    constexpr int kBatch = 32;
    for (int k0 = 0; k0 <= n; k0 += kBatch) {
      for (int k = k0; k < std::min(k0 + kBatch, n + 1); ++k)
        check_hip(hfjk::launch_pchol_step(density, n, k, a, e, pst, s), "pivoted Cholesky launch");
      check_hip(hipMemcpyAsync(&host_st, pst, sizeof(host_st), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
      if (host_st.stop) break;
    }
    check_hip(hipMemsetAsync(stats, 0, sizeof(unsigned long long), s), "hipMemsetAsync");
    check_hip(hfjk::launch_factor_residual(density, n, a, pst, stats, s), "factor residual launch");
  }
  const int chol_rank = host_st.stop ? host_st.rank : -1;
  unsigned long long chol_bits = 0;
  check_hip(hipMemcpyAsync(&chol_bits, stats, sizeof(chol_bits), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  double residual;
  std::memcpy(&residual, &chol_bits, sizeof(double));
  if (residual <= 1e-11 * mag && chol_rank >= 0 && chol_rank <= n) {
    if (chol_rank > 0) {
      Stage st(ctx, s, "density_factor");
      check_hip(hfjk::launch_compact_factor(a, n, chol_rank, cf, sg, s), "compact factor launch");
    }
    return chol_rank;
  }
  {
    Stage st(ctx, s, "density_eigh");
    check_hip(hipMemcpyAsync(a, density, sizeof(double) * nn, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    rocblas_handle h = blas_handle(ctx);
    if (rocblas_set_stream(h, s) != rocblas_status_success) fail(HFJK_ERR_SOLVER, "rocblas_set_stream failed");
    // The density is symmetric, so its row-major storage is its
    // column-major storage; eigenvectors come back column-major.
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, e, info) !=
        rocblas_status_success)
      fail(HFJK_ERR_SOLVER, "rocsolver_dsyevd failed to launch");
  }
  std::vector<double> vals(n);
  rocblas_int host_info = 0;
  check_hip(hipMemcpyAsync(vals.data(), w, sizeof(double) * n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipMemcpyAsync(&host_info, info, sizeof(host_info), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (host_info != 0) fail(HFJK_ERR_SOLVER, "rocsolver_dsyevd did not converge (info " + std::to_string(host_info) + ")");
  double top = 0.0;
  for (double v : vals) top = std::max(top, std::fabs(v));
  const double cut = 1e-12 * top;
  int nneg = 0, npos = 0;
  for (int j = 0; j < n && vals[j] < -cut; ++j) ++nneg;
  for (int j = n - 1; j >= nneg && vals[j] > cut; --j) ++npos;
  *nneg_out = nneg;
  if (nneg + npos > 0) {
    Stage st(ctx, s, "density_factor");
    check_hip(hfjk::launch_split_factor(a, w, n, nneg, npos, cf, sg, s), "split factor launch");
  }
  return nneg + npos;
}

}  // namespace

extern "C" {

int hfjk_exchange_density(hfjk_ctx* ctx, const double* density, double* exchange, int precision, void* workspace,
                          size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    if (density == nullptr || exchange == nullptr) fail(HFJK_ERR_INPUT, "null matrix pointer");
    const int n = ctx->nao;
    const int kprec = precision == HFJK_FP64 ? HFJK_FP64 : HFJK_FP32;
    const Layout L = make_layout(ctx, n, kprec);
    check_workspace(L, workspace, workspace_bytes);
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    double* cf = at<double>(workspace, L.dens_c);
    double* sg = at<double>(workspace, L.dens_signs);
    double asym = 0.0;
    int nneg = 0;
    const int r = density_factor(ctx, density, workspace, L, cf, sg, s, &asym, &nneg);
    if (r < 0) {
      // Not symmetric: the Hermitian factorisation of S + i A and two SYRKs.
      exchange_general(ctx, density, exchange, precision, workspace, workspace_bytes, L.dens_end, s);
      return;
    }
    if (r == 0) {
      check_hip(hipMemsetAsync(exchange, 0, sizeof(double) * n * n, s), "hipMemsetAsync");
      return;
    }
    // All signs +1 (a positive semidefinite density): the unsigned SYRK.
    run_exchange(ctx, cf, r, nneg > 0 ? sg : nullptr, exchange, precision, workspace, make_layout(ctx, r, kprec), s);
  });
}

int hfjk_density_factor(hfjk_ctx* ctx, const double* density, double* factor, double* signs, int* rank,
                        int* negative, void* workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    if (density == nullptr || factor == nullptr || signs == nullptr || rank == nullptr || negative == nullptr)
      fail(HFJK_ERR_INPUT, "null pointer");
    const Layout L = make_layout(ctx, 1, HFJK_FP64);
    check_workspace(L, workspace, workspace_bytes);
    begin_call(ctx);
    *rank = density_factor(ctx, density, workspace, L, factor, signs, (hipStream_t)stream, nullptr, negative);
  });
}

int hfjk_general_workspace_bytes(hfjk_ctx* ctx, int precision, int aux_chunk, size_t* bytes) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (bytes == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    const int pc = aux_chunk > 0 ? std::min(aux_chunk, ctx->naux) : ctx->naux;
    *bytes = general_layout(ctx, precision, ctx->nao, pc, 0).total;
  });
}

int hfjk_exchange_general(hfjk_ctx* ctx, const double* density, double* exchange, int precision, void* workspace,
                          size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    check_precision(precision);
    if (ctx->naux <= 0) fail(HFJK_ERR_INPUT, "no fitted tensor is bound (call hfjk_bind)");
    if (density == nullptr || exchange == nullptr) fail(HFJK_ERR_INPUT, "null matrix pointer");
    if (workspace == nullptr || ((uintptr_t)workspace & 255) != 0)
      fail(HFJK_ERR_WORKSPACE, "the workspace is a 256-byte aligned device buffer");
    begin_call(ctx);
    exchange_general(ctx, density, exchange, precision, workspace, workspace_bytes, 0, (hipStream_t)stream);
  });
}

int hfjk_set_tuning(hfjk_ctx* ctx, int param, int value) {
  return guarded([&] {
    check_ctx(ctx);
    if (value < 0) fail(HFJK_ERR_INPUT, "tuning values are non-negative");
    switch (param) {
      case HFJK_TUNE_X_BN:
      case HFJK_TUNE_X16_BN:
        if (value != 0 && value != 32 && value != 64 && value != 128 && value != 256)
          fail(HFJK_ERR_INPUT, "the half-transform column tile is 0, 32, 64, 128 or 256");
        (param == HFJK_TUNE_X_BN ? ctx->x_bn : ctx->x16_bn) = value;
        break;
      case HFJK_TUNE_SYRK_SPLITS:
        if (value > 65536) fail(HFJK_ERR_INPUT, "at most 65536 SYRK splits");
        ctx->syrk_splits = value;
        break;
      case HFJK_TUNE_X_BM:
        if (value != 0 && value != 64 && value != 128)
          fail(HFJK_ERR_INPUT, "the half-transform row tile is 0, 64 or 128");
        ctx->x_bm = value;
        break;
      case HFJK_TUNE_SYRK_KIND:
        if (value > 3) fail(HFJK_ERR_INPUT, "the SYRK kernel kind is 0, 1, 2 or 3");
        ctx->syrk_kind = value;
        break;
      case HFJK_TUNE_X_FULL:
        if (value > 1) fail(HFJK_ERR_INPUT, "HFJK_TUNE_X_FULL is 0 (use a bound full copy) or 1 (ignore it)");
        ctx->x_full = value;
        break;
      case HFJK_TUNE_GENERAL_CHUNK:
        ctx->general_chunk = value;
        break;
      case HFJK_TUNE_SIDE_GROUPS:
        ctx->side_groups = value;
        break;
      case HFJK_TUNE_COULOMB_SERIAL:
        if (value > 1) fail(HFJK_ERR_INPUT, "HFJK_TUNE_COULOMB_SERIAL is 0 (beside the SYRK) or 1 (after it)");
        ctx->coulomb_serial = value;
        break;
      default:
        fail(HFJK_ERR_INPUT, "unknown tuning parameter " + std::to_string(param));
    }
  });
}

int hfjk_get_tuning(hfjk_ctx* ctx, int param, int* value) {
  return guarded([&] {
    check_ctx(ctx);
    if (value == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    switch (param) {
      case HFJK_TUNE_X_BN: *value = ctx->x_bn; break;
      case HFJK_TUNE_X16_BN: *value = ctx->x16_bn; break;
      case HFJK_TUNE_SYRK_SPLITS: *value = ctx->syrk_splits; break;
      case HFJK_TUNE_X_BM: *value = ctx->x_bm; break;
      case HFJK_TUNE_SYRK_KIND: *value = ctx->syrk_kind; break;
      case HFJK_TUNE_X_FULL: *value = ctx->x_full; break;
      case HFJK_TUNE_GENERAL_CHUNK: *value = ctx->general_chunk; break;
      case HFJK_TUNE_SIDE_GROUPS: *value = ctx->side_groups; break;
      case HFJK_TUNE_COULOMB_SERIAL: *value = ctx->coulomb_serial; break;
      default: fail(HFJK_ERR_INPUT, "unknown tuning parameter " + std::to_string(param));
    }
  });
}

int hfjk_profile(hfjk_ctx* ctx, int enable) {
  return guarded([&] {
    check_ctx(ctx);
    ctx->profile = enable != 0;
    ctx->n_stages = 0;
  });
}

int hfjk_profile_read(hfjk_ctx* ctx, const char** names, float* ms, int max) {
  int count = 0;
  const int status = guarded([&] {
    check_ctx(ctx);
    for (size_t i = 0; i < ctx->n_stages; ++i) {
      auto& t = ctx->stages[i];
      check_hip(hipEventSynchronize(t.stop), "hipEventSynchronize");
      if ((int)i < max) {
        float v = 0.f;
        check_hip(hipEventElapsedTime(&v, t.start, t.stop), "hipEventElapsedTime");
        if (names) names[i] = t.name.c_str();
        if (ms) ms[i] = v;
      }
    }
    count = (int)ctx->n_stages;
    ctx->n_stages = 0;
  });
  return status == HFJK_OK ? count : -status;
}

}  // extern "C"

// ------------------------------------------------------------------ k-point DF-J/K
namespace {

struct KpLayout {
  // Coulomb: weighted inner products [nk][naux] and charges [naux] (complex).
  size_t inner, charges;
  // Exchange: the density copy and its eigenpairs, the realified factor and
  // its column signs, the realified block, X, the SYRK partials, the [2n][2n]
  // real result, a complex [n][n] product for the Madelung term.
  size_t dz, w, e, info, stats, ct, sg, bh, x, part, kh, tz;
  size_t total;
  int npad2, ldo, xrows, npairs, nsplit;
};

KpLayout kp_layout(const hfjk_ctx* ctx, int nk, int naux, int nao) {
  KpLayout L{};
  const size_t n = (size_t)nao;
  L.npad2 = hfjk::n_tiles(2 * nao) * hfjk::kTile;
  L.ldo = hfjk::occ_ld(nao);
  L.xrows = hfjk::x_rows(2 * nao);
  L.npairs = hfjk::syrk_pairs(2 * nao);
  L.nsplit = ctx->syrk_splits > 0 ? ((ctx->syrk_splits + 7) / 8) * 8 : syrk_splits(L.npairs, naux);
  const size_t tile2 = (size_t)hfjk::syrk_tile() * hfjk::syrk_tile();
  size_t off = 0;
  L.inner = off; off = align_up(off + (size_t)nk * naux * 16);
  L.charges = off; off = align_up(off + (size_t)naux * 16);
  L.dz = off; off = align_up(off + n * n * 16);
  L.w = off; off = align_up(off + n * 8);
  L.e = off; off = align_up(off + n * 8);
  L.info = off; off = align_up(off + 8);
  L.stats = off; off = align_up(off + 16);
  L.ct = off; off = align_up(off + (size_t)L.ldo * L.npad2 * 8);
  L.sg = off; off = align_up(off + (size_t)L.ldo * 8);
  L.bh = off; off = align_up(off + (size_t)naux * L.npad2 * L.npad2 * 8);
  L.x = off; off = align_up(off + (size_t)naux * L.xrows * L.ldo * 8);
  L.part = off; off = align_up(off + (size_t)L.nsplit * L.npairs * tile2 * 8);
  L.kh = off; off = align_up(off + 4 * n * n * 8);
  L.tz = off; off = align_up(off + n * n * 16);
  L.total = off;
  return L;
}

void kp_check(int nk, int naux, int nao) {
  if (nk < 1 || naux < 1 || nao < 1) fail(HFJK_ERR_INPUT, "nk, naux and nao are positive");
}

}  // namespace

extern "C" {

int hfjk_kpoint_workspace_bytes(hfjk_ctx* ctx, int nk, int naux, int nao, size_t* bytes) {
  return guarded([&] {
    check_ctx(ctx);
    kp_check(nk, naux, nao);
    if (bytes == nullptr) fail(HFJK_ERR_INPUT, "null output pointer");
    *bytes = kp_layout(ctx, nk, naux, nao).total;
  });
}

int hfjk_kpoint_coulomb(hfjk_ctx* ctx, int nk, int naux, int nao, int zero_transfer, const void* tensor,
                        const double* signature, const double* weights, const void* density, void* coulomb,
                        void* workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    kp_check(nk, naux, nao);
    if (tensor == nullptr || signature == nullptr || weights == nullptr || density == nullptr || coulomb == nullptr)
      fail(HFJK_ERR_INPUT, "null pointer");
    const KpLayout L = kp_layout(ctx, nk, naux, nao);
    check_workspace_size(L.total, workspace, workspace_bytes);
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    const long long nn = (long long)nao * nao;
    // Zero-transfer blocks: B[k][k] of [nk][nk][naux] blocks with signs [nk][nk][naux], or the
    // [nk][naux] blocks of a Coulomb-only fitting with one sign per aux function.
    const long long kstride = zero_transfer ? (long long)naux * nn : (long long)(nk + 1) * naux * nn;
    const long long sig_kstride = zero_transfer ? 0 : (long long)(nk + 1) * naux;
    {
      Stage st(ctx, s, "kpoint_charges");
      check_hip(hfjk::launch_kp_charges(tensor, kstride, nk, naux, nao, density, weights, at<void>(workspace, L.inner),
                                        at<void>(workspace, L.charges), s),
                "k-point charges launch");
    }
    {
      Stage st(ctx, s, "kpoint_coulomb");
      check_hip(hfjk::launch_kp_coulomb(tensor, kstride, nk, naux, nao, signature, sig_kstride,
                                        at<void>(workspace, L.charges), coulomb, s),
                "k-point coulomb launch");
    }
  });
}

int hfjk_kpoint_exchange(hfjk_ctx* ctx, int nk, int naux, int nao, const void* tensor, const double* signature,
                         const double* weights, const void* overlap, double madelung, const void* density,
                         void* exchange, void* workspace, size_t workspace_bytes, void* stream) {
  return guarded([&] {
    check_ctx(ctx);
    kp_check(nk, naux, nao);
    if (tensor == nullptr || signature == nullptr || weights == nullptr || density == nullptr || exchange == nullptr)
      fail(HFJK_ERR_INPUT, "null pointer");
    if (madelung != 0.0 && overlap == nullptr) fail(HFJK_ERR_INPUT, "the Madelung term needs the overlaps");
    const KpLayout L = kp_layout(ctx, nk, naux, nao);
    check_workspace_size(L.total, workspace, workspace_bytes);
    begin_call(ctx);
    hipStream_t s = (hipStream_t)stream;
    const int n = nao;
    const long long nn = (long long)n * n;
    std::vector<double> w(nk);
    check_hip(hipMemcpyAsync(w.data(), weights, sizeof(double) * nk, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (int k = 0; k < nk; ++k)
      if (!(w[k] >= 0.0)) fail(HFJK_ERR_INPUT, "k-point weights are non-negative");
    check_hip(hipMemsetAsync(exchange, 0, sizeof(double) * 2 * nk * nn, s), "hipMemsetAsync");
    auto* dz = at<rocblas_double_complex>(workspace, L.dz);
    double* ev = at<double>(workspace, L.w);
    double* e = at<double>(workspace, L.e);
    rocblas_int* info = at<rocblas_int>(workspace, L.info);
    unsigned long long* stats = at<unsigned long long>(workspace, L.stats);
    double* ct = at<double>(workspace, L.ct);
    double* sg = at<double>(workspace, L.sg);
    double* bh = at<double>(workspace, L.bh);
    double* x = at<double>(workspace, L.x);
    void* part = at<void>(workspace, L.part);
    double* kh = at<double>(workspace, L.kh);
    std::vector<double> vals(n);
    for (int k2 = 0; k2 < nk; ++k2) {
      const char* d2 = (const char*)density + 16 * k2 * nn;
      unsigned long long host_stats[2];
      {
        Stage st(ctx, s, "kpoint_density_eigh");
        check_hip(hipMemsetAsync(stats, 0, 2 * sizeof(unsigned long long), s), "hipMemsetAsync");
        check_hip(hfjk::launch_kp_hermitian_stats(d2, n, stats, s), "hermitian stats launch");
        check_hip(hipMemcpyAsync(dz, d2, 16 * nn, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        rocblas_handle h = blas_handle(ctx);
        if (rocblas_set_stream(h, s) != rocblas_status_success) fail(HFJK_ERR_SOLVER, "rocblas_set_stream failed");
        // The column-major view of the row-major Hermitian D(k) is conj(D(k)); its
        // eigenvectors are conjugated when the factor is formed.
        if (rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n, dz, n, ev, e, info) !=
            rocblas_status_success)
          fail(HFJK_ERR_SOLVER, "rocsolver_zheevd failed to launch");
      }
      rocblas_int host_info = 0;
      check_hip(hipMemcpyAsync(host_stats, stats, sizeof(host_stats), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      check_hip(hipMemcpyAsync(vals.data(), ev, sizeof(double) * n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      check_hip(hipMemcpyAsync(&host_info, info, sizeof(host_info), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
      double asym, mag;
      std::memcpy(&asym, &host_stats[0], sizeof(double));
      std::memcpy(&mag, &host_stats[1], sizeof(double));
      if (!(asym <= 1e-12 * (mag > 0 ? mag : 1.0)))
        fail(HFJK_ERR_INPUT, "the density of k-point " + std::to_string(k2) + " is not Hermitian (max |D - D^H| = " +
                                 std::to_string(asym) + ")");
      if (host_info != 0)
        fail(HFJK_ERR_SOLVER, "rocsolver_zheevd did not converge (info " + std::to_string(host_info) + ")");
      double top = 0.0;
      for (double v : vals) top = std::max(top, std::fabs(v));
      const double cut = 1e-12 * top;
      int nneg = 0, npos = 0;
      for (int j = 0; j < n && vals[j] < -cut; ++j) ++nneg;
      for (int j = n - 1; j >= nneg && vals[j] > cut; --j) ++npos;
      const int r = nneg + npos;
      if (r == 0 || w[k2] == 0.0) continue;
      const int ldo = hfjk::occ_ld(r);
      {
        Stage st(ctx, s, "kpoint_factor");
        check_hip(hfjk::launch_kp_factor(dz, ev, n, nneg, npos, w[k2], ldo, ct, sg, s), "k-point factor launch");
      }
      for (int k1 = 0; k1 < nk; ++k1) {
        const char* block = (const char*)tensor + 16 * ((long long)k1 * nk + k2) * naux * nn;
        const double* sig = signature + ((long long)k1 * nk + k2) * naux;
        {
          Stage st(ctx, s, "kpoint_realify");
          check_hip(hfjk::launch_kp_realify(block, naux, n, bh, s), "k-point realify launch");
        }
        {
          Stage st(ctx, s, "kpoint_half_transform");
          check_hip(hfjk::launch_half_transform_full(bh, HFJK_FP64, naux, 2 * n, ct, ldo, x, s),
                    "k-point half transform launch");
        }
        {
          Stage st(ctx, s, "kpoint_syrk");
          check_hip(hfjk::launch_syrk(x, HFJK_FP64, naux, 2 * n, ldo, sg, true, L.nsplit, part, 0, s, sig),
                    "k-point syrk launch");
          check_hip(hfjk::launch_syrk_reduce(part, HFJK_FP64, L.nsplit, 2 * n, kh, s), "k-point reduce launch");
          check_hip(hfjk::launch_kp_fold(kh, n, true, (char*)exchange + 16 * k1 * nn, s), "k-point fold launch");
        }
      }
    }
    if (madelung != 0.0) {
      Stage st(ctx, s, "kpoint_madelung");
      void* tz = at<void>(workspace, L.tz);
      for (int k = 0; k < nk; ++k) {
        const char* sk = (const char*)overlap + 16 * k * nn;
        const char* dk = (const char*)density + 16 * k * nn;
        check_hip(hfjk::launch_kp_zgemm(sk, dk, n, 1.0, false, tz, s), "k-point zgemm launch");
        check_hip(hfjk::launch_kp_zgemm(tz, sk, n, madelung, true, (char*)exchange + 16 * k * nn, s),
                  "k-point zgemm launch");
      }
    }
  });
}

}  // extern "C"
