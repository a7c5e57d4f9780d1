// Internal declarations of the MI355X DF-J/K library (gfx950, CDNA4).
//
// Storage of the fitted tensor B[aux, ao, ao] (symmetric in the two AO
// indices): each B[P] is cut into 32 x 32 tiles, zero-padded past nao; only
// the tiles (r, c) with r >= c are kept, at tile index r (r + 1) / 2 + c,
// each row-major and complete (diagonal tiles hold both triangles).
// One aux function therefore spans packed_elems(nao) = nt (nt + 1) / 2 * 1024
// elements, nt = ceil(nao / 32). The FP64 copy is the master; the FP32 copy
// and the BF16 hi/lo planes (hi = bf16(fp32(b)), lo = bf16(fp32(b) - hi))
// are derived from it for the mixed-precision modes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>
#include <vector>

#include "../../include/hfjk.h"

namespace hfjk {

constexpr int kTile = 32;        // layout tile edge
constexpr int kTileElems = kTile * kTile;
constexpr int kBlock = 64;       // row block of the exchange kernels
constexpr int kThreads = 256;    // threads per workgroup of every kernel

inline int n_tiles(int nao) { return (nao + kTile - 1) / kTile; }
inline long long n_tile_pairs(int nao) {
  long long nt = n_tiles(nao);
  return nt * (nt + 1) / 2;
}
inline long long packed_elems(int nao) { return n_tile_pairs(nao) * kTileElems; }
// Rows of the half-transformed X per aux function: a multiple of kBlock.
inline int x_rows(int nao) { return ((nao + kBlock - 1) / kBlock) * kBlock; }
// Leading dimension of C and X: nocc rounded up to 32.
inline int occ_ld(int nocc) { return ((nocc + 31) / 32) * 32; }

// Blocks of a grid-stride launch over `count` elements. A dispatch holds at
// most 2^32 work-items per dimension, so flat launches are capped and the
// kernels loop.
inline unsigned grid_for(long long count) {
  long long blocks = (count + kThreads - 1) / kThreads;
  if (blocks < 1) blocks = 1;
  if (blocks > (1LL << 20)) blocks = 1LL << 20;
  return (unsigned)blocks;
}

struct Error {
  int code;
  std::string message;
};

struct StageTimer {
  std::string name;
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
};

}  // namespace hfjk

struct hfjk_ctx {
  int device = 0;
  // Bound tensor (borrowed device pointers).
  int naux = 0;
  int nao = 0;
  const double* b64 = nullptr;
  const float* b32 = nullptr;
  const uint16_t* b16 = nullptr;  // hi plane [naux][P], then lo plane [naux][P]
  // Stage timing of the last call when profiling is on.
  bool profile = false;
  std::vector<hfjk::StageTimer> stages;
  size_t n_stages = 0;
  // rocSOLVER / rocBLAS handle for the density factorisation (lazy).
  void* blas = nullptr;
  // Tuning (0 = default heuristics).
  int x_bn = 0;
  int x16_bn = 0;
  int syrk_splits = 0;
  int x_bm = 0;  // 0 = 64
  int syrk_kind = 0;  // 0 = default, 1 = register-staged, 2/3 = direct LDS loads (two / one workgroups per CU)
  // Borrowed full row-major copies of B ([naux][npad][npad], FP64 and FP32) and their use (0 = when
  // bound, 1 = never).
  const double* b64_full = nullptr;
  const float* b32_full = nullptr;
  const uint16_t* b16_full = nullptr;  // hi plane [naux][npad][npad], then the lo plane
  int x_full = 0;
  int general_chunk = 0;  // aux functions per chunk of the non-symmetric exchange (0 = fit)
  int side_groups = 0;    // grid cap of the Coulomb kernels beside the SYRK in hfjk_jk_occ (0 = default)
  int coulomb_serial = 0; // hfjk_jk_occ: 1 = the Coulomb half after the SYRK on the launch stream
  int cus = 256;          // compute units of the device
  // Second stream of hfjk_jk_occ (default priority, created on first use):
  // the Coulomb half runs there beside the SYRK, fenced by fork/join events.
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr;
  hipEvent_t join = nullptr;
  // Device address of the library's fixed scratch array on this device
  // (hfjk_exchange_occ_noalloc), looked up at context creation.
  void* scratch = nullptr;
};

namespace hfjk {
// Persistent work fetch: the workgroup's next item of phase `ph`. Items are
// dealt in eight shares [total x / 8, total (x + 1) / 8), one per XCD (x =
// blockIdx.x % 8, so an instance's items stay in one XCD's L2); once its own
// share is exhausted a workgroup takes items from the other shares in turn,
// so no XCD idles while another still has work. Returns the item and sets
// share_end to the end of the share it came from; both are `total` once
// every share is exhausted. All threads of the workgroup must call it
// together.
__device__ inline long long batch_next(const long long* prefix, unsigned long long* counters, int ninst, int ph,
                                       long long& share_end) {
  __shared__ long long s_item, s_end;
  const long long total = prefix[(long long)ph * (ninst + 1) + ninst];
  if (threadIdx.x == 0) {
    const int x0 = blockIdx.x & 7;
    long long item = total, end = total;
    for (int d = 0; d < 8; ++d) {
      const int x = (x0 + d) & 7;
      const long long lo = total * x / 8, hi = total * (x + 1) / 8;
      if (lo >= hi) continue;
      const long long it = lo + (long long)atomicAdd(&counters[ph * 8 + x], 1ULL);
      if (it < hi) {
        item = it;
        end = hi;
        break;
      }
    }
    s_item = item;
    s_end = end;
  }
  __syncthreads();
  const long long item = s_item;
  share_end = s_end;
  __syncthreads();
  return item;
}

// Instance i with prefix[i] <= item < prefix[i + 1] (prefix non-decreasing).
__device__ inline int batch_find(const long long* prefix, int ninst, long long item) {
  int lo = 0, hi = ninst;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}
}  // namespace hfjk

// Sets the calling thread's hfjk_last_error message (capi.hip).
void hfjk_internal_set_error(const char* message);

namespace hfjk {

// Kernel launchers (gen.hip). Each returns a hipError_t.
hipError_t launch_generate(uint64_t seed, int q0, int naux, int nao, double* out, hipStream_t s);
hipError_t launch_pack(const double* full, int naux, int nao, double* out,
                       unsigned long long* asym_bits, hipStream_t s);
hipError_t launch_split(const double* in, long long count, int precision, void* out, hipStream_t s);
hipError_t launch_normal_matrix(uint64_t seed, int rows, int cols, uint32_t tag, double* out, hipStream_t s);

// Coulomb (jk.hip).
hipError_t launch_weights(const double* density, int nao, double* w, hipStream_t s);
hipError_t launch_gamma(const void* b, int precision, long long p, int naux, int ntiles_total,
                        const double* w, double* partial, hipStream_t s);
hipError_t launch_gamma_reduce(const double* partial, int nchunks, int naux, double* gamma, hipStream_t s);
// max_groups > 0 caps the grid of the two Coulomb kernels below (each
// workgroup then loops over several items): beside the SYRK in
// hfjk_jk_occ, a capped grid keeps CU slots free for the SYRK.
hipError_t launch_coulomb_accumulate(const void* b, int precision, long long p, int naux, int nchunks,
                                     int nsplit, const double* gamma, double* partial, int max_groups, hipStream_t s);
hipError_t launch_coulomb_unpack(const double* partial, int nsplit, long long p, int nao,
                                 double* out, hipStream_t s);

// Exchange (jk.hip).
hipError_t launch_orbitals(const double* c, int nao, int nocc, int precision, void* out, hipStream_t s);
// Pivoted Cholesky of a symmetric PSD density into l (nao * nao doubles:
// L by columns; diag: nao doubles of scratch), one column per step with its
// state on the device: init, then steps k = 0..nao (each a no-op once the
// state says stop; step nao only stops), then max |D - L L^T| over the
// first st->rank columns into *residual (bits of a double, atomic max:
// clear it first).
struct PcholState {
  int p, stop, rank, pad;
  double piv, cut;
};
hipError_t launch_pchol_init(const double* density, int nao, double tol, double* diag, PcholState* st,
                             hipStream_t s);
hipError_t launch_pchol_step(const double* density, int nao, int k, double* l, double* diag, PcholState* st,
                             hipStream_t s);
hipError_t launch_factor_residual(const double* density, int nao, const double* l, const PcholState* st,
                                  unsigned long long* residual, hipStream_t s);
hipError_t launch_compact_factor(const double* l, int nao, int r, double* c, double* signs, hipStream_t s);
// bn = column tile (32, 64, 128, 256) or 0 for the default for ldo; bm = row
// tile (64 or 128; 128 applies to column tiles up to 128).
hipError_t launch_half_transform(const void* b, int precision, long long p, int naux, int nao,
                                 const void* c, int ldo, void* x, int bn, int bm, hipStream_t s);
// kind: 0 = default, 1 = register-staged FP64/FP32 kernel, 2 = direct LDS
// loads (FP64/FP32 with ldo <= 2048; otherwise the register-staged kernel
// runs; two workgroups per CU for unsigned columns), 3 = direct LDS loads with
// one workgroup per CU.
// is_signed = false promises signs[i] = 1 for every occupied column, which
// lets the direct-LDS kernel skip them.
// Partials: FP64 keeps one 128 x 128 tile per split and lower tile pair;
// the FP32-accumulating modes (FP32, BF16x3) one per run of
// syrk_run_aux(precision, ldo) aux functions (at most kSyrkRunProducts
// products per FP32 accumulator, whatever the split count).
// syrk_slots = the number of partial tiles per pair; launch_syrk_reduce sums
// them in FP64 in a fixed order.
constexpr int kSyrkRunProducts = 4096;
int syrk_run_aux(int precision, int ldo);
int syrk_slots(int precision, int naux, int ldo, int nsplit);
// auxsign (optional, [naux]): one more factor per aux function on the signed
// FP64/FP32 direct-LDS kernels; shift (a multiple of 32 below ldo): the
// second operand's columns rotated, sum_k X[m][k] S_k X[n][(k + shift) mod
// ldo], on the signed direct-LDS kernels of every precision
// (hipErrorInvalidValue for a launch that would not run them).
hipError_t launch_syrk(const void* x, int precision, int naux, int nao, int ldo, const double* signs,
                       bool is_signed, int nsplit, void* partial, int kind, hipStream_t s,
                       const double* auxsign = nullptr, int shift = 0, bool x_ilv = false);
// Whether launch_syrk runs the BF16X3 256 x 256-tile SYRK for these
// arguments (only that kernel reads X interleaved: x_ilv).
bool syrk16_tiles256(int nao, int ldo, bool is_signed, int kind, int shift);
// K (+)= the symmetric tiles sym plus the antisymmetric tiles anti: lower
// tiles give K[m][n] = sym + anti, their mirror K[n][m] = sym - anti
// (diagonal tiles computed whole). accumulate adds to K instead of writing.
hipError_t launch_syrk_reduce2(const void* sym, const void* anti, int precision, int nslot, int nao,
                               bool accumulate, double* out, hipStream_t s);
hipError_t launch_syrk_reduce(const void* partial, int precision, int nslot, int nao, double* out,
                              hipStream_t s);
hipError_t launch_occ_signs(const double* signs_in, int nocc, int ldo, double* signs_out, hipStream_t s);

// BF16x3 exchange (xk.hip): operands as bf16 hi/lo planes; launch_syrk
// takes HFJK_BF16X3 for the split-product kernel. launch_orbitals_bf16x3
// writes C^T [ldo][npad] as the hi plane then the lo plane, or interleaved
// (per row npad / 32 blocks of 32 hi then 32 lo, the layout of the bf16
// full copy of B) for the half-transform from that copy.
hipError_t launch_orbitals_bf16x3(const double* c, int nao, int nocc, void* out, hipStream_t s,
                                  bool interleaved = false);
// b_lo: the lo planes of these aux functions (default: naux p elements after b).
hipError_t launch_half_transform_bf16x3(const uint16_t* b, long long p, int naux, int nao, const void* ct, int ldo,
                                        void* x, int bn, hipStream_t s, const uint16_t* b_lo = nullptr);
// Coulomb charges of D = C C^T from the half-transformed X (xk.hip):
// gamma[q] = sum_{m, i} X_q[m][i] cpad[m][i], cpad = the zero-padded
// orbitals [npad][ldo] of launch_orbitals (FP64 when X is FP64, FP32 for
// FP32 and BF16x3 X); parts holds naux * charge_groups(nao) doubles.
int charge_groups(int nao);
hipError_t launch_charges_from_x(const void* x, int precision, int naux, int nao, int ldo, const void* cpad,
                                 double* parts, double* gamma, int max_groups, hipStream_t s, bool x_ilv = false);
// Half-transform from a full row-major copy of B (xk.hip):
// launch_expand_full writes bfull[q][npad][npad] in the precision's type
// (npad = 32 ceil(nao / 32)) from the tiled packed FP64 tensor; for BF16X3
// bfull[q][npad][npad / 32][64]: each 32-column block of a row as the bf16
// hi parts of its elements, then their lo parts (the element split of
// launch_split), so one 128-byte line holds both parts of 32 elements;
// launch_orbitals_t writes C^T [ldo][npad] in FP64 or FP32 (BF16X3 takes
// launch_orbitals_bf16x3's interleaved C^T);
// launch_half_transform_full computes X = B C on the direct-LDS ring (the
// bf16 ring for BF16X3, X as hi/lo planes; BF16X3 on 256 x 256 tiles when
// x16_tiles256(ldo, bn16): ldo > 128 and bn16 not asking for a column tile
// below 256, and then with x_ilv X interleaved as [naux][xrows][ldo / 32][64],
// hi then lo per 32-column block, for the 256 x 256-tile SYRK and the
// charges).
hipError_t launch_expand_full(const double* packed, long long p, int naux, int nao, int precision, void* out,
                              hipStream_t s);
hipError_t launch_orbitals_t(const double* c, int nao, int nocc, int precision, void* ct, hipStream_t s);
hipError_t launch_half_transform_full(const void* bfull, int precision, int naux, int nao, const void* ct, int ldo,
                                      void* x, hipStream_t s, int bn16 = 0, bool x_ilv = false);
bool x16_tiles256(int ldo, int bn16);
// k-point DF-J/K (kpoint.hip): complex values as interleaved double pairs.
// launch_kp_charges: inner [nk][naux] = w_k <B_k[P], D_k> (conjugated B) and
// charges [naux] = sum_k inner, for blocks B_k at b + k kstride (elements);
// launch_kp_coulomb: J_k = sum_P sig_k[P] charges[P] B_k[P] (sig_k at sig + k sig_kstride);
// launch_kp_realify: the [naux][npad2][npad2] real form [[Re, -Im], [Im, Re]] of a
// complex [naux][nao][nao] block (npad2 = 32 ceil(2 nao / 32));
// launch_kp_factor: C^T [ldo][npad2] of [Re C; Im C], C = conj(U) sqrt(weight |l|)
// over the nneg lowest and npos highest eigenpairs (U column-major from
// rocSOLVER zheevd of the column-major view of a row-major Hermitian D), and
// the column signs; launch_kp_fold: K (+)= (Kh_TL + Kh_BR) + i (Kh_BL - Kh_TR)
// of a [2 nao][2 nao] real Kh; launch_kp_zgemm: c (+)= alpha a b ([nao][nao]);
// launch_kp_hermitian_stats: max |D - D^H| and max |D| (atomic max of double
// bits into stats[0..1]; clear first).
hipError_t launch_kp_charges(const void* b, long long kstride, int nk, int naux, int nao, const void* density,
                             const double* weights, void* inner, void* charges, hipStream_t s);
hipError_t launch_kp_coulomb(const void* b, long long kstride, int nk, int naux, int nao, const double* sig,
                             long long sig_kstride, const void* charges, void* out, hipStream_t s);
hipError_t launch_kp_realify(const void* b, int naux, int nao, double* out, hipStream_t s);
hipError_t launch_kp_factor(const void* eigvecs, const double* eigvals, int nao, int nneg, int npos, double weight,
                            int ldo, double* ct, double* signs, hipStream_t s);
hipError_t launch_kp_fold(const double* kh, int nao, bool accumulate, void* k, hipStream_t s);
hipError_t launch_kp_zgemm(const void* a, const void* b, int nao, double alpha, bool accumulate, void* c,
                           hipStream_t s);
hipError_t launch_kp_hermitian_stats(const void* d, int nao, unsigned long long* stats, hipStream_t s);
// The non-symmetric exchange (kpoint.hip): H = S + i A of a real D (complex
// [nao][nao]); the real orbitals [nao][2 hb] = [Re V |l|^1/2, 0, Im V |l|^1/2, 0]
// of retained eigenpairs j0 .. j0 + rb - 1 of H (the nneg lowest, then the
// npos highest; U column-major from zheevd of conj(H)) and the column signs
// of the symmetric and the column-shifted (cross) SYRK.
hipError_t launch_hermitian_of_real(const double* d, int nao, void* h, hipStream_t s);
hipError_t launch_general_factor(const void* eigvecs, const double* eigvals, int nao, int nneg, int npos, int j0,
                                 int rb, int hb, double* c, double* sym, double* cross, hipStream_t s);

// Lower-triangle 128 x 128 output tiles of the SYRK and the tile edge.
int syrk_pairs(int nao);
int syrk_tile();

// ---------------------------------------------------------------- batched
// One instance of a batched build: its shapes and the element offsets of
// its buffers (input tensors, outputs and workspace regions).
struct BatchInst {
  int nao, naux, nocc, ldo;
  int nt, xrows, npairs, nsplit;
  int nbm, nbn;       // half-transform row / column blocks of the instance's tile
  int bm, qg;         // half-transform tile rows (64 or 128; BF16x3: 128); aux functions per item (FP64 / FP32)
  int st, bn;         // SYRK tile edge (64 or 128); half-transform column tile (32, 64, 128 or 256)
  long long p;        // packed elements per aux function
  long long b_off;    // batch tensor (element units of the bound copy)
  long long c_off;    // orbitals [nao][nocc] (float64)
  long long d_off;    // density, J and K [nao][nao] (float64)
  long long w_off;    // fp64 workspace: packed weights [p]
  long long jp_off;   // fp64 workspace: packed J [p]
  long long gp_off;   // fp64 workspace: charge partials [ntiles][naux]
  long long g_off;    // fp64 workspace: charges [naux]
  long long sg_off;   // fp64 workspace: signs [ldo]
  long long cp_off;   // mode workspace: padded orbitals (elements of the mode's type; bf16: per plane)
  long long x_off;    // mode workspace: X (bf16: per plane)
  long long kp_off;   // mode workspace: SYRK partials (FP64 or FP32 elements)
  long long q_off;    // first charge of the instance in a caller's flat charge vector (sum of earlier naux)
};

enum BatchPhase {
  kPhWeights = 0,
  kPhGamma,
  kPhGammaReduce,
  kPhCoulomb,
  kPhCoulombUnpack,
  kPhOrbitals,
  kPhXform,
  kPhSyrk,
  kPhSyrkReduce,
  kPhCharges,  // charges from the half-transformed X (D = C C^T)
  kPhSyrkRing,  // SYRK items of the instances with 128-row tiles (FP64 / FP32 direct-LDS ring)
  kPhCount
};

// Device view of a batch plan: instances, per-phase item prefix sums
// ([kPhCount][ninst + 1]) and per-phase, per-XCD work counters ([kPhCount][8]).
struct BatchView {
  const BatchInst* insts;
  const long long* prefix;
  unsigned long long* counters;
  int ninst;
};

// Launchers of the batched persistent kernels (batch.hip, xk.hip). `grid`
// is a multiple of 8 workgroups; mode buffers are typed by `precision`.
hipError_t launch_batch_reset(BatchView v, hipStream_t s);
hipError_t launch_batch_coulomb_occ(BatchView v, int precision, const void* b, long long b_plane, const double* c,
                                    const void* xbase, long long x_plane, double* ws64, double* j, int grid,
                                    hipStream_t s);
hipError_t launch_batch_coulomb_charges(BatchView v, int precision, const void* b, long long b_plane,
                                        const double* charges, double* ws64, double* j, int grid, hipStream_t s);
hipError_t launch_batch_coulomb(BatchView v, int precision, const void* b, long long b_plane, const double* d,
                                double* ws64, double* j, int grid, hipStream_t s);
hipError_t launch_batch_orbitals(BatchView v, int precision, const double* c, double* ws64, void* cbase,
                                 long long c_plane, int grid, hipStream_t s);
hipError_t launch_batch_xform(BatchView v, int precision, int bn, const void* b, long long b_plane, const void* cbase,
                              long long c_plane, void* xbase, long long x_plane, int grid, hipStream_t s);
hipError_t launch_batch_syrk(BatchView v, int precision, const void* xbase, long long x_plane, const double* ws64,
                             void* kbase, int grid, hipStream_t s);
hipError_t launch_batch_syrk_reduce(BatchView v, int precision, const void* kbase, double* k, int grid,
                                    hipStream_t s);
// Rows of an instance's half-transform tile: FP64 / FP32 128 when the instance has at least 128
// X rows and a column tile of 64, else 64; BF16x3 128.
int batch_tile_rows(int precision, int xrows, int bn);

// Density factorisation helpers (jk.hip).
hipError_t launch_split_factor(const double* vecs, const double* vals, int nao, int first, int count,
                               double* c, double* signs, hipStream_t s);

}  // namespace hfjk
