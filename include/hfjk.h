/*
 * hfjk.h — C ABI of the MI355X (gfx950) density-fitted Coulomb/exchange build.
 *
 * Replaces the J/K contractions of the reference's CoulombExchange protocol
 * (hafnium/scf/coulomb_exchange.py:53-71) as implemented by
 * DensityFittedCoulombExchange (hafnium/scf/coulomb_exchange.py:124-190):
 *   coulomb(D)  J = sum_P B[P] <B[P], D>            (:176-182)
 *   exchange(D) K = sum_P B[P] D B[P]               (:184-190)
 * and the occupied-orbital factorisation the reference lists as a later
 * optimisation behind the same interface (:129-133):
 *   K = sum_P (B[P] C)(B[P] C)^T                     for D = C C^T.
 *
 * Conventions
 *  - Every function returns an int status: HFJK_OK (0) or an HFJK_ERR_*
 *    code; nothing throws across the ABI (the reference's rule for its own
 *    native layer: crates/hafnium-core/src/lib.rs:22). hfjk_last_error()
 *    returns the message of the calling thread's last failure.
 *  - All array pointers are DEVICE pointers, borrowed for the duration of the
 *    call (or, for hfjk_bind, until the next bind / destroy); the caller owns
 *    every buffer. Matrices are row-major float64 unless stated otherwise.
 *  - `stream` is a hipStream_t (NULL = the default stream). Calls enqueue
 *    work and return without synchronising unless stated otherwise.
 *  - Inputs are never modified; outputs are fully overwritten.
 *
 * Storage of B ("tiled packed"): each B[P] is cut into 32 x 32 tiles,
 * zero-padded past nao; the tiles (r, c) with r >= c are stored at tile index
 * r (r + 1) / 2 + c, each row-major and complete. hfjk_packed_elems(nao)
 * elements per aux function. B must be symmetric in its two AO indices, as
 * every fitted tensor is ((P|uv) = (P|vu), CONVENTIONS.md:208-212 of the
 * reference); hfjk_pack measures the asymmetry of a full tensor.
 */
#ifndef HFJK_H
#define HFJK_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define HFJK_API __attribute__((visibility("default")))
#else
#define HFJK_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define HFJK_OK 0
#define HFJK_ERR_INPUT 1     /* shape/argument error (the reference's InputError) */
#define HFJK_ERR_BACKEND 2   /* unsupported on this device or not bound (BackendError) */
#define HFJK_ERR_HIP 3       /* HIP runtime or launch failure */
#define HFJK_ERR_WORKSPACE 4 /* workspace smaller than hfjk_workspace_bytes */
#define HFJK_ERR_SOLVER 5    /* rocSOLVER failure in the density factorisation */

/* Arithmetic of the contractions. FP64 everywhere; FP32 products on MFMA
 * with FP32 accumulation inside a tile and FP64 across tiles; BF16X3 = each
 * FP32 operand split into bf16 hi + lo, hi*hi + hi*lo + lo*hi on MFMA with
 * FP32 accumulation, FP64 across tiles. */
#define HFJK_FP64 0
#define HFJK_FP32 1
#define HFJK_BF16X3 2

typedef struct hfjk_ctx hfjk_ctx;

/* Library version (major*10000 + minor*100 + patch). */
HFJK_API int hfjk_version(void);

/* Message of the calling thread's last failing call ("" if none). */
HFJK_API const char *hfjk_last_error(void);

/* One context per device. No counterpart in the reference (its builder is a
 * frozen dataclass holding B: coulomb_exchange.py:124-145). */
HFJK_API int hfjk_create(hfjk_ctx **ctx, int device);
HFJK_API int hfjk_destroy(hfjk_ctx *ctx);

/* Elements per aux function in the tiled packed storage. */
HFJK_API long long hfjk_packed_elems(int nao);

/* Full [naux][nao][nao] float64 -> tiled packed [naux][packed_elems].
 * stats (device, two float64, may be NULL) receives max |B[P,u,v] -
 * B[P,v,u]| and max |B[P,u,v]| over the tensor. Stands in for the construction-time checks of
 * DensityFittedCoulombExchange.__post_init__ (coulomb_exchange.py:142-145). */
HFJK_API int hfjk_pack(hfjk_ctx *ctx, const double *full, int naux, int nao, double *packed,
              double *stats, void *stream);

/* Seeded synthetic fitted tensor: aux functions q0 .. q0+naux-1 of
 * B[q, m, n] = z(m, n, q) exp(-|m - n| / 20), z ~ N(0, 1) from Philox4x32-10
 * keyed by `seed` (the generator of hafnium_amd/synthetic.py, bit for bit),
 * written in tiled packed storage. No counterpart in the reference. */
HFJK_API int hfjk_generate(hfjk_ctx *ctx, uint64_t seed, int q0, int naux, int nao, double *packed,
                  void *stream);

/* Seeded N(0, 1) matrix [rows][cols] with Philox counter (r, c, 0xFFFFFFFF,
 * tag); the Gaussian matrix of the synthetic occupied orbitals. */
HFJK_API int hfjk_normal_matrix(hfjk_ctx *ctx, uint64_t seed, int rows, int cols, uint32_t tag,
                       double *out, void *stream);

/* Derived copies of a packed FP64 tensor of `count` elements:
 * precision HFJK_FP32 -> float[count]; HFJK_BF16X3 -> uint16 bf16 hi
 * plane [count] followed by lo plane [count]. */
HFJK_API int hfjk_split(hfjk_ctx *ctx, const double *packed, long long count, int precision, void *out,
               void *stream);

/* Borrow the tiled packed tensor(s) for later calls. Any of b64, b32, b16
 * may be NULL; a later call in a precision whose copy is NULL fails with
 * HFJK_ERR_BACKEND. Mirrors DensityFittedCoulombExchange(df_tensor)
 * (coulomb_exchange.py:124-145). */
HFJK_API int hfjk_bind(hfjk_ctx *ctx, int naux, int nao, const double *b64, const float *b32,
              const uint16_t *b16);

/* Optional full row-major copies of the bound tensor for the
 * half-transforms: hfjk_expand_full writes full[q][npad][npad] in precision
 * HFJK_FP64 (double), HFJK_FP32 (float) or HFJK_BF16X3 (uint16
 * [q][npad][npad / 32][64]: each 32-column block of a row as the bf16 hi
 * parts of its elements, then their lo parts; 2 naux npad^2 values)
 * (npad = 32 ceil(nao / 32), both triangles, zero past nao) from the tiled
 * packed FP64 tensor; hfjk_bind_full borrows it for that precision (NULL
 * unbinds; hfjk_bind unbinds all three). With it the half-transform of that
 * precision streams both operands straight into LDS (twice the memory of the
 * packed copy for 13 % (FP64), 7 % (FP32) and about 25 % (BF16X3, 256 x 256
 * tiles past 128 occupied columns) less time at nao 1500, naux 5000). No
 * counterpart in the reference (its B is always full, [naux, nao, nao]:
 * coulomb_exchange.py:124-145). */
HFJK_API int hfjk_expand_full(hfjk_ctx *ctx, const double *packed, int naux, int nao, int precision, void *full,
                              void *stream);
HFJK_API int hfjk_bind_full(hfjk_ctx *ctx, int precision, const void *full);

/* Workspace bytes that hfjk_coulomb / hfjk_exchange_occ /
 * hfjk_exchange_density need for the bound tensor, `nocc` occupied columns
 * (use nao for hfjk_exchange_density) and `precision`. */
HFJK_API int hfjk_workspace_bytes(hfjk_ctx *ctx, int nocc, int precision, size_t *bytes);

/* J = sum_P B[P] <B[P], D>; density [nao][nao], coulomb [nao][nao].
 * Replaces DensityFittedCoulombExchange.coulomb (coulomb_exchange.py:176-182). */
HFJK_API int hfjk_coulomb(hfjk_ctx *ctx, const double *density, double *coulomb, int precision,
                 void *workspace, size_t workspace_bytes, void *stream);

/* J = sum_P charges[P] B[P] for given charges (device float64 [naux]): the
 * Coulomb matrix of a density known only by its charges against B, e.g. a
 * predicted fitted density d (charges = L^-1 (P|Q) d, (P|Q) = L L^T). Same
 * workspace as hfjk_coulomb. No counterpart in the reference. */
HFJK_API int hfjk_coulomb_charges(hfjk_ctx *ctx, const double *charges, double *coulomb, int precision,
                                  void *workspace, size_t workspace_bytes, void *stream);

/* K = sum_P (B[P] C) S (B[P] C)^T with C = orbitals [nao][nocc] and
 * S = diag(signs) (signs: device float64 [nocc] of +1/-1, or NULL for all
 * +1). The occupied-orbital form of DensityFittedCoulombExchange.exchange
 * (coulomb_exchange.py:129-133, :184-190). */
HFJK_API int hfjk_exchange_occ(hfjk_ctx *ctx, const double *orbitals, int nocc, const double *signs,
                      double *exchange, int precision, void *workspace, size_t workspace_bytes,
                      void *stream);

/* hfjk_exchange_occ without a workspace argument and without any device
 * allocation. No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Its kernels use a fixed 256 MiB scratch array that
 * the library holds on each GPU until the process ends (the layout of
 * hfjk_exchange_occ, with 8 SYRK splits when the usual count does not fit);
 * a larger problem fails with HFJK_ERR_WORKSPACE. Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * threads, other contexts) race. Nothing in the scratch must start at zero. */
HFJK_API int hfjk_exchange_occ_noalloc(hfjk_ctx *ctx, const double *orbitals, int nocc, const double *signs,
                                       double *exchange, int precision, void *stream);

/* J and K of the density D = C C^T of C = orbitals [nao][nocc] in one call:
 * coulomb_exchange_matrices (coulomb_exchange.py:397-404) for an aufbau
 * density, i.e. hfjk_coulomb(D) and hfjk_exchange_occ(C, NULL). The charges
 * gamma_P = <B[P], D> = sum C o (B[P] C) are formed from the half-transformed
 * X the exchange leaves in the workspace, one pass over X instead of one over
 * B. Workspace: hfjk_workspace_bytes(nocc). */
HFJK_API int hfjk_jk_occ(hfjk_ctx *ctx, const double *orbitals, int nocc, double *coulomb, double *exchange,
                         int precision, void *workspace, size_t workspace_bytes, void *stream);

/* K = sum_P B[P] D B[P] for any real density [nao][nao]. Drop-in for
 * DensityFittedCoulombExchange.exchange (coulomb_exchange.py:184-190).
 * Symmetric to 1e-12 relative: D = C S C^T by pivoted Cholesky when D is
 * positive semidefinite (S = 1), otherwise D = V diag(l) V^T by rocSOLVER
 * dsyevd, eigenpairs with |l| <= 1e-12 max|l| dropped, C = V |l|^1/2,
 * S = sign(l); then hfjk_exchange_occ in `precision`.
 * Otherwise (non-symmetric) hfjk_exchange_general's factorisation of
 * S + i A and its two SYRKs, in `precision`, in as many aux passes as the
 * workspace beyond the density region needs.
 * A complex density is two calls, on its real and imaginary parts.
 * Synchronises `stream` once or twice (the symmetry check, the retained
 * rank). */
HFJK_API int hfjk_exchange_density(hfjk_ctx *ctx, const double *density, double *exchange, int precision,
                          void *workspace, size_t workspace_bytes, void *stream);

/* The two halves of hfjk_exchange_density with workspaces sized for what
 * they do (hfjk_exchange_density sizes its workspace for full rank, nocc =
 * nao, the bound for any density). hfjk_density_factor: for a symmetric
 * density (1e-12 relative) the signed factor D = C diag(signs) C^T into
 * factor [nao][rank] (row-major; a buffer of nao * nao doubles) and
 * signs [rank], *rank = the retained rank (0 for D = 0), *negative = the
 * number of -1 signs (the first columns); *rank = -1 for a non-symmetric
 * density, nothing written. A positive semidefinite density is factorised
 * by pivoted Cholesky (accepted when max |D - C C^T| <= 1e-11 max |D|, all
 * signs +1), any other symmetric one by dsyevd (eigenvalues above 1e-12 of
 * the largest kept). Synchronises `stream`. K then is
 * hfjk_exchange_occ(factor, rank, negative ? signs : NULL) with the
 * workspace of hfjk_workspace_bytes(ctx, rank, precision); it takes a
 * workspace of at least hfjk_workspace_bytes(ctx, 1, HFJK_FP64) bytes.
 * hfjk_exchange_general: K of any real density, in `precision`, on the
 * library's kernels: with S, A the symmetric and antisymmetric parts,
 * S + i A = V diag(l) V^H (rocSOLVER zheevd; |l| <= 1e-12 max |l| dropped)
 * and X' = B [Re V|l|^1/2, Im V|l|^1/2]: B S B is the signed SYRK of X',
 * B A B its SYRK with the second operand's two column halves swapped and
 * signed. The workspace holds as many aux functions per pass as fit;
 * hfjk_general_workspace_bytes gives the bytes for passes of `aux_chunk`
 * aux functions (<= 0: all of them at once). Synchronises `stream` once.
 * Both replace the same reference call, DensityFittedCoulombExchange.exchange
 * (coulomb_exchange.py:184-190). */
HFJK_API int hfjk_density_factor(hfjk_ctx *ctx, const double *density, double *factor, double *signs, int *rank,
                                 int *negative, void *workspace, size_t workspace_bytes, void *stream);
HFJK_API int hfjk_general_workspace_bytes(hfjk_ctx *ctx, int precision, int aux_chunk, size_t *bytes);
HFJK_API int hfjk_exchange_general(hfjk_ctx *ctx, const double *density, double *exchange, int precision,
                                   void *workspace, size_t workspace_bytes, void *stream);

/* ---- Batched J/K: many instances of uneven size, one persistent launch per phase ----
 * No counterpart in the reference (it builds one molecule's J/K at a time:
 * coulomb_exchange.py:176-190 per call). For instance i:
 *   J_i = sum_P B_i[P] <B_i[P], D_i>,   K_i = sum_P (B_i[P] C_i)(B_i[P] C_i)^T.
 * B_i is tiled packed at element offset b_offset (a multiple of 1024) of one
 * batch tensor (FP64, FP32, or BF16 hi plane followed by the lo plane, each
 * plane hfjk_batch_sizes()->tensor_elems long); C_i [nao][nocc] at c_offset of
 * the orbital buffer; D_i, J_i, K_i [nao][nao] at d_offset of their buffers. */
typedef struct {
  int nao;
  int naux;
  int nocc;
  long long b_offset;
  long long c_offset;
  long long d_offset;
} hfjk_batch_instance;
typedef struct hfjk_batch hfjk_batch;
/* Plan a batch (host only): shapes, SYRK splits, offsets. tile = half-transform
 * column tile (0 = default; 32, 64, 128, 256), chosen by the autotuner. */
HFJK_API int hfjk_batch_create(hfjk_ctx *ctx, const hfjk_batch_instance *instances, int ninst,
                               int precision, int tile, hfjk_batch **batch);
HFJK_API int hfjk_batch_workspace_bytes(const hfjk_batch *batch, size_t *bytes);
/* Elements of one tensor plane, of the orbital buffer and of each matrix buffer. */
HFJK_API int hfjk_batch_sizes(const hfjk_batch *batch, long long *tensor_elems,
                              long long *orbital_elems, long long *matrix_elems);
/* Upload the plan into the workspace (synchronises `stream` once). */
HFJK_API int hfjk_batch_bind(hfjk_batch *batch, void *workspace, size_t bytes, void *stream);
/* Enqueue one batched J/K build; no host synchronisation and no allocation,
 * so a run can be captured in a hipGraph. Runs on one bound workspace must not
 * overlap. `density` may be NULL: then every instance's density is
 * D = C C^T of its orbitals and the charges gamma_P = <B_P, D> come from the
 * half-transformed X_P = B_P C (sum of C o X_P, one pass over X instead of
 * one over B), as hfjk_jk_occ does for one instance. */
HFJK_API int hfjk_batch_run(hfjk_batch *batch, const void *tensor, const double *orbitals,
                            const double *density, double *coulomb, double *exchange, void *stream);
/* J_i = sum_P charges_i[P] B_i[P] for given charges (device float64; instance
 * i's naux_i values start at the sum of the earlier instances' naux): the
 * Coulomb matrix of a density known by its charges against B, e.g. a
 * predicted fitted density (charges = L^-1 (P|Q) d with (P|Q) = L L^T). */
HFJK_API int hfjk_batch_coulomb_charges(hfjk_batch *batch, const void *tensor, const double *charges,
                                        double *coulomb, void *stream);
HFJK_API int hfjk_batch_destroy(hfjk_batch *batch);

/* ---- Gaussian integrals (host) ---------------------------------------------
 * The SCF inputs that the reference's Rust engine produces
 * (crates/hafnium-core/src/integrals/one_electron.rs, two_electron.rs:74-310
 * via hafnium/integrals/density_fitting.py:58-85): overlap, kinetic energy,
 * nuclear attraction, (P|Q) and (P|uv), over contracted real solid-harmonic
 * Gaussians (m = -l..l, p = x, y, z; functions contraction-major then m;
 * coefficients of normalised primitives; unit-norm functions). Host arrays,
 * row-major; l <= 5. */
typedef struct {
  int nshell;
  const int *l;          /* [nshell] */
  const int *nprim;      /* [nshell] */
  const int *ncontr;     /* [nshell] */
  const double *centers; /* [nshell][3], Bohr */
  const double *exps;    /* primitives of all shells, concatenated */
  const double *coefs;   /* per shell [ncontr][nprim], concatenated */
} hfjk_basis;
HFJK_API int hfjk_basis_nfunc(const hfjk_basis *basis, int *nfunc);
/* S, T, V [n][n] (any output may be NULL); V over point charges at coords [natom][3]. */
HFJK_API int hfjk_int1e(const hfjk_basis *basis, int natom, const double *charges, const double *coords,
                        double *overlap, double *kinetic, double *nuclear);
/* (P|Q) [naux][naux]. */
HFJK_API int hfjk_int2c2e(const hfjk_basis *aux, double *metric);
/* (P|uv) [naux][n][n]. */
HFJK_API int hfjk_int3c2e(const hfjk_basis *aux, const hfjk_basis *orbital, double *out);
/* The same (P|Q) and (P|uv) computed on the GPU of `ctx` into device buffers
 * (row-major, [naux][naux] and [naux][n][n]); the basis structs are host
 * memory. The parallel three-centre assembly of the reference's engine
 * (crates/hafnium-core/src/integrals/two_electron.rs:212-310, :944; used by
 * hafnium/integrals/density_fitting.py:58-85) as one 64-lane workgroup per
 * shell triple. l <= 5, total angular momentum of a triple <= 12.
 * Synchronises `stream` (the shell tables are freed on return). */
HFJK_API int hfjk_int2c2e_device(hfjk_ctx *ctx, const hfjk_basis *aux, double *metric, void *stream);
HFJK_API int hfjk_int3c2e_device(hfjk_ctx *ctx, const hfjk_basis *aux, const hfjk_basis *orbital, double *out,
                                 void *stream);
/* (uv|wx) [n][n][n][n], chemists' notation: the four-centre integrals of the
 * reference's two_electron.eri / IntegralEngine.coulomb_four_center
 * (hafnium/integrals/two_electron.py:26-41, engine.py:280) that
 * ExactCoulombExchange holds (scf/coulomb_exchange.py:74-121), e.g. for the
 * spherical atoms of the density guess (scf/atoms.py:55-73). n^4 doubles. */
HFJK_API int hfjk_int2e(const hfjk_basis *basis, double *out);

/* ---- FP64 dense steps of the SCF around the J/K build ------------------
 * Replace the NumPy/array-API calls of the reference's SCF driver
 * (hafnium/scf/driver.py:248-329, hamiltonian.py:331-370, diis.py:70-106,
 * orthogonalization.py:16-42, density_fitting.py:97-107). All row-major
 * FP64 device buffers; results are deterministic run to run. */

/* C = alpha op(A) op(B) + beta C (rocBLAS dgemm); op = transpose if trans_*. */
HFJK_API int hfjk_dgemm(hfjk_ctx *ctx, int trans_a, int trans_b, int m, int n, int k, double alpha,
                        const double *a, int lda, const double *b, int ldb, double beta, double *c,
                        int ldc, void *stream);
/* Symmetric eigendecomposition (rocSOLVER dsyevd) of a [n][n]: eigenvalues
 * ascending in w[n]; on exit row j of a is the j-th eigenvector. work[n] and
 * info[1] are device scratch (info != 0: not converged). */
HFJK_API int hfjk_dsyevd(hfjk_ctx *ctx, int n, double *a, double *w, double *work, int *info,
                         void *stream);
/* Cholesky a = L L^T (rocSOLVER dpotrf): the lower triangle of a receives L;
 * the strict upper triangle keeps its input. info[1] device (> 0: not PD). */
HFJK_API int hfjk_dpotrf(hfjk_ctx *ctx, int n, double *a, int *info, void *stream);
/* b [m][n] <- L^-1 b with L the lower triangle of l [m][m] (rocBLAS dtrsm). */
HFJK_API int hfjk_dtrsm_lower(hfjk_ctx *ctx, int m, int n, const double *l, double *b, void *stream);
/* out = a x + b y + c z elementwise over n (y, z may be NULL). */
HFJK_API int hfjk_combine(hfjk_ctx *ctx, long long n, double a, const double *x, double b,
                          const double *y, double c, const double *z, double *out, void *stream);
/* out = m - m^T for m [n][n] (out distinct from m). */
HFJK_API int hfjk_antisymmetrize(hfjk_ctx *ctx, int n, const double *m, double *out, void *stream);
/* Fixed-order dot product: *out (device) = sum x[i] y[i]; partial is device
 * scratch of hfjk_dot_workspace() doubles. */
HFJK_API long long hfjk_dot_workspace(void);
HFJK_API int hfjk_ddot(hfjk_ctx *ctx, long long n, const double *x, const double *y, double *partial,
                       double *out, void *stream);

/* Tuning parameters of the kernels (0 = the library's default choice).
 * HFJK_TUNE_X_BN: column tile of the FP64/FP32 half-transform (32, 64, 128, 256).
 * HFJK_TUNE_X16_BN: column tile of the BF16x3 half-transform (32, 64, 128, 256;
 * from a bound full copy of B: 256 or 0 = 256 x 256 tiles past 128 occupied
 * columns, else the 128 x 128 ring).
 * HFJK_TUNE_SYRK_SPLITS: aux-slice splits of the SYRK (rounded up to a multiple of 8).
 * HFJK_TUNE_X_BM: row tile of the FP64/FP32 half-transform (64, 128; 128 is
 * used with column tiles up to 128).
 * HFJK_TUNE_SYRK_KIND: SYRK kernel: 1 = operands staged through registers
 * into one LDS buffer (every precision), 2 = operands loaded straight into a
 * ring of LDS stages with 128 x 128 tiles (occupied widths up to 2048;
 * FP64/FP32: two workgroups per CU unless the columns carry signs; BF16X3:
 * one), 3 = the FP64/FP32 ring with one workgroup per CU; 0 = the library's
 * choice (the ring where it applies; BF16X3 without signs or column shift:
 * its 256 x 256-tile variant).
 * HFJK_TUNE_X_FULL: 0 = the half-transform uses the full copy of B
 * of its precision when one is bound (hfjk_bind_full), 1 = it ignores it.
 * HFJK_TUNE_GENERAL_CHUNK: most aux functions per chunk of the non-symmetric
 * exchange of hfjk_exchange_density (0 = as many as the workspace holds).
 * HFJK_TUNE_SIDE_GROUPS: most workgroups of each Coulomb kernel that
 * hfjk_jk_occ runs beside the SYRK (0 = the library's choice: two per
 * compute unit in FP64, no cap in FP32 and BF16X3, whose shorter SYRK would
 * not hide the capped kernels).
 * HFJK_TUNE_COULOMB_SERIAL: hfjk_jk_occ runs its Coulomb half 0 = on a
 * second stream beside the SYRK, 1 = on the launch stream after the SYRK.
 * Set by the project's autotuner (hafnium_amd/autotune.py); the workspace size
 * depends on HFJK_TUNE_SYRK_SPLITS, so query it after tuning. No counterpart
 * in the reference. */
#define HFJK_TUNE_X_BN 1
#define HFJK_TUNE_X16_BN 2
#define HFJK_TUNE_SYRK_SPLITS 3
#define HFJK_TUNE_X_BM 4
#define HFJK_TUNE_SYRK_KIND 5
#define HFJK_TUNE_X_FULL 6
#define HFJK_TUNE_GENERAL_CHUNK 7
#define HFJK_TUNE_SIDE_GROUPS 8
#define HFJK_TUNE_COULOMB_SERIAL 9
HFJK_API int hfjk_set_tuning(hfjk_ctx *ctx, int param, int value);
HFJK_API int hfjk_get_tuning(hfjk_ctx *ctx, int param, int *value);

/* Stage timing: when enabled, every later call records a HIP event pair on
 * its stream around each kernel stage it launches (at most 4096 stages are
 * kept). hfjk_profile_read synchronises on the recorded events, returns up
 * to `max` (name, milliseconds) pairs in launch order and forgets them; the
 * return value is the number of stages recorded (or a negative error). The
 * names stay valid until the next call on the context. */
HFJK_API int hfjk_profile(hfjk_ctx *ctx, int enable);
HFJK_API int hfjk_profile_read(hfjk_ctx *ctx, const char **names, float *ms, int max);

/* ---- Batched FP64 dense steps (a batched SCF over many molecules) --------
 * No counterpart in the reference (its SCF driver solves one molecule at a
 * time: hafnium/scf/driver.py:154-213). Instances share one padded size n
 * (zero-padded), laid out at a fixed stride; all pointers are device
 * pointers, row-major. */

/* For i < count: C_i = alpha op(A_i) op(B_i) + beta C_i, X_i = x + i * stride_x
 * (rocBLAS dgemm_strided_batched). */
HFJK_API int hfjk_dgemm_batched(hfjk_ctx *ctx, int trans_a, int trans_b, int m, int n, int k, double alpha,
                                const double *a, int lda, long long stride_a, const double *b, int ldb,
                                long long stride_b, double beta, double *c, int ldc, long long stride_c, int count,
                                void *stream);
/* Doubles of device scratch hfjk_dsyev_batched needs. */
HFJK_API long long hfjk_dsyev_batched_work(int n, int count);
/* Eigendecomposition of count symmetric [n][n] matrices at stride n*n:
 * eigenvalues ascending in w[count][n]; row j of each matrix becomes its j-th
 * eigenvector. method 0: divide and conquer (rocSOLVER dsyevd), 1: Jacobi
 * (dsyevj). info[count] device (!= 0: that instance did not converge). */
HFJK_API int hfjk_dsyev_batched(hfjk_ctx *ctx, int n, int count, double *a, double *w, double *work, int *info,
                                int method, void *stream);
/* The lowest eigenpairs of count symmetric [n][n] matrices at stride n*n
 * (row-major, n <= 512), one workgroup per matrix: Householder
 * tridiagonalisation, multisection on the Sturm count for the nvec lowest
 * eigenvalues (w[i][0 .. nvec), ascending; w is [count][n], its other
 * entries untouched), inverse iteration and orthonormalisation for their
 * eigenvectors, which become rows 0 .. nvec - 1 of each matrix (its other
 * rows are left as scratch). One launch, no allocation: work is device
 * scratch of hfjk_eigh_lowest_batched_work(n, count, nvec) doubles; info[i]
 * is set to 0. Replaces the rocSOLVER call of hfjk_dsyev_batched where only
 * the occupied orbitals are needed (the batched SCF); no counterpart in the
 * reference, whose diagonalisation is numpy.linalg.eigh
 * (hafnium/scf/driver.py). */
HFJK_API long long hfjk_eigh_lowest_batched_work(int n, int count, int nvec);
HFJK_API int hfjk_eigh_lowest_batched(hfjk_ctx *ctx, int n, int count, int nvec, double *a, double *w, double *work,
                                      int *info, void *stream);
/* out[i] = <x_i, y_i> over len elements, fixed summation order. */
HFJK_API int hfjk_ddot_batched(hfjk_ctx *ctx, int count, long long len, const double *x, long long stride_x,
                               const double *y, long long stride_y, double *out, void *stream);
/* DIIS overlaps of the newest stored error with every slot: hist is
 * [nslots][count][len]; out[i][a] = <hist[newest][i], hist[a][i]> (fixed
 * summation order; nslots <= 64). */
HFJK_API int hfjk_history_dots_batched(hfjk_ctx *ctx, int count, long long len, int nslots, const double *hist,
                                       int newest, double *out, void *stream);
/* DIIS coefficients of count instances, on the device (replaces the host
 * solve_diis of hafnium/scf/diis.py:82-106 in the batched loop): ov is
 * [count][nslots][nslots] (2 <e_a, e_b>, kept by the caller across calls);
 * row and column `newest` are first set from dots ([count][nslots], the
 * output of hfjk_history_dots_batched, doubled). Over the host list
 * valid[0..nvalid) of slots (distinct, in the caller's order), the bordered
 * system [[B / s, -1], [-1^T, 0]] (s the largest diagonal of B) is solved
 * through its eigenpairs with |w| <= 1e-14 max |w| dropped; coef
 * ([count][nslots]) gets the solution on the valid slots, zero elsewhere.
 * nslots <= 16. */
HFJK_API int hfjk_diis_batched(hfjk_ctx *ctx, int count, int nslots, double *ov, const double *dots, int newest,
                               const int *valid, int nvalid, double *coef, void *stream);
/* The DIIS extrapolation of count instances (diis.py:55-79):
 * out[i] = sum_a coef[i][a] hist[a][i] over len elements, hist
 * [nslots][count][len], coef [count][nslots] (as hfjk_diis_batched writes
 * them; zero-coefficient slots are not read), out [count][len]. */
HFJK_API int hfjk_diis_combine_batched(hfjk_ctx *ctx, int count, long long len, int nslots, const double *hist,
                                       const double *coef, double *out, void *stream);
/* out_i = m_i - m_i^T for count [n][n] matrices at stride n*n. */
HFJK_API int hfjk_antisymmetrize_batched(hfjk_ctx *ctx, int n, int count, const double *m, double *out,
                                         void *stream);
/* a_i <- (a_i + a_i^T) / 2 in place for count [n][n] matrices at stride n*n
 * (removes the rounding asymmetry of a fitted tensor L^-1 (P|uv), whose
 * packed storage keeps one triangle). */
HFJK_API int hfjk_symmetrize_batched(hfjk_ctx *ctx, int n, long long count, double *a, void *stream);
/* Canonical orthogonaliser from eigenpairs of the overlaps (rows of vecs):
 * x_i[u][j] = vecs_i[j][u] / sqrt(vals_i[j]) where vals_i[j] > threshold, else 0;
 * kept[i][j] = 1 / 0 accordingly (orthogonalization.py:16-42 of the reference). */
HFJK_API int hfjk_orthogonalizer_batched(hfjk_ctx *ctx, int n, int count, const double *vecs, const double *vals,
                                         double threshold, double *x, int *kept, void *stream);
/* a_i[j][j] += value for every j with kept[i][j] == 0 (moves the discarded
 * and padded directions above every occupied level). */
HFJK_API int hfjk_shift_discarded_batched(hfjk_ctx *ctx, int n, int count, const int *kept, double value, double *a,
                                          void *stream);
/* a_i[r][c] = 0 for c >= keep[i] (count [rows][cols] matrices at stride rows*cols). */
HFJK_API int hfjk_mask_columns_batched(hfjk_ctx *ctx, int rows, int cols, int count, const int *keep, double *a,
                                       void *stream);
/* out_i = flags[i] ? a_i : b_i for count blocks of len doubles (flags device int). */
HFJK_API int hfjk_select_batched(hfjk_ctx *ctx, int count, long long len, const int *flags, const double *a,
                                 const double *b, double *out, void *stream);
/* Copies between padded (r, c) at i * stride + r * ld + c and compact
 * offsets[i] + r * cols[i] + c, for r < rows[i], c < cols[i] (device int /
 * long long arrays; cols[i] <= ld, rows[i] * ld <= stride, rows[i] * cols[i]
 * <= max_elems). to_compact = 1: padded -> compact, 0: compact -> padded. */
HFJK_API int hfjk_copy_batched(hfjk_ctx *ctx, int count, const int *rows, const int *cols, const long long *offsets,
                               long long max_elems, double *padded, int ld, long long stride, double *compact,
                               int to_compact, void *stream);

/* ---- Learned-guess MLP kernels (no counterpart in the reference: its
 * guesses are the core Hamiltonian and the superposition of atomic
 * densities, hafnium/scf/driver.py:216-246 and hafnium/scf/atoms.py:76-119).
 * Arithmetic: GEMM operands are rounded to bf16 (nearest even) on load,
 * products accumulate in FP32; everything else is FP32. Device pointers,
 * row-major storage. */

/* c[m][n] = op(a) op(b) (+ c if accumulate) (+ bias[n]); h_silu (optional,
 * same leading dimension as c) receives silu(c). op(a) is m x k: a is stored
 * m x k (trans_a = 0) or k x m (trans_a = 1) with leading dimension lda, as
 * fp32 or (a_bf16 = 1) bf16; storage row r of a is read from row a_rows[r]
 * when a_rows is given (every index must address a stored row). op(b) is
 * k x n: b stored k x n (trans_b = 0) or n x k (trans_b = 1). */
HFJK_API int hfjk_mlp_gemm(hfjk_ctx *ctx, int m, int n, int k, const void *a, int a_bf16, int lda, int trans_a,
                           const int *a_rows, const void *b, int b_bf16, int ldb, int trans_b, const float *bias,
                           float *c, int ldc, int accumulate, float *h_silu, void *stream);
/* grad_z = grad_h * silu'(z), elementwise over count. */
HFJK_API int hfjk_mlp_silu_backward(hfjk_ctx *ctx, long long count, const float *grad_h, const float *z,
                                    float *grad_z, void *stream);
/* out[j] = sum_i x[i][j] over an m x n matrix (fixed summation order). */
HFJK_API int hfjk_mlp_column_sum(hfjk_ctx *ctx, int m, int n, const float *x, int ldx, float *out, void *stream);
/* loss[0] = scale sum_ij w_j (y_ij - t_ij)^2 and (grad optional, leading
 * dimension ldy) grad_ij = 2 scale w_j (y_ij - t_ij); target row i is
 * target_rows[i] when given, col_weights NULL means w = 1. */
HFJK_API int hfjk_mlp_mse(hfjk_ctx *ctx, int m, int n, const float *y, int ldy, const float *target, int ldt,
                          const int *target_rows, const float *col_weights, float scale, float *grad, float *loss,
                          void *stream);
/* Adam with decoupled weight decay over count parameters:
 * m1 = b1 m1 + (1 - b1) g, m2 = b2 m2 + (1 - b2) g^2,
 * p -= lr_t m1 / (sqrt(m2) + eps) + lr weight_decay p, and params_bf16
 * (optional) = bf16(p). lr_t carries the bias corrections of the step. The
 * hyperparameters are double so that 1 - beta is formed exactly; the update
 * itself runs in FP32. */
HFJK_API int hfjk_mlp_adam(hfjk_ctx *ctx, long long count, float *params, const float *grads, float *moment1,
                           float *moment2, uint16_t *params_bf16, double lr, double lr_t, double beta1,
                           double beta2, double eps, double weight_decay, void *stream);
/* dst = bf16(src) (nearest even) over count elements. */
HFJK_API int hfjk_mlp_to_bf16(hfjk_ctx *ctx, long long count, const float *src, uint16_t *dst, void *stream);

/* ---- k-point DF-J/K (FP64; complex arrays as interleaved (re, im) double pairs, row-major) ----
 * Replace KPointDensityFittedCoulombExchange.coulomb / .exchange
 * (coulomb_exchange.py:279-297) and KPointDensityFittedCoulomb.coulomb
 * (coulomb_exchange.py:375-383). tensor: B [nk][nk][naux][nao][nao] of
 * momentum transfer k2 - k1 with signature [nk][nk][naux] (+1/-1), or with
 * zero_transfer = 1 the zero-transfer blocks [nk][naux][nao][nao] with one
 * sign per aux function [naux]; weights [nk] (device float64, >= 0); density,
 * overlap and the outputs [nk][nao][nao].
 * hfjk_kpoint_coulomb: c_P = sum_k w_k sum_xw conj(B[k,k,P,x,w]) D(k)[x,w],
 *   J(k) = sum_P s[k,k,P] c_P B[k,k,P].
 * hfjk_kpoint_exchange: K(k) = sum_k' w_k' sum_P s[k,k',P] B[k,k',P] D(k')
 *   B[k,k',P]^H + madelung S(k) D(k) S(k) for Hermitian D(k') (1e-12
 *   relative, else HFJK_ERR_INPUT): D(k') = V diag(l) V^H by rocSOLVER zheevd
 *   (|l| <= 1e-12 max |l| dropped), and per (k, k') block the real form
 *   [[Re B, -Im B], [Im B, Re B]] against [Re C; Im C], C = V |l|^1/2
 *   sqrt(w_k'), through the FP64 direct-LDS half-transform and the signed
 *   SYRK with per-aux signs. Synchronises `stream` once per k-point.
 * Workspace: hfjk_kpoint_workspace_bytes(nk, naux, nao) for both. */
HFJK_API int hfjk_kpoint_workspace_bytes(hfjk_ctx *ctx, int nk, int naux, int nao, size_t *bytes);
HFJK_API int hfjk_kpoint_coulomb(hfjk_ctx *ctx, int nk, int naux, int nao, int zero_transfer, const void *tensor,
                                 const double *signature, const double *weights, const void *density,
                                 void *coulomb, void *workspace, size_t workspace_bytes, void *stream);
HFJK_API int hfjk_kpoint_exchange(hfjk_ctx *ctx, int nk, int naux, int nao, const void *tensor,
                                  const double *signature, const double *weights, const void *overlap,
                                  double madelung, const void *density, void *exchange, void *workspace,
                                  size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HFJK_H */
