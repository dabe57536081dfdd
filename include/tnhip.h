/*
 * tnhip.h — C ABI of libtnhip.so, the MI355X (gfx950) MPS contraction library.
 *
 * Conventions
 *   - Every entry point returns int: 0 = TNH_OK, < 0 = error code. The message of
 *     the last error on a context is returned by tnh_last_error(ctx). Nothing aborts
 *     and no C++ exception crosses this boundary.
 *   - Tensors live on the device inside tnh_buf handles owned by a context. Their
 *     dims are column-major (first index fastest), exactly as Julia passes arrays.
 *   - Complex tensors are stored planar on the device: all real parts, then all
 *     imaginary parts. Host transfers accept interleaved complex (Julia ComplexF64).
 *   - Host pointers are borrowed for the duration of a call only.
 *   - One context = one HIP stream. Calls on one context are serialised by the
 *     caller; separate contexts may be used from separate threads.
 *   - MPS site tensors are (chi_left, d, chi_right); MPO site tensors are
 *     (D_left, d_out(bra), d_in(ket), D_right); environment blocks are
 *     (chi_bra, D, chi_ket).
 *
 * Reference citations are to lcauser/TensorNetworks.jl (file:line).
 */
#ifndef TNHIP_H
#define TNHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tnh_ctx tnh_ctx;
typedef struct tnh_buf tnh_buf;
typedef struct tnh_dmrg tnh_dmrg;
typedef struct tnh_tebd tnh_tebd;
typedef struct tnh_qjmc tnh_qjmc;

/* Element types. Device storage of C128/C64 is planar. */
enum {
  TNH_F64 = 0,
  TNH_F32 = 1,
  TNH_C128 = 2,
  TNH_C64 = 3
};

/* Arithmetic used by the matrix-core contractions. Accumulation is always at
 * least FP32; TNH_PREC_F64 runs FP64 MFMA end to end. */
enum {
  TNH_PREC_F64 = 0,
  TNH_PREC_F32 = 1,
  TNH_PREC_BF16X3 = 2
};

enum {
  TNH_OK = 0,
  TNH_ERR_ARG = -1,
  TNH_ERR_HIP = -2,
  TNH_ERR_SHAPE = -3,
  TNH_ERR_NOMEM = -4,
  TNH_ERR_NUMERIC = -5,
  TNH_ERR_NOCONV = -6,
  TNH_ERR_UNSUPPORTED = -7
};

/* ------------------------------------------------------------------ context */

/* No reference counterpart (the reference is one CPU process): creates a HIP
 * stream and workspace on `device`. */
int tnh_ctx_create(int device, tnh_ctx **out);
int tnh_ctx_destroy(tnh_ctx *ctx);
/* Message of the last failing call on ctx ("" if none). Valid until the next call. */
const char *tnh_last_error(tnh_ctx *ctx);
int tnh_ctx_synchronize(tnh_ctx *ctx);
/* Library version string, e.g. "tnhip 0.1 gfx950". */
const char *tnh_version(void);

/* Kernel timing (no counterpart in the reference, which has no profiling): while
 * enabled, HIP events on the context's stream bracket every launch of the kernel
 * families "gemm_f64", "gemm_f32", "gemm_bf16x3" (matrix-core GEMMs) and "mix" (MPO /
 * gate mixing stage), and the intervals "jacobi" (one block-Jacobi eigensolve of a split,
 * its "flops" field counting outer sweeps) and "dmrg_env", "dmrg_lanczos", "dmrg_split"
 * (the phases of each bond step of tnh_dmrg_sweep). tnh_profile_read synchronises and
 * returns the count, the summed time and the summed algorithmic work (flops = 2 *
 * multiply-adds; bytes = operands read once + results written once). */
int tnh_profile_enable(tnh_ctx *ctx, int on);
int tnh_profile_reset(tnh_ctx *ctx);
int tnh_profile_read(tnh_ctx *ctx, const char *family, int64_t *launches, double *total_ms,
                     double *flops, double *bytes);

/* GEMM tile selection (no counterpart in the reference, whose contractions call BLAS
 * through TensorOperations). mode 1 (the default): the first product of each size class
 * (arithmetic mode, operand types, transposes, vector paths, and M, N, K rounded to four
 * classes per octave) with at least 2^28 flops, both sides >= 96 and no batching times
 * every candidate tile (one untimed and two timed launches each, on a scratch output; the
 * stream is synchronised once); the fastest is kept for the class if it beats the size
 * rule's tile by at least 3%, else the rule's tile (near-ties stay stable); smaller or
 * batched products, and products issued while the stream is being captured, take the
 * size rule. mode 0: always the size rule. mode 2: forget the tuned classes. Tiles
 * differ only in the output block a workgroup computes, never in the summation order,
 * so results do not depend on the choice. */
int tnh_gemm_autotune(tnh_ctx *ctx, int mode);
/* The tuned classes as text, one line each: mode, transposes, the timed product's
 * M N K, the chosen and the rule's tile, and the milliseconds of every candidate.
 * Writes at most len bytes (NUL-terminated); *needed receives the full size. */
int tnh_gemm_tuning_report(tnh_ctx *ctx, char *buf, int64_t len, int64_t *needed);

/* ------------------------------------------------------------------ buffers */

/* Allocate an uninitialised device tensor. ndim <= 8. */
int tnh_buf_create(tnh_ctx *ctx, int dtype, int ndim, const int64_t *dims, tnh_buf **out);
/* Allocate a device tensor of `dtype` and fill it from host memory of
 * `host_dtype` (converting precision; interleaved -> planar for complex; a real
 * device dtype from a complex host array keeps the real part and fails with
 * TNH_ERR_NUMERIC if any imaginary part is non-zero). */
int tnh_buf_upload(tnh_ctx *ctx, const void *host, int host_dtype, int dtype, int ndim,
                   const int64_t *dims, tnh_buf **out);
/* Copy a device tensor to host memory of `host_dtype` (converting). */
int tnh_buf_download(tnh_ctx *ctx, const tnh_buf *buf, void *host, int host_dtype);
/* Wrap caller-owned device memory (not freed by the library). */
int tnh_buf_wrap(tnh_ctx *ctx, void *device_ptr, int dtype, int ndim, const int64_t *dims,
                 tnh_buf **out);
int tnh_buf_free(tnh_ctx *ctx, tnh_buf *buf);
/* dims must hold 8 entries. */
int tnh_buf_shape(const tnh_buf *buf, int *dtype, int *ndim, int64_t *dims);
void *tnh_buf_device_ptr(const tnh_buf *buf);
/* Device-to-device copy with dtype conversion into a new buffer. */
int tnh_buf_convert(tnh_ctx *ctx, const tnh_buf *src, int dtype, tnh_buf **out);
/* Frobenius norm of a device tensor: norm(psi) at the centre, src/structures/mps/gmps.jl:29-37. */
int tnh_buf_norm(tnh_ctx *ctx, const tnh_buf *buf, double *norm);
/* In-place scaling by (re + i im) (promotes a real tensor to complex when im != 0):
 * psi[center] *= x, gmps.jl:46-51 and abstractmps.jl:100-108. */
int tnh_buf_scale(tnh_ctx *ctx, tnh_buf *buf, double re, double im);
/* A[a,o,b] <- sum_i G[o,i] A[a,i,b] for a (chi_l, d, chi_r) site tensor and a (d, d)
 * operator; *A is replaced (its dtype may become complex). Replaces applyop!,
 * src/structures/mps/mps.jl:141-151, and one-site gates of applygate,
 * src/structures/mps/gatelist.jl:137-157 (rng = 1). */
int tnh_apply_site_op(tnh_ctx *ctx, tnh_buf **A, const tnh_buf *G);

/* ------------------------------------------------------- hot-path operators */

/* Pairwise contraction primitive: column-major matrix product on device buffers,
 *   C = alpha * op(A) * op(B) + beta * C,   op in {0: N, 1: T, 2: conjugate transpose},
 * with A, B, C viewed as matrices of the given leading dimensions (elements).
 * Replaces the permute-plus-GEMM lowering of contract(x, y, idxs1, idxs2, conjx, conjy),
 * src/tensors.jl:13-18, for contractions whose index groups are already contiguous.
 * Real or planar-complex buffers (C complex when A or B is); alpha, beta real. */
int tnh_gemm(tnh_ctx *ctx, int opA, int opB, int64_t M, int64_t N, int64_t K, double alpha,
             const tnh_buf *A, int64_t lda, const tnh_buf *B, int64_t ldb, double beta,
             tnh_buf *C, int64_t ldc, int precision);

/* H_eff matvec of two-site DMRG:
 *   out[a,o1,o2,b'] = coeff * sum L[a,w,a'] W1[w,o1,i1,w1] W2[w1,o2,i2,w2] R[b',w2,b] x[a',i1,i2,b]
 * Replaces product(::ProjMPS, A, direction, nsites=2), src/structures/mps/projmps.jl:103-144
 * (rank-2, non-squared branch :107-134, coefficient :144). `out` must have x's shape
 * and dtype. Evaluated as the GEMM chain L*x -> (W1 W2) -> *R^T in the given precision. */
int tnh_heff_apply(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                   const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double coeff_re,
                   double coeff_im, int precision);

/* The same matvec as tnh_heff_apply (same kernels, bitwise the same result) without any
 * device allocation: no counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Every
 * temporary comes from a static 64 MiB device array per GPU, held until the process ends;
 * shapes whose scratch bound exceeds it fail with TNH_ERR_NOMEM before any work is queued
 * (FP64 real operands: chi up to ~450 at D = 5, d = 2). All calls on one GPU share that
 * array, so calls that overlap on one GPU (other contexts, other threads) race. */
int tnh_heff_apply_noalloc(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                           const tnh_buf *R, const tnh_buf *x, tnh_buf *out, double coeff_re,
                           double coeff_im, int precision);

/* Left environment update L'[b,w',b'] = sum L[a,w,a'] conj(A[a,s,b]) W[w,s,t,w'] A[a',t,b'].
 * Replaces buildleft!(::ProjMPS, idx), src/structures/mps/projmps.jl:50-66.
 * *Lout is (re)allocated when NULL or of the wrong shape. Pass Lprev = NULL for the
 * edge block ones(1,1,1) (src/structures/mps/abstractprojmps.jl:33-36). */
int tnh_env_left(tnh_ctx *ctx, const tnh_buf *Lprev, const tnh_buf *A, const tnh_buf *W,
                 tnh_buf **Lout, int precision);
/* Right environment update R'[a,w,a'] = sum A[a',t,b'] R[b,w',b'] W[w,s,t,w'] conj(A[a,s,b]).
 * Replaces buildright!(::ProjMPS, idx), src/structures/mps/projmps.jl:74-95. */
int tnh_env_right(tnh_ctx *ctx, const tnh_buf *Rprev, const tnh_buf *A, const tnh_buf *W,
                  tnh_buf **Rout, int precision);
/* Fully contracted <A|L W R|A> at one site. Replaces calculate(::ProjMPS),
 * src/structures/mps/projmps.jl:192-216. NULL L or R means the edge block. */
int tnh_env_calculate(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *A, const tnh_buf *W,
                      const tnh_buf *R, double coeff, double *re, double *im);

/* Lowest eigenpair of H_eff by restarted Lanczos (Krylov-Schur, full
 * re-orthogonalisation, FP64 vectors). Replaces the KrylovKit call
 * eigsolve(Heff, A0, 1, :SR; maxiter, krylovdim, ishermitian=true, tol) at
 * src/algorithms/mps/dmrg.jl:50-53. `theta` holds A0 on entry and the normalised
 * Ritz vector on exit (dtype F64 or C128). The matvecs use `precision`. */
int tnh_eigsolve_ground(tnh_ctx *ctx, const tnh_buf *L, const tnh_buf *W1, const tnh_buf *W2,
                        const tnh_buf *R, tnh_buf *theta, int krylovdim, int maxiter,
                        double tol, double coeff, int precision, double *energy,
                        int *nmatvec);

/* Truncated SVD of a 2-D device matrix M (rows x cols) with the truncation rule of
 * svd(x, idx; cutoff, maxdim, mindim), src/tensors.jl:168-227 (:202-215):
 * maxdim = 0 means unlimited; k = max(min(maxdim, k_cut), min(mindim, n)) >= 1 where
 * k_cut is the smallest k whose discarded weight sum_{i>k} s_i^2 / sum s_i^2 <= cutoff.
 * Primary path: FP64 Gram matrix + Jacobi eigensolver; fallback (on non-convergence
 * or non-finite values, mirroring gesdd -> gesvd at src/tensors.jl:190-195):
 * one-sided Jacobi on M. Outputs U (rows x k), S (k, F64), Vt (k x cols), each
 * (re)allocated. `discarded` receives sum_{i>k} s_i^2 / sum s_i^2, `fallback` 1 if
 * the fallback ran. Any of U/S/Vt may be NULL when not wanted. force_fallback != 0
 * runs the fallback directly. */
int tnh_svd_trunc(tnh_ctx *ctx, const tnh_buf *M, double cutoff, int maxdim, int mindim,
                  int force_fallback, tnh_buf **U, tnh_buf **S, tnh_buf **Vt, int *k,
                  double *discarded, int *fallback);

/* Split a two-site tensor theta (chi_l, d, d, chi_r) into two site tensors.
 * Replaces replacesites!(psi, A, site, direction, normalize; cutoff, maxdim, mindim),
 * src/structures/mps/gmps.jl:199-267, for nsites = 2, rank 1:
 *   direction = 0 (sweeping right): A1 (chi_l,d,k) left-orthonormal, A2 = S*V (k,d,chi_r);
 *   direction = 1 (sweeping left):  A1 = U*S (chi_l,d,k), A2 (k,d,chi_r) right-orthonormal.
 * normalize != 0 rescales the centre tensor to unit norm (gmps.jl:263-265).
 * If schmidt != NULL, up to schmidt_cap kept singular values are written there. */
int tnh_replacesites2(tnh_ctx *ctx, const tnh_buf *theta, int direction, int normalize,
                      double cutoff, int maxdim, int mindim, tnh_buf **A1, tnh_buf **A2,
                      int *k, double *discarded, double *schmidt, int schmidt_cap);

/* Move the orthogonality centre of a rank-1 MPS of N device site tensors (replaced in
 * place, bond dimensions may change) from `center` (0 = not set) to `new_center`
 * (1-based), by truncated SVDs. Replaces movecenter!(psi, idx; cutoff, maxdim, mindim),
 * src/structures/mps/gmps.jl:90-112, with moveleft!/moveright! at :60-82. */
int tnh_mps_movecenter(tnh_ctx *ctx, int N, tnh_buf **sites, int center, int new_center,
                       double cutoff, int maxdim, int mindim);

/* theta (chi_l, d, d, chi_r) = A1 (chi_l,d,chi) * A2 (chi,d,chi_r): contract(A1, A2, 3, 1),
 * src/algorithms/mps/dmrg.jl:44-47 and src/structures/mps/gatelist.jl:142-145. */
int tnh_two_site_theta(tnh_ctx *ctx, const tnh_buf *A1, const tnh_buf *A2, tnh_buf **theta);

/* Two-site gate application followed by the truncated split (no normalisation).
 * Replaces applygate(psi, site, gate, direction; cutoff, maxdim, mindim),
 * src/structures/mps/gatelist.jl:137-171 (rng = 2, rank 1):
 *   theta'[a,o1,o2,b] = sum G[o1,i1,o2,i2] theta[a,i1,i2,b]   (G is (d,d,d,d)).
 * A1 and A2 are replaced in place by the split factors. */
int tnh_applygate2(tnh_ctx *ctx, tnh_buf **A1, tnh_buf **A2, const tnh_buf *G, int direction,
                   double cutoff, int maxdim, int mindim, int *k, double *discarded);
/* applygate for a gate on nsites >= 2 sites (gatelist.jl:137-171 with rng = nsites; the
 * gates trotterize makes for Hamiltonians whose terms span more than two sites): A[0..nsites-1]
 * are the window's FP64/C128 site tensors, replaced in place; G has 2 nsites indices
 * G[o1, i1, ..., on, in] (nsites <= 4: tensors have at most 8 indices; d^nsites <= 64). theta = A[0] ... A[n-1], theta' = G theta, then the
 * r-site branch of replacesites! without normalisation (gmps.jl:215-257): nsites - 1 truncated
 * splits, left to right for direction 0 (the centre ends on the last site), right to left
 * for direction 1 (the centre ends on the first). k (may be NULL) receives the nsites - 1
 * kept bonds, *discarded the largest discarded weight. */
int tnh_applygate_sites(tnh_ctx *ctx, int nsites, tnh_buf **A, const tnh_buf *G, int direction,
                        double cutoff, int maxdim, int mindim, int *k, double *discarded);
/* replacesites!(psi, theta, site, direction, normalize; cutoff, maxdim, mindim) for
 * nsites in [2, 6] (gmps.jl:199-267, the r-site branch :215-257): theta is the FP64/C128
 * (chi_l, d, ..., d, chi_r) tensor of nsites sites; A[0..nsites-1] receive the new site
 * tensors (an entry may be NULL or a buffer to reuse; the results are library-owned). The
 * splits and the outputs k, discarded are those of tnh_applygate_sites; normalize divides the
 * new centre (the last site for direction 0, the first for direction 1) by its norm. */
int tnh_replacesites_sites(tnh_ctx *ctx, int nsites, const tnh_buf *theta, int direction,
                           int normalize, double cutoff, int maxdim, int mindim, tnh_buf **A,
                           int *k, double *discarded);

/* Local expectation values <psi| O_t |psi> (unnormalised) of nterms operator-list terms
 * over the N device site tensors of psi. Replaces inner(st, psi, oplist, psi),
 * src/structures/mps/mps.jl:87-134 (coefficients are applied by the caller). Term t acts
 * on sites first_site[t] .. first_site[t] + range[t] - 1 (1-based); `ops` holds, term
 * after term, range[t] d x d operator matrices O[s_out, s_in], column-major, interleaved
 * complex (identity where a term has no operator). out receives 2 * nterms doubles
 * (re, im). */
int tnh_inner_oplist(tnh_ctx *ctx, int N, const tnh_buf *const *psi, int nterms,
                     const int *first_site, const int *range, const double *ops, double *out);
/* The same between two states: <psi| O_t |phi> (unnormalised, coefficients applied by the
 * caller) for the N FP64/C128 site tensors of psi and of phi (same N and physical dim;
 * bonds may differ). Replaces inner(st, psi, oplist, phi), mps.jl:87-134, with the
 * overlap blocks of ProjMPS(psi, phi; rank=1) (:90) from both ends and each term's operators
 * applied to phi's sites; terms and `ops` as in tnh_inner_oplist. */
int tnh_inner_oplist_pair(tnh_ctx *ctx, int N, const tnh_buf *const *psi,
                          const tnh_buf *const *phi, int nterms, const int *first_site,
                          const int *range, const double *ops, double *out);
/* inner(phi, O, psi) and inner(phi, psi) (dot), src/structures/mps/mpo.jl:181-217: the N
 * FP64/C128 site tensors of the bra phi and of the ket psi (bonds may differ) and, unless W is
 * NULL, the N tensors (D_l, d, d, D_r) of one MPO between them. *re, *im receive
 * <phi|W|psi> (<phi|psi> without W), unnormalised. Several MPOs are folded into one first
 * (tnh_mpo_mpo_product, exact with cutoff 0); an MPO left of the bra enters as its adjoint,
 * as the reference rearranges its arguments (:190-200). */
int tnh_inner_mpo(tnh_ctx *ctx, int N, const tnh_buf *const *phi, const tnh_buf *const *W,
                  const tnh_buf *const *psi, double *re, double *im);

/* <psi|H|psi> for an MPO H = W[0..N-1] (inner(psi, H, psi), src/structures/mps/mpo.jl:181-217,
 * no normalisation) into (energy_re, energy_im), and, when dW is not NULL, its gradient
 * with respect to every MPO tensor: dW[i] (shape (D_l, d, d, D_r), allocated or reused like
 * tnh_env_left's output) holds dE/dW_i[w, s, t, w'] = sum L[a,w,a'] conj(A[a,s,b]) A[a',t,b']
 * R[b,w',b'], so that E = sum W_i . dW_i for every i and couplings entering W linearly get
 * their gradient by the chain rule. The reference has no differentiation; this is the
 * backward pass of the environment sweep of buildleft!/buildright! (projmps.jl:50-95).
 * The forward sweep keeps left environments only every `block` sites and the backward
 * sweep recomputes them one block at a time. FP64. */
int tnh_mpo_expect_grad(tnh_ctx *ctx, int N, tnh_buf *const *psi, tnh_buf *const *W, int block,
                        double *energy_re, double *energy_im, tnh_buf **dW);

/* ------------------------------------------------------------------- MPOs */

/* MPO of an operator list: replaces MPO(st, H; cutoff, maxdim, mindim),
 * src/structures/mps/mpo.jl:323-459. The finite-state automaton (channel 1 = nothing placed,
 * the last = term complete, one channel per bond that a multi-site term crosses, identities
 * in a term's gaps) is built on the host, then the SVD compression sweeps of :442-457 run on
 * the device (left to right W_s <- U S, W_{s+1} <- V W_{s+1}; right to left W_s <- S V,
 * W_{s-1} <- W_{s-1} U; each split truncated by (cutoff, maxdim, mindim) on one-sided-Jacobi
 * singular values). cutoff < 0 skips the compression. Term t acts on sites
 * first_site[t] .. first_site[t] + range[t] - 1 (1-based) with range[t] d x d operator
 * matrices in `ops` (column-major O[s_out, s_in], interleaved complex, identity where the
 * term has no operator: the layout of tnh_inner_oplist) and coefficient
 * coeffs[2t] + i coeffs[2t + 1]. W receives N new (D_l, d, d, D_r) tensors, F64 when every
 * entry is real, else C128. */
int tnh_mpo_from_oplist(tnh_ctx *ctx, int N, int d, int nterms, const int *first_site, const int *range,
                        const double *ops, const double *coeffs, double cutoff, int maxdim, int mindim,
                        tnh_buf **W);
/* The compression sweeps of MPO(st, H), src/structures/mps/mpo.jl:442-457, on N device MPO
 * tensors (handles replaced in place, bond dimensions may shrink). */
int tnh_mpo_compress(tnh_ctx *ctx, int N, tnh_buf **W, double cutoff, int maxdim, int mindim);
/* phi = O psi: replaces MPOMPSProduct(O, psi; cutoff, maxdim, mindim),
 * src/structures/mps/mpo.jl:125-147 (applyMPO(O, psi), :105-123):
 *   phi[i][(w,a), s, (w',b)] = sum_t O[i][w,s,t,w'] psi[i][a,t,b]   (w fastest in each bond),
 * then the centre moves from site 1 to N without truncation and back to site 1 with
 * (cutoff, maxdim, mindim). phi receives N new device tensors; FP64 / C128 operands. */
int tnh_mpo_apply(tnh_ctx *ctx, int N, const tnh_buf *const *O, const tnh_buf *const *psi, double cutoff,
                  int maxdim, int mindim, tnh_buf **phi);
/* O = O1 O2: replaces MPOMPOProduct(O1, O2; cutoff, maxdim, mindim), mpo.jl:149-169:
 *   O[i][(w1,w2), o, t, (w1',w2')] = sum_k O1[i][w1,o,k,w1'] O2[i][w2,k,t,w2'],
 * then the same two centre sweeps as tnh_mpo_apply (sites viewed as (D_l, d*d, D_r)). */
int tnh_mpo_mpo_product(tnh_ctx *ctx, int N, const tnh_buf *const *O1, const tnh_buf *const *O2,
                        double cutoff, int maxdim, int mindim, tnh_buf **O);

/* ------------------------------------------------ quantum-jump trajectories */

/* One QJMC problem: the Trotter gates of qjmc_gates (src/algorithms/mps/qjmc.jl:9-26, i.e.
 * trotterize of -iH - sum c^2 L^H L / 2), single-site jump operators, one-site observables
 * and the initial product state. Matrices are column-major, interleaved complex. */
typedef struct tnh_qjmc_problem {
  int N, d;                     /* sites, physical dimension */
  int nlayers;                  /* gate layers of the GateList, applied in order */
  const int *layer_count;       /* gates per layer */
  const int *gate_site;         /* per gate (all layers in order): first site, 1-based */
  const int *gate_range;        /* per gate: 1 (d x d) or 2 (G[o1,i1,o2,i2], d^4) */
  const int64_t *gate_offset;   /* per gate: offset into gate_data in complex numbers */
  const double *gate_data;
  int njump;                    /* jump terms c_t L_t on site jump_site[t] */
  const int *jump_site;
  const double *jump_coeff;     /* c_t (the rate is c_t^2 <L_t^H L_t>) */
  const double *jump_op;        /* d x d per term */
  int nobs;                     /* observables <O_o> on site obs_site[o], measured at t = 0 */
  const int *obs_site;          /* and every `savesteps` steps (QJMCOperators) */
  const double *obs_op;         /* d x d per observable */
  const double *init_state;     /* d complex amplitudes per site */
  double dt;
  int nsteps, savesteps;
  double cutoff;
  int maxdim, mindim;
} tnh_qjmc_problem;

/* Run ntraj trajectories of qjmc_simulation (qjmc.jl:28-167, classical branch) with the
 * step loop on the host side of this library and every tensor operation on the GPU:
 * applygates!, normalise, emission rates, jump test exp(-sum(rates) dt) and channel
 * choice, jump by applyop! and the centre sweeps of :103-107. Trajectory t consumes its
 * uniforms uniforms[t * nuniform ...] in the reference's order: per step one discarded
 * draw (:64), the jump test (:97), and the channel (:99) when a jump happens. nworkers
 * host threads each own a context (one HIP stream) on ctx's device and take trajectories
 * in turn. Outputs: meas[t][save][obs] (re, im) with nsteps / savesteps + 1 saves,
 * njumps[t], the first maxjumps jump indices (1-based) and steps per trajectory, and the
 * final maximum bond dimension maxbond[t]. */
int tnh_qjmc_run(tnh_ctx *ctx, const tnh_qjmc_problem *problem, int ntraj, const double *uniforms,
                 int64_t nuniform, int nworkers, double *meas, int *njumps, int *jumps,
                 int *jump_steps, int maxjumps, int *maxbond);

/* The same trajectories and outputs as tnh_qjmc_run (same algorithm and uniform order),
 * advanced together on ctx's stream: all site tensors of all trajectories stay resident in
 * fixed-capacity slots, and every launch gives each trajectory one workgroup that applies its
 * list of primitives (centre moves, two-site gate splits, one-site operators,
 * normalisation). Narrow slots (bond dimension at most min(16, 32 / d)) keep each split's
 * Gram eigensolve in LDS; trajectories that outgrow them continue inside this call, each from
 * the start of the step it outgrew them in (replayed on its own uniforms), on wide slots
 * (bond dimension 64, then up to 512 / d, capped by maxdim and the chain's largest Schmidt
 * rank): a 1024-thread workgroup per trajectory with the divide-and-conquer Gram eigensolve. Replaces the per-trajectory loop over
 * qjmc_simulation (qjmc.jl:28-167) for dissipative chains with uneven bond dimensions. A
 * trajectory whose split would keep more than the largest slot capacity (or whose
 * divide and conquer fails) stops there: its maxbond entry is set to -1 and its other
 * outputs are undefined, while the other trajectories complete (the call returns TNH_OK);
 * rerun the marked ones with tnh_qjmc_run, which has no such limit (d <= 8), on their own rows
 * of uniforms. */
int tnh_qjmc_run_batched(tnh_ctx *ctx, const tnh_qjmc_problem *problem, int ntraj,
                         const double *uniforms, int64_t nuniform, double *meas, int *njumps,
                         int *jumps, int *jump_steps, int maxjumps, int *maxbond);

/* ---------------------------------------------------- device-resident DMRG */

typedef struct tnh_dmrg_params {
  int krylovdim;  /* dmrg.jl:4 default 3 */
  int maxiter;    /* dmrg.jl:5 default 2 */
  double lanczos_tol; /* dmrg.jl:53 (1e-14) */
  double cutoff;  /* dmrg.jl:17 default 1e-12 */
  int maxdim;     /* dmrg.jl:18 default 1000 */
  int mindim;     /* dmrg.jl:19 default 1 */
  int precision;  /* TNH_PREC_* for the matvecs and environment updates */
  int nsites;     /* dmrg.jl:3: 2 (or 0) for the two-site update, 1 for the one-site update */
  int nonhermitian; /* dmrg.jl:6 ishermitian = false: Arnoldi for the Ritz value of smallest
                       real part instead of the Lanczos (needs a complex state) */
} tnh_dmrg_params;

/* Device-resident state of dmrg(psi, H::GMPS...), src/algorithms/mps/dmrg.jl:128-154:
 * copies the N site tensors of psi (any dtype; stored F64 or C128) and the nterms
 * MPOs (site tensors mpo[t*N + i]) with coefficients coeffs[t]; moves the centre of
 * psi to site 1 and builds all right blocks (projmps.jl:16-42). nterms may be 0 when
 * rank-1 terms are added before the first sweep. */
int tnh_dmrg_create(tnh_ctx *ctx, int N, const tnh_buf *const *psi, int nterms,
                    const tnh_buf *const *mpo, const double *coeffs, tnh_dmrg **out);
/* Add a rank-1 Hamiltonian term coeff |phi><phi|: ProjMPS(phi, psi; rank=2, squared=true,
 * coeff), which dmrg(psi, Hs...) builds for each rank-1 argument (dmrg.jl:144-145), with
 * overlap blocks (projmps.jl:50-95), the product coeff * conj(P) * sum(P .* theta) of
 * projmps.jl:135-144 (P = project(...), :152-179) and calculate = coeff <phi|psi>
 * (:192-216). phi: N site tensors (chi_l, d, chi_r) of psi's physical dimension, copied in
 * (a complex phi needs a state created from complex tensors). The overlap blocks are built
 * at the state's current centre and kept in FP64 in every precision; the term joins the
 * sum after the MPO terms. */
int tnh_dmrg_add_projector(tnh_dmrg *st, const tnh_buf *const *phi, double coeff);
int tnh_dmrg_destroy(tnh_dmrg *st);
/* One directional pass (the reference's "sweep", dmrg.jl:35-63): for each bond,
 * move the environment centre, form theta, eigsolve, replacesites!. direction 0 =
 * left-to-right. With p->nsites = 1 each site in turn is updated alone and the centre
 * then moves one site on by an untruncated split, normalised (replacesites!'s one-site
 * branch, gmps.jl:204-213): bond dimensions do not change. Returns the last Ritz value, the max bond dim and the max
 * discarded weight of the pass. In a TNH_PREC_F32 pass the block-Jacobi Gram
 * eigensolves of the splits (Gram sides above 512) stop at 1e-9 max_i G_ii, the scale
 * below which the pass's own FP32 rounding already blurs theta; FP64 passes use the FP64
 * rounding floor, and a following FP64 pass restores full accuracy. */
int tnh_dmrg_sweep(tnh_dmrg *st, int direction, const tnh_dmrg_params *p, double *energy,
                   int *maxbonddim, double *max_discarded);
/* The last Ritz value of the last pass (dmrg.jl:55 `cost = eig[1]`), complex: with
 * nonhermitian set the eigenvalue of smallest real part may have an imaginary part. */
int tnh_dmrg_last_ritz(tnh_dmrg *st, double *re, double *im);
/* Current energy <psi|H|psi> at the centre: calculate(Hs), dmrg.jl:22 (a rank-1 term
 * contributes coeff Re<phi|psi>, as its calculate does). */
int tnh_dmrg_energy(tnh_dmrg *st, int precision, double *energy);
/* Borrow site tensor `site` (1-based). Valid until the next sweep. */
int tnh_dmrg_site(tnh_dmrg *st, int site, const tnh_buf **out);
int tnh_dmrg_center(tnh_dmrg *st, int *center);
/* Bond dims chi_0..chi_N into dims[N+1]. */
int tnh_dmrg_bonddims(tnh_dmrg *st, int64_t *dims);

#ifdef __cplusplus
}
#endif

#endif /* TNHIP_H */
