/*
 * skgemm.h - C ABI of the MI355X-native FP64 Stream-K GEMM / SYRK library.
 *
 * This is the only surface host code links against.  Every entry point is
 * extern "C", takes plain pointers and sizes, and returns an int status
 * (SKGEMM_OK == 0, negative on error; skgemm_last_error() gives the text of
 * the calling thread's last error).  No HIP or torch types appear here: a
 * stream is passed as void* (a hipStream_t), a context as an opaque handle.
 *
 * Reference surface replaced (shriramch/stream-k-gemm, C++ class templates
 * with static members, host pointers, void returns):
 *   gemm<T>::compute_test / compute_streamk_packed / compute_streamk_packed_core
 *       gemm_crsp.cpp:261,316,369         -> skgemm_dgemm(_host)
 *   gemm<T>::compute_packed               gemm_opt.cpp:314  -> skgemm_dgemm(_host)
 *   gemm<T>::pack_A / pack_B              gemm_crsp.cpp:342,355 -> skgemm_pack
 *   sme_syrk<T>::compute                  include/sme_syrk.hpp:137,
 *                                         syrk_crsp.cpp:306 -> skgemm_dsyrk(_host)
 *   sme_gemm_trp_3d<T>::compute           include/sme_trp_3d.hpp:194
 *                                                         -> skgemm_dgemm(_host)
 *   linalg_pack_c16_f64 / unpack_c16      include/sme_pack.hpp:319-352
 *                                                         -> skgemm_pack
 *   sme_pack<T>::band_to_grid_tile /      include/sme_pack.hpp:178-313
 *     grid_to_band_tile                                   -> skgemm_transpose
 *   Hamiltonian::nloc_project_vectors_packed(_16to16)
 *                                         include/nloc_packed_opt2.hpp:229,453
 *                                                         -> skgemm_nloc_*
 *   Stencil_method::Packed::calc_laplacian_extended_packed_16to32 / 16to16
 *                                         include/stencil_extended_packed.hpp:47,276
 *                                                         -> skgemm_stencil_apply
 *   copy_interior_to_extended_packed + single-domain halo
 *                                         include/halo_packed.hpp:86,250
 *                                                         -> skgemm_halo_extend
 *   hbm_alloc / hbm_calloc / hbm_free     include/hbm_alloc.hpp:21-50
 *                                                         -> skgemm_dev_alloc/free
 *   run_fmopa_loop peak probe             sme_peak_microbench.cpp:16-60
 *                                                         -> skgemm_peak_probe
 * The header-only C++ facade include/skgemm.hpp keeps the reference's class
 * names and signatures on top of these functions.
 *
 * Operand layouts.  An operand is addressed by its free index o (m for A,
 * n for B, n for D) and its other index r (k for A and B, m for D):
 *   SKGEMM_ROW_MAJOR  A: m*ld + k      B: k*ld + n      D: m*ld + n
 *   SKGEMM_COL_MAJOR  A: m + k*ld      B: k + n*ld      D: m + n*ld
 *   SKGEMM_PACK16     (o/16)*ld*16 + r*16 + o%16   (ld >= extent of r)
 *   SKGEMM_PACK32     (o/32)*ld*32 + r*32 + o%32
 * PACK16 on A is the reference's PA [M/16][K][16] and the rotation's
 * grid-packed input; PACK32 on B is PB [N/32][K][32]; PACK16 on the SYRK
 * operand is P16(M,K) [ceil(M/16)][K][16]; PACK16 on D is the rotation's
 * band-packed output [nb/16][Nd][16].  Unlike the reference, any M, N, K >= 0
 * is accepted (packed panels past the extent are never read or written).
 *
 * Threading: a context owns one HIP stream, a device workspace and the tuning
 * table.  Calls on one context must not overlap from several host threads;
 * use one context per thread.  Device-pointer calls are asynchronous on the
 * context's stream; *_host calls copy in, run and copy out synchronously.
 */
#ifndef SKGEMM_H_
#define SKGEMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKGEMM_OK 0
#define SKGEMM_ERR_ARG (-1)       /* invalid argument (shape, ld, layout, null) */
#define SKGEMM_ERR_HIP (-2)       /* a HIP runtime call failed */
#define SKGEMM_ERR_NOMEM (-3)     /* device allocation failed or budget exceeded */
#define SKGEMM_ERR_TIMEOUT (-4)   /* a Stream-K fix-up wait hit its spin limit */
#define SKGEMM_ERR_NOTSUP (-5)    /* unsupported combination */
#define SKGEMM_ERR_STATE (-6)     /* call on a destroyed / wrong context */

enum skgemm_layout {
  SKGEMM_ROW_MAJOR = 0,
  SKGEMM_COL_MAJOR = 1,
  SKGEMM_PACK16 = 2,
  SKGEMM_PACK32 = 3
};

/* Precision of a GEMM/SYRK call. */
enum skgemm_precision {
  SKGEMM_FP64 = 0,        /* plain FP64 MFMA (v_mfma_f64_16x16x4) */
  SKGEMM_OZAKI_INT8 = 1   /* Ozaki split: INT8 slices on INT8 MFMA, FP64 sum */
};

typedef struct skgemm_ctx skgemm_ctx;

/* ---- library / context ------------------------------------------------- */

/* Text of the calling thread's last error ("" if none). */
const char* skgemm_last_error(void);

/* Library version string. */
const char* skgemm_version(void);

/* Create a context on HIP device `device`: a non-blocking stream at default
 * priority, device workspace and flags for Stream-K fix-up.
 * No counterpart in the reference (it has no device state). */
int skgemm_create(int device, skgemm_ctx** out);
int skgemm_destroy(skgemm_ctx* ctx);

/* Use caller's stream (hipStream_t) instead of the context's own; NULL
 * restores the context's stream. */
int skgemm_set_stream(skgemm_ctx* ctx, void* stream);
int skgemm_synchronize(skgemm_ctx* ctx);

/* Bound on the s_sleep rounds a Stream-K owner waits for a peer's partial
 * before it gives up and the call reports SKGEMM_ERR_TIMEOUT (default 2^26,
 * many seconds).  Exists so a broken co-residency assumption surfaces as an
 * error instead of a hang. */
int skgemm_set_spin_limit(skgemm_ctx* ctx, int64_t limit);

/* Failure-path test hook for the plain FP64 kernel's waiting fix-up.  1:
 * the workgroups that finish a split tile store their partials but never
 * raise their ready flags, so every owner's wait runs out (the spin limit is
 * clamped to 64 rounds meanwhile) and the call reports SKGEMM_ERR_TIMEOUT;
 * results are then undefined.  0 (default): normal.  No effect on the
 * wait-free fix-up.  No counterpart in the reference (it has no bounded
 * waits); for tests that force a timeout deterministically. */
int skgemm_set_fixup_hold(skgemm_ctx* ctx, int on);

/* Stream-K fix-up of the plain FP64 kernel on this context (single,
 * batched and host-pointer forms).  0 (default): the workgroup that starts
 * a split tile waits for the partials of the workgroups that finish it,
 * which needs every workgroup of the launch resident at once.  1
 * (wait-free): every contributor stores its partial and counts itself in;
 * the last to arrive adds the partials in the same order and writes the
 * tile, so nothing waits and launches on several streams may share the
 * device in any proportion.  Results are bit-identical between the modes;
 * the wait-free mode holds twice the partial-tile workspace.  No
 * counterpart in the reference (its threads join at an OpenMP barrier,
 * gemm_crsp.cpp:285). */
int skgemm_set_waitfree(skgemm_ctx* ctx, int on);

/* Which entries a SYRK on this context writes (plain, split form, batched
 * and their host forms):
 *   SKGEMM_SYRK_LOWER (default): the lower triangle i >= j of the row-major
 *     C only; nothing above the diagonal is written.
 *   SKGEMM_SYRK_DIAG_BLOCKS: the lower triangle plus the whole of every
 *     16 x 16 diagonal block (i / 16 == j / 16), as the reference's store
 *     does: it writes full 16 x 16 blocks of every tile pair tm1 <= tm2 of
 *     its column-major C, the diagonal pairs included
 *     (include/sme_syrk.hpp:57-109,137-149; syrk_crsp.cpp:137-244,306-348).
 *     Entries above the diagonal inside those blocks get the symmetric
 *     value alpha (A^T A)_ij + beta C_ij.  Unlike the reference, nothing
 *     outside the N x N matrix is written when N % 16 != 0.
 * The C++ facade's sme_syrk<double>::compute selects the second mode. */
#define SKGEMM_SYRK_LOWER 0
#define SKGEMM_SYRK_DIAG_BLOCKS 1
int skgemm_set_syrk_store(skgemm_ctx* ctx, int mode);

/* Device facts used by the scheduler: CU count, XCD count, LDS per CU. */
int skgemm_device_info(skgemm_ctx* ctx, int* num_cus, int* num_xcds,
                       int* lds_bytes_per_cu, int* clock_mhz);

/* ---- device memory (include/hbm_alloc.hpp:21-50) ----------------------- */

/* hbm_alloc / hbm_calloc / hbm_free on HBM of the context's device. */
int skgemm_dev_alloc(skgemm_ctx* ctx, size_t bytes, int zero, void** out);
int skgemm_dev_free(skgemm_ctx* ctx, void* ptr);
int skgemm_memcpy_h2d(skgemm_ctx* ctx, void* dst, const void* src, size_t bytes);
int skgemm_memcpy_d2h(skgemm_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- plain FP64 GEMM --------------------------------------------------- */

/* D = alpha * A * B + beta * D + gamma  (device pointers, async).
 * A is M x K, B is K x N, D is M x N with the layouts above.  gamma is the
 * reference's additive scalar (gemm_crsp.cpp:8, :297); beta is the BLAS
 * term the reference lacks (beta == 0 never reads D).
 * gemm_crsp.cpp:261-311 (compute_streamk_packed_core), gemm_opt.cpp:314-337
 * (compute_packed), include/sme_trp_3d.hpp:194-215 (rotation, alpha=1,
 * beta=gamma=0, A PACK16, B ROW_MAJOR, D PACK16). */
int skgemm_dgemm(skgemm_ctx* ctx, int layout_a, int layout_b, int layout_d,
                 int M, int N, int K, double alpha, const double* A,
                 int64_t lda, const double* B, int64_t ldb, double beta,
                 double* D, int64_t ldd, double gamma);

/* Host-pointer form: copies A, B and D to the device (D also when beta == 0:
 * packed and padded layouts hold elements outside the matrix that come back
 * unchanged), runs, copies D back, synchronously.  gemm_crsp.cpp:369-386
 * (compute_test). */
int skgemm_dgemm_host(skgemm_ctx* ctx, int layout_a, int layout_b,
                      int layout_d, int M, int N, int K, double alpha,
                      const double* A, int64_t lda, const double* B,
                      int64_t ldb, double beta, double* D, int64_t ldd,
                      double gamma);

/* ---- plain FP64 SYRK --------------------------------------------------- */

/* C = alpha * A^T * A + beta * C on the lower triangle (i >= j) of the
 * row-major N x N matrix C (== the reference's column-major upper
 * triangle), A is K x N in layout layout_a (B-operand convention: ROW_MAJOR
 * k*ld+n, COL_MAJOR k+n*ld, PACK16 = P16).  Entries above the diagonal are
 * not written, except inside the 16 x 16 diagonal blocks under
 * SKGEMM_SYRK_DIAG_BLOCKS (skgemm_set_syrk_store).  include/sme_syrk.hpp:137-149, syrk_crsp.cpp:306-348.
 * Unlike the reference (quirk: alpha applied to the last KC block only,
 * include/sme_syrk.hpp:123-133) alpha scales the whole product. */
int skgemm_dsyrk(skgemm_ctx* ctx, int layout_a, int N, int K, double alpha,
                 const double* A, int64_t lda, double beta, double* C,
                 int64_t ldc);
int skgemm_dsyrk_host(skgemm_ctx* ctx, int layout_a, int N, int K,
                      double alpha, const double* A, int64_t lda, double beta,
                      double* C, int64_t ldc);

/* Allocation-free form of skgemm_dsyrk (same arguments, same result to
 * 1e-13 relative Frobenius: the same plan, on the even Stream-K split).  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call makes no hipMalloc, hipFree
 * or other allocating call: the fix-up partials, flags and K-chunk sums use
 * a 128 MiB device array of the library's, one per device, held until the
 * process ends.  A problem whose plan needs more (flags + grid partial
 * tiles + kchunks * N * N doubles; e.g. N = 2048 with K chunks) is rejected
 * with SKGEMM_ERR_NOMEM and a message giving both sizes;
 * skgemm_set_kchunks(ctx, 1) lowers the need.  All such calls on one device
 * share the array: calls that overlap in time on one device (several
 * streams or contexts) race and must be serialised by the caller.
 * Asynchronous on the context's stream; a fix-up timeout is reported by the
 * next skgemm_synchronize. */
int skgemm_dsyrk_noalloc(skgemm_ctx* ctx, int layout_a, int N, int K,
                         double alpha, const double* A, int64_t lda,
                         double beta, double* C, int64_t ldc);

/* ---- Ozaki split form (INT8 slices, INT32 accumulation, FP64 sum) ------ */

/* Same contracts as skgemm_dgemm / skgemm_dsyrk (beta/gamma supported) with
 * the product computed in the Ozaki split form (scheme II, CRT variant):
 * power-of-two scaling of rows of A and columns of B from their 2-norms,
 * each scaled operand cut into `slices` INT8 residue planes modulo
 * pairwise-coprime moduli <= 256, one INT8 MFMA product per plane with
 * exact INT32 accumulation (K longer than 131072 is cut into chunks whose
 * residues are added before the CRT; K up to 32 such chunks, 4194304, else
 * SKGEMM_ERR_NOTSUP), and an exact CRT reconstruction rounded once to FP64.  The result does not depend on the schedule
 * (bit-identical for any plan).
 * slices: 8..16 moduli; 0 = 14 (each scaled row keeps about 54 bits of its
 * norm); SKGEMM_SLICES_AUTO = certifying: 14 moduli, then more while the
 * bound (below) exceeds 1e-13, the contract tolerance (the count its
 * scaling predicts, then one at a time; D is restored between attempts
 * when beta != 0), or plain FP64 when 16 moduli cannot reach it.  The
 * certifying mode reads its bound back (it is synchronous and cannot be
 * captured into a graph); skgemm_split_info gives the count it used.
 * Each row of A (column of B) is scaled by s = 2^e m / 8, m in 8..15, the
 * largest such value that keeps its 2-norm below 2^(h-1) (h = 55 at 14
 * moduli), and rounded to integers.
 * *err_bound (may be NULL; non-NULL makes the call synchronous) receives a
 * rigorous a-posteriori bound on ||C - AB||_F / ||AB||_F from the norms of
 * the actual rounding errors of the scaled inputs (7.4e-14 on Psi 65536 x
 * 1024 with 14 moduli, where the measured error is ~1.4e-16; Cauchy-Schwarz
 * on the error vectors makes it pessimistic by about sqrt(K)), or -1 when
 * the call was computed by the plain
 * FP64 path instead: a row of A or column of B could not be scaled (squared
 * norm not finite or outside [2^-1000, 2^1000]), or, certifying, 16 moduli
 * could not reach 1e-13.  Device pointers.  No counterpart in the
 * reference (its path is FP64 only). */
#define SKGEMM_SLICES_AUTO (-1)
int skgemm_dgemm_ozaki(skgemm_ctx* ctx, int layout_a, int layout_b,
                       int layout_d, int M, int N, int K, double alpha,
                       const double* A, int64_t lda, const double* B,
                       int64_t ldb, double beta, double* D, int64_t ldd,
                       double gamma, int slices, double* err_bound);
int skgemm_dsyrk_ozaki(skgemm_ctx* ctx, int layout_a, int N, int K,
                       double alpha, const double* A, int64_t lda, double beta,
                       double* C, int64_t ldc, int slices, double* err_bound);
/* Host-pointer forms of the two calls above (copy in, run, copy out). */
int skgemm_dgemm_ozaki_host(skgemm_ctx* ctx, int layout_a, int layout_b,
                            int layout_d, int M, int N, int K, double alpha,
                            const double* A, int64_t lda, const double* B,
                            int64_t ldb, double beta, double* D, int64_t ldd,
                            double gamma, int slices, double* err_bound);
int skgemm_dsyrk_ozaki_host(skgemm_ctx* ctx, int layout_a, int N, int K,
                            double alpha, const double* A, int64_t lda,
                            double beta, double* C, int64_t ldc, int slices,
                            double* err_bound);

/* ---- layout conversion (include/sme_pack.hpp, gemm_crsp.cpp:342-366) --- */

/* Copy a rows x cols matrix (element (r, o) in the free-index convention
 * above: r = k for A/B, o = m/n) from layout src_layout to dst_layout on
 * the device.  pack_A = (COL_MAJOR -> PACK16) with o = m,
 * pack_B = (ROW_MAJOR -> PACK32) with o = n, linalg_pack_c16_f64 =
 * (COL_MAJOR [Nd,nb] -> PACK16 over grids).  `o_extent` x `r_extent`. */
int skgemm_pack(skgemm_ctx* ctx, int src_layout, const double* src,
                int64_t ld_src, int dst_layout, double* dst, int64_t ld_dst,
                int o_extent, int r_extent);

/* Copy with the roles of the two indices exchanged: dst element (r = o0,
 * o = r0) = src element (o0, r0), so dst has o extent r_extent and r extent
 * o_extent (ld_dst is checked against those).  With src PACK16 of (o = band
 * b, r = grid g, ld_src = Nd) and dst PACK16 (ld_dst = nb) this is
 * band_to_grid_tile: band-packed [nb/16][Nd][16] -> grid-packed
 * [Nd/16][nb][16] (include/sme_pack.hpp:178-244); the same call with the
 * extents exchanged is grid_to_band_tile (include/sme_pack.hpp:246-313).
 * Plain layouts give an ordinary out-of-place transpose.  Device pointers,
 * asynchronous; src and dst must not overlap. */
int skgemm_transpose(skgemm_ctx* ctx, int src_layout, const double* src,
                     int64_t ld_src, int dst_layout, double* dst,
                     int64_t ld_dst, int o_extent, int r_extent);

/* ---- batched / grouped launch ------------------------------------------ */

/* One problem of a batch.  kind: 0 = GEMM D = alpha*A*B + beta*D + gamma,
 * 1 = SYRK C = alpha*A^T*A + beta*C (lower; M = N, C in D/ldd, layout_a;
 * B, ldb, layout_b, layout_d and gamma unused).  Device pointers. */
typedef struct skgemm_problem {
  int kind;
  int layout_a, layout_b, layout_d;
  int M, N, K;
  double alpha, beta, gamma;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* D;
  int64_t ldd;
} skgemm_problem;

/* Run `count` independent problems (outputs must not overlap) on ctx's
 * stream.  Plain FP64: ONE persistent Stream-K launch over the flattened
 * MAC-iteration space of all of them, one tile shape for the batch; the
 * problem table lives in device memory and is refreshed before each
 * launch, so with use_graph != 0 one captured graph per (tile, grid) is
 * replayed for any batch.  SKGEMM_OZAKI_INT8: the split-form sequence of
 * each problem in turn; with use_graph != 0 the whole sequence is one
 * graph, cached per exact problem list (pointers included) and recaptured
 * when a workspace buffer was reallocated.  Asynchronous.  No counterpart
 * in the reference. */
int skgemm_batched(skgemm_ctx* ctx, const skgemm_problem* problems, int count,
                   int precision, int slices, int use_graph);

/* ---- non-local projectors (include/nloc_packed_opt2.hpp:229-655) -------- */

/* result += V_nloc psi on band-packed orbitals, the reference's
 * nloc_project_vectors_packed (16-band-packed psi -> 32-band-packed
 * result, :229-447) and nloc_project_vectors_packed_16to16 (:453-655): per
 * projector p, X_p = dv chi_p^T psi[idx_p, :], X_p *= gamma_p,
 * result[idx_p, :] += chi_p X_p.
 * A projector set on Nd grid points is uploaded once.  Host arrays,
 * projectors concatenated in order: nrow[p] grid indices in
 * index_all[sum_{q<p} nrow_q ..] (in [0, Nd)), chi_p column-major
 * nrow x ncol in chi_all[sum_{q<p} nrow_q ncol_q ..], gamma_p (ncol) in
 * gamma_all[sum_{q<p} ncol_q ..].  As in the reference, gamma scales a
 * projector only when its atom_index differs from the previous
 * projector's (the first projector always); projectors with nrow or ncol
 * 0 contribute nothing; a grid point may belong to several projectors (the
 * contributions are added in projector order, without atomics).  The
 * reference's MPI_Allreduce of X over domain ranks has no counterpart (one
 * GPU holds the whole grid). */
typedef struct skgemm_nloc skgemm_nloc;
int skgemm_nloc_create(skgemm_ctx* ctx, int Nd, int num_proj, const int* nrow,
                       const int* ncol, const int* atom_index,
                       const int* index_all, const double* chi_all,
                       const double* gamma_all, skgemm_nloc** out);
/* Device pointers, asynchronous on the context's stream: psi16
 * [ceil(nb/16)][Nd][16] (16-byte aligned), result [ceil(nb/T)][Nd][T] with
 * T = result_tile (32 or 16).  The handle's workspace grows with nb. */
int skgemm_nloc_apply(skgemm_nloc* h, int nb, double dv, const double* psi16,
                      double* result, int result_tile);
int skgemm_nloc_destroy(skgemm_nloc* h);

/* ---- real-space stencil (include/stencil_extended_packed.hpp,
 *      include/halo_packed.hpp) ------------------------------------------- */

/* Local grid: 16-band-packed [ceil(nb/16)][K][16], K = ni nj nk, grid point
 * g = i + j ni + k ni nj.  Extended grid: the same with
 * K_ex = (ni + 2 FDn)(nj + 2 FDn)(nk + 2 FDn) and the interior at offset
 * (FDn, FDn, FDn).  1 <= FDn <= 8. */
enum {
  SKGEMM_HALO_ZERO = 0,     /* ghost points are 0 (isolated domain) */
  SKGEMM_HALO_PERIODIC = 1, /* ghost points wrap around (one periodic domain) */
  SKGEMM_HALO_KEEP = 2      /* interior copy only, ghost points untouched
                               (copy_interior_to_extended_packed) */
};
/* Interior copy (copy_interior_to_extended_packed, halo_packed.hpp:86-123)
 * and the halo of one domain: what fill_domain_par_ex_arr_packed
 * (halo_packed.hpp:250-398) receives from the neighbouring ranks is, for a
 * single periodic domain, the wrapped interior (every ghost point, corners
 * included).  Whole 16-band rows move.  Device pointers, asynchronous. */
int skgemm_halo_extend(skgemm_ctx* ctx, int ni, int nj, int nk, int nb,
                       int FDn, const double* local16, double* ext16,
                       int boundary);
/* H psi on the interior (calc_laplacian_extended_packed_16to32 /
 * _16to16, stencil_extended_packed.hpp:47-197, :276-402):
 *   out[g, b] = psi[g, b] (A_diag[g] + coef_0)
 *             + sum_{r=1..FDn} coef_x[r] (psi[g + r x] + psi[g - r x]) + y + z
 * read from the extended grid, written band-packed by out_tile (32 or 16),
 * bands b < nb only.  coef_x/y/z: host arrays of FDn + 1 (index 0 unused);
 * ext16, A_diag (K) and out: device pointers.  Asynchronous. */
int skgemm_stencil_apply(skgemm_ctx* ctx, int ni, int nj, int nk, int nb,
                         int FDn, const double* ext16, const double* A_diag,
                         const double* coef_x, const double* coef_y,
                         const double* coef_z, double coef_0, double* out,
                         int out_tile);

/* ---- tuning ------------------------------------------------------------- */

/* Force a tile configuration (index into the kernel table) and a persistent
 * grid size for subsequent calls on ctx; tile < 0 / grid <= 0 restore the
 * autotuner's choice.  The reference's analogue is rebuilding with
 * -DMC/-DNC/-DKC (sub_gemm.sh:31-52). */
int skgemm_set_tuning(skgemm_ctx* ctx, int tile, int grid);
/* Force the number of K chunks (one unit = one tile over one chunk; chunk
 * sums are reduced in chunk order); 0 restores the planner's choice, which
 * uses one chunk per XCD when a problem has fewer tiles than CUs and a long
 * K.  The reference's analogue is its split of K over threads
 * (gemm_crsp.cpp:264-279). */
int skgemm_set_kchunks(skgemm_ctx* ctx, int kchunks);

/* Number of tile configurations and their shapes. */
int skgemm_num_tiles(void);
int skgemm_tile_info(int tile, int* bm, int* bn, int* bk, int* waves);
/* Autotuner: times every candidate (tile configuration x persistent grid
 * size x K-chunk count) with real launches on scratch operands of this
 * shape and layout (layout_b is ignored for SYRK), records the fastest in
 * the context's table, and appends it to `cache_path` when not NULL (one
 * text line per shape: kind la lb M N K tile grid kchunks us).  The
 * reference's analogue is its -DMC/-DNC/-DKC rebuild sweep
 * (sub_gemm.sh:31-52, sub_syrk.sh:30-48).  kind 0 GEMM / 1 SYRK.  A GEMM
 * is tuned with a column-major D when layout_a is column-major (else a
 * row-major D); a GEMM with a column-major D is launched as its transpose
 * D^T = B^T A^T (row-major D^T, the same memory), and tuned plans, the cache
 * file and skgemm_query_plan describe that launched problem. */
int skgemm_autotune(skgemm_ctx* ctx, int kind, int layout_a, int layout_b,
                    int M, int N, int K, const char* cache_path,
                    int* best_tile, int* best_grid, int* best_kchunks,
                    double* best_us);
/* Load tuned plans written by skgemm_autotune. */
int skgemm_load_tuning(skgemm_ctx* ctx, const char* cache_path);
/* The plan the next call on this shape and layout would use. */
int skgemm_query_plan(skgemm_ctx* ctx, int kind, int layout_a, int layout_b,
                      int M, int N, int K, int* tile, int* grid,
                      int* kchunks, int64_t* tiles, int64_t* iters);

/* ---- measurement -------------------------------------------------------- */

/* Elapsed milliseconds of the last compute launch on this context, timed
 * with HIP events recorded on the stream the kernel ran on. */
int skgemm_last_kernel_ms(skgemm_ctx* ctx, double* ms);

/* Per-launch kernel timing without host synchronisation: after
 * skgemm_timing_enable(ctx, capacity) every compute launch on ctx is
 * bracketed by a pair of HIP events recorded on the launch stream (up to
 * `capacity` launches; 0 disables).  skgemm_timing_read waits for the last
 * recorded launch, writes up to `max` durations in launch order (ms) and
 * starts a new recording.  Each launch records one event pair (the ring's
 * while it has room, else the context's own), created without the
 * system-scope fence, so timing adds no cache flush between launches;
 * skgemm_timing_enable clears the last-launch time. */
int skgemm_timing_enable(skgemm_ctx* ctx, int capacity);
int skgemm_timing_read(skgemm_ctx* ctx, double* ms, int max, int* count);

/* Split-form stage timing: with on != 0, every split-form call records one
 * event pair per stage in the timing ring instead of one per call: (1) the
 * norms, scales and residue planes of the operands, (2) the INT8 products,
 * (3) the CRT pass with the predicated FP64 fallback; skgemm_timing_read
 * then returns three durations per such call, in that order.  (The
 * pipelined SYRK of SKGEMM_OZ_PIPE records one pair per call.) */
int skgemm_timing_stages(skgemm_ctx* ctx, int on);
/* Modulus count of the last split-form call on ctx (0 when it was computed
 * by the plain FP64 path). */
int skgemm_split_info(skgemm_ctx* ctx, int* moduli);

/* MFMA peak probe: back-to-back independent MFMAs on every SIMD, no memory
 * traffic.  which: 0 = v_mfma_f64_16x16x4_f64, 1 = v_mfma_i32_32x32x32_i8.
 * Returns achieved TFLOP/s (TOP/s for INT8).  sme_peak_microbench.cpp:16-60. */
int skgemm_peak_probe(skgemm_ctx* ctx, int which, double* tflops);
/* HBM copy probe: device-to-device streaming copy of `bytes`; GB/s counts
 * read + write bytes. */
int skgemm_copy_probe(skgemm_ctx* ctx, size_t bytes, double* gbps);

/* Fill device memory with seeded inputs: uniform(lo, hi) or normal(0, sigma)
 * from a counter-based SplitMix64 stream (element i depends only on seed, i). */
int skgemm_fill_uniform(skgemm_ctx* ctx, double* x, int64_t n, uint64_t seed,
                        double lo, double hi);
int skgemm_fill_normal(skgemm_ctx* ctx, double* x, int64_t n, uint64_t seed,
                       double sigma);
/* x[:, j] *= 10^u_j, u_j = uniform(0, umax) from stream `seed`, for a column-
 * major rows x cols matrix with leading dimension ld (dynamic-range inputs). */
int skgemm_scale_columns_pow10(skgemm_ctx* ctx, double* x, int rows, int cols,
                               int64_t ld, uint64_t seed, double umax);

/* ---- serving ------------------------------------------------------------ */

/* Continuous-batching request loop over resident systems.  See
 * skgemm_serve.h-style docs in skgemm_serve_* below. */
typedef struct skgemm_server skgemm_server;

/* Create a server with an HBM budget for per-request outputs and workspace
 * (bytes), at most `max_batch` problems per launch and `max_inflight`
 * launches in flight (separate streams, default priority). */
int skgemm_server_create(int device, size_t hbm_budget, int max_batch,
                         int max_inflight, int precision, int slices,
                         skgemm_server** out);
int skgemm_server_destroy(skgemm_server* s);
/* Register a resident system: Psi (Nd x Ns, column-major, ld = Nd) already
 * in device memory, owned by the caller.  Returns its id. */
int skgemm_server_add_system(skgemm_server* s, const double* psi, int Nd,
                             int Ns, int* id);
/* Submit a request: kind 0 = overlap Ms = Psi^T Psi (lower, row-major Ns x
 * Ns written to out), kind 1 = rotation out = Psi * Q (Q device Ns x Ns
 * row-major; out Nd x Ns column-major).  `out` is a device buffer owned by
 * the caller.  Thread-safe.  Returns a ticket. */
int skgemm_server_submit(skgemm_server* s, int system, int kind,
                         const double* q, double* out, int64_t* ticket);
/* Block until ticket completes; *latency_us = completion - submit. */
int skgemm_server_wait(skgemm_server* s, int64_t ticket, double* latency_us);
/* Aggregate statistics since creation. */
int skgemm_server_stats(skgemm_server* s, int64_t* completed, int64_t* launches,
                        double* flops, double* busy_s, int* peak_inflight);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* SKGEMM_H_ */
