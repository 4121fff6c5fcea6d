/*
 * mpk.h -- C ABI of libmpk.so, the MI355X (gfx950) Krylov core.
 *
 * The 59 mp_* entry points carry the names, arity and argument kinds of the
 * reference's C ABI (src/capi.mojo, declared for ctypes in
 * python/mojopetsc4py/_lib.py:23-87): every argument is an int64 (an address
 * or an integer) or a double, and nothing is thrown across the boundary.
 * What changes is where the buffers live:
 *
 *   [D]  device address (hipMalloc'd, e.g. by mpk_malloc); index arrays on
 *        the device are int32 (row pointers, columns, permutations, factor
 *        index arrays), values float64;
 *   [H]  host address; host index arrays are int64 as in the reference.
 *
 * Scalar results (dot, norms, min/max) are written to [H] slots and the call
 * returns after they are ready.  Vector work is queued on the library's
 * stream; a call that returns data to the host synchronises it.  Callers own
 * all memory, as in the reference; the library keeps only its own scratch
 * (reduction partials, sweep buffers, solver state).
 *
 * Errors: the first failure is recorded; mpk_status() returns its code
 * (0 = none) and mpk_last_error() its message.
 */
#ifndef MPK_H
#define MPK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ runtime
 * No counterpart in the reference (it has no device): context, memory and
 * error plumbing for the device buffers the mp_* calls take. */
int64_t mpk_init(int64_t device);                        /* -1: current device */
int64_t mpk_status(void);
int64_t mpk_last_error(int64_t buf, int64_t len);       /* [H] char buffer */
void mpk_clear_error(void);
int64_t mpk_device_name(int64_t buf, int64_t len);      /* [H] char buffer */
int64_t mpk_malloc(int64_t bytes);                       /* returns [D] or 0 */
void mpk_free(int64_t ptr);
void mpk_memcpy_htod(int64_t dst, int64_t src, int64_t bytes);  /* synchronous */
void mpk_memcpy_dtoh(int64_t dst, int64_t src, int64_t bytes);  /* synchronous */
void mpk_memcpy_dtod(int64_t dst, int64_t src, int64_t bytes);  /* queued */
void mpk_memset(int64_t dst, int64_t value, int64_t bytes);     /* queued */
void mpk_synchronize(void);
/* The order of every floating-point reduction (dot, norms, sums, MDot) that
 * the mp_* calls and the solvers compute.  0 (the default): a fixed
 * tile-tree order, the same bits on every run.  1: the reference's SIMD
 * order (src/simdvec.mojo:332-473 with SW = 4, BLK = 16, src/simd.mojo:18,
 * 32-33; _mdot_basis, src/ksp.mojo:956-1062), computed by one workgroup as
 * 16 serial chains per reduction -- the reference's bits, at the cost of a
 * dependency chain of n / 16 steps per reduction.  The reference itself has
 * no such switch: its order is the one its build host's SIMD width gives.
 * MPK_REDUCTION=ref|dev sets the initial value.  Returns the previous order;
 * any other value is an error (status 4) and leaves the order unchanged. */
int64_t mpk_set_reduction_order(int64_t order);
int64_t mpk_get_reduction_order(void);
/* Bounded sweep waits (no counterpart in the reference, whose sweeps are
 * serial loops that never wait).  A triangular or SOR sweep on the device
 * waits for rows other workgroups publish; each such wait gives up after
 * `polls` polls (default 2^24, at least about a second), or within 256 polls
 * once mpk_sweep_abort(1) has raised the abort word, takes a NaN for the value
 * and lets its kernel run to the end.  The library then reports status 10 at
 * its next synchronisation (mpk_synchronize, mpk_status, any solve).  A
 * well-formed plan never gives up.  mpk_set_sweep_wait_limit returns the
 * previous limit (a value outside 1..2^32-1 is error 4); mpk_sweep_abort(0)
 * lowers the word again, returns 0, or -1 without a device.  The abort word
 * is host memory the running kernels read, so another host thread may raise
 * it while a sweep runs. */
int64_t mpk_set_sweep_wait_limit(int64_t polls);
int64_t mpk_sweep_abort(int64_t on);
/* The number of Krylov solves since the library loaded that ran as one
 * launch of one workgroup (small systems in the tile-tree order: n <= 16384
 * with wave-plan sweeps, n <= 4096 with a preconditioner without sweeps;
 * MPK_SMALL_ROWS=k lowers the limit, 0 routes every solve to the general
 * path).  A diagnostic with no counterpart in the reference: both paths give
 * the same bits. */
int64_t mpk_small_solves(void);
/* average ms of `reps` launches of the fused CG product (w = A p over the
 * sliced copy with the tile partials of p.w, k_cg_wdot), HIP events on the
 * library stream; -1 without a sliced copy.  A measurement entry with no
 * counterpart in the reference. */
double mpk_time_cg_wdot(int64_t p, int64_t w, int64_t irow, int64_t nrow, int64_t soff, int64_t sv, int64_t sc,
                        int64_t reps);
/* average ms of `reps` SpMV launches (over the sliced copy when soff != 0),
 * HIP events on the library stream */
double mpk_time_spmv(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow, int64_t soff,
                     int64_t sv, int64_t sc, int64_t reps);

/* ------------------------------------------------------------------ Mat
 * capi.mojo:32-65 -> ported.mojo:380-461.  irow/icol [D] int32, data [D]. */
void mp_mat_mult(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow);
void mp_mat_mult_add(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow);
void mp_mat_mult_transpose(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t nrow,
                           int64_t ncols);
void mp_mat_get_diagonal(int64_t diag, int64_t irow, int64_t icol, int64_t data, int64_t nrow);
void mp_mat_scale(int64_t data, int64_t nnz, double alpha);
void mp_mat_zero_entries(int64_t data, int64_t nnz);
/* No counterpart in the reference: the transpose map of a CSR pattern, kept
 * by the caller so repeated transposed products skip rebuilding it.  [D]
 * int32 trp[ncols + 1], trow[nnz], tpos[nnz]; returns 0 on success. */
int64_t mpk_csr_transpose(int64_t irow, int64_t icol, int64_t nrow, int64_t ncols, int64_t trp, int64_t trow,
                          int64_t tpos);
void mpk_mat_mult_transpose_t(int64_t x, int64_t y, int64_t data, int64_t ncols, int64_t trp, int64_t trow,
                              int64_t tpos);
/* No counterpart in the reference: Mat assembly on the device.
 * mpk_mat_fold folds a log of nw entry writes -- rows, cols [H] int64, vals
 * [H] float64, adds [H] uint8 (1: add, 0: insert), in call order -- into CSR
 * arrays with the per-entry semantics of the reference's Mat
 * (python/mojopetsc4py/PETSc/Mat.py:383-439): an insert replaces, an add
 * accumulates onto the value so far (0.0 if none), every entry ever written
 * stays, columns ascending.  Out [D]: rp int32[nrows + 1], ci int32[nw],
 * v float64[nw] (room for nw entries).  Returns the entry count, or -1
 * (status 4: a write outside the matrix, or nw / nrows / ncols past int32).
 * mpk_dense_to_csr: the nonzeros (!= 0.0, NaN included) of dense [D] float64
 * nrows x ncols row-major (fewer than 2^31 elements) as rp [D] int32
 * [nrows + 1], ci [D] int32, v [D] float64; returns the count, or minus it
 * (rp written) when ci or v is 0 or cap is smaller, or -1 on error. */
int64_t mpk_mat_fold(int64_t rows, int64_t cols, int64_t vals, int64_t adds, int64_t nw, int64_t nrows,
                     int64_t ncols, int64_t rp, int64_t ci, int64_t v);
int64_t mpk_dense_to_csr(int64_t dense, int64_t nrows, int64_t ncols, int64_t rp, int64_t ci, int64_t v,
                         int64_t cap);
/* No counterpart in the reference: the sliced (64-row, column-major) copy of a
 * CSR matrix that the products read with coalesced loads; same row sums bit
 * for bit.  [D] soff[ceil(nrow/64) + 1] int32, sv[padded] float64, sc of
 * 6 * padded bytes, padded = mpk_sell_size().  sc holds the int32 column of
 * each slot, then (when every entry lies within 32767 columns of its row)
 * the same slots as int16 offsets from the row, which the products read;
 * soff[ceil(nrow/64)] is then -1 - padded instead of padded.  Offsets and
 * slot indices are int32: mpk_sell_build refuses (status 4, returns -1) a
 * copy of more than 2^31 - 1 padded slots, which the row kernel then
 * multiplies.  mpk_sell_size_host: the padded count from [H] int64 row
 * pointers, or -1 past that limit; needs no device. */
int64_t mpk_sell_size(int64_t irow, int64_t nrow);
int64_t mpk_sell_size_host(int64_t irow, int64_t nrow);
int64_t mpk_sell_build(int64_t irow, int64_t icol, int64_t data, int64_t nrow, int64_t soff, int64_t sv,
                       int64_t sc);
void mpk_sell_mult(int64_t x, int64_t y, int64_t irow, int64_t soff, int64_t sv, int64_t sc, int64_t nrow,
                   int64_t add);

/* ------------------------------------------------------------------ Vec
 * capi.mojo:69-238 -> ported.mojo:75-335.  Vectors [D]; val, z, idx, sval,
 * rval, v [H]; mp_vec_maxpy's alpha [H] doubles and x [H] array of [D]. */
void mp_vec_pointwise_mult(int64_t w, int64_t x, int64_t y, int64_t n);
void mp_vec_pointwise_divide(int64_t w, int64_t x, int64_t y, int64_t n);
void mp_vec_pointwise_min(int64_t w, int64_t x, int64_t y, int64_t n);
void mp_vec_pointwise_max(int64_t w, int64_t x, int64_t y, int64_t n);
void mp_vec_pointwise_max_abs(int64_t w, int64_t x, int64_t y, int64_t n);
void mp_vec_norm(int64_t x, int64_t ntype, int64_t z, int64_t n);
void mp_vec_dot(int64_t x, int64_t y, int64_t val, int64_t n);
/* No counterpart in the reference: mp_vec_dot for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  x, y [D]; the result goes to val [D] (one
 * double), queued on the library's stream with no host wait.  Its reduction
 * partials live in a fixed 512 KB device array (65,536 doubles) that each
 * GPU holds from mpk_init until the process ends, so n is at most
 * 65,536 x 2048 = 134,217,728 (status 4 past that, nothing queued).
 * mpk_init must have run (status 4 otherwise): the call never initialises
 * the library, allocates or frees.  Same bits as mp_vec_dot in either
 * reduction order.  All calls on one GPU share the array: calls that
 * overlap on it (from several host threads at once) race.  Returns 0, or -1
 * on error. */
int64_t mpk_vec_dot_noalloc(int64_t x, int64_t y, int64_t val, int64_t n);
void mp_vec_tdot(int64_t x, int64_t y, int64_t val, int64_t n);
void mp_vec_sum(int64_t x, int64_t val, int64_t n);
void mp_vec_mean(int64_t x, int64_t val, int64_t n);
void mp_vec_dot_norm2(int64_t x, int64_t y, int64_t sval, int64_t rval, int64_t n);
void mp_vec_max(int64_t x, int64_t idx, int64_t z, int64_t n);
void mp_vec_min(int64_t x, int64_t idx, int64_t z, int64_t n);
void mp_vec_normalize(int64_t x, int64_t v, int64_t n);
void mp_vec_reciprocal(int64_t x, int64_t n);
void mp_vec_exp(int64_t x, int64_t n);
void mp_vec_log(int64_t x, int64_t n);
void mp_vec_sqrtabs(int64_t x, int64_t n);
void mp_vec_chop(int64_t x, int64_t n, double tol);
void mp_vec_abs(int64_t x, int64_t n);
void mp_vec_conjugate(int64_t x, int64_t n);
void mp_vec_copy(int64_t x, int64_t y, int64_t n);
void mp_vec_swap(int64_t x, int64_t y, int64_t n);
void mp_vec_set(int64_t x, double alpha, int64_t n);
void mp_vec_zero_entries(int64_t x, int64_t n);
void mp_vec_scale(int64_t x, double alpha, int64_t n);
void mp_vec_shift(int64_t x, double alpha, int64_t n);
void mp_vec_axpy(int64_t y, double alpha, int64_t x, int64_t n);
void mp_vec_aypx(int64_t y, double alpha, int64_t x, int64_t n);
void mp_vec_axpby(int64_t y, double alpha, double beta, int64_t x, int64_t n);
void mp_vec_axpbypcz(int64_t z, double alpha, double beta, double gamma, int64_t x, int64_t y, int64_t n);
void mp_vec_waxpy(int64_t w, double alpha, int64_t x, int64_t y, int64_t n);
void mp_vec_maxpy(int64_t y, int64_t nv, int64_t alpha, int64_t x, int64_t n);
void mp_vec_max_pointwise_divide(int64_t x, int64_t y, int64_t rval, int64_t n);

/* ------------------------------------------------------------------ PC
 * capi.mojo:242-379. */
/* ported.mojo:354-377; irow/icol/diag [D] int32; returns the diagDense flag */
int64_t mp_mat_get_diagonal_markers(int64_t irow, int64_t icol, int64_t diag, int64_t nrow);
/* ported.mojo:502-567; all arrays [D] */
void mp_pc_jacobi_setup(int64_t diag, int64_t jtype, int64_t useabs, int64_t fixdiag, double scale, int64_t irow,
                        int64_t icol, int64_t data, int64_t nrow, int64_t didx);
void mp_pc_jacobi_setup_symmetric(int64_t diagsqrt, int64_t jtype, int64_t irow, int64_t icol, int64_t data,
                                  int64_t nrow);
void mp_pc_jacobi_apply(int64_t y, int64_t x, int64_t diag, int64_t n);
/* ported.mojo:633-748 with its = lits = 1 (capi.mojo:272-280); all arrays [D] */
void mp_pc_sor_apply(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t dmark, int64_t idiag,
                     int64_t mdiag, int64_t t, double omega, double fshift, int64_t flag, int64_t nrow);
/* ported.mojo:864-899; factor arrays [D] int32 (layout of factor.mojo:31-67) */
void mp_pc_ilu_apply(int64_t b, int64_t x, int64_t tmp, int64_t bi, int64_t bj, int64_t ba, int64_t bdiag,
                     int64_t r, int64_t c, int64_t nrow);
/* ported.mojo:902-952; factor arrays [D] int32; returns 0 */
int64_t mp_pc_icc_apply(int64_t b, int64_t x, int64_t ui, int64_t uj, int64_t ua, int64_t udiag, int64_t c,
                        int64_t scratch, int64_t nrow);
/* ported.mojo:957-1124, factor.mojo:31-412: set up on the host, all [H] int64 */
int64_t mp_pc_ilu_numeric(int64_t ba, int64_t rtmp, int64_t irow, int64_t icol, int64_t data, int64_t bi,
                          int64_t bj, int64_t bdiag, int64_t r, int64_t ic, int64_t nrow);
void mp_pc_icc_numeric(int64_t ua, int64_t rtmp, int64_t il, int64_t c2r, int64_t irow, int64_t icol, int64_t data,
                       int64_t ui, int64_t uj, int64_t udiag, int64_t r, int64_t riip, int64_t nrow);
int64_t mp_ilu_symbolic_ilu0(int64_t bi, int64_t bj, int64_t bdiag, int64_t adiag, int64_t irow, int64_t icol,
                             int64_t nrow);
int64_t mp_ilu_symbolic(int64_t bi, int64_t bj, int64_t bjlvl, int64_t bdiag, int64_t irow, int64_t icol,
                        int64_t r, int64_t ic, int64_t levels, int64_t cap, int64_t scratch, int64_t nrow);
void mp_ilu_contiguous_lu(int64_t bjsrc, int64_t bi, int64_t bdiag, int64_t bj, int64_t n);
int64_t mp_icc_symbolic(int64_t ui, int64_t uj, int64_t ujlvl, int64_t udiag, int64_t irow, int64_t icol,
                        int64_t r, int64_t riip, int64_t levels, int64_t cap, int64_t scratch, int64_t nrow,
                        int64_t scratchlen);
void mp_icc_contiguous_cholesky(int64_t ujsrc, int64_t ui, int64_t udiag, int64_t bi, int64_t bdiag, int64_t bj,
                                int64_t n);
void mp_ordering_rcm(int64_t perm, int64_t level, int64_t visited, int64_t queue, int64_t irow, int64_t icol,
                     int64_t nrow);
/* MatILUFactorNumeric_SeqAIJ (ported.mojo:957-1037) on the device, in place
 * of mp_pc_ilu_numeric's host pass, with the same bits.  mpk_ilu_levels
 * (host, needs no device): level sets of the factor's L part from bi, bj,
 * bdiag [H] int64; lrow [H] int32[nrow] receives the rows level by level,
 * lptr [H] int64[nlev + 1] the level bounds; returns nlev, or -2 when some
 * row lacks its own diagonal slot or ascending L / U columns (then only the
 * serial mp_pc_ilu_numeric reproduces the reference).  mpk_ilu_numeric: A as
 * irow, icol [D] int32 and data [D] float64; r, ic (the row permutation and
 * its inverse), bi, bj, bdiag and lrow [D] int32; lptr [H] int64; ba [D]
 * float64[bdiag[0] + 1] receives the factor.  Returns 0, or -1 on error. */
int64_t mpk_ilu_levels(int64_t bi, int64_t bj, int64_t bdiag, int64_t nrow, int64_t lrow, int64_t lptr);
int64_t mpk_ilu_numeric(int64_t nrow, int64_t irow, int64_t icol, int64_t data, int64_t r, int64_t ic, int64_t bi,
                        int64_t bj, int64_t bdiag, int64_t lrow, int64_t lptr, int64_t nlev, int64_t ba);
/* MatICCFactorNumeric_SeqAIJ (ported.mojo:1040-1124) on the device, in place
 * of mp_pc_icc_numeric's host pass, with the same bits.  mpk_icc_visits
 * (host, needs no device): the pass's row chains replayed without
 * arithmetic.  bi, bj [H] int64: the layout of mp_icc_contiguous_cholesky;
 * irow, icol, rip, riip [H] int64: A and the permutation with its inverse.
 * Out [H]: vptr int32[nrow + 1], vrow and vpos int32[bi[nrow]] (the rows
 * that update row k, in order, with the position of their entry (i, k)),
 * lrow int32[nrow] and lptr int64[nlev + 1] (level sets); returns nlev, or
 * -2 when a row lacks ascending strict columns or a diagonal slot, or its
 * result would depend on values left over from earlier rows (then only the
 * serial mp_pc_icc_numeric reproduces the reference).
 * mpk_icc_numeric: irow, icol [D] int32 and data [D] float64 (A); rip, riip,
 * bi, bj, vptr, vrow, vpos, lrow [D] int32; lptr [H] int64; u0 [D]
 * float64[bi[nrow]] scratch; ua [D] float64[bi[nrow]] receives the factor.
 * Returns 0, or -1 on error. */
int64_t mpk_icc_visits(int64_t bi, int64_t bj, int64_t irow, int64_t icol, int64_t rip, int64_t riip, int64_t nrow,
                       int64_t vptr, int64_t vrow, int64_t vpos, int64_t lrow, int64_t lptr);
int64_t mpk_icc_numeric(int64_t nrow, int64_t irow, int64_t icol, int64_t data, int64_t rip, int64_t riip,
                        int64_t bi, int64_t bj, int64_t vptr, int64_t vrow, int64_t vpos, int64_t lrow, int64_t lptr,
                        int64_t nlev, int64_t u0, int64_t ua);
/* No counterpart in the reference: the transpose of an IC factor's strict
 * part, which the device forward sweep gathers over.  [D] int32 arrays:
 * trp[nrow + 1], trow[m], tpos[m], tmid[nrow], m = mpk_icc_transpose_size().
 * Returns the number of "late" entries (strict columns below the diagonal,
 * which the reference's level-of-fill pattern can hold), or -1. */
int64_t mpk_icc_transpose_size(int64_t ui, int64_t udiag, int64_t nrow);
int64_t mpk_icc_transpose(int64_t ui, int64_t uj, int64_t udiag, int64_t nrow, int64_t trp, int64_t trow,
                          int64_t tpos, int64_t tmid);
/* No counterpart in the reference: the IC(0) analysis with the natural
 * ordering on the device, giving exactly the arrays of mp_icc_symbolic at
 * level 0 + mp_icc_contiguous_cholesky, mpk_icc_transpose and mpk_icc_visits
 * (whose replay of the row chains becomes a sort by a closed-form key).
 * irow, icol [D] int32: A.  mpk_ic0_size returns bi[nrow] (each factor row
 * holds A's columns above the diagonal and its diagonal slot), or -2 when a
 * row of A is not strictly ascending within 0..nrow-1, a column of the
 * strict part holds over 256 entries, or bi[nrow] passes int32 (then the
 * host passes apply), or -1 on error.  mpk_ic0_analyse writes, with
 * m = bi[nrow] - nrow, [D] int32: bi[nrow + 1], bj[bi[nrow]], bdiag[nrow + 1]
 * (the contiguous Cholesky layout); trp[nrow + 1], trow[m], tpos[m],
 * tmid[nrow] (the transpose, no late entries); vrow[m], vpos[m] (the visits,
 * vptr being trp); lrow[nrow]; and lptr [H] int64[nlev + 1] (room for
 * nrow + 1).  Returns nlev, or -1. */
int64_t mpk_ic0_size(int64_t irow, int64_t icol, int64_t nrow);
/* No counterpart in the reference: the ILU(0) analysis with the natural
 * ordering on the device, giving exactly the arrays of mp_ilu_symbolic_ilu0
 * and mpk_ilu_levels.  irow, icol [D] int32: A.  mpk_ilu0_size returns the
 * factor's entry count (bdiag[0] + 1), or -2 when a row of A is not strictly
 * ascending within 0..nrow-1 or lacks its diagonal, or the count passes int32
 * (then the host passes apply), or -1 on error.  mpk_ilu0_analyse writes
 * [D] int32 bi[nrow + 1], bj[count], bdiag[nrow + 1] (the contiguous LU
 * layout) and lrow[nrow], and lptr [H] int64[nlev + 1] (room for nrow + 1).
 * Returns nlev, or -1. */
int64_t mpk_ilu0_size(int64_t irow, int64_t icol, int64_t nrow);
int64_t mpk_ilu0_analyse(int64_t irow, int64_t icol, int64_t nrow, int64_t bi, int64_t bj, int64_t bdiag,
                         int64_t lrow, int64_t lptr);
int64_t mpk_ic0_analyse(int64_t irow, int64_t icol, int64_t nrow, int64_t bi, int64_t bj, int64_t bdiag, int64_t trp,
                        int64_t trow, int64_t tpos, int64_t tmid, int64_t vrow, int64_t vpos, int64_t lrow,
                        int64_t lptr);
/* No counterpart in the reference: block plans for the two sweeps of the PC a
 * descriptor describes (SOR, ILU, IC).  A sweep's rows are cut into blocks of
 * consecutive rows; a workgroup runs a block level by level (levels of the
 * dependencies inside the block) with the block's values in LDS.  The plan
 * copies the sweep's coefficients, so it is rebuilt with every new factor.
 * mpk_pc_schedule_size returns the bytes to allocate; mpk_pc_schedule writes
 * the plans into plan [D] (cap bytes) and returns the deepest block's level
 * count (0: the PC has no sweeps; -1: error).  With pci[11] = plan (unused by
 * the reference, 0 = plain row order) every apply runs its sweeps over the
 * plans; the arithmetic of each row is unchanged. */
int64_t mpk_pc_schedule_size(int64_t n, int64_t pci, int64_t pcf);
int64_t mpk_pc_schedule(int64_t n, int64_t pci, int64_t pcf, int64_t plan, int64_t cap);
/* No counterpart in the reference: the host part of mpk_pc_schedule for one
 * sweep.  off [H] int64[n+1], col [H] int32, coef [H] float64: every entry of
 * every row in the row's own order; forward 1 visits rows 0..n-1, 0 visits
 * n-1..0; bmax: most rows per block.  out [H] (cap bytes) receives the plan
 * (layout in mpk_sweep.h).  Returns its size, or minus the size needed.
 * mpk_sweep_plan_bound: an upper bound of that size.  Needs no device. */
int64_t mpk_sweep_plan_bound(int64_t n, int64_t m, int64_t bmax);
int64_t mpk_sweep_plan(int64_t n, int64_t forward, int64_t off, int64_t col, int64_t coef, int64_t bmax, int64_t out,
                       int64_t cap);
/* No counterpart in the reference: y = M^-1 x through a descriptor, as the
 * solvers apply it (KSP_PCApply, ksp.mojo:155-198, left side). x, y [D]. */
void mpk_pc_apply(int64_t n, int64_t pci, int64_t pcf, int64_t x, int64_t y);
/* mp_pc_icc_apply with that transpose supplied (no per-call rebuild) */
int64_t mpk_pc_icc_apply_t(int64_t b, int64_t x, int64_t ui, int64_t uj, int64_t ua, int64_t udiag, int64_t c,
                           int64_t trp, int64_t trow, int64_t tpos, int64_t tmid, int64_t nlate, int64_t nrow);

/* ------------------------------------------------------------------ KSP
 * capi.mojo:383-424 -> ksp.mojo.  irow/icol/data, b, x, work, vv, prevec [D];
 * cfg, cfgi, pci, pcf, hs [H].  cfgi[13..15] (unused by the reference) may
 * hold a sliced copy of the matrix (soff, sv, sc, see mpk_sell_build); when
 * they are 0 the solve builds one for its duration.  pci holds [D] addresses: Jacobi {2, diag,
 * diagsqrt}; SOR {4, irow, icol, data, dmark, idiag, mdiag, work, ..., flag at
 * 10}; ILU {5, bi, bj, ba, bdiag, r, c}; IC {6, bi, bj, ba, bdiag, r, trp,
 * trow, tpos, tmid, nlate} with the transpose made by mpk_icc_transpose;
 * pci[11] may hold block plans from mpk_pc_schedule (SOR, ILU, IC). */
void mp_ksp_solve_cg(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf, int64_t pstmp,
                     int64_t work, int64_t cfg, int64_t cfgi, int64_t b, int64_t x);
void mp_ksp_solve_bcgs(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                       int64_t pstmp, int64_t work, int64_t cfg, int64_t cfgi, int64_t b, int64_t x);
void mp_ksp_solve_gmres(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                        int64_t pstmp, int64_t work, int64_t vv, int64_t hs, int64_t cfg, int64_t cfgi, int64_t b,
                        int64_t x);
void mp_ksp_solve_fgmres(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t pci, int64_t pcf,
                         int64_t pstmp, int64_t work, int64_t vv, int64_t prevec, int64_t hs, int64_t cfg,
                         int64_t cfgi, int64_t b, int64_t x);

/* ------------------------------------------------------------------ mixed precision
 * No counterpart in the reference: solves whose inner Krylov iteration, SpMV
 * and preconditioner run on FP32 vectors -- matrix and preconditioner values
 * in FP32 (prec 32) or in FP16 with one FP32 scale per row (prec 16), FP32
 * accumulation -- inside FP64 iterative refinement; dot products and norms
 * sum in FP64 and the GMRES least squares is FP64.
 * mpk_mixed_create: the FP64 matrix [D] (irow, icol, data) with its sliced
 * copy (soff, sv, sc from mpk_sell_build) and a PC descriptor (pci [H], pcf
 * [H]; ilu or sor with plans from mpk_pc_schedule, or none).  inner: 0 gmres
 * (left preconditioned, classical Gram-Schmidt, restart <= 64), 1 bcgs
 * (right preconditioned).  tri_iters 0: the preconditioner's sweeps run in
 * dependency order over its block plans, exactly; k > 0: each sweep is
 * approximated by k Jacobi passes over its entries (a truncated Neumann
 * series; bandwidth-bound, no ordering).  Returns a handle, 0 on error.  The caller keeps
 * the matrix, its sliced copy and the PC alive while the handle lives.
 * mpk_mixed_solve: x [D] from b [D]; cfg [H] float64 in {rtol, inner_rtol,
 * max_outer, max_inner, gmres refinement (0 never, 1 if needed, 2 always),
 * x_is_zero}, out cfg[8] final ||b - A x|| / ||b|| (FP64), cfg[9] outer
 * steps, cfg[10] inner steps, cfg[11] ||b||, cfg[12] 1 if converged.
 * Returns 0, or -1 on error. */
int64_t mpk_mixed_create(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t soff, int64_t sv, int64_t sc,
                         int64_t pci, int64_t pcf, int64_t prec, int64_t inner, int64_t restart, int64_t tri_iters);
int64_t mpk_mixed_solve(int64_t handle, int64_t b, int64_t x, int64_t cfg);
void mpk_mixed_free(int64_t handle);

/* ------------------------------------------------------------------ batched solves
 * No counterpart in the reference: many independent small systems solved by
 * one persistent launch, a workgroup per system at a time (largest first):
 * FP64 iterative refinement around an FP32 inner CG + Jacobi (kind 0) or
 * GMRES(20) + ILU(0) (kind 1, classical Gram-Schmidt), FP64 dots, scalars and
 * least squares.  All arrays [D], concatenated over the systems (N rows in
 * all); sys int64[nsys][8] = {row0, n, kind, flv offset, forward levels, blv
 * offset, backward levels, ring}; order int32[nsys]; rp int32[N+1] and ci int32
 * (global columns), v64 float64 / v32 float32 values; jd32 float32[N] Jacobi
 * inverse diagonal; the ILU(0) sweeps as single-block plans (see
 * mpk_sweep_plan): prow, cnt int32[N], dep int32[4N] (position slots), coef
 * float32[4N], level starts flv / blv int32, dinv32 float32[N] (inverted
 * diagonal, in backward position order), f2b int32[N] (forward position ->
 * backward position of the same row); sys[7] = 1 marks a system whose plans
 * keep every read within 2048 positions of its level's end and whose dep
 * slots are then position mod 2048.  nmax: the largest n.  The caller keeps
 * the arrays alive.
 * mpk_batch_solve: x [D] from b [D], stats [D] float64[nsys][4] = {relres,
 * outer, inner, converged}; cfg [H] float64 {rtol, inner_rtol, max_outer,
 * max_inner, workgroup (128/256/512), grid, use_graph, variant}: variant an
 * index into mpk_batch_variants' list whose workgroup is cfg[4], or -1 for
 * the default tiles at that workgroup; with use_graph the launch is captured
 * once into a HIP graph and replayed (again when the variant, grid or
 * arguments change).  0 or -1.
 * mpk_batch_variants: the compiled launch variants, four int64 each into out
 * (at most cap of them): workgroup size, rows a thread takes per pass of the
 * product (a wave covers 64 times that), the product's entries loaded per
 * round, and the elements a thread takes per step of the Gram-Schmidt sums;
 * returns their count.  No counterpart in the reference (its solves are
 * serial CPU loops); the autotuner (mpk/batch.py) times them. */
int64_t mpk_batch_create(int64_t nsys, int64_t sys, int64_t order, int64_t rp, int64_t ci, int64_t v64, int64_t v32,
                         int64_t jd32, int64_t fprow, int64_t fcnt, int64_t fdep, int64_t fcoef, int64_t flv,
                         int64_t bprow, int64_t bcnt, int64_t bdep, int64_t bcoef, int64_t blv, int64_t dinv32,
                         int64_t f2b, int64_t nmax);
int64_t mpk_batch_solve(int64_t handle, int64_t b, int64_t x, int64_t stats, int64_t cfg);
void mpk_batch_free(int64_t handle);
int64_t mpk_batch_variants(int64_t out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* MPK_H */
