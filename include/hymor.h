/*
 * hymor.h -- C ABI of libhymor_hip.so, the MI355X (gfx950) engine for HYMOR's
 * linear-stability hot path: the Taylor action w = T_p(A*dt)^n * v of the
 * linearised compressible Navier-Stokes Jacobian, its adjoint, the Chu-norm
 * sandwich and the energy-tracked integration.
 *
 * Plain C types only (pointers, sizes, scalars): callable from Julia
 * (ccall), Python (ctypes) or C.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative HYMOR_ERR_* code on
 *     failure; hymor_last_error() then holds a message (thread-local).
 *   - Vectors cross the ABI in the reference's variable-major order
 *     [rho(all cells); rho_u; rho_v; rho_E; extra DOFs], length n.  Complex
 *     vectors are interleaved (re, im) pairs, the memory layout of Julia's
 *     Vector{ComplexF64}.  Batches are column-major n x nrhs with a leading
 *     dimension, the layout of a Julia Matrix.
 *   - The library owns all device memory of an operator handle and one HIP
 *     stream per handle.  Host-pointer entry points return after the result
 *     is in host memory.  *_dev entry points take device pointers, enqueue on
 *     the handle's stream and return at once; hymor_op_synchronize() waits
 *     and reports asynchronous errors (the reference's CUDA.synchronize(),
 *     utils/Stability_analysis/Eigenvalues/EIGENVALUES.jl:174).
 *   - A handle is used by one host thread at a time (ARPACK calls its matvec
 *     sequentially, EIGENVALUES.jl:153-159).
 */
#ifndef HYMOR_H
#define HYMOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HYMOR_OK 0
#define HYMOR_ERR_ARG (-1)     /* invalid argument or shape */
#define HYMOR_ERR_HIP (-2)     /* HIP runtime error */
#define HYMOR_ERR_NOMEM (-3)   /* host or device allocation failed */
#define HYMOR_ERR_NODEV (-4)   /* no usable GPU */

/* element type of vectors and operator values */
#define HYMOR_F64 0            /* Float64 */
#define HYMOR_C128 1           /* ComplexF64, interleaved (re, im) */

/* storage precision of the block values; accumulation is always FP64 */
#define HYMOR_PREC_F64 0
#define HYMOR_PREC_F32 1

/* Taylor recursion form */
#define HYMOR_FORM_NORMALISED 0 /* term = (A*term)/j; vout += term
                                   (CONSTRUCT_C.jl:84-91, EXPONENTIAL!) */
#define HYMOR_FORM_POWERS 1     /* v = A*v; vout += v/j!
                                   (ARNOLDI_POWER_GPU.jl:70-79, POWER_METHOD!) */
#define HYMOR_FORM_HORNER 2     /* the same polynomial by Horner's rule,
                                   w = v + (A*dt/j) w for j = order..1: three
                                   vector passes per term instead of four (no
                                   reference counterpart) */

typedef struct hymor_op hymor_op;
typedef struct hymor_pattern hymor_pattern;

/* 1 if a GPU the library can use is present, else 0.  Replaces the
 * reference's `use_gpu = CUDA.functional()` (EIGENVALUES.jl:119). */
int hymor_device_available(void);

/* Message for the last failing call on this thread ("" if none). */
const char *hymor_last_error(void);

/* Build a device operator from a CSC matrix in the reference's layout.
 * Replaces `CuSparseMatrixCSC(A * dt)` (EIGENVALUES.jl:139,
 * TRANSIENT_GROWTH_DOWNSTREAM.jl:78-80, FREESTREAM_RECEPTIVITY.jl:83-85).
 *   n, colptr[n+1], rowval[nnz], nzval[nnz] : SparseMatrixCSC fields;
 *       index_base 1 for Julia arrays as-is, 0 for 0-based arrays.
 *   dtype        : HYMOR_F64 (nzval double) or HYMOR_C128 (interleaved).
 *   scale        : values are multiplied by it at upload (the reference's
 *                  A*dt); pass 1.0 and use the per-call dt otherwise.
 *   vars_per_cell: b (4 in the reference); nchi*neta cells; unknowns past
 *                  b*nchi*neta (shock, freestream DOFs) are "extra".
 *   value_prec   : HYMOR_PREC_F64 or HYMOR_PREC_F32 storage of block values.
 * The transpose (adjoint) operator is built at the same time, replacing
 * `sparse(A_dt')` (TRANSIENT_GROWTH_DOWNSTREAM.jl:75). */
int hymor_op_create_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int index_base, int dtype, double scale,
                        int vars_per_cell, int64_t nchi, int64_t neta,
                        int value_prec, hymor_op **out);

/* Build the extended receptivity operator
 *     L_ = [L  repeat(B,1,N);  0  kron(diag(-i*w), I_p)]
 * of LINEARIZE_L_.jl:58-67 -- what `CuSparseMatrixCSC(A_ * dt)` uploads in
 * FREESTREAM_RECEPTIVITY.jl:83-85 -- from ONE per-wavenumber system and the
 * wavenumbers, with B kept once on the device at every vector width: the
 * forward operator applies B to the column sums of the N freestream blocks,
 * the adjoint evaluates B' once for the N blocks (the same arithmetic the
 * library uses for an assembled L_ uploaded by hymor_op_create_csc).
 *   n_sys, colptr, rowval, nzval, index_base, dtype: the CSC of
 *       S = [L B; 0 0] (n_sys = n_l + p, reference var-major order, flow
 *       DOFs then shock DOFs, then the p freestream DOFs of one wavenumber);
 *       rows n_l.. must be empty.  Real (HYMOR_F64) or complex values.
 *   n_l: flow plus shock DOFs; nwave = N, w[N]: the wavenumbers.
 *   scale multiplies every value and the -i*w diagonal (the reference's
 *   A_*dt).  The handle is complex and of size n_l + N p; it serves every
 *   entry point that takes an operator, except the per-column diagonal of
 *   hymor_expm_taylor_batched_coldiag_dev. */
int hymor_op_create_receptivity(int64_t n_sys, const int64_t *colptr, const int64_t *rowval,
                                const void *nzval, int index_base, int dtype, double scale,
                                int vars_per_cell, int64_t nchi, int64_t neta, int64_t n_l,
                                int64_t nwave, const double *w, int value_prec, hymor_op **out);

/* Build a device operator from cell-major block-CSR arrays (0-based), the
 * library's native layout: brow_ptr[ncell+1], bcol[nnzb] (int32),
 * bval[nnzb*b*b] (row-major blocks, real), plus nx extra entries in
 * cell-major global indices (xval real or complex per dtype).  No reference
 * counterpart: a faster upload for callers that already hold blocks.
 * with_adjoint = 0 skips building the transpose. */
int hymor_op_create_bsr(int64_t nchi, int64_t neta, int vars_per_cell, int64_t n_extra,
                        const int64_t *brow_ptr, const int32_t *bcol, const double *bval,
                        int64_t nx, const int64_t *xrow, const int64_t *xcol,
                        const void *xval, int dtype, double scale, int value_prec,
                        int with_adjoint, hymor_op **out);

/* Release all device memory of a handle (CUDA.unsafe_free! + reclaim,
 * EIGENVALUES.jl:175-180). */
int hymor_op_destroy(hymor_op *op);

/* Sizes of a handle: out[0]=n, [1]=b, [2]=ncell, [3]=nnzb, [4]=padded nnzb,
 * [5]=extra nnz, [6]=extra rows, [7]=is_complex, [8]=value_prec,
 * [9]=device bytes held.  nout entries are written (at most 10). */
int hymor_op_info(const hymor_op *op, int64_t *out, int nout);

/* Switches the block values the kernels read to value_prec in place.  An
 * operator uploaded with HYMOR_PREC_F64 keeps its FP64 values and casts an
 * FP32 copy on the device on its first switch to HYMOR_PREC_F32 (both copies
 * then stay resident); one uploaded with HYMOR_PREC_F32 cannot switch to
 * FP64 (error).  Drops the captured Taylor graph.  No counterpart in the
 * reference, which builds one CuSparseMatrixCSC per precision
 * (EIGENVALUES.jl:139); it lets a mixed-precision eigensolve pay one upload. */
int hymor_op_set_value_prec(hymor_op *op, int value_prec);

/* y = A*x (transpose != 0: y = A'*x, the adjoint).  The reference's
 * mul!(y, A, x, 1.0, 0.0) on CuSparseMatrixCSC (ARNOLDI_POWER_GPU.jl:73). */
int hymor_spmv(hymor_op *op, int transpose, const void *x, void *y, int dtype);

/* Device-pointer form of hymor_spmv (asynchronous). */
int hymor_spmv_dev(hymor_op *op, int transpose, const void *d_x, void *d_y, int dtype);

/* y = T_order(A*dt)^nsub * x, where T_p is the degree-p Taylor polynomial
 * of exp.  order=5, form=HYMOR_FORM_POWERS is POWER_METHOD!
 * (ARNOLDI_POWER_GPU.jl:64-82) behind ARNOLDI_POWER_GPU (:36-52);
 * transpose != 0 applies the adjoint operator.  x and y are host vectors. */
int hymor_expm_taylor(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                      int form, const void *x, void *y, int dtype);

/* Device-pointer form of hymor_expm_taylor (asynchronous). */
int hymor_expm_taylor_dev(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                          int form, const void *d_x, void *d_y, int dtype);

/* hymor_expm_taylor_dev that allocates nothing: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The sweep works in a
 * library-owned static device array of 64 MiB per GPU (held until the
 * process ends) instead of the handle's workspace, and never captures or
 * replays a graph of its own, so the call may itself be captured.  Limit:
 * n * (1 for HYMOR_F64, 2 for HYMOR_C128) <= 1 677 696; larger inputs fail
 * with HYMOR_ERR_ARG.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (different handles on different streams, or threads)
 * race and must be serialised by the caller.  Asynchronous, on the handle's
 * stream. */
int hymor_expm_taylor_dev_noalloc(hymor_op *op, int transpose, int order, int64_t nsub,
                                  double dt, int form, const void *d_x, void *d_y, int dtype);

/* Batched action on nrhs vectors at once: X, Y are column-major n x nrhs
 * (leading dimensions ldx, ldy >= n), each column one vector of the
 * unbatched call.  The operator's blocks are read once per Taylor term for
 * all columns (an SpMM).  nrhs*(1 or 2 for complex) <= 64; internally the
 * width is padded to a power of two.  Used for many Krylov vectors or
 * streamwise wavenumbers at once (no reference counterpart: the reference
 * applies one vector per ARPACK callback, ARNOLDI_POWER_GPU.jl:36-52). */
int hymor_expm_taylor_batched(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                              int form, int64_t nrhs, const void *X, int64_t ldx, void *Y,
                              int64_t ldy, int dtype);

int hymor_expm_taylor_batched_dev(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                                  int form, int64_t nrhs, const void *d_X, int64_t ldx, void *d_Y,
                                  int64_t ldy, int dtype);

/* Factored receptivity action (no reference counterpart in this form: the
 * reference applies the extended operator [L repeat(B,1,N); 0
 * kron(diag(-i*w), I_p)] of LINEARIZE_L_.jl:58-67 to one long vector).  op
 * is one per-wavenumber system [L B; 0 0] (rows [row0, row1) are the p
 * freestream DOFs); column c of X (complex, nrhs <= 32) is evolved under
 * [L B; 0 diag[c] I]: diag holds one complex value per column (-i*w_c), and
 * its conjugate is used for transpose.  By linearity the N columns together
 * carry exactly the extended-operator action.  op may hold real values
 * (HYMOR_F64 upload) or complex ones; the vectors are complex either way.
 * Synchronous. */
int hymor_expm_taylor_batched_coldiag_dev(hymor_op *op, int transpose, int order, int64_t nsub,
                                          double dt, int form, int64_t nrhs, const double *diag,
                                          int64_t row0, int64_t row1, const void *d_X,
                                          int64_t ldx, void *d_Y, int64_t ldy);

/* Gradient of J = <wbar, T_order(dt*A)^nsub x> with respect to the
 * operator's entries -- the pullback a ChainRules rrule needs for the
 * operator argument, e.g. the sensitivity of a transient-growth gain to the
 * Jacobian (no reference counterpart: the reference takes no derivatives;
 * the gain's gradient with respect to the vector needs only
 * hymor_chu_sandwich).  Real operators.  x, wbar: host vectors (var-major,
 * length n).  The gradient is returned for every entry of the pattern
 * colptr[n+1], rowval (CSC, index_base 0 or 1: the upload's own arrays give
 * it on A's pattern; entries outside it are allowed): grad[e] = dJ/dA_ij for
 * entry e = (rowval[e], its column), A the operator as stored (the uploaded
 * values times `scale`).  Checkpointed: the forward sweep keeps the state
 * every ckpt_every substeps and the backward sweep recomputes one group of
 * ckpt_every states at a time from its checkpoint, so beside the handle's
 * workspace the call holds ceil(nsub/ckpt_every) + (2*order + 1)*ckpt_every
 * + order vectors of n doubles and 16 bytes per pattern entry.  ckpt_every 1
 * keeps every state and recomputes none (nsub + (2*order + 1)*16 + order
 * vectors).  The entries' sums are grouped per segment of states, so
 * gradients for different ckpt_every agree to rounding, not bit for bit.
 * Synchronous. */
int hymor_expm_taylor_vjp_values(hymor_op *op, int order, int64_t nsub, double dt,
                                 int64_t ckpt_every, const double *x, const double *wbar,
                                 const int64_t *colptr, const int64_t *rowval, int index_base,
                                 double *grad);

/* A gradient pattern kept on the device: the CSC pattern (colptr, rowval,
 * index_base as above) of an operator of op's shape, mapped to the device
 * layout once (8 bytes per entry; for block sizes 4 and 8 also grouped by
 * block for the sampled product: 8 + 4*b*b bytes per block of at least b
 * entries, 12 per other entry), for callers that differentiate on the
 * same pattern repeatedly (an optimisation loop over the Jacobian).  No
 * reference counterpart. */
int hymor_pattern_create(const hymor_op *op, const int64_t *colptr, const int64_t *rowval,
                         int index_base, hymor_pattern **out);
int hymor_pattern_destroy(hymor_pattern *pat);
int hymor_pattern_nnz(const hymor_pattern *pat, int64_t *nnz);

/* hymor_expm_taylor_vjp_values on a device-resident pattern with device
 * vectors: d_x, d_wbar (var-major, n doubles), d_grad (the pattern's nnz
 * doubles, in its CSC order) -- no pattern, vector or gradient crosses to
 * the host.  Synchronous. */
int hymor_expm_taylor_vjp_values_dev(hymor_op *op, const hymor_pattern *pat, int order,
                                     int64_t nsub, double dt, int64_t ckpt_every,
                                     const double *d_x, const double *d_wbar, double *d_grad);

/* Chu-norm sandwich y = T_p(A'dt)^n * C * T_p(A dt)^n * x with the zeroing
 * of CONSTRUCT_C (CONSTRUCT_C.jl:33-58 around EXPONENTIAL! :78-111):
 * x[0:n_set0) is zeroed before and y[0:n_set0) after (var-major indices).
 * C is a handle built from R'*M*R (TRANSIENT_GROWTH_DOWNSTREAM.jl:51).
 * EXPONENTIAL!'s polynomial is evaluated by Horner's rule (equal to
 * rounding; three vector passes per term instead of four). */
int hymor_chu_sandwich(hymor_op *A, hymor_op *C, int order, int64_t nsub, double dt,
                       int64_t n_set0, const void *x, void *y, int dtype);

int hymor_chu_sandwich_dev(hymor_op *A, hymor_op *C, int order, int64_t nsub, double dt,
                           int64_t n_set0, const void *d_x, void *d_y, int dtype);

/* The sandwich on nrhs device columns at once (one batched sweep each way),
 * optionally between two operators of A's grid:
 *   Y = P_out * Z * T_p(A'dt)^n * C * T_p(A dt)^n * Z * P_in * X,
 * Z the zeroing of var-major rows [0, n_set0); P_in / P_out NULL for the
 * identity.  With P_in = L^-H and P_out = L^-1 for the Cholesky factor L of
 * the right-hand matrix (R'M_hat R, TRANSIENT_GROWTH_DOWNSTREAM.jl:52), the
 * generalized problem the reference hands to eigsh (:63-110, and
 * FREESTREAM_RECEPTIVITY.jl:88-112) becomes a standard symmetric one whose
 * blocks a block Krylov method can expand.  nrhs * (2 if complex) <= 64;
 * C and the factors must use A's batch row layout (the default).
 * Asynchronous on A's stream. */
int hymor_chu_sandwich_batched_dev(hymor_op *A, hymor_op *C, hymor_op *P_in, hymor_op *P_out,
                                   int order, int64_t nsub, double dt, int64_t n_set0,
                                   int64_t nrhs, const void *d_X, int64_t ldx, void *d_Y,
                                   int64_t ldy, int dtype);

/* Wait for all work on the handle's stream; reports asynchronous errors. */
int hymor_op_synchronize(hymor_op *op);

/* Device memory helpers for the *_dev entry points. */
int hymor_dev_alloc(int64_t bytes, void **d_ptr);
int hymor_dev_free(void *d_ptr);
int hymor_memcpy_h2d(void *d_dst, const void *h_src, int64_t bytes);
int hymor_memcpy_d2h(void *h_dst, const void *d_src, int64_t bytes);
int hymor_memcpy_d2d(void *d_dst, const void *d_src, int64_t bytes);
/* len doubles at d_x set to pseudo-random values, uniform in [-1, 1), a
 * function of (seed, index) only (splitmix64): Krylov start blocks made on
 * the device.  Synchronous. */
int hymor_dev_fill_uniform(void *d_x, int64_t len, uint64_t seed);

/* Timing probe: launches the fused Taylor-term kernel `reps` times on the
 * handle's stream (a middle term of `form`: powers / normalised gather x,
 * write the new term and update vout; Horner gathers w, reads v and writes
 * the new w) between two HIP events; returns the mean duration per launch
 * in ms and the kernel's algorithmic bytes per launch.  nv = doubles per
 * row (1 real, 2 complex, up to 64 for batches; a power of two).  A term
 * that a sweep runs as one launch with an assembled receptivity operator's
 * replicated rows fused in is timed as such (the forward Horner term after
 * one pass forming the column sums of its x), its bytes counting the
 * replicated rows' compressed entries. */
int hymor_bench_term_kernel(hymor_op *op, int transpose, int nv, int form, int reps,
                            double *ms_per_launch, double *bytes_per_launch);

/* Autotuner, no reference counterpart: times the launch variants of the
 * single-vector Taylor-term kernel (one wave per slice with 2 or 4 waves per
 * group; split-slot with 2 or 4 waves per slice; grouped loads with 3 or 9
 * waves per slice; at nv >= 2 also grouped loads with the x blocks staged
 * through LDS; plain or non-temporal block loads) on this operator, `reps`
 * launches each, and keeps the fastest for both directions.  nv: 1, 2 or 4
 * doubles per row.  The timed terms are the block kernel alone: on an
 * assembled receptivity operator whose replicated rows the default tuning
 * fuses into the term kernel (one launch per term), a pick other than the
 * default runs them as launches of their own. */
int hymor_op_autotune(hymor_op *op, int nv, int reps, double *best_ms);

/* enable != 0: the Taylor sweep of hymor_expm_taylor(_dev) and the batched
 * forms is captured into a HIP graph on first use and replayed while
 * (transpose, order, nsub, dt, form, width) repeat -- one graph launch
 * instead of order*nsub kernel launches.  Off by default.  While a
 * rocprofiler tool is configured for the process (ROCP_TOOL_LIBRARIES set,
 * as rocprofv3 does) the sweep is launched kernel by kernel instead: the
 * tool faulted the host inside hipGraphLaunch of such sweeps. */
int hymor_op_set_graphs(hymor_op *op, int enable);

/* Kernel for batches of 16+ doubles per row: 0 vector FMA (k_spmm),
 * 1 FP64 matrix cores (k_spmm_mfma), 2 / 3 matrix cores with two / one
 * cells per wave and prefetched epilogue operands (k_spmm_mfma_mc),
 * 4 / 5 matrix cores with all operand loads of a large / small group of
 * slots issued together (k_spmm_mfma_w), 6 (the default) matrix cores with
 * the batch held in the operand ("tile") order, 16-byte accesses per lane:
 * at b = 8 two cells per 16x16x4 product, rows 8-15 carrying the second
 * cell of a pair (k_spmm_pair; 8 doubles per row: k_spmm_tile), at b = 4
 * the 4x4x4 FP64 matrix cores with four cells per wave, one cell per
 * product block (k_spmm_quad4c), 7 the 4x4x4 FP64 matrix cores with
 * no idle rows: at b = 8 the tile order (k_spmm_quad), at b = 4 the
 * row-major batch and the single-vector value order (k_spmm_quad4, one
 * cell per wave), 8 one cell per wave on the 16x16x4 products (k_spmm_tile;
 * the same as 6 at b = 4).  All accumulate in FP64. */
int hymor_op_set_batch_kernel(hymor_op *op, int use_mfma);

/* Times the nine batch kernels at width nv (8, 16, 32 or 64) and keeps the
 * fastest; returns its id and ms per launch.  As hymor_op_autotune, the
 * timed terms leave out a receptivity operator's fused replicated rows,
 * which only batch kernels 6 and 8 at b = 4 take into their launch. */
int hymor_op_autotune_batch(hymor_op *op, int nv, int reps, int *best_kernel, double *best_ms);

/* Read / force the chosen variant (split 0 k_term | 1 k_term_split |
 * 2 k_term_split_p | 3 k_term_lds (k_term_split_p at one double per row),
 * waves, nt 0|1). */
int hymor_op_get_tuning(const hymor_op *op, int *split, int *waves, int *nt);
int hymor_op_set_tuning(hymor_op *op, int split, int waves, int nt);

/* Calibration probe, no reference counterpart: a streaming copy of `bytes`
 * (multiple of 16) between two fresh device buffers, launched `reps` times;
 * returns ms per launch.  Used to measure achievable HBM bandwidth and to
 * scale the FETCH_SIZE / WRITE_SIZE profiler counters. */
int hymor_bench_copy(int64_t bytes, int reps, double *ms_per_launch);

/* Timing probe, no reference counterpart: a streaming read of `bytes`
 * (non-temporal 16-byte loads, one partial sum per thread), the read-only
 * bandwidth ceiling for the value stream of the term kernels. */
int hymor_bench_read(int64_t bytes, int reps, double *ms_per_launch);

/* Timing probe, no reference counterpart: streaming reads and writes in the
 * proportion read_bytes : write_bytes (16 384 chunks, each read then
 * written by one workgroup; read_bytes >= 16 MiB), the ceiling for a
 * kernel whose HBM traffic has that mix.  Returns ms per launch and the
 * bytes one launch moves (the requested counts rounded down to whole
 * 1 KiB runs per chunk). */
int hymor_bench_stream_mix(int64_t read_bytes, int64_t write_bytes, int reps,
                           double *ms_per_launch, double *bytes_moved);

/* ------------------------------------------------------------------------
 * Energy-tracked integration (Freestream_receptivity/
 * LINEAR_INTEGRATION_AND_GAINS.jl:375-477, LINEAR_INTEGRATION_AND_ENERGY_GPU
 * and _NO_SAMPLES).  The Chu weights C_w = R'M_wR are block-diagonal per
 * cell; the library keeps them as b x b blocks (plus a diagonal on extra
 * DOFs) and forms the energies on the device, so only scalars (and the
 * optional snapshots) cross to the host.
 * ------------------------------------------------------------------------ */
typedef struct hymor_weights hymor_weights;

/* nmat (1..8) weight matrices given as CSC arrays in the reference layout
 * (C_no_exp_all, _p, _k, _S at LINEAR_INTEGRATION_AND_GAINS.jl:63-111).
 * Fails if an entry lies off the per-cell block diagonal, or off the
 * diagonal on extra DOFs. */
int hymor_weights_create_csc(int64_t n, int nmat, const int64_t *const *colptr,
                             const int64_t *const *rowval, const double *const *nzval,
                             int index_base, int vars_per_cell, int64_t nchi, int64_t neta,
                             hymor_weights **out);

int hymor_weights_destroy(hymor_weights *W);

/* n_t Taylor steps of T_order(A*dt) from x; after each step i (and for the
 * start, i = 0) E[m*(n_t+1) + i] = real(v)' * C_m * real(v).  x_final
 * (optional) receives the last state.  If n_samples > 0, the state after
 * every step i with i % n_skip == 0 is copied to samples[count*n ...] for
 * the first n_samples such steps (the reference's Vf snapshots, :402-409).
 * The reference leaves E[:,1] = 0 and fills it on the host (:168-173);
 * here entry 0 holds that same initial energy.  Concurrent calls sharing
 * one hymor_weights handle race on its reduction buffer. */
int hymor_energy_track(hymor_op *A, const hymor_weights *W, int order, int64_t n_t, double dt,
                       int form, const void *x, int dtype, double *E, void *x_final,
                       int64_t n_skip, int64_t n_samples, void *samples);

/* ------------------------------------------------------------------------
 * Device-resident Krylov basis.  Replaces the host-side basis ARPACK keeps
 * (Arpack.eigs at EIGENVALUES.jl:159, TRANSIENT_GROWTH_DOWNSTREAM.jl:103,
 * FREESTREAM_RECEPTIVITY.jl:107), whose every matvec crosses PCIe twice
 * (ARNOLDI_POWER_GPU.jl:42,44).  The basis V (n x (m_max+1), column-major,
 * var-major vectors) stays in HBM; the host keeps only the small Hessenberg
 * or tridiagonal matrix.  Columns are device pointers usable with the *_dev
 * actions.  All calls are synchronous.
 * ------------------------------------------------------------------------ */
typedef struct hymor_krylov hymor_krylov;

int hymor_krylov_create(int64_t n, int m_max, int dtype, hymor_krylov **out);
int hymor_krylov_destroy(hymor_krylov *K);

/* Device address of column j (0 <= j <= m_max). */
int hymor_krylov_column(hymor_krylov *K, int j, void **d_col);

/* Upload / download columns (host vectors, var-major). */
int hymor_krylov_set_column(hymor_krylov *K, int j, const void *x);
int hymor_krylov_get_columns(hymor_krylov *K, int j0, int ncols, void *out);

/* Classical Gram-Schmidt with one reorthogonalisation (CGS2): w (device,
 * length n) is orthogonalised against columns 0..j-1, normalised and stored
 * as column j.  h_out receives the j projection coefficients (complex:
 * interleaved) followed by the norm, i.e. column j-1 of the Hessenberg
 * matrix when w = Op(V[:, j-1]).  A vector whose norm after projection is
 * at most 64 * DBL_EPSILON times its norm before (numerically in the span
 * of the basis) is dependent: column j is set to zero and the norm is
 * reported as 0. */
int hymor_krylov_orthonormalize(hymor_krylov *K, int j, void *d_w, double *h_out);

/* Block form of hymor_krylov_orthonormalize for block Krylov methods (no
 * single counterpart in the reference, whose ARPACK expands one vector at a
 * time): the s <= 16 device columns W (leading dimension ldw) are
 * orthogonalised against basis columns 0..j-1 by block classical
 * Gram-Schmidt done twice, each projection one pass over the basis for all
 * s vectors (V^T W on the FP64 matrix cores), then orthonormalised among
 * themselves in order into columns j..j+s-1.  h_out receives the
 * (j+s) x s coefficient matrix, column-major: W = V[:, :j+s] h_out; a zero
 * diagonal entry marks a dependent vector (the same relative test as
 * hymor_krylov_orthonormalize; its column is zeroed).  W is overwritten.
 * Real bases only. */
int hymor_krylov_orthonormalize_block(hymor_krylov *K, int j, int s, void *d_W, int64_t ldw,
                                      double *h_out);

/* Restart: V[:, 0:k] <- V[:, 0:m] * Q with Q column-major m x k (host,
 * complex interleaved for complex bases). */
int hymor_krylov_combine(hymor_krylov *K, int m, int k, const double *Q);

/* Ritz vectors: out (host, n x k column-major) = V[:, 0:m] * Q; V is left
 * unchanged. */
int hymor_krylov_ritz(hymor_krylov *K, int m, int k, const double *Q, void *out);

int hymor_krylov_copy_column(hymor_krylov *K, int src, int dst);

/* Timing probe, no reference counterpart: the three kernels of one block
 * Gram-Schmidt projection (hymor_krylov_orthonormalize_block) on a basis V
 * of j columns and a block W of s <= 16 columns, length n, filled on the
 * device; each kernel launched `reps` times between HIP events.
 *   out[0] ms per launch of the V^T W partial products (FP64 matrix cores),
 *   out[1] ms of their reduction, out[2] ms of W -= V h (FP64 matrix cores),
 *   out[3] V^T W algorithmic bytes ((j + s) n 8), out[4] its flops
 *   (2 n j s), out[5] the matrix-core flops either product kernel issues
 *   (512 n ceil(j / 16): the products always span 16 basis and 16 block
 *   columns), out[6] W -= V h algorithmic bytes ((j + 2 s) n 8), out[7] its
 *   flops (2 n j s). */
int hymor_bench_krylov_projection(int64_t n, int j, int s, int reps, double *out);

#ifdef __cplusplus
}
#endif

#endif /* HYMOR_H */
