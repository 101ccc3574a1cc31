/*
 * hymor_oracle.c -- CPU restatement of HYMOR's exp(A*dt) Taylor action.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library
 * (hymor_amd/csrc).  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path never calls it.
 *
 * Every routine works on the reference's own operator format: a Julia
 * SparseMatrixCSC{Float64,Int64} (or ComplexF64) with 1-based colptr/rowval,
 * state vectors in variable-major order [rho; rho_u; rho_v; rho_E; shock].
 * Arithmetic follows the reference's evaluation order so that results agree
 * to the last few ulps:
 *   - SpMV: SparseArrays' generic mul!(y, A, x, 1.0, 0.0) for CSC is a column
 *     scatter: y .= 0; for col j, a = x[j]; for k in nzrange(A,j):
 *     y[rowval[k]] += nzval[k]*a  (Julia does not contract a*b+c into fma,
 *     so this file is built with -ffp-contract=off).
 *   - Complex products use (ac - bd, ad + bc), Julia's Complex{Float64} '*'.
 *
 * Parity pins (see tests/test_oracle_*.py): closed forms (diagonal, nilpotent,
 * zero, block-circulant via FFT), SciPy expm_multiply / dense polynomials,
 * dot-product adjoint identities, and brute force on tiny operators.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* SpMV, Julia SparseArrays column-scatter order (a15, third-party mul!).     */
/* ------------------------------------------------------------------------- */

/* y = A*x for real CSC, 1-based.  SparseArrays mul! with alpha=1, beta=0. */
void oc_csc_mul_f64(int64_t n, const int64_t *colptr, const int64_t *rowval,
                    const double *nzval, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double a = x[j] * 1.0;          /* alpha*x[j], alpha = 1.0 */
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k)
            y[rowval[k] - 1] += nzval[k] * a;
    }
}

/* y = A*x for complex CSC (interleaved re,im as Julia's ComplexF64). */
void oc_csc_mul_c128(int64_t n, const int64_t *colptr, const int64_t *rowval,
                     const double *nzval, const double *x, double *y)
{
    for (int64_t i = 0; i < 2 * n; ++i) y[i] = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double xr = x[2 * j], xi = x[2 * j + 1];
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
            const double ar = nzval[2 * k], ai = nzval[2 * k + 1];
            const int64_t r = rowval[k] - 1;
            const double pr = ar * xr - ai * xi;
            const double pi = ar * xi + ai * xr;
            y[2 * r] += pr;
            y[2 * r + 1] += pi;
        }
    }
}

/* Explicit (conjugate) transpose, the reference's sparse(A_dt')
 * (TRANSIENT_GROWTH_DOWNSTREAM.jl:75, FREESTREAM_RECEPTIVITY.jl:80).
 * Output CSC arrays must hold n+1 / nnz entries.  is_complex selects
 * ComplexF64 values; the imaginary part is negated ("'" is the adjoint). */
void oc_csc_adjoint(int64_t n, const int64_t *colptr, const int64_t *rowval,
                    const double *nzval, int is_complex,
                    int64_t *tcolptr, int64_t *trowval, double *tnzval)
{
    const int64_t nnz = colptr[n] - 1;
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < nnz; ++k) cnt[rowval[k]]++;   /* rowval is 1-based */
    tcolptr[0] = 1;
    for (int64_t i = 0; i < n; ++i) tcolptr[i + 1] = tcolptr[i] + cnt[i + 1];
    int64_t *pos = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) pos[i] = tcolptr[i] - 1;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t k = colptr[j] - 1; k < colptr[j + 1] - 1; ++k) {
            const int64_t r = rowval[k] - 1;
            const int64_t d = pos[r]++;
            trowval[d] = j + 1;
            if (is_complex) {
                tnzval[2 * d] = nzval[2 * k];
                tnzval[2 * d + 1] = -nzval[2 * k + 1];
            } else {
                tnzval[d] = nzval[k];
            }
        }
    }
    free(pos);
    free(cnt);
}

/* ------------------------------------------------------------------------- */
/* a1 / a12: POWER_METHOD!  (Eigenvalues/ARNOLDI_POWER_GPU.jl:64-82) and      */
/* LINEAR_INTEGRATION_GPU (Postprocessing/PLOT_MODES.jl:330-351), with the    */
/* Taylor order as a parameter (the reference hard-codes 5 in the first).     */
/* v is advanced in place.  vout, term are caller scratch of length n.        */
/* ------------------------------------------------------------------------- */
void oc_power_method_f64(int64_t n, const int64_t *colptr, const int64_t *rowval,
                         const double *nzval, int order, int64_t timesteps,
                         double *v, double *vout, double *term)
{
    memcpy(vout, v, (size_t)n * sizeof(double));                 /* :66 */
    for (int64_t i = 0; i < timesteps; ++i) {                     /* :70 */
        int64_t factorial_val = 1;
        for (int j = 1; j <= order; ++j) {
            oc_csc_mul_f64(n, colptr, rowval, nzval, v, term);    /* :73 */
            memcpy(v, term, (size_t)n * sizeof(double));          /* :74 */
            factorial_val = j * factorial_val;                     /* :75 */
            const double c = 1.0 / (double)factorial_val;
            for (int64_t q = 0; q < n; ++q) vout[q] = vout[q] + v[q] * c;  /* :76 */
        }
        memcpy(v, vout, (size_t)n * sizeof(double));              /* :78 */
    }
    memcpy(v, vout, (size_t)n * sizeof(double));                  /* :80 */
}

void oc_power_method_c128(int64_t n, const int64_t *colptr, const int64_t *rowval,
                          const double *nzval, int order, int64_t timesteps,
                          double *v, double *vout, double *term)
{
    const size_t bytes = (size_t)n * 2 * sizeof(double);
    memcpy(vout, v, bytes);
    for (int64_t i = 0; i < timesteps; ++i) {
        int64_t factorial_val = 1;
        for (int j = 1; j <= order; ++j) {
            oc_csc_mul_c128(n, colptr, rowval, nzval, v, term);
            memcpy(v, term, bytes);
            factorial_val = j * factorial_val;
            const double c = 1.0 / (double)factorial_val;
            for (int64_t q = 0; q < 2 * n; ++q) vout[q] = vout[q] + v[q] * c;
        }
        memcpy(v, vout, bytes);
    }
    memcpy(v, vout, bytes);
}

/* ------------------------------------------------------------------------- */
/* a4: EXPONENTIAL!  (Transient_growth_downstream/CONSTRUCT_C.jl:78-111)      */
/* v <- T_p(At)^n * C * T_p(A)^n * v with the normalised recursion            */
/* term = (A*term)/j.  At is the explicit adjoint CSC.  C is any CSC          */
/* (the Chu weight R'MR).  The reference fixes order = 5.                     */
/* ------------------------------------------------------------------------- */
static void taylor_normalised_f64(int64_t n, const int64_t *cp, const int64_t *rv,
                                  const double *nz, int order, int64_t timesteps,
                                  double *v, double *vout, double *term)
{
    for (int64_t i = 0; i < timesteps; ++i) {
        memcpy(term, vout, (size_t)n * sizeof(double));           /* :85 */
        for (int j = 1; j <= order; ++j) {
            oc_csc_mul_f64(n, cp, rv, nz, term, v);                /* :87 */
            memcpy(term, v, (size_t)n * sizeof(double));           /* :88 */
            const double c = 1.0 / (double)j;
            for (int64_t q = 0; q < n; ++q) term[q] = term[q] * c; /* :89 */
            for (int64_t q = 0; q < n; ++q) vout[q] = vout[q] + term[q]; /* :90 */
        }
    }
}

static void taylor_normalised_c128(int64_t n, const int64_t *cp, const int64_t *rv,
                                   const double *nz, int order, int64_t timesteps,
                                   double *v, double *vout, double *term)
{
    const size_t bytes = (size_t)n * 2 * sizeof(double);
    for (int64_t i = 0; i < timesteps; ++i) {
        memcpy(term, vout, bytes);
        for (int j = 1; j <= order; ++j) {
            oc_csc_mul_c128(n, cp, rv, nz, term, v);
            memcpy(term, v, bytes);
            const double c = 1.0 / (double)j;
            for (int64_t q = 0; q < 2 * n; ++q) term[q] = term[q] * c;
            for (int64_t q = 0; q < 2 * n; ++q) vout[q] = vout[q] + term[q];
        }
    }
}

void oc_exponential_f64(int64_t n,
                        const int64_t *a_cp, const int64_t *a_rv, const double *a_nz,
                        const int64_t *t_cp, const int64_t *t_rv, const double *t_nz,
                        const int64_t *c_cp, const int64_t *c_rv, const double *c_nz,
                        int order, int64_t timesteps,
                        double *v, double *vout, double *term)
{
    memcpy(vout, v, (size_t)n * sizeof(double));                        /* :83 */
    taylor_normalised_f64(n, a_cp, a_rv, a_nz, order, timesteps, v, vout, term);
    oc_csc_mul_f64(n, c_cp, c_rv, c_nz, vout, v);                        /* :95 */
    memcpy(vout, v, (size_t)n * sizeof(double));                        /* :96 */
    taylor_normalised_f64(n, t_cp, t_rv, t_nz, order, timesteps, v, vout, term);
    memcpy(v, vout, (size_t)n * sizeof(double));                        /* :109 */
}

void oc_exponential_c128(int64_t n,
                         const int64_t *a_cp, const int64_t *a_rv, const double *a_nz,
                         const int64_t *t_cp, const int64_t *t_rv, const double *t_nz,
                         const int64_t *c_cp, const int64_t *c_rv, const double *c_nz,
                         int c_is_complex, int order, int64_t timesteps,
                         double *v, double *vout, double *term)
{
    const size_t bytes = (size_t)n * 2 * sizeof(double);
    memcpy(vout, v, bytes);
    taylor_normalised_c128(n, a_cp, a_rv, a_nz, order, timesteps, v, vout, term);
    if (c_is_complex) {
        oc_csc_mul_c128(n, c_cp, c_rv, c_nz, vout, v);
    } else {
        /* real weight times complex vector: SparseArrays promotes the product
         * elementwise, a*x = (a*xr, a*xi) */
        for (int64_t i = 0; i < 2 * n; ++i) v[i] = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            const double xr = vout[2 * j], xi = vout[2 * j + 1];
            for (int64_t k = c_cp[j] - 1; k < c_cp[j + 1] - 1; ++k) {
                const int64_t r = c_rv[k] - 1;
                v[2 * r] += c_nz[k] * xr;
                v[2 * r + 1] += c_nz[k] * xi;
            }
        }
    }
    memcpy(vout, v, bytes);
    taylor_normalised_c128(n, t_cp, t_rv, t_nz, order, timesteps, v, vout, term);
    memcpy(v, vout, bytes);
}

/* ------------------------------------------------------------------------- */
/* a11: LINEAR_INTEGRATION_AND_ENERGY_GPU_NO_SAMPLES                           */
/* (Freestream_receptivity/LINEAR_INTEGRATION_AND_GAINS.jl:436-477).           */
/* Taylor stepping in the POWER_METHOD! form, then per step the four energies */
/* E_w[i+1] = real(v)' * (C_w * real(v)) for w in {all, p, k, S}.  Entry 0 of */
/* each series stays 0 as in the reference (the caller fills it, :168-173).  */
/* E must hold 4*(timesteps+1) doubles: E[w*(timesteps+1) + i].               */
/* The reference uses BLAS dot; here the sum runs in index order.             */
/* x is complex (interleaved) when is_complex, else real.                     */
/* ------------------------------------------------------------------------- */
void oc_linear_integration_energy(int64_t n, const int64_t *a_cp, const int64_t *a_rv,
                                  const double *a_nz, int is_complex, int order,
                                  int64_t timesteps,
                                  const int64_t *const *w_cp, const int64_t *const *w_rv,
                                  const double *const *w_nz, int n_weights,
                                  double *x, double *E)
{
    const int64_t comp = is_complex ? 2 : 1;
    double *v = x;
    double *vout = (double *)malloc((size_t)(n * comp) * sizeof(double));
    double *term = (double *)malloc((size_t)(n * comp) * sizeof(double));
    double *vr = (double *)malloc((size_t)n * sizeof(double));
    double *cv = (double *)malloc((size_t)n * sizeof(double));
    const size_t bytes = (size_t)(n * comp) * sizeof(double);
    for (int w = 0; w < n_weights; ++w)
        for (int64_t i = 0; i <= timesteps; ++i) E[w * (timesteps + 1) + i] = 0.0;
    memcpy(vout, v, bytes);                                           /* :448 */
    for (int64_t i = 0; i < timesteps; ++i) {                          /* :451 */
        int64_t factorial_val = 1;
        for (int j = 1; j <= order; ++j) {
            if (is_complex) oc_csc_mul_c128(n, a_cp, a_rv, a_nz, v, term);
            else oc_csc_mul_f64(n, a_cp, a_rv, a_nz, v, term);        /* :454 */
            memcpy(v, term, bytes);
            factorial_val = j * factorial_val;
            const double c = 1.0 / (double)factorial_val;
            for (int64_t q = 0; q < n * comp; ++q) vout[q] = vout[q] + v[q] * c;
        }
        memcpy(v, vout, bytes);                                        /* :459 */
        for (int64_t q = 0; q < n; ++q) vr[q] = v[q * comp];           /* :464-468 */
        for (int w = 0; w < n_weights; ++w) {                          /* :470-473 */
            oc_csc_mul_f64(n, w_cp[w], w_rv[w], w_nz[w], vr, cv);
            double s = 0.0;
            for (int64_t q = 0; q < n; ++q) s += vr[q] * cv[q];
            E[w * (timesteps + 1) + i + 1] = s;
        }
    }
    free(vout); free(term); free(vr); free(cv);
}

/* ------------------------------------------------------------------------- */
/* CPU baseline, row-parallel: the same Taylor action over a CSR copy of the  */
/* operator (row-gather SpMV, one pass fused with the Taylor update), OpenMP  */
/* over rows.  Not a statement of the reference; it is the strongest plain    */
/* CPU form of the same arithmetic, used only as a timing baseline.           */
/* CSR arrays are 0-based int64.  Returns the thread count used.              */
/* ------------------------------------------------------------------------- */
int oc_taylor_csr_omp(int64_t n, const int64_t *rowptr, const int64_t *colind,
                      const double *val, int order, int64_t timesteps,
                      double *v, double *w0, double *w1, double *vout, int nthreads)
{
    int used = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
    }
#endif
    double *src = w0, *dst = w1;
    for (int64_t i = 0; i < timesteps; ++i) {
        const double *x = v;
        for (int j = 1; j <= order; ++j) {
            const double c = 1.0 / (double)j;
            double *out = dst;
#pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < n; ++r) {
                double s = 0.0;
                for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += val[k] * x[colind[k]];
                s *= c;
                out[r] = s;
                vout[r] = (j == 1 ? v[r] : vout[r]) + s;
            }
            x = out;
            double *t = src; src = dst; dst = t;
            (void)t;
        }
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; ++r) v[r] = vout[r];
    }
    return used;
}
