// Device kernels of libhymor_hip (gfx950, wave64).
//
// Layout recap (host side builds it, see convert.cpp):
//   - vectors are cell-major: row = cell*b + var for the b*ncell flow
//     unknowns, then the extra DOFs; NV doubles per row (1 real, 2 complex
//     (re,im), 2*nrhs for batches).
//   - A's flow-flow part is sliced ELL over cells: a slice is CPS = 64/b
//     cells, one wavefront per slice, lane = (cell q, block row r).  Slot t
//     of cell q is block slice_ptr[s] + t*CPS + q, so each slot of a slice is
//     one contiguous CPS*b*b-value run: lanes read their block row with
//     16-byte loads and the wavefront covers the run without gaps.
//   - everything else (shock bands, freestream coupling, imaginary parts) is
//     a short CSR over the listed rows, one wavefront per row.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hymor {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Logical tile of a workgroup such that each of the 8 XCDs (which receive
// workgroups round-robin by blockIdx) works on one contiguous range of
// tiles: neighbouring cells, and the x entries they gather, stay in one
// XCD's L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t nwg) {
    const int64_t per = nwg >> 3, rem = nwg & 7;
    const int64_t xcd = bid & 7, idx = bid >> 3;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

template <int N>
__device__ __forceinline__ void load_f64(const double *__restrict__ p, double *a) {
    static_assert(N % 2 == 0, "pairs");
#pragma unroll
    for (int k = 0; k < N; k += 2) {
        const d2 v = *reinterpret_cast<const d2 *>(p + k);
        a[k] = v.x;
        a[k + 1] = v.y;
    }
}

template <int N>
__device__ __forceinline__ void load_vals(const double *__restrict__ p, double *a) {
    load_f64<N>(p, a);
}

template <int N>
__device__ __forceinline__ void load_vals(const float *__restrict__ p, double *a) {
    static_assert(N % 4 == 0, "quads");
#pragma unroll
    for (int k = 0; k < N; k += 4) {
        const f4 v = *reinterpret_cast<const f4 *>(p + k);
        a[k] = (double)v.x;
        a[k + 1] = (double)v.y;
        a[k + 2] = (double)v.z;
        a[k + 3] = (double)v.w;
    }
}

// One Taylor term over the block part:  y = A*x (+ yx from the extra CSR),
// t = c1*y, tout = t (if tout), vout = vin + c2*t (if vout).
// c1/c2 select the recursion: normalised (c1 = dt/j, c2 = 1) or powers
// (c1 = dt, c2 = 1/j!).  With tout = y and vout = null this is a plain SpMV.
// vin may equal vout (in-place accumulation of the own row) or x (first term
// of a substep); x is never written by this kernel.
template <int B, typename VT, int NV>
__global__ __launch_bounds__(256) void k_term(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t nslice, int64_t ncell,
    const double *__restrict__ x, const double *__restrict__ yx,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t slice = xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6);
    if (slice >= nslice) return;
    const int q = lane / B, r = lane % B;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int32_t *cp = col + p0 + q;
    const VT *vp = val + (p0 + q) * BB + r * B;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll 3
    for (int t = 0; t < width; ++t) {
        const int64_t c = cp[(int64_t)t * CPS];
        double a[B];
        load_vals<B>(vp + (int64_t)t * CPS * BB, a);
        double xv[B * NV];
        load_f64<B * NV>(x + c * (B * NV), xv);
#pragma unroll
        for (int k = 0; k < B; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = fma(a[k], xv[k * NV + v], acc[v]);
    }
    const int64_t cell = slice * CPS + q;
    if (cell >= ncell) return;
    const int64_t row = cell * B + r;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double y = acc[v];
        if (yx) y += yx[row * NV + v];
        const double t = c1 * y;
        if (tout) tout[row * NV + v] = t;
        if (vout) vout[row * NV + v] = fma(c2, t, vin[row * NV + v]);
    }
}

// Extra CSR rows, one wavefront per listed row.  Flow rows (row < nflow)
// leave their partial sum in yx for k_term; extra-DOF rows finish the term
// here.  CPLX: values are complex (re,im) and NV counts doubles of complex
// vectors ((re,im) pairs); otherwise values are real and every one of the
// NV components is an independent right-hand side.
template <int NV, bool CPLX>
__global__ __launch_bounds__(256) void k_extra(
    const int64_t *__restrict__ rows, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val, int64_t nrows,
    int64_t nflow, const double *__restrict__ x, double *__restrict__ yx,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= nrows) return;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    const int64_t e1 = ptr[w + 1];
    for (int64_t e = ptr[w] + lane; e < e1; e += 64) {
        const int64_t c = col[e];
        if constexpr (CPLX) {
            const double ar = val[2 * e], ai = val[2 * e + 1];
#pragma unroll
            for (int v = 0; v < NV; v += 2) {
                const double xr = x[c * NV + v], xi = x[c * NV + v + 1];
                acc[v] = fma(ar, xr, fma(-ai, xi, acc[v]));
                acc[v + 1] = fma(ar, xi, fma(ai, xr, acc[v + 1]));
            }
        } else {
            const double a = val[e];
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = fma(a, x[c * NV + v], acc[v]);
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[v] += __shfl_xor(acc[v], off, 64);
    if (lane == 0) {
        const int64_t row = rows[w];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (row < nflow) {
                yx[row * NV + v] = acc[v];
            } else {
                const double t = c1 * acc[v];
                if (tout) tout[row * NV + v] = t;
                if (vout) vout[row * NV + v] = fma(c2, t, vin[row * NV + v]);
            }
        }
    }
}

// var-major (ABI) -> cell-major (internal).  x_vm holds nrhs columns of n
// elements with leading dimension ld (elements); an element is ncomp
// doubles.  Internal rows hold nrhs*ncomp doubles.  Var-major indices below
// nzero read as 0 (CONSTRUCT_C's x[1:N_set_0] .= 0).
// Internal rows are `width` doubles wide; components past nrhs*ncomp are
// padding and read as 0.
__global__ void k_perm_in(const double *__restrict__ x_vm, int64_t ld, double *__restrict__ out,
                          int64_t n, int64_t ncell, int b, int nrhs, int ncomp, int64_t nzero,
                          int width) {
    const int64_t valid = (int64_t)nrhs * ncomp;
    const int64_t total = n * width;
    const int64_t nflow = ncell * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / width;
        const int64_t rem = i - row * width;
        if (rem >= valid) {
            out[i] = 0.0;
            continue;
        }
        const int64_t rhs = rem / ncomp, comp = rem - rhs * ncomp;
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        out[i] = vm < nzero ? 0.0 : x_vm[(rhs * ld + vm) * ncomp + comp];
    }
}

__global__ void k_perm_out(const double *__restrict__ in, double *__restrict__ y_vm, int64_t ld,
                           int64_t n, int64_t ncell, int b, int nrhs, int ncomp, int64_t nzero,
                           int width) {
    const int64_t total = n * nrhs * ncomp;
    const int64_t nflow = ncell * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t rhs = i / (n * ncomp);
        const int64_t rem = i - rhs * n * ncomp;
        const int64_t vm = rem / ncomp, comp = rem - vm * ncomp;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        y_vm[(rhs * ld + vm) * ncomp + comp] =
            vm < nzero ? 0.0 : in[row * width + rhs * ncomp + comp];
    }
}

}  // namespace hymor

namespace hymor {

// Quadratic forms q_m = re(v)' W_m re(v) for nmat block-diagonal weights.
// Items 0..ncell-1 are cells (b x b block per cell and weight, row-major,
// layout W[m][cell][i][k]); items ncell.. are extra-DOF rows with a
// diagonal weight Wx[m][row - nflow].  Only the real part of v is used
// (component 0 of each NV-double row), as the reference forms real(v).
// Each workgroup writes one partial per weight; k_reduce_partials sums them
// in a fixed order, so results do not depend on scheduling.
template <int B, int NV, int NMAX>
__global__ __launch_bounds__(256) void k_quad_blockdiag(
    const double *__restrict__ v, const double *__restrict__ W, const double *__restrict__ Wx,
    int nmat, int64_t ncell, int64_t n_extra, double *__restrict__ partial) {
    __shared__ double red[NMAX][4];
    const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double q[NMAX];
#pragma unroll
    for (int m = 0; m < NMAX; ++m) q[m] = 0.0;
    if (item < ncell) {
        double vc[B];
#pragma unroll
        for (int k = 0; k < B; ++k) vc[k] = v[(item * B + k) * NV];
#pragma unroll
        for (int m = 0; m < NMAX; ++m) {
            if (m >= nmat) break;
            const double *w = W + ((int64_t)m * ncell + item) * (B * B);
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < B; ++i) {
                double wi[B];
                load_f64<B>(w + i * B, wi);
                double r = 0.0;
#pragma unroll
                for (int k = 0; k < B; ++k) r = fma(wi[k], vc[k], r);
                s = fma(vc[i], r, s);
            }
            q[m] = s;
        }
    } else if (item < ncell + n_extra) {
        const int64_t e = item - ncell;
        const double ve = v[(ncell * B + e) * NV];
#pragma unroll
        for (int m = 0; m < NMAX; ++m) {
            if (m >= nmat) break;
            q[m] = Wx[(int64_t)m * n_extra + e] * ve * ve;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < NMAX; ++m) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) q[m] += __shfl_xor(q[m], off, 64);
        if (lane == 0) red[m][wave] = q[m];
    }
    __syncthreads();
    if (threadIdx.x < nmat) {
        const int m = threadIdx.x;
        partial[(int64_t)m * gridDim.x + blockIdx.x] = (red[m][0] + red[m][1]) + (red[m][2] + red[m][3]);
    }
}

// out[m*ostride] = sum of partial[m][0..np) in a fixed order; one block per m.
__global__ __launch_bounds__(256) void k_reduce_partials(const double *__restrict__ partial,
                                                         int64_t np, double *__restrict__ out,
                                                         int64_t ostride) {
    __shared__ double red[4];
    const int m = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < np; i += blockDim.x) s += partial[(int64_t)m * np + i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[(int64_t)m * ostride] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace hymor

namespace hymor {

// Streaming copy, 16 bytes per lane per iteration: the reference point for
// achievable HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE
// counters (a copy of S bytes fetches S and writes S).
__global__ __launch_bounds__(256) void k_copy(const d2 *__restrict__ in, d2 *__restrict__ out,
                                              int64_t n2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

}  // namespace hymor

namespace hymor {

// Batched Taylor term: R right-hand sides per row (R doubles, rhs fastest).
// One wavefront per cell; lane = (row group g, rhs r) with r = lane % R and
// g = lane / R, each lane owning RPL = B*R/64 consecutive block rows.  Per
// block, a lane reads its RPL block rows (broadcast across the R rhs lanes)
// and, for each k, one double of X (the R lanes of a group read R
// consecutive doubles, one 128-byte line for R = 16); the block's column
// index is wave-uniform.  Same epilogue as k_term.
template <int B, typename VT, int R>
__global__ __launch_bounds__(256) void k_spmm(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    const double *__restrict__ yx, double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int G = 64 / R;
    static_assert(64 % R == 0 && B % G == 0, "unsupported B/R");
    constexpr int RPL = B / G;
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = __builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int r = lane % R, g = lane / R;
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    double acc[RPL];
#pragma unroll
    for (int i = 0; i < RPL; ++i) acc[i] = 0.0;
    for (int t = 0; t < width; ++t) {
        const int64_t blk = p0 + (int64_t)t * CPS + q;
        const int64_t c = col[blk];
        const VT *ap = val + blk * BB + (g * RPL) * B;
        double a[RPL * B];
        load_vals<RPL * B>(ap, a);
        const double *xp = x + c * (B * R) + r;
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const double xk = xp[k * R];
#pragma unroll
            for (int i = 0; i < RPL; ++i) acc[i] = fma(a[i * B + k], xk, acc[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int64_t idx = (cell * B + g * RPL + i) * R + r;
        double y = acc[i];
        if (yx) y += yx[idx];
        const double t = c1 * y;
        if (tout) tout[idx] = t;
        if (vout) vout[idx] = fma(c2, t, vin[idx]);
    }
}

}  // namespace hymor
