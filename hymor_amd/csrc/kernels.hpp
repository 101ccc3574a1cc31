// Device kernels of libhymor_hip (gfx950, wave64).
//
// Layout recap (host side builds it, see convert.cpp):
//   - vectors are cell-major: row = cell*b + var for the b*ncell flow
//     unknowns, then the extra DOFs; NV doubles per row (1 real, 2 complex
//     (re,im), 2*nrhs for batches).
//   - A's flow-flow part is sliced ELL over cells: a slice is CPS = 64/b
//     cells, one wavefront per slice, lane = (cell q, block row r).  Slot t
//     of a slice is one contiguous CPS*b*b-value run starting at block
//     slice_ptr[s] + t*CPS; inside it value k of lane (q, r) sits at
//     ((k / VPL) * 64 + lane) * VPL + k % VPL, VPL = 16 bytes / value size,
//     so every wave-wide 16-byte load covers one contiguous 1 KB run.
//   - everything else (shock bands, freestream coupling, imaginary parts) is
//     a short CSR over the listed rows, one wavefront per row.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

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

// ---------------------------------------------------------------- tile layout
// Batched rows of R doubles (R = 8, 16, 32, 64) can be kept in the matrix
// cores' operand order instead of row-major: inside a cell's B*R doubles,
// variable var = 4*kc + kk and column j = W*nt + n (W = min(R, 16)) sit at
//     ((kk*NT + nt)*W + n)*KC + kc          (KC = B/4, NT = R/W),
// so the KC values a lane (n, kk) feeds to (or receives from) one
// v_mfma_f64_16x16x4f64 per k-chunk are adjacent: one 16-byte access per
// lane instead of KC 8-byte ones.  Extra-DOF rows (row >= nflow) stay
// row-major.
__device__ __forceinline__ int64_t tile_pos(int64_t row, int j, int b, int r) {
    const int w = r < 16 ? r : 16;
    const int nt_ = r / w, kc_ = b >> 2;
    const int64_t cell = row / b;
    const int var = (int)(row - cell * b);
    const int kk = var & 3, kc = var >> 2;
    const int nt = j / w, n = j - nt * w;
    return cell * b * r + ((int64_t)(kk * nt_ + nt) * w + n) * kc_ + kc;
}

// position of (row, j) in an internal batched buffer of width r; tile_b = 0
// selects row-major, else the tile layout above with b = tile_b
__device__ __forceinline__ int64_t buf_pos(int64_t row, int j, int r, int tile_b, int64_t nflow) {
    return (tile_b && row < nflow) ? tile_pos(row, j, tile_b, r) : row * r + j;
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

// Block values are streamed once per term; NT = true marks the loads
// non-temporal so they do not displace the gathered x entries in L2.
template <int N, bool NT = false>
__device__ __forceinline__ void load_vals(const double *__restrict__ p, double *a) {
#pragma unroll
    for (int k = 0; k < N; k += 2) {
        const d2 v = NT ? __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p + k))
                        : *reinterpret_cast<const d2 *>(p + k);
        a[k] = v.x;
        a[k + 1] = v.y;
    }
}

template <int N, bool NT = false>
__device__ __forceinline__ void load_vals(const float *__restrict__ p, double *a) {
    static_assert(N % 4 == 0, "quads");
#pragma unroll
    for (int k = 0; k < N; k += 4) {
        const f4 v = NT ? __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p + k))
                        : *reinterpret_cast<const f4 *>(p + k);
        a[k] = (double)v.x;
        a[k + 1] = (double)v.y;
        a[k + 2] = (double)v.z;
        a[k + 3] = (double)v.w;
    }
}

// Values per 16-byte chunk of the lane-interleaved block layout.
template <typename VT>
struct Vpl {
    static constexpr int value = 16 / (int)sizeof(VT);
};

// Block row of lane `lane` in one slot (base = first value of the slot):
// B values, chunk kc at base + (kc*64 + lane)*VPL.
template <int B, typename VT, bool NT = false>
__device__ __forceinline__ void load_lane_row(const VT *__restrict__ base, int lane, double *a) {
    constexpr int V = Vpl<VT>::value;
#pragma unroll
    for (int kc = 0; kc < B / V; ++kc)
        load_vals<V, NT>(base + ((int64_t)kc * 64 + lane) * V, a + kc * V);
}

// One Taylor term over the block part:  y = A*x (the extra CSR rows add
// theirs afterwards, k_extra), t = c1*y, tout = t (if tout),
// vout = vin + c2*t (if vout).
// c1/c2 select the recursion: normalised (c1 = dt/j, c2 = 1) or powers
// (c1 = dt, c2 = 1/j!).  With tout = y and vout = null this is a plain SpMV.
// vin may equal vout (in-place accumulation of the own row) or x (first term
// of a substep); x is never written by this kernel.
template <int B, typename VT, int NV, bool NT = false>
__global__ __launch_bounds__(256) void k_term(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t nslice, int64_t ncell,
    const double *__restrict__ x,
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
    const VT *vp = val + p0 * BB;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll 3
    for (int t = 0; t < width; ++t) {
        const int64_t c = cp[(int64_t)t * CPS];
        double a[B];
        load_lane_row<B, VT, NT>(vp + (int64_t)t * CPS * BB, lane, a);
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
        const double t = c1 * y;
        if (tout) tout[row * NV + v] = t;
        if (vout) vout[row * NV + v] = fma(c2, t, vin[row * NV + v]);
    }
}

// Replicated freestream blocks of the receptivity operator (see HostReplica):
// kind 1 -- listed rows read virtual columns c >= vbase from sum[(c - vbase)
// * NV + v], the column sums over the nrep copies, formed with the diagonal
// rows [c0, c0 + nrep p) by k_rep_sum before k_extra runs; kind 2 -- nshared
// shared rows (vptr/vcol/vval) each finished into rows c0 + k p + q with their
// own diagonal, in k_extra's last workgroups.
struct ReplicaDev {
    int kind = 0;
    int64_t c0 = 0, p = 0, nrep = 0;
    int64_t vbase = INT64_MAX;      // kind 1: first virtual column (n)
    const double *sum = nullptr;    // kind 1: p * NV column sums
    const double *diag = nullptr;   // nrep * p diagonal values (1 or 2 doubles)
    const int64_t *vptr = nullptr;  // kind 2: shared rows
    const int32_t *vcol = nullptr;
    const double *vval = nullptr;
};

// The term kernels' fused replicated rows, FR = (complex extra values ? 2 : 1)
// + (adjoint ? 4 : 0); 0: none.
template <int FR>
struct FusedKind {
    static constexpr bool on = FR != 0;
    static constexpr bool cplx = (FR & 3) == 2;
    static constexpr bool adj = (FR & 4) != 0;
    static constexpr bool fwd = on && !adj;
};

// Adjoint (kind 2): workgroups [0, nfs_wg) evaluate the shared rows and
// finish their copies (extra_replica_rows, as k_extra would after the block
// kernel; they read only x's flow rows), the others run the block term; no
// column sums, no epilogue entries.
// Forward: a Horner term with the replicated rows of a kind 1 operator fused
// into the term kernel's launch: workgroups [0, nfs_wg) finish the replicated rows
// and form the column sums of the new rows into s_out (the sums the next
// term needs); the others add each flow row's compressed entries times the
// sums s_in of this term's x (formed by the previous term, or once before
// the sweep) in their epilogue.  The flow rows' entries are grouped by
// slice of 64 rows: slice_tab[s] is the slice's block or -1, block k's lane
// l owns entries [lane_ptr[65 k + l], lane_ptr[65 k + l + 1]) of ecol (the
// column q of the sums) and eval (1 or 2 doubles each).
struct FusedRepDev {
    int64_t nfs_wg = 0;
    int64_t nflow = 0;   // adjoint: the shared rows read x's flow rows
    // slices are taken in the order rot, rot + 1, ..., wrapping: the first
    // slice with entries leads, so the slices whose epilogue does more work
    // are dispatched first instead of forming the launch's tail
    int64_t rot = 0;
    ReplicaDev R;
    const double *s_in = nullptr;
    double *s_out = nullptr;
    const int32_t *slice_tab = nullptr, *lane_ptr = nullptr, *ecol = nullptr;
    const double *eval = nullptr;
};

// x value (row c, double v) as the extra rows read it: with replicated
// blocks kept once (REP) a virtual column reads their column sum.  The
// address is selected, not branched on, so that a lane's loads of several
// entries can be in flight together; TILE = false (no tile-layout rows in x,
// tile_b == 0) leaves out buf_pos's test.
template <bool REP, bool TILE = true>
__device__ __forceinline__ double xread(const double *__restrict__ x, const ReplicaDev &R,
                                        int64_t c, int v, int nv, int tile_b, int64_t nflow) {
    const double *px = x + (TILE ? buf_pos(c, v, nv, tile_b, nflow) : c * nv + v);
    if constexpr (REP) {
        const double *ps = R.sum + (c - R.vbase) * nv + v;
        px = c >= R.vbase ? ps : px;
    }
    return *px;
}

// acc += row entries [e, e1) times x, for a lane owning NC doubles of the
// row from v0: the loads of G entries are issued before their products, so
// a short row costs ceil(len / G) memory round trips instead of len; the
// products are added in entry order (the same sum as one entry at a time).
template <int NV, bool CPLX, int G, bool REP, bool TILE = true>
__device__ __forceinline__ void row_dot(double *acc, int64_t e, int64_t e1,
                                        const int32_t *__restrict__ col,
                                        const double *__restrict__ val,
                                        const double *__restrict__ x, const ReplicaDev &R, int v0,
                                        int tile_b, int64_t nflow) {
    constexpr int NC = CPLX ? 2 : 1;
    for (; e < e1; e += G) {
        double a[G][NC], xv[G][NC];
        int32_t cg[G];
        // columns and values first, then the x entries (one wait for the
        // columns); past the row's end the loads repeat entry e
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int64_t eg = e + g < e1 ? e + g : e;
            cg[g] = col[eg];
#pragma unroll
            for (int i = 0; i < NC; ++i) a[g][i] = val[eg * NC + i];
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int i = 0; i < NC; ++i)
                xv[g][i] = xread<REP, TILE>(x, R, cg[g], v0 + i, NV, tile_b, nflow);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (e + g < e1) {
                if constexpr (CPLX) {
                    acc[0] = fma(a[g][0], xv[g][0], fma(-a[g][1], xv[g][1], acc[0]));
                    acc[1] = fma(a[g][0], xv[g][1], fma(a[g][1], xv[g][0], acc[1]));
                } else {
                    acc[0] = fma(a[g][0], xv[g][0], acc[0]);
                }
            }
        }
    }
}

// Lanes of a replicated row group: LPR lanes over the row's doubles (NC
// each), KS lanes over the copies, VPW groups per wave.
template <int NV, bool CPLX>
struct RepLanes {
    static constexpr int NC = CPLX ? 2 : 1;
    static constexpr int LPR = CPLX ? (NV >= 2 ? NV / 2 : 1) : NV;
    static constexpr int KS = (64 / LPR) >= 8 ? 8 : 64 / LPR;
    static constexpr int VPW = 64 / (LPR * KS);
};

// Kind 1, before k_extra: for shared column q, S[q] = sum_k x[c0 + k p + q]
// (copies split over KS lanes, then a fixed-order lane exchange), and each
// row c0 + k p + q -- which holds only its diagonal -- finished as
// extra_short_rows would: t = c1 * d * x, tout = t, vout = vin + c2 * t.
// With OUT_SUM the sums are of the new rows vout instead of x: the x of the
// next Horner term (vout must then be set).  w: this wave's index.
template <int NV, bool CPLX, bool OUT_SUM>
__device__ __forceinline__ void rep_sum_wave(const ReplicaDev &R, int64_t w,
                                             const double *__restrict__ x,
                                             double *__restrict__ sum, double *__restrict__ tout,
                                             const double *vin, double *vout, double c1,
                                             double c2) {
    using L = RepLanes<NV, CPLX>;
    constexpr int NC = L::NC, LPR = L::LPR, KS = L::KS, VPW = L::VPW;
    constexpr int VS = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int64_t q = w * VPW + lane / (LPR * KS);
    const bool live = q < R.p;
    const int ks = (lane / LPR) % KS;
    const int v0 = (lane % LPR) * NC;
    double s[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) s[i] = 0.0;
    if (live) {
        // the copies in chunks of KU, loads first
        constexpr int KU = 4;
        for (int64_t k0 = ks; k0 < R.nrep; k0 += KS * KU) {
            double xv[KU][NC], dv[KU][VS], vv[KU][NC];
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const int64_t k = k0 + (int64_t)u * KS < R.nrep ? k0 + (int64_t)u * KS : k0;
                const int64_t row = R.c0 + k * R.p + q;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    xv[u][i] = x[row * NV + v0 + i];
                    vv[u][i] = vout ? vin[row * NV + v0 + i] : 0.0;
                }
#pragma unroll
                for (int i = 0; i < VS; ++i) dv[u][i] = R.diag[(k * R.p + q) * VS + i];
            }
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const int64_t k = k0 + (int64_t)u * KS;
                if (k >= R.nrep) break;
                const int64_t row = R.c0 + k * R.p + q;
                double acc[NC];
                if constexpr (!OUT_SUM) {
#pragma unroll
                    for (int i = 0; i < NC; ++i) s[i] += xv[u][i];
                }
                if constexpr (CPLX) {
                    acc[0] = fma(dv[u][0], xv[u][0], fma(-dv[u][1], xv[u][1], 0.0));
                    acc[1] = fma(dv[u][0], xv[u][1], fma(dv[u][1], xv[u][0], 0.0));
                } else {
                    acc[0] = fma(dv[u][0], xv[u][0], 0.0);
                }
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const int64_t o = row * NV + v0 + i;
                    const double t = c1 * acc[i];
                    if (tout) tout[o] = t;
                    if constexpr (OUT_SUM) {
                        const double wn = fma(c2, t, vv[u][i]);
                        vout[o] = wn;
                        s[i] += wn;
                    } else {
                        if (vout) vout[o] = fma(c2, t, vv[u][i]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int off = LPR; off < LPR * KS; off <<= 1) s[i] += __shfl_xor(s[i], off, 64);
    if (live && ks == 0) {
#pragma unroll
        for (int i = 0; i < NC; ++i) sum[q * NV + v0 + i] = s[i];
    }
}

template <int NV, bool CPLX>
__global__ __launch_bounds__(256) void k_rep_sum(ReplicaDev R, const double *__restrict__ x,
                                                 double *__restrict__ sum,
                                                 double *__restrict__ tout, const double *vin,
                                                 double *vout, double c1, double c2) {
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    rep_sum_wave<NV, CPLX, false>(R, w, x, sum, tout, vin, vout, c1, c2);
}

// Kind 2, in k_extra: shared row q evaluated once per lane group (the KS
// lanes of a group each evaluate it and finish every KS-th copy), then
// finished into rows c0 + k p + q with each row's diagonal added last -- the
// arithmetic extra_short_rows does on each unshared row.
template <int NV, bool CPLX, bool TILE = true>
__device__ __forceinline__ void extra_replica_rows(
    int64_t w, const ReplicaDev &R, int64_t nflow, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    int tile_b) {
    using L = RepLanes<NV, CPLX>;
    constexpr int NC = L::NC, LPR = L::LPR, KS = L::KS, VPW = L::VPW;
    constexpr int VS = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int64_t q = w * VPW + lane / (LPR * KS);
    if (q >= R.p) return;
    const int ks = (lane / LPR) % KS;
    const int v0 = (lane % LPR) * NC;
    double sh[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) sh[i] = 0.0;
    row_dot<NV, CPLX, 4, false, TILE>(sh, R.vptr[q], R.vptr[q + 1], R.vcol, R.vval, x, R, v0,
                                      tile_b, nflow);
    // the copies in chunks of KU, loads first
    constexpr int KU = 2;
    for (int64_t k0 = ks; k0 < R.nrep; k0 += KS * KU) {
        double xv[KU][NC], dv[KU][VS], vv[KU][NC];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int64_t k = k0 + (int64_t)u * KS < R.nrep ? k0 + (int64_t)u * KS : k0;
            const int64_t row = R.c0 + k * R.p + q;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                xv[u][i] = x[row * NV + v0 + i];
                vv[u][i] = vout ? vin[row * NV + v0 + i] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < VS; ++i) dv[u][i] = R.diag[(k * R.p + q) * VS + i];
        }
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int64_t k = k0 + (int64_t)u * KS;
            if (k >= R.nrep) break;
            const int64_t row = R.c0 + k * R.p + q;
            double acc[NC];
            if constexpr (CPLX) {
                acc[0] = fma(dv[u][0], xv[u][0], fma(-dv[u][1], xv[u][1], sh[0]));
                acc[1] = fma(dv[u][0], xv[u][1], fma(dv[u][1], xv[u][0], sh[1]));
            } else {
                acc[0] = fma(dv[u][0], xv[u][0], sh[0]);
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const int64_t o = row * NV + v0 + i;
                const double t = c1 * acc[i];
                if (tout) tout[o] = t;
                if (vout) vout[o] = fma(c2, t, vv[u][i]);
            }
        }
    }
}

// One extra CSR row with many entries on the NW waves of a workgroup.  Flow
// rows (row < nflow) add their extra entries' part to the block kernel's
// output (which runs first); extra-DOF rows finish the term here.  CPLX: values
// are complex (re,im) and NV counts doubles of complex vectors ((re,im)
// pairs); otherwise values are real and every one of the NV components is
// an independent right-hand side.
// A lane owns NC doubles of the row (one complex pair when CPLX): LPR lanes
// cover the row's NV doubles, so the LPR lanes of an entry read its x row as
// adjacent doubles, and a wave takes EPW = 64 / LPR entries per step, the
// workgroup NW * EPW.  Partial sums meet by lane exchange inside each wave,
// then through LDS in wave order.
// cdiag (optional): a complex diagonal that differs per column (per pair
// (re, im) of the NV doubles), added on rows [cd0, cd1) -- the -i*omega_k
// block of the factored receptivity systems, one wavenumber per column.
template <int NV, bool CPLX, int NW, bool REP, bool TILE, bool OVW = false>
__device__ __forceinline__ void extra_long_row(
    int64_t ri, const int64_t *__restrict__ rows, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val,
    int64_t nflow, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    const double *__restrict__ cdiag, int64_t cd0, int64_t cd1, int tile_b,
    const ReplicaDev &R) {
    constexpr int LPR = CPLX ? NV / 2 : NV;    // lanes per entry
    constexpr int EPW = 64 / LPR;              // entries per wave and step
    constexpr int NC = CPLX ? 2 : 1;           // doubles per lane
    static_assert(!CPLX || NV >= 2, "complex rows need two doubles");
    static_assert(64 % LPR == 0, "lanes per entry divide the wave");
    __shared__ double part[NW > 1 ? NW - 1 : 1][64];
    // one wave per row (NW = 1) may run inside a larger workgroup: its
    // rows are the waves' own
    const int lane = threadIdx.x & 63, wv = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
    const int j = lane / LPR;                  // entry of the step
    const int v0 = (lane % LPR) * NC;          // first of this lane's doubles
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] = 0.0;
    // the operands of the epilogue's update, loaded before the row's loop
    // (only this workgroup writes the row in this launch)
    const int64_t row = rows[ri];
    double old_t[NC], old_v[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int v = v0 + q;
        old_t[q] = old_v[q] = 0.0;
        if (j == 0 && wv == 0) {
            if (row < nflow) {
                const int64_t o = buf_pos(row, v, NV, tile_b, nflow);
                if (OVW) {   // the whole row is here: start from vin
                    if (vout) old_v[q] = vin[o];
                } else {
                    if (tout) old_t[q] = tout[o];
                    if (vout) old_v[q] = vout[o];
                }
            } else if (vout) {
                old_v[q] = vin[row * NV + v];
            }
        }
    }
    // A lane's entries lie S apart; G of them are in flight at a time: their
    // column, value and x loads are issued before the products, which are
    // added in entry order (the same sum as one entry at a time).  A row of
    // the shock bands (3 600 entries at 400 chi stations) then costs one
    // round of dependent loads per lane instead of fourteen.
    constexpr int G = (NC == 1 && LPR <= 4) ? 8 : 4;
    constexpr int64_t S = (int64_t)NW * EPW;
    const int64_t e1 = ptr[ri + 1];
    for (int64_t e = ptr[ri] + wv * EPW + j; e < e1; e += G * S) {
        double a[G][NC], xv[G][NC];
        int32_t cg[G];
        // columns and values of the G entries first, then their x entries:
        // one wait for all the columns (past the row's end the loads repeat
        // entry e, so nothing outside the row is read)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int64_t eg = e + g * S < e1 ? e + g * S : e;
            cg[g] = col[eg];
#pragma unroll
            for (int i = 0; i < NC; ++i) a[g][i] = val[eg * NC + i];
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int i = 0; i < NC; ++i)
                xv[g][i] = xread<REP, TILE>(x, R, cg[g], v0 + i, NV, tile_b, nflow);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (e + g * S < e1) {
                if constexpr (CPLX) {
                    acc[0] = fma(a[g][0], xv[g][0], fma(-a[g][1], xv[g][1], acc[0]));
                    acc[1] = fma(a[g][0], xv[g][1], fma(a[g][1], xv[g][0], acc[1]));
                } else {
                    acc[0] = fma(a[g][0], xv[g][0], acc[0]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
        for (int off = 32; off >= LPR; off >>= 1) acc[q] += __shfl_xor(acc[q], off, 64);
    if constexpr (NW > 1) {
        if (wv > 0 && j == 0) {
#pragma unroll
            for (int q = 0; q < NC; ++q) part[wv - 1][v0 + q] = acc[q];
        }
        __syncthreads();
        if (wv > 0) return;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i)
#pragma unroll
            for (int q = 0; q < NC; ++q) acc[q] += part[i][v0 + q];
    }
    if (j != 0) return;
    if constexpr (NV % 2 == 0) {
        if (cdiag && row >= cd0 && row < cd1) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                // the pair (re, im) this double belongs to
                const int v = v0 + q, vr = v & ~1;
                const double dr = cdiag[vr], di = cdiag[vr + 1];
                const double xr = x[row * NV + vr], xi = x[row * NV + vr + 1];
                acc[q] = (v & 1) ? fma(dr, xi, fma(di, xr, acc[q]))
                                 : fma(dr, xr, fma(-di, xi, acc[q]));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int v = v0 + q;
        if (row < nflow) {
            // flow row: the block kernel has written its part already
            const int64_t o = buf_pos(row, v, NV, tile_b, nflow);
            const double t = c1 * acc[q];
            if (tout) tout[o] = old_t[q] + t;
            if (vout) vout[o] = fma(c2, t, old_v[q]);
        } else {
            const double t = c1 * acc[q];
            if (tout) tout[row * NV + v] = t;
            if (vout) vout[row * NV + v] = fma(c2, t, old_v[q]);
        }
    }
}

// Extra CSR rows with few entries (the -i*omega diagonal of the freestream
// DOFs, rows touching one shock DOF): no cross-lane reduction, each lane
// owns one (row, column) pair -- a complex column pair when CPLX -- and
// loops over the row's entries; RPW = 64 / lanes-per-row rows per wave.
// Wave w takes rows [r0 + w RPW, r0 + (w + 1) RPW) of [r0, nrows); same
// outputs as extra_long_row.
template <int NV, bool CPLX, bool REP, bool TILE>
__device__ __forceinline__ void extra_short_rows(
    int64_t w, const int64_t *__restrict__ rows, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val, int64_t r0, int64_t nrows,
    int64_t nflow, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    const double *__restrict__ cdiag, int64_t cd0, int64_t cd1, int tile_b,
    const ReplicaDev &R) {
    constexpr int LPR = CPLX ? (NV >= 2 ? NV / 2 : 1) : NV;   // lanes per row
    constexpr int RPW = 64 / LPR;                             // rows per wave
    constexpr int NC = CPLX ? 2 : 1;                          // doubles per lane
    static_assert(64 % LPR == 0, "lanes per row divide the wave");
    const int lane = threadIdx.x & 63;
    const int64_t ri = r0 + w * RPW + lane / LPR;
    if (ri >= nrows) return;
    const int v0 = (lane % LPR) * NC;    // first of this lane's doubles in the row
    const int64_t row = rows[ri];
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] = 0.0;
    // one entry at a time (the unreplicated operator's short rows hold a
    // single entry or a few); four loads at a time on the replicated form's
    // rows of compressed B entries
    row_dot<NV, CPLX, REP ? 4 : 1, REP, TILE>(acc, ptr[ri], ptr[ri + 1], col, val, x, R, v0,
                                              tile_b, nflow);
    // the per-column complex diagonal acts on (re, im) pairs of the row
    // whether the operator's values are complex (a lane owns a pair) or real
    // (a lane owns one double of a pair)
    if constexpr (NV % 2 == 0) {
        if (cdiag && row >= cd0 && row < cd1) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const int v = v0 + q, vr = v & ~1;
                const double dr = cdiag[vr], di = cdiag[vr + 1];
                const double xr = x[row * NV + vr], xi = x[row * NV + vr + 1];
                acc[q] = (v & 1) ? fma(dr, xi, fma(di, xr, acc[q]))
                                 : fma(dr, xr, fma(-di, xi, acc[q]));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int v = v0 + q;
        if (row < nflow) {
            // flow row: the block kernel has written its part already
            const int64_t o = buf_pos(row, v, NV, tile_b, nflow);
            const double t = c1 * acc[q];
            if (tout) tout[o] += t;
            if (vout) vout[o] = fma(c2, t, vout[o]);
        } else {
            const double t = c1 * acc[q];
            if (tout) tout[row * NV + v] = t;
            if (vout) vout[row * NV + v] = fma(c2, t, vin[row * NV + v]);
        }
    }
}

// A long real row at NV = 2..64 doubles per row with LPR lanes per entry,
// each lane loading a contiguous chunk of CH = NV / LPR doubles of the
// entry's x row (16-byte loads from CH = 2 up).  The LPR lanes of an entry
// cover its x row in one cache line, and G entries per lane are in flight
// (G * CH about 16 doubles): with extra_long_row's NV lanes per entry and
// four entries in flight, a 3 600-entry shock row at NV = 8 takes 28
// rounds of dependent loads per 4-wave workgroup.  A lane's entries lie
// 64*NW/LPR apart.  The chunk sums are reduced across the wave by
// butterfly shuffles over the lanes of one chunk position and across waves
// through LDS in wave order.  No per-column diagonal.  OVW: a flow row is
// written whole (its block entries are in the row), not added to what the
// block kernel wrote.
template <int NV, int NW, bool TILE, int LPR, bool OVW = false>
__device__ __forceinline__ void extra_long_row_lpe(
    int64_t ri, const int64_t *__restrict__ rows, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val, int64_t nflow,
    const double *__restrict__ x, double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2, int tile_b) {
    static_assert(NV >= 2 && NV <= 64 && NV % LPR == 0 && LPR <= 64, "LPR divides NV <= 64");
    constexpr int CH = NV / LPR;
    constexpr int G = CH == 1 ? 16 : (CH == 2 ? 8 : (CH == 4 ? 4 : 2));
    constexpr int EPW = 64 / LPR;
    constexpr int64_t S = (int64_t)NW * EPW;
    __shared__ double part[NW > 1 ? NW - 1 : 1][NV];
    // one wave per row (NW = 1) may run inside a larger workgroup: its
    // rows are the waves' own
    const int lane = threadIdx.x & 63, wv = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
    const int j = lane / LPR, v0 = (lane % LPR) * CH;
    const int64_t row = rows[ri];
    // lanes [0, LPR) of wave 0 write doubles [v0, v0 + CH): old values first
    double old_t[CH], old_v[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        old_t[i] = old_v[i] = 0.0;
        if (wv == 0 && j == 0) {
            if (row < nflow) {
                const int64_t o = buf_pos(row, v0 + i, NV, tile_b, nflow);
                if (OVW) {
                    if (vout) old_v[i] = vin[o];
                } else {
                    if (tout) old_t[i] = tout[o];
                    if (vout) old_v[i] = vout[o];
                }
            } else if (vout) {
                old_v[i] = vin[row * NV + v0 + i];
            }
        }
    }
    double acc[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) acc[i] = 0.0;
    const int64_t e1 = ptr[ri + 1];
    for (int64_t e = ptr[ri] + wv * EPW + j; e < e1; e += G * S) {
        int32_t cg[G];
        double a[G];
        // past the row's end the loads repeat entry e (not added)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int64_t eg = e + g * S < e1 ? e + g * S : e;
            cg[g] = col[eg];
            a[g] = val[eg];
        }
        double xv[G][CH];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if constexpr (TILE || CH == 1) {
#pragma unroll
                for (int i = 0; i < CH; ++i)
                    xv[g][i] = x[TILE ? buf_pos(cg[g], v0 + i, NV, tile_b, nflow)
                                      : (int64_t)cg[g] * NV + v0 + i];
            } else {
                load_f64<CH>(x + (int64_t)cg[g] * NV + v0, xv[g]);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (e + g * S < e1) {
#pragma unroll
                for (int i = 0; i < CH; ++i) acc[i] = fma(a[g], xv[g][i], acc[i]);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int off = 32; off >= LPR; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
    if constexpr (NW > 1) {
        if (wv > 0 && j == 0) {
#pragma unroll
            for (int i = 0; i < CH; ++i) part[wv - 1][v0 + i] = acc[i];
        }
        __syncthreads();
        if (wv > 0) return;
#pragma unroll
        for (int k = 0; k < NW - 1; ++k)
#pragma unroll
            for (int i = 0; i < CH; ++i) acc[i] += part[k][v0 + i];
    }
    if (j != 0) return;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        const double t = c1 * acc[i];
        if (row < nflow) {
            // flow row: the block kernel has written its part already
            const int64_t o = buf_pos(row, v0 + i, NV, tile_b, nflow);
            if (tout) tout[o] = old_t[i] + t;
            if (vout) vout[o] = fma(c2, t, old_v[i]);
        } else {
            if (tout) tout[row * NV + v0 + i] = t;
            if (vout) vout[row * NV + v0 + i] = fma(c2, t, old_v[i]);
        }
    }
}

// The extra CSR rows in one launch of NW-wave workgroups: workgroups
// [0, nlong) take the long rows [0, nlong) (more than 32 entries), one each,
// in contiguous ranges per XCD (neighbouring rows gather the same x rows --
// the receptivity B rows of adjacent cells -- which then stay in one L2);
// the next nshort_wg take the short rows [nlong, nrows), RPW per wave; the
// rest (kind 2 replicas) the shared rows, VPW per wave.
// LPE > 0: real rows at 2..64 doubles per row, no per-column diagonal: the
// long rows with LPE lanes per entry (extra_long_row_lpe).
template <int NV, bool CPLX, int NW, bool REP, int LPE = 0>
__global__ __launch_bounds__(64 * NW) void k_extra(
    const int64_t *__restrict__ rows, const int64_t *__restrict__ ptr,
    const int32_t *__restrict__ col, const double *__restrict__ val, int64_t nlong,
    int64_t nrows, int64_t nshort_wg, int64_t nflow, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    const double *__restrict__ cdiag, int64_t cd0, int64_t cd1, int tile_b, ReplicaDev R) {
    const int64_t bid = blockIdx.x;
    // tile_b is uniform: one specialised body per layout, so the x reads of
    // the row-major case carry no layout test
    auto body = [&](auto tile) {
        constexpr bool TILE = decltype(tile)::value;
        if (bid < nlong) {
            if constexpr (LPE > 0 && !CPLX && !REP && NV >= 2)
                extra_long_row_lpe<NV, NW, TILE, (LPE > 0 ? LPE : 1)>(xcd_tile(bid, nlong), rows, ptr, col, val, nflow,
                                                 x, tout, vin, vout, c1, c2, tile_b);
            else
                extra_long_row<NV, CPLX, NW, REP, TILE>(xcd_tile(bid, nlong), rows, ptr, col,
                                                        val, nflow, x, tout, vin, vout, c1, c2,
                                                        cdiag, cd0, cd1, tile_b, R);
        } else if (bid < nlong + nshort_wg)
            extra_short_rows<NV, CPLX, REP, TILE>((bid - nlong) * NW + (threadIdx.x >> 6), rows,
                                                  ptr, col, val, nlong, nrows, nflow, x, tout,
                                                  vin, vout, c1, c2, cdiag, cd0, cd1, tile_b, R);
        else if constexpr (REP)
            extra_replica_rows<NV, CPLX, TILE>((bid - nlong - nshort_wg) * NW +
                                                   (threadIdx.x >> 6),
                                               R, nflow, x, tout, vin, vout, c1, c2, tile_b);
    };
    if (tile_b)
        body(std::true_type{});
    else
        body(std::false_type{});
}

// var-major (ABI) -> cell-major (internal).  x_vm holds nrhs columns of n
// elements with leading dimension ld (elements); an element is ncomp
// doubles.  Internal rows hold nrhs*ncomp doubles.  Var-major indices below
// nzero read as 0 (CONSTRUCT_C's x[1:N_set_0] .= 0).
// Internal rows are `width` doubles wide; components past nrhs*ncomp are
// padding and read as 0.  tile_b != 0: flow rows in the tile layout.
__global__ void k_perm_in(const double *__restrict__ x_vm, int64_t ld, double *__restrict__ out,
                          int64_t n, int64_t ncell, int b, int nrhs, int ncomp, int64_t nzero,
                          int width, int tile_b) {
    const int64_t valid = (int64_t)nrhs * ncomp;
    const int64_t total = n * width;
    const int64_t nflow = ncell * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / width;
        const int64_t rem = i - row * width;
        const int64_t o = buf_pos(row, (int)rem, width, tile_b, nflow);
        if (rem >= valid) {
            out[o] = 0.0;
            continue;
        }
        const int64_t rhs = rem / ncomp, comp = rem - rhs * ncomp;
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        out[o] = vm < nzero ? 0.0 : x_vm[(rhs * ld + vm) * ncomp + comp];
    }
}

__global__ void k_perm_out(const double *__restrict__ in, double *__restrict__ y_vm, int64_t ld,
                           int64_t n, int64_t ncell, int b, int nrhs, int ncomp, int64_t nzero,
                           int width, int tile_b) {
    const int64_t total = n * nrhs * ncomp;
    const int64_t nflow = ncell * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t rhs = i / (n * ncomp);
        const int64_t rem = i - rhs * n * ncomp;
        const int64_t vm = rem / ncomp, comp = rem - vm * ncomp;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        y_vm[(rhs * ld + vm) * ncomp + comp] =
            vm < nzero ? 0.0 : in[buf_pos(row, (int)(rhs * ncomp + comp), width, tile_b, nflow)];
    }
}

// Zero, in an internal (cell-major) buffer of `width` doubles per row, the
// rows whose var-major index is below nzero: CONSTRUCT_C's zeroing applied
// between two internal operators.  Only flow rows can map below nzero when
// nzero <= nflow; the loop covers every row so any nzero <= n works.
__global__ void k_zero_vm_prefix(double *__restrict__ buf, int64_t n, int64_t ncell, int b,
                                 int64_t nzero, int width, int tile_b) {
    const int64_t total = n * width;
    const int64_t nflow = ncell * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / width;
        const int j = (int)(i - row * width);
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        if (vm < nzero) buf[buf_pos(row, j, width, tile_b, nflow)] = 0.0;
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
// Each thread moves 4 x 16 bytes per iteration (four loads in flight
// before the stores) so the copy is limited by HBM, not by latency.
__global__ __launch_bounds__(256) void k_copy(const d2 *__restrict__ in, d2 *__restrict__ out,
                                              int64_t n2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const d2 a = in[i], b = in[i + stride], c = in[i + 2 * stride], d = in[i + 3 * stride];
        out[i] = a;
        out[i + stride] = b;
        out[i + 2 * stride] = c;
        out[i + 3 * stride] = d;
    }
    for (; i < n2; i += stride) out[i] = in[i];
}

// Streaming read of n2 16-byte chunks, one partial sum stored per thread:
// the read-only ceiling the value stream of the Taylor-term kernels is
// measured against.  Workgroup g reads the contiguous run [g*chunk,
// (g+1)*chunk), sixteen non-temporal loads in flight per thread, 2048
// workgroups: the fastest of a sweep over contiguous-chunk and grid-strided
// reads, 2048-32768 workgroups, 4-16 loads per thread, with and without the
// non-temporal hint, on an 828 MiB buffer (6.5 TB/s in that sweep, 5.2-6.3
// for the others; 6.4-6.5 TB/s between runs).
__global__ __launch_bounds__(256) void k_read(const d2 *__restrict__ in, double *__restrict__ out,
                                              int64_t n2, int64_t chunk) {
    constexpr int U = 16;
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t e = b0 + chunk < n2 ? b0 + chunk : n2;
    double acc0 = 0.0, acc1 = 0.0;
    int64_t i = b0 + threadIdx.x;
    for (; i + (U - 1) * 256 < e; i += U * 256) {
        d2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(in + i + u * 256);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc0 += v[u].x;
            acc1 += v[u].y;
        }
    }
    for (; i < e; i += 256) {
        const d2 a = __builtin_nontemporal_load(in + i);
        acc0 += a.x;
        acc1 += a.y;
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = acc0 + acc1;
}

// Streaming reads and writes in a fixed proportion: chunk c (workgroups
// grid-strided over nchunk chunks) reads rin 16-byte items of `in` (eight
// non-temporal loads in flight per thread) and then writes rout items of
// `out` with each thread's partial sum, so that across the launch HBM
// serves reads and writes in the ratio rin : rout -- the ceiling for a
// kernel whose traffic has that mix (a Horner term writes its new vectors
// beside the values and vectors it reads).
__global__ __launch_bounds__(256) void k_stream_mix(const d2 *__restrict__ in, int64_t rin,
                                                    d2 *__restrict__ out, int64_t rout,
                                                    int64_t nchunk) {
    constexpr int U = 8;
    for (int64_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const d2 *src = in + c * rin;
        d2 acc = d2{0.0, 0.0};
        int64_t i = threadIdx.x;
        for (; i + (U - 1) * 256 < rin; i += U * 256) {
            d2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * 256);
#pragma unroll
            for (int u = 0; u < U; ++u) acc += v[u];
        }
        for (; i < rin; i += 256) acc += __builtin_nontemporal_load(src + i);
        d2 *dst = out + c * rout;
        for (int64_t k = threadIdx.x; k < rout; k += 256) dst[k] = acc;
    }
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
    double *__restrict__ tout, const double *vin, double *vout,
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
        const VT *sb = val + (p0 + (int64_t)t * CPS) * BB;   // slot base
        double a[RPL * B];
#pragma unroll
        for (int i = 0; i < RPL; ++i)
            load_lane_row<B, VT>(sb, (int)(q * B + g * RPL + i), a + i * B);
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
        const double t = c1 * y;
        if (tout) tout[idx] = t;
        if (vout) vout[idx] = fma(c2, t, vin[idx]);
    }
}

}  // namespace hymor

namespace hymor {

// ------------------------------------------------------------ Krylov basis
// V is column-major n x m (leading dimension ld, in elements); an element is
// one double (real) or an interleaved (re, im) pair (CPLX).

// Partial dot products h_j = sum_rows conj(V[row, j]) * w[row] for columns
// j0..j0+JT-1 (those < m); each workgroup writes one partial per column
// (and per component for complex): partial[(j*ncomp + c) * gridDim.x + bid].
template <bool CPLX, int JT>
__global__ __launch_bounds__(256) void k_dots(const double *__restrict__ V, int64_t ld, int m,
                                              int j0, const double *__restrict__ w, int64_t n,
                                              double *__restrict__ partial) {
    constexpr int NC = CPLX ? 2 : 1;
    __shared__ double red[JT * NC][4];
    double acc[JT * NC];
#pragma unroll
    for (int t = 0; t < JT * NC; ++t) acc[t] = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
         row += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (CPLX) {
            const double wr = w[2 * row], wi = w[2 * row + 1];
#pragma unroll
            for (int t = 0; t < JT; ++t) {
                if (j0 + t >= m) break;
                const double *v = V + 2 * ((int64_t)(j0 + t) * ld + row);
                const double vr = v[0], vi = v[1];
                acc[2 * t] = fma(vr, wr, fma(vi, wi, acc[2 * t]));          // re(conj(v) w)
                acc[2 * t + 1] = fma(vr, wi, fma(-vi, wr, acc[2 * t + 1]));  // im(conj(v) w)
            }
        } else {
            const double wv = w[row];
#pragma unroll
            for (int t = 0; t < JT; ++t) {
                if (j0 + t >= m) break;
                acc[t] = fma(V[(int64_t)(j0 + t) * ld + row], wv, acc[t]);
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < JT * NC; ++t) {
        double a = acc[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[t][wave] = a;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < JT * NC; t += blockDim.x) {
        const int j = j0 + t / NC;
        if (j < m)
            partial[((int64_t)j * NC + t % NC) * gridDim.x + blockIdx.x] =
                (red[t][0] + red[t][1]) + (red[t][2] + red[t][3]);
    }
}

// w[row] -= sum_j V[row, j] * h[j]  (h on the device, m entries)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_sub_vh(const double *__restrict__ V, int64_t ld, int m,
                                                const double *__restrict__ h,
                                                double *__restrict__ w, int64_t n) {
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
         row += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (CPLX) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < m; ++j) {
                const double *v = V + 2 * ((int64_t)j * ld + row);
                const double hr = h[2 * j], hi = h[2 * j + 1];
                sr = fma(v[0], hr, fma(-v[1], hi, sr));
                si = fma(v[0], hi, fma(v[1], hr, si));
            }
            w[2 * row] -= sr;
            w[2 * row + 1] -= si;
        } else {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s = fma(V[(int64_t)j * ld + row], h[j], s);
            w[row] -= s;
        }
    }
}

// out[:, c] = sum_j V[:, j] * Q[j, c] for c < k (Q column-major m x k on the
// device, complex interleaved when CPLX).  out must not alias V.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_basis_combine(const double *__restrict__ V, int64_t ld,
                                                       int m, const double *__restrict__ Q, int k,
                                                       double *__restrict__ out, int64_t ldo,
                                                       int64_t n) {
    // outputs in chunks of CK: each chunk reads the row's m basis values
    // once (not once per output column), Q's entries are the same for all
    // threads; every output is summed over j in order
    constexpr int CK = 16;
    for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n;
         row += (int64_t)gridDim.x * blockDim.x) {
        for (int c0 = 0; c0 < k; c0 += CK) {
            const int nc = k - c0 < CK ? k - c0 : CK;
            if constexpr (CPLX) {
                double sr[CK], si[CK];
#pragma unroll
                for (int c = 0; c < CK; ++c) sr[c] = si[c] = 0.0;
                for (int j = 0; j < m; ++j) {
                    const double *v = V + 2 * ((int64_t)j * ld + row);
                    const double vr = v[0], vi = v[1];
#pragma unroll
                    for (int c = 0; c < CK; ++c) {
                        if (c >= nc) break;
                        const double *q = Q + 2 * ((int64_t)(c0 + c) * m + j);
                        sr[c] = fma(vr, q[0], fma(-vi, q[1], sr[c]));
                        si[c] = fma(vr, q[1], fma(vi, q[0], si[c]));
                    }
                }
#pragma unroll
                for (int c = 0; c < CK; ++c) {
                    if (c >= nc) break;
                    out[2 * ((int64_t)(c0 + c) * ldo + row)] = sr[c];
                    out[2 * ((int64_t)(c0 + c) * ldo + row) + 1] = si[c];
                }
            } else {
                double s[CK];
#pragma unroll
                for (int c = 0; c < CK; ++c) s[c] = 0.0;
                for (int j = 0; j < m; ++j) {
                    const double vj = V[(int64_t)j * ld + row];
#pragma unroll
                    for (int c = 0; c < CK; ++c) {
                        if (c >= nc) break;
                        s[c] = fma(vj, Q[(int64_t)(c0 + c) * m + j], s[c]);
                    }
                }
#pragma unroll
                for (int c = 0; c < CK; ++c) {
                    if (c >= nc) break;
                    out[(int64_t)(c0 + c) * ldo + row] = s[c];
                }
            }
        }
    }
}

// x[i] uniform in [-1, 1): splitmix64 (Steele, Lea and Flood, OOPSLA 2014)
// of the counter seed + i * golden-ratio increment, top 53 bits
__global__ __launch_bounds__(256) void k_fill_uniform(double *__restrict__ x, int64_t len,
                                                      uint64_t seed) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[i] = (double)(z >> 11) * 0x1.0p-52 - 1.0;
    }
}

// x[0 : ncomp*n) *= s (real scale)
__global__ __launch_bounds__(256) void k_scale(double *__restrict__ x, int64_t len, double s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s;
}

// x[0 : len) = 0 (a kernel, so that it may sit inside a captured graph)
__global__ __launch_bounds__(256) void k_zero(double *__restrict__ x, int64_t len) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

}  // namespace hymor

namespace hymor {

// Split-slot form of k_term: one workgroup per slice, NW waves dividing the
// slice's block slots round-robin (wave w takes slots w, w+NW, ...), partial
// sums combined through LDS in wave order (deterministic), epilogue by wave
// 0.  Shorter waves than k_term (about 1/NW of a slice each): more loads in
// flight per wave and a smaller idle tail at the end of each launch.
template <int B, typename VT, int NV, int NW, bool NT>
__global__ __launch_bounds__(64 * NW) void k_term_split(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t nslice, int64_t ncell,
    const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    __shared__ double part[NW > 1 ? NW - 1 : 1][NV][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int64_t slice = xcd_tile(blockIdx.x, gridDim.x);
    if (slice >= nslice) return;
    const int q = lane / B, r = lane % B;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int32_t *cp = col + p0 + q;
    const VT *vp = val + p0 * BB;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll 2
    for (int t = w; t < width; t += NW) {
        const int64_t c = cp[(int64_t)t * CPS];
        double a[B];
        load_lane_row<B, VT, NT>(vp + (int64_t)t * CPS * BB, lane, a);
        double xv[B * NV];
        load_f64<B * NV>(x + c * (B * NV), xv);
#pragma unroll
        for (int k = 0; k < B; ++k)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] = fma(a[k], xv[k * NV + v], acc[v]);
    }
    if constexpr (NW > 1) {
        if (w > 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) part[w - 1][v][lane] = acc[v];
        }
        __syncthreads();
        if (w > 0) return;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += part[i][v][lane];
    }
    const int64_t cell = slice * CPS + q;
    if (cell >= ncell) return;
    const int64_t row = cell * B + r;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double y = acc[v];
        const double t = c1 * y;
        if (tout) tout[row * NV + v] = t;
        if (vout) vout[row * NV + v] = fma(c2, t, vin[row * NV + v]);
    }
}

// The extra entries of a real operator whose flow rows couple only to
// extra DOFs (the shock bands; HostShockPanels) fused into the
// single-vector term kernel.  Workgroups in launch order:
//   [0, nsh_wg): the nrows extra-DOF rows, one workgroup per row
//     (extra_long_row);
//   the next npan_wg: the npan slices with a panel (pslices), one each,
//     dealt round-robin over the XCDs: their block term plus the panel, for
//     every column c of it pval[64 c + lane] * x[pcol[c]] on row lane;
//   the rest: the block term of every other slice, in xcd_tile order (a
//     slice with a panel returns at once).
// nsh_wg and npan_wg are multiples of 8, so the last group keeps the XCD
// mapping of an unfused launch.  The leading groups' work lasts longer per
// workgroup and starts first, under the bulk of the block term.
struct ShockFuseDev {
    int64_t nsh_wg = 0, nrows = 0, nflow = 0;
    int64_t npan_wg = 0, npan = 0;
    const int32_t *pslices = nullptr;
    const int64_t *rows = nullptr, *ptr = nullptr;
    const int32_t *col = nullptr;
    const double *val = nullptr;
    const int32_t *slice_tab = nullptr;
    const int64_t *cptr = nullptr;
    const int32_t *pcol = nullptr;
    const double *pval = nullptr;
};

// A shock operator's batched term in one launch (k_term_lds at 2..4 and
// k_spmm_quad4c at 8..64 doubles per row): leading workgroups, one row
// each, finish the extra-DOF rows (CSR sh_*: [0, nsh)) and, whole (block and
// shock entries), every flow row of the slices with shock columns (CSR fr_*:
// [nsh, nsh + nfr)) -- they read only x, so they overlap the block term --
// and the block kernel skips those slices (skip[slice] >= 0).  nlead_wg (a
// multiple of 8) workgroups lead.
struct ShockBatchDev {
    int64_t nlead_wg = 0, nsh = 0, nfr = 0, nflow = 0;
    const int64_t *sh_rows = nullptr, *sh_ptr = nullptr;
    const int32_t *sh_col = nullptr;
    const double *sh_val = nullptr;
    const int64_t *fr_rows = nullptr, *fr_ptr = nullptr;
    const int32_t *fr_col = nullptr;
    const double *fr_val = nullptr;
    const int32_t *skip = nullptr;
};

// k_term_split with the loads of each wave grouped: the block columns of
// the wave's next S slots and their values are issued together, then the x
// entries they select, and wave 0 issues its epilogue operand (vin)
// before the slot loop -- two waits on memory per group instead of two per
// slot pair plus one at the end.  With NW * S = 9 a 3x3-stencil row has no
// idle slot; an idle slot (shorter or longer rows) contributes exact zeros.
// Same per-wave summation order and LDS combination as k_term_split.
template <typename VT>
struct Chunk;
template <>
struct Chunk<double> {
    typedef d2 type;
};
template <>
struct Chunk<float> {
    typedef f4 type;
};

//
// SH: the shock bands fused in (ShockFuseDev; one real vector): no k_extra
// launch follows.
template <int B, typename VT, int NV, int NW, int S, bool NT, bool SH = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_term_split_p(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t nslice, int64_t ncell,
    const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    ShockFuseDev F) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int V = Vpl<VT>::value;
    constexpr int NC = B / V;  // 16-byte chunks per lane and slot
    typedef typename Chunk<VT>::type CT;
    static_assert(!SH || NV == 1, "the shock bands act on one real vector");
    __shared__ double part[NW > 1 ? NW - 1 : 1][NV][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int64_t bid = blockIdx.x, nwg = gridDim.x;
    if constexpr (SH) {
        // the extra-DOF rows first: they read only x, so they overlap the
        // block term instead of following it
        if (bid < F.nsh_wg) {
            if (bid < F.nrows)
                extra_long_row<NV, false, NW, false, false>(bid, F.rows, F.ptr, F.col, F.val,
                                                           F.nflow, x, tout, vin, vout, c1, c2,
                                                           nullptr, 0, 0, 0, ReplicaDev{});
            return;
        }
        bid -= F.nsh_wg;
        nwg -= F.nsh_wg;
    }
    int64_t slice;
    if constexpr (SH) {
        if (bid < F.npan_wg) {
            if (bid >= F.npan) return;
            slice = F.pslices[bid];
        } else {
            slice = xcd_tile(bid - F.npan_wg, nwg - F.npan_wg);
            if (slice < nslice && F.slice_tab[slice] >= 0) return;   // a leading workgroup's
        }
    } else {
        slice = xcd_tile(bid, nwg);
    }
    if (slice >= nslice) return;
    const int q = lane / B, r = lane % B;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int32_t *cp = col + p0 + q;
    const VT *vp = val + p0 * BB;
    const int64_t cell = slice * CPS + q;
    const int64_t row = (cell < ncell ? cell : ncell - 1) * B + r;
    double pre_v[NV];
    {
        const double *pv = vout ? vin : x;
#pragma unroll
        for (int v = 0; v < NV; ++v) pre_v[v] = w == 0 ? pv[row * NV + v] : 0.0;
    }
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    // slots t = w, w + NW, ... of this wave, S at a time
    const int nmine = width > w ? (width - w + NW - 1) / NW : 0;
    for (int g = 0; g < nmine; g += S) {
        int32_t c[S];
        CT raw[S][NC];
        // columns first: the x loads wait for them, not for the values
        // (the memory counter retires loads in issue order), and no load
        // sits behind a branch, so no wait is placed before the x loads
        // beyond the columns'.  A slot past the wave's last one re-reads
        // that slot (cache hits) and is zeroed before the product.
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int i = g + u < nmine ? g + u : nmine - 1;
            c[u] = cp[(int64_t)(w + i * NW) * CPS];
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int i = g + u < nmine ? g + u : nmine - 1;
            const VT *base = vp + (int64_t)(w + i * NW) * CPS * BB;
#pragma unroll
            for (int kc = 0; kc < NC; ++kc) {
                const CT *pc = reinterpret_cast<const CT *>(base + ((int64_t)kc * 64 + lane) * V);
                raw[u][kc] = NT ? __builtin_nontemporal_load(pc) : *pc;
            }
        }
        double xv[S][B * NV];
#pragma unroll
        for (int u = 0; u < S; ++u) load_f64<B * NV>(x + (int64_t)c[u] * (B * NV), xv[u]);
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const bool live = g + u < nmine;
#pragma unroll
            for (int kc = 0; kc < NC; ++kc)
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const int k = kc * V + e;
                    const double ak = live ? (double)raw[u][kc][e] : 0.0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] = fma(ak, xv[u][k * NV + v], acc[v]);
                }
        }
    }
    if constexpr (SH) {
        // the slice's panel: wave w takes columns w, w + NW, ..., PG at a
        // time (values and x loads first; the column and its x entry are
        // the same for every lane)
        const int32_t pk = F.slice_tab[slice];
        if (pk >= 0) {
            constexpr int PG = 16;
            const int64_t ce = F.cptr[pk + 1];
            for (int64_t cb = F.cptr[pk] + w; cb < ce; cb += (int64_t)NW * PG) {
                double a[PG], xe[PG];
#pragma unroll
                for (int g = 0; g < PG; ++g) {
                    const int64_t c = cb + (int64_t)g * NW < ce ? cb + (int64_t)g * NW : cb;
                    a[g] = F.pval[c * 64 + lane];
                    xe[g] = x[F.pcol[c]];
                }
#pragma unroll
                for (int g = 0; g < PG; ++g)
                    if (cb + (int64_t)g * NW < ce) acc[0] = fma(a[g], xe[g], acc[0]);
            }
        }
    }
    if constexpr (NW > 1) {
        if (w > 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) part[w - 1][v][lane] = acc[v];
        }
        __syncthreads();
        if (w > 0) return;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += part[i][v][lane];
    }
    if (cell >= ncell) return;
    // the vout stores first, behind one wait for the epilogue operands (a
    // store between two such waits would be waited for as well)
    if (vout) {
#pragma unroll
        for (int v = 0; v < NV; ++v) vout[row * NV + v] = fma(c2, c1 * acc[v], pre_v[v]);
    }
    if (tout) {
#pragma unroll
        for (int v = 0; v < NV; ++v) tout[row * NV + v] = c1 * acc[v];
    }
}

// Orders one wave's LDS accesses around a hand-off between its lanes: the
// LDS unit runs a wave's instructions in issue order, so keeping the
// compiler from moving accesses across this point is all that is needed.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// k_term_split_p for NV >= 2 doubles per row with the gathered x blocks
// staged through LDS.  In k_term_split_p every lane of a cell loads the
// cell's whole B*NV-double x block (the B lanes of a cell read the same
// addresses); here the B lanes load it once as contiguous 16-byte chunks
// (NV/2 per lane), write it to the wave's LDS area and read it back with
// broadcast reads: B times fewer vector-memory instructions for x.  Cell
// areas are padded by one chunk so the CPS cells of a read fall in
// different banks.  Same slot order, FMA order and LDS combination as
// k_term_split_p, hence the same bits.
//
// FR != 0: the receptivity operator's replicated rows fused in (FusedRepDev,
// FusedKind: forward Horner terms of a kind 1 operator, or any term of a
// kind 2 adjoint): no k_extra launch follows.
//
// SH: a shock operator's extra-DOF rows in leading workgroups
// (ShockBatchDev; one wave per slice).
template <int B, typename VT, int NV, int NW, int S, bool NT, int FR = 0, bool SH = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(1, 4))) void k_term_lds(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t nslice, int64_t ncell,
    const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2,
    FusedRepDev F, ShockBatchDev H) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int V = Vpl<VT>::value;
    constexpr int NC = B / V;       // 16-byte value chunks per lane and slot
    constexpr int XC = B * NV / 2;  // 16-byte chunks per x block
    constexpr int CPL = NV / 2;     // x chunks each lane loads
    constexpr int XS = XC + 1;      // padded chunk stride of a cell's area
    static_assert(NV >= 2 && NV % 2 == 0, "x blocks move as pairs of doubles");
    typedef typename Chunk<VT>::type CT;
    using FK = FusedKind<FR>;
    static_assert(!(FK::fwd && FK::cplx) || NV == 2, "complex entries act on one complex vector");
    static_assert(!SH || (NW == 1 && FR == 0), "shock rows: one wave per slice, no replicas");
    __shared__ double part[NW > 1 ? NW - 1 : 1][NV][64];
    __shared__ d2 xs[NW][S][CPS * XS];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int64_t bid = blockIdx.x, nwg = gridDim.x;
    if constexpr (SH) {
        // the leading rows first: they read only x
        if (bid < H.nlead_wg) {
            if (bid < H.nsh)
                extra_long_row_lpe<NV, 1, false, NV / 2, true>(
                    bid, H.sh_rows, H.sh_ptr, H.sh_col, H.sh_val, H.nflow, x, tout, vin, vout,
                    c1, c2, 0);
            else if (bid < H.nsh + H.nfr)
                extra_long_row_lpe<NV, 1, false, NV / 2, true>(
                    bid - H.nsh, H.fr_rows, H.fr_ptr, H.fr_col, H.fr_val, H.nflow, x, tout, vin,
                    vout, c1, c2, 0);
            return;
        }
        bid -= H.nlead_wg;
        nwg -= H.nlead_wg;
    }
    if constexpr (FR != 0) {
        // the replicated rows first (nfs_wg is a multiple of 8, so the
        // slices below keep the XCD mapping of an unfused launch)
        if (bid < F.nfs_wg) {
            if constexpr (FK::adj)
                extra_replica_rows<NV, FK::cplx, false>(bid * NW + w, F.R, F.nflow, x, tout,
                                                        vin, vout, c1, c2, 0);
            else
                rep_sum_wave<NV, FK::cplx, true>(F.R, bid * NW + w, x, F.s_out, tout, vin, vout,
                                                 c1, c2);
            return;
        }
        bid -= F.nfs_wg;
        nwg -= F.nfs_wg;
    }
    int64_t slice = xcd_tile(bid, nwg);
    if constexpr (FK::fwd) {
        slice += F.rot;   // rot < nslice == nwg
        if (slice >= nwg) slice -= nwg;
    }
    if (slice >= nslice) return;
    if constexpr (SH) {
        if (H.skip[slice] >= 0) return;   // its rows are leading rows
    }
    const int q = lane / B, r = lane % B;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int32_t *cp = col + p0 + q;
    const VT *vp = val + p0 * BB;
    const int64_t cell = slice * CPS + q;
    const int64_t row = (cell < ncell ? cell : ncell - 1) * B + r;
    double pre_v[NV];
    {
        const double *pv = vout ? vin : x;
#pragma unroll
        for (int v = 0; v < NV; ++v) pre_v[v] = w == 0 ? pv[row * NV + v] : 0.0;
    }
    // this lane's range of fused extra entries, fetched with the epilogue
    // operands so that the epilogue starts with them at hand
    int32_t fe0 = 0, fe1 = 0;
    if constexpr (FK::fwd) {
        const int32_t blk = F.slice_tab[slice];
        if (blk >= 0 && w == 0) {
            fe0 = F.lane_ptr[(int64_t)blk * 65 + lane];
            fe1 = F.lane_ptr[(int64_t)blk * 65 + lane + 1];
        }
    }
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    d2 *mine = &xs[w][0][0];
    const int nmine = width > w ? (width - w + NW - 1) / NW : 0;
    for (int g = 0; g < nmine; g += S) {
        int32_t c[S];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int i = g + u < nmine ? g + u : nmine - 1;
            c[u] = cp[(int64_t)(w + i * NW) * CPS];
        }
        // the values (the HBM stream) go out with the columns; the x chunks
        // follow once the columns land.  Nothing may move above the x loads
        // (converting a value early would wait for every value first).
        CT raw[S][NC];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int i = g + u < nmine ? g + u : nmine - 1;
            const VT *base = vp + (int64_t)(w + i * NW) * CPS * BB;
#pragma unroll
            for (int kc = 0; kc < NC; ++kc) {
                const CT *pc = reinterpret_cast<const CT *>(base + ((int64_t)kc * 64 + lane) * V);
                raw[u][kc] = NT ? __builtin_nontemporal_load(pc) : *pc;
            }
        }
        d2 xr[S][CPL];
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int i = 0; i < CPL; ++i)
                xr[u][i] = *reinterpret_cast<const d2 *>(x + (int64_t)c[u] * (B * NV) +
                                                         (r * CPL + i) * 2);
        __builtin_amdgcn_sched_barrier(0);
        wave_lds_sync();  // the previous group's reads are issued
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int i = 0; i < CPL; ++i) mine[u * CPS * XS + q * XS + r * CPL + i] = xr[u][i];
        wave_lds_sync();
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const bool live = g + u < nmine;
            const d2 *xb = mine + u * CPS * XS + q * XS;
#pragma unroll
            for (int kc = 0; kc < NC; ++kc) {
                // the rows of this value chunk first: their reads in flight
                // together before the FMAs
                d2 xx[V * (NV / 2)];
#pragma unroll
                for (int i = 0; i < V * (NV / 2); ++i) xx[i] = xb[kc * V * (NV / 2) + i];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const double ak = live ? (double)raw[u][kc][e] : 0.0;
#pragma unroll
                    for (int vp2 = 0; vp2 < NV / 2; ++vp2) {
                        acc[2 * vp2] = fma(ak, xx[e * (NV / 2) + vp2].x, acc[2 * vp2]);
                        acc[2 * vp2 + 1] = fma(ak, xx[e * (NV / 2) + vp2].y, acc[2 * vp2 + 1]);
                    }
                }
            }
        }
    }
    if constexpr (NW > 1) {
        if (w > 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) part[w - 1][v][lane] = acc[v];
        }
        __syncthreads();
        if (w > 0) return;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += part[i][v][lane];
    }
    if (cell >= ncell) return;
    if constexpr (FK::fwd) {
        // this row's compressed extra entries times the column sums of x, in
        // groups of 8: the columns and values of a group, then its sums, are
        // in flight together; products added in entry order
        constexpr int G = 8, VS = FK::cplx ? 2 : 1;
        for (int32_t e0 = fe0; e0 < fe1; e0 += G) {
            int32_t qc[G];
            double a[G][VS];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int32_t e = e0 + u < fe1 ? e0 + u : fe1 - 1;
                qc[u] = F.ecol[e];
#pragma unroll
                for (int i = 0; i < VS; ++i) a[u][i] = F.eval[(int64_t)e * VS + i];
            }
            double sv[G][NV];
#pragma unroll
            for (int u = 0; u < G; ++u)
#pragma unroll
                for (int v = 0; v < NV; ++v) sv[u][v] = F.s_in[(int64_t)qc[u] * NV + v];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (e0 + u >= fe1) break;
                if constexpr (FK::cplx) {
                    acc[0] = fma(a[u][0], sv[u][0], fma(-a[u][1], sv[u][1], acc[0]));
                    acc[1] = fma(a[u][0], sv[u][1], fma(a[u][1], sv[u][0], acc[1]));
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] = fma(a[u][0], sv[u][v], acc[v]);
                }
            }
        }
    }
    // the vout stores first, behind one wait for the epilogue operands (a
    // store between two such waits would be waited for as well)
    if (vout) {
#pragma unroll
        for (int v = 0; v < NV; ++v) vout[row * NV + v] = fma(c2, c1 * acc[v], pre_v[v]);
    }
    if (tout) {
#pragma unroll
        for (int v = 0; v < NV; ++v) tout[row * NV + v] = c1 * acc[v];
    }
}

}  // namespace hymor

namespace hymor {

typedef double v4d __attribute__((ext_vector_type(4)));

// Batched Taylor term on the FP64 matrix cores.  One wavefront per cell;
// per block slot and k-chunk of 4 the block product Y(b x 16 rhs) +=
// A_blk(b x 4) * X(4 x 16 rhs) is one v_mfma_f64_16x16x4f64 per 16-rhs tile
// (rows >= b are zero padding).  Operand layout of that instruction on
// gfx950 (measured): A[m][k] from lane m + 16k, B[k][n] from lane n + 16k,
// D[m][n] in lane n + 16*(m % 4), register m / 4.  So lane (m = lane % 16,
// kk = lane / 16) supplies block value A[m][4kc+kk] and x value
// X[4kc+kk][16nt + m]: every lane loads one value per operand instead of a
// whole block row, and the x loads are 4 x 128-byte lines per instruction.
template <int B, typename VT, int R>
__global__ __launch_bounds__(256) void k_spmm_mfma(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int NT = R / 16;
    constexpr int KC = B / 4;
    constexpr int V = Vpl<VT>::value;
    // slots whose operands are loaded together before their MFMAs: enough
    // loads in flight to cover memory latency within one wave
    constexpr int U = NT == 1 ? 9 : (NT == 2 ? 4 : 2);
    static_assert(R % 16 == 0 && B % 4 == 0, "unsupported B/R");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = __builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int m = lane & 15, kk = lane >> 4;
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    v4d acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < width; t0 += U) {
        double a[U][KC], bx[U][KC][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u;
            if (t < width) {
                const int64_t c = col[p0 + (int64_t)t * CPS + q];
                const VT *sb = val + (p0 + (int64_t)t * CPS) * BB;
#pragma unroll
                for (int kc = 0; kc < KC; ++kc) {
                    const int k = 4 * kc + kk;
                    a[u][kc] = m < B ? (double)sb[((int64_t)(k / V) * 64 + q * B + m) * V + k % V]
                                     : 0.0;
                    const double *xp = x + (c * B + k) * R + m;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) bx[u][kc][nt] = xp[nt * 16];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (t0 + u < width) {
#pragma unroll
                for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][kc], bx[u][kc][nt],
                                                                       acc[nt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
        for (int r = 0; r < KC; ++r) {   // output rows kk + 4r < B
            const int64_t idx = (cell * B + kk + 4 * r) * R + nt * 16 + m;
            double y = acc[nt][r];
            const double t = c1 * y;
            if (tout) tout[idx] = t;
            if (vout) vout[idx] = fma(c2, t, vin[idx]);
        }
    }
}

}  // namespace hymor

namespace hymor {

// k_spmm_mfma with more memory work in flight per wave: CPW cells per wave
// (consecutive cells of one slice), their U-slot operand loads issued
// together, and the epilogue's vin reads issued at wave start.  Same
// arithmetic, operand layout and summation order as k_spmm_mfma.
template <int B, typename VT, int R, int U, int CPW>
__global__ __launch_bounds__(256) void k_spmm_mfma_mc(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int NT = R / 16;
    constexpr int KC = B / 4;
    constexpr int V = Vpl<VT>::value;
    static_assert(R % 16 == 0 && B % 4 == 0 && CPS % CPW == 0, "unsupported B/R/CPW");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell0 = (int64_t)__builtin_amdgcn_readfirstlane(
                              (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6))) *
                          CPW;
    if (cell0 >= ncell) return;
    const int m = lane & 15, kk = lane >> 4;
    const int64_t slice = cell0 / CPS, q0 = cell0 - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    int ncw = (int)(ncell - cell0 < CPW ? ncell - cell0 : CPW);
    // epilogue operands first: independent of everything below
    double pre_v[CPW][NT][KC];
#pragma unroll
    for (int ci = 0; ci < CPW; ++ci)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < KC; ++r) {
                const int64_t idx = ((cell0 + ci) * B + kk + 4 * r) * R + nt * 16 + m;
                const bool ok = ci < ncw;
                pre_v[ci][nt][r] = (ok && vout) ? vin[idx] : 0.0;
            }
    v4d acc[CPW][NT];
#pragma unroll
    for (int ci = 0; ci < CPW; ++ci)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[ci][nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < width; t0 += U) {
        double a[U][CPW][KC], bx[U][CPW][KC][NT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u;
            if (t < width) {
                const VT *sb = val + (p0 + (int64_t)t * CPS) * BB;
#pragma unroll
                for (int ci = 0; ci < CPW; ++ci) {
                    const int64_t q = q0 + ci;
                    const int64_t c = col[p0 + (int64_t)t * CPS + q];
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc) {
                        const int k = 4 * kc + kk;
                        a[u][ci][kc] =
                            m < B ? (double)sb[((int64_t)(k / V) * 64 + q * B + m) * V + k % V] : 0.0;
                        const double *xp = x + (c * B + k) * R + m;
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt) bx[u][ci][kc][nt] = xp[nt * 16];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (t0 + u < width) {
#pragma unroll
                for (int ci = 0; ci < CPW; ++ci)
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[ci][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[u][ci][kc], bx[u][ci][kc][nt], acc[ci][nt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int ci = 0; ci < CPW; ++ci) {
        if (ci >= ncw) break;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < KC; ++r) {
                const int64_t idx = ((cell0 + ci) * B + kk + 4 * r) * R + nt * 16 + m;
                const double y = acc[ci][nt][r];
                const double t = c1 * y;
                if (tout) tout[idx] = t;
                if (vout) vout[idx] = fma(c2, t, pre_v[ci][nt][r]);
            }
    }
}

// k_spmm_mfma with every operand load of a group of W slots issued before
// the first product and no branch around any load, so a wave waits for
// memory once per group instead of once per operand.  Rows m >= B of the
// 16-row tile read a duplicate row (m mod B) and slots past the row's width
// read the row's last slot; both are zeroed by a select before the MFMA.
// Block values are streamed with non-temporal loads (read once per launch);
// x, re-read by the neighbouring cells, stays cacheable.  Same arithmetic,
// operand layout and summation order as k_spmm_mfma (a zeroed slot adds
// exact zeros).
template <int B, typename VT, int R, int W>
__global__ __launch_bounds__(256) void k_spmm_mfma_w(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int NT = R / 16;
    constexpr int KC = B / 4;
    constexpr int V = Vpl<VT>::value;
    static_assert(R % 16 == 0 && (B == 4 || B == 8), "unsupported B/R");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int m = lane & 15, kk = lane >> 4;
    const int mr = m & (B - 1);
    const bool live_row = m < B;
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    // epilogue operands, read from a valid array when the caller passes none
    const double *pv = vout ? vin : x;
    double pre_v[NT][KC];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < KC; ++r) {
            const int64_t idx = (cell * B + kk + 4 * r) * R + nt * 16 + m;
            pre_v[nt][r] = pv[idx];
        }
    v4d acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = v4d{0.0, 0.0, 0.0, 0.0};
    static_assert(W <= 64, "one column per lane");
    for (int t0 = 0; t0 < width; t0 += W) {
        // the group's block columns with one vector load (lane u < W reads
        // slot t0 + u), broadcast by readlane: one wait instead of one per
        // scalar load, whose out-of-order returns force a full wait each
        const int tl = t0 + (lane < W ? lane : 0);
        const int32_t cl = col[p0 + (int64_t)(tl < width ? tl : width - 1) * CPS + q];
        int32_t c[W];
#pragma unroll
        for (int u = 0; u < W; ++u) c[u] = __builtin_amdgcn_readlane(cl, u);
        VT araw[W][KC];
        double bx[W][KC][NT];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int t = t0 + u < width ? t0 + u : width - 1;
            const VT *sb = val + (p0 + (int64_t)t * CPS) * BB;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const int k = 4 * kc + kk;
                araw[u][kc] = __builtin_nontemporal_load(
                    sb + ((int64_t)(k / V) * 64 + q * B + mr) * V + k % V);
                const double *xp = x + ((int64_t)c[u] * B + k) * R + m;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bx[u][kc][nt] = xp[nt * 16];
            }
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const bool live = live_row && (t0 + u < width);
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const double a = live ? (double)araw[u][kc] : 0.0;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bx[u][kc][nt], acc[nt], 0,
                                                                   0, 0);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < KC; ++r) {
            const int64_t idx = (cell * B + kk + 4 * r) * R + nt * 16 + m;
            const double y = acc[nt][r];
            const double t = c1 * y;
            if (tout) tout[idx] = t;
            if (vout) vout[idx] = fma(c2, t, pre_v[nt][r]);
        }
}

// Block values in the tile kernel's order: inside each block, value (m, k)
// with k = 4*kc + kk at (m*4 + kk)*KC + kc (a lane's KC values adjacent).
// Block values placed in the sliced, lane-interleaved layout on the device:
// padded block d takes block src[d] of raw (b x b, row-major; read
// transposed for the adjoint) times scale, as VT, value k of lane
// L = q*b + r of a slot at ((k/V)*64 + L)*V + k%V with V = 16/sizeof(VT);
// src[d] < 0 is a zero padding block.  A slot holds 64*b values (its CPS
// blocks), so slot g, lane group q is padded block g*CPS + q.
template <typename VT>
__global__ void k_slice_vals(const double *__restrict__ raw, const int64_t *__restrict__ src,
                             VT *__restrict__ dst, int64_t total, int b, int transpose,
                             double scale) {
    constexpr int V = Vpl<VT>::value;
    const int bb = b * b, cps = 64 / b, slot_len = 64 * b;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t slot = i / slot_len;
        const int j = (int)(i - slot * slot_len);
        const int kq = j / (64 * V), rem = j - kq * 64 * V;
        const int lane = rem / V, k = kq * V + rem % V;
        const int q = lane / b, r = lane - q * b;
        const int64_t sb = src[slot * cps + q];
        double v = 0.0;
        if (sb >= 0) v = raw[sb * bb + (transpose ? k * b + r : r * b + k)] * scale;
        dst[i] = (VT)v;
    }
}

// FP64 values in the two-per-lane interleave (value k of lane L of a slot
// at ((k/2)*64 + L)*2 + k%2) rewritten as FP32 in the four-per-lane
// interleave ((k/4)*64 + L)*4 + k%4: the layout an FP32 upload builds.  A
// slot holds 64*b values with b a multiple of 4, so slots are whole 256-value
// runs and the index map needs no block size.
__global__ void k_vals_to_f32(const double *__restrict__ src, float *__restrict__ dst,
                              int64_t total) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = 4 * (i >> 8) + (i & 3);
        const int64_t lane = (i & 255) >> 2;
        dst[i] = (float)src[((k >> 1) * 64 + lane) * 2 + (k & 1)];
    }
}

// Rewritten on the device from the lane-interleaved layout of the same
// slots; one thread per value.
template <typename VT>
__global__ void k_tile_vals(const VT *__restrict__ src, VT *__restrict__ dst, int64_t total,
                            int b) {
    constexpr int V = Vpl<VT>::value;
    const int cps = 64 / b, bb = b * b, kc_ = b >> 2;
    const int64_t slot_len = (int64_t)cps * bb;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t slot = i / slot_len;
        const int within = (int)(i - slot * slot_len);
        const int q = within / bb, rem = within - q * bb;
        const int kc = rem % kc_, tmp = rem / kc_;
        const int kk = tmp & 3, m = tmp >> 2;
        const int k = kk + 4 * kc;
        dst[i] = src[slot * slot_len + ((int64_t)(k / V) * 64 + q * b + m) * V + k % V];
    }
}

template <typename VT, int KC>
struct TileVals;
template <>
struct TileVals<float, 2> {
    typedef float __attribute__((ext_vector_type(2))) type;  // f2
};
template <>
struct TileVals<double, 2> {
    typedef d2 type;
};
template <typename VT>
struct TileVals<VT, 1> {
    typedef VT type;
};

template <int KC>
struct TileX;
template <>
struct TileX<2> {
    typedef d2 type;
};
template <>
struct TileX<1> {
    typedef double type;
};

typedef float f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double lane_elem(double v, int) { return v; }
__device__ __forceinline__ double lane_elem(float v, int) { return (double)v; }
__device__ __forceinline__ double lane_elem(const d2 &v, int i) { return v[i]; }
__device__ __forceinline__ double lane_elem(const f2 &v, int i) { return (double)v[i]; }

// Batched Taylor term on the FP64 matrix cores with x, vin, tout, vout
// in the tile layout and block values in TileVals order: per slot one
// value load and NT x loads per lane (16 bytes each at B = 8), the
// operands of KC x NT v_mfma_f64_16x16x4f64; the epilogue reads and writes
// 16 bytes per lane.  One cell per wave; the group's W block columns come
// from one vector load.  For R = 8 lanes n >= 8 duplicate columns n - 8 and
// their results are dropped; rows m >= B read row m mod B and are zeroed.
// Summation order: slots ascending, k-chunks ascending (as k_spmm_mfma).
template <int B, typename VT, int R, int W>
__global__ __launch_bounds__(256) void k_spmm_tile(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ valt, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int CPS = 64 / B;
    constexpr int BB = B * B;
    constexpr int WN = R < 16 ? R : 16;
    constexpr int NT = R / WN;
    constexpr int KC = B / 4;
    typedef typename TileVals<VT, KC>::type AT;
    typedef typename TileX<KC>::type XT;
    static_assert(B == 4 || B == 8, "block size 4 or 8");
    static_assert(R == 8 || R == 16 || R == 32 || R == 64, "width 8..64");
    static_assert(W <= 64, "one column per lane");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int n = lane & 15, kk = lane >> 4;
    const int nr = n & (WN - 1);                  // column this lane reads
    const int mr = n & (B - 1);                   // block row this lane reads
    const bool live_row = n < B;
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int64_t cbase = cell * B * R;
    // epilogue operands (a valid array stands in for a missing one)
    const double *pv = vout ? vin : x;
    XT pre_v[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int64_t o = cbase + ((int64_t)(kk * NT + nt) * WN + nr) * KC;
        pre_v[nt] = *reinterpret_cast<const XT *>(pv + o);
    }
    v4d acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < width; t0 += W) {
        const int tl = t0 + (lane < W ? lane : 0);
        const int32_t cl = col[p0 + (int64_t)(tl < width ? tl : width - 1) * CPS + q];
        int32_t c[W];
#pragma unroll
        for (int u = 0; u < W; ++u) c[u] = __builtin_amdgcn_readlane(cl, u);
        AT araw[W];
        XT bx[W][NT];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int t = t0 + u < width ? t0 + u : width - 1;
            const VT *sb = valt + (p0 + (int64_t)t * CPS + q) * BB + (mr * 4 + kk) * KC;
            araw[u] = __builtin_nontemporal_load(reinterpret_cast<const AT *>(sb));
            const double *xp = x + (int64_t)c[u] * B * R + ((int64_t)kk * NT * WN + nr) * KC;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bx[u][nt] = *reinterpret_cast<const XT *>(xp + (int64_t)nt * WN * KC);
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const bool live = live_row && (t0 + u < width);
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const double a = live ? lane_elem(araw[u], kc) : 0.0;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, lane_elem(bx[u][nt], kc),
                                                                   acc[nt], 0, 0, 0);
            }
        }
    }
    if (n >= WN) return;
    // the vout stores first, behind one wait for the epilogue operands (a
    // store between two such waits would be waited for as well)
    if (vout) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            XT vv;
#pragma unroll
            for (int r = 0; r < KC; ++r) {
                const double w = fma(c2, c1 * acc[nt][r], lane_elem(pre_v[nt], r));
                if constexpr (KC == 1) vv = w;
                else vv[r] = w;
            }
            *reinterpret_cast<XT *>(vout + cbase + ((int64_t)(kk * NT + nt) * WN + n) * KC) = vv;
        }
    }
    if (tout) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            XT tv;
#pragma unroll
            for (int r = 0; r < KC; ++r) {
                if constexpr (KC == 1) tv = c1 * acc[nt][r];
                else tv[r] = c1 * acc[nt][r];
            }
            *reinterpret_cast<XT *>(tout + cbase + ((int64_t)(kk * NT + nt) * WN + n) * KC) = tv;
        }
    }
}

// k_spmm_tile for b = 8 with the two cells of a pair (2p, 2p + 1) in one
// wave: rows 0-7 of each v_mfma_f64_16x16x4f64 are cell 2p's block rows,
// rows 8-15 cell 2p + 1's (k_spmm_tile leaves them as padding).  One
// product (per k-chunk and 16-vector tile) serves a block column of either
// cell: the pair's columns are the union of its cells' (bsr_pair_map;
// 12 for two neighbours of a 3x3 stencil, against 18 products and 18 x
// loads for the two cells apart), and a cell without a block in a column
// contributes zero rows without loading values.  Lane (n, kk): A operand
// row n of the pair (cell n / 8, block row n % 8), B operand x (variable
// 4 kc + kk, vector 16 nt + n) of the column, D rows kk, 4 + kk (cell 2p)
// and 8 + kk, 12 + kk (cell 2p + 1).  Summation order: the pair's columns
// ascending, k-chunks ascending.
template <typename VT, int R, int W>
__global__ __launch_bounds__(256) void k_spmm_pair(
    const int64_t *__restrict__ slice_ptr, const int64_t *__restrict__ pptr,
    const int32_t *__restrict__ pcol, const int32_t *__restrict__ pslot,
    const VT *__restrict__ valt, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int B = 8, CPS = 8, BB = 64, KC = 2, WN = 16;
    constexpr int NT = R / WN;
    typedef typename TileVals<VT, KC>::type AT;
    typedef d2 XT;
    static_assert(R == 16 || R == 32 || R == 64, "width 16..64");
    static_assert(W <= 64, "one column per lane");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t pair = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    const int64_t cell0 = 2 * pair;
    if (cell0 >= ncell) return;
    const bool has1 = cell0 + 1 < ncell;
    const int n = lane & 15, kk = lane >> 4;
    const int half = n >> 3;                     // the cell whose row this lane feeds
    const int mr = n & 7;                        // its block row
    const int64_t slice = cell0 / CPS;           // a pair never straddles slices
    const int q = (int)(cell0 - slice * CPS) + half;
    const int64_t p0 = slice_ptr[slice];
    const int64_t e0 = pptr[pair], e1 = pptr[pair + 1];
    const int ncol = (int)(e1 - e0);
    const int la = (mr * 4 + kk) * KC;           // lane's value offset in a block
    const int lx = (kk * NT * WN + n) * KC;      // lane's x offset in a cell's batch
    const double *pv = vout ? vin : x;
    // epilogue operands of both cells (the second cell's only when present)
    const int64_t cb0 = cell0 * B * R, cb1 = (has1 ? cell0 + 1 : cell0) * B * R;
    XT pre0[NT], pre1[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int64_t o = (int64_t)(kk * NT + nt) * WN * KC + n * KC;
        pre0[nt] = *reinterpret_cast<const XT *>(pv + cb0 + o);
        pre1[nt] = *reinterpret_cast<const XT *>(pv + cb1 + o);
    }
    v4d acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < ncol; t0 += W) {
        const int tl = t0 + (lane < W ? lane : 0);
        const int64_t ei = e0 + (tl < ncol ? tl : ncol - 1);
        const int32_t cl = pcol[ei], sl = pslot[ei];
        int32_t c[W], sp[W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            c[u] = __builtin_amdgcn_readlane(cl, u);
            sp[u] = __builtin_amdgcn_readlane(sl, u);
        }
        AT araw[W];
        XT bx[W][NT];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            // this lane's cell's slot for the column (none: zero rows)
            const int t = half ? (int)(((uint32_t)sp[u] >> 16) & 0xFFFF) : (sp[u] & 0xFFFF);
            const bool have = t != 0xFFFF && t0 + u < ncol;
            AT a{};
            if (have)
                a = __builtin_nontemporal_load(reinterpret_cast<const AT *>(
                    valt + (p0 + (int64_t)t * CPS + q) * BB + la));
            araw[u] = a;
            const double *xp = x + (int64_t)c[u] * B * R + lx;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bx[u][nt] = *reinterpret_cast<const XT *>(xp + (int64_t)nt * WN * KC);
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const double a = lane_elem(araw[u], kc);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, lane_elem(bx[u][nt], kc),
                                                                   acc[nt], 0, 0, 0);
            }
        }
    }
    // vout first (one wait for the epilogue operands), then tout
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 1 && !has1) break;
        const int64_t cb = h ? cb1 : cb0;
        if (vout) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const XT pv2 = h ? pre1[nt] : pre0[nt];
                XT vv;
                vv[0] = fma(c2, c1 * acc[nt][2 * h], pv2[0]);
                vv[1] = fma(c2, c1 * acc[nt][2 * h + 1], pv2[1]);
                *reinterpret_cast<XT *>(vout + cb + (int64_t)(kk * NT + nt) * WN * KC + n * KC) = vv;
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 1 && !has1) break;
        const int64_t cb = h ? cb1 : cb0;
        if (tout) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                XT tv;
                tv[0] = c1 * acc[nt][2 * h];
                tv[1] = c1 * acc[nt][2 * h + 1];
                *reinterpret_cast<XT *>(tout + cb + (int64_t)(kk * NT + nt) * WN * KC + n * KC) = tv;
            }
        }
    }
}

// ------------------------------------------------------- block Gram-Schmidt
// Partial products of V^T W for a block of s <= 16 new Krylov vectors W
// against ncol basis columns V, on the FP64 matrix cores: one
// v_mfma_f64_16x16x4f64 per 16 basis columns x 16 block columns x 4 rows,
// the 4-row index being a dummy summation index.  Lane (m = lane % 16,
// kk = lane / 16) supplies V[row, c0 + m] and W[row, m] for rows
// r + 4 kk + i (i = 0..3, adjacent: two 16-byte loads each when ALIGNED).
// Workgroup g covers rows [g*rows_per, (g+1)*rows_per) (rows_per a multiple
// of 64), each of its four waves a quarter of them; a wave computes CT
// 16-column tiles of V at a time against its rows of W, so W is read once
// per CT tiles and every wave has work whatever ncol is.  The four waves'
// products meet in LDS in wave order; partial [(g*ctiles*16 + c)*16 + n]
// holds the workgroup's sum for basis column c, block column n.
template <bool ALIGNED, int CT>
__global__ __launch_bounds__(256) void k_vtw_mfma(const double *__restrict__ V, int64_t ldv,
                                                  int ncol, const double *__restrict__ W,
                                                  int64_t ldw, int s, int64_t n,
                                                  int64_t rows_per, int ctiles,
                                                  double *__restrict__ partial) {
    __shared__ double red[3][CT][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kk = lane >> 4;
    const int64_t g = blockIdx.x;
    const int64_t rw = rows_per / 4;
    const int64_t r0 = g * rows_per + wave * rw;
    const int64_t r1 = r0 + rw < n ? r0 + rw : n;
    const int64_t rfull = r1 > r0 ? r0 + ((r1 - r0) / 16) * 16 : r0;
    const bool wv = m < s;
    const double *wp = W + (int64_t)(wv ? m : 0) * ldw;
    for (int ct0 = 0; ct0 < ctiles; ct0 += CT) {
        v4d acc[CT];
        const double *vp[CT];
        bool cv[CT];
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            acc[u] = v4d{0.0, 0.0, 0.0, 0.0};
            const int c = (ct0 + u) * 16 + m;
            cv[u] = ct0 + u < ctiles && c < ncol;
            vp[u] = V + (int64_t)(cv[u] ? c : 0) * ldv;
        }
        for (int64_t r = r0; r < rfull; r += 16) {
            const int64_t rr = r + 4 * kk;
            double b[4];
            if constexpr (ALIGNED) {
                const d2 b0 = *reinterpret_cast<const d2 *>(wp + rr);
                const d2 b1 = *reinterpret_cast<const d2 *>(wp + rr + 2);
                b[0] = b0.x; b[1] = b0.y; b[2] = b1.x; b[3] = b1.y;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = wp[rr + i];
            }
            double a[CT][4];
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                if constexpr (ALIGNED) {
                    const d2 a0 = *reinterpret_cast<const d2 *>(vp[u] + rr);
                    const d2 a1 = *reinterpret_cast<const d2 *>(vp[u] + rr + 2);
                    a[u][0] = a0.x; a[u][1] = a0.y; a[u][2] = a1.x; a[u][3] = a1.y;
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) a[u][i] = vp[u][rr + i];
                }
            }
#pragma unroll
            for (int u = 0; u < CT; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(cv[u] ? a[u][i] : 0.0,
                                                                  wv ? b[i] : 0.0, acc[u], 0, 0, 0);
        }
        // tail rows (< 16, only in the last wave with rows): clamped reads, zeroed
        for (int64_t r = rfull; r < r1; r += 4) {
            const int64_t row = r + kk;
            const bool ok = row < r1;
            const int64_t rc = ok ? row : r1 - 1;
            const double b = (wv && ok) ? wp[rc] : 0.0;
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                const double a = (cv[u] && ok) ? vp[u][rc] : 0.0;
                acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u], 0, 0, 0);
            }
        }
        if (wave > 0) {
#pragma unroll
            for (int u = 0; u < CT; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) red[wave - 1][u][r][lane] = acc[u][r];
        }
        __syncthreads();
        if (wave == 0) {
            // D[c][n] sits in lane n + 16 (c % 4), register c / 4
#pragma unroll
            for (int u = 0; u < CT; ++u) {
                if (ct0 + u >= ctiles) break;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    partial[((g * ctiles + ct0 + u) * 16 + 4 * r + kk) * 16 + m] =
                        ((acc[u][r] + red[0][u][r][lane]) + red[1][u][r][lane]) +
                        red[2][u][r][lane];
            }
        }
        __syncthreads();
    }
}

// h[nn*ncol + c] = sum over workgroups of the partials above: one wave per
// entry, lane l summing groups l, l+64, ... in order, then a fixed
// butterfly across lanes (deterministic)
__global__ __launch_bounds__(256) void k_reduce_vtw(const double *__restrict__ partial,
                                                    int64_t ngroups, int ctiles, int ncol, int s,
                                                    double *__restrict__ h) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= ncol * s) return;
    const int c = i % ncol, nn = i / ncol;
    double acc = 0.0;
    for (int64_t g = lane; g < ngroups; g += 64) acc += partial[((g * ctiles) * 16 + c) * 16 + nn];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) h[i] = acc;
}

// W[:, :s] -= V[:, :ncol] h on the FP64 matrix cores, formed transposed:
// (V h)^T = h^T V^T, one v_mfma_f64_16x16x4f64 per 16 block columns x 16
// rows x 4 basis columns with A[m][k] = h[c0 + k, m] (staged in LDS; h is
// column-major, h[m * ncol + c]) and B[k][n] = V[r + n, c0 + k].  Lane
// (n, kk) loads V[r + n] of basis column c0 + kk -- 16 consecutive rows per
// column, 128-byte runs -- and afterwards holds (V h)[r + n, 4 reg + kk] in
// register reg, so the read-modify-write of W runs along rows as well.  One
// wave per 16-row tile (grid-stride); the loads of KU groups of 4 basis
// columns are issued before their products.
template <int KU>
__global__ __launch_bounds__(256) void k_sub_vh_mfma(const double *__restrict__ V, int64_t ldv,
                                                     int ncol, const double *__restrict__ h,
                                                     double *__restrict__ W, int64_t ldw, int s,
                                                     int64_t n) {
    extern __shared__ double hs[];
    for (int i = threadIdx.x; i < ncol * s; i += blockDim.x) hs[i] = h[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int nl = lane & 15, kk = lane >> 4;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t ntiles = (n + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
         tile < ntiles; tile += nwaves) {
        const int64_t row = tile * 16 + nl;
        const bool rok = row < n;
        v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
        for (int c0 = 0; c0 < ncol; c0 += 4 * KU) {
            double a[KU], b[KU];
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const int c = c0 + 4 * u + kk;
                const bool cok = c < ncol;
                b[u] = (cok && rok) ? V[(int64_t)c * ldv + row] : 0.0;
                a[u] = (cok && nl < s) ? hs[nl * ncol + c] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < KU; ++u)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
        }
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int m = 4 * reg + kk;
            if (m < s && rok) W[(int64_t)m * ldw + row] -= acc[reg];
        }
    }
}

// ------------------------------------------------------ 4x4x4 matrix cores
// v_mfma_f64_4x4x4f64 on gfx950 (measured): four independent 4x4x4
// products per instruction; block blk takes A[blk][m][k] from lane
// m + 4 blk + 16 k and B[blk][k][n] from lane n + 4 blk + 16 k, and leaves
// D[blk][m][n] in lane n + 4 blk + 16 m.  For b = 8 an 8 x R output tile
// splits into 4 x 4 tiles with no idle rows, where the 16 x 16 x 4 form
// leaves rows 8..15 idle.
//
// Value order ("quad") inside each 8 x 8 block: lane group g = m + 4 k
// (m, k < 4) holds A[4 rq + m][4 kc + k] for (kc, rq) at g*4 + kc*2 + rq, so
// a lane's four A operands are one 16-byte (FP32) or 32-byte (FP64) load,
// and the two of one k-chunk one 8- or 16-byte load.
template <typename VT>
__global__ void k_quad_vals(const VT *__restrict__ src, VT *__restrict__ dst, int64_t total) {
    constexpr int V = Vpl<VT>::value;
    constexpr int CPS = 8;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t blkidx = i >> 6;
        const int within = (int)(i & 63);
        const int64_t slot = blkidx / CPS;
        const int q = (int)(blkidx - slot * CPS);
        const int g = within >> 2, kc = (within >> 1) & 1, rq = within & 1;
        const int m = g & 3, k = g >> 2;
        const int row = 4 * rq + m, colk = 4 * kc + k;
        dst[i] = src[slot * CPS * 64 + ((int64_t)(colk / V) * 64 + q * 8 + row) * V + colk % V];
    }
}

template <typename VT>
struct Quad4;
template <>
struct Quad4<float> {
    typedef f4 raw;
    typedef f2 half;
    __device__ __forceinline__ static raw load(const float *p) {
        return __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p));
    }
    __device__ __forceinline__ static half load_half(const float *p) {
        return __builtin_nontemporal_load(reinterpret_cast<const f2 *>(p));
    }
    __device__ __forceinline__ static double get(const raw &r, int i) { return (double)r[i]; }
    __device__ __forceinline__ static double get_half(const half &r, int i) { return (double)r[i]; }
};
template <>
struct Quad4<double> {
    struct raw {
        d2 lo, hi;
    };
    __device__ __forceinline__ static raw load(const double *p) {
        return raw{__builtin_nontemporal_load(reinterpret_cast<const d2 *>(p)),
                   __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p + 2))};
    }
    typedef d2 half;
    __device__ __forceinline__ static half load_half(const double *p) {
        return __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p));
    }
    __device__ __forceinline__ static double get(const raw &r, int i) {
        return i < 2 ? r.lo[i] : r.hi[i - 2];
    }
    __device__ __forceinline__ static double get_half(const half &r, int i) { return r[i]; }
};

// Batched Taylor term for b = 8 on v_mfma_f64_4x4x4f64, vectors in the tile
// order (R doubles per row, R = 8..64) and values in the quad order.
//   R >= 16: block blk = column quarter of a 16-column tile; per slot and
//     tile 4 instructions (rq, kc), accumulators acc[nt][rq]; the lane's
//     outputs (rows kk and kk + 4) are one 16-byte store.
//   R = 8: blk = (column half nq, k-chunk kc); per slot 2 instructions (rq),
//     the two k-chunks' partial sums added across lanes l and l ^ 8 at the
//     end; lanes with kc = 0 store.
// One cell per wave; the group's W block columns from one vector load.
// Summation order: slots ascending, k-chunks in order (R >= 16) or the two
// k-chunk sums added last (R = 8).
template <typename VT, int R, int W>
__global__ __launch_bounds__(256) void k_spmm_quad(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ valq, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int B = 8, CPS = 8, BB = 64;
    constexpr int NT = R >= 16 ? R / 16 : 1;
    constexpr int WN = R >= 16 ? 16 : 8;
    static_assert(R == 8 || R == 16 || R == 32 || R == 64, "width 8..64");
    static_assert(W <= 64, "one column per lane");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int m4 = lane & 3, k4 = lane >> 4;
    const int g = m4 + 4 * k4;
    const int kcl = (lane >> 3) & 1;            // R = 8: this lane's k-chunk
    const int jl = R >= 16 ? (lane & 15) : (lane & 7);
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int64_t cbase = cell * B * R;
    const double *pv = vout ? vin : x;
    d2 pre_v[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int64_t o = cbase + ((int64_t)(k4 * NT + nt) * WN + jl) * 2;
        pre_v[nt] = *reinterpret_cast<const d2 *>(pv + o);
    }
    double acc[NT][2];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt][0] = acc[nt][1] = 0.0;
    for (int t0 = 0; t0 < width; t0 += W) {
        const int tl = t0 + (lane < W ? lane : 0);
        const int32_t cl = col[p0 + (int64_t)(tl < width ? tl : width - 1) * CPS + q];
        int32_t c[W];
#pragma unroll
        for (int u = 0; u < W; ++u) c[u] = __builtin_amdgcn_readlane(cl, u);
        // R >= 16: all four operands (kc, rq); R = 8: the two of this lane's kc
        typedef typename std::conditional<(R >= 16), typename Quad4<VT>::raw,
                                          typename Quad4<VT>::half>::type AR;
        AR a[W];
        d2 bx[W][NT];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int t = t0 + u < width ? t0 + u : width - 1;
            const VT *ap = valq + (p0 + (int64_t)t * CPS + q) * BB + g * 4;
            if constexpr (R >= 16) a[u] = Quad4<VT>::load(ap);
            else a[u] = Quad4<VT>::load_half(ap + kcl * 2);
            const double *xp = x + (int64_t)c[u] * B * R;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                bx[u][nt] = *reinterpret_cast<const d2 *>(
                    xp + ((int64_t)(k4 * NT + nt) * WN + jl) * 2);
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const bool live = t0 + u < width;
            if constexpr (R >= 16) {
#pragma unroll
                for (int rq = 0; rq < 2; ++rq)
#pragma unroll
                    for (int kc = 0; kc < 2; ++kc) {
                        const double av = live ? Quad4<VT>::get(a[u], kc * 2 + rq) : 0.0;
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[nt][rq] = __builtin_amdgcn_mfma_f64_4x4x4f64(
                                av, bx[u][nt][kc], acc[nt][rq], 0, 0, 0);
                    }
            } else {
                const double xv = kcl ? bx[u][0].y : bx[u][0].x;
#pragma unroll
                for (int rq = 0; rq < 2; ++rq) {
                    const double av = live ? Quad4<VT>::get_half(a[u], rq) : 0.0;
                    acc[0][rq] = __builtin_amdgcn_mfma_f64_4x4x4f64(av, xv, acc[0][rq], 0, 0, 0);
                }
            }
        }
    }
    if constexpr (R < 16) {
#pragma unroll
        for (int rq = 0; rq < 2; ++rq) acc[0][rq] += __shfl_xor(acc[0][rq], 8, 64);
        if (kcl) return;
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int64_t o = cbase + ((int64_t)(k4 * NT + nt) * WN + jl) * 2;
        d2 tv, vv;
#pragma unroll
        for (int rq = 0; rq < 2; ++rq) {
            const double y = acc[nt][rq];
            const double t = c1 * y;
            tv[rq] = t;
            vv[rq] = fma(c2, t, pre_v[nt][rq]);
        }
        if (tout) *reinterpret_cast<d2 *>(tout + o) = tv;
        if (vout) *reinterpret_cast<d2 *>(vout + o) = vv;
    }
}

// Batched Taylor term for b = 4 on v_mfma_f64_4x4x4f64, one wave per cell,
// R = 8..64 vectors per row in row-major order (at b = 4 the tile order is
// row-major).  Per slot the cell's 4x4 block times its 4 x R slice of x is
// R/4 block products, four per instruction: lane (n, blk, k) = n + 4 blk +
// 16 k supplies A[m = n][k] (the block value, read in the lane-interleaved
// order: no second copy of the values) and x[k][16 i + 4 blk + n]; the
// product of column group blk lands in lane n + 4 blk + 16 m, so lane l
// owns row m = l / 16 and columns 16 i + 4 blk + n.  No padding rows (the
// 16x16x4 form idles 12 of 16 rows at b = 4).  R = 8: groups 2, 3 repeat
// groups 0, 1 and are dropped.  W slots' loads are issued together; the
// group's block columns come from one vector load.  Summation order: slots
// ascending.
template <typename VT, int R, int W>
__global__ __launch_bounds__(256) void k_spmm_quad4(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2) {
    constexpr int B = 4, CPS = 16, BB = 16;
    constexpr int V = Vpl<VT>::value;
    constexpr int NI = R >= 16 ? R / 16 : 1;
    static_assert(R == 8 || R == 16 || R == 32 || R == 64, "width 8..64");
    static_assert(W <= 64, "one column per lane");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t cell = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_tile(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6)));
    if (cell >= ncell) return;
    const int n4 = lane & 3, blk = (lane >> 2) & 3, k4 = lane >> 4;
    const int g = R >= 16 ? 4 * blk + n4 : 4 * (blk & 1) + n4;
    const int64_t slice = cell / CPS, q = cell - slice * CPS;
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    const int64_t cbase = cell * B * R;
    const int aoff = ((k4 / V) * 64 + (int)q * 4 + n4) * V + k4 % V;
    const double *pv = vout ? vin : x;
    double pre_v[NI], acc[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int64_t o = cbase + (int64_t)k4 * R + 16 * i + g;
        pre_v[i] = pv[o];
        acc[i] = 0.0;
    }
    for (int t0 = 0; t0 < width; t0 += W) {
        const int tl = t0 + (lane < W ? lane : 0);
        const int32_t cl = col[p0 + (int64_t)(tl < width ? tl : width - 1) * CPS + q];
        int32_t c[W];
#pragma unroll
        for (int u = 0; u < W; ++u) c[u] = __builtin_amdgcn_readlane(cl, u);
        VT a[W];
        double bx[W][NI];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int t = t0 + u < width ? t0 + u : width - 1;
            a[u] = __builtin_nontemporal_load(val + (p0 + (int64_t)t * CPS) * BB + aoff);
            const double *xp = x + (int64_t)c[u] * B * R + (int64_t)k4 * R + g;
#pragma unroll
            for (int i = 0; i < NI; ++i) bx[u][i] = xp[16 * i];
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const double av = t0 + u < width ? (double)a[u] : 0.0;
#pragma unroll
            for (int i = 0; i < NI; ++i)
                acc[i] = __builtin_amdgcn_mfma_f64_4x4x4f64(av, bx[u][i], acc[i], 0, 0, 0);
        }
    }
    if (R < 16 && blk >= 2) return;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int64_t o = cbase + (int64_t)k4 * R + 16 * i + g;
        const double y = acc[i];
        const double t = c1 * y;
        if (tout) tout[o] = t;
        if (vout) vout[o] = fma(c2, t, pre_v[i]);
    }
}

// b = 4 batches on the 4x4x4 FP64 matrix cores with four cells per wave:
// block blk of v_mfma_f64_4x4x4f64 is cell q0 + blk of a group of four
// (a slice holds four groups), so one instruction covers 4 cells x 4 rows x
// 4 vectors and each slot costs one value load (the four cells' blocks are
// 64 adjacent values of the slice) and 32 bytes of x per lane and 16
// vectors.  Lane (lo, blk, hi) = lane % 4, lane / 4 % 4, lane / 16 feeds
// A = value (row lo, input hi) of cell blk, B = x (input hi of cell blk's
// block column, vectors GW i + VPL lo + g) to product g of vector group i,
// and receives D = row hi, vectors GW i + VPL lo + g (VPL = 4 vectors per
// lane and GW = 16 per group at R >= 16, 2 and 8 at R = 8): a lane's VPL x
// doubles feed its VPL products and its VPL results are adjacent in the row.
// Slots in groups of W, ascending; the group's 4 W block columns come from
// one vector load and reach their lanes by ds_bpermute.
//
// FR != 0: the receptivity operator's replicated rows fused in (FusedRepDev,
// FusedKind, as k_term_lds; with complex extra values the R doubles of a
// row are R / 2 complex values): workgroups [0, nfs_wg) take the replicated
// rows, the others (one 16-cell slice per 4-wave workgroup) the block term,
// forward adding their rows' entries in the epilogue.
//
// SH: a shock operator's extra-DOF rows in leading workgroups
// (ShockBatchDev).
template <typename VT, int R, int W, int FR = 0, bool SH = false>
__global__ __launch_bounds__(256) void k_spmm_quad4c(
    const int64_t *__restrict__ slice_ptr, const int32_t *__restrict__ col,
    const VT *__restrict__ val, int64_t ncell, const double *__restrict__ x,
    double *__restrict__ tout, const double *vin, double *vout,
    double c1, double c2, FusedRepDev F, ShockBatchDev SB) {
    constexpr int B = 4, CPS = 16, BB = 16;
    constexpr int V = Vpl<VT>::value;
    constexpr int VPL = R >= 16 ? 4 : 2;          // vectors of a lane's x chunk
    constexpr int GW = 4 * VPL;                   // vectors per group of products
    constexpr int NG = R / GW;
    constexpr int H = VPL / 2;                    // 16-byte pieces of a chunk
    static_assert(R == 8 || R == 16 || R == 32 || R == 64, "width 8..64");
    static_assert(4 * W <= 64, "four columns per slot");
    const int lane = threadIdx.x & 63;
    const int64_t wpb = blockDim.x >> 6;
    using FK = FusedKind<FR>;
    static_assert(!SH || FR == 0, "shock bands and replicas do not mix");
    int64_t bid = blockIdx.x, nwg = gridDim.x;
    if constexpr (SH) {
        // the leading rows first, one per workgroup: they read only x
        if (bid < SB.nlead_wg) {
            const bool shr = bid < SB.nsh;
            const int64_t ri = shr ? bid : bid - SB.nsh;
            const int64_t *rows = shr ? SB.sh_rows : SB.fr_rows;
            const int64_t *ptr = shr ? SB.sh_ptr : SB.fr_ptr;
            const int32_t *lc = shr ? SB.sh_col : SB.fr_col;
            const double *lv = shr ? SB.sh_val : SB.fr_val;
            if (shr || ri < SB.nfr)
                extra_long_row_lpe<R, 4, false, R / 2, true>(ri, rows, ptr, lc, lv, SB.nflow, x,
                                                             tout, vin, vout, c1, c2, 0);
            return;
        }
        bid -= SB.nlead_wg;
        nwg -= SB.nlead_wg;
    }
    if constexpr (FK::on) {
        // the replicated rows first (nfs_wg a multiple of 8: the slices keep
        // the XCD mapping of an unfused launch)
        if (bid < F.nfs_wg) {
            const int64_t wv = bid * wpb + (threadIdx.x >> 6);
            if constexpr (FK::adj)
                extra_replica_rows<R, FK::cplx, false>(wv, F.R, F.nflow, x, tout, vin, vout, c1,
                                                       c2, 0);
            else
                rep_sum_wave<R, FK::cplx, true>(F.R, wv, x, F.s_out, tout, vin, vout, c1, c2);
            return;
        }
        bid -= F.nfs_wg;
        nwg -= F.nfs_wg;
    }
    int64_t tile = xcd_tile(bid, nwg);
    if constexpr (FK::fwd) {
        tile += F.rot;   // in slices (one per workgroup); rot < nwg
        if (tile >= nwg) tile -= nwg;
    }
    const int64_t cell0 = 4 * (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(tile * wpb + (threadIdx.x >> 6)));
    if (cell0 >= ncell) return;
    const int lo = lane & 3, blk = (lane >> 2) & 3, hi = lane >> 4;
    const int64_t slice = cell0 / CPS;
    if constexpr (SH) {
        if (SB.skip[slice] >= 0) return;   // its rows are leading rows
    }
    const int q0 = (int)(cell0 - slice * CPS);
    const int64_t p0 = slice_ptr[slice];
    const int width = (int)((slice_ptr[slice + 1] - p0) / CPS);
    // lane's value in a slot: row lo, input hi of cell q0 + blk
    const int aoff = ((hi / V) * 64 + (q0 + blk) * 4 + lo) * V + hi % V;
    // lane's output (and epilogue operand) offset: cell q0 + blk (clamped for
    // the partial last group, whose extra cells are not written), row hi
    const int64_t mycell = cell0 + blk < ncell ? cell0 + blk : ncell - 1;
    const int64_t obase = (mycell * B + hi) * R + VPL * lo;
    const double *pv = vout ? vin : x;
    d2 pre_v[NG][H];
#pragma unroll
    for (int i = 0; i < NG; ++i)
#pragma unroll
        for (int h = 0; h < H; ++h) {
            pre_v[i][h] = *reinterpret_cast<const d2 *>(pv + obase + GW * i + 2 * h);
        }
    double acc[NG][VPL];
#pragma unroll
    for (int i = 0; i < NG; ++i)
#pragma unroll
        for (int g = 0; g < VPL; ++g) acc[i][g] = 0.0;
    for (int t0 = 0; t0 < width; t0 += W) {
        // lane 4u + b < 4W holds the column of slot t0 + u of cell q0 + b
        const int lu = lane < 4 * W ? lane >> 2 : 0;
        const int tl = t0 + lu < width ? t0 + lu : width - 1;
        const int32_t cl = col[p0 + (int64_t)tl * CPS + q0 + (lane & 3)];
        VT a[W];
        d2 bx[W][NG][H];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int t = t0 + u < width ? t0 + u : width - 1;
            a[u] = __builtin_nontemporal_load(val + (p0 + (int64_t)t * CPS) * BB + aoff);
            const int32_t c = __builtin_amdgcn_ds_bpermute(4 * (4 * u + blk), cl);
            const double *xp = x + ((int64_t)c * B + hi) * R + VPL * lo;
#pragma unroll
            for (int i = 0; i < NG; ++i)
#pragma unroll
                for (int h = 0; h < H; ++h)
                    bx[u][i][h] = *reinterpret_cast<const d2 *>(xp + GW * i + 2 * h);
        }
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const double av = t0 + u < width ? (double)a[u] : 0.0;
#pragma unroll
            for (int i = 0; i < NG; ++i)
#pragma unroll
                for (int g = 0; g < VPL; ++g)
                    acc[i][g] = __builtin_amdgcn_mfma_f64_4x4x4f64(av, bx[u][i][g >> 1][g & 1],
                                                                   acc[i][g], 0, 0, 0);
        }
    }
    if (cell0 + blk >= ncell) return;
    if constexpr (FK::fwd) {
        // row (cell0 + blk) * 4 + hi: its compressed extra entries times the
        // column sums of x, for this lane's doubles of the row
        const int64_t row = (cell0 + blk) * B + hi;
        const int32_t fb = F.slice_tab[row >> 6];
        if (fb >= 0) {
            const int32_t *lp = F.lane_ptr + (int64_t)fb * 65 + (row & 63);
            const int32_t e1 = lp[1];
            for (int32_t e = lp[0]; e < e1; ++e) {
                const double *sp = F.s_in + (int64_t)F.ecol[e] * R + VPL * lo;
                if constexpr (FK::cplx) {
                    const double a0 = F.eval[2 * (int64_t)e], a1 = F.eval[2 * (int64_t)e + 1];
#pragma unroll
                    for (int i = 0; i < NG; ++i)
#pragma unroll
                        for (int h = 0; h < H; ++h) {
                            const d2 s = *reinterpret_cast<const d2 *>(sp + GW * i + 2 * h);
                            acc[i][2 * h] = fma(a0, s[0], fma(-a1, s[1], acc[i][2 * h]));
                            acc[i][2 * h + 1] = fma(a0, s[1], fma(a1, s[0], acc[i][2 * h + 1]));
                        }
                } else {
                    const double a = F.eval[e];
#pragma unroll
                    for (int i = 0; i < NG; ++i)
#pragma unroll
                        for (int h = 0; h < H; ++h) {
                            const d2 s = *reinterpret_cast<const d2 *>(sp + GW * i + 2 * h);
                            acc[i][2 * h] = fma(a, s[0], acc[i][2 * h]);
                            acc[i][2 * h + 1] = fma(a, s[1], acc[i][2 * h + 1]);
                        }
                }
            }
        }
    }
    // the vout stores first, behind one wait for the epilogue operands (a
    // store between two such waits would be waited for as well)
    if (vout) {
#pragma unroll
        for (int i = 0; i < NG; ++i)
#pragma unroll
            for (int h = 0; h < H; ++h) {
                d2 vv;
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    vv[e] = fma(c2, c1 * acc[i][2 * h + e], pre_v[i][h][e]);
                *reinterpret_cast<d2 *>(vout + obase + GW * i + 2 * h) = vv;
            }
    }
    if (tout) {
#pragma unroll
        for (int i = 0; i < NG; ++i)
#pragma unroll
            for (int h = 0; h < H; ++h) {
                d2 tv;
#pragma unroll
                for (int e = 0; e < 2; ++e) tv[e] = c1 * acc[i][2 * h + e];
                *reinterpret_cast<d2 *>(tout + obase + GW * i + 2 * h) = tv;
            }
    }
}

}  // namespace hymor

namespace hymor {

// ------------------------------------------------------ operator gradient
// The gradient of J = <lambda, T_p(dt A)^n x> with respect to A's entries.
// One substep x_{k+1} = S x_k, S = sum_{j<=p} (dt A)^j / j!, contributes
//     dt * sum_{m<p} u_m z_m'   with u_m = (dt A')^m lambda_{k+1},
//     z_m = sum_{r<p-m} y_r / (m+r+1)!,  y_r = (dt A)^r x_k,
// (d(A^j) = sum_m A^m dA A^(j-1-m)), and lambda_k = S' lambda_{k+1}.

// Up to kMaxVjp vectors of n doubles with one coefficient each.
constexpr int kMaxVjp = 21;
struct VecList {
    const double *v[kMaxVjp];
    double c[kMaxVjp];
    int cnt;
};

// Z column m (at Z + m n) = sum_{r < p - m} Y.c[m + r + 1] * y_r, m = 0..p-1,
// for y_r = Y.v[r] and Y.c[j] = 1 / j!: the right factors of one substep's
// contribution (terms added in ascending r).
__global__ void k_vjp_z(VecList Y, int p, int64_t n, double *__restrict__ Z) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int m = 0; m < p; ++m) {
            double z = 0.0;
            for (int r = 0; r + m < p; ++r) z = fma(Y.c[m + r + 1], Y.v[r][i], z);
            Z[m * n + i] = z;
        }
    }
}

// Columns of a batched buffer (width w, layout buf_pos with tile_b) from the
// cnt single vectors X.v[0..cnt) (cell-major, n doubles each); columns from
// cnt on are zero padding.
__global__ void k_cols_to_batch(VecList X, int cnt, int64_t n, int w, int tile_b,
                                int64_t nflow, double *__restrict__ out) {
    const int64_t total = n * w;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / w;
        const int j = (int)(i - row * w);
        out[buf_pos(row, j, w, tile_b, nflow)] = j < cnt ? X.v[j][row] : 0.0;
    }
}

// k_vjp_z for cnt states at once: state i's y_r is column i of the batched
// buffer Yb[r] (r < p, width w, layout buf_pos), and its z_m, m < p, go to
// Z + (i p + m) n.  Coefficients c[j] = 1 / j! (terms added in ascending r,
// as k_vjp_z adds them).
struct BatchList {
    const double *v[kMaxVjp];
    double c[kMaxVjp];
};
__global__ void k_vjp_z_batch(BatchList Yb, int p, int cnt, int64_t n, int w, int tile_b,
                              int64_t nflow, double *__restrict__ Z) {
    const int64_t total = n * cnt;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = t / cnt;
        const int i = (int)(t - row * cnt);
        const int64_t o = buf_pos(row, i, w, tile_b, nflow);
        for (int m = 0; m < p; ++m) {
            double z = 0.0;
            for (int r = 0; r + m < p; ++r) z = fma(Yb.c[m + r + 1], Yb.v[r][o], z);
            Z[((int64_t)i * p + m) * n + row] = z;
        }
    }
}

// out = sum_{j < L.cnt} L.c[j] * L.v[j] (terms added in ascending j); out is
// none of the inputs.
__global__ void k_lincomb(VecList L, int64_t n, double *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < L.cnt; ++j) s = fma(L.c[j], L.v[j][i], s);
        out[i] = s;
    }
}

// Sampled product on a pattern: for its nnz entries (row er[e], column
// ec[e], cell-major), grad[e] = (init ? 0 : grad[e]) + alpha * sum_{k < kc}
// U[k n + er[e]] * Z[k n + ec[e]] (U, Z column-major n x kc; the products
// added in ascending k).  A thread per entry; the workgroups take
// contiguous ranges of entries per XCD, so the U and Z rows that
// neighbouring entries share stay in one L2.  With eidx, list entry i is
// the pattern's entry eidx[i] (the entries k_sddmm_blocks leaves).
__global__ __launch_bounds__(256) void k_sddmm_pattern(
    const int32_t *__restrict__ er, const int32_t *__restrict__ ec, int64_t nnz,
    const double *__restrict__ U, const double *__restrict__ Z, int64_t n, int kc, double alpha,
    int init, const int32_t *__restrict__ eidx, double *__restrict__ grad) {
    const int64_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const int64_t r = er[i], c = ec[i];
    const int64_t e = eidx ? (int64_t)eidx[i] : i;
    double acc = 0.0;
    int k = 0;
    for (; k + 4 <= kc; k += 4) {
        double u[4], z[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u[q] = U[(int64_t)(k + q) * n + r];
            z[q] = Z[(int64_t)(k + q) * n + c];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = fma(u[q], z[q], acc);
    }
    for (; k < kc; ++k) acc = fma(U[(int64_t)k * n + r], Z[(int64_t)k * n + c], acc);
    grad[e] = (init ? 0.0 : grad[e]) + alpha * acc;
}

// The sampled product over the b x b blocks of a pattern (B = b): thread t
// takes block t, whose rows are those of cell bcr[2t] and columns those of
// cell bcr[2t + 1] (cell-major: B consecutive indices each), and whose
// entries are slot[t B^2 + i B + j] for row i, column j of the block (-1:
// none).  Each slot gets the sum k_sddmm_pattern forms for it -- the same
// products, added in ascending k -- with the block's B values of U and B of
// Z per column k read once for its up to B^2 entries instead of once per
// entry.
template <int B>
__global__ __launch_bounds__(256) void k_sddmm_blocks(
    const int32_t *__restrict__ bcr, const int32_t *__restrict__ slot, int64_t nblk,
    const double *__restrict__ U, const double *__restrict__ Z, int64_t n, int kc, double alpha,
    int init, double *__restrict__ grad) {
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int64_t rb = (int64_t)bcr[2 * t] * B, cb = (int64_t)bcr[2 * t + 1] * B;
    double acc[B][B];
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
        for (int j = 0; j < B; ++j) acc[i][j] = 0.0;
    for (int k = 0; k < kc; ++k) {
        const double *u = U + (int64_t)k * n + rb, *z = Z + (int64_t)k * n + cb;
        double uv[B], zv[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
            uv[i] = u[i];
            zv[i] = z[i];
        }
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int j = 0; j < B; ++j) acc[i][j] = fma(uv[i], zv[j], acc[i][j]);
    }
    const int32_t *sl = slot + t * (B * B);
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int32_t e = sl[i * B + j];
            if (e >= 0) grad[e] = (init ? 0.0 : grad[e]) + alpha * acc[i][j];
        }
}

}  // namespace hymor
