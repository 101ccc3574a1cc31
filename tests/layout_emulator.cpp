// CPU emulation of the device layout and the kernels' index arithmetic.
//
// Built by tests/test_layout_cpu.py with g++ together with
// hymor_amd/csrc/convert.cpp (the library's own host conversion).  It runs
// csc_to_blocks -> bsr_to_sliced and then replays k_extra / k_term's exact
// index math on the CPU, checking every access against the buffer it reads,
// so that a layout or indexing mistake shows up here rather than as a GPU
// memory fault.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hymor_internal.hpp"

using namespace hymor;

namespace {
int g_oob = 0;
template <typename T>
const T &at(const std::vector<T> &v, int64_t i) {
    if (i < 0 || i >= (int64_t)v.size()) {
        ++g_oob;
        static T zero{};
        return zero;
    }
    return v[(size_t)i];
}
}  // namespace

extern "C" int emu_spmv_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                            const double *nzval, int base, int is_complex, int b, int64_t ncell,
                            int transpose, const double *x_vm, double *y_vm) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, base, is_complex != 0, 1.0, b, ncell, fwd, adj, xf, xa);
    const HostBSR &A = transpose ? adj : fwd;
    const HostExtra &X = transpose ? xa : xf;
    HostSliced s;
    bsr_to_sliced(A, s);
    interleave_lanes(s, 2);   // FP64 storage: 2 values per 16-byte chunk
    const int NV = is_complex ? 2 : 1;
    const int64_t nflow = (int64_t)b * ncell;
    // perm_in
    std::vector<double> x((size_t)n * NV), y((size_t)n * NV, 0.0), yx((size_t)nflow * NV, 0.0);
    for (int64_t row = 0; row < n; ++row) {
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        for (int v = 0; v < NV; ++v) x[(size_t)(row * NV + v)] = x_vm[vm * NV + v];
    }
    // k_extra
    for (int64_t w = 0; w < (int64_t)X.rows.size(); ++w) {
        double acc[2] = {0, 0};
        for (int64_t e = at(X.ptr, w); e < at(X.ptr, w + 1); ++e) {
            const int64_t c = at(X.col, e);
            if (is_complex) {
                const double ar = at(X.val, 2 * e), ai = at(X.val, 2 * e + 1);
                const double xr = at(x, c * NV), xi = at(x, c * NV + 1);
                acc[0] += ar * xr - ai * xi;
                acc[1] += ar * xi + ai * xr;
            } else {
                acc[0] += at(X.val, e) * at(x, c);
            }
        }
        const int64_t row = at(X.rows, w);
        for (int v = 0; v < NV; ++v) {
            if (row < nflow) {
                if (row * NV + v >= (int64_t)yx.size()) ++g_oob;
                else yx[(size_t)(row * NV + v)] = acc[v];
            } else {
                if (row * NV + v >= (int64_t)y.size()) ++g_oob;
                else y[(size_t)(row * NV + v)] = acc[v];
            }
        }
    }
    // k_term: one "wave" per slice, lane = (q, r)
    const int CPS = 64 / b, BB = b * b;
    for (int64_t slice = 0; slice < s.nslice; ++slice) {
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        if ((at(s.slice_ptr, slice + 1) - p0) % CPS) ++g_oob;
        for (int lane = 0; lane < 64; ++lane) {
            const int q = lane / b, r = lane % b;
            double acc[2] = {0, 0};
            for (int64_t t = 0; t < width; ++t) {
                const int64_t blk = p0 + q + t * CPS;
                const int64_t c = at(s.col, blk);
                if (c < 0 || c >= ncell) ++g_oob;
                const int64_t sb = (p0 + t * CPS) * BB;   // slot base
                for (int k = 0; k < b; ++k) {
                    const double a = at(s.val, sb + ((k / 2) * 64 + lane) * 2 + k % 2);
                    for (int v = 0; v < NV; ++v) acc[v] += a * at(x, (c * b + k) * NV + v);
                }
            }
            const int64_t cell = slice * CPS + q;
            if (cell >= ncell) continue;
            const int64_t row = cell * b + r;
            for (int v = 0; v < NV; ++v)
                y[(size_t)(row * NV + v)] = acc[v] + at(yx, row * NV + v);
        }
    }
    // perm_out
    for (int64_t vm = 0; vm < n; ++vm) {
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        for (int v = 0; v < NV; ++v) y_vm[vm * NV + v] = y[(size_t)(row * NV + v)];
    }
    return g_oob;
}

// emu_spmv_csc with the replicated freestream blocks compressed as the
// library does at upload (compress_column_replicas on the forward operator,
// compress_row_replicas on the adjoint with the forward's period), and the
// compressed form's arithmetic replayed: forward, S[q] = sum_k x[c0 + k p + q]
// and the diagonal rows [c0, n) where S is formed, virtual columns n + q of
// the remaining rows reading S; adjoint, each shared row q evaluated once and
// finished into rows c0 + k p + q with their own diagonal.  info[0..3] =
// kind (0 if nothing was compressed), c0, p, nrep.
extern "C" int emu_spmv_csc_rep(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                const double *nzval, int base, int is_complex, int b,
                                int64_t ncell, int transpose, const double *x_vm, double *y_vm,
                                int64_t *info) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, base, is_complex != 0, 1.0, b, ncell, fwd, adj, xf, xa);
    HostReplica rf, ra, *rep = nullptr;
    const bool found = compress_column_replicas(xf, n, rf);
    if (found && !transpose) rep = &rf;
    if (found && transpose && compress_row_replicas(xa, n, rf.c0, rf.p, rf.nrep, ra)) rep = &ra;
    info[0] = rep ? rep->kind : 0;
    info[1] = rep ? rep->c0 : 0;
    info[2] = rep ? rep->p : 0;
    info[3] = rep ? rep->nrep : 0;
    const HostBSR &A = transpose ? adj : fwd;
    const HostExtra &X = transpose ? xa : xf;
    HostSliced s;
    bsr_to_sliced(A, s);
    interleave_lanes(s, 2);
    const int NV = is_complex ? 2 : 1;
    const int vs = is_complex ? 2 : 1;
    const int64_t nflow = (int64_t)b * ncell;
    std::vector<double> x((size_t)n * NV), y((size_t)n * NV, 0.0), yx((size_t)nflow * NV, 0.0);
    for (int64_t row = 0; row < n; ++row) {
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        for (int v = 0; v < NV; ++v) x[(size_t)(row * NV + v)] = x_vm[vm * NV + v];
    }
    auto cmac = [&](double *acc, const double *a, const double *xv) {
        if (is_complex) {
            acc[0] += a[0] * xv[0] - a[1] * xv[1];
            acc[1] += a[0] * xv[1] + a[1] * xv[0];
        } else {
            acc[0] += a[0] * xv[0];
        }
    };
    std::vector<double> S;
    if (rep && rep->kind == 1) {
        // the pre-pass: column sums, and the diagonal rows finished there
        S.assign((size_t)(rep->p * NV), 0.0);
        for (int64_t q = 0; q < rep->p; ++q)
            for (int64_t k = 0; k < rep->nrep; ++k) {
                const int64_t row = rep->c0 + k * rep->p + q;
                double xv[2] = {at(x, row * NV), NV > 1 ? at(x, row * NV + 1) : 0.0};
                for (int v = 0; v < NV; ++v) S[(size_t)(q * NV + v)] += xv[v];
                double acc[2] = {0, 0};
                const double dg[2] = {at(rep->diag, (row - rep->c0) * vs),
                                      vs > 1 ? at(rep->diag, (row - rep->c0) * vs + 1) : 0.0};
                cmac(acc, dg, xv);
                for (int v = 0; v < NV; ++v) y[(size_t)(row * NV + v)] = acc[v];
            }
    }
    // k_extra: listed rows (virtual columns read S)
    for (int64_t w = 0; w < (int64_t)X.rows.size(); ++w) {
        double acc[2] = {0, 0};
        for (int64_t e = at(X.ptr, w); e < at(X.ptr, w + 1); ++e) {
            const int64_t c = at(X.col, e);
            double a[2] = {at(X.val, e * vs), vs > 1 ? at(X.val, e * vs + 1) : 0.0};
            double xv[2];
            if (c >= n) {
                xv[0] = at(S, (c - n) * NV);
                xv[1] = NV > 1 ? at(S, (c - n) * NV + 1) : 0.0;
            } else {
                xv[0] = at(x, c * NV);
                xv[1] = NV > 1 ? at(x, c * NV + 1) : 0.0;
            }
            cmac(acc, a, xv);
        }
        const int64_t row = at(X.rows, w);
        for (int v = 0; v < NV; ++v) {
            if (row < nflow) {
                if (row * NV + v >= (int64_t)yx.size()) ++g_oob;
                else yx[(size_t)(row * NV + v)] = acc[v];
            } else {
                if (row * NV + v >= (int64_t)y.size()) ++g_oob;
                else y[(size_t)(row * NV + v)] = acc[v];
            }
        }
    }
    if (rep && rep->kind == 2) {
        // shared rows: evaluated once, finished into every replica
        for (int64_t q = 0; q < rep->p; ++q) {
            double sh[2] = {0, 0};
            for (int64_t e = at(rep->vptr, q); e < at(rep->vptr, q + 1); ++e) {
                const int64_t c = at(rep->vcol, e);
                const double a[2] = {at(rep->vval, e * vs), vs > 1 ? at(rep->vval, e * vs + 1) : 0.0};
                const double xv[2] = {at(x, c * NV), NV > 1 ? at(x, c * NV + 1) : 0.0};
                cmac(sh, a, xv);
            }
            for (int64_t k = 0; k < rep->nrep; ++k) {
                const int64_t row = rep->c0 + k * rep->p + q;
                double acc[2] = {sh[0], sh[1]};
                const double dg[2] = {at(rep->diag, (row - rep->c0) * vs),
                                      vs > 1 ? at(rep->diag, (row - rep->c0) * vs + 1) : 0.0};
                const double xv[2] = {at(x, row * NV), NV > 1 ? at(x, row * NV + 1) : 0.0};
                cmac(acc, dg, xv);
                for (int v = 0; v < NV; ++v) {
                    if (row * NV + v >= (int64_t)y.size()) ++g_oob;
                    else y[(size_t)(row * NV + v)] = acc[v];
                }
            }
        }
    }
    // k_term on the blocks, adding the extra rows' part of flow rows
    const int CPS = 64 / b, BB = b * b;
    for (int64_t slice = 0; slice < s.nslice; ++slice) {
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        for (int lane = 0; lane < 64; ++lane) {
            const int q = lane / b, r = lane % b;
            double acc[2] = {0, 0};
            for (int64_t t = 0; t < width; ++t) {
                const int64_t blk = p0 + q + t * CPS;
                const int64_t c = at(s.col, blk);
                if (c < 0 || c >= ncell) ++g_oob;
                const int64_t sb = (p0 + t * CPS) * BB;
                for (int k = 0; k < b; ++k) {
                    const double a = at(s.val, sb + ((k / 2) * 64 + lane) * 2 + k % 2);
                    for (int v = 0; v < NV; ++v) acc[v] += a * at(x, (c * b + k) * NV + v);
                }
            }
            const int64_t cell = slice * CPS + q;
            if (cell >= ncell) continue;
            const int64_t row = cell * b + r;
            for (int v = 0; v < NV; ++v)
                y[(size_t)(row * NV + v)] = acc[v] + at(yx, row * NV + v);
        }
    }
    for (int64_t vm = 0; vm < n; ++vm) {
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        for (int v = 0; v < NV; ++v) y_vm[vm * NV + v] = y[(size_t)(row * NV + v)];
    }
    return g_oob;
}

// A Horner sweep T_order(A dt)^nsub x with the replicated rows fused into
// the term kernel, at nv doubles per row (x_vm[vm * nv + d]; complex
// operators: nv / 2 complex values): the prologue's column sums (k_rep_sum
// with no outputs), then per term the fused kernel's index math -- the
// replicated rows by rep_sum_wave's lane mapping summing the new rows, the
// flow rows' entries from the slice tables in the epilogue, the slices in
// rotated order -- with the two sum buffers alternating and the work
// buffers picked as run_taylor_in picks them.  nv = 2: k_term_lds<FR>'s
// lanes (one per row); nv = 8..64 (b = 4): k_spmm_quad4c<FR>'s lanes (row
// (cell0 + blk) * 4 + hi, doubles GW i + VPL lo + g), its block part
// computed from the BSR.  info[0] = 1 when the tables were built (nothing
// is computed otherwise), info[1] their blocks, info[2] their entries.
extern "C" int emu_horner_fused(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                const double *nzval, int base, int is_complex, int b,
                                int64_t ncell, int order, int nsub, double dt, int nv,
                                const double *x_vm, double *y_vm, int64_t *info) {
    g_oob = 0;
    info[0] = info[1] = info[2] = 0;
    if (!(nv == 2 || (b == 4 && (nv == 8 || nv == 16 || nv == 32 || nv == 64)))) return -1;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, base, is_complex != 0, 1.0, b, ncell, fwd, adj, xf, xa);
    HostReplica rep;
    const int64_t nflow = (int64_t)b * ncell;
    if (!compress_column_replicas(xf, n, rep)) return 0;
    HostFusedRep F;
    if (!build_fused_tables(xf, rep, n, nflow, F)) return 0;
    info[0] = 1;
    info[1] = (int64_t)F.lane_ptr.size() / 65;
    info[2] = (int64_t)F.ecol.size();
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, 2);
    constexpr int KU = 4;
    const int NV = nv;
    const bool cplx = is_complex != 0;
    const int vs = cplx ? 2 : 1, NC = cplx ? 2 : 1, LPR = cplx ? NV / 2 : NV;
    const int KS = (64 / LPR) >= 8 ? 8 : 64 / LPR;
    const int VPW = 64 / (LPR * KS);
    std::vector<double> in((size_t)n * NV), pool[4], S[2];
    for (auto &p : pool) p.assign((size_t)n * NV, 0.0);
    for (auto &v : S) v.assign((size_t)rep.p * NV, 0.0);
    for (int64_t row = 0; row < n; ++row) {
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        for (int v = 0; v < NV; ++v) in[(size_t)(row * NV + v)] = x_vm[vm * NV + v];
    }
    auto put = [&](std::vector<double> &buf, int64_t i, double val) {
        if (i < 0 || i >= (int64_t)buf.size()) ++g_oob;
        else buf[(size_t)i] = val;
    };
    // rep_sum_wave over all waves; vout null: sums of x only (the prologue)
    auto rep_waves = [&](bool out_sum, const std::vector<double> &x, const std::vector<double> *vin,
                         std::vector<double> *vout, double c1, double c2, std::vector<double> &sum) {
        const int64_t nw = (rep.p + VPW - 1) / VPW;
        for (int64_t w = 0; w < nw; ++w) {
            double sl[64][2] = {};
            for (int lane = 0; lane < 64; ++lane) {
                const int64_t q = w * VPW + lane / (LPR * KS);
                const int ks = (lane / LPR) % KS, v0 = (lane % LPR) * NC;
                if (q >= rep.p) continue;
                for (int64_t k0 = ks; k0 < rep.nrep; k0 += KS * KU) {
                    double xv[KU][2], dv[KU][2], vv[KU][2];
                    for (int u = 0; u < KU; ++u) {
                        const int64_t k = k0 + (int64_t)u * KS < rep.nrep ? k0 + (int64_t)u * KS : k0;
                        const int64_t row = rep.c0 + k * rep.p + q;
                        for (int i = 0; i < NC; ++i) {
                            xv[u][i] = at(x, row * NV + v0 + i);
                            vv[u][i] = vout ? at(*vin, row * NV + v0 + i) : 0.0;
                        }
                        for (int i = 0; i < vs; ++i) dv[u][i] = at(rep.diag, (k * rep.p + q) * vs + i);
                    }
                    for (int u = 0; u < KU; ++u) {
                        const int64_t k = k0 + (int64_t)u * KS;
                        if (k >= rep.nrep) break;
                        const int64_t row = rep.c0 + k * rep.p + q;
                        double acc[2];
                        if (!out_sum)
                            for (int i = 0; i < NC; ++i) sl[lane][i] += xv[u][i];
                        if (cplx) {
                            acc[0] = dv[u][0] * xv[u][0] - dv[u][1] * xv[u][1];
                            acc[1] = dv[u][0] * xv[u][1] + dv[u][1] * xv[u][0];
                        } else {
                            acc[0] = dv[u][0] * xv[u][0];
                        }
                        for (int i = 0; i < NC; ++i) {
                            const double wn = c2 * (c1 * acc[i]) + vv[u][i];
                            if (vout) put(*vout, row * NV + v0 + i, wn);
                            if (out_sum) sl[lane][i] += wn;
                        }
                    }
                }
            }
            for (int off = LPR; off < LPR * KS; off <<= 1) {
                double t[64][2];
                for (int l = 0; l < 64; ++l)
                    for (int i = 0; i < 2; ++i) t[l][i] = sl[l][i] + sl[l ^ off][i];
                for (int l = 0; l < 64; ++l)
                    for (int i = 0; i < 2; ++i) sl[l][i] = t[l][i];
            }
            for (int lane = 0; lane < 64; ++lane) {
                const int64_t q = w * VPW + lane / (LPR * KS);
                const int ks = (lane / LPR) % KS, v0 = (lane % LPR) * NC;
                if (q < rep.p && ks == 0)
                    for (int i = 0; i < NC; ++i) put(sum, q * NV + v0 + i, sl[lane][i]);
            }
        }
    };
    // entries of `row` times the sums, added to the doubles d0 + 0, 1 of acc
    auto add_entries = [&](int64_t row, const std::vector<double> &Sin, int64_t d0, double *a2) {
        const int32_t fb = at(F.slice_tab, row >> 6);
        if (fb < 0) return;
        const int64_t e1 = at(F.lane_ptr, (int64_t)fb * 65 + (row & 63) + 1);
        for (int64_t e = at(F.lane_ptr, (int64_t)fb * 65 + (row & 63)); e < e1; ++e) {
            const int64_t qc = at(F.ecol, e);
            const double s0 = at(Sin, qc * NV + d0), s1 = at(Sin, qc * NV + d0 + 1);
            if (cplx) {
                const double a0 = at(F.eval, 2 * e), a1 = at(F.eval, 2 * e + 1);
                a2[0] += a0 * s0 - a1 * s1;
                a2[1] += a0 * s1 + a1 * s0;
            } else {
                const double a = at(F.eval, e);
                a2[0] += a * s0;
                a2[1] += a * s1;
            }
        }
    };
    const int CPS = 64 / b, BB = b * b;
    int64_t rot = 0;   // slices from the first one with entries on, wrapping
    while (rot < s.nslice && at(F.slice_tab, rot) < 0) ++rot;
    if (rot == s.nslice) rot = 0;
    int slot = 0;
    const std::vector<double> *cur = &in;
    if (nsub > 0) rep_waves(false, *cur, nullptr, nullptr, 0.0, 0.0, S[slot]);
    for (int sub = 0; sub < nsub; ++sub) {
        std::vector<double> *pp[2] = {nullptr, nullptr};
        int np = 0;
        for (auto &p : pool)
            if (&p != cur && np < 2) pp[np++] = &p;
        const std::vector<double> *w = cur;
        for (int j = order; j >= 1; --j) {
            std::vector<double> &wn = *pp[(order - j) & 1];
            const double c1 = dt / (double)j, c2 = 1.0;
            const std::vector<double> &Sin = S[slot];
            rep_waves(true, *w, cur, &wn, c1, c2, S[slot ^ 1]);
            for (int64_t t = 0; t < s.nslice; ++t) {
                int64_t slice = t + rot;
                if (slice >= s.nslice) slice -= s.nslice;
                if (NV == 2) {
                    // k_term_lds: lane (q, r) of the slice owns row slice * 64 + lane
                    const int64_t p0 = at(s.slice_ptr, slice);
                    const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
                    for (int lane = 0; lane < 64; ++lane) {
                        const int q = lane / b, r = lane % b;
                        double acc[2] = {0, 0};
                        for (int64_t tt = 0; tt < width; ++tt) {
                            const int64_t c = at(s.col, p0 + q + tt * CPS);
                            if (c < 0 || c >= ncell) ++g_oob;
                            const int64_t sb = (p0 + tt * CPS) * BB;
                            for (int k = 0; k < b; ++k) {
                                const double a = at(s.val, sb + ((k / 2) * 64 + lane) * 2 + k % 2);
                                for (int v = 0; v < 2; ++v) acc[v] += a * at(*w, (c * b + k) * 2 + v);
                            }
                        }
                        const int64_t cell = slice * CPS + q;
                        if (cell >= ncell) continue;
                        const int64_t row = cell * b + r;
                        add_entries(row, Sin, 0, acc);
                        for (int v = 0; v < 2; ++v)
                            put(wn, row * 2 + v, c2 * (c1 * acc[v]) + at(*cur, row * 2 + v));
                    }
                } else {
                    // k_spmm_quad4c: workgroup `slice`, four waves of four cells
                    const int VPL = NV >= 16 ? 4 : 2, GW = 4 * VPL, NG = NV / GW, H = VPL / 2;
                    for (int wv = 0; wv < 4; ++wv) {
                        const int64_t cell0 = 4 * (slice * 4 + wv);
                        if (cell0 >= ncell) continue;
                        for (int lane = 0; lane < 64; ++lane) {
                            const int lo = lane & 3, blk = (lane >> 2) & 3, hi = lane >> 4;
                            const int64_t cell = cell0 + blk;
                            if (cell >= ncell) continue;
                            const int64_t row = cell * 4 + hi;
                            for (int i = 0; i < NG; ++i)
                                for (int h = 0; h < H; ++h) {
                                    const int64_t d0 = GW * i + VPL * lo + 2 * h;
                                    double acc[2] = {0, 0};
                                    for (int64_t k = at(fwd.ptr, cell); k < at(fwd.ptr, cell + 1); ++k) {
                                        const int64_t c = at(fwd.col, k);
                                        for (int jj = 0; jj < 4; ++jj) {
                                            const double a = at(fwd.val, k * 16 + hi * 4 + jj);
                                            for (int e = 0; e < 2; ++e)
                                                acc[e] += a * at(*w, (c * 4 + jj) * NV + d0 + e);
                                        }
                                    }
                                    add_entries(row, Sin, d0, acc);
                                    for (int e = 0; e < 2; ++e)
                                        put(wn, row * NV + d0 + e,
                                            c2 * (c1 * acc[e]) + at(*cur, row * NV + d0 + e));
                                }
                        }
                    }
                }
            }
            slot ^= 1;
            w = &wn;
        }
        cur = w;
    }
    for (int64_t vm = 0; vm < n; ++vm) {
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        for (int v = 0; v < NV; ++v) y_vm[vm * NV + v] = at(*cur, row * NV + v);
    }
    return g_oob;
}

// Batched form: k_perm_in (padded width) -> k_spmm's index math for R
// doubles per row -> k_perm_out.  Real operators without extras.
extern "C" int emu_spmm_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                            const double *nzval, int b, int64_t ncell, int nrhs, int R,
                            const double *X_vm, double *Y_vm) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, 2);
    const int64_t nflow = (int64_t)b * ncell;
    std::vector<double> x((size_t)n * R), y((size_t)n * R, 0.0);
    // k_perm_in with width R, ncomp 1
    for (int64_t i = 0; i < n * R; ++i) {
        const int64_t row = i / R, rem = i % R;
        if (rem >= nrhs) { x[(size_t)i] = 0.0; continue; }
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        x[(size_t)i] = X_vm[rem * n + vm];
    }
    const int CPS = 64 / b, BB = b * b, G = 64 / R, RPL = b / G;
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t slice = cell / CPS, q = cell - slice * CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        for (int lane = 0; lane < 64; ++lane) {
            const int r = lane % R, g = lane / R;
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t t = 0; t < width; ++t) {
                const int64_t blk = p0 + t * CPS + q;
                const int64_t c = at(s.col, blk);
                const int64_t sb = (p0 + t * CPS) * BB;
                for (int k = 0; k < b; ++k) {
                    const double xk = at(x, c * (b * R) + r + k * R);
                    for (int i = 0; i < RPL; ++i) {
                        const int64_t ln = q * b + g * RPL + i;
                        acc[i] += at(s.val, sb + ((k / 2) * 64 + ln) * 2 + k % 2) * xk;
                    }
                }
            }
            for (int i = 0; i < RPL; ++i) {
                const int64_t idx = (cell * b + g * RPL + i) * R + r;
                if (idx < 0 || idx >= (int64_t)y.size()) ++g_oob;
                else y[(size_t)idx] = acc[i];
            }
        }
    }
    // k_perm_out
    for (int64_t i = 0; i < n * nrhs; ++i) {
        const int64_t rhs = i / n, vm = i % n;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        Y_vm[rhs * n + vm] = at(y, row * R + rhs);
    }
    return g_oob;
}

// k_spmm_mfma: per cell, lanes supply MFMA operands exactly as the kernel
// indexes them; the instruction is replayed with the gfx950 operand layout
// of v_mfma_f64_16x16x4f64 (A[m][k] from lane m+16k, B[k][n] from lane
// n+16k, D[m][n] in lane n+16*(m%4), register m/4).  vpl: values per
// 16-byte chunk of the block storage (2 FP64, 4 FP32; FP32 values are
// emulated in double).
extern "C" int emu_spmm_mfma_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                 const double *nzval, int b, int64_t ncell, int nrhs, int R,
                                 int vpl, const double *X_vm, double *Y_vm) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    const int64_t nflow = (int64_t)b * ncell;
    std::vector<double> x((size_t)n * R), y((size_t)n * R, 0.0);
    for (int64_t i = 0; i < n * R; ++i) {
        const int64_t row = i / R, rem = i % R;
        if (rem >= nrhs) { x[(size_t)i] = 0.0; continue; }
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        x[(size_t)i] = X_vm[rem * n + vm];
    }
    const int CPS = 64 / b, BB = b * b, NT = R / 16, KC = b / 4;
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t slice = cell / CPS, q = cell - slice * CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        std::vector<double> D((size_t)NT * 256, 0.0);  // D[nt][m][n]
        for (int64_t t = 0; t < width; ++t) {
            const int64_t c = at(s.col, p0 + t * CPS + q);
            const int64_t sb = (p0 + t * CPS) * BB;
            for (int kc = 0; kc < KC; ++kc) {
                double A[16][4], Bm[NT > 0 ? 4 : 1][16 * 4];
                double opa[64], opb[4][64];
                for (int lane = 0; lane < 64; ++lane) {
                    const int m = lane & 15, kk = lane >> 4, k = 4 * kc + kk;
                    opa[lane] = m < b ? at(s.val, sb + ((int64_t)(k / vpl) * 64 + q * b + m) * vpl + k % vpl) : 0.0;
                    for (int nt = 0; nt < NT; ++nt)
                        opb[nt][lane] = at(x, (c * b + k) * R + m + nt * 16);
                }
                for (int mm = 0; mm < 16; ++mm)
                    for (int k = 0; k < 4; ++k) A[mm][k] = opa[mm + 16 * k];
                (void)Bm;
                for (int nt = 0; nt < NT; ++nt)
                    for (int mm = 0; mm < 16; ++mm)
                        for (int nn = 0; nn < 16; ++nn) {
                            double sacc = 0.0;
                            for (int k = 0; k < 4; ++k) sacc += A[mm][k] * opb[nt][nn + 16 * k];
                            D[(size_t)nt * 256 + mm * 16 + nn] += sacc;
                        }
            }
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int m = lane & 15, kk = lane >> 4;
            for (int nt = 0; nt < NT; ++nt)
                for (int r = 0; r < KC; ++r) {
                    // register r of this lane holds D[kk + 4r][m]
                    const double v = D[(size_t)nt * 256 + (kk + 4 * r) * 16 + m];
                    const int64_t idx = (cell * b + kk + 4 * r) * R + nt * 16 + m;
                    if (idx < 0 || idx >= (int64_t)y.size()) ++g_oob;
                    else y[(size_t)idx] = v;
                }
        }
    }
    for (int64_t i = 0; i < n * nrhs; ++i) {
        const int64_t rhs = i / n, vm = i % n;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        Y_vm[rhs * n + vm] = at(y, row * R + rhs);
    }
    return g_oob;
}

// Tile path: k_perm_in / k_perm_out with tile_b = b, k_tile_vals, and
// k_spmm_tile's per-lane operand fetches (KC adjacent values per lane),
// MFMA operand placement and epilogue stores, for R = 8, 16, 32, 64.
namespace {
int64_t tile_pos_h(int64_t row, int j, int b, int r) {
    const int w = r < 16 ? r : 16;
    const int nt_ = r / w, kc_ = b >> 2;
    const int64_t cell = row / b;
    const int var = (int)(row - cell * b);
    const int kk = var & 3, kc = var >> 2;
    const int nt = j / w, n = j - nt * w;
    return cell * b * r + ((int64_t)(kk * nt_ + nt) * w + n) * kc_ + kc;
}
}  // namespace

extern "C" int emu_spmm_tile_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                 const double *nzval, int b, int64_t ncell, int nrhs, int R,
                                 int vpl, const double *X_vm, double *Y_vm) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    const int64_t nflow = (int64_t)b * ncell;
    // k_perm_in (tile layout for flow rows)
    std::vector<double> x((size_t)n * R, -1e300), y((size_t)n * R, 0.0);
    for (int64_t i = 0; i < n * R; ++i) {
        const int64_t row = i / R, rem = i % R;
        const int64_t o = row < nflow ? tile_pos_h(row, (int)rem, b, R) : row * R + rem;
        if (o < 0 || o >= (int64_t)x.size()) { ++g_oob; continue; }
        if (rem >= nrhs) { x[(size_t)o] = 0.0; continue; }
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        x[(size_t)o] = X_vm[rem * n + vm];
    }
    // k_tile_vals
    const int CPS = 64 / b, BB = b * b, KC = b / 4;
    const int64_t slot_len = (int64_t)CPS * BB;
    const int64_t total = (int64_t)s.val.size();
    std::vector<double> valt((size_t)total, 0.0);
    for (int64_t i = 0; i < total; ++i) {
        const int64_t slot = i / slot_len;
        const int within = (int)(i - slot * slot_len);
        const int q = within / BB, rem = within - q * BB;
        const int kc = rem % KC, tmp = rem / KC;
        const int kk = tmp & 3, m = tmp >> 2;
        const int k = kk + 4 * kc;
        valt[(size_t)i] = at(s.val, slot * slot_len + ((int64_t)(k / vpl) * 64 + q * b + m) * vpl + k % vpl);
    }
    const int WN = R < 16 ? R : 16, NT = R / WN;
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t slice = cell / CPS, q = cell - slice * CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        const int64_t cbase = cell * b * R;
        std::vector<double> D((size_t)NT * 256, 0.0);  // D[nt][m][n]
        for (int64_t t = 0; t < width; ++t) {
            const int64_t c = at(s.col, p0 + t * CPS + q);
            // per lane: KC adjacent values and NT x KC adjacent x entries
            double opa[64][2], opb[4][64][2];
            for (int lane = 0; lane < 64; ++lane) {
                const int nn = lane & 15, kk = lane >> 4;
                const int nr = nn & (WN - 1), mr = nn & (b - 1);
                const int64_t sb = (p0 + t * CPS + q) * BB + (mr * 4 + kk) * KC;
                const int64_t xb = c * b * R + ((int64_t)kk * NT * WN + nr) * KC;
                for (int kc = 0; kc < KC; ++kc) {
                    opa[lane][kc] = nn < b ? at(valt, sb + kc) : 0.0;
                    for (int nt = 0; nt < NT; ++nt)
                        opb[nt][lane][kc] = at(x, xb + (int64_t)nt * WN * KC + kc);
                }
            }
            for (int kc = 0; kc < KC; ++kc)
                for (int nt = 0; nt < NT; ++nt)
                    for (int mm = 0; mm < 16; ++mm)
                        for (int nn = 0; nn < 16; ++nn) {
                            double sacc = 0.0;
                            // A[m][k] from lane m + 16k, B[k][n] from lane n + 16k
                            for (int k = 0; k < 4; ++k)
                                sacc += opa[mm + 16 * k][kc] * opb[nt][nn + 16 * k][kc];
                            D[(size_t)nt * 256 + mm * 16 + nn] += sacc;
                        }
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int nn = lane & 15, kk = lane >> 4;
            if (nn >= WN) continue;
            for (int nt = 0; nt < NT; ++nt) {
                const int64_t o = cbase + ((int64_t)(kk * NT + nt) * WN + nn) * KC;
                for (int r = 0; r < KC; ++r) {
                    // register r of lane n + 16 kk holds D[kk + 4r][n]
                    const double v = D[(size_t)nt * 256 + (kk + 4 * r) * 16 + nn];
                    if (o + r < 0 || o + r >= (int64_t)y.size()) ++g_oob;
                    else y[(size_t)(o + r)] = v;
                }
            }
        }
    }
    // k_perm_out
    for (int64_t i = 0; i < n * nrhs; ++i) {
        const int64_t rhs = i / n, vm = i % n;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        const int64_t o = row < nflow ? tile_pos_h(row, (int)rhs, b, R) : row * R + rhs;
        Y_vm[rhs * n + vm] = at(y, o);
    }
    return g_oob;
}

// Quad path (b = 8): k_quad_vals, and k_spmm_quad's lane operand fetches
// and v_mfma_f64_4x4x4f64 placement (block blk: A[blk][m][k] from lane
// m + 4 blk + 16 k, B[blk][k][n] from lane n + 4 blk + 16 k, D[blk][m][n] in
// lane n + 4 blk + 16 m), in the tile vector layout.
extern "C" int emu_spmm_quad_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                 const double *nzval, int64_t ncell, int nrhs, int R, int vpl,
                                 const double *X_vm, double *Y_vm) {
    g_oob = 0;
    const int b = 8;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    const int64_t nflow = (int64_t)b * ncell;
    std::vector<double> x((size_t)n * R, -1e300), y((size_t)n * R, 0.0);
    for (int64_t i = 0; i < n * R; ++i) {
        const int64_t row = i / R, rem = i % R;
        const int64_t o = row < nflow ? tile_pos_h(row, (int)rem, b, R) : row * R + rem;
        if (o < 0 || o >= (int64_t)x.size()) { ++g_oob; continue; }
        if (rem >= nrhs) { x[(size_t)o] = 0.0; continue; }
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        x[(size_t)o] = X_vm[rem * n + vm];
    }
    // k_quad_vals
    const int CPS = 8;
    const int64_t total = (int64_t)s.val.size();
    std::vector<double> valq((size_t)total, 0.0);
    for (int64_t i = 0; i < total; ++i) {
        const int64_t blkidx = i >> 6;
        const int within = (int)(i & 63);
        const int64_t slot = blkidx / CPS;
        const int q = (int)(blkidx - slot * CPS);
        const int g = within >> 2, kc = (within >> 1) & 1, rq = within & 1;
        const int m = g & 3, k = g >> 2;
        const int row = 4 * rq + m, colk = 4 * kc + k;
        valq[(size_t)i] = at(s.val, slot * CPS * 64 + ((int64_t)(colk / vpl) * 64 + q * 8 + row) * vpl + colk % vpl);
    }
    const int NT = R >= 16 ? R / 16 : 1, WN = R >= 16 ? 16 : 8;
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t slice = cell / CPS, q = cell - slice * CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        const int64_t cbase = cell * b * R;
        double acc[4][2][64];          // [nt][rq][lane]
        for (int nt = 0; nt < 4; ++nt)
            for (int rq = 0; rq < 2; ++rq)
                for (int l = 0; l < 64; ++l) acc[nt][rq][l] = 0.0;
        for (int64_t t = 0; t < width; ++t) {
            const int64_t c = at(s.col, p0 + t * CPS + q);
            double a[64][4], bx[64][4][2];
            for (int l = 0; l < 64; ++l) {
                const int m4 = l & 3, k4 = l >> 4, g = m4 + 4 * k4;
                const int jl = R >= 16 ? (l & 15) : (l & 7);
                for (int e = 0; e < 4; ++e) a[l][e] = at(valq, (p0 + t * CPS + q) * 64 + g * 4 + e);
                for (int nt = 0; nt < NT; ++nt)
                    for (int e = 0; e < 2; ++e)
                        bx[l][nt][e] = at(x, c * b * R + ((int64_t)(k4 * NT + nt) * WN + jl) * 2 + e);
            }
            // one 4x4x4 instruction: operands per lane -> D per lane
            auto mfma4 = [](const double *opa, const double *opb, double *d) {
                for (int blk = 0; blk < 4; ++blk)
                    for (int m = 0; m < 4; ++m)
                        for (int nn = 0; nn < 4; ++nn) {
                            double sacc = 0.0;
                            for (int k = 0; k < 4; ++k)
                                sacc += opa[m + 4 * blk + 16 * k] * opb[nn + 4 * blk + 16 * k];
                            d[nn + 4 * blk + 16 * m] += sacc;
                        }
            };
            if (R >= 16) {
                for (int rq = 0; rq < 2; ++rq)
                    for (int kc = 0; kc < 2; ++kc)
                        for (int nt = 0; nt < NT; ++nt) {
                            double opa[64], opb[64];
                            for (int l = 0; l < 64; ++l) {
                                opa[l] = a[l][kc * 2 + rq];
                                opb[l] = bx[l][nt][kc];
                            }
                            mfma4(opa, opb, acc[nt][rq]);
                        }
            } else {
                for (int rq = 0; rq < 2; ++rq) {
                    double opa[64], opb[64];
                    for (int l = 0; l < 64; ++l) {
                        const int kcl = (l >> 3) & 1;
                        opa[l] = a[l][kcl * 2 + rq];
                        opb[l] = bx[l][0][kcl];
                    }
                    mfma4(opa, opb, acc[0][rq]);
                }
            }
        }
        if (R < 16)
            for (int rq = 0; rq < 2; ++rq) {
                double tmp[64];
                for (int l = 0; l < 64; ++l) tmp[l] = acc[0][rq][l] + acc[0][rq][l ^ 8];
                for (int l = 0; l < 64; ++l) acc[0][rq][l] = tmp[l];
            }
        for (int l = 0; l < 64; ++l) {
            if (R < 16 && ((l >> 3) & 1)) continue;
            const int k4 = l >> 4;
            const int jl = R >= 16 ? (l & 15) : (l & 7);
            for (int nt = 0; nt < NT; ++nt) {
                const int64_t o = cbase + ((int64_t)(k4 * NT + nt) * WN + jl) * 2;
                for (int rq = 0; rq < 2; ++rq) {
                    if (o + rq < 0 || o + rq >= (int64_t)y.size()) ++g_oob;
                    else y[(size_t)(o + rq)] = acc[nt][rq][l];
                }
            }
        }
    }
    for (int64_t i = 0; i < n * nrhs; ++i) {
        const int64_t rhs = i / n, vm = i % n;
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        const int64_t o = row < nflow ? tile_pos_h(row, (int)rhs, b, R) : row * R + rhs;
        Y_vm[rhs * n + vm] = at(y, o);
    }
    return g_oob;
}

// k_vals_to_f32's index map replayed over an FP64 (two-per-lane) interleave
// and compared with the host's own four-per-lane interleave of the same
// slots: returns the number of mismatching or out-of-range values.
extern "C" int emu_vals_to_f32_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                   const double *nzval, int b, int64_t ncell) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    HostSliced s2, s4;
    bsr_to_sliced(fwd, s2);
    bsr_to_sliced(fwd, s4);
    interleave_lanes(s2, 2);
    interleave_lanes(s4, 4);
    const int64_t total = s2.slice_ptr[s2.nslice] * b * b;
    int bad = (int64_t)s2.val.size() == total ? 0 : 1;
    for (int64_t i = 0; i < total; ++i) {
        const int64_t k = 4 * (i >> 8) + (i & 3);
        const int64_t lane = (i & 255) >> 2;
        const float v = (float)at(s2.val, ((k >> 1) * 64 + lane) * 2 + (k & 1));
        if (v != (float)at(s4.val, i)) ++bad;
    }
    return bad + g_oob;
}

// The upload path's layout: bsr_to_sliced_map (+ bsr_transpose_map for the
// adjoint) and k_slice_vals' index map replayed on the CPU, against the host
// layout bsr_to_sliced + interleave_lanes (of bsr_transpose for the
// adjoint).  Returns the number of differing structure entries or values.
extern "C" int emu_slice_vals_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                  const double *nzval, int b, int64_t ncell, int vpl,
                                  int transpose) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    HostBSR ref_bsr;
    if (transpose) bsr_transpose(fwd, ref_bsr);
    else ref_bsr = fwd;
    HostSliced ref;
    bsr_to_sliced(ref_bsr, ref);
    interleave_lanes(ref, vpl);
    HostSliced s;
    std::vector<int64_t> pad_src;
    if (transpose) {
        std::vector<int64_t> tptr, tsrc;
        std::vector<int32_t> tcol;
        bsr_transpose_map(ncell, fwd.ptr, fwd.col, tptr, tcol, tsrc);
        bsr_to_sliced_map(b, ncell, tptr, tcol, &tsrc, s, pad_src);
    } else {
        bsr_to_sliced_map(b, ncell, fwd.ptr, fwd.col, nullptr, s, pad_src);
    }
    int bad = 0;
    if (s.slice_ptr != ref.slice_ptr || s.col != ref.col || s.nslice != ref.nslice) ++bad;
    const int bb = b * b, cps = 64 / b, slot_len = 64 * b;
    const int64_t total = s.slice_ptr[s.nslice] * bb;
    if ((int64_t)ref.val.size() != total) ++bad;
    for (int64_t i = 0; i < total; ++i) {
        const int64_t slot = i / slot_len;
        const int j = (int)(i - slot * slot_len);
        const int kq = j / (64 * vpl), rem = j - kq * 64 * vpl;
        const int lane = rem / vpl, k = kq * vpl + rem % vpl;
        const int q = lane / b, r = lane - q * b;
        const int64_t sb = at(pad_src, slot * cps + q);
        double v = 0.0;
        if (sb >= 0) v = at(fwd.val, sb * bb + (transpose ? k * b + r : r * b + k));
        if (v != at(ref.val, i)) ++bad;
    }
    return bad + g_oob;
}

// k_spmm_quad4 (b = 4, row-major batch = tile order at b = 4) replayed on
// the CPU with the measured v_mfma_f64_4x4x4f64 operand placement: lane
// (n, blk, k) = n + 4 blk + 16 k supplies A[m = n][k] and X[k][16 i + g],
// g = 4 blk + n (R = 8: blk & 1); D[blk][m][n] lands in lane n + 4 blk + 16 m.
extern "C" int emu_spmm_quad4_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                  const double *nzval, int64_t ncell, int nrhs, int R, int vpl,
                                  const double *X_vm, double *Y_vm) {
    g_oob = 0;
    const int b = 4, CPS = 16, BB = 16;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    std::vector<double> x((size_t)n * R, 0.0), y((size_t)n * R, 0.0);
    for (int64_t row = 0; row < n; ++row)
        for (int j = 0; j < nrhs; ++j) {
            const int64_t vm = (row % b) * ncell + row / b;
            x[(size_t)(row * R + j)] = X_vm[j * n + vm];
        }
    const int NI = R >= 16 ? R / 16 : 1;
    for (int64_t cell = 0; cell < ncell; ++cell) {
        const int64_t slice = cell / CPS, q = cell - slice * CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        const int64_t cbase = cell * b * R;
        double acc[4][64];
        for (int i = 0; i < 4; ++i)
            for (int l = 0; l < 64; ++l) acc[i][l] = 0.0;
        for (int64_t t = 0; t < width; ++t) {
            const int64_t c = at(s.col, p0 + t * CPS + q);
            for (int i = 0; i < NI; ++i) {
                double opa[64], opb[64];
                for (int l = 0; l < 64; ++l) {
                    const int n4 = l & 3, blk = (l >> 2) & 3, k4 = l >> 4;
                    const int g = R >= 16 ? 4 * blk + n4 : 4 * (blk & 1) + n4;
                    const int aoff = ((k4 / vpl) * 64 + (int)q * 4 + n4) * vpl + k4 % vpl;
                    opa[l] = at(s.val, (p0 + t * CPS) * BB + aoff);
                    opb[l] = at(x, c * b * R + (int64_t)k4 * R + 16 * i + g);
                }
                for (int blk = 0; blk < 4; ++blk)
                    for (int m = 0; m < 4; ++m)
                        for (int nn = 0; nn < 4; ++nn) {
                            double sacc = 0.0;
                            for (int k = 0; k < 4; ++k)
                                sacc += opa[m + 4 * blk + 16 * k] * opb[nn + 4 * blk + 16 * k];
                            acc[i][nn + 4 * blk + 16 * m] += sacc;
                        }
            }
        }
        for (int l = 0; l < 64; ++l) {
            const int n4 = l & 3, blk = (l >> 2) & 3, k4 = l >> 4;
            if (R < 16 && blk >= 2) continue;
            const int g = R >= 16 ? 4 * blk + n4 : 4 * (blk & 1) + n4;
            for (int i = 0; i < NI; ++i) {
                const int64_t o = cbase + (int64_t)k4 * R + 16 * i + g;
                if (o < 0 || o >= (int64_t)y.size()) { ++g_oob; continue; }
                y[(size_t)o] = acc[i][l];
            }
        }
    }
    for (int64_t row = 0; row < n; ++row)
        for (int j = 0; j < nrhs; ++j) {
            const int64_t vm = (row % b) * ncell + row / b;
            Y_vm[j * n + vm] = y[(size_t)(row * R + j)];
        }
    return g_oob;
}

// k_spmm_quad4c (b = 4, four cells per wave) replayed on the CPU: block blk
// of v_mfma_f64_4x4x4f64 is cell 4 grp + blk; lane (lo, blk, hi) = lo + 4 blk
// + 16 hi supplies A[m = lo][k = hi] of its cell and X[k = hi][16 i + 4 lo + g]
// of its cell's block column to product g; D[blk][m][n] lands in lane
// n + 4 blk + 16 m.  Slots in groups of W; lane 4u + b of the group's column
// load holds slot t0 + u of cell b and is read back by lane index
// 4 (4u + blk) (ds_bpermute's byte address).  The cells past ncell in the
// last group are computed (padded slice) but not written.
extern "C" int emu_spmm_quad4c_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                   const double *nzval, int64_t ncell, int nrhs, int R, int W,
                                   int vpl, const double *X_vm, double *Y_vm) {
    g_oob = 0;
    const int b = 4, CPS = 16, BB = 16;
    if ((R != 8 && R % 16 != 0) || 4 * W > 64 || W < 1) return -2;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    std::vector<double> x((size_t)n * R, 0.0), y((size_t)n * R, 0.0);
    for (int64_t row = 0; row < n; ++row)
        for (int j = 0; j < nrhs; ++j) {
            const int64_t vm = (row % b) * ncell + row / b;
            x[(size_t)(row * R + j)] = X_vm[j * n + vm];
        }
    const int VPL = R >= 16 ? 4 : 2, GW = 4 * VPL, NG = R / GW;
    const int64_t ngrp = (ncell + 3) / 4;
    for (int64_t grp = 0; grp < ngrp; ++grp) {
        const int64_t cell0 = 4 * grp;
        const int64_t slice = cell0 / CPS;
        const int q0 = (int)(cell0 - slice * CPS);
        const int64_t p0 = at(s.slice_ptr, slice);
        const int width = (int)((at(s.slice_ptr, slice + 1) - p0) / CPS);
        std::vector<double> acc((size_t)NG * 4 * 64, 0.0);   // [i][g][lane]
        for (int t0 = 0; t0 < width; t0 += W) {
            int32_t cl[64];
            for (int l = 0; l < 64; ++l) {
                const int lu = l < 4 * W ? l >> 2 : 0;
                const int tl = t0 + lu < width ? t0 + lu : width - 1;
                cl[l] = at(s.col, p0 + (int64_t)tl * CPS + q0 + (l & 3));
            }
            for (int u = 0; u < W; ++u) {
                const int t = t0 + u < width ? t0 + u : width - 1;
                double opa[64];
                double opb[64][16];   // [lane][4 i + g]
                for (int l = 0; l < 64; ++l) {
                    const int lo = l & 3, blk = (l >> 2) & 3, hi = l >> 4;
                    const int aoff = ((hi / vpl) * 64 + (q0 + blk) * 4 + lo) * vpl + hi % vpl;
                    const double a = at(s.val, (p0 + (int64_t)t * CPS) * BB + aoff);
                    opa[l] = t0 + u < width ? a : 0.0;
                    const int src = 4 * u + blk;      // bpermute source lane
                    if (src < 0 || src >= 64) { ++g_oob; continue; }
                    const int64_t c = cl[src];
                    for (int i = 0; i < NG; ++i)
                        for (int g = 0; g < VPL; ++g)
                            opb[l][4 * i + g] = at(x, (c * b + hi) * R + VPL * lo + GW * i + g);
                }
                for (int i = 0; i < NG; ++i)
                    for (int g = 0; g < VPL; ++g)
                        for (int blk = 0; blk < 4; ++blk)
                            for (int m = 0; m < 4; ++m)
                                for (int nn = 0; nn < 4; ++nn) {
                                    double sacc = 0.0;
                                    for (int k = 0; k < 4; ++k)
                                        sacc += opa[m + 4 * blk + 16 * k] *
                                                opb[nn + 4 * blk + 16 * k][4 * i + g];
                                    acc[((size_t)i * 4 + g) * 64 + nn + 4 * blk + 16 * m] += sacc;
                                }
            }
        }
        for (int l = 0; l < 64; ++l) {
            const int lo = l & 3, blk = (l >> 2) & 3, hi = l >> 4;
            if (cell0 + blk >= ncell) continue;
            const int64_t obase = ((cell0 + blk) * b + hi) * R + VPL * lo;
            for (int i = 0; i < NG; ++i)
                for (int g = 0; g < VPL; ++g) {
                    const int64_t o = obase + GW * i + g;
                    if (o < 0 || o >= (int64_t)y.size()) { ++g_oob; continue; }
                    y[(size_t)o] = acc[((size_t)i * 4 + g) * 64 + l];
                }
        }
    }
    for (int64_t row = 0; row < n; ++row)
        for (int j = 0; j < nrhs; ++j) {
            const int64_t vm = (row % b) * ncell + row / b;
            Y_vm[j * n + vm] = y[(size_t)(row * R + j)];
        }
    return g_oob;
}

// k_extra (long extra rows) replayed on the CPU: one workgroup of NW waves
// per listed row; lane (j, i) = j LPR + i of wave wv takes entries
// ptr[row] + wv EPW + j + k NW EPW and the NC doubles from v0 = i NC of the
// row's NV (row-major rows: the extra-DOF layout); partial sums are combined
// by the xor butterfly over offsets 32 .. LPR inside each wave, then over
// the waves in order.  x / y hold n rows of NV doubles; the listed rows are
// written to y (flow rows included: the emulation has no yx split).
extern "C" int emu_extra_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                             const double *nzval, int is_complex, int b, int64_t ncell, int NV,
                             int NW, const double *x, double *y) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, is_complex != 0, 1.0, b, ncell, fwd, adj, xf, xa);
    const bool cplx = is_complex != 0;
    if (cplx && NV < 2) return -2;
    const int LPR = cplx ? NV / 2 : NV, EPW = 64 / LPR, NC = cplx ? 2 : 1;
    if (64 % LPR != 0) return -3;
    std::vector<double> xv(x, x + n * NV);
    for (int64_t ri = 0; ri < (int64_t)xf.rows.size(); ++ri) {
        std::vector<double> wsum((size_t)NW * 64 * 2, 0.0);   // [wave][lane][q]
        for (int wv = 0; wv < NW; ++wv)
            for (int lane = 0; lane < 64; ++lane) {
                const int j = lane / LPR, v0 = (lane % LPR) * NC;
                double acc[2] = {0, 0};
                for (int64_t e = at(xf.ptr, ri) + wv * EPW + j; e < at(xf.ptr, ri + 1);
                     e += (int64_t)NW * EPW) {
                    const int64_t c = at(xf.col, e);
                    if (cplx) {
                        const double ar = at(xf.val, 2 * e), ai = at(xf.val, 2 * e + 1);
                        const double xr = at(xv, c * NV + v0), xi = at(xv, c * NV + v0 + 1);
                        acc[0] = ar * xr + (-ai * xi + acc[0]);
                        acc[1] = ar * xi + (ai * xr + acc[1]);
                    } else {
                        acc[0] = at(xf.val, e) * at(xv, c * NV + v0) + acc[0];
                    }
                }
                for (int q = 0; q < NC; ++q) wsum[((size_t)wv * 64 + lane) * 2 + q] = acc[q];
            }
        // butterfly inside each wave: every lane ends with its group's sum
        for (int wv = 0; wv < NW; ++wv)
            for (int off = 32; off >= LPR; off >>= 1) {
                std::vector<double> nxt(64 * 2);
                for (int lane = 0; lane < 64; ++lane)
                    for (int q = 0; q < 2; ++q)
                        nxt[lane * 2 + q] = wsum[((size_t)wv * 64 + lane) * 2 + q] +
                                            wsum[((size_t)wv * 64 + (lane ^ off)) * 2 + q];
                for (int lane = 0; lane < 64; ++lane)
                    for (int q = 0; q < 2; ++q) wsum[((size_t)wv * 64 + lane) * 2 + q] = nxt[lane * 2 + q];
            }
        const int64_t row = at(xf.rows, ri);
        for (int lane = 0; lane < LPR; ++lane) {   // j == 0 lanes of wave 0
            const int v0 = lane * NC;
            for (int q = 0; q < NC; ++q) {
                double sacc = wsum[(size_t)lane * 2 + q];
                for (int wv = 1; wv < NW; ++wv) sacc += wsum[((size_t)wv * 64 + lane) * 2 + q];
                const int64_t o = row * NV + v0 + q;
                if (o < 0 || o >= n * NV) { ++g_oob; continue; }
                y[o] = sacc;
            }
        }
    }
    return g_oob;
}

// k_spmm_pair (b = 8, two cells per product) replayed on the CPU: the pair
// map of the library's host code (bsr_pair_map), tile-ordered values and
// vectors, lane operand fetches (A row n of the pair: cell n / 8, block row
// n % 8, zero where that cell has no block in the column; B as
// k_spmm_tile), D rows 0-7 to the pair's first cell and 8-15 to its second,
// slot groups of W columns.
extern "C" int emu_spmm_pair_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                 const double *nzval, int64_t ncell, int nrhs, int R, int W,
                                 int vpl, const double *X_vm, double *Y_vm) {
    g_oob = 0;
    const int b = 8;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, 1, false, 1.0, b, ncell, fwd, adj, xf, xa);
    if (!xf.rows.empty()) return -1;
    if (R < 16 || W < 1 || W > 64) return -2;
    HostSliced s;
    bsr_to_sliced(fwd, s);
    interleave_lanes(s, vpl);
    std::vector<int64_t> pptr;
    std::vector<int32_t> pcol, pslot;
    bsr_pair_map(ncell, fwd.ptr, fwd.col, pptr, pcol, pslot);
    const int64_t nflow = (int64_t)b * ncell;
    std::vector<double> x((size_t)n * R, -1e300), y((size_t)n * R, 0.0);
    for (int64_t i = 0; i < n * R; ++i) {
        const int64_t row = i / R, rem = i % R;
        const int64_t o = tile_pos_h(row, (int)rem, b, R);
        if (o < 0 || o >= (int64_t)x.size()) { ++g_oob; continue; }
        if (rem >= nrhs) { x[(size_t)o] = 0.0; continue; }
        x[(size_t)o] = X_vm[rem * n + (row % b) * ncell + row / b];
    }
    const int CPS = 8, BB = 64, KC = 2, WN = 16, NT = R / WN;
    const int64_t slot_len = (int64_t)CPS * BB;
    const int64_t total = (int64_t)s.val.size();
    std::vector<double> valt((size_t)total, 0.0);
    for (int64_t i = 0; i < total; ++i) {
        const int64_t slot = i / slot_len;
        const int within = (int)(i - slot * slot_len);
        const int q = within / BB, rem = within - q * BB;
        const int kc = rem % KC, tmp = rem / KC;
        const int kk = tmp & 3, m = tmp >> 2;
        const int k = kk + 4 * kc;
        valt[(size_t)i] = at(s.val, slot * slot_len + ((int64_t)(k / vpl) * 64 + q * b + m) * vpl + k % vpl);
    }
    const int64_t npair = (ncell + 1) / 2;
    if ((int64_t)pptr.size() != npair + 1) return -3;
    for (int64_t pair = 0; pair < npair; ++pair) {
        const int64_t cell0 = 2 * pair;
        const bool has1 = cell0 + 1 < ncell;
        const int64_t slice = cell0 / CPS;
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t e0 = at(pptr, pair), e1 = at(pptr, pair + 1);
        const int ncol = (int)(e1 - e0);
        std::vector<double> D((size_t)NT * 256, 0.0);
        for (int t0 = 0; t0 < ncol; t0 += W) {
            int32_t cl[64], sl[64];
            for (int lane = 0; lane < 64; ++lane) {
                const int tl = t0 + (lane < W ? lane : 0);
                const int64_t ei = e0 + (tl < ncol ? tl : ncol - 1);
                cl[lane] = at(pcol, ei);
                sl[lane] = at(pslot, ei);
            }
            for (int u = 0; u < W; ++u) {
                const int64_t c = cl[u];
                if (c < 0 || c >= ncell) { ++g_oob; continue; }
                double opa[64][2], opb[4][64][2];
                for (int lane = 0; lane < 64; ++lane) {
                    const int nn = lane & 15, kk = lane >> 4, half = nn >> 3, mr = nn & 7;
                    const int q = (int)(cell0 - slice * CPS) + half;
                    const int t = half ? (int)(((uint32_t)sl[u] >> 16) & 0xFFFF) : (sl[u] & 0xFFFF);
                    const bool have = t != 0xFFFF && t0 + u < ncol;
                    for (int kc = 0; kc < KC; ++kc) {
                        opa[lane][kc] = have ? at(valt, (p0 + (int64_t)t * CPS + q) * BB + (mr * 4 + kk) * KC + kc) : 0.0;
                        for (int nt = 0; nt < NT; ++nt)
                            opb[nt][lane][kc] = at(x, c * b * R + ((int64_t)(kk * NT + nt) * WN + nn) * KC + kc);
                    }
                }
                for (int kc = 0; kc < KC; ++kc)
                    for (int nt = 0; nt < NT; ++nt)
                        for (int mm = 0; mm < 16; ++mm)
                            for (int nn = 0; nn < 16; ++nn) {
                                double sacc = 0.0;
                                for (int k = 0; k < 4; ++k)
                                    sacc += opa[mm + 16 * k][kc] * opb[nt][nn + 16 * k][kc];
                                D[(size_t)nt * 256 + mm * 16 + nn] += sacc;
                            }
            }
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int nn = lane & 15, kk = lane >> 4;
            for (int h = 0; h < 2; ++h) {
                if (h == 1 && !has1) continue;
                const int64_t cb = (cell0 + h) * b * R;
                for (int nt = 0; nt < NT; ++nt)
                    for (int r = 0; r < KC; ++r) {
                        // register 2h + r of lane n + 16 kk holds D[kk + 4 (2h + r)][n]
                        const double v = D[(size_t)nt * 256 + (kk + 4 * (2 * h + r)) * 16 + nn];
                        const int64_t o = cb + ((int64_t)(kk * NT + nt) * WN + nn) * KC + r;
                        if (o < 0 || o >= (int64_t)y.size()) ++g_oob;
                        else y[(size_t)o] = v;
                    }
            }
        }
    }
    for (int64_t i = 0; i < n * nrhs; ++i) {
        const int64_t rhs = i / n, vm = i % n;
        const int64_t row = (vm % ncell) * b + vm / ncell;
        Y_vm[rhs * n + vm] = at(y, tile_pos_h(row, (int)rhs, b, R));
    }
    return g_oob;
}

// The receptivity operator built from its parts (receptivity_from_parts on
// one per-wavenumber system S = [L B; 0 0] and the wavenumbers w) against
// the compressed form the library finds in the assembled L_ (both CSCs
// 1-based, complex, var-major): returns 0 when the forward and adjoint extra
// CSRs, both replica descriptions and the fused Horner tables are equal
// entry for entry, else the number of the first part that differs (1
// forward extras, 2 forward replica, 3 adjoint extras, 4 adjoint replica, 5
// fused tables, 6 the assembled operator was not compressed, 7 the block
// parts differ).
namespace {
bool same_extra(const HostExtra &a, const HostExtra &b) {
    return a.is_complex == b.is_complex && a.rows == b.rows && a.ptr == b.ptr &&
           a.col == b.col && a.val == b.val;
}
bool same_rep(const HostReplica &a, const HostReplica &b) {
    return a.kind == b.kind && a.c0 == b.c0 && a.p == b.p && a.nrep == b.nrep &&
           a.diag == b.diag && a.vptr == b.vptr && a.vcol == b.vcol && a.vval == b.vval;
}
}  // namespace

extern "C" int emu_receptivity_parts_equal(int64_t n_ext, const int64_t *colptr,
                                           const int64_t *rowval, const double *nzval,
                                           int64_t n_sys, const int64_t *s_colptr,
                                           const int64_t *s_rowval, const double *s_nzval,
                                           int b, int64_t ncell, int64_t n_l, int64_t nwave,
                                           const double *w, double scale) {
    HostBSR fwd, adj, sfwd, sadj;
    HostExtra xf, xa, sxf, sxa;
    csc_to_blocks(n_ext, colptr, rowval, nzval, 1, true, scale, b, ncell, fwd, adj, xf, xa);
    csc_to_blocks(n_sys, s_colptr, s_rowval, s_nzval, 1, true, scale, b, ncell, sfwd, sadj, sxf,
                  sxa);
    if (fwd.ptr != sfwd.ptr || fwd.col != sfwd.col || fwd.val != sfwd.val ||
        adj.ptr != sadj.ptr || adj.col != sadj.col || adj.val != sadj.val)
        return 7;
    HostReplica rf, ra, pf, pa;
    if (!compress_column_replicas(xf, n_ext, rf)) return 6;
    if (!compress_row_replicas(xa, n_ext, rf.c0, rf.p, rf.nrep, ra)) return 6;
    receptivity_from_parts(n_l, n_sys - n_l, nwave, w, scale, sxf, sxa, pf, pa);
    if (!same_extra(xf, sxf)) return 1;
    if (!same_rep(rf, pf)) return 2;
    if (!same_extra(xa, sxa)) return 3;
    if (!same_rep(ra, pa)) return 4;
    HostFusedRep fa, fp;
    const bool ba = build_fused_tables(xf, rf, n_ext, (int64_t)b * ncell, fa);
    const bool bp = build_fused_tables(sxf, pf, n_ext, (int64_t)b * ncell, fp);
    if (ba != bp || fa.slice_tab != fp.slice_tab || fa.lane_ptr != fp.lane_ptr ||
        fa.ecol != fp.ecol || fa.eval != fp.eval)
        return 5;
    return 0;
}

// xcd_balanced_order (the long extra rows' order at upload) for m rows of
// len[i] entries: out[new position] = old index.
extern "C" void emu_xcd_balanced_order(int64_t m, const int64_t *len, int64_t *out) {
    const std::vector<int64_t> l(len, len + m);
    const std::vector<int64_t> o = xcd_balanced_order(l);
    for (int64_t i = 0; i < m; ++i) out[i] = o[(size_t)i];
}

// One real SpMV through the shock bands' fused form (build_shock_fuse, as
// k_term_split_p<..., SH> takes it): the block term per slice lane, plus the
// slice's panel (columns pcol[cptr[k] .. cptr[k+1]), values pval[64 c +
// lane]), and the extra-DOF rows (their CSR).  info[0] = 1 if the fused
// form was built, info[1] = panels, info[2] = panel columns, info[3] =
// extra-DOF rows.  Returns the out-of-bounds access count.
extern "C" int emu_spmv_shock(int64_t n, const int64_t *colptr, const int64_t *rowval,
                              const double *nzval, int base, int b, int64_t ncell, int transpose,
                              const double *x_vm, double *y_vm, int64_t *info) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, base, false, 1.0, b, ncell, fwd, adj, xf, xa);
    const HostBSR &A = transpose ? adj : fwd;
    const HostExtra &X = transpose ? xa : xf;
    const int64_t nflow = (int64_t)b * ncell;
    HostExtra rows;
    HostShockPanels pan;
    const bool ok = build_shock_fuse(X, nflow, rows, pan);
    info[0] = ok ? 1 : 0;
    info[1] = ok ? (int64_t)pan.cptr.size() - 1 : 0;
    info[2] = ok ? (int64_t)pan.pcol.size() : 0;
    info[3] = ok ? (int64_t)rows.rows.size() : 0;
    if (!ok) return 0;
    HostSliced s;
    bsr_to_sliced(A, s);
    interleave_lanes(s, 2);
    std::vector<double> x((size_t)n), y((size_t)n, 0.0);
    for (int64_t row = 0; row < n; ++row) {
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        x[(size_t)row] = x_vm[vm];
    }
    // extra-DOF rows (leading workgroups)
    for (int64_t i = 0; i < (int64_t)rows.rows.size(); ++i) {
        double acc = 0.0;
        for (int64_t e = at(rows.ptr, i); e < at(rows.ptr, i + 1); ++e)
            acc += at(rows.val, e) * at(x, at(rows.col, e));
        const int64_t row = at(rows.rows, i);
        if (row < nflow || row >= n) ++g_oob;
        else y[(size_t)row] = acc;
    }
    // slices: block term, then the panel
    const int CPS = 64 / b, BB = b * b;
    if ((int64_t)pan.slice_tab.size() != s.nslice) ++g_oob;
    for (int64_t slice = 0; slice < s.nslice; ++slice) {
        const int64_t p0 = at(s.slice_ptr, slice);
        const int64_t width = (at(s.slice_ptr, slice + 1) - p0) / CPS;
        const int32_t pk = at(pan.slice_tab, slice);
        for (int lane = 0; lane < 64; ++lane) {
            const int q = lane / b, r = lane % b;
            double acc = 0.0;
            for (int64_t t = 0; t < width; ++t) {
                const int64_t c = at(s.col, p0 + q + t * CPS);
                const int64_t sb = (p0 + t * CPS) * BB;
                for (int k = 0; k < b; ++k)
                    acc += at(s.val, sb + ((k / 2) * 64 + lane) * 2 + k % 2) * at(x, c * b + k);
            }
            if (pk >= 0)
                for (int64_t c = at(pan.cptr, pk); c < at(pan.cptr, pk + 1); ++c)
                    acc += at(pan.pval, c * 64 + lane) * at(x, at(pan.pcol, c));
            const int64_t cell = slice * CPS + q;
            if (cell >= ncell) continue;
            y[(size_t)(cell * b + r)] = acc;
        }
    }
    for (int64_t vm = 0; vm < n; ++vm) {
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        y_vm[vm] = y[(size_t)row];
    }
    return g_oob;
}

namespace {
template <typename T>
T &atw(std::vector<T> &v, int64_t i) {
    if (i < 0 || i >= (int64_t)v.size()) {
        ++g_oob;
        static T zero{};
        return zero;
    }
    return v[(size_t)i];
}
}  // namespace

// A batched term over a shock operator at R doubles per row, in one launch:
// the leading workgroups finish the rows of build_shock_lead (the extra-DOF
// rows, then every flow row of the slices with shock columns, whole) and
// the block kernel runs the other slices.  Every row is written exactly
// once, every access is bounds-checked.  bsr_form = 1 builds the leading
// rows as the BSR upload does for the adjoint (the caller's raw blocks
// through bsr_transpose_map, read transposed) and checks them against the
// CSC upload's (blocks transposed on the host): same columns, same values.
// The FP32 copy must hold the block entries rounded to FP32 and the extra
// entries unchanged.  x_vm, y_vm: R columns of n (var-major).  info:
// [fused?, leading rows, skipped slices, FP32 mismatches, BSR-form
// mismatches].  Returns the out-of-bounds count.
extern "C" int emu_spmv_shock_batched(int64_t n, const int64_t *colptr, const int64_t *rowval,
                                      const double *nzval, int base, int b, int64_t ncell,
                                      int transpose, int R, int bsr_form, const double *x_vm,
                                      double *y_vm, int64_t *info) {
    g_oob = 0;
    HostBSR fwd, adj;
    HostExtra xf, xa;
    csc_to_blocks(n, colptr, rowval, nzval, base, false, 1.0, b, ncell, fwd, adj, xf, xa);
    const HostBSR &A = transpose ? adj : fwd;
    const HostExtra &X = transpose ? xa : xf;
    const int64_t nflow = (int64_t)b * ncell;
    HostExtra rows;
    HostShockPanels pan;
    const bool ok = build_shock_fuse(X, nflow, rows, pan);
    for (int i = 0; i < 5; ++i) info[i] = 0;
    info[0] = ok ? 1 : 0;
    if (!ok) return 0;
    HostShockLead lead;
    build_shock_lead(b, A.ptr, A.col, A.val.data(), nullptr, false, 1.0, X, rows, pan, nflow,
                     lead);
    const HostExtra &SH = lead.sh, &L = lead.fr, &L32 = lead.fr32;
    info[1] = (int64_t)(SH.rows.size() + L.rows.size());
    if (L32.rows != L.rows || L32.ptr != L.ptr || L32.col != L.col) ++info[3];
    for (size_t i = 0; i < L.rows.size(); ++i)
        for (int64_t e = at(L.ptr, (int64_t)i); e < at(L.ptr, (int64_t)i + 1); ++e) {
            const double v = at(L.val, e), v32 = at(L32.val, e);
            // block entries (flow row, flow column) rounded, extra entries unchanged
            const bool block = L.rows[i] < nflow && at(L.col, e) < nflow;
            if (v32 != (block ? (double)(float)v : v)) ++info[3];
        }
    if (bsr_form && transpose) {
        // the BSR upload's adjoint: raw forward blocks through the transpose map
        std::vector<int64_t> tptr, tsrc;
        std::vector<int32_t> tcol;
        bsr_transpose_map(ncell, fwd.ptr, fwd.col, tptr, tcol, tsrc);
        HostShockLead lb;
        build_shock_lead(b, tptr, tcol, fwd.val.data(), &tsrc, true, 1.0, X, rows, pan, nflow, lb);
        if (lb.fr.rows != L.rows || lb.fr.ptr != L.ptr || lb.fr.col != L.col ||
            lb.fr.val != L.val || lb.sh.rows != SH.rows || lb.sh.val != SH.val)
            ++info[4];
    }
    std::vector<double> x((size_t)(n * R)), y((size_t)(n * R), 0.0);
    std::vector<int> writes((size_t)n, 0);
    for (int64_t row = 0; row < n; ++row) {
        const int64_t vm = row < nflow ? (row % b) * ncell + row / b : row;
        for (int v = 0; v < R; ++v) x[(size_t)(row * R + v)] = x_vm[(int64_t)v * n + vm];
    }
    // leading workgroups, one row each, as both kernels map them: the
    // extra-DOF rows, then the whole flow rows
    const int64_t nsh = (int64_t)SH.rows.size(), nfr = (int64_t)L.rows.size();
    const int64_t nlead_wg = (nsh + nfr + 7) / 8 * 8;
    std::vector<int> done_sh((size_t)nsh, 0), done_fr((size_t)nfr, 0);
    auto run_row = [&](const HostExtra &C, int64_t i) {
        const int64_t row = at(C.rows, i);
        if (row < 0 || row >= n) ++g_oob;
        atw(writes, row) += 1;
        for (int v = 0; v < R; ++v) {
            double acc = 0.0;
            for (int64_t e = at(C.ptr, i); e < at(C.ptr, i + 1); ++e)
                acc += at(C.val, e) * at(x, (int64_t)at(C.col, e) * R + v);
            atw(y, row * R + v) = acc;
        }
    };
    for (int64_t bid = 0; bid < nlead_wg; ++bid) {
        if (bid < nsh) {
            atw(done_sh, bid) += 1;
            run_row(SH, bid);
        } else if (bid - nsh < nfr) {
            atw(done_fr, bid - nsh) += 1;
            run_row(L, bid - nsh);
        }
    }
    for (int d : done_sh)
        if (d != 1) ++g_oob;
    for (int d : done_fr)
        if (d != 1) ++g_oob;
    // the block kernel over the slices it does not skip
    const int64_t nslice = (nflow + 63) / 64;
    if ((int64_t)pan.slice_tab.size() != nslice) ++g_oob;
    for (int64_t slice = 0; slice < nslice; ++slice) {
        if (at(pan.slice_tab, slice) >= 0) {
            ++info[2];
            continue;
        }
        for (int64_t row = slice * 64; row < slice * 64 + 64 && row < nflow; ++row) {
            const int64_t cell = row / b;
            const int r = (int)(row % b);
            atw(writes, row) += 1;
            for (int v = 0; v < R; ++v) {
                double acc = 0.0;
                for (int64_t k = at(A.ptr, cell); k < at(A.ptr, cell + 1); ++k) {
                    const int64_t c = at(A.col, k);
                    for (int j = 0; j < b; ++j)
                        acc += at(A.val, k * b * b + r * b + j) * at(x, (c * b + j) * R + v);
                }
                atw(y, row * R + v) = acc;
            }
        }
    }
    for (int64_t row = 0; row < n; ++row)
        if (writes[(size_t)row] != 1) ++g_oob;
    for (int64_t vm = 0; vm < n; ++vm) {
        const int64_t row = vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
        for (int v = 0; v < R; ++v) y_vm[(int64_t)v * n + vm] = y[(size_t)(row * R + v)];
    }
    return g_oob;
}
