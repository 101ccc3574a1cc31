// Host-side layout conversion: the reference's variable-major CSC
// (a Julia SparseMatrixCSC from LINEARIZE_EQUATIONS_NO_DISCONTINUITY.jl:276-285)
// into cell-major block storage for the HIP kernels.
//
// Index map: var-major global index vm = k*ncell + c (variable k of cell c)
// for vm < b*ncell; cell-major index cm = c*b + k.  Extra DOFs keep their
// index.  Entries whose row and column are both flow unknowns become b x b
// blocks (real part); everything else goes to the "extra" CSR.
#include "hymor_internal.hpp"

#include <algorithm>
#include <exception>
#include <functional>
#include <thread>
#include <numeric>
#include <stdexcept>

namespace hymor {

namespace {

struct BEnt {
    int64_t cell_r;
    int32_t i, j;
    double v;
};

inline int64_t vm_to_cm(int64_t vm, int64_t ncell, int b, int64_t nflow) {
    return vm < nflow ? (vm % ncell) * b + vm / ncell : vm;
}

}  // namespace

namespace {

// Host threads for the upload's conversion: the hardware's, at most 16,
// and at most one per 4 096 cell columns.
int host_threads(int64_t ncell) {
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t by_size = std::max<int64_t>(1, ncell / 4096);
    return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)(hw ? hw : 1), 16, by_size}));
}

// f(t) for t = 0 .. nth-1, each on a thread of its own (t = 0 on the
// caller's); an exception in any is rethrown here.
void parallel_for(int nth, const std::function<void(int)> &f) {
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> err((size_t)nth);
    for (int t = 1; t < nth; ++t)
        pool.emplace_back([&, t] {
            try {
                f(t);
            } catch (...) {
                err[(size_t)t] = std::current_exception();
            }
        });
    try {
        f(0);
    } catch (...) {
        err[0] = std::current_exception();
    }
    for (std::thread &th : pool) th.join();
    for (const std::exception_ptr &e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace

void csc_to_blocks(int64_t n, const int64_t *colptr, const int64_t *rowval,
                   const double *nzval, int index_base, bool is_complex, double scale,
                   int b, int64_t ncell, HostBSR &fwd, HostBSR &adj,
                   HostExtra &xf, HostExtra &xa) {
    const int64_t nflow = (int64_t)b * ncell;
    const int bb = b * b;
    const int64_t base = index_base;
    // basic validation of the CSC arrays
    if (colptr[0] != base) throw std::invalid_argument("colptr[0] must equal the index base");
    for (int64_t j = 0; j < n; ++j)
        if (colptr[j + 1] < colptr[j]) throw std::invalid_argument("colptr is not monotone");
    const int64_t nnz = colptr[n] - base;
    for (int64_t k = 0; k < nnz; ++k)
        if (rowval[k] < base || rowval[k] >= n + base)
            throw std::invalid_argument("rowval out of range");

    const int vstride = is_complex ? 2 : 1;

    // block-compressed columns (BSC) over cell columns, contiguous ranges of
    // cell columns per host thread, concatenated in order (the same arrays
    // as one pass over the columns)
    const int nth = host_threads(ncell);
    struct Part {
        std::vector<int64_t> cnt;        // blocks per cell column of the range
        std::vector<int64_t> row;        // block rows
        std::vector<double> val;         // b*b per block
        std::vector<int64_t> xr, xc;     // extra entries
        std::vector<double> xv;
    };
    std::vector<Part> parts((size_t)nth);
    auto c_lo = [&](int t) { return ncell * t / nth; };
    parallel_for(nth, [&](int t) {
        Part &P = parts[(size_t)t];
        std::vector<BEnt> tmp;
        const int64_t c0 = c_lo(t), c1 = c_lo(t + 1);
        P.cnt.assign((size_t)(c1 - c0), 0);
        for (int64_t cc = c0; cc < c1; ++cc) {
            tmp.clear();
            for (int k = 0; k < b; ++k) {
                const int64_t vmc = (int64_t)k * ncell + cc;
                const int64_t cmc = cc * b + k;
                for (int64_t e = colptr[vmc] - base; e < colptr[vmc + 1] - base; ++e) {
                    const int64_t vmr = rowval[e] - base;
                    const int64_t cmr = vm_to_cm(vmr, ncell, b, nflow);
                    const double re = nzval[e * vstride] * scale;
                    const double im = is_complex ? nzval[e * vstride + 1] * scale : 0.0;
                    if (cmr < nflow) {
                        tmp.push_back({cmr / b, (int32_t)(cmr % b), k, re});
                        if (is_complex && im != 0.0) {
                            P.xr.push_back(cmr); P.xc.push_back(cmc);
                            P.xv.push_back(0.0); P.xv.push_back(im);
                        }
                    } else {
                        P.xr.push_back(cmr); P.xc.push_back(cmc);
                        P.xv.push_back(re);
                        if (is_complex) P.xv.push_back(im);
                    }
                }
            }
            std::stable_sort(tmp.begin(), tmp.end(),
                             [](const BEnt &a, const BEnt &c) { return a.cell_r < c.cell_r; });
            int64_t last = -1, nb = 0;
            for (const BEnt &e : tmp) {
                if (e.cell_r != last) {
                    P.row.push_back(e.cell_r);
                    P.val.resize(P.val.size() + bb, 0.0);
                    last = e.cell_r;
                    ++nb;
                }
                // duplicates are summed, as sparse(I,J,V) does
                P.val[P.val.size() - bb + e.i * b + e.j] += e.v;
            }
            P.cnt[(size_t)(cc - c0)] = nb;
        }
    });
    std::vector<int64_t> bsc_ptr(ncell + 1, 0);
    std::vector<int64_t> part_blk((size_t)nth + 1, 0);   // first block of each range
    for (int t = 0; t < nth; ++t) {
        const Part &P = parts[(size_t)t];
        for (int64_t i = 0; i < (int64_t)P.cnt.size(); ++i)
            bsc_ptr[c_lo(t) + i + 1] = bsc_ptr[c_lo(t) + i] + P.cnt[(size_t)i];
        part_blk[(size_t)t + 1] = part_blk[(size_t)t] + (int64_t)P.row.size();
    }
    const int64_t nnzb = bsc_ptr[ncell];
    std::vector<int64_t> xr, xc;
    std::vector<double> xv;
    for (const Part &P : parts) {
        xr.insert(xr.end(), P.xr.begin(), P.xr.end());
        xc.insert(xc.end(), P.xc.begin(), P.xc.end());
        xv.insert(xv.end(), P.xv.begin(), P.xv.end());
    }
    // columns of extra DOFs: every entry is extra
    for (int64_t vmc = nflow; vmc < n; ++vmc) {
        for (int64_t e = colptr[vmc] - base; e < colptr[vmc + 1] - base; ++e) {
            const int64_t vmr = rowval[e] - base;
            xr.push_back(vm_to_cm(vmr, ncell, b, nflow));
            xc.push_back(vmc);
            xv.push_back(nzval[e * vstride] * scale);
            if (is_complex) xv.push_back(nzval[e * vstride + 1] * scale);
        }
    }

    // adjoint BSR = BSC with each block transposed (real parts: no conjugate)
    adj.b = b;
    adj.ncell = ncell;
    adj.ptr = bsc_ptr;
    adj.col.resize(nnzb);
    adj.val.resize((size_t)nnzb * bb);
    parallel_for(nth, [&](int t) {
        const Part &P = parts[(size_t)t];
        const int64_t k0 = part_blk[(size_t)t];
        for (int64_t q = 0; q < (int64_t)P.row.size(); ++q) {
            const int64_t k = k0 + q;
            adj.col[k] = (int32_t)P.row[(size_t)q];
            for (int i = 0; i < b; ++i)
                for (int j = 0; j < b; ++j)
                    adj.val[(size_t)k * bb + j * b + i] = P.val[(size_t)q * bb + i * b + j];
        }
    });
    // forward BSR = BSC regrouped by block row (columns ascending): each
    // range counts its blocks per block row; a block row's blocks are placed
    // range by range, so the order is that of one pass over the columns
    std::vector<std::vector<int64_t>> rc((size_t)nth);
    parallel_for(nth, [&](int t) {
        std::vector<int64_t> &c = rc[(size_t)t];
        c.assign((size_t)ncell, 0);
        for (const int64_t r : parts[(size_t)t].row) c[(size_t)r]++;
    });
    fwd.b = b;
    fwd.ncell = ncell;
    fwd.ptr.assign(ncell + 1, 0);
    for (int64_t r = 0; r < ncell; ++r) {
        int64_t acc = fwd.ptr[r];
        for (int t = 0; t < nth; ++t) {
            const int64_t c = rc[(size_t)t][(size_t)r];
            rc[(size_t)t][(size_t)r] = acc;   // this range's first slot in row r
            acc += c;
        }
        fwd.ptr[r + 1] = acc;
    }
    fwd.col.resize(nnzb);
    fwd.val.resize((size_t)nnzb * bb);
    parallel_for(nth, [&](int t) {
        const Part &P = parts[(size_t)t];
        std::vector<int64_t> &pos = rc[(size_t)t];
        int64_t q = 0;
        for (int64_t cc = c_lo(t); cc < c_lo(t + 1); ++cc) {
            for (int64_t k = bsc_ptr[cc]; k < bsc_ptr[cc + 1]; ++k, ++q) {
                const int64_t d = pos[(size_t)P.row[(size_t)q]]++;
                fwd.col[d] = (int32_t)cc;
                std::copy(P.val.begin() + (size_t)q * bb, P.val.begin() + (size_t)(q + 1) * bb,
                          fwd.val.begin() + (size_t)d * bb);
            }
        }
    });
    extra_from_coo(n, nflow, xr, xc, xv, is_complex, false, xf);
    extra_from_coo(n, nflow, xr, xc, xv, is_complex, true, xa);
}

void bsr_transpose(const HostBSR &a, HostBSR &t) {
    const int b = a.b, bb = b * b;
    const int64_t ncell = a.ncell, nnzb = a.nnzb();
    t.b = b;
    t.ncell = ncell;
    t.ptr.assign(ncell + 1, 0);
    for (int64_t k = 0; k < nnzb; ++k) t.ptr[(int64_t)a.col[k] + 1]++;
    for (int64_t c = 0; c < ncell; ++c) t.ptr[c + 1] += t.ptr[c];
    t.col.resize(nnzb);
    t.val.resize((size_t)nnzb * bb);
    std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
    for (int64_t r = 0; r < ncell; ++r) {
        for (int64_t k = a.ptr[r]; k < a.ptr[r + 1]; ++k) {
            const int64_t d = pos[a.col[k]]++;
            t.col[d] = (int32_t)r;
            for (int i = 0; i < b; ++i)
                for (int j = 0; j < b; ++j)
                    t.val[(size_t)d * bb + j * b + i] = a.val[(size_t)k * bb + i * b + j];
        }
    }
}

void bsr_transpose_map(int64_t ncell, const std::vector<int64_t> &ptr,
                       const std::vector<int32_t> &col, std::vector<int64_t> &tptr,
                       std::vector<int32_t> &tcol, std::vector<int64_t> &tsrc) {
    // the block order of bsr_transpose: inside a transposed row, ascending
    // forward row
    const int64_t nnzb = ptr[ncell];
    tptr.assign(ncell + 1, 0);
    for (int64_t k = 0; k < nnzb; ++k) tptr[(int64_t)col[k] + 1]++;
    for (int64_t c = 0; c < ncell; ++c) tptr[c + 1] += tptr[c];
    tcol.resize(nnzb);
    tsrc.resize(nnzb);
    std::vector<int64_t> pos(tptr.begin(), tptr.end() - 1);
    for (int64_t r = 0; r < ncell; ++r)
        for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int64_t d = pos[col[k]]++;
            tcol[d] = (int32_t)r;
            tsrc[d] = k;
        }
}

void bsr_pair_map(int64_t ncell, const std::vector<int64_t> &ptr,
                  const std::vector<int32_t> &col, std::vector<int64_t> &pptr,
                  std::vector<int32_t> &pcol, std::vector<int32_t> &pslot) {
    const int64_t npair = (ncell + 1) / 2;
    pptr.assign((size_t)npair + 1, 0);
    pcol.clear();
    pslot.clear();
    struct Ent {
        int32_t c, ta, tb;
    };
    std::vector<Ent> u;
    for (int64_t p = 0; p < npair; ++p) {
        u.clear();
        for (int side = 0; side < 2; ++side) {
            const int64_t cell = 2 * p + side;
            if (cell >= ncell) break;
            const int64_t k0 = ptr[(size_t)cell], k1 = ptr[(size_t)cell + 1];
            if (k1 - k0 >= kPairNone) throw std::invalid_argument("block row too long for the pair map");
            for (int64_t k = k0; k < k1; ++k) {
                const int32_t c = col[(size_t)k], t = (int32_t)(k - k0);
                bool found = false;
                for (Ent &e : u)
                    if (e.c == c && (side == 0 ? e.ta : e.tb) == kPairNone) {
                        (side == 0 ? e.ta : e.tb) = t;
                        found = true;
                        break;
                    }
                if (!found) u.push_back(side == 0 ? Ent{c, t, kPairNone} : Ent{c, kPairNone, t});
            }
        }
        std::stable_sort(u.begin(), u.end(), [](const Ent &a, const Ent &b2) { return a.c < b2.c; });
        for (const Ent &e : u) {
            pcol.push_back(e.c);
            pslot.push_back((int32_t)((uint32_t)e.ta | ((uint32_t)e.tb << 16)));
        }
        pptr[(size_t)p + 1] = (int64_t)pcol.size();
    }
}

void bsr_to_sliced_map(int b, int64_t ncell, const std::vector<int64_t> &ptr,
                       const std::vector<int32_t> &col, const std::vector<int64_t> *src,
                       HostSliced &s, std::vector<int64_t> &pad_src) {
    // the slices, slots and padding of bsr_to_sliced, without the values
    const int cps = 64 / b;
    s.b = b;
    s.cps = cps;
    s.ncell = ncell;
    s.nslice = (ncell + cps - 1) / cps;
    s.nnzb = ptr[ncell];
    s.slice_ptr.assign(s.nslice + 1, 0);
    for (int64_t sl = 0; sl < s.nslice; ++sl) {
        int64_t w = 0;
        for (int q = 0; q < cps; ++q) {
            const int64_t cell = sl * cps + q;
            if (cell < ncell) w = std::max(w, ptr[cell + 1] - ptr[cell]);
        }
        s.slice_ptr[sl + 1] = s.slice_ptr[sl] + w * cps;
    }
    const int64_t tot = s.slice_ptr[s.nslice];
    s.col.assign(tot, 0);
    s.val.clear();
    pad_src.assign(tot, -1);
    for (int64_t sl = 0; sl < s.nslice; ++sl) {
        const int64_t w = (s.slice_ptr[sl + 1] - s.slice_ptr[sl]) / cps;
        for (int q = 0; q < cps; ++q) {
            const int64_t cell = sl * cps + q;
            const int64_t len = cell < ncell ? ptr[cell + 1] - ptr[cell] : 0;
            const int32_t self = (int32_t)(cell < ncell ? cell : 0);
            for (int64_t t = 0; t < w; ++t) {
                const int64_t d = s.slice_ptr[sl] + t * cps + q;
                if (t < len) {
                    const int64_t k = ptr[cell] + t;
                    s.col[d] = col[k];
                    pad_src[d] = src ? (*src)[k] : k;
                } else {
                    s.col[d] = self;
                }
            }
        }
    }
}

void extra_from_coo(int64_t n, int64_t nflow, const std::vector<int64_t> &r,
                    const std::vector<int64_t> &c, const std::vector<double> &v,
                    bool is_complex, bool adjoint, HostExtra &out) {
    const int vs = is_complex ? 2 : 1;
    const int64_t m = (int64_t)r.size();
    const std::vector<int64_t> &rr = adjoint ? c : r;
    const std::vector<int64_t> &cc = adjoint ? r : c;
    // entries in (row, column) order, ties in input order: a stable counting
    // sort by row, then a stable sort by column inside each row
    std::vector<int64_t> idx(m), start((size_t)n + 1, 0);
    for (int64_t k = 0; k < m; ++k) start[(size_t)rr[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) start[(size_t)i + 1] += start[(size_t)i];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t k = 0; k < m; ++k) idx[(size_t)fill[(size_t)rr[k]]++] = k;
    }
    for (int64_t i = 0; i < n; ++i) {
        const auto b0 = idx.begin() + start[(size_t)i], b1 = idx.begin() + start[(size_t)i + 1];
        if (b1 - b0 > 1 && !std::is_sorted(b0, b1, [&](int64_t a, int64_t b2) {
                return cc[a] < cc[b2];
            }))
            std::stable_sort(b0, b1, [&](int64_t a, int64_t b2) { return cc[a] < cc[b2]; });
    }
    out.is_complex = is_complex;
    out.rows.clear();
    out.ptr.assign(1, 0);
    out.col.clear();
    out.val.clear();
    out.col.reserve(m);
    out.val.reserve((size_t)m * vs);
    // merge entry rows with every extra-DOF row (those always need a result)
    int64_t e = 0;
    int64_t next_extra = nflow;
    while (e < m || next_extra < n) {
        const int64_t re = e < m ? rr[idx[e]] : INT64_MAX;
        const int64_t row = std::min(re, next_extra < n ? next_extra : INT64_MAX);
        if (row == next_extra) ++next_extra;
        while (e < m && rr[idx[e]] == row) {
            const int64_t k = idx[e];
            out.col.push_back((int32_t)cc[k]);
            out.val.push_back(v[(size_t)k * vs]);
            if (is_complex) out.val.push_back(adjoint ? -v[(size_t)k * vs + 1] : v[(size_t)k * vs + 1]);
            ++e;
        }
        out.rows.push_back(row);
        out.ptr.push_back((int64_t)out.col.size());
    }
}

void bsr_to_sliced(const HostBSR &a, HostSliced &s) {
    const int b = a.b, bb = b * b;
    const int cps = 64 / b;
    s.b = b;
    s.cps = cps;
    s.ncell = a.ncell;
    s.nslice = (a.ncell + cps - 1) / cps;
    s.nnzb = a.nnzb();
    s.slice_ptr.assign(s.nslice + 1, 0);
    for (int64_t sl = 0; sl < s.nslice; ++sl) {
        int64_t w = 0;
        for (int q = 0; q < cps; ++q) {
            const int64_t cell = sl * cps + q;
            if (cell < a.ncell) w = std::max(w, a.ptr[cell + 1] - a.ptr[cell]);
        }
        s.slice_ptr[sl + 1] = s.slice_ptr[sl] + w * cps;
    }
    const int64_t tot = s.slice_ptr[s.nslice];
    s.col.assign(tot, 0);
    s.val.assign((size_t)tot * bb, 0.0);
    for (int64_t sl = 0; sl < s.nslice; ++sl) {
        const int64_t w = (s.slice_ptr[sl + 1] - s.slice_ptr[sl]) / cps;
        for (int q = 0; q < cps; ++q) {
            const int64_t cell = sl * cps + q;
            const int64_t len = cell < a.ncell ? a.ptr[cell + 1] - a.ptr[cell] : 0;
            const int32_t self = (int32_t)(cell < a.ncell ? cell : 0);
            for (int64_t t = 0; t < w; ++t) {
                const int64_t d = s.slice_ptr[sl] + t * cps + q;
                if (t < len) {
                    const int64_t k = a.ptr[cell] + t;
                    s.col[d] = a.col[k];
                    std::copy(a.val.begin() + (size_t)k * bb, a.val.begin() + (size_t)(k + 1) * bb,
                              s.val.begin() + (size_t)d * bb);
                } else {
                    s.col[d] = self;
                }
            }
        }
    }
}

void interleave_lanes(HostSliced &s, int vpl) {
    // Inside each slot of a slice the CPS blocks (64 block rows = 64 lanes)
    // are re-ordered so that value k of lane (q, r) sits at
    // ((k / vpl) * 64 + q*b + r) * vpl + k % vpl: the lanes' 16-byte chunks
    // for one k-range are consecutive, so each wave-wide 16-byte load reads
    // one contiguous 1 KB run.
    const int b = s.b, bb = b * b, cps = s.cps;
    const int64_t nslots = s.slice_ptr[s.nslice] / cps;
    std::vector<double> tmp((size_t)cps * bb);
    for (int64_t g = 0; g < nslots; ++g) {
        double *base = s.val.data() + (size_t)g * cps * bb;
        for (int q = 0; q < cps; ++q)
            for (int r = 0; r < b; ++r)
                for (int k = 0; k < b; ++k)
                    tmp[(size_t)(((k / vpl) * 64 + q * b + r) * vpl + k % vpl)] =
                        base[q * bb + r * b + k];
        std::copy(tmp.begin(), tmp.end(), base);
    }
}

namespace {

// Index of the listed row `row` in x, or -1.
int64_t listed_index(const HostExtra &x, int64_t row) {
    auto it = std::lower_bound(x.rows.begin(), x.rows.end(), row);
    return (it != x.rows.end() && *it == row) ? (int64_t)(it - x.rows.begin()) : -1;
}

bool same_value(const HostExtra &x, int64_t e, int64_t f, int vs) {
    for (int q = 0; q < vs; ++q)
        if (x.val[(size_t)(e * vs + q)] != x.val[(size_t)(f * vs + q)]) return false;
    return true;
}

// Copy of x keeping listed rows [0, keep_rows) only, with the entries each
// row's `edit` callback emits (col, value index) in place of its own.
template <typename Edit>
HostExtra rebuild(const HostExtra &x, int64_t keep_rows, Edit &&edit) {
    const int vs = x.is_complex ? 2 : 1;
    HostExtra out;
    out.is_complex = x.is_complex;
    out.ptr.assign(1, 0);
    for (int64_t i = 0; i < keep_rows; ++i) {
        edit(i, [&](int32_t col, int64_t e) {
            out.col.push_back(col);
            for (int q = 0; q < vs; ++q) out.val.push_back(x.val[(size_t)(e * vs + q)]);
        });
        out.rows.push_back(x.rows[(size_t)i]);
        out.ptr.push_back((int64_t)out.col.size());
    }
    return out;
}

}  // namespace

bool compress_column_replicas(HostExtra &x, int64_t n, HostReplica &rep) {
    const int vs = x.is_complex ? 2 : 1;
    const int64_t nr = (int64_t)x.rows.size();
    // trailing run of rows [c0, n), listed consecutively, each holding at
    // most its own diagonal entry
    int64_t i0 = nr;
    while (i0 > 0) {
        const int64_t i = i0 - 1, row = x.rows[(size_t)i];
        const int64_t len = x.ptr[(size_t)i + 1] - x.ptr[(size_t)i];
        if (row != n - (nr - i)) break;
        if (len > 1 || (len == 1 && x.col[(size_t)x.ptr[(size_t)i]] != row)) break;
        i0 = i;
    }
    if (i0 == nr) return false;
    const int64_t c0 = x.rows[(size_t)i0], M = n - c0;
    if (M < 2 || n + M >= (int64_t)INT32_MAX) return false;
    // per earlier row, where its entries in columns >= c0 start (sorted rows)
    std::vector<int64_t> tail((size_t)i0);
    bool any = false;
    for (int64_t i = 0; i < i0; ++i) {
        const int64_t b0 = x.ptr[(size_t)i], b1 = x.ptr[(size_t)i + 1];
        int64_t e = b0;
        while (e < b1 && x.col[(size_t)e] < c0) ++e;
        tail[(size_t)i] = e;
        any = any || e < b1;
    }
    if (!any) return false;
    // the largest number of copies nrep >= 2 (p = M / nrep) under which
    // every row's tail is nrep copies of its first p columns' entries
    int64_t nrep = 0, p = 0;
    for (int64_t N = M; N >= 2 && !nrep; --N) {
        if (M % N) continue;
        const int64_t pp = M / N;
        bool ok = true;
        for (int64_t i = 0; i < i0 && ok; ++i) {
            const int64_t t0 = tail[(size_t)i], len = x.ptr[(size_t)i + 1] - t0;
            if (len == 0) continue;
            if (len % N) { ok = false; break; }
            const int64_t m = len / N;
            for (int64_t j = 0; j < m && ok; ++j) {
                const int64_t off = x.col[(size_t)(t0 + j)] - c0;
                if (off >= pp) { ok = false; break; }
                for (int64_t k = 1; k < N; ++k) {
                    const int64_t e = t0 + k * m + j;
                    if (x.col[(size_t)e] - c0 != off + k * pp || !same_value(x, e, t0 + j, vs)) {
                        ok = false;
                        break;
                    }
                }
            }
        }
        if (ok) {
            nrep = N;
            p = pp;
        }
    }
    if (!nrep) return false;
    rep = HostReplica();
    rep.kind = 1;
    rep.c0 = c0;
    rep.p = p;
    rep.nrep = nrep;
    rep.diag.assign((size_t)(M * vs), 0.0);
    for (int64_t i = i0; i < nr; ++i)
        if (x.ptr[(size_t)i + 1] > x.ptr[(size_t)i])
            for (int q = 0; q < vs; ++q)
                rep.diag[(size_t)((x.rows[(size_t)i] - c0) * vs + q)] =
                    x.val[(size_t)(x.ptr[(size_t)i] * vs + q)];
    x = rebuild(x, i0, [&](int64_t i, auto &&emit) {
        const int64_t b0 = x.ptr[(size_t)i], t0 = tail[(size_t)i];
        for (int64_t e = b0; e < t0; ++e) emit(x.col[(size_t)e], e);
        const int64_t m = (x.ptr[(size_t)i + 1] - t0) / nrep;
        for (int64_t j = 0; j < m; ++j)
            emit((int32_t)(n + x.col[(size_t)(t0 + j)] - c0), t0 + j);
    });
    return true;
}

bool build_fused_tables(const HostExtra &xc, const HostReplica &rep, int64_t n, int64_t nflow,
                        HostFusedRep &out) {
    out = HostFusedRep();
    if (rep.kind != 1 || rep.c0 != nflow || nflow <= 0 || rep.p < 1) return false;
    const int vs = xc.is_complex ? 2 : 1;
    const int64_t nslice = (nflow + 63) / 64, nr = (int64_t)xc.rows.size();
    if (nslice >= (int64_t)INT32_MAX) return false;
    for (int64_t i = 1; i < nr; ++i)
        if (xc.rows[(size_t)i] <= xc.rows[(size_t)i - 1]) return false;
    HostFusedRep f;
    f.slice_tab.assign((size_t)nslice, -1);
    int64_t i = 0;
    while (i < nr) {
        const int64_t s = xc.rows[(size_t)i] / 64;
        if (xc.rows[(size_t)i] < 0 || xc.rows[(size_t)i] >= nflow) return false;
        f.slice_tab[(size_t)s] = (int32_t)(f.lane_ptr.size() / 65);
        const size_t base = f.lane_ptr.size();
        f.lane_ptr.resize(base + 65);
        for (int l = 0; l < 64; ++l) {
            f.lane_ptr[base + (size_t)l] = (int32_t)f.ecol.size();
            if (i < nr && xc.rows[(size_t)i] == 64 * s + l) {
                for (int64_t e = xc.ptr[(size_t)i]; e < xc.ptr[(size_t)i + 1]; ++e) {
                    const int64_t q = (int64_t)xc.col[(size_t)e] - n;
                    if (q < 0 || q >= rep.p) return false;   // not a column sum
                    f.ecol.push_back((int32_t)q);
                    for (int v = 0; v < vs; ++v) f.eval.push_back(xc.val[(size_t)(e * vs + v)]);
                }
                ++i;
                if ((int64_t)f.ecol.size() >= (int64_t)INT32_MAX) return false;
            }
        }
        f.lane_ptr[base + 64] = (int32_t)f.ecol.size();
        // rows ascending: anything left in this slice is out of order
        if (i < nr && xc.rows[(size_t)i] / 64 == s) return false;
    }
    out = std::move(f);
    return true;
}

bool compress_row_replicas(HostExtra &x, int64_t n, int64_t c0, int64_t p, int64_t nrep,
                           HostReplica &rep) {
    const int vs = x.is_complex ? 2 : 1;
    if (c0 < 0 || p < 1 || nrep < 2 || c0 + p * nrep != n) return false;
    const int64_t i0 = listed_index(x, c0);
    const int64_t nr = (int64_t)x.rows.size();
    if (i0 < 0 || nr - i0 != n - c0) return false;   // rows c0..n-1 listed consecutively
    // off-diagonal part of a row: entries before its diagonal (columns < c0
    // only), and the diagonal entry itself if any
    auto split = [&](int64_t i, int64_t &b0, int64_t &b1, int64_t &dg) {
        b0 = x.ptr[(size_t)i];
        const int64_t e1 = x.ptr[(size_t)i + 1], row = x.rows[(size_t)i];
        int64_t e = b0;
        while (e < e1 && x.col[(size_t)e] < c0) ++e;
        b1 = e;
        dg = -1;
        if (e < e1) {
            if (x.col[(size_t)e] != row || e + 1 != e1) return false;
            dg = e;
        }
        return true;
    };
    for (int64_t q = 0; q < p; ++q) {
        int64_t a0, a1, ad;
        if (!split(i0 + q, a0, a1, ad)) return false;
        for (int64_t k = 1; k < nrep; ++k) {
            int64_t b0, b1, bd;
            if (!split(i0 + k * p + q, b0, b1, bd)) return false;
            if (b1 - b0 != a1 - a0) return false;
            for (int64_t j = 0; j < a1 - a0; ++j)
                if (x.col[(size_t)(b0 + j)] != x.col[(size_t)(a0 + j)] ||
                    !same_value(x, b0 + j, a0 + j, vs))
                    return false;
        }
    }
    rep = HostReplica();
    rep.kind = 2;
    rep.c0 = c0;
    rep.p = p;
    rep.nrep = nrep;
    rep.diag.assign((size_t)((n - c0) * vs), 0.0);
    rep.vptr.assign(1, 0);
    for (int64_t i = i0; i < nr; ++i) {
        int64_t b0, b1, dg;
        split(i, b0, b1, dg);
        if (dg >= 0)
            for (int q = 0; q < vs; ++q)
                rep.diag[(size_t)((x.rows[(size_t)i] - c0) * vs + q)] = x.val[(size_t)(dg * vs + q)];
        if (i - i0 < p) {
            for (int64_t e = b0; e < b1; ++e) {
                rep.vcol.push_back(x.col[(size_t)e]);
                for (int q = 0; q < vs; ++q) rep.vval.push_back(x.val[(size_t)(e * vs + q)]);
            }
            rep.vptr.push_back((int64_t)rep.vcol.size());
        }
    }
    x = rebuild(x, i0, [&](int64_t i, auto &&emit) {
        for (int64_t e = x.ptr[(size_t)i]; e < x.ptr[(size_t)i + 1]; ++e) emit(x.col[(size_t)e], e);
    });
    return true;
}


void receptivity_from_parts(int64_t n_l, int64_t p, int64_t nwave, const double *w, double scale,
                            HostExtra &xf, HostExtra &xa, HostReplica &rf, HostReplica &ra) {
    if (!xf.is_complex || !xa.is_complex)
        throw std::invalid_argument("the receptivity operator is complex");
    if (n_l < 1 || p < 1 || nwave < 1) throw std::invalid_argument("bad receptivity sizes");
    const int64_t n = n_l + nwave * p;
    // forward: S's rows n_l .. n_l + p - 1 are the empty lower block; the
    // listed rows below n_l keep their entries, B's columns renamed to the
    // virtual columns n + q that read the column sums
    int64_t keep = 0;
    while (keep < (int64_t)xf.rows.size() && xf.rows[(size_t)keep] < n_l) ++keep;
    for (int64_t i = keep; i < (int64_t)xf.rows.size(); ++i)
        if (xf.ptr[(size_t)i + 1] > xf.ptr[(size_t)i])
            throw std::invalid_argument(
                "rows n_l.. of the per-wavenumber system must be empty (its freestream "
                "diagonal is given by w)");
    xf = rebuild(xf, keep, [&](int64_t i, auto &&emit) {
        for (int64_t e = xf.ptr[(size_t)i]; e < xf.ptr[(size_t)i + 1]; ++e) {
            const int64_t c = xf.col[(size_t)e];
            emit((int32_t)(c < n_l ? c : n + c - n_l), e);
        }
    });
    rf = HostReplica();
    rf.kind = 1;
    rf.c0 = n_l;
    rf.p = p;
    rf.nrep = nwave;
    rf.diag.assign((size_t)(2 * nwave * p), 0.0);
    for (int64_t k = 0; k < nwave; ++k)
        for (int64_t q = 0; q < p; ++q) rf.diag[(size_t)(2 * (k * p + q) + 1)] = -w[k] * scale;
    // adjoint: S''s rows n_l + q (every extra-DOF row is listed) are B''s
    // rows, kept once as the shared rows of the N copies
    const int64_t i0 = listed_index(xa, n_l);
    if (i0 < 0 || (int64_t)xa.rows.size() - i0 != p)
        throw std::invalid_argument("adjoint extras do not list the freestream rows");
    ra = HostReplica();
    ra.kind = 2;
    ra.c0 = n_l;
    ra.p = p;
    ra.nrep = nwave;
    ra.vptr.assign(1, 0);
    for (int64_t q = 0; q < p; ++q) {
        const int64_t i = i0 + q;
        if (xa.rows[(size_t)i] != n_l + q)
            throw std::invalid_argument("adjoint extras do not list the freestream rows");
        for (int64_t e = xa.ptr[(size_t)i]; e < xa.ptr[(size_t)i + 1]; ++e) {
            if (xa.col[(size_t)e] >= n_l)
                throw std::invalid_argument(
                    "columns n_l.. of the per-wavenumber system must be empty below row n_l");
            ra.vcol.push_back(xa.col[(size_t)e]);
            ra.vval.push_back(xa.val[(size_t)(2 * e)]);
            ra.vval.push_back(xa.val[(size_t)(2 * e + 1)]);
        }
        ra.vptr.push_back((int64_t)ra.vcol.size());
    }
    ra.diag.assign((size_t)(2 * nwave * p), 0.0);
    for (int64_t k = 0; k < nwave; ++k)
        for (int64_t q = 0; q < p; ++q) ra.diag[(size_t)(2 * (k * p + q) + 1)] = w[k] * scale;
    xa = rebuild(xa, i0, [&](int64_t i, auto &&emit) {
        for (int64_t e = xa.ptr[(size_t)i]; e < xa.ptr[(size_t)i + 1]; ++e)
            emit(xa.col[(size_t)e], e);
    });
}


std::vector<int64_t> xcd_balanced_order(const std::vector<int64_t> &len) {
    const int64_t m = (int64_t)len.size();
    std::vector<int64_t> order((size_t)m);
    std::iota(order.begin(), order.end(), 0);
    if (m < 16) return order;
    const auto mm = std::minmax_element(len.begin(), len.end());
    if (*mm.second <= *mm.first + *mm.first / 8) return order;   // contiguous split is balanced
    // xcd_tile's ranges: the first m % 8 hold m / 8 + 1 rows, the rest m / 8
    const int64_t per = m / 8, rem = m % 8;
    std::vector<int64_t> by_len = order;
    std::stable_sort(by_len.begin(), by_len.end(),
                     [&](int64_t a, int64_t b) { return len[(size_t)a] > len[(size_t)b]; });
    std::vector<std::vector<int64_t>> bin(8);
    int64_t load[8] = {};
    for (const int64_t r : by_len) {
        int best = -1;
        for (int k = 0; k < 8; ++k) {
            const int64_t cap = per + (k < rem ? 1 : 0);
            if ((int64_t)bin[(size_t)k].size() < cap && (best < 0 || load[k] < load[best]))
                best = k;
        }
        bin[(size_t)best].push_back(r);
        load[best] += len[(size_t)r];
    }
    int64_t o = 0;
    for (auto &bk : bin) {
        std::sort(bk.begin(), bk.end());
        for (const int64_t r : bk) order[(size_t)o++] = r;
    }
    return order;
}


bool build_shock_fuse(const HostExtra &x, int64_t nflow, HostExtra &rows_out,
                      HostShockPanels &pan) {
    rows_out = HostExtra();
    pan = HostShockPanels();
    if (x.is_complex || nflow <= 0) return false;
    const int64_t nr = (int64_t)x.rows.size();
    int64_t nflow_entries = 0;
    for (int64_t i = 0; i < nr; ++i) {
        const int64_t row = x.rows[(size_t)i];
        if (row >= nflow) continue;
        for (int64_t e = x.ptr[(size_t)i]; e < x.ptr[(size_t)i + 1]; ++e)
            if (x.col[(size_t)e] < nflow) return false;
        nflow_entries += x.ptr[(size_t)i + 1] - x.ptr[(size_t)i];
    }
    // the extra-DOF rows, in their order
    HostExtra r;
    r.is_complex = false;
    r.ptr.assign(1, 0);
    for (int64_t i = 0; i < nr; ++i) {
        const int64_t row = x.rows[(size_t)i];
        if (row < nflow) continue;
        for (int64_t e = x.ptr[(size_t)i]; e < x.ptr[(size_t)i + 1]; ++e) {
            r.col.push_back(x.col[(size_t)e]);
            r.val.push_back(x.val[(size_t)e]);
        }
        r.rows.push_back(row);
        r.ptr.push_back((int64_t)r.col.size());
    }
    if (nflow_entries == 0 && r.rows.empty()) return false;
    // panels: the listed flow rows grouped by slice of 64 rows (ascending)
    const int64_t nslice = (nflow + 63) / 64;
    HostShockPanels pp;
    pp.slice_tab.assign((size_t)nslice, -1);
    pp.cptr.assign(1, 0);
    int64_t i = 0;
    while (i < nr && x.rows[(size_t)i] < nflow) {
        const int64_t sl = x.rows[(size_t)i] / 64;
        int64_t j = i;
        std::vector<int32_t> cols;
        while (j < nr && x.rows[(size_t)j] < nflow && x.rows[(size_t)j] / 64 == sl) {
            for (int64_t e = x.ptr[(size_t)j]; e < x.ptr[(size_t)j + 1]; ++e)
                cols.push_back(x.col[(size_t)e]);
            ++j;
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        if (!cols.empty()) {
            const int64_t base = (int64_t)pp.pcol.size();
            pp.slice_tab[(size_t)sl] = (int32_t)(pp.cptr.size() - 1);
            pp.pcol.insert(pp.pcol.end(), cols.begin(), cols.end());
            pp.pval.resize(pp.pcol.size() * 64, 0.0);
            for (int64_t k = i; k < j; ++k) {
                const int lane = (int)(x.rows[(size_t)k] - 64 * sl);
                for (int64_t e = x.ptr[(size_t)k]; e < x.ptr[(size_t)k + 1]; ++e) {
                    const int64_t pos =
                        std::lower_bound(cols.begin(), cols.end(), x.col[(size_t)e]) - cols.begin();
                    pp.pval[(size_t)((base + pos) * 64 + lane)] += x.val[(size_t)e];
                }
            }
            pp.cptr.push_back((int64_t)pp.pcol.size());
            if ((int64_t)pp.pcol.size() >= (int64_t)INT32_MAX) return false;
        }
        i = j;
    }
    if ((int64_t)pp.pval.size() > 8 * nflow_entries) return false;
    rows_out = std::move(r);
    pan = std::move(pp);
    return true;
}

void build_shock_lead(int b, const std::vector<int64_t> &bptr, const std::vector<int32_t> &bcol,
                      const double *bval, const std::vector<int64_t> *bsrc, bool transpose,
                      double scale, const HostExtra &x, const HostExtra &rows_x,
                      const HostShockPanels &pan, int64_t nflow, HostShockLead &out) {
    out = HostShockLead();
    out.sh = rows_x;
    HostExtra &r64 = out.fr, &r32 = out.fr32;
    r64.ptr.assign(1, 0);
    r32.ptr.assign(1, 0);
    auto push = [&](int32_t c, double v64, double v32) {
        r64.col.push_back(c);
        r64.val.push_back(v64);
        r32.col.push_back(c);
        r32.val.push_back(v32);
    };
    const int bb = b * b;
    for (size_t sl = 0; sl < pan.slice_tab.size(); ++sl) {
        if (pan.slice_tab[sl] < 0) continue;
        for (int64_t row = (int64_t)sl * 64; row < (int64_t)sl * 64 + 64 && row < nflow; ++row) {
            const int64_t cell = row / b;
            const int r = (int)(row % b);
            for (int64_t k = bptr[(size_t)cell]; k < bptr[(size_t)cell + 1]; ++k) {
                const int64_t src = bsrc ? (*bsrc)[(size_t)k] : k;
                const int64_t c = bcol[(size_t)k];
                for (int j = 0; j < b; ++j) {
                    const double v =
                        bval[src * bb + (transpose ? j * b + r : r * b + j)] * scale;
                    push((int32_t)(c * b + j), v, (double)(float)v);
                }
            }
            // the row's extra entries (x lists rows ascending)
            const auto it = std::lower_bound(x.rows.begin(), x.rows.end(), row);
            if (it != x.rows.end() && *it == row) {
                const size_t i = (size_t)(it - x.rows.begin());
                for (int64_t e = x.ptr[i]; e < x.ptr[i + 1]; ++e)
                    push(x.col[(size_t)e], x.val[(size_t)e], x.val[(size_t)e]);
            }
            r64.rows.push_back(row);
            r32.rows.push_back(row);
            r64.ptr.push_back((int64_t)r64.col.size());
            r32.ptr.push_back((int64_t)r32.col.size());
        }
    }
}

}  // namespace hymor
