// libhymor_hip: C ABI, operator handles and launch logic.
//
// Reference path being replaced (all in utils/Stability_analysis/):
//   Eigenvalues/ARNOLDI_POWER_GPU.jl:36-82        POWER_METHOD! + adapter
//   Transient_growth_downstream/CONSTRUCT_C.jl:33-111  EXPONENTIAL! + zeroing
//   cuSPARSE SpMV on CuSparseMatrixCSC (mul!, ARNOLDI_POWER_GPU.jl:73)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <climits>
#include <memory>
#include <vector>

#include "hymor.h"
#include "hymor_internal.hpp"
#include "kernels.hpp"

using namespace hymor;

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error {
    explicit HipError(const std::string &m) : std::runtime_error(m) {}
};
struct ArgError : std::invalid_argument {
    explicit ArgError(const std::string &m) : std::invalid_argument(m) {}
};

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (line " + \
                           std::to_string(__LINE__) + ")");                            \
    } while (0)

#define ARG_CHECK(cond, msg)              \
    do {                                  \
        if (!(cond)) throw ArgError(msg); \
    } while (0)

// Launch variant of the single-vector Taylor-term kernel, chosen per
// operator by hymor_op_autotune (defaults: k_term_split, 4 waves per slice,
// non-temporal value loads -- the usual pick on large operators).
struct TermTuning {
    int split = 3;   // 0: k_term (one wave per slice), 1: k_term_split,
                     // 2: k_term_split_p (grouped loads), 3: k_term_lds
                     // (grouped loads, x staged through LDS; widths 2 and 4 only,
                     // the others run 2)
    int waves = 1;   // k_term: waves per workgroup; the others: waves per slice
                     // (k_term_split_p: 3 unless 9)
    int nt = 1;      // non-temporal loads of the block values
    int spmm_mfma = 6;  // batches of >= 16 doubles per row: 0 vector-FMA k_spmm,
                        // 1 k_spmm_mfma, 2/3 k_spmm_mfma_mc (2 or 1 cells per wave),
                        // 4/5 k_spmm_mfma_w (large / small load groups),
                        // 6 tile layout: b = 8 k_spmm_pair at 16+, k_spmm_tile at 8;
                        // b = 4 k_spmm_quad4c at 16+, k_term_lds at 8,
                        // 7 k_spmm_quad (4x4x4 matrix cores; b = 4: k_spmm_quad4),
                        // 8 k_spmm_tile (one cell per wave, tile layout)
};

struct DevSliced {
    TermTuning tune;
    int b = 0, cps = 0, prec = HYMOR_PREC_F64;
    int64_t ncell = 0, nslice = 0, nnzb = 0, nnzb_pad = 0;
    int64_t *slice_ptr = nullptr;
    int32_t *col = nullptr;
    void *val = nullptr;    // the values the kernels read: val64 or val32
    void *val64 = nullptr;  // FP64 lane-interleaved values (an FP64 upload)
    void *val32 = nullptr;  // FP32 values (an FP32 upload, or cast from val64)
    void *val_t = nullptr;  // the same values in k_spmm_tile's order (made on first use)
    void *val_q = nullptr;  // ... in k_spmm_quad's order (b = 8, made on first use)
    // b = 8: block columns of cell pairs for k_spmm_pair (bsr_pair_map)
    int64_t npair = 0;
    int64_t *pptr = nullptr;
    int32_t *pcol = nullptr, *pslot = nullptr;
};

struct DevExtra {
    bool is_complex = false;
    int64_t nrows = 0, nnz = 0;
    int64_t nlong = 0;   // listed rows [0, nlong) have more than kShortRow entries
    int64_t *rows = nullptr, *ptr = nullptr;
    int32_t *col = nullptr;
    double *val = nullptr;
    // replicated freestream blocks (HostReplica), kept once
    ReplicaDev rep;
    double *rep_sum = nullptr;     // kind 1: p * 64 doubles of column sums
    int64_t rep_nnz = 0;           // kind 2: entries of the shared rows
    // kind 1 forward with every extra entry on a flow row (HostFusedRep):
    // the tables of the fused Horner term and its two column-sum buffers
    // (p * 64 doubles each: up to 64 doubles per row)
    bool fused = false;
    int64_t fr_rot = 0;   // first slice with entries
    int32_t *fr_slice = nullptr, *fr_lane = nullptr, *fr_col = nullptr;
    double *fr_val = nullptr;
    double *fr_sum[2] = {nullptr, nullptr};
};

// The shock bands in the single-vector term kernel's launch
// (build_shock_fuse / ShockFuseDev): the extra-DOF rows and the slice panels.
struct DevShock {
    bool on = false;
    DevExtra rows;
    int64_t npan = 0;
    int32_t *pslices = nullptr;   // the slices with a panel, ascending
    int32_t *slice_tab = nullptr, *pcol = nullptr;
    int64_t *cptr = nullptr;
    double *pval = nullptr;
    int64_t npcol = 0;
    // batched terms: the rows the batch kernel's leading workgroups finish
    // (build_shock_lead): extra-DOF rows, whole flow rows in FP64 and with
    // FP32-rounded block entries
    DevExtra lead_sh, lead_fr, lead_fr32;
};

// Where the block values of the rows in build_shock_lead come from.
struct BlockSrc {
    const std::vector<int64_t> *ptr = nullptr;
    const std::vector<int32_t> *col = nullptr;
    const double *val = nullptr;
    const std::vector<int64_t> *src = nullptr;
    bool transpose = false;
    double scale = 1.0;
};


}  // namespace

struct hymor_op {
    int64_t n = 0, nchi = 0, neta = 0, ncell = 0, nflow = 0;
    int b = 0, prec = HYMOR_PREC_F64;
    bool is_complex = false, has_adj = false;
    DevSliced fwd, adj;
    DevExtra xf, xa;
    // the same extra rows with the receptivity operator's replicated
    // freestream blocks kept once (rep.kind != 0 when found), used at
    // kRepMinWidth or more doubles per row
    DevExtra xf_rep, xa_rep;
    // built from one per-wavenumber system (hymor_op_create_receptivity):
    // only the compressed extras exist, used at every width
    bool rep_only = false;
    // the extra rows of a real operator whose flow rows couple only to
    // extra DOFs, in the single-vector term kernel's launch (forward,
    // adjoint); DevShock::on when found
    DevShock shf, sha;
    hipStream_t stream = nullptr;
    int device = 0;
    // workspace: [0] stage (var-major), [1] in0, [2] va, [3] vb, [4] t0, [5] t1, [6] spare
    int ws_nv = 0;
    double *ws[7] = {};
    int64_t dev_bytes = 0;
    std::vector<void *> allocs;
    // per-column complex diagonal on extra rows [cd0, cd1) for the current
    // batched call (factored receptivity); null otherwise
    double *cdiag = nullptr, *cdiag_adj = nullptr;
    int64_t cd0 = 0, cd1 = 0;
    bool cdiag_on = false;
    // captured Taylor sweep (internal buffers only), replayed while the
    // parameters repeat; cleared when buffers or kernel variants change
    bool use_graphs = false;
    struct {
        bool valid = false;
        bool adj = false;
        int order = 0, form = 0, nv = 0;
        int64_t nsub = 0;
        double dt = 0.0;
        hipGraphExec_t exec = nullptr;
        const double *result = nullptr;
    } graph;
};

// Block-diagonal weights (the Chu matrices R'MR): one b x b block per cell
// and weight, a diagonal on the extra DOFs.
struct hymor_weights {
    int b = 0, nmat = 0;
    int64_t nchi = 0, neta = 0, ncell = 0, n = 0, n_extra = 0;
    double *blocks = nullptr;  // nmat * ncell * b * b
    double *xdiag = nullptr;   // nmat * n_extra
    double *partial = nullptr; // nmat * nblocks of the quadratic-form kernel
    int64_t nblocks = 0;
    std::vector<void *> allocs;
};

// Device-resident Krylov basis: V is n x (m_max+1), column-major, in the
// ABI's var-major order, so its columns feed the *_dev actions directly.
struct hymor_krylov {
    int64_t n = 0;
    int m_max = 0, ncomp = 1;
    double *V = nullptr;        // (m_max+1) columns
    double *tmp = nullptr;      // (m_max+1) columns, for the restart combination
    double *h = nullptr;        // 2*(m_max+1) coefficients
    double *h2 = nullptr;
    double *q = nullptr;        // (m_max+1)^2 complex, restart matrix
    double *partial = nullptr;  // 2*(m_max+1) * kDotBlocks
    // block Gram-Schmidt (made on first use): partial products and the
    // two projection passes' coefficients
    double *bpartial = nullptr, *hb = nullptr, *hb2 = nullptr;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
};

namespace {

constexpr int kDotBlocks = 512;   // workgroups of the tall-skinny dot kernel
constexpr int kDotTile = 16;      // basis columns per dot launch
constexpr int kMaxWeights = 8;   // matches k_quad_blockdiag's NMAX

void *dalloc_raw(std::vector<void *> &allocs, size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    allocs.push_back(p);
    return p;
}

void *dalloc(hymor_op *op, size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    op->allocs.push_back(p);
    op->dev_bytes += (int64_t)bytes;
    return p;
}

template <typename T>
T *upload(hymor_op *op, const std::vector<T> &h) {
    T *d = (T *)dalloc(op, h.size() * sizeof(T));
    if (!h.empty()) HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Device memory for one call (upload staging), freed on scope exit.
struct DevTemp {
    void *p = nullptr;
    explicit DevTemp(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 256) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw std::bad_alloc();
        }
    }
    ~DevTemp() {
        if (p) (void)hipFree(p);
    }
    DevTemp(const DevTemp &) = delete;
    DevTemp &operator=(const DevTemp &) = delete;
};

// d from a BSR structure (ptr, col) whose block k takes its values from
// block src[k] (k when src is null) of raw_dev, b x b row-major doubles on
// the device, read transposed when `transpose`, times scale.  The host
// builds the slices and the per-slot source indices; k_slice_vals places,
// scales and casts the values on the device.
void build_sliced(hymor_op *op, int b, int64_t ncell, const std::vector<int64_t> &ptr,
                  const std::vector<int32_t> &col, const std::vector<int64_t> *src,
                  bool transpose, const double *raw_dev, double scale, int prec, DevSliced &d) {
    HostSliced s;
    std::vector<int64_t> pad_src;
    bsr_to_sliced_map(b, ncell, ptr, col, src, s, pad_src);
    d.b = s.b;
    d.cps = s.cps;
    d.prec = prec;
    d.ncell = s.ncell;
    d.nslice = s.nslice;
    d.nnzb = s.nnzb;
    d.nnzb_pad = s.slice_ptr[s.nslice];
    d.slice_ptr = upload(op, s.slice_ptr);
    d.col = upload(op, s.col);
    if (b == 8) {
        std::vector<int64_t> pptr;
        std::vector<int32_t> pcol, pslot;
        bsr_pair_map(ncell, ptr, col, pptr, pcol, pslot);
        d.npair = (int64_t)pptr.size() - 1;
        d.pptr = upload(op, pptr);
        d.pcol = upload(op, pcol);
        d.pslot = upload(op, pslot);
    }
    const int64_t total = d.nnzb_pad * b * b;
    DevTemp dsrc(pad_src.size() * sizeof(int64_t));
    if (!pad_src.empty())
        HIP_TRY(hipMemcpy(dsrc.p, pad_src.data(), pad_src.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
    if (prec == HYMOR_PREC_F32) {
        d.val = d.val32 = dalloc(op, (size_t)total * sizeof(float));
        hipLaunchKernelGGL(k_slice_vals<float>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           raw_dev, (const int64_t *)dsrc.p, (float *)d.val32, total, b,
                           transpose ? 1 : 0, scale);
    } else {
        d.val = d.val64 = dalloc(op, (size_t)total * sizeof(double));
        hipLaunchKernelGGL(k_slice_vals<double>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           raw_dev, (const int64_t *)dsrc.p, (double *)d.val64, total, b,
                           transpose ? 1 : 0, scale);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(op->stream));   // before dsrc (and raw) are freed
}

// d from a host BSR whose values are final (scaled)
void upload_sliced(hymor_op *op, const HostBSR &a, int prec, DevSliced &d) {
    DevTemp raw(a.val.size() * sizeof(double));
    if (!a.val.empty())
        HIP_TRY(hipMemcpy(raw.p, a.val.data(), a.val.size() * sizeof(double),
                          hipMemcpyHostToDevice));
    build_sliced(op, a.b, a.ncell, a.ptr, a.col, nullptr, false, (const double *)raw.p, 1.0, prec,
                 d);
}

// frees one tracked allocation
void dfree(hymor_op *op, void *p, size_t bytes) {
    if (!p) return;
    for (size_t k = 0; k < op->allocs.size(); ++k)
        if (op->allocs[k] == p) {
            (void)hipFree(p);
            op->allocs.erase(op->allocs.begin() + (long)k);
            op->dev_bytes -= (int64_t)bytes;
            return;
        }
}

// Rows of the extra CSR with more than this many entries get a workgroup
// each (lanes per entry, cross-lane sums); shorter ones lanes per row, no
// reduction (both in k_extra).
constexpr int64_t kShortRow = 32;
// waves per workgroup of k_extra (one workgroup per long row)
constexpr int kLongWaves = 4;

// The listed rows reordered long rows first (each group keeps its order),
// so the two kernels take contiguous ranges.
void upload_extra(hymor_op *op, const HostExtra &h, DevExtra &d) {
    const int64_t nr = (int64_t)h.rows.size();
    const int vs = h.is_complex ? 2 : 1;
    std::vector<int64_t> order;
    order.reserve((size_t)nr);
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t r = 0; r < nr; ++r)
            if ((h.ptr[r + 1] - h.ptr[r] > kShortRow) == (pass == 0)) order.push_back(r);
    {
        // the long rows spread over the XCDs by entry count
        std::vector<int64_t> lens;
        for (const int64_t r : order)
            if (h.ptr[r + 1] - h.ptr[r] > kShortRow) lens.push_back(h.ptr[r + 1] - h.ptr[r]);
        const std::vector<int64_t> perm = xcd_balanced_order(lens);
        std::vector<int64_t> head(order.begin(), order.begin() + (long)lens.size());
        for (size_t i = 0; i < perm.size(); ++i) order[i] = head[(size_t)perm[i]];
    }
    std::vector<int64_t> rows((size_t)nr), ptr((size_t)nr + 1, 0);
    std::vector<int32_t> col(h.col.size());
    std::vector<double> val(h.val.size());
    int64_t nlong = 0;
    for (int64_t i = 0; i < nr; ++i) {
        const int64_t r = order[(size_t)i], b0 = h.ptr[r], len = h.ptr[r + 1] - b0;
        if (len > kShortRow) ++nlong;
        rows[(size_t)i] = h.rows[(size_t)r];
        ptr[(size_t)i + 1] = ptr[(size_t)i] + len;
        std::copy(h.col.begin() + b0, h.col.begin() + b0 + len, col.begin() + ptr[(size_t)i]);
        std::copy(h.val.begin() + b0 * vs, h.val.begin() + (b0 + len) * vs,
                  val.begin() + ptr[(size_t)i] * vs);
    }
    d.is_complex = h.is_complex;
    d.nrows = nr;
    d.nlong = nlong;
    d.nnz = (int64_t)h.col.size();
    d.rows = upload(op, rows);
    d.ptr = upload(op, ptr);
    d.col = upload(op, col);
    d.val = upload(op, val);
}

// The replicated blocks found by compress_column_replicas /
// compress_row_replicas for the extra CSR d (uploaded by upload_extra).
void upload_replica(hymor_op *op, const HostReplica &h, DevExtra &d) {
    d.rep = ReplicaDev();
    if (h.kind == 0) return;
    d.rep.kind = h.kind;
    d.rep.c0 = h.c0;
    d.rep.p = h.p;
    d.rep.nrep = h.nrep;
    d.rep.diag = upload(op, h.diag);
    if (h.kind == 1) {
        d.rep.vbase = op->n;
        d.rep_sum = (double *)dalloc(op, (size_t)h.p * 64 * sizeof(double));
        d.rep.sum = d.rep_sum;
    } else {
        d.rep.vptr = upload(op, h.vptr);
        d.rep.vcol = upload(op, h.vcol);
        d.rep.vval = upload(op, h.vval);
        d.rep_nnz = (int64_t)h.vcol.size();
    }
}

// The compressed extra rows and their replicas (nothing when none found).
void upload_replicated(hymor_op *op, const HostExtra &h, const HostReplica &rep, DevExtra &d) {
    if (rep.kind == 0) return;
    upload_extra(op, h, d);
    upload_replica(op, rep, d);
}

// The fused Horner term's tables (build_fused_tables) for the compressed
// forward extras d, when the operator has them.
void upload_fused(hymor_op *op, const HostExtra &h, const HostReplica &rep, DevExtra &d) {
    HostFusedRep f;
    if (rep.kind != 1 || !build_fused_tables(h, rep, op->n, op->nflow, f)) return;
    d.fr_rot = 0;
    for (size_t s = 0; s < f.slice_tab.size(); ++s)
        if (f.slice_tab[s] >= 0) {
            d.fr_rot = (int64_t)s;
            break;
        }
    d.fr_slice = upload(op, f.slice_tab);
    d.fr_lane = upload(op, f.lane_ptr);
    d.fr_col = upload(op, f.ecol);
    d.fr_val = upload(op, f.eval);
    for (double *&s : d.fr_sum) s = (double *)dalloc(op, (size_t)rep.p * 64 * sizeof(double));
    d.fused = true;
}

// The shock bands' fused form (build_shock_fuse) of the extra CSR h, when
// the operator has it.
void upload_shock(hymor_op *op, const HostExtra &h, const BlockSrc &bs, DevShock &d) {
    HostExtra rows;
    HostShockPanels pan;
    d = DevShock();
    if (!build_shock_fuse(h, op->nflow, rows, pan)) return;
    HostShockLead lead;
    build_shock_lead(op->b, *bs.ptr, *bs.col, bs.val, bs.src, bs.transpose, bs.scale, h, rows,
                     pan, op->nflow, lead);
    upload_extra(op, lead.sh, d.lead_sh);
    upload_extra(op, lead.fr, d.lead_fr);
    upload_extra(op, lead.fr32, d.lead_fr32);
    upload_extra(op, rows, d.rows);
    d.slice_tab = upload(op, pan.slice_tab);
    d.cptr = upload(op, pan.cptr);
    d.pcol = upload(op, pan.pcol);
    d.pval = upload(op, pan.pval);
    d.npcol = (int64_t)pan.pcol.size();
    std::vector<int32_t> ps;
    for (size_t k = 0; k < pan.slice_tab.size(); ++k)
        if (pan.slice_tab[k] >= 0) ps.push_back((int32_t)k);
    d.npan = (int64_t)ps.size();
    d.pslices = upload(op, ps);
    d.on = true;
}

// Compresses the replicated freestream blocks of a receptivity operator in
// both directions' extra CSRs (xa may be null: no adjoint).
void find_replicas(int64_t n, HostExtra &xf, HostExtra *xa, HostReplica &rf, HostReplica &ra) {
    rf = HostReplica();
    ra = HostReplica();
    if (!compress_column_replicas(xf, n, rf)) return;
    if (xa && !compress_row_replicas(*xa, n, rf.c0, rf.p, rf.nrep, ra)) ra = HostReplica();
}

void drop_graph(hymor_op *op) {
    if (op->graph.exec) (void)hipGraphExecDestroy(op->graph.exec);
    op->graph.exec = nullptr;
    op->graph.valid = false;
}

void release(hymor_op *op) {
    if (!op) return;
    drop_graph(op);
    for (void *p : op->allocs) (void)hipFree(p);
    op->allocs.clear();
    if (op->stream) (void)hipStreamDestroy(op->stream);
    op->stream = nullptr;
}

void ensure_tile(hymor_op *op);
int tile_b_of(const hymor_op *op, int nv);

void ensure_ws_only(hymor_op *op, int nv);

// workspace for nv doubles per row, plus the tile-ordered values when the
// width runs k_spmm_tile
void ensure_ws(hymor_op *op, int nv) {
    ensure_ws_only(op, nv);
    if (tile_b_of(op, nv)) ensure_tile(op);
}

void ensure_ws_only(hymor_op *op, int nv) {
    if (op->ws_nv >= nv) return;
    drop_graph(op);
    // free the old set (tracked in allocs): rebuild the list without them
    for (int i = 0; i < 7; ++i) {
        if (!op->ws[i]) continue;
        for (size_t k = 0; k < op->allocs.size(); ++k)
            if (op->allocs[k] == op->ws[i]) {
                (void)hipFree(op->ws[i]);
                op->allocs.erase(op->allocs.begin() + (long)k);
                break;
            }
        op->ws[i] = nullptr;
    }
    const size_t vb = (size_t)op->n * nv * sizeof(double);
    for (int i = 0; i < 7; ++i) {
        op->ws[i] = (double *)dalloc(op, vb);
        HIP_TRY(hipMemsetAsync(op->ws[i], 0, vb, op->stream));
    }
    HIP_TRY(hipStreamSynchronize(op->stream));
    op->ws_nv = nv;
}

// Batched widths of 8+ doubles per row run the tile-layout kernels under
// batch kernels 6 and 7; their internal buffers then use the tile layout for
// flow rows (row-major itself at b = 4).  Returns the block size to pass as
// tile_b, or 0 (row-major).
int tile_b_of(const hymor_op *op, int nv) {
    return (nv >= 8 && op->fwd.tune.spmm_mfma >= 6) ? op->b : 0;
}

void quad_vals_of(hymor_op *op, DevSliced &s) {
    if (s.val_q || s.nnzb_pad == 0 || s.b != 8) return;
    const int64_t total = s.nnzb_pad * 64;
    const size_t esz = s.prec == HYMOR_PREC_F32 ? sizeof(float) : sizeof(double);
    s.val_q = dalloc(op, (size_t)total * esz);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
    if (s.prec == HYMOR_PREC_F32)
        hipLaunchKernelGGL(k_quad_vals<float>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           (const float *)s.val, (float *)s.val_q, total);
    else
        hipLaunchKernelGGL(k_quad_vals<double>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           (const double *)s.val, (double *)s.val_q, total);
    HIP_TRY(hipGetLastError());
}

void tile_vals_of(hymor_op *op, DevSliced &s) {
    if (s.val_t || s.nnzb_pad == 0) return;
    const int64_t total = s.nnzb_pad * s.b * s.b;
    const size_t esz = s.prec == HYMOR_PREC_F32 ? sizeof(float) : sizeof(double);
    s.val_t = dalloc(op, (size_t)total * esz);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
    if (s.prec == HYMOR_PREC_F32)
        hipLaunchKernelGGL(k_tile_vals<float>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           (const float *)s.val, (float *)s.val_t, total, s.b);
    else
        hipLaunchKernelGGL(k_tile_vals<double>, dim3((unsigned)grid), dim3(256), 0, op->stream,
                           (const double *)s.val, (double *)s.val_t, total, s.b);
    HIP_TRY(hipGetLastError());
}

// The value copies the chosen tile-layout kernel reads (k_spmm_tile: tile
// order; k_spmm_quad at b = 8: quad order), for both directions; called
// before any launch (or graph capture) of those kernels.
void ensure_tile(hymor_op *op) {
    if (op->ncell == 0) return;
    // at b = 4 every batch kernel of 6 and 7 (k_term_lds, k_spmm_quad4c,
    // k_spmm_quad4) reads the lane-interleaved values
    if (op->b == 4) return;
    const bool quad = op->fwd.tune.spmm_mfma == 7 && op->b == 8;
    auto missing = [&](const DevSliced &s) { return quad ? !s.val_q : !s.val_t; };
    if (!missing(op->fwd) && !(op->has_adj && missing(op->adj))) return;
    drop_graph(op);
    for (DevSliced *s : {&op->fwd, &op->adj}) {
        if (s == &op->adj && !op->has_adj) continue;
        if (quad) quad_vals_of(op, *s);
        else tile_vals_of(op, *s);
    }
    HIP_TRY(hipStreamSynchronize(op->stream));
}

// ---------------------------------------------------------------- launches

template <int B, typename VT, int NV, int NW, bool NT>
void launch_split(const DevSliced &s, hipStream_t st, const double *x,
                  double *tout, const double *vin, double *vout, double c1, double c2) {
    hipLaunchKernelGGL((k_term_split<B, VT, NV, NW, NT>), dim3((unsigned)s.nslice), dim3(64 * NW),
                       0, st, s.slice_ptr, s.col, (const VT *)s.val, s.nslice, s.ncell, x,
                       tout, vin, vout, c1, c2);
}

// split == 2: k_term_split_p with NW = 3 or 9 waves per slice and slot
// groups of S = 9 / NW, so one group per wave covers a 3x3-stencil row
template <int B, typename VT, int NV, int NW, bool NT>
void launch_split_p(const DevSliced &s, hipStream_t st, const double *x,
                    double *tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int S = (9 + NW - 1) / NW;
    hipLaunchKernelGGL((k_term_split_p<B, VT, NV, NW, S, NT>), dim3((unsigned)s.nslice),
                       dim3(64 * NW), 0, st, s.slice_ptr, s.col, (const VT *)s.val, s.nslice,
                       s.ncell, x, tout, vin, vout, c1, c2, ShockFuseDev{});
}

template <typename T>
struct TypeTag {
    using type = T;
};

// One real vector's term with the shock bands in the same launch
// (k_term_split_p with SH): the default single-vector tuning's kernel, three
// waves per slice, plus one leading workgroup per extra-DOF row.
template <int B, typename VT, bool NT>
void launch_shock_t(const DevSliced &s, const DevShock &sh, int64_t nflow, hipStream_t st,
                    const double *x, double *tout, const double *vin, double *vout, double c1,
                    double c2) {
    ShockFuseDev F;
    F.nrows = sh.rows.nrows;
    F.nsh_wg = (F.nrows + 7) / 8 * 8;
    F.nflow = nflow;
    F.rows = sh.rows.rows;
    F.ptr = sh.rows.ptr;
    F.col = sh.rows.col;
    F.val = sh.rows.val;
    F.slice_tab = sh.slice_tab;
    F.cptr = sh.cptr;
    F.pcol = sh.pcol;
    F.pval = sh.pval;
    F.npan = sh.npan;
    F.npan_wg = (sh.npan + 7) / 8 * 8;
    F.pslices = sh.pslices;
    hipLaunchKernelGGL((k_term_split_p<B, VT, 1, 3, 3, NT, true>),
                       dim3((unsigned)(F.nsh_wg + F.npan_wg + s.nslice)), dim3(64 * 3), 0, st,
                       s.slice_ptr,
                       s.col, (const VT *)s.val, s.nslice, s.ncell, x, tout, vin, vout, c1, c2,
                       F);
}

ShockBatchDev shock_batch_dev(const DevShock &sh, int64_t nflow, bool f32) {
    const DevExtra &F = f32 ? sh.lead_fr32 : sh.lead_fr;
    ShockBatchDev H;
    H.nsh = sh.lead_sh.nrows;
    H.nfr = F.nrows;
    // a multiple of 8: the slices keep their XCDs
    H.nlead_wg = (H.nsh + H.nfr + 7) / 8 * 8;
    H.nflow = nflow;
    H.sh_rows = sh.lead_sh.rows;
    H.sh_ptr = sh.lead_sh.ptr;
    H.sh_col = sh.lead_sh.col;
    H.sh_val = sh.lead_sh.val;
    H.fr_rows = F.rows;
    H.fr_ptr = F.ptr;
    H.fr_col = F.col;
    H.fr_val = F.val;
    H.skip = sh.slice_tab;
    return H;
}

// Whether a batched term of nv doubles per row runs in one launch over a
// shock operator (the extra-DOF rows and the shock-column slices' whole rows
// in the batch kernel's leading workgroups, those slices skipped): k_term_lds
// with one wave per slice at 2 and 4, the default b = 4 batch kernel
// (k_spmm_quad4c) at 8..64, whose row-major buffers the leading rows write.
bool shock_batched_ok(const hymor_op *op, const DevSliced &s, int nv) {
    if (nv == 2 || nv == 4) return s.tune.split == 3 && s.tune.waves == 1;
    if (nv >= 8 && nv <= 64)
        return op->b == 4 && (s.tune.spmm_mfma == 6 || s.tune.spmm_mfma == 8);
    return false;
}

template <int B, typename VT, int NV, bool NT>
void launch_shock_lds_t(const DevSliced &s, const ShockBatchDev &H, hipStream_t st,
                        const double *x, double *tout, const double *vin, double *vout,
                        double c1, double c2) {
    hipLaunchKernelGGL((k_term_lds<B, VT, NV, 1, 9, NT, 0, true>),
                       dim3((unsigned)(H.nlead_wg + s.nslice)), dim3(64), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.nslice, s.ncell, x, tout, vin,
                       vout, c1, c2, FusedRepDev{}, H);
}

template <typename VT, int R, int W>
void launch_shock_quad4c_t(const DevSliced &s, const ShockBatchDev &H, hipStream_t st,
                           const double *x, double *tout, const double *vin, double *vout,
                           double c1, double c2) {
    const int64_t grid = ((s.ncell + 3) / 4 + 3) / 4;
    hipLaunchKernelGGL((k_spmm_quad4c<VT, R, W, 0, true>),
                       dim3((unsigned)(H.nlead_wg + grid)), dim3(256), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x, tout, vin, vout, c1,
                       c2, FusedRepDev{}, H);
}

template <int B, typename VT, bool NT>
void launch_shock_batched_t(const DevSliced &s, const DevShock &sh, int64_t nflow, int nv,
                            hipStream_t st, const double *x, double *tout, const double *vin,
                            double *vout, double c1, double c2) {
    const ShockBatchDev H = shock_batch_dev(sh, nflow, s.prec == HYMOR_PREC_F32);
    if (nv == 2) return launch_shock_lds_t<B, VT, 2, NT>(s, H, st, x, tout, vin, vout, c1, c2);
    if (nv == 4) return launch_shock_lds_t<B, VT, 4, NT>(s, H, st, x, tout, vin, vout, c1, c2);
    if constexpr (B == 4) {
        // the slot groups launch_block_part uses at each width
        switch (nv) {
        case 8: launch_shock_quad4c_t<VT, 8, 9>(s, H, st, x, tout, vin, vout, c1, c2); return;
        case 16: launch_shock_quad4c_t<VT, 16, 9>(s, H, st, x, tout, vin, vout, c1, c2); return;
        case 32: launch_shock_quad4c_t<VT, 32, 3>(s, H, st, x, tout, vin, vout, c1, c2); return;
        case 64: launch_shock_quad4c_t<VT, 64, 3>(s, H, st, x, tout, vin, vout, c1, c2); return;
        default: break;
        }
    }
    throw ArgError("shock rows: unsupported vector width (internal)");
}

void launch_shock_batched(const DevSliced &s, const DevShock &sh, int64_t nflow, int nv,
                          hipStream_t st, const double *x, double *tout, const double *vin,
                          double *vout, double c1, double c2) {
    const bool f32 = s.prec == HYMOR_PREC_F32, nt = s.tune.nt != 0;
    auto go = [&](auto b, auto vt, auto ntc) {
        constexpr int B = decltype(b)::value;
        using VT = typename decltype(vt)::type;
        constexpr bool NT = decltype(ntc)::value;
        launch_shock_batched_t<B, VT, NT>(s, sh, nflow, nv, st, x, tout, vin, vout, c1, c2);
    };
    auto by_nt = [&](auto b, auto vt) {
        if (nt) go(b, vt, std::true_type{});
        else go(b, vt, std::false_type{});
    };
    auto by_prec = [&](auto b) {
        if (f32) by_nt(b, TypeTag<float>{});
        else by_nt(b, TypeTag<double>{});
    };
    if (s.b == 4) by_prec(std::integral_constant<int, 4>{});
    else if (s.b == 8) by_prec(std::integral_constant<int, 8>{});
    else throw ArgError("vars_per_cell must be 4 or 8");
}

void launch_shock(const DevSliced &s, const DevShock &sh, int64_t nflow, hipStream_t st,
                  const double *x, double *tout, const double *vin, double *vout, double c1,
                  double c2) {
    const bool f32 = s.prec == HYMOR_PREC_F32, nt = s.tune.nt != 0;
    auto go = [&](auto b, auto vt, auto ntc) {
        constexpr int B = decltype(b)::value;
        using VT = typename decltype(vt)::type;
        constexpr bool NT = decltype(ntc)::value;
        launch_shock_t<B, VT, NT>(s, sh, nflow, st, x, tout, vin, vout, c1, c2);
    };
    auto by_nt = [&](auto b, auto vt) {
        if (nt) go(b, vt, std::true_type{});
        else go(b, vt, std::false_type{});
    };
    auto by_prec = [&](auto b) {
        if (f32) by_nt(b, TypeTag<float>{});
        else by_nt(b, TypeTag<double>{});
    };
    if (s.b == 4) by_prec(std::integral_constant<int, 4>{});
    else if (s.b == 8) by_prec(std::integral_constant<int, 8>{});
    else throw ArgError("vars_per_cell must be 4 or 8");
}

// split == 3: k_term_lds, as split 2 with the x blocks staged through LDS
template <int B, typename VT, int NV, int NW, bool NT>
void launch_lds(const DevSliced &s, hipStream_t st, const double *x,
                double *tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int S = (9 + NW - 1) / NW;
    hipLaunchKernelGGL((k_term_lds<B, VT, NV, NW, S, NT>), dim3((unsigned)s.nslice),
                       dim3(64 * NW), 0, st, s.slice_ptr, s.col, (const VT *)s.val, s.nslice,
                       s.ncell, x, tout, vin, vout, c1, c2, FusedRepDev{}, ShockBatchDev{});
}

template <int B, typename VT, int NV, bool NT>
void launch_term_nt(const DevSliced &s, hipStream_t st, const double *x,
                    double *tout, const double *vin, double *vout, double c1, double c2) {
    const TermTuning &tu = s.tune;
    if constexpr (NV >= 2 && (NV <= 4 || (B == 4 && NV == 8))) {
        if (tu.split == 3) {
            switch (tu.waves) {
            case 1: launch_lds<B, VT, NV, 1, NT>(s, st, x, tout, vin, vout, c1, c2); break;
            case 2: launch_lds<B, VT, NV, 2, NT>(s, st, x, tout, vin, vout, c1, c2); break;
            case 9: launch_lds<B, VT, NV, 9, NT>(s, st, x, tout, vin, vout, c1, c2); break;
            default: launch_lds<B, VT, NV, 3, NT>(s, st, x, tout, vin, vout, c1, c2); break;
            }
            return;
        }
    }
    if (tu.split >= 2) {
        if (tu.waves == 9) launch_split_p<B, VT, NV, 9, NT>(s, st, x, tout, vin, vout, c1, c2);
        else launch_split_p<B, VT, NV, 3, NT>(s, st, x, tout, vin, vout, c1, c2);
        return;
    }
    if (tu.split) {
        if (tu.waves == 2) launch_split<B, VT, NV, 2, NT>(s, st, x, tout, vin, vout, c1, c2);
        else launch_split<B, VT, NV, 4, NT>(s, st, x, tout, vin, vout, c1, c2);
        return;
    }
    // small operators: one wave per group so the slices spread over all CUs
    const int wpb = s.nslice >= 8192 ? tu.waves : 1;
    const int64_t grid = (s.nslice + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_term<B, VT, NV, NT>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.nslice, s.ncell, x, tout, vin,
                       vout, c1, c2);
}

template <int B, typename VT, int NV>
void launch_term_t(const DevSliced &s, hipStream_t st, const double *x,
                   double *tout, const double *vin, double *vout, double c1, double c2) {
    if (s.tune.nt) launch_term_nt<B, VT, NV, true>(s, st, x, tout, vin, vout, c1, c2);
    else launch_term_nt<B, VT, NV, false>(s, st, x, tout, vin, vout, c1, c2);
}

template <int B, typename VT, int R>
void launch_spmm_t(const DevSliced &s, hipStream_t st, const double *x,
                   double *tout, const double *vin, double *vout, double c1, double c2) {
    const int wpb = 4;
    const int64_t grid = (s.ncell + wpb - 1) / wpb;
    if constexpr (R % 16 == 0) {
        if (s.tune.spmm_mfma == 1) {
            hipLaunchKernelGGL((k_spmm_mfma<B, VT, R>), dim3((unsigned)grid), dim3(64 * wpb), 0,
                               st, s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x, tout,
                               vin, vout, c1, c2);
            return;
        }
        if (s.tune.spmm_mfma >= 4) {
            // 4: all loads of the row's slots in one group (W = 9 covers the
            // 3x3 stencil) at width 16, fewer at 32/64 to bound registers;
            // 5: smaller groups
            constexpr int W4 = R == 16 ? 9 : (R == 32 ? 5 : 3);
            constexpr int W5 = R == 16 ? 5 : (R == 32 ? 3 : 2);
            if (s.tune.spmm_mfma == 4)
                hipLaunchKernelGGL((k_spmm_mfma_w<B, VT, R, W4>), dim3((unsigned)grid),
                                   dim3(64 * wpb), 0, st, s.slice_ptr, s.col, (const VT *)s.val,
                                   s.ncell, x, tout, vin, vout, c1, c2);
            else
                hipLaunchKernelGGL((k_spmm_mfma_w<B, VT, R, W5>), dim3((unsigned)grid),
                                   dim3(64 * wpb), 0, st, s.slice_ptr, s.col, (const VT *)s.val,
                                   s.ncell, x, tout, vin, vout, c1, c2);
            return;
        }
        if (s.tune.spmm_mfma >= 2) {
            // 2: two cells per wave at width 16 (one above, to stay within the
            // register file), 3: one cell per wave; slot groups of U
            constexpr int U = R == 16 ? 3 : (R == 32 ? 2 : 1);
            if (s.tune.spmm_mfma == 2 && R == 16) {
                const int64_t waves = (s.ncell + 1) / 2;
                hipLaunchKernelGGL((k_spmm_mfma_mc<B, VT, R, U, 2>),
                                   dim3((unsigned)((waves + wpb - 1) / wpb)), dim3(64 * wpb), 0,
                                   st, s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x,
                                   tout, vin, vout, c1, c2);
            } else {
                hipLaunchKernelGGL((k_spmm_mfma_mc<B, VT, R, U, 1>), dim3((unsigned)grid),
                                   dim3(64 * wpb), 0, st, s.slice_ptr, s.col, (const VT *)s.val,
                                   s.ncell, x, tout, vin, vout, c1, c2);
            }
            return;
        }
    }
    hipLaunchKernelGGL((k_spmm<B, VT, R>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x, tout, vin, vout, c1,
                       c2);
}

// Kernel choice per (block size, row width): lane-per-row k_term up to 4
// doubles per row (8 for b = 4), the rhs-parallel k_spmm above.
template <int B, typename VT, int R>
void launch_tile_t(const DevSliced &s, hipStream_t st, const double *x,
                   double *tout, const double *vin, double *vout, double c1, double c2) {
    if (!s.val_t) throw ArgError("tile-ordered values missing (internal)");
    // slot groups: the 3x3 stencil's 9 slots in one group up to width 16,
    // 3 at 32 / 64 to bound the registers holding operands in flight (C5:
    // groups of 5 or 9 at 32, 2 or 5 at 64 measured 0-11 % slower)
    constexpr int W = R <= 16 ? 9 : 3;
    const int wpb = 4;
    const int64_t grid = (s.ncell + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_tile<B, VT, R, W>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val_t, s.ncell, x, tout, vin, vout,
                       c1, c2);
}

template <typename VT, int R>
void launch_pair_t(const DevSliced &s, hipStream_t st, const double *x,
                   double *tout, const double *vin, double *vout, double c1, double c2) {
    if (!s.val_t || !s.pptr) throw ArgError("pair layout missing (internal)");
    // column groups: a 3x3-stencil pair's 12 columns in one group at 16
    // vectors, 6 and 4 at 32 and 64 (C5: 2-11 % faster than the
    // neighbouring choices measured)
    constexpr int W = R <= 16 ? 12 : (R == 32 ? 6 : 4);
    const int wpb = 4;
    const int64_t grid = (s.npair + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_pair<VT, R, W>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.pptr, s.pcol, s.pslot, (const VT *)s.val_t, s.ncell, x,
                       tout, vin, vout, c1, c2);
}

template <typename VT, int R>
void launch_quad_t(const DevSliced &s, hipStream_t st, const double *x,
                   double *tout, const double *vin, double *vout, double c1, double c2) {
    if (!s.val_q) throw ArgError("quad-ordered values missing (internal)");
    constexpr int W = R <= 16 ? 9 : (R == 32 ? 5 : 3);
    const int wpb = 4;
    const int64_t grid = (s.ncell + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_quad<VT, R, W>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val_q, s.ncell, x, tout, vin, vout,
                       c1, c2);
}

template <typename VT, int R>
void launch_quad4_t(const DevSliced &s, hipStream_t st, const double *x,
                    double *tout, const double *vin, double *vout, double c1, double c2) {
    constexpr int W = R <= 32 ? 9 : 5;
    const int wpb = 4;
    const int64_t grid = (s.ncell + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_quad4<VT, R, W>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x, tout, vin, vout,
                       c1, c2);
}

template <typename VT, int R, int W>
void launch_quad4c_t(const DevSliced &s, hipStream_t st, const double *x,
                     double *tout, const double *vin, double *vout, double c1, double c2) {
    const int wpb = 4;
    const int64_t ngrp = (s.ncell + 3) / 4;
    const int64_t grid = (ngrp + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_quad4c<VT, R, W>), dim3((unsigned)grid), dim3(64 * wpb), 0, st,
                       s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x, tout, vin, vout,
                       c1, c2, FusedRepDev{}, ShockBatchDev{});
}

template <int B, typename VT>
void launch_block_part(const DevSliced &s, int nv, hipStream_t st, const double *x,
                       double *tout, const double *vin, double *vout, double c1,
                       double c2) {
    if constexpr (B == 4) {
        // (batch kernel 8, one cell per wave, is 6 at b = 4)
        const bool dflt = s.tune.spmm_mfma == 6 || s.tune.spmm_mfma == 8;
        // four cells per wave on the 4x4x4 matrix cores (one-cell waves do
        // too little work at b = 4); slot groups of 9 at 8 and 16 doubles,
        // 3 at 32 and 64 (C4: 2-10 % faster than groups of 5 / 2 there,
        // 12-15 % slower at 16).  At 8 doubles per row this beats the
        // LDS-staged vector-FMA term kernel by 25-27 %.
        if (nv >= 8 && dflt) {
            switch (nv) {
            case 8: launch_quad4c_t<VT, 8, 9>(s, st, x, tout, vin, vout, c1, c2); return;
            case 16: launch_quad4c_t<VT, 16, 9>(s, st, x, tout, vin, vout, c1, c2); return;
            case 32: launch_quad4c_t<VT, 32, 3>(s, st, x, tout, vin, vout, c1, c2); return;
            case 64: launch_quad4c_t<VT, 64, 3>(s, st, x, tout, vin, vout, c1, c2); return;
            default: throw ArgError("unsupported vector width");
            }
        }
        if (nv >= 8 && s.tune.spmm_mfma == 7) {
            switch (nv) {
            case 8: launch_quad4_t<VT, 8>(s, st, x, tout, vin, vout, c1, c2); return;
            case 16: launch_quad4_t<VT, 16>(s, st, x, tout, vin, vout, c1, c2); return;
            case 32: launch_quad4_t<VT, 32>(s, st, x, tout, vin, vout, c1, c2); return;
            case 64: launch_quad4_t<VT, 64>(s, st, x, tout, vin, vout, c1, c2); return;
            default: throw ArgError("unsupported vector width");
            }
        }
    }
    if constexpr (B == 8) {
        // 16+ vectors under the default: two cells per matrix-core product
        if (nv >= 16 && s.tune.spmm_mfma == 6) {
            switch (nv) {
            case 16: launch_pair_t<VT, 16>(s, st, x, tout, vin, vout, c1, c2); return;
            case 32: launch_pair_t<VT, 32>(s, st, x, tout, vin, vout, c1, c2); return;
            case 64: launch_pair_t<VT, 64>(s, st, x, tout, vin, vout, c1, c2); return;
            default: throw ArgError("unsupported vector width");
            }
        }
        if (nv >= 8 && s.tune.spmm_mfma == 7) {
            switch (nv) {
            case 8: launch_quad_t<VT, 8>(s, st, x, tout, vin, vout, c1, c2); return;
            case 16: launch_quad_t<VT, 16>(s, st, x, tout, vin, vout, c1, c2); return;
            case 32: launch_quad_t<VT, 32>(s, st, x, tout, vin, vout, c1, c2); return;
            case 64: launch_quad_t<VT, 64>(s, st, x, tout, vin, vout, c1, c2); return;
            default: throw ArgError("unsupported vector width");
            }
        }
    }
    if constexpr (B == 8) {
        // (b = 4 returned above for every batch of 8+ under kernels 6 and 7)
        if (nv >= 8 && s.tune.spmm_mfma >= 6) {
            switch (nv) {
            case 8: launch_tile_t<B, VT, 8>(s, st, x, tout, vin, vout, c1, c2); return;
            case 16: launch_tile_t<B, VT, 16>(s, st, x, tout, vin, vout, c1, c2); return;
            case 32: launch_tile_t<B, VT, 32>(s, st, x, tout, vin, vout, c1, c2); return;
            case 64: launch_tile_t<B, VT, 64>(s, st, x, tout, vin, vout, c1, c2); return;
            default: throw ArgError("unsupported vector width");
            }
        }
    }
    switch (nv) {
    case 1: launch_term_t<B, VT, 1>(s, st, x, tout, vin, vout, c1, c2); break;
    case 2: launch_term_t<B, VT, 2>(s, st, x, tout, vin, vout, c1, c2); break;
    case 4: launch_term_t<B, VT, 4>(s, st, x, tout, vin, vout, c1, c2); break;
    case 8:
        if constexpr (B == 8) launch_spmm_t<B, VT, 8>(s, st, x, tout, vin, vout, c1, c2);
        else launch_term_t<B, VT, 8>(s, st, x, tout, vin, vout, c1, c2);
        break;
    case 16: launch_spmm_t<B, VT, 16>(s, st, x, tout, vin, vout, c1, c2); break;
    case 32: launch_spmm_t<B, VT, 32>(s, st, x, tout, vin, vout, c1, c2); break;
    case 64: launch_spmm_t<B, VT, 64>(s, st, x, tout, vin, vout, c1, c2); break;
    default: throw ArgError("unsupported vector width");
    }
}

void launch_block(const DevSliced &s, int nv, hipStream_t st, const double *x,
                  double *tout, const double *vin, double *vout, double c1, double c2) {
    const bool f32 = s.prec == HYMOR_PREC_F32;
    if (s.b == 4) {
        if (f32) launch_block_part<4, float>(s, nv, st, x, tout, vin, vout, c1, c2);
        else launch_block_part<4, double>(s, nv, st, x, tout, vin, vout, c1, c2);
    } else if (s.b == 8) {
        if (f32) launch_block_part<8, float>(s, nv, st, x, tout, vin, vout, c1, c2);
        else launch_block_part<8, double>(s, nv, st, x, tout, vin, vout, c1, c2);
    } else {
        throw ArgError("vars_per_cell must be 4 or 8");
    }
}

struct ColDiag {
    const double *d = nullptr;
    int64_t r0 = 0, r1 = 0;
};

template <int NV>
void launch_extra_nv(const DevExtra &xe, int64_t nflow, hipStream_t st, const double *x,
                     double *tout, const double *vin, double *vout, double c1,
                     double c2, const ColDiag &cd, int tile_b) {
    // one launch: a workgroup of kLongWaves waves per long row, then
    // workgroups of kLongWaves waves over the short rows, then (replicated
    // rows of an adjoint receptivity operator) over the shared rows
    constexpr int NW = kLongWaves;
    const int lpr = xe.is_complex ? (NV >= 2 ? NV / 2 : 1) : NV;
    const int64_t rpw = 64 / lpr;
    const int64_t nshort = xe.nrows - xe.nlong;
    const int64_t short_waves = (nshort + rpw - 1) / rpw;
    const int64_t nshort_wg = (short_waves + NW - 1) / NW;
    int64_t nrep_wg = 0;
    if (xe.rep.kind == 2) {
        const int64_t vpw = xe.is_complex ? RepLanes<NV, true>::VPW : RepLanes<NV, false>::VPW;
        nrep_wg = ((xe.rep.p + vpw - 1) / vpw + NW - 1) / NW;
    }
    if (xe.rep.kind && cd.d) throw ArgError("a per-column diagonal on replicated rows (internal)");
    if (xe.rep.kind == 1) {
        // column sums of the replicated blocks, and their diagonal rows
        const int64_t vpw = xe.is_complex ? RepLanes<NV, true>::VPW : RepLanes<NV, false>::VPW;
        const int64_t waves = (xe.rep.p + vpw - 1) / vpw;
        const unsigned g = (unsigned)((waves + 3) / 4);
        if (xe.is_complex) {
            if constexpr (NV >= 2)
                hipLaunchKernelGGL((k_rep_sum<NV, true>), dim3(g), dim3(256), 0, st, xe.rep, x,
                                   xe.rep_sum, tout, vin, vout, c1, c2);
        } else {
            hipLaunchKernelGGL((k_rep_sum<NV, false>), dim3(g), dim3(256), 0, st, xe.rep, x,
                               xe.rep_sum, tout, vin, vout, c1, c2);
        }
        HIP_TRY(hipGetLastError());
    }
    const int64_t grid = xe.nlong + nshort_wg + nrep_wg;
    if (grid == 0) return;
    // the replicated form is a variant of its own, so the other operators'
    // extra rows run without its virtual-column test and shared-row section
    auto go = [&](auto cplx, auto repl, auto lpe) {
        constexpr bool C = decltype(cplx)::value, RP = decltype(repl)::value;
        constexpr int L = decltype(lpe)::value;
        hipLaunchKernelGGL((k_extra<NV, C, NW, RP, L>), dim3((unsigned)grid), dim3(64 * NW), 0,
                           st, xe.rows, xe.ptr, xe.col, xe.val, xe.nlong, xe.nrows, nshort_wg,
                           nflow, x, tout, vin, vout, c1, c2, cd.d, cd.r0, cd.r1, tile_b, xe.rep);
    };
    const bool rep = xe.rep.kind != 0;
    if (xe.is_complex) {
        if constexpr (NV >= 2) {
            if (rep) go(std::true_type{}, std::true_type{}, std::integral_constant<int, 0>{});
            else go(std::true_type{}, std::false_type{}, std::integral_constant<int, 0>{});
        } else {
            throw ArgError("complex extras need two doubles per row");
        }
    } else if (rep) {
        go(std::false_type{}, std::true_type{}, std::integral_constant<int, 0>{});
    } else {
        // real long rows at 2..64 doubles per row: NV/2 lanes per entry,
        // one 16-byte load of x per lane and entry.  On the 400x300-cell
        // shock operator the Chu sandwich at 2, 4 and 8 vectors runs 3, 6
        // and 12 % faster than with NV lanes per entry; one lane per entry
        // is 11 % slower at 8 (16 bytes per cache-line access)
        if constexpr (NV >= 2) {
            if (!cd.d) {
                go(std::false_type{}, std::false_type{}, std::integral_constant<int, NV / 2>{});
                return;
            }
        }
        go(std::false_type{}, std::false_type{}, std::integral_constant<int, 0>{});
    }
}

void launch_extra(const DevExtra &xe, int nv, int64_t nflow, hipStream_t st, const double *x,
                  double *tout, const double *vin, double *vout, double c1,
                  double c2, const ColDiag &cd, int tile_b) {
    switch (nv) {
    case 1: launch_extra_nv<1>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 2: launch_extra_nv<2>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 4: launch_extra_nv<4>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 8: launch_extra_nv<8>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 16: launch_extra_nv<16>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 32: launch_extra_nv<32>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    case 64: launch_extra_nv<64>(xe, nflow, st, x, tout, vin, vout, c1, c2, cd, tile_b); break;
    default: throw ArgError("unsupported vector width");
    }
}

// Doubles per row from which the replicated freestream blocks are applied
// once (C4 receptivity operator, 32 copies: 16 vectors per sweep 11 %
// faster, 1 and 4 vectors 6-14 % slower -- there the extra rows' cost is
// launch and latency, not gathers, and the column sums take one more
// launch per term).
constexpr int kRepMinWidth = 16;

const DevExtra &extra_of(const hymor_op *op, bool adj, int nv) {
    const DevExtra &r = adj ? op->xa_rep : op->xf_rep;
    if ((nv >= kRepMinWidth || op->rep_only) && r.rep.kind) return r;
    return adj ? op->xa : op->xf;
}

int fused_kind(const hymor_op *op, bool adj, int nv);
void launch_term_fused(hymor_op *op, bool adj, hipStream_t st, const double *x, double *tout,
                       const double *vin, double *vout, double c1, double c2, int slot, int nv);

// One Taylor term (or a plain SpMV when vout == nullptr) of `op` on `st`;
// nv doubles per row.
void launch_term(hymor_op *op, bool adj, int nv, hipStream_t st, const double *x, double *tout,
                 const double *vin, double *vout, double c1, double c2) {
    // an adjoint whose only extra rows are its replicated blocks' shared
    // rows: one launch
    if (adj && fused_kind(op, true, nv)) {
        launch_term_fused(op, true, st, x, tout, vin, vout, c1, c2, 0, nv);
        return;
    }
    const DevSliced &s = adj ? op->adj : op->fwd;
    // one real vector under the default tuning (three waves per slice):
    // the shock bands in the term kernel's own launch
    const DevShock &sh = adj ? op->sha : op->shf;
    if (nv == 1 && sh.on && !op->cdiag_on && op->ncell > 0 && s.tune.split >= 2 &&
        s.tune.waves != 9) {
        launch_shock(s, sh, op->nflow, st, x, tout, vin, vout, c1, c2);
        HIP_TRY(hipGetLastError());
        return;
    }
    // batches: the extra-DOF rows and the panel slices' rows in the batch
    // kernel's leading workgroups, one launch
    if (nv >= 2 && sh.on && !op->cdiag_on && op->ncell > 0 && shock_batched_ok(op, s, nv)) {
        launch_shock_batched(s, sh, op->nflow, nv, st, x, tout, vin, vout, c1, c2);
        HIP_TRY(hipGetLastError());
        return;
    }
    const DevExtra &xe = extra_of(op, adj, nv);
    if (op->ncell > 0) {
        launch_block(s, nv, st, x, tout, vin, vout, c1, c2);
        HIP_TRY(hipGetLastError());
    }
    // then the extra rows: extra-DOF rows in full, flow rows adding their
    // extra entries' part to what the block kernel wrote
    if (xe.nrows > 0 || xe.rep.kind) {
        ColDiag cd;
        if (op->cdiag_on) cd = ColDiag{adj ? op->cdiag_adj : op->cdiag, op->cd0, op->cd1};
        launch_extra(xe, nv, op->nflow, st, x, tout, vin, vout, c1, c2, cd,
                     tile_b_of(op, nv));
        HIP_TRY(hipGetLastError());
    }
}

// Which fused term applies (0: none -- the block kernel and the extra rows
// run as separate launches).  Forward: Horner terms of an operator with the
// fused tables (build_fused_tables).  Adjoint: any term of an operator whose
// only extra rows are the shared rows of its replicated blocks (kind 2,
// starting at the first extra DOF).  Both: no per-column diagonal, and
//   1: two doubles per row under the default tuning (k_term_lds, one wave
//      per slice, non-temporal value loads);
//   2: 8-64 doubles per row at b = 4 under the default batch kernel
//      (k_spmm_quad4c).
int fused_kind(const hymor_op *op, bool adj, int nv) {
    if (op->ncell == 0 || op->cdiag_on) return 0;
    if (adj) {
        const DevExtra &xa = op->xa_rep;
        if (!op->has_adj || xa.rep.kind != 2 || xa.nrows != 0 || xa.rep.c0 != op->nflow)
            return 0;
    } else if (!op->xf_rep.fused) {
        return 0;
    }
    const TermTuning &t = op->fwd.tune;
    if (nv == 2) return (t.split == 3 && t.waves == 1 && t.nt == 1) ? 1 : 0;
    if (op->b == 4 && nv >= 8 && nv <= 64 && (t.spmm_mfma == 6 || t.spmm_mfma == 8)) return 2;
    return 0;
}

// Waves of rep_sum_wave over the replicated rows at NV doubles per row.
template <int NV>
int64_t fused_rep_waves_t(const DevExtra &xe) {
    const int64_t vpw = xe.is_complex ? RepLanes<NV, true>::VPW : RepLanes<NV, false>::VPW;
    return (xe.rep.p + vpw - 1) / vpw;
}

int64_t fused_rep_waves(const DevExtra &xe, int nv) {
    switch (nv) {
    case 2: return fused_rep_waves_t<2>(xe);
    case 8: return fused_rep_waves_t<8>(xe);
    case 16: return fused_rep_waves_t<16>(xe);
    case 32: return fused_rep_waves_t<32>(xe);
    case 64: return fused_rep_waves_t<64>(xe);
    default: throw ArgError("fused term: unsupported vector width (internal)");
    }
}

template <int NV>
void launch_fused_prologue_t(const DevExtra &xe, hipStream_t st, const double *x, int slot) {
    const unsigned g = (unsigned)((fused_rep_waves_t<NV>(xe) + 3) / 4);
    if (xe.is_complex)
        hipLaunchKernelGGL((k_rep_sum<NV, true>), dim3(g), dim3(256), 0, st, xe.rep, x,
                           xe.fr_sum[slot], nullptr, nullptr, nullptr, 0.0, 0.0);
    else
        hipLaunchKernelGGL((k_rep_sum<NV, false>), dim3(g), dim3(256), 0, st, xe.rep, x,
                           xe.fr_sum[slot], nullptr, nullptr, nullptr, 0.0, 0.0);
}

// Column sums of x's replicated rows (nv doubles per row) into
// fr_sum[slot], before the first fused term of a sweep.
void launch_fused_prologue(hymor_op *op, hipStream_t st, const double *x, int slot, int nv) {
    const DevExtra &xe = op->xf_rep;
    switch (nv) {
    case 2: launch_fused_prologue_t<2>(xe, st, x, slot); break;
    case 8: launch_fused_prologue_t<8>(xe, st, x, slot); break;
    case 16: launch_fused_prologue_t<16>(xe, st, x, slot); break;
    case 32: launch_fused_prologue_t<32>(xe, st, x, slot); break;
    case 64: launch_fused_prologue_t<64>(xe, st, x, slot); break;
    default: throw ArgError("fused term: unsupported vector width (internal)");
    }
    HIP_TRY(hipGetLastError());
}

template <int B, typename VT, int FR>
void launch_fused_lds_t(const DevSliced &s, const FusedRepDev &F, hipStream_t st,
                        const double *x, double *tout, const double *vin, double *vout,
                        double c1, double c2) {
    hipLaunchKernelGGL((k_term_lds<B, VT, 2, 1, 9, true, FR>),
                       dim3((unsigned)(F.nfs_wg + s.nslice)), dim3(64), 0, st, s.slice_ptr,
                       s.col, (const VT *)s.val, s.nslice, s.ncell, x, tout, vin, vout, c1, c2,
                       F, ShockBatchDev{});
}

template <typename VT, int R, int W, int FR>
void launch_fused_quad4c_t(const DevSliced &s, const FusedRepDev &F, hipStream_t st,
                           const double *x, double *tout, const double *vin, double *vout,
                           double c1, double c2) {
    const int wpb = 4;
    const int64_t ngrp = (s.ncell + 3) / 4;
    const int64_t grid = (ngrp + wpb - 1) / wpb;
    hipLaunchKernelGGL((k_spmm_quad4c<VT, R, W, FR>), dim3((unsigned)(F.nfs_wg + grid)),
                       dim3(64 * wpb), 0, st, s.slice_ptr, s.col, (const VT *)s.val, s.ncell, x,
                       tout, vin, vout, c1, c2, F, ShockBatchDev{});
}

// the slot groups launch_block_part uses at each width
template <typename VT, int FR>
void launch_fused_quad4c(const DevSliced &s, const FusedRepDev &F, int nv, hipStream_t st,
                         const double *x, double *tout, const double *vin, double *vout,
                         double c1, double c2) {
    switch (nv) {
    case 8: launch_fused_quad4c_t<VT, 8, 9, FR>(s, F, st, x, tout, vin, vout, c1, c2); break;
    case 16: launch_fused_quad4c_t<VT, 16, 9, FR>(s, F, st, x, tout, vin, vout, c1, c2); break;
    case 32: launch_fused_quad4c_t<VT, 32, 3, FR>(s, F, st, x, tout, vin, vout, c1, c2); break;
    case 64: launch_fused_quad4c_t<VT, 64, 3, FR>(s, F, st, x, tout, vin, vout, c1, c2); break;
    default: throw ArgError("fused term: unsupported vector width (internal)");
    }
}

template <int FR>
void launch_fused_fr(const DevSliced &s, const FusedRepDev &F, int kind, int nv, hipStream_t st,
                     const double *x, double *tout, const double *vin, double *vout, double c1,
                     double c2) {
    const bool f32 = s.prec == HYMOR_PREC_F32;
    if (kind == 1) {
        if (s.b == 4) {
            if (f32) launch_fused_lds_t<4, float, FR>(s, F, st, x, tout, vin, vout, c1, c2);
            else launch_fused_lds_t<4, double, FR>(s, F, st, x, tout, vin, vout, c1, c2);
        } else if (s.b == 8) {
            if (f32) launch_fused_lds_t<8, float, FR>(s, F, st, x, tout, vin, vout, c1, c2);
            else launch_fused_lds_t<8, double, FR>(s, F, st, x, tout, vin, vout, c1, c2);
        } else {
            throw ArgError("vars_per_cell must be 4 or 8");
        }
    } else {
        if (f32) launch_fused_quad4c<float, FR>(s, F, nv, st, x, tout, vin, vout, c1, c2);
        else launch_fused_quad4c<double, FR>(s, F, nv, st, x, tout, vin, vout, c1, c2);
    }
}

// One term in one launch (fused_kind != 0) at nv doubles per row.  Forward
// (a Horner term w_new = v + c1 L_ w, c2 = 1, tout null): the flow rows take
// the column sums of w from fr_sum[slot], the replicated rows leave those of
// w_new in fr_sum[slot ^ 1].  Adjoint (any term, as launch_term): the
// shared rows in the leading workgroups.
void launch_term_fused(hymor_op *op, bool adj, hipStream_t st, const double *x, double *tout,
                       const double *vin, double *vout, double c1, double c2, int slot,
                       int nv) {
    const DevSliced &s = adj ? op->adj : op->fwd;
    const DevExtra &xe = adj ? op->xa_rep : op->xf_rep;
    const int kind = fused_kind(op, adj, nv);
    if (kind == 0) throw ArgError("fused term does not apply (internal)");
    FusedRepDev F;
    F.R = xe.rep;
    if (adj) {
        F.nflow = op->nflow;
    } else {
        F.rot = xe.fr_rot;
        F.s_in = xe.fr_sum[slot];
        F.s_out = xe.fr_sum[slot ^ 1];
        F.slice_tab = xe.fr_slice;
        F.lane_ptr = xe.fr_lane;
        F.ecol = xe.fr_col;
        F.eval = xe.fr_val;
    }
    // replicated-row waves (rep_sum_wave and extra_replica_rows share their
    // lane layout): one wave per workgroup for k_term_lds, four for
    // k_spmm_quad4c; a multiple of 8 workgroups
    const int64_t waves = fused_rep_waves(xe, nv);
    F.nfs_wg = ((kind == 1 ? waves : (waves + 3) / 4) + 7) / 8 * 8;
    const int fr = (xe.is_complex ? 2 : 1) + (adj ? 4 : 0);
    switch (fr) {
    case 1: launch_fused_fr<1>(s, F, kind, nv, st, x, tout, vin, vout, c1, c2); break;
    case 2: launch_fused_fr<2>(s, F, kind, nv, st, x, tout, vin, vout, c1, c2); break;
    case 5: launch_fused_fr<5>(s, F, kind, nv, st, x, tout, vin, vout, c1, c2); break;
    default: launch_fused_fr<6>(s, F, kind, nv, st, x, tout, vin, vout, c1, c2); break;
    }
    HIP_TRY(hipGetLastError());
}

// width: doubles per internal row (>= nrhs*ncomp; the rest is zero padding)
void launch_perm_in(hymor_op *op, hipStream_t st, const double *x_vm, int64_t ld, double *out,
                    int nrhs, int ncomp, int64_t nzero, int width) {
    const int tile_b = tile_b_of(op, width);
    const int64_t total = op->n * width;
    const int64_t grid = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_perm_in, dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(256), 0, st,
                       x_vm, ld, out, op->n, op->ncell, op->b, nrhs, ncomp, nzero, width, tile_b);
    HIP_TRY(hipGetLastError());
}

void launch_perm_out(hymor_op *op, hipStream_t st, const double *in, double *y_vm, int64_t ld,
                     int nrhs, int ncomp, int64_t nzero, int width) {
    const int tile_b = tile_b_of(op, width);
    const int64_t total = op->n * nrhs * ncomp;
    const int64_t grid = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_perm_out, dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(256), 0, st,
                       in, y_vm, ld, op->n, op->ncell, op->b, nrhs, ncomp, nzero, width, tile_b);
    HIP_TRY(hipGetLastError());
}

// Supported internal row widths (doubles per row): powers of two up to 64.
int pad_width(int64_t valid) {
    for (int w = 1; w <= 64; w *= 2)
        if (valid <= w) return w;
    throw ArgError("nrhs too large: at most 64 doubles per row (64 real or 32 complex rhs)");
}

// The four n*nv-double buffers a Taylor sweep works in besides its input.
struct SweepBufs {
    double *va, *vb, *t0, *t1;
};

// T_order(A*dt)^nsub applied to `in` (cell-major, internal buffer, never
// written), working in `b`.  Returns the buffer that holds the result: one
// of b's four, or `in` if nsub == 0.
const double *run_taylor_in(hymor_op *op, const SweepBufs &b, bool adj, int order, int64_t nsub,
                            double dt, int form, int nv, const double *in) {
    double *va = b.va, *vb = b.vb, *t0 = b.t0, *t1 = b.t1;
    const double *cur = in;
    if (form == HYMOR_FORM_HORNER) {
        // w = v; w = v + (dt/j) A w for j = order, ..., 1.  Each term reads
        // the gathered w and v and writes the new w; v (the substep's input)
        // is never written, w ping-pongs between two other buffers.  With
        // the replicated rows fused into the term (one launch per term), the
        // column sums of each term's x alternate between two buffers: the
        // prologue forms those of the sweep's input, each term those of the
        // w it writes -- the next term's x, or the next substep's v = x.
        // (the adjoint's fused term needs no sums: launch_term takes it)
        const bool fused = !adj && fused_kind(op, adj, nv) != 0;
        int slot = 0;
        if (fused && nsub > 0) launch_fused_prologue(op, op->stream, cur, slot, nv);
        double *const pool[4] = {va, vb, t0, t1};
        for (int64_t s = 0; s < nsub; ++s) {
            double *pp[2] = {nullptr, nullptr};
            int np = 0;
            for (double *p : pool)
                if (p != cur && np < 2) pp[np++] = p;
            const double *w = cur;
            for (int j = order; j >= 1; --j) {
                double *wn = pp[(order - j) & 1];
                if (fused) {
                    launch_term_fused(op, adj, op->stream, w, nullptr, cur, wn, dt / (double)j,
                                      1.0, slot, nv);
                    slot ^= 1;
                } else {
                    launch_term(op, adj, nv, op->stream, w, nullptr, cur, wn, dt / (double)j,
                                1.0);
                }
                w = wn;
            }
            cur = w;
        }
        return cur;
    }
    for (int64_t s = 0; s < nsub; ++s) {
        double *nxt = (cur == va) ? vb : va;
        const double *x = cur;
        int64_t fact = 1;
        for (int j = 1; j <= order; ++j) {
            double c1, c2;
            if (form == HYMOR_FORM_POWERS) {
                fact *= j;
                c1 = dt;
                c2 = 1.0 / (double)fact;
            } else {
                c1 = dt / (double)j;
                c2 = 1.0;
            }
            double *tout = (j < order) ? ((j & 1) ? t0 : t1) : nullptr;
            const double *vin = (j == 1) ? cur : nxt;
            launch_term(op, adj, nv, op->stream, x, tout, vin, nxt, c1, c2);
            x = tout;
        }
        cur = nxt;
    }
    return cur;
}

// run_taylor_in on the operator's own workspace (ws[2..5])
const double *run_taylor(hymor_op *op, bool adj, int order, int64_t nsub, double dt, int form,
                         int nv, const double *in) {
    const SweepBufs b{op->ws[2], op->ws[3], op->ws[4], op->ws[5]};
    return run_taylor_in(op, b, adj, order, nsub, dt, form, nv, in);
}

// True when a rocprofiler tool library is configured for this process
// (rocprofv3 sets ROCP_TOOL_LIBRARIES for the program it runs).  Under
// `rocprofv3 --kernel-trace` the host process took a SIGSEGV inside the
// tool's frames on the first hipGraphLaunch of a captured Taylor sweep,
// while the same captured sweeps replay correctly without the tool; with a
// tool present the sweep is therefore launched kernel by kernel (the same
// kernels, in the same order, on the same stream).
bool profiler_tool_present() {
    static const bool present = [] {
        const char *v = std::getenv("ROCP_TOOL_LIBRARIES");
        return v != nullptr && v[0] != '\0';
    }();
    return present;
}

// run_taylor from op->ws[1], through a captured graph when enabled: the
// sweep touches only internal buffers, so one capture serves every call with
// the same parameters (the caller-facing permutations stay outside).
const double *run_taylor_ws1(hymor_op *op, bool adj, int order, int64_t nsub, double dt, int form,
                             int nv) {
    if (!op->use_graphs || nsub == 0 || profiler_tool_present())
        return run_taylor(op, adj, order, nsub, dt, form, nv, op->ws[1]);
    auto &g = op->graph;
    if (!(g.valid && g.adj == adj && g.order == order && g.nsub == nsub && g.dt == dt &&
          g.form == form && g.nv == nv)) {
        drop_graph(op);
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
        const double *res = nullptr;
        try {
            res = run_taylor(op, adj, order, nsub, dt, form, nv, op->ws[1]);
        } catch (...) {
            (void)hipStreamEndCapture(op->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIP_TRY(hipStreamEndCapture(op->stream, &graph));
        hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(e);
        g.valid = true;
        g.adj = adj;
        g.order = order;
        g.nsub = nsub;
        g.dt = dt;
        g.form = form;
        g.nv = nv;
        g.result = res;
    }
    HIP_TRY(hipGraphLaunch(g.exec, op->stream));
    return g.result;
}

int nv_of(hymor_op *op, int dtype) {
    ARG_CHECK(dtype == HYMOR_F64 || dtype == HYMOR_C128, "dtype must be HYMOR_F64 or HYMOR_C128");
    ARG_CHECK(!(op->is_complex && dtype == HYMOR_F64),
              "complex operator needs complex (HYMOR_C128) vectors");
    return dtype == HYMOR_C128 ? 2 : 1;
}

void check_taylor_args(hymor_op *op, int transpose, int order, int64_t nsub, int form) {
    ARG_CHECK(op != nullptr, "null operator");
    ARG_CHECK(order >= 1 && order <= 20, "order must be in [1, 20]");
    ARG_CHECK(nsub >= 0, "nsub must be >= 0");
    ARG_CHECK(form == HYMOR_FORM_NORMALISED || form == HYMOR_FORM_POWERS ||
                  form == HYMOR_FORM_HORNER,
              "unknown Taylor form");
    ARG_CHECK(!transpose || op->has_adj, "operator was built without its adjoint");
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_err.clear();
        return HYMOR_OK;
    } catch (const ArgError &e) {
        g_err = e.what();
        return HYMOR_ERR_ARG;
    } catch (const std::invalid_argument &e) {
        g_err = e.what();
        return HYMOR_ERR_ARG;
    } catch (const HipError &e) {
        g_err = e.what();
        return HYMOR_ERR_HIP;
    } catch (const std::bad_alloc &) {
        g_err = "allocation failed";
        return HYMOR_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_err = e.what();
        return HYMOR_ERR_ARG;
    }
}

hymor_op *new_op(int64_t nchi, int64_t neta, int b) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        throw HipError("no HIP device available");
    }
    hymor_op *op = new hymor_op();
    op->nchi = nchi;
    op->neta = neta;
    op->ncell = nchi * neta;
    op->b = b;
    op->nflow = (int64_t)b * op->ncell;
    HIP_TRY(hipGetDevice(&op->device));
    HIP_TRY(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    return op;
}

}  // namespace

// ------------------------------------------------------ no-alloc scratch
// Workspace of hymor_expm_taylor_dev_noalloc: one array per GPU, held from
// the library's code-object load to the end of the process.  The Taylor
// sweep needs five n*nv-double buffers (input, two ping-pong, two term
// buffers), each starting on a 256-byte boundary as hipMalloc'd buffers do,
// so the form serves n*nv <= kNoAllocRowDoubles.  Host code uses the device
// address from hipGetSymbolAddress (scratch_base), never the array's name.
// Every call on a GPU shares its array.
constexpr int64_t kNoAllocDoubles = int64_t(8) << 20;   // 64 MiB
constexpr int kNoAllocBufs = 5;
constexpr int64_t kNoAllocRowDoubles = (kNoAllocDoubles / kNoAllocBufs) & ~int64_t(31);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_base[kMaxDevices] = {};

// Device address of `device`'s scratch array, looked up once per device
// with that device current.
double *scratch_base(int device) {
    ARG_CHECK(device >= 0 && device < kMaxDevices, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_base[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != device) HIP_TRY(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        HIP_TRY(e);
        ARG_CHECK(p != nullptr, "no device address for the scratch array");
        g_scratch_base[device] = (double *)p;
    }
    return g_scratch_base[device];
}

}  // namespace

// ==================================================================== C ABI

extern "C" {

int hymor_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n > 0 ? 1 : 0;
}

const char *hymor_last_error(void) { return g_err.c_str(); }

int hymor_op_create_csc(int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int index_base, int dtype, double scale,
                        int vars_per_cell, int64_t nchi, int64_t neta, int value_prec,
                        hymor_op **out) {
    hymor_op *op = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "out is null");
        *out = nullptr;
        ARG_CHECK(colptr && rowval && nzval, "null CSC array");
        ARG_CHECK(index_base == 0 || index_base == 1, "index_base must be 0 or 1");
        ARG_CHECK(dtype == HYMOR_F64 || dtype == HYMOR_C128, "bad dtype");
        ARG_CHECK(vars_per_cell == 4 || vars_per_cell == 8, "vars_per_cell must be 4 or 8");
        ARG_CHECK(value_prec == HYMOR_PREC_F64 || value_prec == HYMOR_PREC_F32, "bad value_prec");
        ARG_CHECK(nchi > 0 && neta > 0, "grid must be non-empty");
        ARG_CHECK(n >= vars_per_cell * nchi * neta, "n smaller than vars_per_cell*nchi*neta");
        ARG_CHECK(n < (int64_t)INT32_MAX, "n too large for int32 column indices");
        HostBSR fwd, adj;
        HostExtra xf, xa;
        csc_to_blocks(n, colptr, rowval, (const double *)nzval, index_base, dtype == HYMOR_C128,
                      scale, vars_per_cell, nchi * neta, fwd, adj, xf, xa);
        HostExtra xfc = xf, xac = xa;
        HostReplica rf, ra;
        find_replicas(n, xfc, &xac, rf, ra);
        op = new_op(nchi, neta, vars_per_cell);
        op->n = n;
        op->prec = value_prec;
        op->is_complex = dtype == HYMOR_C128;
        upload_sliced(op, fwd, value_prec, op->fwd);
        upload_sliced(op, adj, value_prec, op->adj);
        upload_extra(op, xf, op->xf);
        upload_extra(op, xa, op->xa);
        upload_replicated(op, xfc, rf, op->xf_rep);
        upload_fused(op, xfc, rf, op->xf_rep);
        upload_replicated(op, xac, ra, op->xa_rep);
        if (rf.kind == 0) {
            BlockSrc bf, ba;
            bf.ptr = &fwd.ptr;
            bf.col = &fwd.col;
            bf.val = fwd.val.data();
            ba.ptr = &adj.ptr;
            ba.col = &adj.col;
            ba.val = adj.val.data();
            upload_shock(op, xf, bf, op->shf);
            upload_shock(op, xa, ba, op->sha);
        }
        op->has_adj = true;
        *out = op;
    });
    if (rc != HYMOR_OK && op) {
        release(op);
        delete op;
    }
    return rc;
}

int hymor_op_create_receptivity(int64_t n_sys, const int64_t *colptr, const int64_t *rowval,
                                const void *nzval, int index_base, int dtype, double scale,
                                int vars_per_cell, int64_t nchi, int64_t neta, int64_t n_l,
                                int64_t nwave, const double *w, int value_prec, hymor_op **out) {
    hymor_op *op = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "out is null");
        *out = nullptr;
        ARG_CHECK(colptr && rowval && nzval && w, "null argument");
        ARG_CHECK(index_base == 0 || index_base == 1, "index_base must be 0 or 1");
        ARG_CHECK(dtype == HYMOR_F64 || dtype == HYMOR_C128, "bad dtype");
        ARG_CHECK(vars_per_cell == 4 || vars_per_cell == 8, "vars_per_cell must be 4 or 8");
        ARG_CHECK(value_prec == HYMOR_PREC_F64 || value_prec == HYMOR_PREC_F32, "bad value_prec");
        ARG_CHECK(nchi > 0 && neta > 0, "grid must be non-empty");
        ARG_CHECK(n_l >= vars_per_cell * nchi * neta, "n_l smaller than vars_per_cell*nchi*neta");
        ARG_CHECK(n_sys > n_l, "the system needs freestream DOFs: n_sys > n_l");
        ARG_CHECK(nwave >= 1, "nwave must be positive");
        const int64_t p = n_sys - n_l;
        ARG_CHECK(nwave <= ((int64_t)INT32_MAX - n_l) / p - 1,
                  "n_l + (nwave + 1) * p too large for int32 column indices");
        for (int64_t k = 0; k < nwave; ++k) ARG_CHECK(std::isfinite(w[k]), "w must be finite");
        // the operator is complex (its -i w diagonal): real values get a zero
        // imaginary part
        const int64_t nnz = colptr[n_sys] - index_base;
        ARG_CHECK(nnz >= 0, "colptr is not monotone");
        std::vector<double> cval;
        const double *vals = (const double *)nzval;
        if (dtype == HYMOR_F64) {
            cval.assign((size_t)(2 * nnz), 0.0);
            for (int64_t e = 0; e < nnz; ++e) cval[(size_t)(2 * e)] = vals[e];
            vals = cval.data();
        }
        HostBSR fwd, adj;
        HostExtra xf, xa;
        csc_to_blocks(n_sys, colptr, rowval, vals, index_base, true, scale, vars_per_cell,
                      nchi * neta, fwd, adj, xf, xa);
        HostReplica rf, ra;
        receptivity_from_parts(n_l, p, nwave, w, scale, xf, xa, rf, ra);
        op = new_op(nchi, neta, vars_per_cell);
        op->n = n_l + nwave * p;
        op->prec = value_prec;
        op->is_complex = true;
        op->rep_only = true;
        upload_sliced(op, fwd, value_prec, op->fwd);
        upload_sliced(op, adj, value_prec, op->adj);
        upload_replicated(op, xf, rf, op->xf_rep);
        upload_fused(op, xf, rf, op->xf_rep);
        upload_replicated(op, xa, ra, op->xa_rep);
        op->has_adj = true;
        *out = op;
    });
    if (rc != HYMOR_OK && op) {
        release(op);
        delete op;
    }
    return rc;
}

int hymor_op_create_bsr(int64_t nchi, int64_t neta, int vars_per_cell, int64_t n_extra,
                        const int64_t *brow_ptr, const int32_t *bcol, const double *bval,
                        int64_t nx, const int64_t *xrow, const int64_t *xcol, const void *xval,
                        int dtype, double scale, int value_prec, int with_adjoint,
                        hymor_op **out) {
    hymor_op *op = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "out is null");
        *out = nullptr;
        ARG_CHECK(vars_per_cell == 4 || vars_per_cell == 8, "vars_per_cell must be 4 or 8");
        ARG_CHECK(nchi > 0 && neta > 0 && n_extra >= 0, "bad sizes");
        ARG_CHECK(dtype == HYMOR_F64 || dtype == HYMOR_C128, "bad dtype");
        ARG_CHECK(value_prec == HYMOR_PREC_F64 || value_prec == HYMOR_PREC_F32, "bad value_prec");
        ARG_CHECK(brow_ptr && bcol && bval, "null BSR array");
        ARG_CHECK(nx == 0 || (xrow && xcol && xval), "null extra array");
        const int b = vars_per_cell, bb = b * b;
        const int64_t ncell = nchi * neta;
        const int64_t n = (int64_t)b * ncell + n_extra;
        ARG_CHECK(n < (int64_t)INT32_MAX, "n too large for int32 column indices");
        ARG_CHECK(brow_ptr[0] == 0, "brow_ptr[0] must be 0");
        HostBSR fwd;
        fwd.b = b;
        fwd.ncell = ncell;
        fwd.ptr.assign(brow_ptr, brow_ptr + ncell + 1);
        for (int64_t c = 0; c < ncell; ++c)
            ARG_CHECK(fwd.ptr[c + 1] >= fwd.ptr[c], "brow_ptr not monotone");
        const int64_t nnzb = fwd.ptr[ncell];
        fwd.col.assign(bcol, bcol + nnzb);
        for (int64_t k = 0; k < nnzb; ++k)
            ARG_CHECK(fwd.col[k] >= 0 && fwd.col[k] < ncell, "bcol out of range");
        const bool cplx = dtype == HYMOR_C128;
        const int vs = cplx ? 2 : 1;
        std::vector<int64_t> r(xrow, xrow + nx), c(xcol, xcol + nx);
        std::vector<double> v((size_t)nx * vs);
        for (int64_t k = 0; k < nx; ++k) {
            ARG_CHECK(r[k] >= 0 && r[k] < n && c[k] >= 0 && c[k] < n, "extra index out of range");
            for (int q = 0; q < vs; ++q) v[(size_t)k * vs + q] = ((const double *)xval)[k * vs + q] * scale;
        }
        const int64_t nflow = (int64_t)b * ncell;
        HostExtra xf, xa;
        extra_from_coo(n, nflow, r, c, v, cplx, false, xf);
        if (with_adjoint) extra_from_coo(n, nflow, r, c, v, cplx, true, xa);
        HostExtra xfc = xf, xac = xa;
        HostReplica rf, ra;
        find_replicas(n, xfc, with_adjoint ? &xac : nullptr, rf, ra);
        op = new_op(nchi, neta, b);
        op->n = n;
        op->prec = value_prec;
        op->is_complex = cplx;
        // the caller's values go to the device once; both directions are
        // laid out (scaled, cast, transposed for the adjoint) from there
        DevTemp raw((size_t)nnzb * bb * sizeof(double));
        if (nnzb)
            HIP_TRY(hipMemcpy(raw.p, bval, (size_t)nnzb * bb * sizeof(double),
                              hipMemcpyHostToDevice));
        build_sliced(op, b, ncell, fwd.ptr, fwd.col, nullptr, false, (const double *)raw.p, scale,
                     value_prec, op->fwd);
        upload_extra(op, xf, op->xf);
        upload_replicated(op, xfc, rf, op->xf_rep);
        upload_fused(op, xfc, rf, op->xf_rep);
        BlockSrc bf;
        bf.ptr = &fwd.ptr;
        bf.col = &fwd.col;
        bf.val = bval;
        bf.scale = scale;
        if (rf.kind == 0) upload_shock(op, xf, bf, op->shf);
        if (with_adjoint) {
            std::vector<int64_t> tptr, tsrc;
            std::vector<int32_t> tcol;
            bsr_transpose_map(ncell, fwd.ptr, fwd.col, tptr, tcol, tsrc);
            build_sliced(op, b, ncell, tptr, tcol, &tsrc, true, (const double *)raw.p, scale,
                         value_prec, op->adj);
            upload_extra(op, xa, op->xa);
            upload_replicated(op, xac, ra, op->xa_rep);
            BlockSrc ba;
            ba.ptr = &tptr;
            ba.col = &tcol;
            ba.val = bval;
            ba.src = &tsrc;
            ba.transpose = true;
            ba.scale = scale;
            if (rf.kind == 0) upload_shock(op, xa, ba, op->sha);
            op->has_adj = true;
        }
        *out = op;
    });
    if (rc != HYMOR_OK && op) {
        release(op);
        delete op;
    }
    return rc;
}

int hymor_op_destroy(hymor_op *op) {
    return guarded([&] {
        if (!op) return;
        if (op->stream) HIP_TRY(hipStreamSynchronize(op->stream));
        release(op);
        delete op;
    });
}

int hymor_op_set_value_prec(hymor_op *op, int value_prec) {
    return guarded([&] {
        ARG_CHECK(op, "null argument");
        ARG_CHECK(value_prec == HYMOR_PREC_F64 || value_prec == HYMOR_PREC_F32, "bad value_prec");
        if (value_prec == op->prec) return;
        for (DevSliced *s : {&op->fwd, &op->adj}) {
            if (s == &op->adj && !op->has_adj) continue;
            if (s->nnzb_pad == 0) { s->prec = value_prec; continue; }
            ARG_CHECK(value_prec == HYMOR_PREC_F32 || s->val64,
                      "the values were uploaded in FP32; FP64 needs a new upload");
        }
        drop_graph(op);
        for (DevSliced *s : {&op->fwd, &op->adj}) {
            if (s == &op->adj && !op->has_adj) continue;
            if (s->nnzb_pad == 0) continue;
            const int64_t total = s->nnzb_pad * s->b * s->b;
            const size_t esz = s->prec == HYMOR_PREC_F32 ? sizeof(float) : sizeof(double);
            // the tile/quad copies hold the old precision: rebuilt on next use
            dfree(op, s->val_t, (size_t)total * esz);
            dfree(op, s->val_q, (size_t)total * esz);
            s->val_t = s->val_q = nullptr;
            if (value_prec == HYMOR_PREC_F32 && !s->val32) {
                s->val32 = dalloc(op, (size_t)total * sizeof(float));
                const int64_t grid =
                    std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
                hipLaunchKernelGGL(k_vals_to_f32, dim3((unsigned)grid), dim3(256), 0, op->stream,
                                   (const double *)s->val64, (float *)s->val32, total);
                HIP_TRY(hipGetLastError());
            }
            s->val = value_prec == HYMOR_PREC_F32 ? s->val32 : s->val64;
            s->prec = value_prec;
        }
        op->prec = value_prec;
        HIP_TRY(hipStreamSynchronize(op->stream));
    });
}

int hymor_op_info(const hymor_op *op, int64_t *out, int nout) {
    return guarded([&] {
        ARG_CHECK(op && out, "null argument");
        const DevExtra &xe = op->rep_only ? op->xf_rep : op->xf;
        const int64_t v[10] = {op->n, op->b, op->ncell, op->fwd.nnzb, op->fwd.nnzb_pad,
                               xe.nnz, xe.nrows, op->is_complex ? 1 : 0, op->prec,
                               op->dev_bytes};
        for (int i = 0; i < nout && i < 10; ++i) out[i] = v[i];
    });
}

int hymor_spmv(hymor_op *op, int transpose, const void *x, void *y, int dtype) {
    return guarded([&] {
        ARG_CHECK(op && x && y, "null argument");
        ARG_CHECK(!transpose || op->has_adj, "operator was built without its adjoint");
        const int nv = nv_of(op, dtype);
        ensure_ws(op, nv);
        const size_t bytes = (size_t)op->n * nv * sizeof(double);
        HIP_TRY(hipMemcpyAsync(op->ws[0], x, bytes, hipMemcpyHostToDevice, op->stream));
        launch_perm_in(op, op->stream, op->ws[0], op->n, op->ws[1], 1, nv, 0, nv);
        launch_term(op, transpose != 0, nv, op->stream, op->ws[1], op->ws[2], nullptr,
                    nullptr, 1.0, 0.0);
        launch_perm_out(op, op->stream, op->ws[2], op->ws[0], op->n, 1, nv, 0, nv);
        HIP_TRY(hipMemcpyAsync(y, op->ws[0], bytes, hipMemcpyDeviceToHost, op->stream));
        HIP_TRY(hipStreamSynchronize(op->stream));
    });
}

int hymor_spmv_dev(hymor_op *op, int transpose, const void *d_x, void *d_y, int dtype) {
    return guarded([&] {
        ARG_CHECK(op && d_x && d_y, "null argument");
        ARG_CHECK(!transpose || op->has_adj, "operator was built without its adjoint");
        const int nv = nv_of(op, dtype);
        ensure_ws(op, nv);
        launch_perm_in(op, op->stream, (const double *)d_x, op->n, op->ws[1], 1, nv, 0, nv);
        launch_term(op, transpose != 0, nv, op->stream, op->ws[1], op->ws[2], nullptr, nullptr,
                    1.0, 0.0);
        launch_perm_out(op, op->stream, op->ws[2], (double *)d_y, op->n, 1, nv, 0, nv);
    });
}

static void check_batch(hymor_op *op, int64_t nrhs, int64_t ldx, int64_t ldy) {
    ARG_CHECK(nrhs >= 1, "nrhs must be >= 1");
    ARG_CHECK(ldx >= op->n && ldy >= op->n, "leading dimension smaller than n");
}

int hymor_expm_taylor_batched_dev(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                                  int form, int64_t nrhs, const void *d_X, int64_t ldx, void *d_Y,
                                  int64_t ldy, int dtype) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(d_X && d_Y, "null matrix");
        check_batch(op, nrhs, ldx, ldy);
        const int ncomp = nv_of(op, dtype);
        const int w = pad_width(nrhs * ncomp);
        ensure_ws(op, w);
        launch_perm_in(op, op->stream, (const double *)d_X, ldx, op->ws[1], (int)nrhs, ncomp, 0, w);
        const double *res = run_taylor_ws1(op, transpose != 0, order, nsub, dt, form, w);
        launch_perm_out(op, op->stream, res, (double *)d_Y, ldy, (int)nrhs, ncomp, 0, w);
    });
}

int hymor_expm_taylor_batched_coldiag_dev(hymor_op *op, int transpose, int order, int64_t nsub,
                                          double dt, int form, int64_t nrhs, const double *diag,
                                          int64_t row0, int64_t row1, const void *d_X,
                                          int64_t ldx, void *d_Y, int64_t ldy) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(d_X && d_Y && diag, "null argument");
        check_batch(op, nrhs, ldx, ldy);
        ARG_CHECK(row0 >= op->nflow && row0 <= row1 && row1 <= op->n,
                  "the diagonal block must lie on extra DOFs: nflow <= row0 <= row1 <= n");
        ARG_CHECK(!op->rep_only,
                  "a receptivity operator carries its wavenumbers: no per-column diagonal");
        const int ncomp = 2;
        const int w = pad_width(nrhs * ncomp);
        ensure_ws(op, w);
        // per-column diagonal (and its conjugate for the adjoint), padded
        std::vector<double> d((size_t)w, 0.0), da((size_t)w, 0.0);
        for (int64_t c = 0; c < nrhs; ++c) {
            d[2 * c] = diag[2 * c];
            d[2 * c + 1] = diag[2 * c + 1];
            da[2 * c] = diag[2 * c];
            da[2 * c + 1] = -diag[2 * c + 1];
        }
        if (!op->cdiag) {
            op->cdiag = (double *)dalloc(op, 64 * sizeof(double));
            op->cdiag_adj = (double *)dalloc(op, 64 * sizeof(double));
        }
        HIP_TRY(hipMemcpyAsync(op->cdiag, d.data(), (size_t)w * sizeof(double),
                               hipMemcpyHostToDevice, op->stream));
        HIP_TRY(hipMemcpyAsync(op->cdiag_adj, da.data(), (size_t)w * sizeof(double),
                               hipMemcpyHostToDevice, op->stream));
        op->cd0 = row0;
        op->cd1 = row1;
        op->cdiag_on = true;
        struct Off {
            hymor_op *o;
            ~Off() { o->cdiag_on = false; }
        } off{op};
        launch_perm_in(op, op->stream, (const double *)d_X, ldx, op->ws[1], (int)nrhs, ncomp, 0, w);
        // not graph-captured: the diagonal buffer is rewritten per call
        const double *res = run_taylor(op, transpose != 0, order, nsub, dt, form, w, op->ws[1]);
        launch_perm_out(op, op->stream, res, (double *)d_Y, ldy, (int)nrhs, ncomp, 0, w);
        HIP_TRY(hipStreamSynchronize(op->stream));
    });
}

int hymor_expm_taylor_batched(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                              int form, int64_t nrhs, const void *X, int64_t ldx, void *Y,
                              int64_t ldy, int dtype) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(X && Y, "null matrix");
        check_batch(op, nrhs, ldx, ldy);
        const int ncomp = nv_of(op, dtype);
        const int w = pad_width(nrhs * ncomp);
        ensure_ws(op, w);
        // ws[0] holds n*w doubles >= n*nrhs*ncomp: stage the columns packed (ld = n)
        const size_t esz = (size_t)ncomp * sizeof(double);
        for (int64_t j = 0; j < nrhs; ++j)
            HIP_TRY(hipMemcpyAsync((char *)op->ws[0] + (size_t)j * op->n * esz,
                                   (const char *)X + (size_t)j * ldx * esz, (size_t)op->n * esz,
                                   hipMemcpyHostToDevice, op->stream));
        launch_perm_in(op, op->stream, op->ws[0], op->n, op->ws[1], (int)nrhs, ncomp, 0, w);
        const double *res = run_taylor_ws1(op, transpose != 0, order, nsub, dt, form, w);
        launch_perm_out(op, op->stream, res, op->ws[0], op->n, (int)nrhs, ncomp, 0, w);
        for (int64_t j = 0; j < nrhs; ++j)
            HIP_TRY(hipMemcpyAsync((char *)Y + (size_t)j * ldy * esz,
                                   (const char *)op->ws[0] + (size_t)j * op->n * esz,
                                   (size_t)op->n * esz, hipMemcpyDeviceToHost, op->stream));
        HIP_TRY(hipStreamSynchronize(op->stream));
    });
}

int hymor_expm_taylor_dev(hymor_op *op, int transpose, int order, int64_t nsub, double dt,
                          int form, const void *d_x, void *d_y, int dtype) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(d_x && d_y, "null vector");
        const int nv = nv_of(op, dtype);
        ensure_ws(op, nv);
        launch_perm_in(op, op->stream, (const double *)d_x, op->n, op->ws[1], 1, nv, 0, nv);
        const double *res = run_taylor_ws1(op, transpose != 0, order, nsub, dt, form, nv);
        launch_perm_out(op, op->stream, res, (double *)d_y, op->n, 1, nv, 0, nv);
    });
}

int hymor_expm_taylor_dev_noalloc(hymor_op *op, int transpose, int order, int64_t nsub,
                                  double dt, int form, const void *d_x, void *d_y, int dtype) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(d_x && d_y, "null vector");
        const int nv = nv_of(op, dtype);
        const int64_t len = op->n * nv;
        const int64_t stride = (len + 31) & ~int64_t(31);   // 256-byte aligned buffers
        ARG_CHECK(stride <= kNoAllocRowDoubles,
                  "hymor_expm_taylor_dev_noalloc: n * (1 real, 2 complex) = " +
                      std::to_string(len) + " exceeds its scratch limit of " +
                      std::to_string(kNoAllocRowDoubles) + " doubles per vector");
        double *s = scratch_base(op->device);
        // every buffer is written in full before it is read (the input by
        // the permutation, the others by the sweep), so the scratch needs no
        // clearing between calls
        double *in = s;
        const SweepBufs b{s + stride, s + 2 * stride, s + 3 * stride, s + 4 * stride};
        launch_perm_in(op, op->stream, (const double *)d_x, op->n, in, 1, nv, 0, nv);
        const double *res = run_taylor_in(op, b, transpose != 0, order, nsub, dt, form, nv, in);
        launch_perm_out(op, op->stream, res, (double *)d_y, op->n, 1, nv, 0, nv);
    });
}

int hymor_expm_taylor(hymor_op *op, int transpose, int order, int64_t nsub, double dt, int form,
                      const void *x, void *y, int dtype) {
    return guarded([&] {
        check_taylor_args(op, transpose, order, nsub, form);
        ARG_CHECK(x && y, "null vector");
        const int nv = nv_of(op, dtype);
        ensure_ws(op, nv);
        const size_t bytes = (size_t)op->n * nv * sizeof(double);
        HIP_TRY(hipMemcpyAsync(op->ws[0], x, bytes, hipMemcpyHostToDevice, op->stream));
        launch_perm_in(op, op->stream, op->ws[0], op->n, op->ws[1], 1, nv, 0, nv);
        const double *res = run_taylor_ws1(op, transpose != 0, order, nsub, dt, form, nv);
        launch_perm_out(op, op->stream, res, op->ws[0], op->n, 1, nv, 0, nv);
        HIP_TRY(hipMemcpyAsync(y, op->ws[0], bytes, hipMemcpyDeviceToHost, op->stream));
        HIP_TRY(hipStreamSynchronize(op->stream));
    });
}

// The caller's CSC pattern (var-major rows and columns, as
// hymor_op_create_csc takes A) in the operator's cell-major indices, one
// (row, column) pair per entry in the pattern's order.
static void pattern_cellmajor(const hymor_op *op, const int64_t *colptr, const int64_t *rowval,
                              int index_base, std::vector<int32_t> &er,
                              std::vector<int32_t> &ec) {
    ARG_CHECK(colptr && rowval, "null pattern");
    ARG_CHECK(index_base == 0 || index_base == 1, "index_base must be 0 or 1");
    const int64_t n = op->n, base = index_base;
    ARG_CHECK(colptr[0] == base, "colptr[0] must equal the index base");
    for (int64_t j = 0; j < n; ++j)
        ARG_CHECK(colptr[j + 1] >= colptr[j], "colptr is not monotone");
    const int64_t nnz = colptr[n] - base;
    // var-major index vm = var * ncell + cell -> cell * b + var, extras unchanged
    std::vector<int32_t> cm((size_t)n);
    for (int64_t vm = 0; vm < n; ++vm)
        cm[(size_t)vm] =
            (int32_t)(vm < op->nflow ? (vm % op->ncell) * op->b + vm / op->ncell : vm);
    er.resize((size_t)nnz);
    ec.resize((size_t)nnz);
    for (int64_t j = 0; j < n; ++j) {
        const int32_t cj = cm[(size_t)j];
        for (int64_t e = colptr[j] - base; e < colptr[j + 1] - base; ++e) {
            const int64_t r = rowval[e] - base;
            ARG_CHECK(r >= 0 && r < n, "rowval out of range");
            er[(size_t)e] = cm[(size_t)r];
            ec[(size_t)e] = cj;
        }
    }
}

// A pattern mapped for the device: its entries as cell-major (row, column)
// pairs and, for operators of block size 4 or 8, the same entries grouped by
// b x b block for k_sddmm_blocks (nblk blocks: bcr, bslot), with the entries
// no block takes (rer, rec, ridx) left to the per-entry kernel.
struct hymor_pattern {
    int64_t n = 0, b = 0, ncell = 0, nnz = 0;
    int32_t *er = nullptr, *ec = nullptr;
    int64_t nblk = 0, nres = 0;
    int32_t *bcr = nullptr, *bslot = nullptr, *rer = nullptr, *rec = nullptr, *ridx = nullptr;
};

// Groups a cell-major pattern's flow entries by b x b block: blocks ordered
// by the cell of their columns, then of their rows (bcr: the two cells;
// slot: b*b entry positions, -1 for none).  A block is kept when it holds at
// least b entries (reading its 2b values per column then costs no more than
// the 2 per entry of the per-entry kernel); the entries of every other block,
// those in extra-DOF rows or columns and a repeated (row, column) stay in
// the per-entry list, in pattern order (rer, rec; ridx their positions).
static void build_pattern_blocks(int b, int64_t ncell, int64_t nflow,
                                 const std::vector<int32_t> &er, const std::vector<int32_t> &ec,
                                 std::vector<int32_t> &bcr, std::vector<int32_t> &slot,
                                 std::vector<int32_t> &rer, std::vector<int32_t> &rec,
                                 std::vector<int32_t> &ridx) {
    const int64_t nnz = (int64_t)er.size();
    std::vector<int64_t> cptr((size_t)ncell + 1, 0);
    std::vector<int32_t> res;
    for (int64_t e = 0; e < nnz; ++e) {
        if (er[(size_t)e] < nflow && ec[(size_t)e] < nflow)
            ++cptr[(size_t)(ec[(size_t)e] / b) + 1];
        else
            res.push_back((int32_t)e);
    }
    for (int64_t c = 0; c < ncell; ++c) cptr[(size_t)c + 1] += cptr[(size_t)c];
    std::vector<int32_t> ord((size_t)cptr[(size_t)ncell]);
    {
        std::vector<int64_t> pos(cptr.begin(), cptr.end() - 1);
        for (int64_t e = 0; e < nnz; ++e)
            if (er[(size_t)e] < nflow && ec[(size_t)e] < nflow)
                ord[(size_t)pos[(size_t)(ec[(size_t)e] / b)]++] = (int32_t)e;
    }
    std::vector<std::pair<int32_t, int32_t>> buf;
    std::vector<int32_t> tile((size_t)(b * b));
    for (int64_t c = 0; c < ncell; ++c) {
        buf.clear();
        for (int64_t q = cptr[(size_t)c]; q < cptr[(size_t)c + 1]; ++q) {
            const int32_t e = ord[(size_t)q];
            buf.emplace_back(er[(size_t)e] / b, e);
        }
        std::sort(buf.begin(), buf.end());
        for (size_t q0 = 0; q0 < buf.size();) {
            size_t q1 = q0;
            while (q1 < buf.size() && buf[q1].first == buf[q0].first) ++q1;
            std::fill(tile.begin(), tile.end(), -1);
            int held = 0;
            for (size_t q = q0; q < q1; ++q) {
                const int32_t e = buf[q].second;
                int32_t &t = tile[(size_t)((er[(size_t)e] % b) * b + ec[(size_t)e] % b)];
                if (t < 0) {
                    t = e;
                    ++held;
                } else {
                    res.push_back(e);
                }
            }
            if (held >= b) {
                bcr.push_back(buf[q0].first);
                bcr.push_back((int32_t)c);
                slot.insert(slot.end(), tile.begin(), tile.end());
            } else {
                for (int32_t e : tile)
                    if (e >= 0) res.push_back(e);
            }
            q0 = q1;
        }
    }
    std::sort(res.begin(), res.end());
    rer.resize(res.size());
    rec.resize(res.size());
    ridx = res;
    for (size_t q = 0; q < res.size(); ++q) {
        rer[q] = er[(size_t)res[q]];
        rec[q] = ec[(size_t)res[q]];
    }
}

static void check_vjp_args(const hymor_op *op, int order, int64_t nsub, int64_t ckpt_every) {
    ARG_CHECK(op != nullptr, "null operator");
    ARG_CHECK(!op->is_complex, "the operator gradient is for real operators");
    ARG_CHECK(order >= 1 && order <= 20, "order must be in [1, 20]");
    ARG_CHECK(nsub >= 0, "nsub must be >= 0");
    ARG_CHECK(ckpt_every >= 1, "ckpt_every must be >= 1");
    ARG_CHECK(op->has_adj, "operator was built without its adjoint");
}

// The gradient of <wbar, T_p(dt A)^nsub x> on the pattern (d_er, d_ec: nnz
// cell-major pairs) into d_grad; x and wbar are var-major device vectors.
// With `blocks` (a device pattern grouped by block), its blocks go through
// k_sddmm_blocks and the rest through the per-entry kernel: the same sums,
// the same bits.  Synchronous (its buffers are released on return).
static void vjp_core(hymor_op *op, int order, int64_t nsub, double dt, int64_t ckpt_every,
                     const double *d_x, const double *d_wbar, const int32_t *d_er,
                     const int32_t *d_ec, int64_t nnz, double *d_grad,
                     const hymor_pattern *blocks = nullptr) {
    hipStream_t st = op->stream;
    if (nnz == 0) return;
    if (nsub == 0) {   // w = x does not depend on A
        HIP_TRY(hipMemsetAsync(d_grad, 0, (size_t)nnz * 8, st));
        HIP_TRY(hipStreamSynchronize(st));
        return;
    }
    const int64_t n = op->n;
    const int p = order;
    const int64_t K = std::min(ckpt_every, nsub);    // substeps per checkpoint
    const int64_t ng = (nsub + K - 1) / K;
    const size_t vb = (size_t)n * sizeof(double);
    // the backward sweep takes the states in segments: a checkpoint group,
    // recomputed from its checkpoint, or, when every state is kept
    // (ckpt_every 1), up to kVjpBatch consecutive kept states, recomputing
    // nothing.  A segment's forward powers y_r = (dt A)^r x_i, r < p, run as
    // p - 1 batched SpMVs over its states (one launch per power instead of
    // one per power and state); the workspace for that width is sized
    // before any work is queued
    constexpr int64_t kVjpBatch = 16;
    const bool keep_all = K == 1 && nsub > 1;
    const int64_t S = keep_all ? std::min(nsub, kVjpBatch) : K;   // states per segment
    const int64_t nseg = (nsub + S - 1) / S;
    const int wb = (S >= 2 && p >= 2) ? pad_width(std::min(S, kVjpBatch)) : 1;
    ensure_ws(op, wb);
    const int tile_b = tile_b_of(op, wb);
    std::unique_ptr<DevTemp> d_yb;
    if (wb > 1) d_yb.reset(new DevTemp((size_t)p * n * wb * sizeof(double)));
    DevTemp d_ck((size_t)ng * vb), d_grp((size_t)std::max<int64_t>(K - 1, 1) * vb);
    DevTemp d_U((size_t)p * S * vb), d_Z((size_t)p * S * vb);
    DevTemp d_Y((size_t)std::max(p - 1, 1) * vb), d_lam(vb), d_up(vb);
    double *ck = (double *)d_ck.p, *grp = (double *)d_grp.p, *U = (double *)d_U.p,
           *Z = (double *)d_Z.p, *Y = (double *)d_Y.p, *lam = (double *)d_lam.p,
           *up = (double *)d_up.p;
    double invf[kMaxVjp];
    invf[0] = 1.0;
    for (int j = 1; j <= p; ++j) invf[j] = invf[j - 1] / (double)j;
    const unsigned gv = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256,
                                                                          65536));
    // forward: the state at the start of every K-substep group
    launch_perm_in(op, st, d_x, n, ck, 1, 1, 0, 1);
    for (int64_t g = 0; g + 1 < ng; ++g) {
        const double *res = run_taylor(op, false, p, K, dt, HYMOR_FORM_HORNER, 1, ck + g * n);
        HIP_TRY(hipMemcpyAsync(ck + (g + 1) * n, res, vb, hipMemcpyDeviceToDevice, st));
    }
    // backward, group by group from the last: the group's states are
    // recomputed from its checkpoint, then lambda steps back through them
    launch_perm_in(op, st, d_wbar, n, lam, 1, 1, 0, 1);
    int init = 1;
    for (int64_t g = nseg - 1; g >= 0; --g) {
        const int64_t cnt = std::min(S, nsub - g * S);
        std::vector<const double *> xs((size_t)cnt);
        if (keep_all) {
            for (int64_t i = 0; i < cnt; ++i) xs[(size_t)i] = ck + (g * S + i) * n;
        } else {
            xs[0] = ck + g * n;
            for (int64_t i = 1; i < cnt; ++i) {
                const double *res =
                    run_taylor(op, false, p, 1, dt, HYMOR_FORM_HORNER, 1, xs[(size_t)i - 1]);
                double *dst = grp + (i - 1) * n;
                HIP_TRY(hipMemcpyAsync(dst, res, vb, hipMemcpyDeviceToDevice, st));
                xs[(size_t)i] = dst;
            }
        }
        const bool batched = wb > 1 && cnt >= 2 && cnt <= wb;
        if (batched) {
            // the states as columns, p - 1 batched SpMVs, then every z_m
            double *yb = (double *)d_yb->p;
            VecList xl{};
            for (int64_t i = 0; i < cnt; ++i) xl.v[i] = xs[(size_t)i];
            xl.cnt = (int)cnt;
            const int64_t tot = n * wb;
            const unsigned gb =
                (unsigned)std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 65536));
            hipLaunchKernelGGL(k_cols_to_batch, dim3(gb), dim3(256), 0, st, xl, (int)cnt, n, wb,
                               tile_b, op->nflow, yb);
            HIP_TRY(hipGetLastError());
            BatchList yl{};
            yl.v[0] = yb;
            for (int r = 1; r < p; ++r) {
                double *yr = yb + (size_t)r * n * wb;
                launch_term(op, false, wb, st, yl.v[r - 1], yr, nullptr, nullptr, dt, 0.0);
                yl.v[r] = yr;
            }
            for (int j = 0; j <= p; ++j) yl.c[j] = invf[j];
            hipLaunchKernelGGL(k_vjp_z_batch, dim3(gb), dim3(256), 0, st, yl, p, (int)cnt, n, wb,
                               tile_b, op->nflow, Z);
            HIP_TRY(hipGetLastError());
        }
        for (int64_t i = cnt - 1; i >= 0; --i) {
            double *Uk = U + (size_t)i * p * n, *Zk = Z + (size_t)i * p * n;
            if (!batched) {
                // y_r = (dt A)^r x_k, then the z_m
                VecList ys{};
                ys.v[0] = xs[(size_t)i];
                for (int r = 1; r < p; ++r) {
                    double *yr = Y + (size_t)(r - 1) * n;
                    launch_term(op, false, 1, st, ys.v[r - 1], yr, nullptr, nullptr, dt, 0.0);
                    ys.v[r] = yr;
                }
                for (int j = 0; j <= p; ++j) ys.c[j] = invf[j];
                ys.cnt = p;
                hipLaunchKernelGGL(k_vjp_z, dim3(gv), dim3(256), 0, st, ys, p, n, Zk);
                HIP_TRY(hipGetLastError());
            }
            // u_m = (dt A')^m lambda, m < p; u_p into up
            HIP_TRY(hipMemcpyAsync(Uk, lam, vb, hipMemcpyDeviceToDevice, st));
            for (int m = 1; m <= p; ++m) {
                double *um = m < p ? Uk + (size_t)m * n : up;
                launch_term(op, true, 1, st, Uk + (size_t)(m - 1) * n, um, nullptr, nullptr,
                            dt, 0.0);
            }
            // lambda_k = S' lambda_{k+1} = sum_{m <= p} u_m / m!
            VecList us{};
            for (int m = 0; m <= p; ++m) {
                us.v[m] = m < p ? Uk + (size_t)m * n : up;
                us.c[m] = invf[m];
            }
            us.cnt = p + 1;
            hipLaunchKernelGGL(k_lincomb, dim3(gv), dim3(256), 0, st, us, n, lam);
            HIP_TRY(hipGetLastError());
        }
        const int kc = (int)(cnt * p);
        if (blocks && blocks->nblk > 0) {
            const unsigned gbk = (unsigned)((blocks->nblk + 255) / 256);
            if (op->b == 4)
                hipLaunchKernelGGL(k_sddmm_blocks<4>, dim3(gbk), dim3(256), 0, st, blocks->bcr,
                                   blocks->bslot, blocks->nblk, (const double *)U,
                                   (const double *)Z, n, kc, dt, init, d_grad);
            else
                hipLaunchKernelGGL(k_sddmm_blocks<8>, dim3(gbk), dim3(256), 0, st, blocks->bcr,
                                   blocks->bslot, blocks->nblk, (const double *)U,
                                   (const double *)Z, n, kc, dt, init, d_grad);
            HIP_TRY(hipGetLastError());
            if (blocks->nres > 0) {
                const unsigned gr = (unsigned)((blocks->nres + 255) / 256);
                hipLaunchKernelGGL(k_sddmm_pattern, dim3(gr), dim3(256), 0, st, blocks->rer,
                                   blocks->rec, blocks->nres, (const double *)U,
                                   (const double *)Z, n, kc, dt, init,
                                   (const int32_t *)blocks->ridx, d_grad);
                HIP_TRY(hipGetLastError());
            }
        } else {
            const unsigned ge = (unsigned)((nnz + 255) / 256);
            hipLaunchKernelGGL(k_sddmm_pattern, dim3(ge), dim3(256), 0, st, d_er, d_ec, nnz,
                               (const double *)U, (const double *)Z, n, kc, dt, init,
                               (const int32_t *)nullptr, d_grad);
            HIP_TRY(hipGetLastError());
        }
        init = 0;
    }
    HIP_TRY(hipStreamSynchronize(st));
}

int hymor_expm_taylor_vjp_values(hymor_op *op, int order, int64_t nsub, double dt,
                                 int64_t ckpt_every, const double *x, const double *wbar,
                                 const int64_t *colptr, const int64_t *rowval, int index_base,
                                 double *grad) {
    return guarded([&] {
        check_vjp_args(op, order, nsub, ckpt_every);
        ARG_CHECK(x && wbar && grad, "null argument");
        std::vector<int32_t> er, ec;
        pattern_cellmajor(op, colptr, rowval, index_base, er, ec);
        const int64_t nnz = (int64_t)er.size();
        if (nnz == 0) return;
        const size_t vb = (size_t)op->n * sizeof(double);
        DevTemp d_er((size_t)nnz * 4), d_ec((size_t)nnz * 4), d_grad((size_t)nnz * 8);
        DevTemp d_x(vb), d_w(vb);
        hipStream_t st = op->stream;
        HIP_TRY(hipMemcpyAsync(d_er.p, er.data(), (size_t)nnz * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_ec.p, ec.data(), (size_t)nnz * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_x.p, x, vb, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_w.p, wbar, vb, hipMemcpyHostToDevice, st));
        vjp_core(op, order, nsub, dt, ckpt_every, (const double *)d_x.p, (const double *)d_w.p,
                 (const int32_t *)d_er.p, (const int32_t *)d_ec.p, nnz, (double *)d_grad.p);
        HIP_TRY(hipMemcpyAsync(grad, d_grad.p, (size_t)nnz * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    });
}

int hymor_pattern_create(const hymor_op *op, const int64_t *colptr, const int64_t *rowval,
                         int index_base, hymor_pattern **out) {
    hymor_pattern *pat = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(op && out, "null argument");
        *out = nullptr;
        std::vector<int32_t> er, ec;
        pattern_cellmajor(op, colptr, rowval, index_base, er, ec);
        pat = new hymor_pattern();
        pat->n = op->n;
        pat->b = op->b;
        pat->ncell = op->ncell;
        pat->nnz = (int64_t)er.size();
        auto up = [](int32_t **d, const std::vector<int32_t> &h) {
            if (h.empty()) return;
            HIP_TRY(hipMalloc((void **)d, h.size() * 4));
            HIP_TRY(hipMemcpy(*d, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        };
        up(&pat->er, er);
        up(&pat->ec, ec);
        if ((op->b == 4 || op->b == 8) && pat->nnz > 0 && pat->nnz <= INT32_MAX) {
            std::vector<int32_t> bcr, bslot, rer, rec, ridx;
            build_pattern_blocks(op->b, op->ncell, op->nflow, er, ec, bcr, bslot, rer, rec,
                                 ridx);
            pat->nblk = (int64_t)bcr.size() / 2;
            pat->nres = (int64_t)ridx.size();
            up(&pat->bcr, bcr);
            up(&pat->bslot, bslot);
            up(&pat->rer, rer);
            up(&pat->rec, rec);
            up(&pat->ridx, ridx);
        }
        *out = pat;
    });
    if (rc != HYMOR_OK && pat) {
        for (int32_t *d : {pat->er, pat->ec, pat->bcr, pat->bslot, pat->rer, pat->rec, pat->ridx})
            (void)hipFree(d);
        delete pat;
    }
    return rc;
}

int hymor_pattern_destroy(hymor_pattern *pat) {
    return guarded([&] {
        if (!pat) return;
        for (int32_t *d : {pat->er, pat->ec, pat->bcr, pat->bslot, pat->rer, pat->rec, pat->ridx})
            if (d) HIP_TRY(hipFree(d));
        delete pat;
    });
}

int hymor_pattern_nnz(const hymor_pattern *pat, int64_t *nnz) {
    return guarded([&] {
        ARG_CHECK(pat && nnz, "null argument");
        *nnz = pat->nnz;
    });
}

int hymor_expm_taylor_vjp_values_dev(hymor_op *op, const hymor_pattern *pat, int order,
                                     int64_t nsub, double dt, int64_t ckpt_every,
                                     const double *d_x, const double *d_wbar, double *d_grad) {
    return guarded([&] {
        check_vjp_args(op, order, nsub, ckpt_every);
        ARG_CHECK(pat != nullptr, "null pattern");
        ARG_CHECK(pat->n == op->n && pat->b == op->b && pat->ncell == op->ncell,
                  "the pattern was made for an operator of another shape");
        ARG_CHECK(d_x && d_wbar && (d_grad || pat->nnz == 0), "null argument");
        vjp_core(op, order, nsub, dt, ckpt_every, d_x, d_wbar, pat->er, pat->ec, pat->nnz,
                 d_grad, pat);
    });
}

static void sandwich_core(hymor_op *A, hymor_op *C, int order, int64_t nsub, double dt,
                          int64_t n_set0, const double *x_vm, double *y_vm, int nv) {
    launch_perm_in(A, A->stream, x_vm, A->n, A->ws[1], 1, nv, n_set0, nv);
    // EXPONENTIAL!'s polynomial (CONSTRUCT_C.jl:84-105), evaluated by
    // Horner's rule: three vector passes per term instead of four
    const double *r1 = run_taylor(A, false, order, nsub, dt, HYMOR_FORM_HORNER, nv, A->ws[1]);
    // weight: ws[6] = C * r1 (C's extras use C's own partial-sum buffer)
    launch_term(C, false, nv, A->stream, r1, A->ws[6], nullptr, nullptr, 1.0, 0.0);
    const double *r2 = run_taylor(A, true, order, nsub, dt, HYMOR_FORM_HORNER, nv, A->ws[6]);
    launch_perm_out(A, A->stream, r2, y_vm, A->n, 1, nv, n_set0, nv);
}

static void check_sandwich(hymor_op *A, hymor_op *C, int order, int64_t nsub, int64_t n_set0) {
    check_taylor_args(A, 1, order, nsub, HYMOR_FORM_NORMALISED);
    ARG_CHECK(C != nullptr, "null weight operator");
    ARG_CHECK(C->n == A->n && C->b == A->b && C->ncell == A->ncell,
              "weight operator shape differs from A");
    ARG_CHECK(n_set0 >= 0 && n_set0 <= A->n, "n_set0 out of range");
}

int hymor_chu_sandwich_dev(hymor_op *A, hymor_op *C, int order, int64_t nsub, double dt,
                           int64_t n_set0, const void *d_x, void *d_y, int dtype) {
    return guarded([&] {
        check_sandwich(A, C, order, nsub, n_set0);
        ARG_CHECK(d_x && d_y, "null vector");
        const int nv = nv_of(A, dtype);
        ARG_CHECK(!(C->is_complex && nv == 1), "complex weight needs complex vectors");
        ensure_ws(A, nv);
        ensure_ws(C, nv);
        sandwich_core(A, C, order, nsub, dt, n_set0, (const double *)d_x, (double *)d_y, nv);
    });
}

int hymor_chu_sandwich_batched_dev(hymor_op *A, hymor_op *C, hymor_op *P_in, hymor_op *P_out,
                                   int order, int64_t nsub, double dt, int64_t n_set0,
                                   int64_t nrhs, const void *d_X, int64_t ldx, void *d_Y,
                                   int64_t ldy, int dtype) {
    return guarded([&] {
        check_sandwich(A, C, order, nsub, n_set0);
        ARG_CHECK(d_X && d_Y, "null matrix");
        check_batch(A, nrhs, ldx, ldy);
        const int ncomp = nv_of(A, dtype);
        ARG_CHECK(!(C->is_complex && ncomp == 1), "complex weight needs complex vectors");
        const int w = pad_width(nrhs * ncomp);
        hymor_op *ops[3] = {C, P_in, P_out};
        for (hymor_op *o : ops) {
            if (!o) continue;
            ARG_CHECK(o->n == A->n && o->b == A->b && o->ncell == A->ncell,
                      "weight or factor operator shape differs from A");
            ARG_CHECK(!(o->is_complex && ncomp == 1), "complex operator needs complex vectors");
            ARG_CHECK(tile_b_of(o, w) == tile_b_of(A, w),
                      "weight or factor operator uses another batch row layout than A");
        }
        ensure_ws(A, w);
        for (hymor_op *o : ops)
            if (o) ensure_ws(o, w);
        hipStream_t st = A->stream;
        const int tile_b = tile_b_of(A, w);
        auto zero_prefix = [&](double *buf) {
            if (n_set0 == 0) return;
            const int64_t total = A->n * w;
            const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
            hipLaunchKernelGGL(k_zero_vm_prefix, dim3((unsigned)grid), dim3(256), 0, st, buf, A->n,
                               A->ncell, A->b, n_set0, w, tile_b);
            HIP_TRY(hipGetLastError());
        };
        // x -> P_in x, zeroed (ws[1] -> ws[0]; the zeroing rides on the
        // permutation when there is no P_in)
        launch_perm_in(A, st, (const double *)d_X, ldx, A->ws[1], (int)nrhs, ncomp,
                       P_in ? 0 : n_set0, w);
        const double *in = A->ws[1];
        if (P_in) {
            launch_term(P_in, false, w, st, A->ws[1], A->ws[0], nullptr, nullptr, 1.0, 0.0);
            zero_prefix(A->ws[0]);
            in = A->ws[0];
        }
        const double *r1 = run_taylor(A, false, order, nsub, dt, HYMOR_FORM_HORNER, w, in);
        launch_term(C, false, w, st, r1, A->ws[6], nullptr, nullptr, 1.0, 0.0);
        const double *r2 = run_taylor(A, true, order, nsub, dt, HYMOR_FORM_HORNER, w, A->ws[6]);
        if (P_out) {
            // r2 is one of A's own buffers (ws[2..6]): zeroed in place
            zero_prefix(const_cast<double *>(r2));
            launch_term(P_out, false, w, st, r2, A->ws[1], nullptr, nullptr, 1.0, 0.0);
            r2 = A->ws[1];
        }
        launch_perm_out(A, st, r2, (double *)d_Y, ldy, (int)nrhs, ncomp, P_out ? 0 : n_set0, w);
    });
}

int hymor_chu_sandwich(hymor_op *A, hymor_op *C, int order, int64_t nsub, double dt,
                       int64_t n_set0, const void *x, void *y, int dtype) {
    return guarded([&] {
        check_sandwich(A, C, order, nsub, n_set0);
        ARG_CHECK(x && y, "null vector");
        const int nv = nv_of(A, dtype);
        ARG_CHECK(!(C->is_complex && nv == 1), "complex weight needs complex vectors");
        ensure_ws(A, nv);
        ensure_ws(C, nv);
        const size_t bytes = (size_t)A->n * nv * sizeof(double);
        HIP_TRY(hipMemcpyAsync(A->ws[0], x, bytes, hipMemcpyHostToDevice, A->stream));
        sandwich_core(A, C, order, nsub, dt, n_set0, A->ws[0], A->ws[0], nv);
        HIP_TRY(hipMemcpyAsync(y, A->ws[0], bytes, hipMemcpyDeviceToHost, A->stream));
        HIP_TRY(hipStreamSynchronize(A->stream));
    });
}

int hymor_op_synchronize(hymor_op *op) {
    return guarded([&] {
        ARG_CHECK(op != nullptr, "null operator");
        HIP_TRY(hipStreamSynchronize(op->stream));
        HIP_TRY(hipGetLastError());
    });
}

int hymor_dev_alloc(int64_t bytes, void **d_ptr) {
    return guarded([&] {
        ARG_CHECK(d_ptr && bytes >= 0, "bad argument");
        *d_ptr = nullptr;
        HIP_TRY(hipMalloc(d_ptr, bytes > 0 ? (size_t)bytes : 256));
    });
}

int hymor_dev_free(void *d_ptr) {
    return guarded([&] {
        if (d_ptr) HIP_TRY(hipFree(d_ptr));
    });
}

int hymor_dev_fill_uniform(void *d_x, int64_t len, uint64_t seed) {
    return guarded([&] {
        ARG_CHECK(d_x && len >= 0, "bad argument");
        if (len == 0) return;
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((len + 255) / 256, 16384));
        hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)grid), dim3(256), 0, 0, (double *)d_x,
                           len, seed);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    });
}

int hymor_memcpy_h2d(void *d_dst, const void *h_src, int64_t bytes) {
    return guarded([&] {
        ARG_CHECK(d_dst && h_src && bytes >= 0, "bad argument");
        HIP_TRY(hipMemcpy(d_dst, h_src, (size_t)bytes, hipMemcpyHostToDevice));
    });
}

int hymor_memcpy_d2d(void *d_dst, const void *d_src, int64_t bytes) {
    return guarded([&] {
        ARG_CHECK(d_dst && d_src && bytes >= 0, "bad argument");
        HIP_TRY(hipMemcpy(d_dst, d_src, (size_t)bytes, hipMemcpyDeviceToDevice));
    });
}

int hymor_memcpy_d2h(void *h_dst, const void *d_src, int64_t bytes) {
    return guarded([&] {
        ARG_CHECK(h_dst && d_src && bytes >= 0, "bad argument");
        HIP_TRY(hipMemcpy(h_dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost));
    });
}

int hymor_op_autotune(hymor_op *op, int nv, int reps, double *best_ms) {
    return guarded([&] {
        ARG_CHECK(op && best_ms, "null argument");
        ARG_CHECK(nv == 1 || nv == 2 || nv == 4, "autotune covers the k_term widths 1, 2, 4");
        ARG_CHECK(!(op->is_complex && nv == 1), "complex operator needs nv >= 2");
        ARG_CHECK(reps >= 1, "reps must be >= 1");
        ensure_ws(op, nv);
        const TermTuning cands[] = {{0, 4, 0, 1}, {0, 4, 1, 1}, {0, 2, 0, 1}, {1, 2, 0, 1},
                                    {1, 2, 1, 1}, {1, 4, 0, 1}, {1, 4, 1, 1}, {2, 3, 0, 1},
                                    {2, 3, 1, 1}, {2, 9, 0, 1}, {2, 9, 1, 1}, {3, 3, 0, 1},
                                    {3, 3, 1, 1}, {3, 9, 1, 1}, {3, 1, 0, 1}, {3, 1, 1, 1},
                                    {3, 2, 1, 1}};
        const int ncand = nv >= 2 ? 17 : 11;   // k_term_lds needs two doubles per row
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        double best = 1e300;
        TermTuning pick = op->fwd.tune;
        const int batch_kernel = op->fwd.tune.spmm_mfma;   // not this autotuner's choice
        for (int ci = 0; ci < ncand; ++ci) {
            TermTuning c = cands[ci];
            c.spmm_mfma = batch_kernel;
            op->fwd.tune = c;
            launch_term(op, false, nv, op->stream, op->ws[4], op->ws[5], op->ws[2], op->ws[2], 0.5,
                        1.0);
            HIP_TRY(hipEventRecord(e0, op->stream));
            for (int i = 0; i < reps; ++i)
                launch_term(op, false, nv, op->stream, op->ws[4], op->ws[5], op->ws[2], op->ws[2],
                            0.5, 1.0);
            HIP_TRY(hipEventRecord(e1, op->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            if (ms / reps < best) {
                best = ms / reps;
                pick = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        op->fwd.tune = pick;
        op->adj.tune = pick;
        drop_graph(op);
        *best_ms = best;
    });
}

int hymor_op_set_graphs(hymor_op *op, int enable) {
    return guarded([&] {
        ARG_CHECK(op, "null operator");
        HIP_TRY(hipStreamSynchronize(op->stream));
        op->use_graphs = enable != 0;
        if (!op->use_graphs) drop_graph(op);
    });
}

int hymor_op_set_batch_kernel(hymor_op *op, int use_mfma) {
    return guarded([&] {
        ARG_CHECK(op, "null operator");
        ARG_CHECK(use_mfma >= 0 && use_mfma <= 8, "batch kernel must be 0..8");
        op->fwd.tune.spmm_mfma = use_mfma;
        op->adj.tune.spmm_mfma = op->fwd.tune.spmm_mfma;
        drop_graph(op);
    });
}

int hymor_op_autotune_batch(hymor_op *op, int nv, int reps, int *best_kernel, double *best_ms) {
    return guarded([&] {
        ARG_CHECK(op && best_kernel && best_ms, "null argument");
        ARG_CHECK(nv == 8 || nv == 16 || nv == 32 || nv == 64,
                  "batch autotune covers widths 8, 16, 32, 64");
        ARG_CHECK(reps >= 1, "reps must be >= 1");
        ensure_ws(op, nv);
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        double best = 1e300;
        int pick = op->fwd.tune.spmm_mfma;
        for (int kern = 0; kern <= 8; ++kern) {
            op->fwd.tune.spmm_mfma = kern;
            if (kern >= 6) ensure_tile(op);
            launch_term(op, false, nv, op->stream, op->ws[4], op->ws[5], op->ws[2], op->ws[2], 0.5,
                        1.0);
            HIP_TRY(hipEventRecord(e0, op->stream));
            for (int i = 0; i < reps; ++i)
                launch_term(op, false, nv, op->stream, op->ws[4], op->ws[5], op->ws[2], op->ws[2],
                            0.5, 1.0);
            HIP_TRY(hipEventRecord(e1, op->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            if (ms / reps < best) {
                best = ms / reps;
                pick = kern;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        op->fwd.tune.spmm_mfma = pick;
        op->adj.tune.spmm_mfma = pick;
        drop_graph(op);
        *best_kernel = pick;
        *best_ms = best;
    });
}

int hymor_op_get_tuning(const hymor_op *op, int *split, int *waves, int *nt) {
    return guarded([&] {
        ARG_CHECK(op && split && waves && nt, "null argument");
        *split = op->fwd.tune.split;
        *waves = op->fwd.tune.waves;
        *nt = op->fwd.tune.nt;
    });
}

int hymor_op_set_tuning(hymor_op *op, int split, int waves, int nt) {
    return guarded([&] {
        ARG_CHECK(op, "null operator");
        ARG_CHECK((split == 0 && (waves == 1 || waves == 2 || waves == 4)) ||
                      (split == 1 && (waves == 2 || waves == 4)) ||
                      (split == 2 && (waves == 3 || waves == 9)) ||
                      (split == 3 && (waves == 1 || waves == 2 || waves == 3 || waves == 9)),
                  "k_term takes 1, 2 or 4 waves per group, k_term_split 2 or 4 waves per "
                  "slice, k_term_split_p 3 or 9, k_term_lds 1, 2, 3 or 9");
        ARG_CHECK(nt == 0 || nt == 1, "nt must be 0 or 1");
        op->fwd.tune.split = split;
        op->fwd.tune.waves = waves;
        op->fwd.tune.nt = nt;
        op->adj.tune = op->fwd.tune;
        drop_graph(op);
    });
}

int hymor_bench_term_kernel(hymor_op *op, int transpose, int nv, int form, int reps,
                            double *ms_per_launch, double *bytes_per_launch) {
    return guarded([&] {
        ARG_CHECK(op && ms_per_launch && bytes_per_launch, "null argument");
        ARG_CHECK(form == HYMOR_FORM_NORMALISED || form == HYMOR_FORM_POWERS ||
                      form == HYMOR_FORM_HORNER,
                  "unknown Taylor form");
        const bool horner = form == HYMOR_FORM_HORNER;
        ARG_CHECK(nv >= 1 && nv <= 64 && (nv & (nv - 1)) == 0, "nv must be a power of two <= 64");
        ARG_CHECK(!(op->is_complex && nv == 1), "complex operator needs nv >= 2");
        ARG_CHECK(reps >= 1, "reps must be >= 1");
        ARG_CHECK(!transpose || op->has_adj, "operator was built without its adjoint");
        ensure_ws(op, nv);
        const DevSliced &s = transpose ? op->adj : op->fwd;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        // a middle term of the form: powers / normalised gather x, write the
        // term and update vout in place; Horner gathers w, reads v, writes
        // the new w (in one launch with the replicated rows fused in where a
        // sweep does so: the sums of w formed once, into slot 0).  One warm
        // launch, then the timed ones
        const bool fused = horner && !transpose && fused_kind(op, false, nv) != 0;
        if (fused) launch_fused_prologue(op, op->stream, op->ws[4], 0, nv);
        auto term = [&] {
            if (fused)
                launch_term_fused(op, false, op->stream, op->ws[4], nullptr, op->ws[5],
                                  op->ws[2], 0.5, 1.0, 0, nv);
            else if (horner)
                launch_term(op, transpose != 0, nv, op->stream, op->ws[4], nullptr, op->ws[5],
                            op->ws[2], 0.5, 1.0);
            else
                launch_term(op, transpose != 0, nv, op->stream, op->ws[4], op->ws[5], op->ws[2],
                            op->ws[2], 0.5, 1.0);
        };
        term();
        HIP_TRY(hipEventRecord(e0, op->stream));
        for (int i = 0; i < reps; ++i) term();
        HIP_TRY(hipEventRecord(e1, op->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_launch = (double)ms / reps;
        const double sval = op->prec == HYMOR_PREC_F32 ? 4.0 : 8.0;
        const double bb = (double)s.b * s.b;
        // blocks (values + int32 column) + block-row pointers + x read,
        // term written, vout read and written (Horner: w read, v read, new w
        // written)
        double bytes = (double)s.nnzb * (bb * sval + 4.0) + (double)(s.ncell + 1) * 4.0 +
                       (horner ? 3.0 : 4.0) * (double)op->n * nv * 8.0;
        // (the fused terms read the compressed entries)
        const bool fused_adj = transpose && fused_kind(op, true, nv) != 0;
        const DevExtra &xe = fused ? op->xf_rep
                                   : (fused_adj ? op->xa_rep : extra_of(op, transpose != 0, nv));
        const double vsz = xe.is_complex ? 16.0 : 8.0;
        bytes += (double)xe.nnz * (vsz + 4.0);
        // replicated blocks: their diagonal, and the shared rows (kind 2)
        if (xe.rep.kind) bytes += (double)xe.rep.nrep * xe.rep.p * vsz;
        if (xe.rep.kind == 2) bytes += (double)xe.rep_nnz * (vsz + 4.0);
        *bytes_per_launch = bytes;
    });
}

}  // extern "C"

// ------------------------------------------------------------ energy tracking

namespace {

void launch_quad(const hymor_weights *W, hipStream_t st, const double *v, int nv, double *out,
                 int64_t ostride) {
    const int64_t items = W->ncell + W->n_extra;
    const int64_t nblk = (items + 255) / 256;
    if (nblk != W->nblocks) throw ArgError("weights partial buffer has the wrong size");
#define HYMOR_QUAD(BV, NVV)                                                                  \
    hipLaunchKernelGGL((k_quad_blockdiag<BV, NVV, kMaxWeights>), dim3((unsigned)nblk),       \
                       dim3(256), 0, st, v, W->blocks, W->xdiag, W->nmat, W->ncell,          \
                       W->n_extra, W->partial)
    if (W->b == 4 && nv == 1) HYMOR_QUAD(4, 1);
    else if (W->b == 4 && nv == 2) HYMOR_QUAD(4, 2);
    else if (W->b == 8 && nv == 1) HYMOR_QUAD(8, 1);
    else if (W->b == 8 && nv == 2) HYMOR_QUAD(8, 2);
    else throw ArgError("unsupported weights block size / vector width");
#undef HYMOR_QUAD
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials, dim3(W->nmat), dim3(256), 0, st, W->partial, nblk, out,
                       ostride);
    HIP_TRY(hipGetLastError());
}

}  // namespace

extern "C" {

int hymor_bench_copy(int64_t bytes, int reps, double *ms_per_launch) {
    return guarded([&] {
        ARG_CHECK(ms_per_launch && reps >= 1 && bytes >= 16 && bytes % 16 == 0,
                  "bytes must be a positive multiple of 16, reps >= 1");
        std::vector<void *> allocs;
        struct Free {
            std::vector<void *> &a;
            ~Free() {
                for (void *p : a) (void)hipFree(p);
            }
        } guard{allocs};
        d2 *in = (d2 *)dalloc_raw(allocs, (size_t)bytes);
        d2 *out = (d2 *)dalloc_raw(allocs, (size_t)bytes);
        HIP_TRY(hipMemset(in, 0, (size_t)bytes));
        const int64_t n2 = bytes / 16;
        const int64_t grid = std::min<int64_t>((n2 + 255) / 256, 4096);
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, nullptr, in, out, n2);
        HIP_TRY(hipEventRecord(e0, nullptr));
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(256), 0, nullptr, in, out, n2);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e1, nullptr));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_launch = (double)ms / reps;
    });
}

int hymor_bench_read(int64_t bytes, int reps, double *ms_per_launch) {
    return guarded([&] {
        ARG_CHECK(ms_per_launch && reps >= 1 && bytes >= 16 && bytes % 16 == 0,
                  "bytes must be a positive multiple of 16, reps >= 1");
        std::vector<void *> allocs;
        struct Free {
            std::vector<void *> &a;
            ~Free() {
                for (void *p : a) (void)hipFree(p);
            }
        } guard{allocs};
        const int64_t n2 = bytes / 16;
        const int64_t grid = std::min<int64_t>((n2 + 255) / 256, 2048);
        d2 *in = (d2 *)dalloc_raw(allocs, (size_t)bytes);
        double *out = (double *)dalloc_raw(allocs, (size_t)grid * 256 * sizeof(double));
        HIP_TRY(hipMemset(in, 0, (size_t)bytes));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        const int64_t chunk = (n2 + grid - 1) / grid;
        hipLaunchKernelGGL(k_read, dim3((unsigned)grid), dim3(256), 0, nullptr, in, out, n2,
                           chunk);
        HIP_TRY(hipEventRecord(e0, nullptr));
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(k_read, dim3((unsigned)grid), dim3(256), 0, nullptr, in, out, n2,
                               chunk);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e1, nullptr));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_launch = (double)ms / reps;
    });
}

int hymor_bench_stream_mix(int64_t read_bytes, int64_t write_bytes, int reps,
                           double *ms_per_launch, double *bytes_moved) {
    return guarded([&] {
        ARG_CHECK(ms_per_launch && bytes_moved && reps >= 1, "null argument or reps < 1");
        // chunks of whole 1 KiB runs (64 items), so no cache line is shared
        // between two chunks
        constexpr int64_t kChunks = 16384, kRun = 64;
        ARG_CHECK(read_bytes >= kChunks * kRun * 16 && write_bytes >= 0,
                  "read_bytes must be >= 16 MiB, write_bytes >= 0");
        const int64_t rin = read_bytes / 16 / kChunks / kRun * kRun;
        const int64_t rout = write_bytes / 16 / kChunks / kRun * kRun;
        std::vector<void *> allocs;
        struct Free {
            std::vector<void *> &a;
            ~Free() {
                for (void *p : a) (void)hipFree(p);
            }
        } guard{allocs};
        d2 *in = (d2 *)dalloc_raw(allocs, (size_t)(kChunks * rin * 16));
        d2 *out = rout ? (d2 *)dalloc_raw(allocs, (size_t)(kChunks * rout * 16)) : nullptr;
        HIP_TRY(hipMemset(in, 0, (size_t)(kChunks * rin * 16)));
        const int64_t grid = 2048;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_stream_mix, dim3((unsigned)grid), dim3(256), 0, nullptr, in, rin,
                           out, rout, kChunks);
        HIP_TRY(hipEventRecord(e0, nullptr));
        for (int i = 0; i < reps; ++i)
            hipLaunchKernelGGL(k_stream_mix, dim3((unsigned)grid), dim3(256), 0, nullptr, in,
                               rin, out, rout, kChunks);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e1, nullptr));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_per_launch = (double)ms / reps;
        *bytes_moved = (double)(kChunks * (rin + rout) * 16);
    });
}

int hymor_weights_create_csc(int64_t n, int nmat, const int64_t *const *colptr,
                             const int64_t *const *rowval, const double *const *nzval,
                             int index_base, int vars_per_cell, int64_t nchi, int64_t neta,
                             hymor_weights **out) {
    hymor_weights *W = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "out is null");
        *out = nullptr;
        ARG_CHECK(nmat >= 1 && nmat <= kMaxWeights, "nmat must be in [1, 8]");
        ARG_CHECK(colptr && rowval && nzval, "null CSC array list");
        ARG_CHECK(index_base == 0 || index_base == 1, "index_base must be 0 or 1");
        ARG_CHECK(vars_per_cell == 4 || vars_per_cell == 8, "vars_per_cell must be 4 or 8");
        ARG_CHECK(nchi > 0 && neta > 0, "grid must be non-empty");
        const int b = vars_per_cell;
        const int64_t ncell = nchi * neta, nflow = (int64_t)b * ncell;
        ARG_CHECK(n >= nflow, "n smaller than vars_per_cell*nchi*neta");
        const int64_t n_extra = n - nflow;
        std::vector<double> blocks((size_t)nmat * ncell * b * b, 0.0);
        std::vector<double> xdiag((size_t)nmat * std::max<int64_t>(n_extra, 1), 0.0);
        for (int m = 0; m < nmat; ++m) {
            ARG_CHECK(colptr[m] && rowval[m] && nzval[m], "null CSC array");
            ARG_CHECK(colptr[m][0] == index_base, "colptr[0] must equal the index base");
            for (int64_t j = 0; j < n; ++j) {
                const int64_t cmc = j < nflow ? (j % ncell) * b + j / ncell : j;
                for (int64_t e = colptr[m][j] - index_base; e < colptr[m][j + 1] - index_base; ++e) {
                    const int64_t i = rowval[m][e] - index_base;
                    ARG_CHECK(i >= 0 && i < n, "rowval out of range");
                    const int64_t cmr = i < nflow ? (i % ncell) * b + i / ncell : i;
                    const double v = nzval[m][e];
                    if (cmr < nflow && cmc < nflow) {
                        ARG_CHECK(cmr / b == cmc / b,
                                  "weight has an entry outside the per-cell block diagonal");
                        blocks[(((size_t)m * ncell + cmr / b) * b + cmr % b) * b + cmc % b] += v;
                    } else {
                        ARG_CHECK(cmr == cmc || v == 0.0,
                                  "weight has an off-diagonal entry on the extra DOFs");
                        if (cmr == cmc) xdiag[(size_t)m * n_extra + (cmr - nflow)] += v;
                    }
                }
            }
        }
        W = new hymor_weights();
        W->b = b;
        W->nmat = nmat;
        W->nchi = nchi;
        W->neta = neta;
        W->ncell = ncell;
        W->n = n;
        W->n_extra = n_extra;
        W->nblocks = (ncell + n_extra + 255) / 256;
        W->blocks = (double *)dalloc_raw(W->allocs, blocks.size() * sizeof(double));
        HIP_TRY(hipMemcpy(W->blocks, blocks.data(), blocks.size() * sizeof(double),
                          hipMemcpyHostToDevice));
        W->xdiag = (double *)dalloc_raw(W->allocs, xdiag.size() * sizeof(double));
        HIP_TRY(hipMemcpy(W->xdiag, xdiag.data(), xdiag.size() * sizeof(double),
                          hipMemcpyHostToDevice));
        W->partial = (double *)dalloc_raw(W->allocs, (size_t)nmat * W->nblocks * sizeof(double));
        *out = W;
    });
    if (rc != HYMOR_OK && W) {
        for (void *p : W->allocs) (void)hipFree(p);
        delete W;
    }
    return rc;
}

int hymor_weights_destroy(hymor_weights *W) {
    return guarded([&] {
        if (!W) return;
        HIP_TRY(hipDeviceSynchronize());
        for (void *p : W->allocs) (void)hipFree(p);
        delete W;
    });
}

int hymor_energy_track(hymor_op *A, const hymor_weights *W, int order, int64_t n_t, double dt,
                       int form, const void *x, int dtype, double *E, void *x_final,
                       int64_t n_skip, int64_t n_samples, void *samples) {
    return guarded([&] {
        check_taylor_args(A, 0, order, n_t, form);
        ARG_CHECK(W != nullptr && x && E, "null argument");
        ARG_CHECK(W->n == A->n && W->b == A->b && W->ncell == A->ncell,
                  "weights shape differs from the operator");
        ARG_CHECK(n_samples == 0 || (samples != nullptr && n_skip >= 1),
                  "sampling needs a buffer and n_skip >= 1");
        const int nv = nv_of(A, dtype);
        ensure_ws(A, nv);
        const int64_t ncol = n_t + 1;
        // per-call series buffer, released when this call returns
        double *dE = (double *)dalloc_raw(A->allocs, (size_t)W->nmat * ncol * sizeof(double));
        struct Free {
            hymor_op *op;
            double *p;
            ~Free() {
                for (size_t k = 0; k < op->allocs.size(); ++k)
                    if (op->allocs[k] == p) {
                        (void)hipFree(p);
                        op->allocs.erase(op->allocs.begin() + (long)k);
                        break;
                    }
            }
        } guard{A, dE};
        const size_t bytes = (size_t)A->n * nv * sizeof(double);
        hipStream_t st = A->stream;
        HIP_TRY(hipMemcpyAsync(A->ws[0], x, bytes, hipMemcpyHostToDevice, st));
        launch_perm_in(A, st, A->ws[0], A->n, A->ws[1], 1, nv, 0, nv);
        const double *cur = A->ws[1];
        launch_quad(W, st, cur, nv, dE, ncol);
        int64_t count = 0;
        for (int64_t i = 1; i <= n_t; ++i) {
            cur = run_taylor(A, false, order, 1, dt, form, nv, cur);
            launch_quad(W, st, cur, nv, dE + i, ncol);
            if (n_samples > 0 && i % n_skip == 0 && count < n_samples) {
                launch_perm_out(A, st, cur, A->ws[0], A->n, 1, nv, 0, nv);
                HIP_TRY(hipMemcpyAsync((char *)samples + (size_t)count * bytes, A->ws[0], bytes,
                                       hipMemcpyDeviceToHost, st));
                ++count;
            }
        }
        if (x_final) {
            launch_perm_out(A, st, cur, A->ws[0], A->n, 1, nv, 0, nv);
            HIP_TRY(hipMemcpyAsync(x_final, A->ws[0], bytes, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipMemcpyAsync(E, dE, (size_t)W->nmat * ncol * sizeof(double),
                               hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    });
}

}  // extern "C"

// ------------------------------------------------------------ Krylov basis

namespace {

// h[0 .. NC*m) <- V[:, 0:m]^H w  (device), deterministic reduction
void basis_dots(hymor_krylov *K, const double *V, int64_t ld, int m, const double *w,
                double *h) {
    const bool cplx = K->ncomp == 2;
    for (int j0 = 0; j0 < m; j0 += kDotTile) {
        if (cplx)
            hipLaunchKernelGGL((k_dots<true, kDotTile>), dim3(kDotBlocks), dim3(256), 0, K->stream,
                               V, ld, m, j0, w, K->n, K->partial);
        else
            hipLaunchKernelGGL((k_dots<false, kDotTile>), dim3(kDotBlocks), dim3(256), 0,
                               K->stream, V, ld, m, j0, w, K->n, K->partial);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_reduce_partials, dim3(m * K->ncomp), dim3(256), 0, K->stream, K->partial,
                       (int64_t)kDotBlocks, h, (int64_t)1);
    HIP_TRY(hipGetLastError());
}

void basis_sub(hymor_krylov *K, const double *V, int64_t ld, int m, const double *h, double *w) {
    const int64_t grid = std::min<int64_t>((K->n + 255) / 256, 4096);
    if (K->ncomp == 2)
        hipLaunchKernelGGL(k_sub_vh<true>, dim3((unsigned)grid), dim3(256), 0, K->stream, V, ld, m,
                           h, w, K->n);
    else
        hipLaunchKernelGGL(k_sub_vh<false>, dim3((unsigned)grid), dim3(256), 0, K->stream, V, ld,
                           m, h, w, K->n);
    HIP_TRY(hipGetLastError());
}

double *col(hymor_krylov *K, double *base, int j) {
    return base + (size_t)j * K->n * K->ncomp;
}

// A vector left with norm nrm after projecting out coefficients of squared
// sum hsq had norm sqrt(hsq + nrm^2) before (Pythagoras over an orthonormal
// basis).  It counts as dependent on the basis when projection removed all
// but a rounding-level part of it: normalising that remainder would add a
// direction made of rounding errors.
constexpr double kDepTol = 64.0 * 2.220446049250313e-16;

bool is_dependent(double nrm, double hsq) {
    return !(nrm > kDepTol * std::sqrt(hsq + nrm * nrm));
}

void launch_zero(hymor_krylov *K, double *p, int64_t len) {
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((len + 255) / 256, 4096));
    hipLaunchKernelGGL(k_zero, dim3((unsigned)grid), dim3(256), 0, K->stream, p, len);
    HIP_TRY(hipGetLastError());
}

constexpr int kVtwGroups = 1024;   // row chunks of the block projection kernel

// Row split of the V^T W partial products: at most kVtwGroups workgroups of
// a multiple of 64 rows each (16 per wave)
struct VtwShape {
    int ctiles;
    int64_t rows_per, groups;
};

VtwShape vtw_shape(const hymor_krylov *K, int ncol) {
    VtwShape v;
    v.ctiles = (ncol + 15) / 16;
    v.rows_per = ((K->n + kVtwGroups - 1) / kVtwGroups + 63) / 64 * 64;
    v.groups = (K->n + v.rows_per - 1) / v.rows_per;
    return v;
}

// K->bpartial <- the per-workgroup partial products of V[:, :ncol]^T W
void launch_vtw(hymor_krylov *K, const VtwShape &v, int ncol, const double *W, int64_t ldw,
                int s) {
    const bool aligned = K->n % 2 == 0 && ldw % 2 == 0 &&
                         ((uintptr_t)K->V % 16) == 0 && ((uintptr_t)W % 16) == 0;
    auto go = [&](auto al, auto ct) {
        constexpr bool AL = decltype(al)::value;
        constexpr int CT = decltype(ct)::value;
        hipLaunchKernelGGL((k_vtw_mfma<AL, CT>), dim3((unsigned)v.groups), dim3(256), 0,
                           K->stream, K->V, K->n, ncol, W, ldw, s, K->n, v.rows_per, v.ctiles,
                           K->bpartial);
    };
    // two column tiles per pass up to 32 basis columns, four beyond
    using Two = std::integral_constant<int, 2>;
    using Four = std::integral_constant<int, 4>;
    if (aligned) {
        if (v.ctiles <= 2) go(std::true_type{}, Two{});
        else go(std::true_type{}, Four{});
    } else {
        if (v.ctiles <= 2) go(std::false_type{}, Two{});
        else go(std::false_type{}, Four{});
    }
    HIP_TRY(hipGetLastError());
}

// h (ncol x s, column-major) <- the sum of those partials, in a fixed order
void launch_reduce_vtw(hymor_krylov *K, const VtwShape &v, int ncol, int s, double *h) {
    hipLaunchKernelGGL(k_reduce_vtw, dim3((unsigned)((ncol * s + 3) / 4)), dim3(256), 0,
                       K->stream, K->bpartial, v.groups, v.ctiles, ncol, s, h);
    HIP_TRY(hipGetLastError());
}

// h (ncol x s, column-major) = V[:, :ncol]^T W, deterministic
void block_dots(hymor_krylov *K, int ncol, const double *W, int64_t ldw, int s, double *h) {
    const VtwShape v = vtw_shape(K, ncol);
    launch_vtw(K, v, ncol, W, ldw, s);
    launch_reduce_vtw(K, v, ncol, s, h);
}

// the block Gram-Schmidt buffers, made on first use
void ensure_block_bufs(hymor_krylov *K) {
    if (K->bpartial) return;
    const int ct = (K->m_max + 1 + 15) / 16;
    K->bpartial = (double *)dalloc_raw(K->allocs, (size_t)kVtwGroups * ct * 256 * sizeof(double));
    K->hb = (double *)dalloc_raw(K->allocs, (size_t)(K->m_max + 1) * 16 * sizeof(double));
    K->hb2 = (double *)dalloc_raw(K->allocs, (size_t)(K->m_max + 1) * 16 * sizeof(double));
}

// W[:, :s] -= V[:, :ncol] h
void block_sub(hymor_krylov *K, int ncol, const double *h, double *W, int64_t ldw, int s) {
    // one wave per 16-row tile, four waves per workgroup
    const int64_t tiles = (K->n + 15) / 16;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 4096));
    const size_t lds = (size_t)ncol * s * sizeof(double);
    hipLaunchKernelGGL(k_sub_vh_mfma<8>, dim3((unsigned)grid), dim3(256), lds, K->stream, K->V,
                       K->n, ncol, h, W, ldw, s, K->n);
    HIP_TRY(hipGetLastError());
}

}  // namespace

extern "C" {

int hymor_krylov_create(int64_t n, int m_max, int dtype, hymor_krylov **out) {
    hymor_krylov *K = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "out is null");
        *out = nullptr;
        ARG_CHECK(n >= 1 && m_max >= 1 && m_max <= 1024, "need n >= 1 and 1 <= m_max <= 1024");
        ARG_CHECK(dtype == HYMOR_F64 || dtype == HYMOR_C128, "bad dtype");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            (void)hipGetLastError();
            throw HipError("no HIP device available");
        }
        K = new hymor_krylov();
        K->n = n;
        K->m_max = m_max;
        K->ncomp = dtype == HYMOR_C128 ? 2 : 1;
        const size_t colb = (size_t)n * K->ncomp * sizeof(double);
        K->V = (double *)dalloc_raw(K->allocs, colb * (m_max + 1));
        K->tmp = (double *)dalloc_raw(K->allocs, colb * (m_max + 1));
        K->h = (double *)dalloc_raw(K->allocs, 2 * (m_max + 1) * sizeof(double));
        K->h2 = (double *)dalloc_raw(K->allocs, 2 * (m_max + 1) * sizeof(double));
        K->q = (double *)dalloc_raw(K->allocs, 2 * (size_t)(m_max + 1) * (m_max + 1) * sizeof(double));
        K->partial = (double *)dalloc_raw(K->allocs, 2 * (size_t)(m_max + 1) * kDotBlocks * sizeof(double));
        HIP_TRY(hipStreamCreateWithFlags(&K->stream, hipStreamNonBlocking));
        HIP_TRY(hipMemsetAsync(K->V, 0, colb * (m_max + 1), K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
        *out = K;
    });
    if (rc != HYMOR_OK && K) {
        for (void *p : K->allocs) (void)hipFree(p);
        if (K->stream) (void)hipStreamDestroy(K->stream);
        delete K;
    }
    return rc;
}

int hymor_krylov_destroy(hymor_krylov *K) {
    return guarded([&] {
        if (!K) return;
        HIP_TRY(hipStreamSynchronize(K->stream));
        for (void *p : K->allocs) (void)hipFree(p);
        (void)hipStreamDestroy(K->stream);
        delete K;
    });
}

int hymor_krylov_column(hymor_krylov *K, int j, void **d_col) {
    return guarded([&] {
        ARG_CHECK(K && d_col, "null argument");
        ARG_CHECK(j >= 0 && j <= K->m_max, "column out of range");
        *d_col = col(K, K->V, j);
    });
}

int hymor_krylov_set_column(hymor_krylov *K, int j, const void *x) {
    return guarded([&] {
        ARG_CHECK(K && x, "null argument");
        ARG_CHECK(j >= 0 && j <= K->m_max, "column out of range");
        HIP_TRY(hipMemcpyAsync(col(K, K->V, j), x, (size_t)K->n * K->ncomp * sizeof(double),
                               hipMemcpyHostToDevice, K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_get_columns(hymor_krylov *K, int j0, int ncols, void *out) {
    return guarded([&] {
        ARG_CHECK(K && out, "null argument");
        ARG_CHECK(j0 >= 0 && ncols >= 0 && j0 + ncols <= K->m_max + 1, "columns out of range");
        HIP_TRY(hipMemcpyAsync(out, col(K, K->V, j0),
                               (size_t)ncols * K->n * K->ncomp * sizeof(double),
                               hipMemcpyDeviceToHost, K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_orthonormalize(hymor_krylov *K, int j, void *d_w, double *h_out) {
    return guarded([&] {
        ARG_CHECK(K && d_w && h_out, "null argument");
        ARG_CHECK(j >= 0 && j <= K->m_max, "j out of range");
        double *w = (double *)d_w;
        const int nc = K->ncomp;
        const int64_t len = K->n * nc;
        // classical Gram-Schmidt, twice (CGS2)
        if (j > 0) {
            basis_dots(K, K->V, K->n, j, w, K->h);
            basis_sub(K, K->V, K->n, j, K->h, w);
            basis_dots(K, K->V, K->n, j, w, K->h2);
            basis_sub(K, K->V, K->n, j, K->h2, w);
        }
        basis_dots(K, w, K->n, 1, w, K->h2 + 2 * (size_t)K->m_max);   // |w|^2 (re part)
        std::vector<double> h1((size_t)nc * j + 2, 0.0), h2((size_t)nc * j + 2, 0.0);
        if (j > 0) {
            HIP_TRY(hipMemcpyAsync(h1.data(), K->h, (size_t)nc * j * sizeof(double),
                                   hipMemcpyDeviceToHost, K->stream));
            HIP_TRY(hipMemcpyAsync(h2.data(), K->h2, (size_t)nc * j * sizeof(double),
                                   hipMemcpyDeviceToHost, K->stream));
        }
        double nrm2 = 0.0;
        HIP_TRY(hipMemcpyAsync(&nrm2, K->h2 + 2 * (size_t)K->m_max, sizeof(double),
                               hipMemcpyDeviceToHost, K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
        double hsq = 0.0;
        for (int i = 0; i < nc * j; ++i) {
            h_out[i] = h1[i] + h2[i];
            hsq += h_out[i] * h_out[i];
        }
        const double nrm = std::sqrt(std::max(nrm2, 0.0));
        const bool dependent = is_dependent(nrm, hsq);
        h_out[nc * j] = dependent ? 0.0 : nrm;
        if (nc == 2) h_out[nc * j + 1] = 0.0;
        double *dst = col(K, K->V, j);
        if (!dependent) {
            const int64_t grid = std::min<int64_t>((len + 255) / 256, 4096);
            hipLaunchKernelGGL(k_scale, dim3((unsigned)grid), dim3(256), 0, K->stream, w, len,
                               1.0 / nrm);
            HIP_TRY(hipGetLastError());
            if (w != dst)
                HIP_TRY(hipMemcpyAsync(dst, w, (size_t)len * sizeof(double),
                                       hipMemcpyDeviceToDevice, K->stream));
        } else {
            // dependent vector: a zero column, reported by a zero norm
            launch_zero(K, dst, len);
        }
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_orthonormalize_block(hymor_krylov *K, int j, int s, void *d_W, int64_t ldw,
                                      double *h_out) {
    return guarded([&] {
        ARG_CHECK(K && d_W && h_out, "null argument");
        ARG_CHECK(K->ncomp == 1, "block orthonormalisation covers real bases");
        ARG_CHECK(s >= 1 && s <= 16, "block of 1..16 vectors");
        ARG_CHECK(j >= 0 && j + s <= K->m_max + 1, "columns j..j+s-1 must fit the basis");
        ARG_CHECK(ldw >= K->n, "leading dimension smaller than n");
        ARG_CHECK((size_t)j * s * sizeof(double) <= 65536,
                  "j * s too large for the projection update (at most 8192 coefficients)");
        double *W = (double *)d_W;
        const int64_t n = K->n;
        const int ldh = j + s;
        std::fill(h_out, h_out + (size_t)ldh * s, 0.0);
        if (j > 0) {
            ensure_block_bufs(K);
            // block CGS2 against the j accepted columns: one pass over V per
            // projection for all s vectors
            block_dots(K, j, W, ldw, s, K->hb);
            block_sub(K, j, K->hb, W, ldw, s);
            block_dots(K, j, W, ldw, s, K->hb2);
            block_sub(K, j, K->hb2, W, ldw, s);
            std::vector<double> a((size_t)j * s), b((size_t)j * s);
            HIP_TRY(hipMemcpyAsync(a.data(), K->hb, a.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, K->stream));
            HIP_TRY(hipMemcpyAsync(b.data(), K->hb2, b.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, K->stream));
            HIP_TRY(hipStreamSynchronize(K->stream));
            for (int i = 0; i < s; ++i)
                for (int c = 0; c < j; ++c)
                    h_out[(size_t)i * ldh + c] = a[(size_t)i * j + c] + b[(size_t)i * j + c];
        }
        // inside the block: CGS2 of vector i against the block's accepted
        // columns j..j+i-1, then normalisation into column j+i
        for (int i = 0; i < s; ++i) {
            double *w = W + (size_t)i * ldw;
            const double *Vb = col(K, K->V, j);
            std::vector<double> h1((size_t)i + 1, 0.0), h2((size_t)i + 1, 0.0);
            if (i > 0) {
                basis_dots(K, Vb, n, i, w, K->h);
                basis_sub(K, Vb, n, i, K->h, w);
                basis_dots(K, Vb, n, i, w, K->h2);
                basis_sub(K, Vb, n, i, K->h2, w);
                HIP_TRY(hipMemcpyAsync(h1.data(), K->h, (size_t)i * sizeof(double),
                                       hipMemcpyDeviceToHost, K->stream));
                HIP_TRY(hipMemcpyAsync(h2.data(), K->h2, (size_t)i * sizeof(double),
                                       hipMemcpyDeviceToHost, K->stream));
            }
            basis_dots(K, w, n, 1, w, K->h2 + 2 * (size_t)K->m_max);
            double nrm2 = 0.0;
            HIP_TRY(hipMemcpyAsync(&nrm2, K->h2 + 2 * (size_t)K->m_max, sizeof(double),
                                   hipMemcpyDeviceToHost, K->stream));
            HIP_TRY(hipStreamSynchronize(K->stream));
            for (int c = 0; c < i; ++c) h_out[(size_t)i * ldh + j + c] = h1[c] + h2[c];
            double hsq = 0.0;
            for (int c = 0; c < j + i; ++c)
                hsq += h_out[(size_t)i * ldh + c] * h_out[(size_t)i * ldh + c];
            const double nrm = std::sqrt(std::max(nrm2, 0.0));
            const bool dependent = is_dependent(nrm, hsq);
            h_out[(size_t)i * ldh + j + i] = dependent ? 0.0 : nrm;
            double *dst = col(K, K->V, j + i);
            if (!dependent) {
                const int64_t grid = std::min<int64_t>((n + 255) / 256, 4096);
                hipLaunchKernelGGL(k_scale, dim3((unsigned)grid), dim3(256), 0, K->stream, w, n,
                                   1.0 / nrm);
                HIP_TRY(hipGetLastError());
                if (w != dst)
                    HIP_TRY(hipMemcpyAsync(dst, w, (size_t)n * sizeof(double),
                                           hipMemcpyDeviceToDevice, K->stream));
            } else {
                // dependent vector: a zero column, reported by a zero norm
                launch_zero(K, dst, n);
            }
        }
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_combine(hymor_krylov *K, int m, int k, const double *Q) {
    return guarded([&] {
        ARG_CHECK(K && Q, "null argument");
        ARG_CHECK(m >= 1 && m <= K->m_max + 1 && k >= 1 && k <= m, "need 1 <= k <= m <= m_max+1");
        const int nc = K->ncomp;
        HIP_TRY(hipMemcpyAsync(K->q, Q, (size_t)nc * m * k * sizeof(double), hipMemcpyHostToDevice,
                               K->stream));
        const int64_t grid = std::min<int64_t>((K->n + 255) / 256, 4096);
        if (nc == 2)
            hipLaunchKernelGGL(k_basis_combine<true>, dim3((unsigned)grid), dim3(256), 0, K->stream,
                               K->V, K->n, m, K->q, k, K->tmp, K->n, K->n);
        else
            hipLaunchKernelGGL(k_basis_combine<false>, dim3((unsigned)grid), dim3(256), 0,
                               K->stream, K->V, K->n, m, K->q, k, K->tmp, K->n, K->n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(K->V, K->tmp, (size_t)k * K->n * nc * sizeof(double),
                               hipMemcpyDeviceToDevice, K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_ritz(hymor_krylov *K, int m, int k, const double *Q, void *out) {
    return guarded([&] {
        ARG_CHECK(K && Q && out, "null argument");
        ARG_CHECK(m >= 1 && m <= K->m_max + 1 && k >= 1 && k <= K->m_max + 1,
                  "need 1 <= m, k <= m_max+1");
        const int nc = K->ncomp;
        HIP_TRY(hipMemcpyAsync(K->q, Q, (size_t)nc * m * k * sizeof(double), hipMemcpyHostToDevice,
                               K->stream));
        const int64_t grid = std::min<int64_t>((K->n + 255) / 256, 4096);
        if (nc == 2)
            hipLaunchKernelGGL(k_basis_combine<true>, dim3((unsigned)grid), dim3(256), 0, K->stream,
                               K->V, K->n, m, K->q, k, K->tmp, K->n, K->n);
        else
            hipLaunchKernelGGL(k_basis_combine<false>, dim3((unsigned)grid), dim3(256), 0,
                               K->stream, K->V, K->n, m, K->q, k, K->tmp, K->n, K->n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(out, K->tmp, (size_t)k * K->n * nc * sizeof(double),
                               hipMemcpyDeviceToHost, K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_krylov_copy_column(hymor_krylov *K, int src, int dst) {
    return guarded([&] {
        ARG_CHECK(K, "null argument");
        ARG_CHECK(src >= 0 && src <= K->m_max && dst >= 0 && dst <= K->m_max, "column out of range");
        if (src == dst) return;
        HIP_TRY(hipMemcpyAsync(col(K, K->V, dst), col(K, K->V, src),
                               (size_t)K->n * K->ncomp * sizeof(double), hipMemcpyDeviceToDevice,
                               K->stream));
        HIP_TRY(hipStreamSynchronize(K->stream));
    });
}

int hymor_bench_krylov_projection(int64_t n, int j, int s, int reps, double *out) {
    hymor_krylov *K = nullptr;
    int rc = guarded([&] {
        ARG_CHECK(out != nullptr, "null argument");
        ARG_CHECK(n >= 16 && j >= 1 && j <= 1024 && s >= 1 && s <= 16 && reps >= 1,
                  "need n >= 16, 1 <= j <= 1024, 1 <= s <= 16, reps >= 1");
        ARG_CHECK((size_t)j * s * sizeof(double) <= 65536, "j * s too large");
        hymor_krylov **pk = &K;
        int e = hymor_krylov_create(n, j + s, HYMOR_F64, pk);
        if (e != HYMOR_OK) throw HipError(g_err);
        ensure_block_bufs(K);
        // basis columns 0..j-1 and the block W (columns j..j+s-1) filled on
        // the device; the values do not change what the kernels do
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n * (j + s) + 255) / 256,
                                                                     16384));
        hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)grid), dim3(256), 0, K->stream, K->V,
                           n * (j + s), (uint64_t)12345);
        HIP_TRY(hipGetLastError());
        double *W = col(K, K->V, j);
        const VtwShape v = vtw_shape(K, j);
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        auto timed = [&](auto &&launch) {
            launch();
            HIP_TRY(hipEventRecord(e0, K->stream));
            for (int i = 0; i < reps; ++i) launch();
            HIP_TRY(hipEventRecord(e1, K->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        // W -= V h with h = 0 leaves W unchanged, so repeated launches keep
        // the same data and the same work
        HIP_TRY(hipMemsetAsync(K->hb2, 0, (size_t)j * s * sizeof(double), K->stream));
        out[0] = timed([&] { launch_vtw(K, v, j, W, n, s); });
        out[1] = timed([&] { launch_reduce_vtw(K, v, j, s, K->hb); });
        out[2] = timed([&] { block_sub(K, j, K->hb2, W, n, s); });
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        out[3] = (double)(j + s) * n * 8.0;
        out[4] = 2.0 * n * j * s;
        out[5] = 512.0 * n * v.ctiles;
        out[6] = (double)(j + 2 * s) * n * 8.0;
        out[7] = 2.0 * n * j * s;
    });
    const std::string msg = g_err;   // destroying K must not clear a failure's message
    if (K) (void)hymor_krylov_destroy(K);
    g_err = msg;
    return rc;
}

}  // extern "C"
