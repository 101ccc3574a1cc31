// Internal structures of libhymor_hip.  Host-side only; kernels see raw
// pointers.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hymor {

// Host block-CSR over cells (0-based, cell-major).  Block k couples cell
// row(k) to cell col[k]; val holds b*b doubles per block, row-major.
struct HostBSR {
    int b = 0;
    int64_t ncell = 0;
    std::vector<int64_t> ptr;   // ncell+1
    std::vector<int32_t> col;   // nnzb
    std::vector<double> val;    // nnzb*b*b
    int64_t nnzb() const { return (int64_t)col.size(); }
};

// Host CSR of the "extra" entries: every entry that touches an extra DOF
// (index >= b*ncell) and, for complex operators, every imaginary part.
// Rows are cell-major global indices; only rows holding entries, plus every
// extra-DOF row, are listed.
struct HostExtra {
    bool is_complex = false;
    std::vector<int64_t> rows;  // listed rows, ascending
    std::vector<int64_t> ptr;   // rows.size()+1
    std::vector<int32_t> col;   // cell-major global column
    std::vector<double> val;    // 1 or 2 doubles per entry
};

// Sliced-ELL device layout of a BSR: slices of CPS = 64/b consecutive cells,
// one wavefront per slice, each lane one (cell, block-row) pair.  Inside a
// slice the blocks are slot-major: slot t of cell q sits at
// slice_ptr[s] + t*CPS + q, so one slot of a slice is a contiguous
// CPS*b*b-value segment.  Short rows are padded with zero blocks that point
// at the row's own cell.
struct HostSliced {
    int b = 0;
    int cps = 0;
    int64_t ncell = 0;
    int64_t nslice = 0;
    std::vector<int64_t> slice_ptr;  // nslice+1, in blocks
    std::vector<int32_t> col;        // padded blocks
    std::vector<double> val;         // padded blocks * b*b (see interleave_lanes)
    int64_t nnzb = 0;                // unpadded
};

// Replicated freestream blocks: the receptivity operator
//   L_ = [L  repeat(B,1,N);  0  kron(diag(-i w), I_p)]
// (LINEARIZE_L_.jl:58-67) repeats B's columns N times, and its adjoint
// repeats B''s rows N times.  Kept once, they are applied once per term:
//   kind 1 (forward): rows [c0, n) carry only their diagonal; every other
//     row's entries in columns [c0, n) repeat with period p over nrep
//     copies, so sum_k B[r, q] x[c0 + k p + q] = B[r, q] S[q] with
//     S[q] = sum_k x[c0 + k p + q].  Copy 0 stays, its columns renamed to
//     virtual columns n + q that read S; rows [c0, n) leave the listing
//     (their diagonal, in `diag`, is applied where S is formed).
//   kind 2 (adjoint): rows c0 + k p + q share their off-diagonal entries
//     across k: kept once as shared row q (vptr/vcol/vval), evaluated once
//     and finished into the nrep rows with their own diagonal (`diag`);
//     rows [c0, n) leave the listing.
struct HostReplica {
    int kind = 0;               // 0 none, 1 column replicas, 2 row replicas
    int64_t c0 = 0, p = 0, nrep = 0;
    std::vector<double> diag;   // nrep * p entries (1 or 2 doubles): rows c0 .. n-1
    std::vector<int64_t> vptr;  // kind 2: p + 1
    std::vector<int32_t> vcol;
    std::vector<double> vval;
};

// Finds and applies kind 1 on a forward extra CSR; false (x unchanged) if
// the operator has no such structure (nrep >= 2 required).
bool compress_column_replicas(HostExtra &x, int64_t n, HostReplica &rep);

// The compressed flow-row entries of a kind 1 forward operator regrouped by
// slice of 64 internal rows (a term kernel's workgroup), so that the term
// kernel can add them in its epilogue from column sums formed by the
// previous term: slice_tab[s] is the slice's block or -1; block k's lane l
// (row 64 s + l) owns entries [lane_ptr[65 k + l], lane_ptr[65 k + l + 1])
// of ecol (column q of the sums) and eval (1 or 2 doubles each).
struct HostFusedRep {
    std::vector<int32_t> slice_tab, lane_ptr, ecol;
    std::vector<double> eval;
};

// Builds `out` from the compressed forward extras xc (compress_column_replicas)
// when the replicated rows are exactly the extra DOFs (rep.c0 == nflow) and
// every listed row is a flow row whose entries all reference column sums
// (column >= n); false (out empty) otherwise.
bool build_fused_tables(const HostExtra &xc, const HostReplica &rep, int64_t n, int64_t nflow,
                        HostFusedRep &out);

// Applies kind 2 with the forward operator's (c0, p, nrep) on its adjoint
// extra CSR; false (x unchanged) if the rows do not repeat.
bool compress_row_replicas(HostExtra &x, int64_t n, int64_t c0, int64_t p, int64_t nrep,
                           HostReplica &rep);

// The receptivity operator L_ = [L  repeat(B,1,N);  0  kron(diag(-i w), I_p)]
// (LINEARIZE_L_.jl:58-67) in the compressed form that compress_column_replicas
// and compress_row_replicas give for its assembled upload, made from the
// extras of ONE per-wavenumber system S = [L B; 0 0] (size n_l + p, from
// csc_to_blocks with the same scale, complex) and the N = nwave wavenumbers
// w, so that no copy of B is ever formed.  In: xf, xa are S's forward and
// adjoint extras; out: L_'s compressed extras (rows below n_l; B's columns
// renamed to the virtual columns n + q, n = n_l + N p), rf (kind 1) and ra
// (kind 2) with c0 = n_l, the period p, N copies and the diagonal
// -i w_k scale (+i w_k scale in the adjoint).  Throws std::invalid_argument if
// S's lower rows or its adjoint's freestream columns hold entries.
void receptivity_from_parts(int64_t n_l, int64_t p, int64_t nwave, const double *w, double scale,
                            HostExtra &xf, HostExtra &xa, HostReplica &rf, HostReplica &ra);

// Order of the m long extra rows (one k_extra workgroup each; len[i]
// entries) such that the 8 contiguous ranges xcd_tile hands to the 8 XCDs
// (sizes m / 8 + 1 for the first m % 8, m / 8 for the rest) hold about the
// same number of entries: when the lengths differ by more than an eighth,
// the rows are dealt longest first to the range with the fewest entries that
// has room, and keep their relative order inside a range; otherwise (and
// below 16 rows) the order is unchanged, so neighbouring rows stay on one
// XCD.  Returns new position -> old index.
std::vector<int64_t> xcd_balanced_order(const std::vector<int64_t> &len);

// The extra entries of a real operator whose flow rows couple only to
// extra DOFs (the shock bands: A12 in the forward operator, A21' in the
// adjoint) in the form the single-vector term kernel takes them in its own
// launch: the extra-DOF rows (rows >= nflow) as their own CSR, finished by
// leading workgroups, and, per slice of 64 rows holding flow-row entries, a
// dense panel: the slice's extra columns (sorted) with 64 values each, one
// per row of the slice (zero where a row has no entry), added by the
// slice's workgroup.  slice_tab[s] is the panel of slice s or -1; panel k
// owns columns [cptr[k], cptr[k+1]) of pcol and values pval[64 c + lane]
// for those columns c.
struct HostShockPanels {
    std::vector<int32_t> slice_tab;
    std::vector<int64_t> cptr;
    std::vector<int32_t> pcol;
    std::vector<double> pval;
};

// Builds the panels and the extra-DOF rows from the extra CSR x of an
// operator with nflow flow rows; false (outputs empty) when x is complex,
// when a flow row has an extra entry in a flow column, when there is nothing
// to fuse, or when the panels would hold more than 8 values per entry they
// replace (rows of a slice with little overlap in their columns).
bool build_shock_fuse(const HostExtra &x, int64_t nflow, HostExtra &rows_out,
                      HostShockPanels &pan);

// The rows a batched term over the shock bands finishes in its batch
// kernel's leading workgroups (cell-major rows and columns): `sh`, the
// extra-DOF rows (rows_x, from build_shock_fuse), and `fr`, every flow row
// of every slice that holds a panel, whole -- its block entries followed by
// its extra entries -- so that the batch kernel skips those slices and no
// launch follows it.  Block values are read as bval[src * b * b + r * b + j]
// for block row r and column j (transpose: bval[src * b * b + j * b + r]),
// src = bsrc[k] for block k when bsrc is given (else k), times scale.  `fr`
// holds them in FP64, `fr32` the same with the block entries rounded to FP32
// (what the batch kernel multiplies when the operator stores FP32).
struct HostShockLead {
    HostExtra sh, fr, fr32;
};

void build_shock_lead(int b, const std::vector<int64_t> &bptr, const std::vector<int32_t> &bcol,
                      const double *bval, const std::vector<int64_t> *bsrc, bool transpose,
                      double scale, const HostExtra &x, const HostExtra &rows_x,
                      const HostShockPanels &pan, int64_t nflow, HostShockLead &out);

void csc_to_blocks(int64_t n, const int64_t *colptr, const int64_t *rowval,
                   const double *nzval, int index_base, bool is_complex, double scale,
                   int b, int64_t ncell, HostBSR &fwd, HostBSR &adj,
                   HostExtra &xf, HostExtra &xa);

void bsr_transpose(const HostBSR &a, HostBSR &t);

void extra_from_coo(int64_t n, int64_t nflow, const std::vector<int64_t> &r,
                    const std::vector<int64_t> &c, const std::vector<double> &v,
                    bool is_complex, bool adjoint, HostExtra &out);

void bsr_to_sliced(const HostBSR &a, HostSliced &s);

// Structure of the transpose of a BSR given by (ptr, col): its ptr, col and,
// per block, the forward block it mirrors (its values read transposed).
void bsr_transpose_map(int64_t ncell, const std::vector<int64_t> &ptr,
                       const std::vector<int32_t> &col, std::vector<int64_t> &tptr,
                       std::vector<int32_t> &tcol, std::vector<int64_t> &tsrc);

// bsr_to_sliced's structure (s.val stays empty) and, per padded block, the
// source block of its values: src[k] for the BSR's block k (k itself when
// src is null), -1 for a zero padding block.  The values are placed on the
// device from these indices.
void bsr_to_sliced_map(int b, int64_t ncell, const std::vector<int64_t> &ptr,
                       const std::vector<int32_t> &col, const std::vector<int64_t> *src,
                       HostSliced &s, std::vector<int64_t> &pad_src);

// Lane-interleaved value order inside every slot (vpl values per 16-byte
// chunk: 2 for FP64 storage, 4 for FP32).
void interleave_lanes(HostSliced &s, int vpl);

// Block columns of cell pairs (2p, 2p + 1), for the kernel that puts both
// cells of a pair into one matrix-core product: pair p lists the union of
// its two cells' block columns, ascending, in pcol[pptr[p] .. pptr[p+1]),
// and pslot holds each column's slot in the two cells' block rows (the
// position k - ptr[cell] of the block in the cell's row, which is its slot
// in the sliced layout): bits 0-15 cell 2p, bits 16-31 cell 2p + 1,
// kPairNone where the cell has no block in that column.
constexpr int32_t kPairNone = 0xFFFF;
void bsr_pair_map(int64_t ncell, const std::vector<int64_t> &ptr,
                  const std::vector<int32_t> &col, std::vector<int64_t> &pptr,
                  std::vector<int32_t> &pcol, std::vector<int32_t> &pslot);

}  // namespace hymor
