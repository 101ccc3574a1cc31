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
