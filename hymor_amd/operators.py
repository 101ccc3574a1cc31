"""Config-shaped synthetic linearised operators and layout conversions.

HYMOR builds the Jacobian A of its finite-volume operator by coloured finite
differences (utils/Stability_analysis/Modal_stability_analysis/
LINEARIZE_EQUATIONS_NO_DISCONTINUITY.jl:42-287).  The resulting matrix is a
Julia ``SparseMatrixCSC{Float64,Int64}`` with

* a variable-major state ``[rho(all cells); rho_u(all); rho_v(all);
  rho_E(all); shock(Nchi)]``, cells numbered ``i + (j-1)*Nchi`` (chi fastest),
  (LINEARIZE_EQUATIONS_NO_DISCONTINUITY.jl:276-285, 323-324);
* a 3x3 cell stencil times 4 variables in the flow-flow block A11
  (:299-366), explicit zeros kept by ``sparse(I,J,V)``;
* when ``perturb_shock`` is set, dense bands A12 (4*Nchi x Nchi), A21
  (Nchi x 8*Nchi) and A22 (Nchi x Nchi) for the shock-position DOFs
  (:139-274, 802-907).

The freestream-receptivity operator extends it (LINEARIZE_L_.jl:58-67):
``L_ = [L  repeat(B,1,N); 0  kron(diag(-i*w), I_p)]``.

No base flow exists in this build, so the *values* here are synthetic: a
convection-diffusion-like block stencil ``kappa(di,dj)*I + eps*G`` with G drawn
N(0,1) from PCG64 and the whole operator rescaled to a requested max row sum.
The *structure* (layout, stencil, shock bands, extension) follows the
reference exactly, which is what the HIP kernels depend on.

Internally the library works cell-major: global index ``cell*b + var`` for the
b*ncell flow unknowns, then the extra DOFs (shock, freestream) in their
original order.  ``BlockOperator`` holds A in that cell-major form: a BSR part
over cell blocks plus an "extra" COO part for every entry touching an extra
DOF (and, for complex operators, every imaginary part).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# 3x3 stencil offsets (di along chi, dj along eta) and their diagonal weights.
_OFFSETS = [(di, dj) for dj in (-1, 0, 1) for di in (-1, 0, 1)]


def _kappa(di: int, dj: int) -> float:
    if di == 0 and dj == 0:
        return -1.0
    if dj == 0:
        return 0.25 - 0.05 * di       # upwind-biased convection along chi
    if di == 0:
        return 0.20                   # diffusion along eta
    return 0.025 * di * dj            # cross-derivative corner terms


def varmajor_to_cellmajor_index(ncell: int, b: int, n: int) -> np.ndarray:
    """perm[vm] = cm for every global index (0-based)."""
    vm = np.arange(n, dtype=np.int64)
    cm = vm.copy()
    flow = vm < b * ncell
    k = vm[flow] // ncell
    c = vm[flow] % ncell
    cm[flow] = c * b + k
    return cm


@dataclass
class BlockOperator:
    """Cell-major block operator: BSR over cells + extra COO entries.

    ``bval[k]`` is the b x b block (row-major) coupling cell ``row(k)`` to cell
    ``bcol[k]``; ``brow_ptr`` is the usual CSR pointer over block rows, with
    block columns sorted ascending inside each row.  ``xrow``/``xcol`` are
    cell-major global indices of the extra entries, ``xval`` their values.
    """

    nchi: int
    neta: int
    b: int
    n_extra: int
    brow_ptr: np.ndarray
    bcol: np.ndarray
    bval: np.ndarray
    xrow: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    xcol: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    xval: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))

    @property
    def ncell(self) -> int:
        return self.nchi * self.neta

    @property
    def n(self) -> int:
        return self.b * self.ncell + self.n_extra

    @property
    def nnzb(self) -> int:
        return int(self.bcol.shape[0])

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.xval) or np.iscomplexobj(self.bval)

    def coo_cellmajor(self):
        """All entries as cell-major COO (rows, cols, vals)."""
        b = self.b
        nb = self.nnzb
        brow = np.repeat(np.arange(self.ncell, dtype=np.int64), np.diff(self.brow_ptr))
        ii = np.arange(b, dtype=np.int64)
        r = (brow[:, None, None] * b + ii[None, :, None]) + np.zeros((1, 1, b), np.int64)
        c = (self.bcol.astype(np.int64)[:, None, None] * b + ii[None, None, :]) + np.zeros((1, b, 1), np.int64)
        rows = np.concatenate([r.reshape(nb * b * b), self.xrow.astype(np.int64)])
        cols = np.concatenate([c.reshape(nb * b * b), self.xcol.astype(np.int64)])
        dt = np.complex128 if self.is_complex else np.float64
        vals = np.concatenate([self.bval.reshape(nb * b * b).astype(dt), self.xval.astype(dt)])
        return rows, cols, vals

    def to_scipy_cellmajor(self):
        import scipy.sparse as sp
        r, c, v = self.coo_cellmajor()
        return sp.csr_matrix((v, (r, c)), shape=(self.n, self.n))

    def to_csc_varmajor(self, one_based: bool = True):
        """Julia-style CSC arrays (colptr, rowval, nzval) in var-major order.

        Duplicates are summed and explicit zeros kept, as ``sparse(I,J,V)``
        does; row indices are sorted inside each column.
        """
        r, c, v = self.coo_cellmajor()
        inv = np.empty(self.n, dtype=np.int64)
        inv[varmajor_to_cellmajor_index(self.ncell, self.b, self.n)] = np.arange(self.n, dtype=np.int64)
        rv = inv[r]
        cv = inv[c]
        order = np.lexsort((rv, cv))
        rv, cv, v = rv[order], cv[order], v[order]
        # sum duplicates
        key = cv * self.n + rv
        uniq, start = np.unique(key, return_index=True)
        if uniq.shape[0] != key.shape[0]:
            v = np.add.reduceat(v, start)
            rv = rv[start]
            cv = cv[start]
        counts = np.bincount(cv, minlength=self.n)
        colptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(counts, out=colptr[1:])
        base = 1 if one_based else 0
        return colptr + base, rv + base, np.ascontiguousarray(v)

    def to_scipy_varmajor(self):
        import scipy.sparse as sp
        colptr, rowval, nzval = self.to_csc_varmajor(one_based=False)
        return sp.csc_matrix((nzval, rowval, colptr), shape=(self.n, self.n))

    def to_csc_varmajor_blocks(self, one_based: bool = True):
        """Same arrays as to_csc_varmajor for operators without extras, built
        by direct placement instead of a global sort (large configs).

        Column (j, C) of the var-major CSC (var j of cell C) holds, for
        i = 0..b-1 and the blocks (R, C) in ascending R, the entries
        (i*ncell + R) -- rows ascending, as SparseMatrixCSC stores them.
        """
        if self.xrow.shape[0] or self.n_extra:
            return self.to_csc_varmajor(one_based)
        b, ncell = self.b, self.ncell
        R = np.repeat(np.arange(ncell, dtype=np.int64), np.diff(self.brow_ptr))
        Cc = self.bcol.astype(np.int64)
        tk = np.argsort(Cc, kind="stable")             # blocks in (C, R) order
        cnt = np.bincount(Cc, minlength=ncell)
        cstart = np.zeros(ncell + 1, dtype=np.int64)
        np.cumsum(cnt, out=cstart[1:])
        rank = np.empty(self.nnzb, dtype=np.int64)
        rank[tk] = np.arange(self.nnzb, dtype=np.int64) - cstart[Cc[tk]]
        colcount = np.tile(cnt * b, b)                  # column j*ncell + C
        colptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(colcount, out=colptr[1:])
        nnz = int(colptr[-1])
        dt = np.complex128 if self.is_complex else np.float64
        nzval = np.empty(nnz, dtype=dt)
        rowval = np.empty(nnz, dtype=np.int64)
        for j in range(b):
            colbase = colptr[j * ncell + Cc] + rank
            for i in range(b):
                pos = colbase + i * cnt[Cc]
                nzval[pos] = self.bval[:, i, j]
                rowval[pos] = i * ncell + R
        base = 1 if one_based else 0
        return colptr + base, rowval + base, nzval

    def to_csr_cellmajor(self):
        """scipy CSR of the cell-major operator (blocks expanded)."""
        import scipy.sparse as sp
        if self.xrow.shape[0] or self.n_extra:
            return self.to_scipy_cellmajor().tocsr()
        b = self.b
        # row (cell*b + i) holds, per block of the cell in order, b entries
        lens = np.diff(self.brow_ptr) * b
        indptr = np.zeros(self.n + 1, dtype=np.int64)
        indptr[1:] = np.cumsum(np.repeat(lens, b))
        brow = np.repeat(np.arange(self.ncell, dtype=np.int64), np.diff(self.brow_ptr))
        # order: for cell, for i, for block k of the cell, for j
        blk_of_cell = np.arange(self.nnzb) - self.brow_ptr[brow]
        nb_cell = np.diff(self.brow_ptr)[brow]
        ii = np.arange(b)
        jj = np.arange(b)
        pos = (indptr[brow * b][:, None, None] + ii[None, :, None] * (nb_cell[:, None, None] * b)
               + blk_of_cell[:, None, None] * b + jj[None, None, :])
        data = np.empty(int(indptr[-1]), dtype=self.bval.dtype)
        cols = np.empty(int(indptr[-1]), dtype=np.int64)
        data[pos.reshape(-1)] = self.bval.reshape(-1)
        cols[pos.reshape(-1)] = (self.bcol.astype(np.int64)[:, None, None] * b
                                 + np.zeros((1, b, 1), np.int64) + jj[None, None, :]).reshape(-1)
        return sp.csr_matrix((data, cols, indptr), shape=(self.n, self.n))

    def scaled(self, s):
        return BlockOperator(self.nchi, self.neta, self.b, self.n_extra, self.brow_ptr, self.bcol,
                             self.bval * s, self.xrow, self.xcol, self.xval * s)


def stencil_operator(nchi: int, neta: int, b: int = 4, seed: int = 0, *,
                     periodic_chi: bool = False, n_shock: int = 0,
                     row_sum: float | None = None, eps: float | None = None,
                     dtype=np.float64) -> BlockOperator:
    """9-cell x b*b block stencil on an nchi x neta grid.

    ``n_shock`` > 0 (must equal nchi, as in the reference) appends the
    shock-position DOFs with the reference's band shapes: A12 dense over the
    4 (b) variables of the eta row nearest the shock, A21 over the two
    nearest eta rows, A22 dense.  ``row_sum`` rescales A so that its largest
    absolute row sum equals that value.
    """
    if n_shock not in (0, nchi):
        raise ValueError("shock DOFs are one per chi station (n_shock == nchi)")
    rng = np.random.Generator(np.random.PCG64(seed))
    ncell = nchi * neta
    if eps is None:
        eps = 0.3 / np.sqrt(9.0 * b)
    ci = np.arange(ncell, dtype=np.int64) % nchi
    cj = np.arange(ncell, dtype=np.int64) // nchi
    rows, cols, kap = [], [], []
    for di, dj in _OFFSETS:
        ni = ci + di
        nj = cj + dj
        if periodic_chi:
            ni = ni % nchi
            ok = (nj >= 0) & (nj < neta)
        else:
            ok = (ni >= 0) & (ni < nchi) & (nj >= 0) & (nj < neta)
        rows.append(np.nonzero(ok)[0])
        cols.append((ni + nj * nchi)[ok])
        kap.append(np.full(int(ok.sum()), _kappa(di, dj)))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    kap = np.concatenate(kap)
    order = np.lexsort((cols, rows))
    rows, cols, kap = rows[order], cols[order], kap[order]
    if periodic_chi and nchi < 3:
        raise ValueError("periodic chi needs nchi >= 3")
    nnzb = rows.shape[0]
    g = rng.standard_normal((nnzb, b, b))
    bval = g * eps
    idx = np.arange(b)
    bval[:, idx, idx] += kap[:, None]
    brow_ptr = np.zeros(ncell + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=ncell), out=brow_ptr[1:])
    op = BlockOperator(nchi, neta, b, n_shock, brow_ptr, cols.astype(np.int32), bval)
    if n_shock:
        nf = b * ncell
        sd = 1.0 / np.sqrt(nchi)
        top = (neta - 1) * nchi + np.arange(nchi)          # cells next to the shock
        top2 = (neta - 2) * nchi + np.arange(nchi) if neta > 1 else top
        shock = nf + np.arange(nchi)
        # A12: the b vars of the top eta row  x  all shock DOFs (dense)
        r12 = (top[:, None] * b + np.arange(b)[None, :]).reshape(-1)
        xr = [np.repeat(r12, nchi)]
        xc = [np.tile(shock, r12.shape[0])]
        # A21: shock DOFs x the b vars of the two nearest eta rows (dense)
        c21 = np.concatenate([(top[:, None] * b + np.arange(b)).reshape(-1),
                              (top2[:, None] * b + np.arange(b)).reshape(-1)])
        c21 = np.unique(c21)
        xr.append(np.repeat(shock, c21.shape[0]))
        xc.append(np.tile(c21, nchi))
        # A22: shock x shock (dense)
        xr.append(np.repeat(shock, nchi))
        xc.append(np.tile(shock, nchi))
        xrow = np.concatenate(xr)
        xcol = np.concatenate(xc)
        xval = rng.standard_normal(xrow.shape[0]) * sd * 0.1
        diag = xrow == xcol
        xval[diag] -= 1.0
        op.xrow, op.xcol, op.xval = xrow, xcol, xval
    if row_sum is not None:
        op = op.scaled(row_sum / max_abs_row_sum(op))
    if dtype == np.complex128:
        op.bval = op.bval.astype(np.complex128)
        op.xval = op.xval.astype(np.complex128)
    return op


def max_abs_row_sum(op: BlockOperator) -> float:
    b = op.b
    rs = np.zeros(op.n)
    brow = np.repeat(np.arange(op.ncell), np.diff(op.brow_ptr))
    contrib = np.abs(op.bval).sum(axis=2)                       # (nnzb, b)
    np.add.at(rs, (brow[:, None] * b + np.arange(b)[None, :]).reshape(-1), contrib.reshape(-1))
    if op.xrow.shape[0]:
        np.add.at(rs, op.xrow, np.abs(op.xval))
    return float(rs.max())


def receptivity_extension(op: BlockOperator, w_infty, seed: int = 0, b_scale: float = 0.1) -> BlockOperator:
    """Extended operator L_ = [L  repeat(B,1,N); 0  kron(diag(-i*w), I_p)].

    Follows LINEARIZE_L_.jl:58-67 (note the -i*w sign).  B maps the
    p = b*Nchi freestream DOFs of one frequency onto the flow rows of the
    cells next to the inflow boundary (top eta row), tridiagonal in chi per
    variable pair, the shape LINEARIZE_B.jl:68-202 produces.  The same B is
    repeated for every frequency.  Values are synthetic N(0, b_scale^2).
    """
    rng = np.random.Generator(np.random.PCG64(10_000 + seed))
    w = np.asarray(w_infty, dtype=np.float64)
    nchi, b = op.nchi, op.b
    p = b * nchi
    n0 = op.n
    top = (op.neta - 1) * nchi + np.arange(nchi)
    # B triplets in cell-major row index / local freestream column (var-major
    # inside the freestream block: col = k_in*Nchi + i', as CONSTRUCT_R_ orders it)
    br, bc = [], []
    for ko in range(b):
        for ki in range(b):
            for d in (-1, 0, 1):
                i = np.arange(nchi)
                i2 = i + d
                ok = (i2 >= 0) & (i2 < nchi)
                br.append(top[ok] * b + ko)
                bc.append(ki * nchi + i2[ok])
    br = np.concatenate(br)
    bc = np.concatenate(bc)
    bv = rng.standard_normal(br.shape[0]) * b_scale
    xr = [op.xrow.astype(np.int64)]
    xc = [op.xcol.astype(np.int64)]
    xv = [op.xval.astype(np.complex128)]
    for k, wk in enumerate(w):
        off = n0 + k * p
        xr.append(br)
        xc.append(bc + off)
        xv.append(bv.astype(np.complex128))
        xr.append(off + np.arange(p))
        xc.append(off + np.arange(p))
        xv.append(np.full(p, -1j * wk))
    return BlockOperator(op.nchi, op.neta, b, op.n_extra + w.shape[0] * p, op.brow_ptr, op.bcol,
                         op.bval.astype(np.complex128), np.concatenate(xr), np.concatenate(xc),
                         np.concatenate(xv))


def marching_operator(nchi: int, neta: int, b: int, seed: int = 0, n_lead: int = 12,
                      lead_top: float = 0.55, lead_gap: float = 0.1, bulk=(-3.0, -1.0),
                      couple: float = 0.15, n_shock: int = 0):
    """3x3-stencil block operator with an exactly known spectrum.

    Every cell couples only to the neighbours with a larger cell index (east,
    and the three cells of the next eta row): a space-marching, one-way
    discretisation.  The blocks towards the other four neighbours are kept in
    the pattern as explicit zeros, so storage, bytes and kernel work are those
    of the full 9-block stencil.  The operator is block upper triangular in
    cell order, so its eigenvalues are exactly those of the diagonal blocks.
    Diagonal block i is S_i diag(l_i) S_i^-1 with S_i = I + E_i, E_i = 0.2 N(0,1)
    scaled to ||E_i||_F <= 0.5 (so cond(S_i) <= 3), and real eigenvalues l_i
    uniform in `bulk`, except `n_lead` cells spread over the
    grid whose leading eigenvalue is lead_top - lead_gap * j (j = 0..n_lead-1).
    Returns (op, eigenvalues of A) with the eigenvalues computed from the
    factors (l_i), not from the stored blocks.

    n_shock = nchi appends shock-position DOFs with the band shapes of
    stencil_operator (A12, A21, A22 dense), kept one-way as well: A12 holds
    explicit zeros, A21 random values, A22 is upper triangular with its
    diagonal drawn from `bulk` (explicit zeros below).  The operator is then
    block lower triangular over [flow; shock], and the spectrum is the flow
    part's plus A22's diagonal.
    """
    if n_shock not in (0, nchi):
        raise ValueError("shock DOFs are one per chi station (n_shock == nchi)")
    rng = np.random.Generator(np.random.PCG64(20_000 + seed))
    ncell = nchi * neta
    ci = np.arange(ncell, dtype=np.int64) % nchi
    cj = np.arange(ncell, dtype=np.int64) // nchi
    rows, cols, live = [], [], []
    for di, dj in _OFFSETS:
        ni, nj = ci + di, cj + dj
        ok = (ni >= 0) & (ni < nchi) & (nj >= 0) & (nj < neta)
        r = np.nonzero(ok)[0]
        c = (ni + nj * nchi)[ok]
        rows.append(r)
        cols.append(c)
        live.append(c > r)
    rows, cols, live = (np.concatenate(v) for v in (rows, cols, live))
    order = np.lexsort((cols, rows))
    rows, cols, live = rows[order], cols[order], live[order]
    bval = np.zeros((rows.shape[0], b, b))
    bval[live] = rng.standard_normal((int(live.sum()), b, b)) * (couple / np.sqrt(b))
    lam = rng.uniform(bulk[0], bulk[1], size=(ncell, b))
    lead_cells = np.linspace(0, ncell - 1, n_lead + 2).astype(np.int64)[1:-1]
    lam[lead_cells, 0] = lead_top - lead_gap * np.arange(n_lead)
    E = 0.2 * rng.standard_normal((ncell, b, b))
    fro = np.sqrt((E * E).sum(axis=(1, 2)))
    E *= np.minimum(1.0, 0.5 / fro)[:, None, None]     # ||E||_2 <= 0.5: cond(S) <= 3
    S = np.eye(b)[None] + E
    D = np.einsum("cij,cj,cjk->cik", S, lam, np.linalg.inv(S))
    diag_pos = np.nonzero(rows == cols)[0]
    bval[diag_pos] = D[rows[diag_pos]]
    brow_ptr = np.zeros(ncell + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=ncell), out=brow_ptr[1:])
    op = BlockOperator(nchi, neta, b, n_shock, brow_ptr, cols.astype(np.int32), bval)
    lam = lam.reshape(-1)
    if n_shock:
        nf = b * ncell
        top = (neta - 1) * nchi + np.arange(nchi)
        top2 = (neta - 2) * nchi + np.arange(nchi) if neta > 1 else top
        shock = nf + np.arange(nchi)
        r12 = (top[:, None] * b + np.arange(b)[None, :]).reshape(-1)
        c21 = np.unique(np.concatenate([(top[:, None] * b + np.arange(b)).reshape(-1),
                                        (top2[:, None] * b + np.arange(b)).reshape(-1)]))
        xr = [np.repeat(r12, nchi), np.repeat(shock, c21.shape[0]), np.repeat(shock, nchi)]
        xc = [np.tile(shock, r12.shape[0]), np.tile(c21, nchi), np.tile(shock, nchi)]
        v12 = np.zeros(r12.shape[0] * nchi)                       # one-way: explicit zeros
        v21 = rng.standard_normal(nchi * c21.shape[0]) * (couple / np.sqrt(c21.shape[0]))
        d22 = rng.uniform(bulk[0], bulk[1], size=nchi)
        A22 = np.triu(rng.standard_normal((nchi, nchi)) * (couple / np.sqrt(nchi)), 1)
        A22[np.arange(nchi), np.arange(nchi)] = d22
        op.xrow = np.concatenate(xr)
        op.xcol = np.concatenate(xc)
        op.xval = np.concatenate([v12, v21, A22.reshape(-1)])
        lam = np.concatenate([lam, d22])
    return op, lam


def circulant_operator(nchi: int, neta: int, b: int, seed: int = 0, row_sum: float | None = None):
    """Doubly periodic, constant-coefficient 9-cell stencil.

    Returns (op, blocks) where blocks[(di,dj)] is the b x b coupling of a cell
    to its neighbour at offset (di, dj).  The operator is block-circulant, so
    the 2-D DFT block-diagonalises it; the oracle uses that for closed forms.
    Needs nchi, neta >= 3 so the 9 offsets map to distinct cells.
    """
    if nchi < 3 or neta < 3:
        raise ValueError("need nchi, neta >= 3")
    rng = np.random.Generator(np.random.PCG64(20_000 + seed))
    eps = 0.3 / np.sqrt(9.0 * b)
    blocks = {}
    for di, dj in _OFFSETS:
        blk = rng.standard_normal((b, b)) * eps
        blk[np.arange(b), np.arange(b)] += _kappa(di, dj)
        blocks[(di, dj)] = blk
    ncell = nchi * neta
    ci = np.arange(ncell) % nchi
    cj = np.arange(ncell) // nchi
    rows, cols, vals = [], [], []
    for (di, dj), blk in blocks.items():
        nb = ((ci + di) % nchi) + ((cj + dj) % neta) * nchi
        rows.append(np.arange(ncell))
        cols.append(nb)
        vals.append(np.broadcast_to(blk, (ncell, b, b)))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], np.ascontiguousarray(vals[order])
    brow_ptr = np.zeros(ncell + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=ncell), out=brow_ptr[1:])
    op = BlockOperator(nchi, neta, b, 0, brow_ptr, cols.astype(np.int32), vals)
    if row_sum is not None:
        s = row_sum / max_abs_row_sum(op)
        op = op.scaled(s)
        blocks = {k: v * s for k, v in blocks.items()}
    return op, blocks


def circulant_spectrum(blocks: dict, nchi: int, neta: int) -> np.ndarray:
    """All eigenvalues of the block-circulant operator of circulant_operator:
    the union over the 2-D wavenumbers (p, q) of the eigenvalues of its
    symbol sum_{(di,dj)} B_{di,dj} exp(2 pi i (di p/nchi + dj q/neta))."""
    p = np.arange(nchi)
    q = np.arange(neta)
    b = next(iter(blocks.values())).shape[0]
    S = np.zeros((neta, nchi, b, b), dtype=np.complex128)
    for (di, dj), B in blocks.items():
        ph = np.exp(2j * np.pi * (di * p[None, :] / nchi + dj * q[:, None] / neta))
        S += ph[:, :, None, None] * B[None, None, :, :]
    return np.linalg.eigvals(S.reshape(-1, b, b)).reshape(-1)


# Config shapes (grid, block size, extra DOFs) used by tests and the bench.
CONFIGS = {
    "C1": dict(nchi=100, neta=80, b=4, n_shock=0, seed=1001),
    "C2": dict(nchi=200, neta=300, b=4, n_shock=200, seed=1002),
    "C3": dict(nchi=400, neta=300, b=4, n_shock=400, seed=1003),
    "C4": dict(nchi=500, neta=250, b=4, n_shock=0, seed=1004),
    "C5": dict(nchi=600, neta=300, b=8, n_shock=0, seed=1005),
}


def config_operator(name: str, row_sum: float = 32.0) -> BlockOperator:
    c = CONFIGS[name]
    return stencil_operator(c["nchi"], c["neta"], c["b"], c["seed"], n_shock=c["n_shock"], row_sum=row_sum)
