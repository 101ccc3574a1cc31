"""Host layout conversion + kernel index math, emulated on the CPU.

tests/layout_emulator.cpp is compiled with g++ against the library's own
hymor_amd/csrc/convert.cpp; it replays the device kernels' index arithmetic
on the converted layout with bounds checks.  Results must match scipy's
SpMV of the reference-layout CSC, with zero out-of-bounds accesses.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import pytest

from hymor_amd import operators as ops

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = tmp_path_factory.mktemp("emu") / "libemu.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC",
                    f"-I{ROOT / 'hymor_amd' / 'csrc'}", str(ROOT / "tests" / "layout_emulator.cpp"),
                    str(ROOT / "hymor_amd" / "csrc" / "convert.cpp"), "-o", str(out)], check=True)
    lib = C.CDLL(str(out))
    lib.emu_spmv_csc.restype = C.c_int
    lib.emu_spmv_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [C.c_int] * 3 + [C.c_int64,
                                                                                  C.c_int, C.c_void_p,
                                                                                  C.c_void_p]
    return lib


def _run(emu, op, transpose, x):
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    cplx = np.iscomplexobj(nzval)
    x = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64)
    y = np.zeros_like(x)
    oob = emu.emu_spmv_csc(op.n, colptr.ctypes.data, rowval.ctypes.data,
                           np.ascontiguousarray(nzval).ctypes.data, 1, int(cplx), op.b, op.ncell,
                           int(transpose), x.ctypes.data, y.ctypes.data)
    return y, oob


CASES = [
    dict(nchi=20, neta=16, b=4, seed=1),
    dict(nchi=12, neta=10, b=4, seed=2, n_shock=12),
    dict(nchi=10, neta=8, b=8, seed=3),
    dict(nchi=9, neta=6, b=4, seed=4, periodic_chi=True),
    dict(nchi=7, neta=5, b=8, seed=5, n_shock=7),
    dict(nchi=3, neta=1, b=4, seed=6),
    dict(nchi=1, neta=1, b=8, seed=7),
]


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("kind", ["shock_b4", "shock_b8", "complex"])
def test_conversion_split_over_host_threads(emu, kind, transpose):
    """Grids large enough that the upload's conversion splits its cell
    columns over host threads (one per 4 096: 19 200 and 9 600 cells): the
    blocks, their transposes, the regrouping by block row and the extra
    rows (shock DOFs; a complex receptivity extension) give scipy's product."""
    if kind == "shock_b4":
        op = ops.stencil_operator(160, 120, 4, seed=8, n_shock=160)
    elif kind == "shock_b8":
        op = ops.stencil_operator(120, 80, 8, seed=9, n_shock=120)
    else:
        op = ops.receptivity_extension(ops.stencil_operator(100, 90, 4, seed=10, n_shock=100),
                                       [0.0, 1.5])
    A = op.to_scipy_varmajor()
    r = np.random.default_rng(3)
    x = r.standard_normal(op.n)
    if np.iscomplexobj(A.data):
        x = x + 1j * r.standard_normal(op.n)
    y, oob = _run(emu, op, transpose, x)
    assert oob == 0
    ref = (A.conj().T if transpose else A) @ x     # the adjoint: conjugate transpose
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['nchi']}x{c['neta']}b{c['b']}")
@pytest.mark.parametrize("transpose", [False, True])
def test_emulated_spmv_matches_scipy(emu, case, transpose):
    c = dict(case)
    op = ops.stencil_operator(c.pop("nchi"), c.pop("neta"), c.pop("b"), c.pop("seed"), **c)
    A = op.to_scipy_varmajor()
    x = np.random.default_rng(0).standard_normal(op.n)
    y, oob = _run(emu, op, transpose, x)
    assert oob == 0
    ref = (A.conj().T if transpose else A) @ x
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("transpose", [False, True])
def test_emulated_complex_receptivity(emu, transpose):
    base = ops.stencil_operator(6, 4, 4, seed=8, n_shock=6)
    L_ = ops.receptivity_extension(base, [0.0, 1.0, 2.5])
    A = L_.to_scipy_varmajor()
    r = np.random.default_rng(1)
    x = r.standard_normal(L_.n) + 1j * r.standard_normal(L_.n)
    y, oob = _run(emu, L_, transpose, x)
    assert oob == 0
    ref = (A.conj().T if transpose else A) @ x
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def _run_rep(emu, op, transpose, x):
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    cplx = np.iscomplexobj(nzval)
    x = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64)
    y = np.zeros_like(x)
    info = np.zeros(4, dtype=np.int64)
    emu.emu_spmv_csc_rep.restype = C.c_int
    oob = emu.emu_spmv_csc_rep(C.c_int64(op.n), C.c_void_p(colptr.ctypes.data),
                               C.c_void_p(rowval.ctypes.data),
                               C.c_void_p(np.ascontiguousarray(nzval).ctypes.data), C.c_int(1),
                               C.c_int(int(cplx)), C.c_int(op.b), C.c_int64(op.ncell),
                               C.c_int(int(transpose)), C.c_void_p(x.ctypes.data),
                               C.c_void_p(y.ctypes.data), C.c_void_p(info.ctypes.data))
    return y, oob, info


@pytest.mark.parametrize("nchi,neta,shock,w", [(6, 4, True, [0.0, 1.0, 2.5]),
                                               (5, 3, False, [0.3, 0.7]),
                                               (8, 5, True, list(np.linspace(0.1, 4.0, 9)))])
@pytest.mark.parametrize("transpose", [False, True])
def test_emulated_receptivity_replicas(emu, nchi, neta, shock, w, transpose):
    """The receptivity operator's N copies of B are found at upload (forward:
    N copies of p = 4 nchi freestream columns after the flow and shock DOFs;
    adjoint: N copies of its rows) and the compressed arithmetic -- column
    sums of the freestream blocks, shared rows finished with their own
    diagonal -- gives scipy's SpMV of the extended operator."""
    base = ops.stencil_operator(nchi, neta, 4, seed=9, n_shock=nchi if shock else 0)
    L_ = ops.receptivity_extension(base, w)
    A = L_.to_scipy_varmajor()
    r = np.random.default_rng(2)
    x = r.standard_normal(L_.n) + 1j * r.standard_normal(L_.n)
    y, oob, info = _run_rep(emu, L_, transpose, x)
    assert oob == 0
    assert list(info) == [2 if transpose else 1, base.n, 4 * nchi, len(w)]
    ref = (A.conj().T if transpose else A) @ x
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: f"{c['nchi']}x{c['neta']}b{c['b']}")
def test_emulated_no_replicas_elsewhere(emu, case):
    """Operators without repeated freestream blocks are left as they are."""
    c = dict(case)
    op = ops.stencil_operator(c.pop("nchi"), c.pop("neta"), c.pop("b"), c.pop("seed"), **c)
    x = np.random.default_rng(3).standard_normal(op.n)
    y, oob, info = _run_rep(emu, op, False, x)
    assert oob == 0 and info[0] == 0
    assert np.allclose(y, op.to_scipy_varmajor() @ x, rtol=1e-13, atol=1e-13)


def test_emulated_single_wavenumber_is_not_compressed(emu):
    """One freestream block (N = 1) has nothing to share."""
    base = ops.stencil_operator(5, 4, 4, seed=10, n_shock=5)
    L_ = ops.receptivity_extension(base, [1.3])
    x = np.random.default_rng(4).standard_normal(L_.n) * (1 + 0j)
    y, oob, info = _run_rep(emu, L_, False, x)
    assert oob == 0 and info[0] == 0
    assert np.allclose(y, L_.to_scipy_varmajor() @ x, rtol=1e-13, atol=1e-13)


def _csc1(A, dtype=np.complex128):
    A = sp.csc_matrix(A, dtype=dtype)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr + 1, dtype=np.int64),
            np.ascontiguousarray(A.indices + 1, dtype=np.int64),
            np.ascontiguousarray(A.data, dtype=dtype))


@pytest.mark.parametrize("nchi,neta,shock,N", [(6, 4, True, 3), (5, 3, False, 2),
                                               (8, 5, False, 32), (7, 4, True, 9)])
def test_receptivity_from_parts_equals_compressed_assembly(emu, nchi, neta, shock, N):
    """hymor_op_create_receptivity's host conversion (receptivity_from_parts,
    from ONE per-wavenumber system [L B; 0 0] and the wavenumbers) yields
    exactly the tables the library derives from the reference's assembled
    L_ = [L repeat(B,1,N); 0 kron(diag(-i w), I)]: block parts, forward and
    adjoint extra CSRs, both replica descriptions (period, copies, -i w
    diagonal, shared rows) and the fused Horner tables, entry for entry.  The
    assembled form's device arithmetic is pinned against the oracle by the
    GPU tests, so the form from parts runs the same arithmetic."""
    from hymor_amd.receptivity import per_omega_system, split_extended
    base = ops.stencil_operator(nchi, neta, 4, seed=70 + N, n_shock=nchi if shock else 0)
    w = np.linspace(0.0, 6.0, N)
    A_ = ops.receptivity_extension(base, w, seed=N).to_scipy_varmajor()
    n_L, p = base.n, 4 * nchi
    L, B = split_extended(A_, n_L, p, w)
    S = per_omega_system(L, B)
    a = _csc1(A_)
    s_ = _csc1(S)
    emu.emu_receptivity_parts_equal.restype = C.c_int
    rc = emu.emu_receptivity_parts_equal(
        C.c_int64(A_.shape[0]), *(C.c_void_p(v.ctypes.data) for v in a),
        C.c_int64(S.shape[0]), *(C.c_void_p(v.ctypes.data) for v in s_),
        C.c_int(4), C.c_int64(base.ncell), C.c_int64(n_L), C.c_int64(N),
        C.c_void_p(np.ascontiguousarray(w).ctypes.data), C.c_double(0.05))
    assert rc == 0


def _xcd_loads(order, lens):
    """Entries per XCD when k_extra's workgroup bid takes long row
    order[xcd_tile(bid, m)] and the hardware deals workgroups round-robin
    over the 8 XCDs (xcd_tile's contiguous ranges replayed)."""
    m = len(order)
    per, rem = m // 8, m % 8
    loads = np.zeros(8)
    for bid in range(m):
        xcd, idx = bid & 7, bid >> 3
        tile = xcd * (per + 1) + idx if xcd < rem else rem * (per + 1) + (xcd - rem) * per + idx
        assert 0 <= tile < m
        loads[xcd] += lens[order[tile]]
    return loads


@pytest.mark.parametrize("case", ["c3_forward", "c3_adjoint", "c2_forward", "ragged",
                                  "receptivity", "few"])
def test_long_extra_rows_spread_over_xcds(emu, case):
    """The long extra rows (one k_extra workgroup each) are ordered at upload
    so that the 8 XCDs get about the same number of entries: at 400 x 300
    cells with 400 shock DOFs the 400 shock rows (3 600 entries) would
    otherwise all fall on the last XCDs' ranges.  Rows of equal length keep
    their order (neighbouring rows on one XCD)."""
    lens = {"c3_forward": [400] * 1600 + [3600] * 400,
            "c3_adjoint": [400] * 3200 + [2000] * 400,
            "c2_forward": [200] * 800 + [1800] * 200,
            "ragged": list(np.random.default_rng(5).integers(33, 5000, 1003)),
            "receptivity": [384] * 2000,
            "few": [40, 5000, 60, 7000, 33]}[case]
    lens = np.asarray(lens, dtype=np.int64)
    out = np.zeros(lens.shape[0], dtype=np.int64)
    emu.emu_xcd_balanced_order(C.c_int64(lens.shape[0]), C.c_void_p(lens.ctypes.data),
                               C.c_void_p(out.ctypes.data))
    assert sorted(out) == list(range(lens.shape[0]))
    if case in ("receptivity", "few"):
        assert list(out) == list(range(lens.shape[0]))
        return
    loads = _xcd_loads(out, lens)
    assert loads.max() <= loads.mean() + lens.max(), loads
    before = _xcd_loads(np.arange(lens.shape[0]), lens)
    assert loads.max() < before.max()


@pytest.mark.parametrize("nchi,neta,b", [(10, 7, 4), (6, 5, 8), (9, 4, 4), (40, 5, 4)])
@pytest.mark.parametrize("transpose", [False, True])
def test_emulated_shock_bands_fused(emu, nchi, neta, b, transpose):
    """The shock bands in the single-vector term kernel's launch: the
    extra-DOF rows as their own CSR and, per slice of 64 rows, a dense panel
    of the flow rows' extra entries over the slice's extra columns.  Built
    for operators with shock bands (forward: A12 rows; adjoint: A21' rows),
    and the replayed SpMV (block term + panel per slice lane + extra-DOF
    rows, every access bounds-checked) equals scipy's."""
    op = ops.stencil_operator(nchi, neta, b, seed=80 + nchi, n_shock=nchi)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    A = op.to_scipy_varmajor()
    x = np.random.default_rng(nchi).standard_normal(op.n)
    y = np.zeros_like(x)
    info = np.zeros(4, dtype=np.int64)
    emu.emu_spmv_shock.restype = C.c_int
    oob = emu.emu_spmv_shock(C.c_int64(op.n), C.c_void_p(colptr.ctypes.data),
                             C.c_void_p(rowval.ctypes.data),
                             C.c_void_p(np.ascontiguousarray(nzval).ctypes.data), C.c_int(1),
                             C.c_int(b), C.c_int64(op.ncell), C.c_int(int(transpose)),
                             C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data),
                             C.c_void_p(info.ctypes.data))
    assert oob == 0
    assert info[0] == 1 and info[3] == nchi and info[1] >= 1
    ref = (A.T if transpose else A) @ x
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("R", [2, 8])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("nchi,neta,b", [(70, 3, 4), (40, 5, 4), (24, 4, 8)])
def test_emulated_shock_rows_batched(emu, nchi, neta, b, transpose, R):
    """A batched term over a shock operator in one launch: leading
    workgroups finish the extra-DOF rows and, whole, every flow row of the
    slices with shock columns; the block kernel skips those slices.  Every
    row is written exactly once, every access is in bounds, the product
    equals scipy's on all R columns; the FP32 copy holds the block entries
    rounded to FP32 and the shock entries unchanged; the BSR upload's
    adjoint (raw blocks through the transpose map) gives the same rows as
    the CSC upload's."""
    op = ops.stencil_operator(nchi, neta, b, seed=70 + nchi, n_shock=nchi)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    A = op.to_scipy_varmajor()
    X = np.asfortranarray(np.random.default_rng(R).standard_normal((op.n, R)))
    Y = np.zeros_like(X)
    info = np.zeros(5, dtype=np.int64)
    emu.emu_spmv_shock_batched.restype = C.c_int
    oob = emu.emu_spmv_shock_batched(
        C.c_int64(op.n), C.c_void_p(colptr.ctypes.data), C.c_void_p(rowval.ctypes.data),
        C.c_void_p(np.ascontiguousarray(nzval).ctypes.data), C.c_int(1), C.c_int(b),
        C.c_int64(op.ncell), C.c_int(int(transpose)), C.c_int(R), C.c_int(1),
        C.c_void_p(X.ctypes.data), C.c_void_p(Y.ctypes.data), C.c_void_p(info.ctypes.data))
    assert oob == 0
    assert info[0] == 1 and info[2] >= 1 and info[3] == 0 and info[4] == 0
    # leading rows: the nchi extra-DOF rows and 64 per skipped slice (fewer
    # in a last, partial slice)
    assert nchi + 64 * (info[2] - 1) < info[1] <= nchi + 64 * info[2]
    ref = (A.T if transpose else A) @ X
    assert np.allclose(Y, ref, rtol=1e-13, atol=1e-12)


def test_emulated_shock_fuse_declines_without_extras(emu):
    """An operator without extra DOFs has nothing to fuse: no fused form."""
    op = ops.stencil_operator(6, 4, 4, seed=90)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    x = np.random.default_rng(1).standard_normal(op.n)
    y = np.zeros_like(x)
    info = np.zeros(4, dtype=np.int64)
    emu.emu_spmv_shock.restype = C.c_int
    assert emu.emu_spmv_shock(C.c_int64(op.n), C.c_void_p(colptr.ctypes.data),
                              C.c_void_p(rowval.ctypes.data),
                              C.c_void_p(np.ascontiguousarray(nzval).ctypes.data), C.c_int(1),
                              C.c_int(4), C.c_int64(op.ncell), C.c_int(0),
                              C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data),
                              C.c_void_p(info.ctypes.data)) == 0
    assert info[0] == 0


def _run_fused(emu, A, b, ncell, X2, order, nsub, dt, nv=2):
    """emu_horner_fused on scipy matrix A (var-major) with nv doubles per
    row: X2 is n x nv/2 complex values (or n complex values at nv = 2) or
    an n x nv real array."""
    A = A.tocsc()
    A.sort_indices()
    cplx = np.iscomplexobj(A.data)
    colptr = (A.indptr.astype(np.int64) + 1)
    rowval = (A.indices.astype(np.int64) + 1)
    nzval = np.ascontiguousarray(A.data, dtype=np.complex128 if cplx else np.float64)
    x = np.ascontiguousarray(X2)
    y = np.zeros_like(x)
    info = np.zeros(3, dtype=np.int64)
    emu.emu_horner_fused.restype = C.c_int
    oob = emu.emu_horner_fused(
        C.c_int64(A.shape[0]), C.c_void_p(colptr.ctypes.data), C.c_void_p(rowval.ctypes.data),
        C.c_void_p(nzval.ctypes.data), C.c_int(1), C.c_int(int(cplx)), C.c_int(b),
        C.c_int64(ncell), C.c_int(order), C.c_int(nsub), C.c_double(dt), C.c_int(nv),
        C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data), C.c_void_p(info.ctypes.data))
    return y, oob, info


@pytest.mark.parametrize("nchi,neta,w", [(6, 4, [0.0, 1.0, 2.5]), (5, 3, [0.3, 0.7]),
                                         (9, 7, list(np.linspace(0.1, 4.0, 9)))])
def test_emulated_fused_horner_sweep(emu, orc, nchi, neta, w):
    """The Horner sweep with the replicated rows fused into the term kernel
    (slice tables in the epilogue, column sums of each new w for the next
    term, two sum buffers alternating across terms and substeps) stays in
    bounds and gives the oracle's POWER_METHOD! on the extended operator."""
    base = ops.stencil_operator(nchi, neta, 4, seed=9)
    L_ = ops.receptivity_extension(base, w)
    A = L_.to_scipy_varmajor()
    r = np.random.default_rng(6)
    x = r.standard_normal(L_.n) + 1j * r.standard_normal(L_.n)
    order, nsub, dt = 5, 3, 0.05
    y, oob, info = _run_fused(emu, A, 4, base.ncell, x, order, nsub, dt)
    assert oob == 0
    # one table entry per entry of B (4 nchi freestream columns)
    nnz_B = A[:base.n, base.n:base.n + 4 * nchi].nnz
    assert info[0] == 1 and info[2] == nnz_B
    ref = orc.power_method(orc.CSC.from_scipy(A).scaled(dt), x, nsub, order)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("nv", [8, 16, 32, 64])
def test_emulated_fused_horner_batched(emu, orc, nv):
    """The fused sweep in the b = 4 batch kernel's lane layout (4 cells per
    wave, each lane a row and a run of its doubles): nv / 2 complex vectors,
    every column against the oracle's POWER_METHOD!, in bounds."""
    base = ops.stencil_operator(9, 7, 4, seed=12)
    L_ = ops.receptivity_extension(base, list(np.linspace(0.2, 5.0, 6)))
    A = L_.to_scipy_varmajor()
    r = np.random.default_rng(nv)
    X = r.standard_normal((L_.n, nv // 2)) + 1j * r.standard_normal((L_.n, nv // 2))
    order, nsub, dt = 4, 2, 0.05
    Y, oob, info = _run_fused(emu, A, 4, base.ncell, X, order, nsub, dt, nv=nv)
    assert oob == 0 and info[0] == 1
    Ao = orc.CSC.from_scipy(A).scaled(dt)
    for v in (0, nv // 2 - 1):
        ref = orc.power_method(Ao, X[:, v], nsub, order)
        assert np.linalg.norm(Y[:, v] - ref) <= 1e-12 * np.linalg.norm(ref), v


def test_emulated_fused_horner_real_values(emu, orc):
    """Real extra values (two real vectors per row): the same sweep with the
    freestream diagonal -w made real."""
    base = ops.stencil_operator(7, 5, 4, seed=11)
    L_ = ops.receptivity_extension(base, [0.4, 1.1, 2.0, 3.3])
    A = L_.to_scipy_varmajor()
    A = (A.real + A.imag).tocsc()
    X = np.random.default_rng(7).standard_normal((L_.n, 2))
    y, oob, info = _run_fused(emu, A, 4, base.ncell, X, 4, 2, 0.1)
    assert oob == 0 and info[0] == 1
    Ao = orc.CSC.from_scipy(A).scaled(0.1)
    for v in range(2):
        ref = orc.power_method(Ao, X[:, v], 2, 4)
        assert np.linalg.norm(y[:, v] - ref) <= 1e-12 * np.linalg.norm(ref)


def test_emulated_fused_declines_other_extras(emu):
    """Shock DOFs between the flow and freestream rows, or no replicated
    blocks at all: no tables (the library keeps its two-launch term)."""
    base = ops.stencil_operator(6, 4, 4, seed=9, n_shock=6)
    L_ = ops.receptivity_extension(base, [0.2, 0.9, 1.7])
    xs = np.zeros(L_.n, dtype=np.complex128)
    _, oob, info = _run_fused(emu, L_.to_scipy_varmajor(), 4, base.ncell, xs, 2, 1, 0.1)
    assert oob == 0 and info[0] == 0
    op = ops.stencil_operator(6, 4, 4, seed=9)
    x = np.zeros((op.n, 2))
    _, oob, info = _run_fused(emu, op.to_scipy_varmajor(), 4, op.ncell, x, 2, 1, 0.1)
    assert oob == 0 and info[0] == 0


@pytest.mark.parametrize("b,R,nrhs", [(4, 16, 16), (4, 16, 11), (4, 32, 32), (4, 64, 40),
                                      (8, 8, 8), (8, 16, 13), (8, 32, 32), (8, 64, 64)])
def test_emulated_batched_spmm(emu, b, R, nrhs):
    emu.emu_spmm_csc.restype = C.c_int
    emu.emu_spmm_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [C.c_int, C.c_int64, C.c_int,
                                                                   C.c_int, C.c_void_p, C.c_void_p]
    op = ops.stencil_operator(11, 7, b, seed=9)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(2).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data, b,
                           op.ncell, nrhs, R, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("b,R,nrhs,vpl", [(4, 16, 16, 2), (4, 32, 20, 4), (8, 16, 16, 2),
                                          (8, 16, 9, 4), (8, 32, 32, 2), (8, 64, 50, 4),
                                          (4, 64, 64, 2)])
def test_emulated_batched_mfma(emu, b, R, nrhs, vpl):
    emu.emu_spmm_mfma_csc.restype = C.c_int
    emu.emu_spmm_mfma_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    op = ops.stencil_operator(9, 7, b, seed=10)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(3).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_mfma_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data,
                                b, op.ncell, nrhs, R, vpl, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("b,R,nrhs,vpl", [(4, 8, 8, 2), (4, 8, 5, 4), (8, 8, 8, 2), (8, 8, 3, 4),
                                          (4, 16, 16, 2), (8, 16, 16, 4), (8, 16, 11, 2),
                                          (8, 32, 32, 2), (4, 32, 20, 4), (8, 64, 50, 4),
                                          (4, 64, 64, 2)])
@pytest.mark.parametrize("wide", [False, True])
def test_emulated_batched_tile(emu, b, R, nrhs, vpl, wide):
    """Tile layout path (permutations, value relayout, lane operand fetches,
    MFMA placement, epilogue) replayed on the CPU with every access checked,
    against scipy; also on rows of 20 blocks with ragged slice widths."""
    emu.emu_spmm_tile_csc.restype = C.c_int
    emu.emu_spmm_tile_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if wide:
        from test_gpu_parity import random_pattern_operator
        op = random_pattern_operator(9, 7, b, 19, seed=b + R)
    else:
        op = ops.stencil_operator(9, 7, b, seed=10)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(3).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_tile_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data,
                                b, op.ncell, nrhs, R, vpl, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("R,nrhs,vpl", [(8, 8, 2), (8, 5, 4), (16, 16, 4), (16, 13, 2),
                                        (32, 32, 2), (64, 40, 4)])
@pytest.mark.parametrize("wide", [False, True])
def test_emulated_batched_quad(emu, R, nrhs, vpl, wide):
    """4x4x4 matrix-core path at b = 8 (quad value order, lane operand
    placement, the R = 8 cross-lane k-chunk sum) replayed on the CPU with
    every access checked, against scipy."""
    emu.emu_spmm_quad_csc.restype = C.c_int
    emu.emu_spmm_quad_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if wide:
        from test_gpu_parity import random_pattern_operator
        op = random_pattern_operator(9, 7, 8, 19, seed=R)
    else:
        op = ops.stencil_operator(9, 7, 8, seed=12)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(4).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_quad_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data,
                                op.ncell, nrhs, R, vpl, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("b,seed", [(4, 1), (8, 2)])
def test_emulated_value_cast_map(emu, b, seed):
    """The on-device FP64 -> FP32 value rewrite (hymor_op_set_value_prec)
    lands every value where an FP32 upload puts it."""
    emu.emu_vals_to_f32_csc.restype = C.c_int
    emu.emu_vals_to_f32_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [C.c_int, C.c_int64]
    op = ops.stencil_operator(11, 9, b, seed=seed, n_shock=11)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    assert emu.emu_vals_to_f32_csc(op.n, colptr.ctypes.data, rowval.ctypes.data,
                                   nzval.ctypes.data, b, op.ncell) == 0


@pytest.mark.parametrize("b", [4, 8])
@pytest.mark.parametrize("vpl", [2, 4])
@pytest.mark.parametrize("transpose", [0, 1])
def test_emulated_device_value_layout(emu, b, vpl, transpose):
    """The operator upload places the block values on the device
    (k_slice_vals over the host's slice map, transposed for the adjoint):
    same structure and values as the host layout it replaces."""
    emu.emu_slice_vals_csc.restype = C.c_int
    emu.emu_slice_vals_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int, C.c_int64, C.c_int, C.c_int]
    from test_gpu_parity import random_pattern_operator
    for op in (ops.stencil_operator(11, 9, b, seed=b, n_shock=11),
               random_pattern_operator(9, 7, b, 19, seed=vpl)):
        colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
        assert emu.emu_slice_vals_csc(op.n, colptr.ctypes.data, rowval.ctypes.data,
                                      nzval.ctypes.data, b, op.ncell, vpl, transpose) == 0


@pytest.mark.parametrize("R,nrhs,vpl", [(8, 8, 2), (8, 5, 4), (16, 16, 4), (16, 13, 2),
                                        (32, 32, 2), (64, 40, 4)])
@pytest.mark.parametrize("wide", [False, True])
def test_emulated_batched_quad4(emu, R, nrhs, vpl, wide):
    """4x4x4 matrix-core path at b = 4 (values read in the lane-interleaved
    order, row-major batch, lane operand placement) replayed on the CPU with
    every access checked, against scipy."""
    emu.emu_spmm_quad4_csc.restype = C.c_int
    emu.emu_spmm_quad4_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if wide:
        from test_gpu_parity import random_pattern_operator
        op = random_pattern_operator(9, 7, 4, 19, seed=R)
    else:
        op = ops.stencil_operator(9, 7, 4, seed=12)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(4).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_quad4_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data,
                                 op.ncell, nrhs, R, vpl, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("R,W,nrhs,vpl", [(16, 9, 16, 4), (16, 5, 13, 2), (32, 5, 32, 4),
                                          (32, 4, 27, 2), (64, 2, 64, 4), (64, 1, 40, 2),
                                          (8, 9, 8, 4), (8, 5, 7, 2)])
@pytest.mark.parametrize("wide", [False, True])
def test_emulated_batched_quad4c(emu, R, W, nrhs, vpl, wide):
    """Four cells per wave on the 4x4x4 matrix cores at b = 4 (block = cell,
    column distribution by lane permute, slot groups of W, a partial last
    group of cells) replayed on the CPU with every access checked, against
    scipy."""
    emu.emu_spmm_quad4c_csc.restype = C.c_int
    emu.emu_spmm_quad4c_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if wide:
        from test_gpu_parity import random_pattern_operator
        op = random_pattern_operator(9, 7, 4, 19, seed=R + W)
    else:
        op = ops.stencil_operator(9, 7, 4, seed=13)   # 63 cells: last group of 3
    assert op.ncell % 4 != 0
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(5).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_quad4c_csc(op.n, colptr.ctypes.data, rowval.ctypes.data,
                                  nzval.ctypes.data, op.ncell, nrhs, R, W, vpl, X.ctypes.data,
                                  Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("NV,NW", [(1, 4), (2, 4), (8, 4), (32, 4), (64, 4), (4, 1), (16, 2)])
@pytest.mark.parametrize("cplx", [False, True])
def test_emulated_long_extra_rows(emu, NV, NW, cplx):
    """k_extra's lane mapping (NW waves per long row, lanes per entry, xor
    butterfly, waves combined in order) replayed on the CPU with every
    access checked: the long rows of the shock bands (real) and of the
    receptivity coupling (complex) against scipy on the same rows."""
    if cplx and NV < 2:
        pytest.skip("complex rows need two doubles")
    emu.emu_extra_csc.restype = C.c_int
    emu.emu_extra_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if cplx:
        base = ops.stencil_operator(6, 4, 4, seed=3)
        op = ops.receptivity_extension(base, np.array([0.5, 1.0, 2.0]))
    else:
        op = ops.stencil_operator(12, 10, 4, seed=2, n_shock=12)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    nrhs = NV // 2 if cplx else NV
    rng = np.random.default_rng(NV)
    Xvm = rng.standard_normal((op.n, nrhs))
    if cplx:
        Xvm = Xvm + 1j * rng.standard_normal((op.n, nrhs))
    # rows in the internal order: flow rows cell-major, extra DOFs after
    b, ncell = op.b, op.ncell
    nflow = b * ncell
    perm = np.arange(op.n)
    r = np.arange(nflow)
    perm[:nflow] = (r % b) * ncell + r // b        # internal row -> var-major index
    Xi = Xvm[perm]
    x = np.ascontiguousarray(Xi.view(np.float64) if cplx else Xi)
    y = np.zeros_like(x)
    oob = emu.emu_extra_csc(op.n, colptr.ctypes.data, rowval.ctypes.data,
                            np.ascontiguousarray(nzval).ctypes.data, int(cplx), b, ncell, NV, NW,
                            x.ctypes.data, y.ctypes.data)
    assert oob == 0
    # the extra entries: every entry outside the flow-flow block part
    A = op.to_scipy_varmajor().tolil()
    A[:nflow, :nflow] = 0
    want = (A.tocsr() @ Xvm)[perm]
    got = y.view(np.complex128) if cplx else y
    assert np.abs(want).max() > 0
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("R,W,nrhs,vpl", [(16, 12, 16, 4), (16, 5, 11, 2), (32, 6, 32, 4),
                                          (64, 6, 50, 2)])
@pytest.mark.parametrize("shape", ["stencil_odd", "wide", "single"])
def test_emulated_batched_pair(emu, R, W, nrhs, vpl, shape):
    """Two cells per matrix-core product at b = 8 (pair map of the host code,
    zero rows where a cell lacks a column, the pair's second cell on rows
    8-15, slot groups of W) replayed on the CPU with every access checked,
    against scipy; an odd cell count leaves the last pair with one cell."""
    emu.emu_spmm_pair_csc.restype = C.c_int
    emu.emu_spmm_pair_csc.argtypes = [C.c_int64] + [C.c_void_p] * 3 + [
        C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    if shape == "wide":
        from test_gpu_parity import random_pattern_operator
        op = random_pattern_operator(9, 7, 8, 19, seed=R + W)
    elif shape == "single":
        op = ops.stencil_operator(1, 1, 8, seed=2)
    else:
        op = ops.stencil_operator(9, 7, 8, seed=14)        # 63 cells
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=True)
    X = np.asfortranarray(np.random.default_rng(6).standard_normal((op.n, nrhs)))
    Y = np.zeros_like(X)
    oob = emu.emu_spmm_pair_csc(op.n, colptr.ctypes.data, rowval.ctypes.data, nzval.ctypes.data,
                                op.ncell, nrhs, R, W, vpl, X.ctypes.data, Y.ctypes.data)
    assert oob == 0
    assert np.allclose(Y, op.to_scipy_varmajor() @ X, rtol=1e-13, atol=1e-13)
