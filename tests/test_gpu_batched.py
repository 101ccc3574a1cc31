"""GPU parity of the batched (SpMM) action: every column must equal the
single-vector action, and the FP64 oracle, at FP64 tolerance (1e-12), for
real and complex operators, with and without extra DOFs, and for batch sizes
that need padding to the next supported width."""
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd.core import DeviceMatrix, DeviceOperator, expm_taylor_batched_dev

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("b,nrhs,shock", [(4, 16, False), (4, 7, True), (8, 16, False),
                                          (8, 8, True), (4, 32, False), (8, 64, False),
                                          (8, 3, False), (4, 40, True)])
def test_batched_matches_oracle_columns(orc, b, nrhs, shock):
    say(f"[batched] b={b} nrhs={nrhs} shock={shock}")
    nchi, neta = 13, 9
    op = ops.stencil_operator(nchi, neta, b, seed=11, n_shock=nchi if shock else 0, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    X = np.random.default_rng(1).standard_normal((op.n, nrhs))
    with DeviceOperator.from_blocks(op) as dev:
        Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1)
        Yt = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1, transpose=True)
        for j in (0, nrhs // 2, nrhs - 1):
            assert relerr(Y[:, j], orc.power_method(A.scaled(0.1), X[:, j], 3)) <= 1e-12
            assert relerr(Yt[:, j], dev.expm_taylor(X[:, j], order=5, nsub=3, dt=0.1,
                                                    transpose=True)) <= 1e-13


@pytest.mark.parametrize("nchi,neta", [(1, 1), (3, 1), (2, 3), (5, 1)])
@pytest.mark.parametrize("b,nrhs", [(4, 16), (4, 32), (4, 64), (8, 16), (4, 8)])
def test_batched_tiny_grids(orc, nchi, neta, b, nrhs):
    """Grids of 1-6 cells: at b = 4 the four-cells-per-wave kernel then has
    one partial group of cells (the only wave, or the last one), and the
    single slice of every kernel is partial."""
    say(f"[batched tiny] {nchi}x{neta} b={b} nrhs={nrhs}")
    op = ops.stencil_operator(nchi, neta, b, seed=nchi + 7 * neta, n_shock=nchi, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    X = np.random.default_rng(nrhs).standard_normal((op.n, nrhs))
    with DeviceOperator.from_blocks(op) as dev:
        Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1)
        Yt = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1, transpose=True)
    At = orc.csc_adjoint(A)
    for j in (0, nrhs - 1):
        assert relerr(Y[:, j], orc.power_method(A.scaled(0.1), X[:, j], 3)) <= 1e-12
        assert relerr(Yt[:, j], orc.power_method(At.scaled(0.1), X[:, j], 3)) <= 1e-12


def test_batched_complex_receptivity(orc):
    say("[batched] complex receptivity operator, 12 rhs")
    base = ops.stencil_operator(8, 6, 4, seed=12, n_shock=8, row_sum=2.0)
    L_ = ops.receptivity_extension(base, [0.0, 1.0, 4.0])
    A = orc.CSC.from_arrays(*L_.to_csc_varmajor())
    r = np.random.default_rng(2)
    X = r.standard_normal((L_.n, 12)) + 1j * r.standard_normal((L_.n, 12))
    with DeviceOperator.from_blocks(L_) as dev:
        Y = dev.expm_taylor_batched(X, order=5, nsub=4, dt=0.05)
    for j in (0, 5, 11):
        assert relerr(Y[:, j], orc.power_method(A.scaled(0.05), X[:, j], 4)) <= 1e-12


def test_batched_f32_and_device_path(orc):
    say("[batched] FP32 values, device-resident matrices")
    op = ops.stencil_operator(20, 12, 8, seed=13, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    X = np.random.default_rng(3).standard_normal((op.n, 16))
    with DeviceOperator.from_blocks(op, value_prec="f32") as dev:
        dX = DeviceMatrix.from_host(X)
        dY = DeviceMatrix(op.n, 16)
        expm_taylor_batched_dev(dev, dX, dY, order=12, nsub=8, dt=1 / 64)
        dev.synchronize()
        Y = dY.download()
        assert np.array_equal(Y, dev.expm_taylor_batched(X, order=12, nsub=8, dt=1 / 64))
    ref = orc.power_method(A.scaled(1 / 64), X[:, 5], 8, order=12)
    assert 1e-12 < relerr(Y[:, 5], ref) <= 2e-6


@pytest.mark.parametrize("mfma", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("b,nrhs,prec", [(8, 16, "f64"), (8, 32, "f32"), (4, 16, "f32"),
                                         (4, 64, "f64"), (8, 64, "f32"), (8, 16, "f32")])
def test_batch_kernels_agree_with_oracle(orc, mfma, b, nrhs, prec):
    say(f"[batched] kernel {mfma} b={b} nrhs={nrhs} {prec}")
    # 13 x 9 cells: an odd cell count exercises the partial last wave
    op = ops.stencil_operator(13, 9, b, seed=31, n_shock=13, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    X = np.random.default_rng(5).standard_normal((op.n, nrhs))
    with DeviceOperator.from_blocks(op, value_prec=prec) as dev:
        dev.set_batch_kernel(mfma)
        Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1)
        if nrhs >= 16:
            kern, ms = dev.autotune_batch(nv=nrhs, reps=2)
            assert kern in range(9) and ms > 0
            Y2 = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1)
            assert relerr(Y2, Y) <= 1e-13
    tol = 1e-12 if prec == "f64" else 2e-6
    for j in (0, nrhs - 1):
        assert relerr(Y[:, j], orc.power_method(A.scaled(0.1), X[:, j], 3)) <= tol


@pytest.mark.parametrize("mfma", [1, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("b,nrhs", [(4, 16), (8, 32), (8, 64), (8, 8), (4, 8)])
def test_batch_kernels_wide_rows(orc, mfma, b, nrhs):
    """Batch kernels on rows of up to 20 blocks with ragged widths inside a
    slice (several load groups per cell)."""
    from test_gpu_parity import random_pattern_operator
    say(f"[batched wide] kernel {mfma} b={b} nrhs={nrhs}")
    op = random_pattern_operator(9, 7, b, 19, seed=10 + b)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    X = np.random.default_rng(6).standard_normal((op.n, nrhs))
    with DeviceOperator.from_blocks(op) as dev:
        dev.set_batch_kernel(mfma)
        Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1)
        Yt = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.1, transpose=True)
    At = orc.csc_adjoint(A)
    for j in (0, nrhs - 1):
        assert relerr(Y[:, j], orc.power_method(A.scaled(0.1), X[:, j], 3)) <= 1e-12
        assert relerr(Yt[:, j], orc.power_method(At.scaled(0.1), X[:, j], 3)) <= 1e-12


def test_calibration_probes():
    """The streaming copy and read probes run and report plausible rates
    (between 0.5 TB/s and the 8 TB/s HBM peak) at 256 MiB."""
    from hymor_amd.calibrate import copy_bandwidth, mixed_bandwidth, read_bandwidth
    c = copy_bandwidth(256, 3)
    r = read_bandwidth(256, 3)
    assert 500 < c["GBps"] < 8000 and 500 < r["GBps"] < 8000, (c, r)
    # mixed stream: 16 384 chunks of whole 1 KiB runs; 200 MiB read and
    # 50 MiB written round down to 12 and 3 runs per chunk
    m = mixed_bandwidth(200 << 20, 50 << 20, 3)
    assert m["bytes_moved"] == 16384 * (12 + 3) * 1024, m
    assert 500 < m["GBps"] < 8000, m
