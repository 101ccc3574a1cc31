"""Host checks of the extended receptivity operator's structure (no GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

from hymor_amd import operators as ops
from hymor_amd.receptivity import split_extended


def _extended(w):
    base = ops.stencil_operator(5, 4, 4, seed=41)
    return base, ops.receptivity_extension(base, w).to_scipy_varmajor()


def test_split_extended_recovers_L_and_B():
    w = [0.0, 1.0, 2.5]
    base, A_ = _extended(w)
    p = 4 * base.nchi
    L, B = split_extended(A_, base.n, p, w)
    assert L.shape == (base.n, base.n) and B.shape == (base.n, p)
    assert abs(L - A_[:base.n, :base.n]).max() == 0
    assert abs(B - A_[:base.n, base.n + 2 * p:base.n + 3 * p]).max() == 0


@pytest.mark.parametrize("where", ["diagonal", "off_diagonal"])
def test_split_extended_rejects_a_wrong_freestream_block(where):
    w = [0.0, 1.0]
    base, A_ = _extended(w)
    A_ = A_.tolil()
    i = base.n + 1
    if where == "diagonal":
        A_[i, i] = 7.0                   # not -i w
    else:
        A_[i, i + 3] = 1e-3              # couples two freestream DOFs
    with pytest.raises(ValueError, match="freestream block"):
        split_extended(sp.csc_matrix(A_), base.n, 4 * base.nchi, w)
