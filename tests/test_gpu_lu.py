"""The library's LU factorisation on arbitrary matrices (rf_selftest_lu)
against LAPACK getrf (scipy.linalg.lu_factor), the factorisation of the
reference's dgesv (fortran/coeffs_and_vels.f90:62,158).  Fiber systems never
pivot, so these matrices are what exercises the row swaps: random ones pivot
in almost every column; column-diagonally-dominant ones with rows exchanged
inside one panel pivot only inside that panel's diagonal block."""
import numpy as np
import pytest
import scipy.linalg as sl

import rigidfibers_kth_amd as rf

pytestmark = pytest.mark.gpu


def _random(n, seed):
    return np.random.default_rng(seed).normal(size=(n, n))


def _dominant_with_block_swaps(n, nb, seed):
    """Column-diagonally-dominant (partial pivoting keeps the diagonal) with
    rows exchanged inside the diagonal blocks of width nb."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, size=(n, n)) / n
    A[np.diag_indices(n)] = 2.0 + rng.uniform(size=n)
    perm = np.arange(n)
    for k0 in range(0, n, nb):
        k1 = min(n, k0 + nb)
        if k1 - k0 >= 2:
            i, j = rng.choice(np.arange(k0, k1), 2, replace=False)
            perm[[i, j]] = perm[[j, i]]
    return A[perm]


def _residual(A, LU, ipiv):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for k, p in enumerate(ipiv):
        PA[[k, p]] = PA[[p, k]]
    return np.abs(PA - L @ U).max() / np.abs(A).max()


@pytest.mark.parametrize("n,nb,valu", [(300, 0, 0), (300, 13, 0), (257, 64, 1), (130, 1, 0)])
def test_column_panel_matches_lapack_on_random_matrices(n, nb, valu):
    A = _random(n, n + nb)
    LU, ipiv, info, fb = rf.selftest_lu(A, 8, nb, column_panel=1, valu=valu)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert info == 0 and fb == 0
    assert np.sum(ipiv != np.arange(n)) > n // 2  # the matrix pivots
    np.testing.assert_array_equal(ipiv, piv_ref)
    assert np.abs(LU - lu_ref).max() <= 1e-11 * np.abs(lu_ref).max()


@pytest.mark.parametrize("n,nb,es", [(300, 0, 8), (301, 40, 8), (300, 0, 4)])
def test_block_panel_redoes_panels_that_pivot_outside_the_block(n, nb, es):
    """Random matrices fail the check in every panel with rows below it:
    each is restored from its copy and redone over whole columns, and the
    result is partial pivoting's."""
    A = _random(n, 7)
    if es == 4:
        A = A.astype(np.float32).astype(np.float64)
    LU, ipiv, info, fb = rf.selftest_lu(A, es, nb, column_panel=0)
    width = nb if nb else 64
    assert info == 0 and fb == (n - 1) // width
    lu_ref, piv_ref = sl.lu_factor(A)
    if es == 8:
        np.testing.assert_array_equal(ipiv, piv_ref)
        assert np.abs(LU - lu_ref).max() <= 1e-11 * np.abs(lu_ref).max()
    else:
        assert np.mean(ipiv == piv_ref) >= 0.95
        assert _residual(A, LU, ipiv) <= 1e-5


@pytest.mark.parametrize("n,nb,es", [(300, 0, 8), (301, 40, 8), (250, 13, 8), (515, 0, 4), (300, 100, 4),
                                     (200, 64, 4)])
def test_block_panel_equals_partial_pivoting_when_pivots_stay_in_the_block(n, nb, es):
    width = nb if nb else 64
    A = _dominant_with_block_swaps(n, width, 3 + n)
    if es == 4:
        A = A.astype(np.float32).astype(np.float64)
    lu_ref, piv_ref = sl.lu_factor(A)
    assert np.sum(piv_ref != np.arange(n)) > 0
    for column_panel in (0, 1):
        LU, ipiv, info, fb = rf.selftest_lu(A, es, nb, column_panel=column_panel)
        assert info == 0 and fb == 0, column_panel
        np.testing.assert_array_equal(ipiv, piv_ref)
        tol = 1e-13 if es == 8 else 2e-6
        assert np.abs(LU - lu_ref).max() <= tol * np.abs(lu_ref).max(), column_panel
        assert _residual(A, LU, ipiv) <= tol


def test_fp32_column_panel_factors_random_matrix():
    n = 300
    A = _random(n, 11).astype(np.float32).astype(np.float64)
    LU, ipiv, info, fb = rf.selftest_lu(A, 4, 0, column_panel=1)
    assert info == 0 and fb == 0
    assert _residual(A, LU, ipiv) <= 1e-5
    # pivots as LAPACK's wherever the FP32 rounding leaves no near-tie
    _, piv_ref = sl.lu_factor(A)
    assert np.mean(ipiv == piv_ref) >= 0.95


def test_zero_column_sets_info():
    n = 100
    A = _random(n, 5)
    A[:, 37] = 0.0
    _, _, info, fb = rf.selftest_lu(A, 8, 0, column_panel=1)
    assert info == 38 and fb == 0  # LAPACK: first k with U(k,k) = 0, 1-based
    _, _, info, fb = rf.selftest_lu(A, 8, 0, column_panel=0)
    assert info == 38 and fb >= 1  # the zero pivot sends its panel to the column search


@pytest.mark.parametrize("n,nb", [(515, 0), (300, 100), (200, 13), (5500, 0)])
def test_fp16_update_factors(n, nb):
    """FP32 factors with fp16-rounded trailing-update operands: the pivots of
    partial pivoting and a backward error of the order of fp16 rounding;
    scaled copies of the matrix (2^-30, 2^30) give the same relative error
    (per-column scaling of U12 keeps fp16 away from underflow and overflow).
    n = 5500 reaches the 256 x 256 tile update (trailing matrix >= 4096 after
    the first 1024-column super-panel)."""
    A = _dominant_with_block_swaps(n, nb if nb else 64, 17 + n)
    A = A.astype(np.float32).astype(np.float64)
    _, piv_ref = sl.lu_factor(A)
    errs = []
    for scale in (1.0, 2.0 ** -30, 2.0 ** 30):
        LU, ipiv, info, fb = rf.selftest_lu(A * scale, 4, nb, column_panel=0, update_fp16=1)
        assert info == 0 and fb == 0
        np.testing.assert_array_equal(ipiv, piv_ref)
        errs.append(_residual(A * scale, LU, ipiv))
    assert max(errs) <= 2e-3 and max(errs) >= 1e-7, errs  # fp16 operands, not FP32
    LU32, _, _, _ = rf.selftest_lu(A, 4, nb, column_panel=0, update_fp16=0)
    assert _residual(A, LU32, piv_ref) <= 1e-5
