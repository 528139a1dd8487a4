"""The oracle's Cahn-Hilliard linear solve (pp_oracle.cpp, or_ch_*) against the reference Solver's
own solve_phase_linear_system_eq25 compiled from its sources: the operator's CSR, the ILDLT(0)
factor (or its failure, and then the diagonal preconditioner), the solution, iteration count and
iterate residual, bit for bit. Independent checks: the operator is symmetric with the pattern of
the 13-point stencil, A 1 = alpha0 1 on periodic grids (Lap 1 = 0), and the ILDLT(0) factor
reproduces A on the pattern when the stencil has no fill across the dropped entries' columns."""
import numpy as np
import pytest

from ding_nsch_core_solver_amd import workloads as W
from tests.ch_cases import small_ch_cases

CASES = small_ch_cases()


@pytest.mark.parametrize("name,cc", CASES, ids=[c[0] for c in CASES])
def test_oracle_equals_reference(oracle_lib, ref_lib, name, cc):
    R = ref_lib.RefCh(cc)
    c_ref, it_ref, itr_ref, _, solver_name = R.solve(cc.alpha0, cc.rhs, cc.c0)
    rp, ci, val, kind, lower, D = R.operator()
    A = oracle_lib.ch_operator(cc)
    assert np.array_equal(rp, A.row_ptr) and np.array_equal(ci, A.col_idx) and np.array_equal(val, A.values)
    ok, lo, Do, _ = oracle_lib.ch_ildlt(cc)
    assert ok == (kind == 1)
    if ok:
        assert np.array_equal(lower, lo) and np.array_equal(D, Do)
    c, it, _, _, itr, pre = oracle_lib.ch_solve(cc, cc.rhs, cc.c0)
    assert np.array_equal(c, c_ref) and it == it_ref and itr == itr_ref
    assert solver_name == {1: "SparsePCG[ILDLT]", 0: "SparsePCG[Diagonal]", 2: "SparsePCG[DiagonalFallback]"}[pre]


def _dense(A):
    M = np.zeros((A.n, A.n))
    for r in range(A.n):
        M[r, A.col_idx[A.row_ptr[r]:A.row_ptr[r + 1]]] = A.values[A.row_ptr[r]:A.row_ptr[r + 1]]
    return M


@pytest.mark.parametrize("px,py", [(False, False), (True, False), (False, True), (True, True)])
def test_operator_closed_forms(oracle_lib, px, py):
    cc = W.ch_case(9, 8, seed=0, periodic_x=px, periodic_y=py)
    A = oracle_lib.ch_operator(cc)
    M = _dense(A)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    assert max(np.diff(A.row_ptr)) <= 13
    if px and py:
        assert np.allclose(M @ np.ones(A.n), cc.alpha0, rtol=1e-12)
    # the matrix is alpha0 I - beta1 L + beta2 L^2 with the 5-point Laplacian L
    n, nx, ny = A.n, cc.nx, cc.ny
    Lm = np.zeros((n, n))
    hx, hy = 1.0 / (cc.lx / nx) ** 2, 1.0 / (cc.ly / ny) ** 2
    for i in range(nx):
        for j in range(ny):
            r = i * ny + j
            for di, dj, h in ((-1, 0, hx), (1, 0, hx), (0, -1, hy), (0, 1, hy)):
                ii, jj = i + di, j + dj
                if px:
                    ii %= nx
                if py:
                    jj %= ny
                if 0 <= ii < nx and 0 <= jj < ny:
                    Lm[r, ii * ny + jj] += h
                    Lm[r, r] -= h
    expect = cc.alpha0 * np.eye(n) - cc.beta1 * Lm + cc.beta2 * Lm @ Lm
    assert np.abs(M - expect).max() <= 1e-12 * np.abs(expect).max()


def test_ildlt_reproduces_the_operator_on_its_pattern_where_exact(oracle_lib):
    """ILDLT(0) keeps A's pattern: (L D L^T)_ij = a_ij for every stored (i, j) the cross terms reach."""
    cc = W.ch_case(12, 10, seed=1)
    A = oracle_lib.ch_operator(cc)
    ok, lower, D, _ = oracle_lib.ch_ildlt(cc)
    assert ok
    n = A.n
    Lf = np.eye(n)
    for r in range(n):
        for p in range(A.row_ptr[r], A.row_ptr[r + 1]):
            if A.col_idx[p] < r:
                Lf[r, A.col_idx[p]] = lower[p]
    P = Lf @ np.diag(D) @ Lf.T
    M = _dense(A)
    mask = M != 0.0
    assert np.abs((P - M)[mask]).max() <= 1e-9 * np.abs(M).max()
