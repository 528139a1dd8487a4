"""PCG monitor log (the scalar fields of the reference's CGMonitor, ch_sparse_krylov.hpp:47-58, called
at 547-555, 569-579, 612-621 and 636-648): the oracle's records against the reference's own
solve_preconditioned_cg with a recording monitor, bit for bit, on the reference's assembled systems
for every boundary kind, both preconditioners and the split operator; and checks that do not need
the reference (record count and indices, the converged and non-converged last records, the
relations between alpha, beta and the residual fields)."""
import numpy as np
import pytest

from ding_nsch_core_solver_amd import workloads as W
from tests.cases import small_cases

CASES = small_cases()


def _ids(cases):
    return [name for name, _ in cases]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("name,case", CASES, ids=_ids(CASES))
def test_oracle_records_match_reference_bitwise(oracle_lib, ref_lib, name, case, kind, split):
    o = oracle_lib.OracleSystem(case, use_split=split)
    r = ref_lib.RefSystem(case, use_split=split)
    ok_o, ok_r = o.factor(kind)[0], r.factor(kind)[0]
    assert ok_o == ok_r
    if not ok_o:
        pytest.skip("factorization fails for this system (both sides agree)")
    got = o.pcg_monitored(case.max_iterations, case.tolerance)
    expect = r.pcg_monitored(case.max_iterations, case.tolerance)
    assert got[5] == expect[5]
    assert got[1] == expect[1]
    assert len(got[4]) == len(expect[4])
    assert got[4].tobytes() == expect[4].tobytes()


def _records(oracle_lib, case, kind=1, max_iterations=None, tol=None):
    s = oracle_lib.OracleSystem(case)
    assert s.factor(kind)[0]
    x, it, rel, ab, rec, err = s.pcg_monitored(max_iterations or case.max_iterations, tol or case.tolerance)
    assert err is None
    return s, it, rel, ab, rec


@pytest.mark.parametrize("name,case", CASES[:6], ids=_ids(CASES[:6]))
def test_record_invariants(oracle_lib, name, case):
    s, it, rel, ab, rec = _records(oracle_lib, case)
    assert len(rec) == it + 1
    assert np.array_equal(rec["iteration"], np.arange(it + 1))
    bnorm = np.sqrt(np.dot(s.rhs, s.rhs))
    # record 0: initial residual, no alpha/beta, preconditioned residual sqrt(r.z) > 0
    assert rec[0]["alpha"] == 0.0 and rec[0]["beta"] == 0.0 and rec[0]["iterate_residual"] == 0.0
    assert rec[0]["preconditioned_residual"] > 0.0
    assert np.isclose(rec[0]["relative_residual"], rec[0]["absolute_residual"] / bnorm, rtol=1e-13, atol=0)
    # converged last record: alpha set, beta and preconditioned residual left at 0, residual below tol
    last = rec[-1]
    assert last["alpha"] > 0.0 and last["beta"] == 0.0 and last["preconditioned_residual"] == 0.0
    assert last["relative_residual"] == rel and last["absolute_residual"] == ab
    assert rel < case.tolerance or ab < case.tolerance
    # regular records: beta_k = (r.z)_k / (r.z)_{k-1} = (prec_k / prec_{k-1})^2, residuals above tol
    mid = rec[1:-1]
    if len(mid):
        prev = rec[:-2]["preconditioned_residual"]
        assert np.allclose(mid["beta"], (mid["preconditioned_residual"] / prev) ** 2, rtol=1e-10, atol=0)
        assert np.all(mid["alpha"] > 0.0) and np.all(mid["beta"] > 0.0)
        assert np.all((mid["relative_residual"] >= case.tolerance) & (mid["absolute_residual"] >= case.tolerance))
        assert np.allclose(mid["relative_residual"], mid["absolute_residual"] / bnorm, rtol=1e-13, atol=0)


def test_alpha_is_the_cg_step_length(oracle_lib):
    """alpha_1 = r0.z0 / p0.A p0 with p0 = z0 = M^-1 r0, from the oracle's own operator pieces."""
    case = W.droplet_case(10, 8, seed=5, density_ratio=20.0, bc_type=(W.BC_DIRICHLET, 0, 0, 0),
                          bc_value=(1.0, 0.0, 0.0, 0.0))
    s, it, rel, ab, rec = _records(oracle_lib, case)
    r0 = s.rhs.copy()
    z0 = s.apply_precond(r0)
    alpha1 = np.dot(r0, z0) / np.dot(z0, s.apply_matrix(z0))
    assert np.isclose(rec[1]["alpha"], alpha1, rtol=1e-12, atol=0)
    assert np.isclose(rec[0]["preconditioned_residual"], np.sqrt(np.dot(r0, z0)), rtol=1e-12, atol=0)
    assert np.isclose(rec[0]["absolute_residual"], np.linalg.norm(r0), rtol=1e-13, atol=0)


def test_zero_rhs_logs_one_record(oracle_lib, ref_lib):
    """Converged before the first iteration: a single record with the initial residual only."""
    case = W.droplet_case(8, 6, seed=0, density_ratio=10.0)
    case.u = np.zeros_like(case.u)
    case.v = np.zeros_like(case.v)
    o = oracle_lib.OracleSystem(case)
    r = ref_lib.RefSystem(case)
    assert o.factor(1)[0] and r.factor(1)[0]
    got = o.pcg_monitored(case.max_iterations, case.tolerance)
    expect = r.pcg_monitored(case.max_iterations, case.tolerance)
    assert got[1] == 0 and len(got[4]) == 1
    assert got[4].tobytes() == expect[4].tobytes()
    assert got[4][0]["preconditioned_residual"] == 0.0


def test_iteration_limit_ends_on_a_regular_record(oracle_lib, ref_lib):
    """Stopped by the iteration limit: every record after the first carries beta (no converged
    record), as the reference's loop ends without the converged branch."""
    case = W.droplet_case(24, 20, seed=2, density_ratio=500.0)
    o = oracle_lib.OracleSystem(case)
    r = ref_lib.RefSystem(case)
    assert o.factor(1)[0] and r.factor(1)[0]
    got = o.pcg_monitored(3, 1e-14)
    expect = r.pcg_monitored(3, 1e-14)
    assert got[1] == 3 and len(got[4]) == 4
    assert np.all(got[4][1:]["beta"] > 0.0)
    assert got[4].tobytes() == expect[4].tobytes()


def test_capacity_truncates_and_counts(oracle_lib):
    case = W.droplet_case(12, 12, seed=1, density_ratio=50.0)
    p, it, res, full = oracle_lib.pressure_solve_monitored("icpcg", case)
    p2, it2, res2, part = oracle_lib.pressure_solve_monitored("icpcg", case, capacity=4)
    assert len(full) == it + 1 and len(part) == 4
    assert part.tobytes() == full[:4].tobytes()
    assert np.array_equal(p, p2) and it == it2


def test_monitored_solve_leaves_the_result_unchanged(oracle_lib):
    for scheme in ("icpcg", "ildlt_pcg"):
        for name, case in CASES[:4]:
            p, it, res = oracle_lib.pressure_solve(scheme, case, oracle_lib.CANONICAL)
            pm, itm, resm, rec = oracle_lib.pressure_solve_monitored(scheme, case, oracle_lib.CANONICAL)
            assert np.array_equal(p, pm) and it == itm and res == resm
            assert len(rec) == it + 1 and rec[-1]["relative_residual"] == res
