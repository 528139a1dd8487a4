"""The oracle on pressure-solve states captured from real timesteps of the reference's own runs:
run in the reference's sequential order it reproduces the reference's solve on each state bit for
bit (pressure_correction_ interior by SHA-256, iteration count, relative residual), for the run's
icpcg scheme and for split_icpcg with the state's real pressure history. The 128x128 state is the
published validation run one step past its step-5000 snapshot, whose history records 133 ICC-PCG
iterations on every step."""
import pytest

from tests.timestep_states import load_state, sha, state_files, state_id

FILES = state_files()


def test_states_present():
    ids = {state_id(p) for p in FILES}
    assert {f"static_droplet_64_step{k}" for k in range(10)} <= ids
    assert {"rising_bubble_192x384_step0", "rising_bubble_192x384_step17",
            "static_droplet_128_restart_step5000"} <= ids


@pytest.mark.parametrize("path", FILES, ids=[state_id(p) for p in FILES])
@pytest.mark.parametrize("scheme", ["icpcg", "split_icpcg"])
def test_oracle_reproduces_reference_on_timestep_state(oracle_lib, path, scheme):
    case, results = load_state(path)
    p, it, res = oracle_lib.pressure_solve(scheme, case, oracle_lib.SEQUENTIAL)
    it_ref, res_ref, sha_ref = results[scheme]
    assert it == it_ref and res == res_ref and sha(p) == sha_ref


def test_restart_state_keeps_the_published_iteration_count():
    """history.csv of the published run: 133 iterations on every step, residuals 8.82e-11..9.93e-11."""
    path = next(p for p in FILES if "128_restart" in p)
    _, results = load_state(path)
    it, res, _ = results["icpcg"]
    assert it == 133 and 8.8e-11 <= res <= 9.93e-11
