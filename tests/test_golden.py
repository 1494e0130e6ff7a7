"""The oracle against vectors written by the reference itself (tests/golden/make_golden.py):
G2 = Newton-Krylov trajectories, G3 = the reference's torch.func implicit-function adjoint."""
import numpy as np
import pytest
import torch

from oracle import fem
from helpers import GOLDEN_SOLVES, check_solution, check_trajectory, golden_max_newton, golden_problem, \
    history_array, load_golden as _load


def _params(spec):
    return fem.precompute_geometry(golden_problem(spec))


@pytest.mark.parametrize("name", GOLDEN_SOLVES)
def test_oracle_reproduces_reference_newton_trajectory(name):
    z, spec = _load(name)
    p = _params(spec)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    records = []
    u, info = fem.newton_krylov_solve(p, th, int(z["n_load_steps"]), float(z["tol"]), golden_max_newton(z),
                                      callback=records.append)
    assert info["converged"] == bool(z["converged"])
    assert info["n_cutbacks"] == int(z["n_cutbacks"])
    if "load_factor" in z.files:
        assert info["load_factor"] == float(z["load_factor"])
        assert np.array_equal([lf for lf, _ in info["snapshots"]], z["snapshot_lf"])
    assert records == info["history"]
    check_solution(u, z)
    check_trajectory(history_array(info), z)
    assert info["total_newton"] == int(z["total_newton"])
    assert abs(info["total_cg"] - int(z["total_cg"])) <= 0.02 * int(z["total_cg"])


def test_cutback_goldens_exercise_the_branches():
    """The fixtures hold what they are for: NaN residuals and a NaN-driven cut-back, cut-backs after
    max_newton with backtracking steps, and a give-up at load factor 0."""
    z, _ = _load("g2_q4_nan_cutback_4")
    h = z["history"]
    assert int(z["n_cutbacks"]) >= 1 and np.isnan(h[:, 3]).any() and bool(z["converged"])
    assert np.all(h[np.isnan(h[:, 3]), 4] == 500)          # NaN PCG runs to the cap, as the reference's does
    z, _ = _load("g2_q4_maxnewton_cutback")
    h = z["history"]
    assert int(z["n_cutbacks"]) >= 1 and (h[:, 5] < 1).any() and bool(z["converged"])
    z, _ = _load("g2_q4_giveup")
    assert not bool(z["converged"]) and float(z["load_factor"]) == 0.0 and int(z["n_cutbacks"]) == 9


def test_oracle_adjoint_matches_reference_adjoint():
    z, spec = _load("g3_q4_compliance_16")
    p = _params(spec)
    th = torch.from_numpy(z["theta"])
    u, g = fem.compute_sensitivity(p, th, lambda w: torch.dot(p["f"], w), n_load_steps=int(z["n_load_steps"]),
                                   tol=float(z["tol"]))
    ref = torch.from_numpy(z["grad"])
    assert float((u - torch.from_numpy(z["u"])).abs().max() / torch.from_numpy(z["u"]).abs().max()) < 1e-10
    assert float((g - ref).abs().max() / ref.abs().max()) < 1e-9
