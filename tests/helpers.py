"""Shared helpers: small problems, seeded states and the reference's golden vectors (tests only)."""
import json
import os

import numpy as np
import torch

from difffea_amd.problem import build_problem

DTYPE = torch.float64
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    """(npz, spec) of a golden file written by tests/golden/make_golden.py (no pickles)."""
    z = np.load(os.path.join(GOLD, name + ".npz"))          # allow_pickle=False (default)
    return z, json.loads(str(z["spec"]))


def golden_problem(spec):
    """params of a golden spec, built by this package's set-up code."""
    sp = dict(spec)
    sp.pop("mesher", None)
    return build_problem(sp)


def golden_max_newton(z):
    return int(z["max_newton"]) if "max_newton" in z.files else 25


def history_array(info):
    return np.array([[r["step"], r["load_factor"], r["iteration"], r["residual"], r["cg_iters"], r["alpha"]]
                     for r in info["history"]])


def check_trajectory(hist, z, res_rtol=5e-2):
    """A Newton history against the reference's: step, iteration and line-search alpha identical, load factors
    to 1e-15, CG counts within a few iterations, residual norms (NaN where the reference's is NaN) within
    res_rtol where they are not round-off-sized."""
    ref = z["history"]
    assert hist.shape == ref.shape
    assert np.array_equal(hist[:, [0, 2, 5]], ref[:, [0, 2, 5]])      # step, iteration, line-search alpha
    assert np.allclose(hist[:, 1], ref[:, 1], rtol=0, atol=1e-15)      # load factors
    # Loose inner solves (eta up to 0.5) amplify round-off: on the quad9 case merely switching the oracle's
    # contraction order (qp -> qp_ew) moves the first Newton residual by 0.3 %.  Round-off differences with
    # the reference (geometry, load integration) therefore move residuals by a few per cent and late CG
    # counts by a few iterations; Newton counts, line-search steps and the solution stay identical.
    assert np.all(np.abs(hist[:, 4] - ref[:, 4]) <= np.maximum(3, 0.06 * ref[:, 4]))
    nan = np.isnan(ref[:, 3])
    assert np.array_equal(np.isnan(hist[:, 3]), nan)
    fin = ref[~nan, 3]
    big = ~nan & (ref[:, 3] > 1e-6 * (fin.max() if fin.size else 0.0)) & (ref[:, 3] > 1e3 * float(z["tol"]))
    assert np.allclose(hist[big, 3], ref[big, 3], rtol=res_rtol)


def check_solution(u, z):
    """u within 1e-10 of the reference's (relative, max norm); exactly zero where the reference gave up at
    load factor 0."""
    ref_u = torch.from_numpy(z["u"])
    if float(ref_u.abs().max()) == 0.0:
        assert float(u.abs().max()) == 0.0
    else:
        assert float((u.cpu() - ref_u).abs().max() / ref_u.abs().max()) < 1e-10


# the original trajectories, and those that exercise the load-step cut-back: after NaN steps (inverted
# elements), after max_newton, and the give-up exit after max_cutbacks halvings; the 100x100 square
GOLDEN_SOLVES = ["g2_q4_disp_20", "g2_q9_force_8", "g2_q4_force_16", "g2_q4_nan_cutback_4", "g2_q4_maxnewton_cutback",
                 "g2_q4_giveup", "g2_q4_disp_100"]


def unit_square(element_type="quad4", n=8, mode="displacement", value=(0.1, None), E=1.0, nu=0.3, **kw):
    spec = dict(element_type=element_type, nx=n, ny=kw.pop("ny", n), E=E, nu=nu, load_mode=mode,
                load_value=value, fixed_side="left", load_side="right", **kw)
    return build_problem(spec)


def cantilever(element_type="quad4", h=0.5, load=(0.0, -40.0), E=2600.0, nu=0.3, **kw):
    """The reference's small test problem: 2x1 rectangle, clamped left, consistent force on the right
    (other materials with their library defaults, as the reference's helper builds them)."""
    spec = dict(element_type=element_type, geometry="rectangle", geom_params=dict(width=2.0, height=1.0), h=h,
                E=E, nu=nu, load_mode="force", load_value=load, fixed_side="left", load_side="right", **kw)
    return build_problem(spec)


def state(params, seed=0, scale=0.02):
    """A non-trivial linearisation point and a random direction (same recipe as the reference tests)."""
    g = torch.Generator().manual_seed(seed)
    n = params["f"].shape[0]
    u = scale * torch.randn(n, dtype=DTYPE, generator=g) * params["m"]
    v = torch.randn(n, dtype=DTYPE, generator=g)
    return u, v


def stretch_state(params, h, seed=0, stretch=0.05):
    """u = (F0 - I) X with F0 = diag(1+stretch, 1), plus N(0, (1e-3 h)^2) per dof."""
    g = torch.Generator().manual_seed(seed)
    X = params["nodes"]
    u = torch.zeros_like(X)
    u[:, 0] = stretch * X[:, 0]
    u = u.reshape(-1) + 1e-3 * h * torch.randn(u.numel(), dtype=DTYPE, generator=g)
    return u


def rel(a, b):
    """max|a-b| / max|b| (the reference's agreement metric, test_tangent_operators.py:26-27)."""
    return float((a.double().cpu() - b.double().cpu()).abs().max() / b.double().cpu().abs().max())


def dense(op, n):
    eye = torch.eye(n, dtype=DTYPE)
    return torch.stack([op(eye[i]) for i in range(n)], dim=1)
