"""The PCG state machine of the HIP kernel, restated on the CPU and pinned against the textbook PCG.

The kernel (`k_pcg`) runs Jacobi-PCG in the Chronopoulos-Gear arrangement, one kernel per iteration: kernel
K_{-1} forms w_0 = A u_0 (u = M r), kernel K_k first takes the scalar step from the previous kernel's sums
(gamma = r.u, delta = w.u, r.r), then updates p, s = A p, x, r and forms w_{k+1}.  This file restates that
sequence -- the order of its tests (convergence of r_k for k >= 1, iteration cap, curvature with the answer
z at k = 0), its recurrences and its bookkeeping of k and the iteration count -- in numpy, and compares it
with the oracle's textbook PCG (the reference's `newton_krylov.py:28-88`): same iteration counts, same exit
kind, same iterates to round-off, on the dense tangent of a small mesh and on constructed matrices.
"""
import numpy as np
import pytest
import torch

from helpers import dense, state, unit_square
from oracle import fem

RUNNING, CONVERGED, CURVATURE, MAXITER = 0, 1, 2, 3


def cg_gear(A, b, dinv, rtol, max_iter):
    """The kernel sequence; returns (x, iterations, status)."""
    x = np.zeros_like(b)
    r = b.copy()
    p = np.zeros_like(b)
    s = np.zeros_like(b)
    bb = float(b @ b)
    if np.sqrt(bb) == 0.0:
        return x, 0, CONVERGED
    # K_{-1}
    u = dinv * r
    w = A @ u
    sums = (float(r @ u), float(w @ u), float(r @ r))
    gam_prev = alpha_prev = 0.0
    k = 0
    while True:
        gam, delta, rr = sums
        if k >= 1 and np.sqrt(rr) <= rtol * np.sqrt(bb):
            return x, k, CONVERGED
        if k >= max_iter:
            return x, max_iter, MAXITER
        beta = gam / gam_prev if k >= 1 else 0.0
        den = delta - beta * gam / alpha_prev if k >= 1 else delta
        if den <= 0.0:          # as the kernel and the reference: a NaN curvature runs on
            return (dinv * r if k == 0 else x), k, CURVATURE
        alpha = gam / den
        u = dinv * r
        p = u + beta * p
        s = w + beta * s
        x = x + alpha * p
        r = r - alpha * s
        u = dinv * r
        w = A @ u
        sums = (float(r @ u), float(w @ u), float(r @ r))
        gam_prev, alpha_prev = gam, alpha
        k += 1


def textbook(A, b, dinv, rtol, max_iter):
    At = torch.from_numpy(A)
    x, k = fem.pcg(lambda v: At @ v, torch.from_numpy(b), lambda v: torch.from_numpy(dinv) * v, rtol, max_iter)
    return x.numpy(), k


def _tangent(n=4, seed=3):
    p = fem.precompute_geometry(unit_square("quad4", n))
    u, v = state(p, seed, scale=0.01)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    A = dense(fem.make_A_operator_qp(u, th, p), u.numel()).numpy()
    dinv = 1.0 / fem.compute_jacobi_diagonal(u, th, p).numpy()
    return A, (v * p["m"]).numpy(), dinv


@pytest.mark.parametrize("rtol", [1e-2, 1e-6, 1e-10])
def test_converged_solves_match_textbook_pcg(rtol):
    A, b, dinv = _tangent()
    x_ref, k_ref = textbook(A, b, dinv, rtol, 1000)
    x, k, st = cg_gear(A, b, dinv, rtol, 1000)
    assert st == CONVERGED and k == k_ref
    assert np.abs(x - x_ref).max() <= 1e-9 * np.abs(x_ref).max()
    assert np.linalg.norm(b - A @ x) <= rtol * np.linalg.norm(b) * (1 + 1e-6)


def test_iteration_cap_and_zero_right_hand_side():
    A, b, dinv = _tangent()
    for cap in (0, 1, 5):
        x_ref, k_ref = textbook(A, b, dinv, 1e-14, cap)
        x, k, st = cg_gear(A, b, dinv, 1e-14, cap)
        assert (st, k, k_ref) == (MAXITER, cap, cap)
        assert np.abs(x - x_ref).max() <= 1e-10 * max(np.abs(x_ref).max(), 1e-300)
    x, k, st = cg_gear(A, np.zeros_like(b), dinv, 1e-8, 10)
    assert (st, k) == (CONVERGED, 0) and not x.any()


def test_negative_curvature_first_and_later_iteration():
    # negative definite: stop at k = 0 with z = M b (reference: `return z if k == 0`)
    A, b, dinv = _tangent()
    x, k, st = cg_gear(-A, b, -dinv, 1e-8, 100)
    x_ref, k_ref = textbook(-A, b, -dinv, 1e-8, 100)
    assert (st, k, k_ref) == (CURVATURE, 0, 0) and np.allclose(x, x_ref, rtol=0, atol=0)
    # indefinite diagonal matrix: positive first direction, negative curvature at a later iteration
    d = np.array([4.0, 3.0, 2.0, -1.0, 5.0])
    b2 = np.array([1.0, 1.0, 1.0, 0.05, 1.0])
    x_ref, k_ref = textbook(np.diag(d), b2, np.ones(5), 1e-12, 50)
    x, k, st = cg_gear(np.diag(d), b2, np.ones(5), 1e-12, 50)
    assert st == CURVATURE and k == k_ref >= 1
    assert np.abs(x - x_ref).max() <= 1e-12 * np.abs(x_ref).max()


@pytest.mark.parametrize("max_iter", [0, 1, 7])
def test_nan_right_hand_side_result_written_without_a_run(max_iter):
    """A Newton residual with a NaN entry (an inverted element) gets the PCG's answer without a launch
    (`newton_krylov.nan_pcg_result`): the oracle's textbook PCG and the kernel sequence return exactly that --
    every entry NaN after max_iter iterations (zeros after none), the cap as the iteration count."""
    from difffea_amd.newton_krylov import nan_pcg_result
    p = fem.precompute_geometry(unit_square("quad4", 4))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    u, _ = state(p, 0)
    A = dense(fem.make_A_operator(u, th, p), u.numel())
    b = torch.randn(u.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    b[5] = float("nan")
    dinv = 1.0 / torch.diagonal(A)
    x_ref, k_ref = fem.pcg(lambda v: A @ v, b, lambda r: dinv * r, rtol=1e-8, max_iter=max_iter)
    x_k, k_k, st = cg_gear(A.numpy(), b.numpy(), dinv.numpy(), 1e-8, max_iter)
    x, k = nan_pcg_result(torch.empty_like(b), max_iter)
    assert k == k_ref == k_k == max_iter
    assert torch.equal(torch.isnan(x), torch.isnan(x_ref)) and torch.isnan(x).all() == (max_iter > 0)
    assert np.array_equal(np.isnan(x_k), torch.isnan(x).numpy())
    if max_iter == 0:
        assert torch.equal(x, x_ref)
