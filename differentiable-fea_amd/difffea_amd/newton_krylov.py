"""Newton-Krylov solve on device-resident state, mirroring `src/difffea/newton_krylov.py`.

* ``pcg(A_fn, b, Minv_fn, rtol, max_iter)``  the closure-based PCG (`newton_krylov.py:28-88`), for any
  operator closures on any device (same algorithm, same early exits, same return values).
* ``newton_krylov_solve(params, theta, ..., tangent="hip", precision="fp64")`` (`newton_krylov.py:186-264`):
  load continuation with cut-backs, Eisenstat-Walker forcing, Armijo line search -- the same state machine
  and ``info`` dictionary -- where every vector operation runs in the HIP library: residual, moduli +
  preconditioner (one fused kernel per Newton step: the Jacobi diagonal, or the inverse nodal 2x2 blocks
  for ``preconditioner="block_jacobi"``), the fused PCG (one kernel per iteration, device scalars, status
  read once per chunk of iterations), line-search trials.
* ``precision="mixed"``: moduli, A(v) and the inner CG vectors in FP32 (dot products accumulated in FP64),
  while the Newton residual, the state u and the line search stay FP64 -- Newton itself is the outer
  refinement loop, so the converged residual meets the same FP64 tolerance.  The inner relative tolerance
  is floored at ``MIXED_RTOL_FLOOR``.
"""
import ctypes
import math
import time

import torch

from . import _lib
from .operators import PRECONDITIONERS, PRECISIONS, TANGENT_OPERATORS, get_mesh, _check_material
from .device import HipSystem, _stream

# Inner tolerance floor of the mixed-precision Newton steps.  On the same state and right-hand side the FP32 PCG
# needs the FP64 iteration count (100x100 Q4 and 200x200 Q9 Newton states, rtol 1e-3 and 1e-4: equal within a few
# iterations), but the FP32 rounding of its iterates leaves a true residual against the FP64 tangent that no
# inner tolerance removes (smooth right-hand side: about 9e-3 at 100x100 Q4, 8e-2 at 256x256, 0.43 at 512x512),
# so each Newton step is accurate to that floor and tight inner solves only chase it (100x100 Q4, second load
# step: 383 and 422 iterations at rtol 1e-3 where the FP64 path needs 252 and 316 at rtol 1e-3 / 3e-5; the same
# solves with FP64 vectors and FP32 moduli need 252 and 352).
# Looser inner solves cost a Newton iteration more and save the tail.  Measured on MI355X, floors 1e-3 / 3e-3 /
# 1e-2 / 3e-2: 100x100 Q4 28.1 / 27.0 / 25.6 / 27.0 ms, O-grid 64x256 Q9 0.591 / 0.567 / 0.507 / 0.597 s, 200x200
# Q9 stretch 0.478 / 0.443 / 0.428 / 0.411 s; the 64-request stream one request after another 26.2 s (1e-3) and
# 25.1 s (1e-2).  Newton still meets the FP64 tolerance (it is the outer refinement).  The mixed adjoint's
# refinement corrections use the same floor: one reduces the FP64 residual by about 1e-2 whatever its own
# tolerance (256x256 Q4 adjoint, floors 1e-4 / 1e-3 / 3e-3 / 1e-2: 4 / 5 / 5 / 6 corrections, 0.067 / 0.073 /
# 0.060 / 0.060 s).
MIXED_RTOL_FLOOR = 1e-2
_N_ALPHAS = 16          # line-search step lengths 2^0 .. 2^-15 kept on the device


def first_round(it, counts, prev_counts):
    """Kernels of the first PCG round of Newton iteration `it` of a load step: an estimate of the CG
    iterations + 2.  The Eisenstat-Walker forcing terms repeat from one load step to the next, and so do the
    iteration counts (100x100 Q4: 4, 11, 18, 68, 97, 125, 194, 250 in the first step; 4, 11, 17, 69, 86, 120,
    188, 252, 318 in the second), so the estimate is the count of the same iteration of the previous load step
    (`prev_counts`) with a margin, else 1.5x the previous iteration's count; a short round costs one more
    status read (a synchronisation, tens of microseconds), a long one a few microseconds per extra kernel."""
    if it < len(prev_counts):
        est = 1.1 * prev_counts[it] + 8
    elif counts:
        est = 1.5 * counts[-1] + 8
    else:
        est = 14
    return int(est) + 2


def nan_pcg_result(x, max_iter):
    """The PCG's result for a right-hand side with a NaN entry, written into x without running the kernels.

    A NaN residual comes from an inverted element at the Newton state (log J with det F <= 0,
    `constitutive_relations.py:46`).  In the reference's PCG (`newton_krylov.py:28-88`) r.z, the curvature and
    the residual norm are then NaN, so the zero right-hand side exit, the curvature exit and the convergence
    test are all False, and x + alpha p with alpha = NaN is NaN in every entry: it returns that x after
    max_iter iterations (x = 0 when max_iter is 0).  The fused kernels reach the same x and count (NaN
    propagates through every product and sum); the line search then rejects every NaN trial and the load step
    is cut back.  Returns (x, max_iter)."""
    x.fill_(float("nan") if max_iter > 0 else 0.0)
    return x, max_iter


def pcg(A_fn, b, Minv_fn, rtol=1e-5, max_iter=500):
    """Textbook Jacobi-PCG with x0 = 0 (returns (x, n_iters))."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = Minv_fn(r)
    p = z.clone()
    rz = torch.dot(r, z)
    nb = torch.norm(b)
    if nb == 0.0:
        return x, 0
    for k in range(max_iter):
        Ap = A_fn(p)
        curv = torch.dot(p, Ap)
        if curv <= 0.0:
            return (z if k == 0 else x), k
        alpha = rz / curv
        x = x + alpha * p
        r = r - alpha * Ap
        if torch.norm(r) <= rtol * nb:
            return x, k + 1
        z = Minv_fn(r)
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter


class HipNewton:
    """The Newton-Krylov state machine of one solve on one ``HipSystem`` (device-resident)."""

    def __init__(self, params, theta, precision="fp64", device=None, preconditioner="jacobi", mesh=None):
        """mesh: a ``HipMesh`` of exactly these params built beforehand (else the cached one of params)."""
        if precision not in PRECISIONS:
            raise ValueError(f"Unknown precision '{precision}'. Choose one of {PRECISIONS}.")
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.params = params
        self.theta = theta.detach()
        self.mesh = mesh if mesh is not None else get_mesh(params, device)
        self.sys = HipSystem(self.mesh, theta.detach(), _check_material(params), params["f"], precision,
                             preconditioner=preconditioner)
        self.device = self.mesh.device
        self.uD = params["u_D"].to(self.device, torch.float64).contiguous()
        nd = self.mesh.n_dof
        self.ucand = torch.empty(nd, dtype=torch.float64, device=self.device)
        self._alpha = torch.empty(1, dtype=torch.float64, device=self.device)
        # the line search's step lengths 2^-j on the device (a trial reads its alpha there, no fill per trial)
        self._alphas = torch.tensor([2.0 ** -j for j in range(_N_ALPHAS)], dtype=torch.float64, device=self.device)
        self._moduli_at = None          # the state the system's moduli were last computed at (see armijo)
        self._rhs_nan = False           # the current Newton residual has a NaN entry (see nan_pcg_result)
        self._spec = None               # the line search's first trial, enqueued behind the PCG (linear_solve)
        self.pcg_not_run = 0            # PCG solves whose result nan_pcg_result wrote
        self.pcg_chunk = 16
        self._step_counts = []          # CG iterations of the Newton iterations of the last load step solved

    # -- building blocks -------------------------------------------------------------------------------------
    def residual(self, u, lf, out=None):
        return self.sys.residual(u, out=self.sys.vec["R"] if out is None else out, lf=lf)

    def norm(self, R):
        return math.sqrt(self.sys.dot_host(R))

    def linear_solve(self, rtol, max_iter, first=None, trial_at=None):
        """PCG on J du = -R with R in sys.vec['R']; returns (du, iterations).  first: kernels of the first
        round of PCG kernels (``first_round``).  trial_at = (u, lf): behind every PCG round the line search's
        first trial u + du and its residual norm are enqueued on the round's answer (``_spec``), so the GPU
        goes on with them while the host reads the PCG status; the last round's are the line search's."""
        s = self.sys
        self._spec = None
        if self._rhs_nan:
            return nan_pcg_result(s.vec["x"], max_iter)
        if s.precision == "mixed":
            rtol = max(rtol, MIXED_RTOL_FLOOR)
        s.pcg_start(rtol=rtol, max_iter=max_iter, neg_res=True)
        after = None
        if trial_at is not None:
            u, lf = trial_at

            def after():
                # the trial and the residual write ucand and vec['R'], which the running PCG does not read
                # (its right-hand side was copied at the start)
                uc = self.trial(u, s.vec["x"], 1.0, self.ucand)
                self._spec = (uc, s.dot_start(self.residual(uc, lf)))
        its, _ = s.pcg_run(max_iter, self.pcg_chunk, first=first, after_round=after)
        return s.vec["x"], its

    def trial(self, u, du, alpha, out):
        j = -math.log2(alpha) if alpha > 0.0 else -1.0
        if j == int(j) and 0 <= j < _N_ALPHAS:
            a = ctypes.c_void_p(self._alphas.data_ptr() + 8 * int(j))
        else:
            self._alpha.fill_(alpha)
            a = ctypes_ptr(self._alpha)
        m = self.mesh
        _lib.check(m.lib.dfea_axpy(ctypes_ptr(self.sys.table), ctypes_ptr(m.vitems), m.n_vblk, 1, ctypes_ptr(u),
                                   ctypes_ptr(du), None, a, ctypes_ptr(out), self.sys.prec_id, stream()), "dfea_axpy")
        return out

    # -- reference state machine ---------------------------------------------------------------------------
    def armijo(self, u, du, rn, lf, c1=1e-4, min_alpha=2.0 ** -5, spec=None):
        """The reference's backtracking line search; spec: the first trial (alpha = 1) already enqueued, as
        (candidate, pending norm) from ``linear_solve``."""
        alpha = 1.0
        while alpha >= min_alpha:
            if alpha == 1.0 and spec is not None:
                uc, pending = spec
            else:
                uc = self.trial(u, du, alpha, self.ucand)
                R = self.residual(uc, lf)
                pending = self.sys.dot_start(R)
            # the next Newton iteration's moduli at the trial state, enqueued before the norm is read: when the
            # trial is accepted (the usual case) the GPU computes them while the host waits for the norm and
            # decides; otherwise they are recomputed where needed (same kernel, same state, same bits)
            self.sys.moduli(uc)
            self._moduli_at = uc
            rc = math.sqrt(pending())
            if rc < (1.0 - c1 * alpha) * rn:          # False for NaN (inverted element)
                self.ucand = u
                return uc, rc, alpha, True
            alpha *= 0.5
        return u, rn, alpha, False

    def at_load(self, u0, lf, tol, max_newton, step_id, cg_rtol, cg_max_iter, callback, verbose):
        self._moduli_at = None
        u = self.sys.lift(u0, self.uD, lf)
        R = self.residual(u, lf)
        r0 = self.norm(R)
        rn = r0
        records = []
        counts, prev = [], self._step_counts
        for it in range(max_newton + 1):
            if rn < tol:
                if verbose:
                    print(f"  converged in {it} Newton iterations (||R|| = {rn:.3e})")
                self._step_counts = counts
                return u, True, records
            if it == max_newton:
                break
            eta = min(0.5, math.sqrt(rn / (r0 + 1e-14))) if cg_rtol is None else cg_rtol
            self._rhs_nan = math.isnan(rn)
            if self._rhs_nan:
                self.pcg_not_run += 1                   # no tangent needed: the PCG's result is known
            elif self._moduli_at is not u:              # else computed after the line search that accepted u
                self.sys.moduli(u)
            self._moduli_at = None
            if self.ucand is u:
                self.ucand = torch.empty_like(u)
            du, k = self.linear_solve(eta, cg_max_iter, first_round(it, counts, prev), trial_at=(u, lf))
            counts.append(int(k))
            u, rn, alpha, ok = self.armijo(u, du, rn, lf, spec=self._spec)
            rec = dict(step=step_id, load_factor=lf, iteration=it, residual=float(rn), cg_iters=int(k), alpha=alpha)
            records.append(rec)
            if callback is not None:
                callback(rec)
            if verbose:
                print(f"  iter {it:2d} | ||R|| = {rn:.4e} | CG = {k:3d} | alpha = {alpha:.4f}")
            if not ok:
                return u, False, records
        return u, False, records

    def solve(self, n_load_steps=5, tol=1e-8, max_newton=25, verbose=False, callback=None, max_cutbacks=8,
              cg_rtol=None, cg_max_iter=500):
        n = self.mesh.n_dof
        u = torch.zeros(n, dtype=torch.float64, device=self.device)
        info = {"converged": False, "load_factor": 0.0, "history": [], "total_newton": 0, "total_cg": 0,
                "n_cutbacks": 0, "snapshots": [(0.0, u.clone())], "precision": self.sys.precision,
                "pcg_not_run": 0}       # Newton iterations with a NaN residual: PCG result written, not computed
        d_nom = 1.0 / n_load_steps
        d = d_nom
        lf = 0.0
        step, in_row = 0, 0
        while lf < 1.0 - 1e-12:
            lf_try = min(1.0, lf + d)
            step += 1
            if verbose:
                print(f"\n=== Load step {step}: load factor {lf:.3f} -> {lf_try:.3f} ===")
            u_try, ok, recs = self.at_load(u, lf_try, tol, max_newton, step, cg_rtol, cg_max_iter, callback, verbose)
            info["history"] += recs
            info["total_newton"] += len(recs)
            info["total_cg"] += sum(r["cg_iters"] for r in recs)
            info["pcg_not_run"] = self.pcg_not_run
            if ok:
                u, lf = u_try.clone(), lf_try
                info["snapshots"].append((lf, u.clone()))
                in_row = 0
                d = min(d_nom, 1.5 * d)
            else:
                in_row += 1
                info["n_cutbacks"] += 1
                if in_row > max_cutbacks:
                    info["load_factor"] = lf
                    return u, info
                d *= 0.5
                if verbose:
                    print(f"  Newton failed -> load increment reduced to {d:.4f}")
        info["converged"] = True
        info["load_factor"] = 1.0
        return u, info


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream():
    return _stream()


@torch.no_grad()
def newton_krylov_solve(params, theta, n_load_steps=5, tol=1e-8, max_newton=25, verbose=True, callback=None,
                        max_cutbacks=8, preconditioner="jacobi", tangent="auto", precision="fp64",
                        cg_rtol=None, cg_max_iter=500, device=None):
    """Solve R(u, theta) = 0 by load continuation + inexact Newton-Krylov; returns (u, info).

    Same signature and ``info`` keys as the reference (``tol`` is absolute on ||R||_2), plus
    ``precision`` ("fp64" | "mixed"), ``cg_rtol`` (fixed inner tolerance instead of Eisenstat-Walker),
    ``cg_max_iter`` (the reference's PCG cap, 500) and ``device``.
    """
    if tangent not in TANGENT_OPERATORS:
        raise ValueError(f"Unknown tangent operator '{tangent}'. Choose one of {TANGENT_OPERATORS}.")
    if preconditioner not in PRECONDITIONERS:
        raise ValueError(f"Unknown preconditioner '{preconditioner}'. Choose one of {PRECONDITIONERS}.")
    if device is None:
        device = theta.device if theta.is_cuda else torch.device("cuda", torch.cuda.current_device())
    solver = HipNewton(params, theta.to(device), precision, device, preconditioner)
    t0 = time.perf_counter()
    u, info = solver.solve(n_load_steps, tol, max_newton, verbose, callback, max_cutbacks, cg_rtol, cg_max_iter)
    torch.cuda.synchronize(solver.device)
    info["seconds"] = time.perf_counter() - t0
    return u, info


def solve_newton_krylov(params, theta, n_load_steps=5, tol=1e-8, max_newton=25, verbose=True, **kw):
    """Convenience wrapper returning only u (`newton_krylov.py:267-274`)."""
    u, _ = newton_krylov_solve(params, theta, n_load_steps=n_load_steps, tol=tol, max_newton=max_newton,
                               verbose=verbose, **kw)
    return u
