"""Write golden vectors from the reference implementation (run ONLY where /root/reference/src exists).

Only the two functions that no closed form or invariant pins are recorded:

* G2 -- the exact Newton / CG trajectory of ``newton_krylov_solve`` (Eisenstat-Walker forcing, Armijo
  steps, load continuation): solution and per-iteration history on small problems
  (`src/difffea/newton_krylov.py:126-264`), including
    - the load-step cut-back after a stalled line search on a state whose lift inverts elements (NaN
      residual from log J, `constitutive_relations.py:46`; NaN through the Jacobi clamp, `operators.py:555`;
      NaN steps rejected by Armijo, `newton_krylov.py:101-102`),
    - the cut-back after ``max_newton`` iterations without convergence, with recovery of the nominal
      increment (`newton_krylov.py:247-260`),
    - the give-up exit after ``max_cutbacks`` halvings in a row (`newton_krylov.py:254-256`),
    - the 100x100 displacement-driven unit square at full size (20,402 dofs);
* G3 -- the reference's torch.func implicit-function adjoint ``compute_sensitivity`` (compliance gradient,
  `src/difffea/sensitivity.py:7-38`).

Usage (this container only; the tests read the .npz files and never import the reference):

    PYTHONPATH=/root/reference/src PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py [name ...]

With names, only those files are written (the manifest keeps the entries of the others).
"""
import json
import os
import sys

import numpy as np

REF_SRC = "/root/reference/src"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    if not os.path.isdir(REF_SRC):
        sys.exit(f"{REF_SRC} not found: golden vectors are generated only where the reference is mounted")
    sys.path.insert(0, REF_SRC)
    import torch
    from difffea.newton_krylov import newton_krylov_solve
    from difffea.operators import global_residual
    from difffea.problem import build_problem
    from difffea.sensitivity import compute_sensitivity

    only = set(sys.argv[1:])
    mpath = os.path.join(OUT, "MANIFEST.json")
    manifest = {"files": {}}
    if only and os.path.exists(mpath):
        with open(mpath) as fh:
            manifest = json.load(fh)
    manifest.update({"torch": torch.__version__, "numpy": np.__version__})

    def spec(et, n, mode, value, E=1.0, nu=0.3, width=1.0, height=1.0, h=None):
        return dict(geometry="rectangle", geom_params=dict(width=width, height=height), element_type=et,
                    h=1.0 / n if h is None else h, mesher="structured", E=E, nu=nu, fixed_side="left",
                    fixed_dofs="xy", load_side="right", load_mode=mode, load_value=value, dtype=torch.float64)

    def cantilever(load, h=0.5):
        # the reference's small test problem (tests/helpers.py:8-16): 2x1, E = 2600, nu = 0.3, left clamped
        return spec("quad4", None, "force", load, E=2600.0, width=2.0, height=1.0, h=h)

    def lift_norm(params):
        # ||K(0) u_D|| over the free dofs, as ||R(d lift)|| / d with d = 1e-6 (the relative-residual scale of
        # displacement-driven problems)
        d = 1e-6
        th = torch.ones(params["conn"].shape[0], dtype=torch.float64)
        lift = torch.where(params["fixed"], d * params["u_D"], torch.zeros_like(params["u_D"]))
        return float(torch.norm(global_residual(lift, th, params, d))) / d

    # ---- G2: Newton trajectories -------------------------------------------------------------------
    # name: (spec, n_load_steps, tol, max_newton); tol None = 1e-10 ||K(0) u_D||
    cases = {
        "g2_q4_disp_20": (spec("quad4", 20, "displacement", (0.1, None)), 3, 1e-11, 25),
        "g2_q9_force_8": (spec("quad9", 8, "force", (0.0, -0.02)), 3, 1e-12, 25),
        "g2_q4_force_16": (spec("quad4", 16, "force", (0.08, 0.0)), 2, 1e-12, 25),
        # one load step pushing the right edge of a 4x4 square 0.5 to the left: the lift inverts the
        # right-hand column, so R is NaN, the line search rejects every step and the increment is halved
        "g2_q4_nan_cutback_4": (spec("quad4", 4, "displacement", (-0.5, None)), 1, 1e-10, 25),
        # a 200 N tip load in one step with at most 8 Newton iterations per level: cut-backs after
        # max_newton, backtracking steps, and the increment growing back by 1.5 after each success
        "g2_q4_maxnewton_cutback": (cantilever((0.0, -200.0)), 1, 2e-7, 8),
        # the same with at most 5 Newton iterations: nine halvings in a row, the solve gives up at lf = 0
        "g2_q4_giveup": (cantilever((0.0, -200.0)), 1, 2e-7, 5),
        # the 100x100 unit square stretched 10 % in 3 displacement steps, 1e-10 relative residual
        "g2_q4_disp_100": (spec("quad4", 100, "displacement", (0.1, None)), 3, None, 25),
    }
    for name, (sp, steps, tol, max_newton) in cases.items():
        if only and name not in only:
            continue
        params = build_problem(sp)
        if tol is None:
            tol = 1e-10 * lift_norm(params)
        theta = torch.ones(params["conn"].shape[0], dtype=torch.float64)
        u, info = newton_krylov_solve(params, theta, n_load_steps=steps, tol=tol, max_newton=max_newton,
                                      verbose=False, preconditioner="jacobi", tangent="qp")
        hist = np.array([[r["step"], r["load_factor"], r["iteration"], r["residual"], r["cg_iters"], r["alpha"]]
                         for r in info["history"]], dtype=np.float64)
        path = os.path.join(OUT, name + ".npz")
        np.savez(path, u=u.numpy(), history=hist, n_load_steps=steps, tol=tol, max_newton=max_newton,
                 converged=info["converged"], load_factor=info["load_factor"], total_newton=info["total_newton"],
                 total_cg=info["total_cg"], n_cutbacks=info["n_cutbacks"],
                 snapshot_lf=np.array([lf for lf, _ in info["snapshots"]]),
                 spec=json.dumps({k: v for k, v in sp.items() if k != "dtype"}))
        manifest["files"][name] = {"kind": "newton_krylov_solve", "n_dof": int(u.numel()),
                                   "converged": bool(info["converged"]), "total_newton": info["total_newton"],
                                   "total_cg": info["total_cg"], "n_cutbacks": info["n_cutbacks"]}
        print(name, info["converged"], info["total_newton"], info["total_cg"], info["n_cutbacks"], flush=True)

    # ---- G3: reference adjoint sensitivity ------------------------------------------------------------
    if only and "g3_q4_compliance_16" not in only:
        with open(mpath, "w") as fh:
            json.dump(manifest, fh, indent=1)
        return
    sp = spec("quad4", 16, "force", (0.0, -0.02))
    params = build_problem(sp)
    n_el = params["conn"].shape[0]
    theta = 0.1 + 0.9 * torch.rand(n_el, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    f = params["f"]
    u, g = compute_sensitivity(params, theta, lambda w: torch.dot(f, w), n_load_steps=2, tol=1e-12)
    np.savez(os.path.join(OUT, "g3_q4_compliance_16.npz"), theta=theta.numpy(), u=u.numpy(), grad=g.numpy(),
             n_load_steps=2, tol=1e-12, spec=json.dumps({k: v for k, v in sp.items() if k != "dtype"}))
    manifest["files"]["g3_q4_compliance_16"] = {"kind": "compute_sensitivity", "n_el": n_el}
    print("g3", float(g.abs().max()), flush=True)

    with open(mpath, "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
