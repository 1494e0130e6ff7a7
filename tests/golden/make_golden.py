"""Write golden vectors from the reference implementation (run ONLY where /root/reference/src exists).

Only the two functions that no closed form or invariant pins are recorded:

* G2 -- the exact Newton / CG trajectory of ``newton_krylov_solve`` (Eisenstat-Walker forcing, Armijo
  steps, load continuation): solution and per-iteration history on small problems
  (`src/difffea/newton_krylov.py:126-264`);
* G3 -- the reference's torch.func implicit-function adjoint ``compute_sensitivity`` (compliance gradient,
  `src/difffea/sensitivity.py:7-38`).

Usage (this container only; the tests read the .npz files and never import the reference):

    PYTHONPATH=/root/reference/src PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py
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
    from difffea.problem import build_problem
    from difffea.sensitivity import compute_sensitivity

    manifest = {"torch": torch.__version__, "numpy": np.__version__, "files": {}}

    def spec(et, n, mode, value, E=1.0, nu=0.3):
        return dict(geometry="rectangle", geom_params=dict(width=1.0, height=1.0), element_type=et, h=1.0 / n,
                    mesher="structured", E=E, nu=nu, fixed_side="left", fixed_dofs="xy", load_side="right",
                    load_mode=mode, load_value=value, dtype=torch.float64)

    # ---- G2: Newton trajectories -------------------------------------------------------------------
    cases = {
        "g2_q4_disp_20": (spec("quad4", 20, "displacement", (0.1, None)), 3, 1e-11),
        "g2_q9_force_8": (spec("quad9", 8, "force", (0.0, -0.02)), 3, 1e-12),
        "g2_q4_force_16": (spec("quad4", 16, "force", (0.08, 0.0)), 2, 1e-12),
    }
    for name, (sp, steps, tol) in cases.items():
        params = build_problem(sp)
        theta = torch.ones(params["conn"].shape[0], dtype=torch.float64)
        u, info = newton_krylov_solve(params, theta, n_load_steps=steps, tol=tol, verbose=False,
                                      preconditioner="jacobi", tangent="qp")
        hist = np.array([[r["step"], r["load_factor"], r["iteration"], r["residual"], r["cg_iters"], r["alpha"]]
                         for r in info["history"]], dtype=np.float64)
        path = os.path.join(OUT, name + ".npz")
        np.savez(path, u=u.numpy(), history=hist, n_load_steps=steps, tol=tol,
                 converged=info["converged"], total_newton=info["total_newton"], total_cg=info["total_cg"],
                 n_cutbacks=info["n_cutbacks"], spec=json.dumps({k: v for k, v in sp.items() if k != "dtype"}))
        manifest["files"][name] = {"kind": "newton_krylov_solve", "n_dof": int(u.numel()),
                                   "total_newton": info["total_newton"], "total_cg": info["total_cg"]}
        print(name, info["converged"], info["total_newton"], info["total_cg"], flush=True)

    # ---- G3: reference adjoint sensitivity ------------------------------------------------------------
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

    with open(os.path.join(OUT, "MANIFEST.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
