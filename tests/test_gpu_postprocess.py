"""Post-processing on the GPU (`postprocess.py:24-137` of the reference) vs the oracle: element fields,
nodal averages, strain energy, min J and support reactions within 1e-11 relative (FP64), on partials and
halo plans; reactions balance the applied load after a GPU solve; an affine deformation of the 400x400
mesh gives the closed-form stress everywhere."""
import math

import pytest
import torch

from helpers import cantilever, rel, state, unit_square
from oracle import fem

pytestmark = pytest.mark.gpu
DEV = "cuda"
CASES = [("quad4", 8), ("quad9", 4), ("quad4", 37), ("quad9", 19)]
KEYS = ("J", "sxx", "syy", "sxy", "mises", "energy_density")


def _setup(et, n, seed=0):
    p = fem.precompute_geometry(unit_square(et, n))
    amp = 0.16 if et == "quad4" else 0.04
    u, _ = state(p, seed, scale=amp / n)
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(seed + 7))
    return p, u, th


def _check_fields(out, ref):
    for k in KEYS:
        assert rel(out[k], ref[k]) < 1e-11, k
        assert rel(out[k + "_n"], ref[k + "_n"]) < 1e-11, k + "_n"
    for k in ("ux", "uy", "u_mag", "ux_n"):
        assert rel(out[k], ref[k]) < 1e-15, k
    assert float(out["strain_energy"]) == pytest.approx(float(ref["strain_energy"]), rel=1e-11)
    assert float(out["min_J"]) == pytest.approx(float(ref["min_J"]), rel=1e-13)


@pytest.mark.parametrize("et,n", CASES)
def test_fields_and_reactions_match_oracle(et, n):
    from difffea_amd.postprocess import compute_fields, support_reactions
    p, u, th = _setup(et, n, seed=2)
    out = compute_fields(u.to(DEV), th.to(DEV), p)
    assert set(out) == set(k for k in fem.compute_fields(u, th, p) if k != "area")
    _check_fields(out, fem.compute_fields(u, th, p))
    r = support_reactions(u.to(DEV), th.to(DEV), p)
    assert rel(r, fem.support_reactions(u, th, p)) < 1e-11
    # repeatable bit for bit
    assert torch.equal(compute_fields(u.to(DEV), th.to(DEV), p)["mises_n"], out["mises_n"])


def test_fields_and_reactions_on_halo_plans():
    from difffea_amd.device import HipMesh, HipSystem
    from difffea_amd.operators import _check_material
    p, u, th = _setup("quad4", 37, seed=3)
    ref = fem.compute_fields(u, th, p)
    mesh = HipMesh(p, DEV, plan_mode="halo")
    assert mesh.plan.n_shared == 0 and len(mesh.plan.hel) > 0
    s = HipSystem(mesh, th, _check_material(p), p["f"], "fp64")
    eout, nout = s.fields(u.to(DEV))
    for i, k in enumerate(("area",) + KEYS):
        assert rel(eout[i], ref[k]) < 1e-11, k
    for i, k in enumerate(KEYS):
        assert rel(nout[i], ref[k + "_n"]) < 1e-11, k
    assert rel(s.reactions(u.to(DEV)), fem.support_reactions(u, th, p)) < 1e-11


@pytest.mark.parametrize("et", ["quad4", "quad9"])
def test_reactions_balance_the_load_after_a_gpu_solve(et):
    """The reference's end-to-end check (`test_solver_end_to_end.py:46-50`) on the GPU path."""
    from difffea_amd.newton_krylov import newton_krylov_solve
    from difffea_amd.postprocess import compute_fields, support_reactions
    p = cantilever(et, load=(0.0, -20.0))
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64)
    u, info = newton_krylov_solve(p, th.to(DEV), 3, 1e-9, verbose=False)
    assert info["converged"]
    reaction = support_reactions(u, th.to(DEV), p).cpu()
    assert torch.allclose(reaction + p["f"].reshape(-1, 2).sum(0), torch.zeros(2, dtype=torch.float64), atol=1e-7)
    fields = compute_fields(u, th.to(DEV), p)
    assert float(fields["min_J"]) > 0 and bool(torch.isfinite(fields["mises"]).all())
    _check_fields(fields, fem.compute_fields(u.cpu(), th, fem.precompute_geometry(p)))


def test_affine_fields_at_full_size():
    """400x400 quad4 (160,000 elements): u = (F0 - I) X gives the closed-form Cauchy stress, energy and J in
    every element and at every node, and reactions equal to minus the free dofs' internal forces."""
    from difffea_amd.postprocess import compute_fields, support_reactions
    from difffea_amd.operators import make_system
    p = unit_square("quad4", 400)
    F0 = torch.tensor([[1.05, 0.02], [-0.01, 0.98]], dtype=torch.float64)
    u = (p["nodes"] @ (F0 - torch.eye(2, dtype=torch.float64)).T).reshape(-1).to(DEV)
    th = torch.ones(p["conn"].shape[0], dtype=torch.float64, device=DEV)
    out = compute_fields(u, th, p)
    mu, lam = float(p["mat_p"][0]), float(p["mat_p"][1])
    J = float(F0[0, 0] * F0[1, 1] - F0[0, 1] * F0[1, 0])
    lnJ = math.log(J)
    B = F0 @ F0.T
    c = dict(J=J, sxx=(mu * (float(B[0, 0]) - 1) + lam * lnJ) / J, syy=(mu * (float(B[1, 1]) - 1) + lam * lnJ) / J,
             sxy=mu * float(B[0, 1]) / J)
    szz = lam * lnJ / J
    c["mises"] = math.sqrt(0.5 * ((c["sxx"] - c["syy"]) ** 2 + (c["syy"] - szz) ** 2 + (szz - c["sxx"]) ** 2)
                           + 3 * c["sxy"] ** 2)
    c["energy_density"] = 0.5 * mu * (float((F0 * F0).sum()) - 2 - 2 * lnJ) + 0.5 * lam * lnJ ** 2
    for k, v in c.items():
        assert float((out[k] - v).abs().max()) < 1e-12 * max(1.0, abs(v)), k
        assert float((out[k + "_n"] - v).abs().max()) < 1e-12 * max(1.0, abs(v)), k
    assert float(out["strain_energy"]) == pytest.approx(c["energy_density"], rel=1e-10)
    # the internal forces of all dofs sum to zero, so the reactions are minus the free dofs' sum, which the
    # residual kernel gives (R = m R_int at load factor 0)
    R = make_system(u, th, p).residual(u, lf=0.0).reshape(-1, 2)
    r = support_reactions(u, th, p)
    assert float((r + R.sum(0)).abs().max()) < 1e-12 * float(R.abs().sum())
    assert float(r.abs().max()) > 1e-3
