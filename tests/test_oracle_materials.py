"""The oracle's material library (`materials.py:65-186` of the reference), pinned on the CPU: stress-free
reference state, objectivity, plane-strain closed forms, limits that reduce to neo-Hookean, fibres slack in
compression, stress and moduli against central differences of the energy, R_int = grad(total energy), and
the linearity in theta that the adjoint gradient kernel relies on."""
import math

import pytest
import torch
from torch.func import grad, hessian

from helpers import DTYPE, state, unit_square
from oracle import fem
from difffea_amd.materials import material_names, material_vector

F_SAMPLE = torch.tensor([[1.12, 0.07], [-0.04, 0.93]], dtype=DTYPE)
CASES = [(n, None) for n in material_names()] + [("fiber_reinforced", {"d": 0.2, "beta_deg": 60.0})]


def _psi(name):
    return fem.material_psi({"material": name})


def _rot(a):
    return torch.tensor([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]], dtype=DTYPE)


@pytest.mark.parametrize("name,vals", CASES)
def test_stress_free_reference_and_objectivity(name, vals):
    p, _ = material_vector(name, vals)
    psi = _psi(name)
    I = torch.eye(2, dtype=DTYPE)
    assert abs(float(psi(I, p))) < 1e-12
    assert float(grad(psi)(I, p).abs().max()) < 1e-10
    assert abs(float(psi(_rot(0.9) @ F_SAMPLE, p) - psi(F_SAMPLE, p))) < 1e-10 * max(1.0, float(psi(F_SAMPLE, p)))


@pytest.mark.parametrize("name,vals", CASES)
def test_stress_and_moduli_match_central_differences(name, vals):
    p, _ = material_vector(name, vals)
    psi = _psi(name)
    F = F_SAMPLE.clone()
    P = grad(psi)(F, p)
    C = hessian(psi)(F, p).reshape(4, 4)
    h = 1e-6
    for k in range(4):
        dF = torch.zeros(4, dtype=DTYPE)
        dF[k] = h
        dF = dF.reshape(2, 2)
        fd = (psi(F + dF, p) - psi(F - dF, p)) / (2 * h)
        assert float(abs(fd - P.reshape(-1)[k])) < 1e-6 * max(1.0, float(P.abs().max()))
        fdP = ((grad(psi)(F + dF, p) - grad(psi)(F - dF, p)) / (2 * h)).reshape(-1)
        assert float((fdP - C[:, k]).abs().max()) < 1e-5 * max(1.0, float(C.abs().max()))
    assert float((C - C.T).abs().max()) < 1e-9 * float(C.abs().max())


def test_plane_strain_closed_forms():
    F = F_SAMPLE
    s = float((F * F).sum())
    J = float(torch.det(F))
    lnJ = math.log(J)
    c10, c01, kap = 400.0, 100.0, 1500.0
    mr = c10 * (J ** (-2 / 3) * (s + 1) - 3) + c01 * (J ** (-4 / 3) * (J * J + s) - 3) + 0.5 * kap * lnJ ** 2
    assert float(_psi("mooney_rivlin")(F, torch.tensor([c10, c01, kap], dtype=DTYPE))) == pytest.approx(mr, rel=1e-13)
    mu, lam = 1000.0, 1500.0
    svk = lam / 8 * (s - 2) ** 2 + mu / 4 * (s * s - 2 * J * J - 2 * s + 2)
    assert float(_psi("saint_venant")(F, torch.tensor([mu, lam], dtype=DTYPE))) == pytest.approx(svk, rel=1e-12)
    nh = float(fem.psi_neo_hookean(F, torch.tensor([mu, lam], dtype=DTYPE)))
    assert float(fem.psi3_neo_hookean(fem.embed_plane_strain(F), torch.tensor([mu, lam], dtype=DTYPE))) == \
        pytest.approx(nh, rel=1e-14)
    # Demiray with gamma -> 0 and HGO with k1 -> 0 are neo-Hookean
    dem = float(_psi("demiray")(F, torch.tensor([mu, 1e-6, lam], dtype=DTYPE)))   # O(gamma) from neo-Hookean
    assert dem == pytest.approx(nh, rel=1e-6)
    p_hgo, _ = material_vector("fiber_reinforced", {"mu": mu, "lmbda": lam, "k1": 1e-12})
    assert float(_psi("fiber_reinforced")(F, p_hgo)) == pytest.approx(nh, rel=1e-12)


def test_fibres_carry_load_only_in_tension():
    p, _ = material_vector("fiber_reinforced", {"beta_deg": 0.0, "d": 0.0})
    nh = lambda F_: fem.psi_neo_hookean(F_, p[:2])
    psi = _psi("fiber_reinforced")
    Fc = torch.tensor([[0.9, 0.0], [0.0, 1.0]], dtype=DTYPE)         # fibres along x compressed
    Ft = torch.tensor([[1.1, 0.0], [0.0, 1.0]], dtype=DTYPE)         # stretched
    assert float(psi(Fc, p)) == float(nh(Fc))
    assert torch.equal(grad(psi)(Fc, p), grad(nh)(Fc))
    E = 1.1 ** 2 - 1.0
    extra = 2 * float(p[2]) / (2 * float(p[3])) * (math.exp(float(p[3]) * E * E) - 1)   # two families
    assert float(psi(Ft, p) - nh(Ft)) == pytest.approx(extra, rel=1e-12)


@pytest.mark.parametrize("name,vals", CASES)
def test_out_of_plane_stress_by_central_differences(name, vals):
    """P3[2,2] from the oracle's automatic differentiation of psi3 (the sigma_zz of the fields)."""
    p, _ = material_vector(name, vals)
    psi3 = fem.PSI3[name]
    F3 = fem.embed_plane_strain(F_SAMPLE)
    P33 = float(grad(psi3)(F3, p)[2, 2])
    h = 1e-6
    d = torch.zeros(3, 3, dtype=DTYPE)
    d[2, 2] = h
    fd = float((psi3(F3 + d, p) - psi3(F3 - d, p)) / (2 * h))
    assert P33 == pytest.approx(fd, rel=1e-6, abs=1e-6)


@pytest.mark.parametrize("name,vals", CASES)
def test_internal_force_gradient_and_linearity_in_theta(name, vals):
    p_ = fem.precompute_geometry(unit_square("quad4", 3, material=name, material_params=vals))
    p = dict(p_, m=torch.ones_like(p_["m"]), f=torch.zeros_like(p_["f"]))
    th = 0.5 + torch.rand(p["conn"].shape[0], dtype=DTYPE, generator=torch.Generator().manual_seed(2))
    N = p["nodes"].shape[0]
    psi = fem.material_psi(p)
    pe = fem.element_material_parameters(th, p)

    def energy(u):
        F = torch.eye(2, dtype=DTYPE) + torch.einsum("eai,eqaj->eqij", u.reshape(N, 2)[p["conn"]], p["dN_dX"])
        e = torch.vmap(torch.vmap(psi, in_dims=(0, None)))(F, pe)
        return (p["w_detJ"] * e).sum()

    u, _ = state(p, 3, scale=0.03)
    R = fem.global_residual(u, th, p)
    assert torch.allclose(grad(energy)(u), R, atol=1e-9 * float(R.abs().max()), rtol=1e-9)
    # every library energy is linear in its theta-scaled parameters: R_int(theta) = theta R_int(1) per element
    R1 = fem.global_residual(u, 2.5 * th, p)
    assert torch.allclose(R1, 2.5 * R, rtol=1e-12, atol=1e-12 * float(R.abs().max()))
