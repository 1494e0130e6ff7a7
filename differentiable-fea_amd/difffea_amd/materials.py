"""The material library of the ``hip`` variant: the parameter table of `src/difffea/materials.py:164-186`
(names, order, defaults and theta-scalable entries, which callers' ``params`` dictionaries carry) with the
kernel ids of ``include/dfea.h`` (``DFEA_MAT_*``).  The energies themselves are the device closed forms of
``csrc/dfea_material.h``.
"""
from dataclasses import dataclass

import torch

from . import _lib

# name -> (kernel id, parameter names, defaults, theta-scalable parameters)
MATERIALS = {
    "neo_hookean": (0, ("mu", "lmbda"), (1000.0, 1500.0), ("mu", "lmbda")),
    "mooney_rivlin": (1, ("c10", "c01", "kappa"), (400.0, 100.0, 1500.0), ("c10", "c01", "kappa")),
    "saint_venant": (2, ("mu", "lmbda"), (1000.0, 1500.0), ("mu", "lmbda")),
    "demiray": (3, ("mu", "gamma", "lmbda"), (500.0, 3.0, 1500.0), ("mu", "lmbda")),
    "fiber_reinforced": (4, ("mu", "lmbda", "k1", "k2", "beta_deg", "d", "second_family"),
                         (400.0, 1200.0, 1500.0, 4.0, 30.0, 0.0, 1.0), ("mu", "lmbda", "k1")),
}


def material_names():
    return list(MATERIALS)


def material_vector(name, values=None, dtype=torch.float64):
    """(p, scalable mask) of a material; ``values`` overrides defaults by name (`materials.py:208-222`)."""
    if name not in MATERIALS:
        raise ValueError(f"Unknown material '{name}'. Choose one of {material_names()}.")
    _, names, defaults, scal = MATERIALS[name]
    vals = dict(zip(names, defaults))
    if values:
        unknown = set(values) - set(vals)
        if unknown:
            raise ValueError(f"Unknown parameters {sorted(unknown)} for '{name}'; valid: {list(names)}")
        vals.update({k: float(v) for k, v in values.items()})
    return (torch.tensor([vals[n] for n in names], dtype=dtype),
            torch.tensor([n in scal for n in names], dtype=torch.bool))


@dataclass(frozen=True)
class MaterialSpec:
    """What a descriptor carries about the material: kernel id, nominal parameters, scalable bits."""
    mid: int
    p: tuple
    scal: int

    @property
    def library(self):
        """True when the kernels need the library instantiations (any material but neo-Hookean)."""
        return self.mid != 0


def material_spec(params):
    """MaterialSpec of a ``params`` dictionary (`materials.py:225-236`: without a ``material`` key the
    problem is neo-Hookean with ``mu0`` / ``lmbda0``)."""
    if "material" not in params:
        return MaterialSpec(0, (float(params["mu0"]), float(params["lmbda0"])) + (0.0,) * (_lib.MAT_NP - 2), 3)
    name = params["material"]
    if name not in MATERIALS:
        raise ValueError(f"Unknown material '{name}'. Choose one of {material_names()}.")
    mid, names, _, _ = MATERIALS[name]
    p = torch.as_tensor(params["mat_p"], dtype=torch.float64).cpu().reshape(-1)
    if p.numel() != len(names):
        raise ValueError(f"material '{name}' has {len(names)} parameters, mat_p has {p.numel()}")
    _, lib_mask = material_vector(name)
    sc = params.get("mat_scalable")
    sc = lib_mask if sc is None else torch.as_tensor(sc).cpu().reshape(-1)
    if not torch.equal(sc.to(torch.bool), lib_mask):
        # every library energy is linear in its scalable parameters, so R_int(u; p(theta)) = theta R_int(u; p0)
        # and the adjoint kernel's gradient is exact; another scaling would break that identity
        raise ValueError(f"the hip kernels scale the parameters {[n for n, s_ in zip(names, lib_mask) if s_]} "
                         f"of '{name}' by theta, as the material library does")
    bits = sum(1 << k for k in range(len(names)) if bool(sc[k]))
    vals = tuple(float(x) for x in p) + (0.0,) * (_lib.MAT_NP - len(names))
    return MaterialSpec(mid, vals, bits)


__all__ = ["MATERIALS", "material_names", "material_vector", "MaterialSpec", "material_spec"]
