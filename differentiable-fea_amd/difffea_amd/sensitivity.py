"""``compute_sensitivity`` where the reference keeps it (`src/difffea/sensitivity.py:7-38`); the adjoint solve and
the hand-written gradient kernel behind it live in ``adjoint``."""
from .adjoint import compute_sensitivity, equilibrium  # noqa: F401
