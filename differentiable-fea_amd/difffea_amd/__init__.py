"""MI355X-native matrix-free Newton-Krylov hot path for Differentiable-FEA (the ``hip`` variant)."""
__version__ = "0.1.0"
