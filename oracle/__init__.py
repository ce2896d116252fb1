"""CPU oracle of the explicit DG Euler residual: test infrastructure only (see binding.py)."""
