"""Named tolerances of the tests (the reference's tests/tolerances.py plus the mixed modes)."""

# The reference's values (tests/tolerances.py:3-36 of the reference).
ENERGY_TOL = 1e-10
STRUCTURE_TOL = 1e-10
DENSITY_FITTING_TOL = 1e-10
FITTED_INTEGRAL_TOL = 1e-2
MO_ENERGY_TOL = 1e-7
EIGEN_TOL = 1e-9

# FP64 GPU J/K against the NumPy oracle: normwise max relative error.
FP64_REL_TOL = 1e-10
# FP32 products on MFMA with FP32 accumulation inside a tile.
FP32_REL_TOL = 2e-5
# BF16 hi/lo split products with FP32 accumulation.
BF16X3_REL_TOL = 2e-4
