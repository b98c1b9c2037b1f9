"""MI355X-native Barnes-Hut force step (gfx950 HIP kernels behind the C ABI of include/bh.h)."""
from .solver import BHContext, NBodySolverHipBH, Points  # noqa: F401
