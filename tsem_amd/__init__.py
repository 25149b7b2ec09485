"""MI355X-native TSEM operator and solver path (hierarchical Jacobi basis on
tetrahedra): index maps, matrix-free element apply, p-multigrid, PCG and
mixed-precision refinement on gfx950, behind the C ABI of include/tsem.h."""
from .lib import TsemError, load  # noqa: F401
from .api import Context, Space, poisson_setup  # noqa: F401
