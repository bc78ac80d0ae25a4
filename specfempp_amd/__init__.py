"""specfempp_amd — MI355X-native SPECFEM++ time-stepping path (2D elastic
isotropic P-SV) behind a C ABI (include/specfempp_amd.h)."""

from .config import (Config2D, Receivers2D, Sources2D, adjoint_case, batched_shots,  # noqa: F401
                     homogeneous_box, layered_box)
from .mesh import Mesh2D, box_mesh, elastic_moduli  # noqa: F401
from .solver import (Context, DomainKernels, Newmark, TimeMarching, simulate,  # noqa: F401
                     simulate_batch)

__version__ = "0.1.0"
