"""diffractom_amd: diffractom's single-phase forward/adjoint operator and SPDHG-Poisson solver, rebuilt for AMD
Instinct MI355X (gfx950) with hand-written HIP kernels behind a C ABI (include/diffractom_hip.h), keeping the
reference's Python call signatures."""
from .crystallography import Material, group_reflections_into_rings
from .grid import Grid, MatrixGrid
from .local_support import LocalSupport
from .operator import SinglePhaseForwardOperator, estimate_L_power
from .local_operator import LocalSinglePhaseForwardOperator
from .fista import FISTAPoisson
from .osem import OSEMPoisson
from .pdhg import PDHGPoisson
from .quadrature import QuadratureRule
from .spdhg import SPDHGPoisson
from .support import fov_support_mask

__all__ = ["Material", "Grid", "MatrixGrid", "SinglePhaseForwardOperator", "LocalSinglePhaseForwardOperator",
           "LocalSupport", "SPDHGPoisson", "OSEMPoisson",
           "PDHGPoisson", "FISTAPoisson",
           "estimate_L_power",
           "QuadratureRule", "fov_support_mask", "group_reflections_into_rings"]
