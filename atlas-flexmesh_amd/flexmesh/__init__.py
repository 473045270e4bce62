"""MI355X (gfx950) quasi-static Newton path: P1-tet Neo-Hookean residual and
tangent, colored CSR assembly, ILU(0)-PCG and a mixed-precision Newton loop,
all in hand-written HIP behind the C ABI of include/flexmesh/fm_api.h."""
from ._lib import (FM_FP64, FM_MIXED, FM_MIXED_PC, FlexMeshError, Solver, field_lognormal, lib, mesh_box,  # noqa: F401
                   mesh_jitter)
from . import configs  # noqa: F401
