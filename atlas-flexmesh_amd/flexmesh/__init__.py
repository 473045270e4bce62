"""MI355X (gfx950) quasi-static Newton path: P1-tet Neo-Hookean residual and
tangent, colored CSR assembly, ILU(0)-PCG and a mixed-precision Newton loop,
all in hand-written HIP behind the C ABI of include/flexmesh/fm_api.h."""
from ._lib import (FM_EDEVICE, FM_EHIP, FM_EINVAL, FM_ENOCONV, FM_ESTATE, FM_FP64, FM_MIXED, FM_MIXED_PC,  # noqa: F401
                   FM_OK, FlexMeshError, Solver, field_lognormal, lib, mesh_box, mesh_jitter)
from . import configs  # noqa: F401
