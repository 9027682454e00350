"""vfvm — MI355X-native hot path of vfvchem-vfvm3d (edge-flux assembly + GMRES).

The compute runs in lib/libvfvm.so (HIP kernels for gfx950 behind the C ABI of
include/vfvm.h); this package is the host-side mirror of the reference's API.
"""
from ._lib import LIB_PATH, VFVChemError, lib  # noqa: F401  (fails loudly without the library)
from .core import (FIELD_CONSTANT, FIELD_FUNCTION, FIELD_UNKNOWN, FP32_REFINED, FP64, K_ASSEMBLE, K_MDOT, K_PRECOND, K_UPDATE,  # noqa: F401
                   K_SPMV, ORTHO_AUTO, ORTHO_CGS, ORTHO_CGS2, PRECOND_ILU0, PRECOND_JACOBI, PRECOND_NONE, WRT_INITIAL, WRT_SCLBND,
                   WRT_SCLDOM, WRT_VECBND, Context, Mesh, solve_opts)
from .core import (OP_SCL_BND_ADVECTION_UNITY, OP_SCL_BND_DIFFUSION, OP_SCL_BND_DIRICHLET, OP_SCL_BND_NEUMANN,  # noqa: F401
                   OP_SCL_BND_PRODUCTION, OP_SCL_BND_SOURCE, OP_SCL_BND_TIME, OP_SCL_DOM_ADVECTION_UNITY,
                   OP_SCL_DOM_DIFFUSION, OP_SCL_DOM_PRODUCTION, OP_SCL_DOM_SOURCE, OP_SCL_DOM_TIME)
from .core import (OP_SCL_BND_ADVECTION, OP_SCL_BND_TIME_PREVIOUS, OP_SCL_BND_TIME_UNITY, OP_SCL_DOM_ADVECTION,  # noqa: F401
                   OP_SCL_DOM_TIME_PREVIOUS, OP_SCL_DOM_TIME_UNITY, OP_VEC_DOM_TIME)
from .core import (FLW_FLUX, FLW_MOM, FLW_RC, OP_SCL_BND_DIVERGENCE, OP_SCL_BND_PRESSURE_STABILIZER,  # noqa: F401
                   OP_SCL_DOM_DIVERGENCE, OP_SCL_DOM_PRESSURE_REFERENCE, OP_SCL_DOM_PRESSURE_STABILIZER, OP_VEC_BND_DIRICHLET,
                   OP_VEC_DOM_ADVECTION, OP_VEC_DOM_DIFFUSION, OP_VEC_DOM_DILATATORY, OP_VEC_DOM_GRADIENT, OP_VEC_DOM_SOURCE,
                   OP_VEC_DOM_TRANSPOSE, PRECOND_BJACOBI)
from .core import (OP_SCL_ITF_ADVECTION, OP_SCL_ITF_ADVECTION_UNITY, OP_SCL_ITF_CONTINUITY, OP_SCL_ITF_DIFFUSION,  # noqa: F401
                   OP_SCL_ITF_PRODUCTION, OP_SCL_ITF_SOURCE, OP_SCL_ITF_TIME, OP_SCL_ITF_TIME_PREVIOUS, OP_SCL_ITF_TIME_UNITY)
from .api import (LinearGmres, SteadyFlow, SteadyHeat, SteadyHeatFlow, SteadyMass, SteadyMassFlow, TransientFlow,  # noqa: F401
                  TransientHeat, TransientHeatFlow, TransientHeatMassFlow, Variables)

__version__ = "0.1.0"
