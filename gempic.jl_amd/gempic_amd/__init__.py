"""gempic_amd: MI355X (gfx950) drop-in for GEMPIC.jl's particle-field path.

The host API mirrors GEMPIC.jl (OneDGrid, ParticleGroup,
ParticleMeshCoupling1D, Maxwell1DFEM, HamiltonianSplitting, the operatorH*
sub-steps, strang_splitting!, solve_poisson!, write_step!). Every compute
call runs hand-written HIP kernels through the C ABI in
include/gempic_hip.h (lib/libgempic_hip.so).
"""
from .core import *  # noqa: F401,F403
from .core import __all__  # noqa: F401
from ._lib import GempicError, InexactError, LIB_PATH  # noqa: F401
from ._lib import load as load_library  # noqa: F401
