"""gfx950-native exterior Burton-Miller BEM path (fused Galerkin P1/DP0 assembly,
blocked complex LU on MFMA, multi-RHS solves, FP64 refinement, field evaluation)
behind the C ABI declared in include/beat_bem.h."""
from .mesh import BoundaryMesh, icosphere, jittered_icosphere, load_gmsh22, fibonacci_sphere  # noqa: F401
from .phasor import (NEGATIVE_TIME_PHASOR, POSITIVE_TIME_PHASOR, burton_miller_coupling,  # noqa: F401
                     neumann_scale, outgoing_wavenumber)

__all__ = ["BoundaryMesh", "icosphere", "jittered_icosphere", "load_gmsh22", "fibonacci_sphere"]
