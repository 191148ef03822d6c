"""MI355X-native element-pair view-factor path (drop-in for RadiativeViewFactor.jl's
``compute_view_factors`` GPU hook).  See ``api`` for the reference-shaped entry
points and ``csrc/`` for the HIP kernels and the C ABI."""
from .mesh import MeshData, SurfaceElement  # noqa: F401
