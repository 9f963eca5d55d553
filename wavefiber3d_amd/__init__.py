"""wavefiber3d_amd -- MI355X-native physical-optics far-field solver.

Drop-in for WaveFiber3d's PO path: surface samples + incidences in, far-field
intensity / complex E grids out, computed by hand-written HIP kernels for
gfx950 behind the C-ABI in include/wf_po.h (libwfpo.so).
"""
from ._lib import WfError, lib  # noqa: F401
from .geometry import SurfaceSamples, fibonacci_sphere, fiber_segment, read_soa_mesh  # noqa: F401
from .po import (  # noqa: F401
    Batch,
    Incidence,
    POSolver,
    output_directions,
    reference_incidence_phi,
    reference_incidence_theta,
    reference_lambda,
)

__all__ = [
    "WfError", "lib", "SurfaceSamples", "fibonacci_sphere", "fiber_segment", "read_soa_mesh",
    "Batch", "Incidence", "POSolver", "output_directions", "reference_incidence_phi",
    "reference_incidence_theta", "reference_lambda",
]
