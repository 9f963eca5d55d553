"""wavefiber3d_amd -- MI355X-native physical-optics far-field solver.

Drop-in for WaveFiber3d's PO path: surface samples + incidences in, far-field
intensity / complex E grids out, computed by hand-written HIP kernels for
gfx950 behind the C-ABI in include/wf_po.h (libwfpo.so).
"""
from ._lib import WfError, lib  # noqa: F401
from .geometry import (  # noqa: F401
    SurfaceSamples,
    fibonacci_sphere,
    fiber_segment,
    heightfield_samples,
    read_heightfield,
    read_soa_mesh,
)
from .po import (  # noqa: F401
    Batch,
    Incidence,
    POSolver,
    average_grid,
    output_directions,
    reference_incidence_phi,
    reference_incidence_theta,
    reference_lambda,
)
from .rough2d import Rough2dIncidence, read_crosssection, rough2d_solve  # noqa: F401

__all__ = [
    "WfError", "lib", "SurfaceSamples", "fibonacci_sphere", "fiber_segment", "read_soa_mesh",
    "heightfield_samples", "read_heightfield", "average_grid",
    "Batch", "Incidence", "POSolver", "output_directions", "reference_incidence_phi",
    "reference_incidence_theta", "reference_lambda", "Rough2dIncidence", "read_crosssection",
    "rough2d_solve",
]
