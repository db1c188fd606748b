"""jaxtrace_amd: MI355X (gfx950) implementation of JAXTrace's particle-advection
and density hot path, behind the reference's own Python API.

Host code holds torch tensors on the ROCm device and calls hand-written HIP
kernels through the C ABI declared in ``include/jaxtrace_amd.h``.
"""
from ._lib import load as _load_library  # noqa: F401
from .fields import (StructuredGridSampler, StructuredTimeSeriesField, OctreeFEMTimeSeriesFieldOptimized,
                     TimeSeriesField)
from .boundary import (
    BoundaryCondition,
    periodic_boundary,
    reflective_boundary,
    clamping_boundary,
    absorbing_boundary_factory,
    mixed_boundary,
    continuous_inlet_boundary_factory,
    apply_periodic,
    clamp_to_bounds,
    reflect_boundary,
    absorbing_boundary,
    generate_inlet_grid_particles,
)
from .tracking import TrackerOptions, Trajectory, ParticleTracker, create_tracker
from .adjoint import track_rk4
from .seeding import random_seeds, uniform_grid_seeds
from .vtk_io import VTKUnstructuredTimeSeriesReader, load_octree_field, read_octree_series
from .density import HashGridNeighbors, SPHDensityEstimator, KDEEstimator, scott_bandwidth, silverman_bandwidth

__version__ = "0.1.0"

__all__ = [
    "TimeSeriesField", "StructuredGridSampler", "StructuredTimeSeriesField", "OctreeFEMTimeSeriesFieldOptimized",
    "BoundaryCondition", "periodic_boundary", "reflective_boundary", "clamping_boundary",
    "absorbing_boundary_factory", "mixed_boundary", "continuous_inlet_boundary_factory",
    "apply_periodic", "clamp_to_bounds", "reflect_boundary", "absorbing_boundary",
    "TrackerOptions", "Trajectory", "ParticleTracker", "create_tracker",
    "track_rk4", "HashGridNeighbors", "SPHDensityEstimator", "KDEEstimator", "scott_bandwidth", "silverman_bandwidth",
    "uniform_grid_seeds", "random_seeds", "generate_inlet_grid_particles",
    "VTKUnstructuredTimeSeriesReader", "load_octree_field", "read_octree_series",
]
