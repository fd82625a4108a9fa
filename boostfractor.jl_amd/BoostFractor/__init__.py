"""BoostFractor on MI355X: drop-in mirror of BoostFractor.jl's public API.

The module layout follows the reference (src/BoostFractor.jl:5-20): the
coordinate/boundary data model, the propagator plug-ins, the recursive
Fourier solvers and the mode-matching transformer.  Every numerical call runs
in libboostfractor_hip.so (hand-written gfx950 HIP kernels) through its C ABI
(include/boostfractor_hip.h); there is no CPU fallback.
"""
from .setup import (CoordinateSystem, DimensionMismatch, SeedCoordinateSystem, SeedSetupBoundaries,
                    SetupBoundaries, get_kspace_coords, wavelength)
from .jlrange import julia_range
from ._lib import ArgumentError, BoundsError, HipError, InexactError, UnsupportedError, Context, context, device_count
from .beam import gauss_profile, unnormalized_match
from .propagators import propagator, propagator1D, propagatorMomentumSpace, propagatorNoTilts
from .recurprop import DancerPlan, cheerleader, dance_intro, dancer, dancer_batch
from .transformer import (ErrorException, Modes, SeedModes, axion_induced_modes, calc_propagation_matrices,
                          e_field_dimensions, field2modes, index, mode, modes2field, propagation_matrix, transformer,
                          TransformerPlan, TransformerBatchPlan, transformer_batch, transformer_trace_back)
from . import analytical1d as Analytical1D  # noqa: N812  (the reference module's name, src/Analytical1D.py)
from .sweep import (SpacingInterpolation, calc_boostfactor_cost, calc_boostfactor_cost_batch, calc_boostfactor_modes,
                    calc_modes, calc_propagation_matrices_grid, construct_prop_matrix_interpolation,
                    copy_setup_boundaries, interpolate_prop_matrix, p_norm, soft_box_penalty)

__all__ = [
    "CoordinateSystem", "SeedCoordinateSystem", "get_kspace_coords", "SetupBoundaries", "SeedSetupBoundaries",
    "wavelength", "julia_range", "DimensionMismatch", "ArgumentError", "HipError", "UnsupportedError", "Context",
    "context", "device_count", "propagator", "propagatorNoTilts", "propagatorMomentumSpace", "propagator1D",
    "dancer", "dancer_batch", "dance_intro", "cheerleader", "DancerPlan", "Modes", "SeedModes", "mode", "index",
    "e_field_dimensions", "axion_induced_modes", "field2modes", "modes2field", "propagation_matrix",
    "calc_propagation_matrices", "transformer", "TransformerPlan", "TransformerBatchPlan", "transformer_batch", "transformer_trace_back", "ErrorException", "InexactError",
    "gauss_profile", "unnormalized_match", "BoundsError", "calc_propagation_matrices_grid", "SpacingInterpolation",
    "construct_prop_matrix_interpolation", "interpolate_prop_matrix", "calc_boostfactor_modes", "calc_modes",
    "calc_boostfactor_cost", "calc_boostfactor_cost_batch", "soft_box_penalty", "p_norm", "copy_setup_boundaries",
    "Analytical1D",
]
