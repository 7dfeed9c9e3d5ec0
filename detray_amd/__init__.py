"""detray_amd: MI355X-native track propagation (detray's propagate hot path)
behind a C ABI, with a thin Python mirror of the reference interface."""

from .adjoint import AdjointConfig, Chi2Propagator, Hits, chi2  # noqa: F401
from .capi import DetrayError, load  # noqa: F401
from .propagator import (  # noqa: F401
    Actors,
    GEV,
    MEV,
    MM,
    TESLA,
    UM,
    Detector,
    LaunchConfig,
    MagneticField,
    PropagationConfig,
    PropagationResult,
    PropagationState,
    Propagator,
    StepTrace,
    generate_tracks,
    material,
)

__all__ = [
    "AdjointConfig", "Chi2Propagator", "Hits", "chi2", "DetrayError", "load", "Detector", "LaunchConfig", "MagneticField", "PropagationConfig",
    "PropagationResult", "PropagationState", "Propagator", "StepTrace", "Actors", "generate_tracks", "material", "MM", "UM", "GEV", "MEV", "TESLA",
]
