"""MI355X-native charged-LJ force engine for RBMD (random batch molecular dynamics).

The computation lives in librbmd_force.so (hand-written HIP kernels for gfx950
behind the C ABI of include/rbmd_force.h); this package is the host-side
mirror of the reference's Force interface plus input generators.
"""
from .force import ForceConfig, LJCutCoulKspaceForce, RBMDError  # noqa: F401

__all__ = ["ForceConfig", "LJCutCoulKspaceForce", "RBMDError"]
