"""MI355X-native explicit compressible-Euler residual loop on face-based polyhedral meshes.

`autan_amd.mesh` builds meshes and their FP64 set-up (include/mxmesh.h); `autan_amd.stepper`
drives the HIP residual, SSP-RK3 step and norms on one GPU through the C ABI of include/mxeuler.h.
"""

__all__ = ["mesh", "stepper"]
