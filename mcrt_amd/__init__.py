"""MI355X-native (gfx950) ray-query core and wavefront path tracer with the
scene / camera / render interface of the MCRT reference.

The compute path is libmcrt_hip.so (hand-written HIP kernels behind the C
ABI in include/mcrt.h); this package holds its ctypes binding, the scene
generators used by the benchmarks and tests, and the build recipe.
"""
from .api import (DEVICE_PTRS, DIFF, GLSY, LGHT, BuildStats, Context, McrtError, RenderParams, Scene, lib,
                  split_hits, tonemap_gamma, write_hdr, write_ppm, Batch, render_batch, obj_read)

__all__ = ["Context", "Scene", "McrtError", "BuildStats", "RenderParams", "lib", "split_hits", "LGHT", "DIFF",
           "GLSY", "DEVICE_PTRS", "write_hdr", "write_ppm", "tonemap_gamma", "Batch", "render_batch",
           "obj_read"]
