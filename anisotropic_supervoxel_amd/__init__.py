"""MI355X-native supervoxel hot path (anisotropic JFA + Lloyd), Python host.

The compute lives in libasv.so (hand-written gfx950 HIP kernels behind the C
ABI of include/asv.h); this package binds it with ctypes and mirrors the
reference's Image/Render stage interface for tests, the benchmark and tools.
"""
from .api import Clip, default_params, plan, synth_video, seeds_placed_for  # noqa: F401
