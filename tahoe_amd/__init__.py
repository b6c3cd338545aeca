"""MI355X-native implementation of Tahoe's continuum-solid Newton step.

The numerical work runs in ``libtahoe_gpu.so`` (hand-written HIP kernels for
gfx950 behind a C ABI declared in ``include/tahoe_gpu.h``); ``tahoe_amd.api``
loads it with ctypes.  ``tahoe_amd.mesh`` holds the seeded mesh generator.
"""
