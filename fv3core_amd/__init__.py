"""fv3core_amd: the FV3 dynamical-core hot path on AMD MI355X (gfx950).

Component classes keep the reference's (ai2cm/fv3core) constructor and call
signatures; their bodies call hand-written HIP kernels through the C ABI in
``include/fv3core_amd.h`` (library ``libfv3core_amd.so``, built in-tree by
``fv3core_amd.build``).
"""

__version__ = "0.1.0"
