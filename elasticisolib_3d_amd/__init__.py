"""MI355X-native elastic isotropic 3-D staggered-grid propagator.

Layers:
  - ``csrc/``        hand-written HIP kernels for gfx950 and the C ABI
                     (``include/elastic_iso_3d.h``)
  - ``propagator``   regular-grid propagator handle over the C ABI
"""
__version__ = "0.1.0"
