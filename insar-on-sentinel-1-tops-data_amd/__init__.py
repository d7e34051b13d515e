"""MI355X-native TOPS burst core: deramp -> 12-tap sinc resample -> reramp ->
coherence, with DEM-assisted coregistration and ESD, behind the reference's
C-ABI (include/tops_abi.h).

The compute lives in libtops_mi355x.so (HIP, gfx950); this package holds its
ctypes bindings (native), the synthetic scenario generator (synth) and the
stage driver over subswaths in host or device memory (pipeline, imported on
demand: it needs torch for device buffers).
"""
from . import native, synth  # noqa: F401

__all__ = ["native", "synth"]
