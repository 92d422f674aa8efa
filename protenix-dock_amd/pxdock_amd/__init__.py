"""pxdock_amd — MI355X (gfx950) docking hot path of Protenix-Dock.

Submodules:
  ffi      ctypes mirror of the include/pxdock.h descriptors
  data     receptor / ligand data model and the reference JSON format
  scoring  scoring-function JSON (pose_selection) -> pxd_score_params
  synth    seeded synthetic receptors, ligands and poses
  lib      binding of libpxdock.so (the HIP product path)
"""
__all__ = ["ffi", "data", "scoring", "synth", "lib"]
