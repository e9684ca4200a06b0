"""gs_hip -- host runtime of the MI355X Gaussian rasteriser (libgs_hip.so).

Submodules:
  _lib     ctypes binding of include/gs_raster.h
  raster   torch-facing forward / backward / kNN / sort entry points
  camera   camera matrices as CF-3DGS builds them (scene/cameras.py)
  scenes   seeded synthetic workloads
"""
