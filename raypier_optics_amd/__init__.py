"""MI355X (gfx950) implementation of raypier's tracing loop and gausslet
field summation, behind raypier.core's Python API.

Modules mirror raypier.core: ctracer, cfaces, cshapes, cdistortions,
cmaterials, cfields, tracer, fields.  Tracing and field summation run in
librpx.so (HIP, built in-tree by build_lib.build()).
"""
__version__ = "0.1.0"
