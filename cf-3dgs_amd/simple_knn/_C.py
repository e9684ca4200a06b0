"""``simple_knn._C.distCUDA2`` on gfx950.

``distCUDA2(points [P,3] float32 on the GPU) -> [P] float32``: the mean of the
squared Euclidean distances to the 3 nearest other points, exactly (the same
value a brute-force search gives).  CF-3DGS clamps it at 1e-7 and uses
log(sqrt(.)) as the isotropic initial scale (scene/gaussian_model_cf.py:194-196).
"""
from __future__ import annotations

try:
    from gs_hip.raster import knn3_meandist as _knn
except ImportError:
    import os as _os
    import sys as _sys
    _sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
    from gs_hip.raster import knn3_meandist as _knn


def distCUDA2(points):
    return _knn(points)
