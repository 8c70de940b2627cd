"""MI355X (gfx950) engine for MDSCTK's knn_rms path.

The engine is native code: hand-written HIP kernels behind the C ABI in
include/mdsctk_knn.h, built into mdsctk_amd/lib/libmdsctk_knn.so, plus the
knn_rms drop-in executable in mdsctk_amd/bin/.  This package only binds and
builds them.
"""
from .knn import KnnContext, KnnError, load  # noqa: F401
