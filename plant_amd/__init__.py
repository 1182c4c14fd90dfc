"""MI355X canonical hub labeling engine (C++ host + HIP kernels for gfx950).

The compute lives in libchl.so behind the C ABI of include/chl.h; plant_amd.chl
is its Python binding.
"""
from . import chl  # noqa: F401
