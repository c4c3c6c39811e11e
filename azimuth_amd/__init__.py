"""azimuth_amd: the azimuth render hot path on MI355X (gfx950).

`native` binds libazimuth_gpu.so (include/azimuth_gpu.h, azimuth_scene.h);
`scene` wraps scene building and GPU contexts; `configs` builds the seeded
procedural workloads.
"""
from . import native  # noqa: F401

__all__ = ["native"]
