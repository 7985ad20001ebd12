"""MI355X-native MHFEM per-time-step path (host side).

``capi`` binds libmhfem_amd.so (include/mhfem.h); ``problems`` and ``meshgen``
build the synthetic inputs of the benchmark configurations.
"""
from . import meshgen, problems  # noqa: F401

__all__ = ["meshgen", "problems", "capi"]
