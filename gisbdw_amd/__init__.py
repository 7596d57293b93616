"""MI355X-native exact treedepth: subset DP and branch-and-bound in HIP.

The compute lives in libtdhip.so (gisbdw_amd/csrc, C ABI in include/tdhip.h);
this package holds the host-side mirror of the reference's classes and the
graph utilities.
"""
from .graphs import Graph  # noqa: F401

__all__ = ["Graph", "api", "graphs"]
