"""MI355X (gfx950) implementation of MCPNet's hot paths: all-pairs B-spline and
adaptive-partitioning mutual information, and the max-min products behind MCP2/3/4.

The compute lives in libmcpgpu.so (HIP kernels behind the C ABI in include/mcpgpu/mcpgpu.h);
this package binds it with ctypes (`Handle`) and holds the seeded synthetic-data generator.
"""
from ._lib import EXPORTS, LIB_PATH, PRECISIONS, Handle, MCPGError, load

__all__ = ["Handle", "MCPGError", "load", "LIB_PATH", "EXPORTS", "PRECISIONS"]
