"""bfgpu — MI355X K-lane batched PHAST over the customized edge-based CCH.

Host-side mirror of butterfly-route's PHAST engine API; the computation runs
in the HIP library libbfgpu.so through the C ABI declared in include/bfgpu.h.
"""
from .graph import (INF, CchGraph, GridSpec, build_grid, cached_grid, weight_decode, weight_encode,
                    weight_width_choose)
from .engine import (K_LANES, BatchedPhastEngine, BatchedPhastResult, BatchedPhastStats, BfgError)
from .customize import (RelaxResult, RelaxStats, customize, original_arc_weights, triangle_relax_lex_parallel,
                        triangle_relax_parallel)

__all__ = ["INF", "CchGraph", "GridSpec", "build_grid", "cached_grid", "weight_decode", "weight_encode",
           "weight_width_choose", "K_LANES", "BatchedPhastEngine", "BatchedPhastResult", "BatchedPhastStats",
           "BfgError", "RelaxResult", "RelaxStats", "customize", "original_arc_weights",
           "triangle_relax_lex_parallel", "triangle_relax_parallel"]
