"""Vanadis DR-HEX8 solve path on MI355X (gfx950).

Python host mirror of the reference's solver interface:
  kernel_dbcg_simple(...)  -- the legacy drop-in (VANADIS.FOR:809-810 ->
                              src_v2026_2_1/cudatest.cu:301-305)
  Context                  -- the device-resident matrix-free context API
  Batch                    -- several contexts solved together (vgpu_batch_*)
  Net, learned             -- learned initial guess for implicit time stepping
  autotune                 -- tile / z-chunk tuner with a JSON cache
  problem                  -- meshes, masks and synthetic problems (host)
The compute path is libvanadis_gpu.so (hand-written HIP); it raises if absent.
"""
from . import autotune, learned, problem  # noqa: F401
from ._lib import (  # noqa: F401
    BREAKDOWN, DR_V1, DR_V2, DR_V3, ERR_ARG, ERR_MESH, ERR_STATE, EXPORTED, FP32, FP64, MIXED,
    NOT_CONVERGED, OK, SCATTER_ATOMIC, SCATTER_COLORED, SCATTER_OWNER, SCHEME_EULER,
    SCHEME_STEADY, SCHEME_THETA, Batch, Context, Net, VgpuError, device_count, kernel_dbcg_simple, load,
)

__all__ = ["Batch", "Context", "Net", "autotune", "kernel_dbcg_simple", "problem", "load",
           "VgpuError"]
