"""erlbm_amd -- MI355X-native E-RLBM collide-stream (D2Q9 / D3Q19).

HIP kernels for gfx950 behind the C ABI of ``include/erlbm.h``
(``liberlbm.so``, built in-tree by ``python -m erlbm_amd.build``); this
package is the Python host side over that ABI.  Importing it does not need a
GPU; using a lattice does, and there is no CPU fallback.
"""
from ._lib import (ARITH_EXACT, ARITH_MOMENT, COL_BGK, COL_EXT_REG, COL_MRT, COL_REG, FP16_SHIFTED, FP32_SHIFTED, FP32_SHIFTED_F32, FP64,
                   POST_COLLISION,
                   STATE, STOP_DENSITY, STOP_ENERGY, ErlbmError, device_count, hbm_copy_bench, lib)
from ._lib import FP_AUTO, INIT_REST, INIT_TAYLOR_GREEN, JOB_DTYPE, RESULT_DTYPE
from .lattice import (Lattice, dist_eq, init_double_shear, init_taylor_green, init_tg_map, obstacle_map_point_params,
                      obstacle_mask, taylor_green_error, tg_map_point_params)

__all__ = [
    "ARITH_EXACT", "ARITH_MOMENT", "COL_BGK", "COL_MRT", "COL_REG", "COL_EXT_REG", "FP64", "FP32_SHIFTED", "FP16_SHIFTED", "FP32_SHIFTED_F32", "STATE",
    "POST_COLLISION", "STOP_DENSITY", "STOP_ENERGY", "ErlbmError", "device_count", "hbm_copy_bench", "lib", "Lattice", "dist_eq",
    "init_double_shear", "init_taylor_green", "init_tg_map", "obstacle_mask", "tg_map_point_params",
    "obstacle_map_point_params", "FP_AUTO", "INIT_REST", "INIT_TAYLOR_GREEN", "JOB_DTYPE", "RESULT_DTYPE", "Batch", "taylor_green_error",
    "obstacle_batch", "obstacle_map_jobs", "run_obstacle_map", "run_tg_map", "tg_batch", "tg_map_jobs",
]

# The batched-map names are imported on first use, so that running the maps
# module as a program (python -m erlbm_amd.maps) does not import it twice.
_MAPS = ("Batch", "obstacle_batch", "obstacle_map_jobs", "run_obstacle_map", "run_tg_map", "tg_batch", "tg_map_jobs")


def __getattr__(name):
    if name in _MAPS:
        from . import maps
        return getattr(maps, name)
    raise AttributeError(f"module 'erlbm_amd' has no attribute {name!r}")
