"""MI355X-native PPHPC per-iteration step.

Python mirror of the C ABI in ``include/pphpc.h`` (the library itself is
``pphpc_amd/lib/libpphpc.so``, hand-written HIP for gfx950 plus a C host
layer). The names follow the reference simulator's interface:
``load_params`` / ``stats_save`` restate ``pp_load_params`` / ``pp_stats_save``
(opencl/pp/pp_common.c:39-240) and :class:`Simulation` wraps the loop of
``ppg_simulate`` (opencl/pp/pp_gpu.c:573-987). :class:`Batch` runs many
replicas of that loop together on one GPU.
"""
from .sim import (  # noqa: F401
    PPError,
    Params,
    Stats,
    Simulation,
    Batch,
    RUNNING,
    DONE,
    FAILED,
    load_params,
    stats_save,
    stats_rows_to_array,
    simulate_noalloc,
    strerror,
    library_path,
    PP_SUCCESS,
    PP_UNKNOWN_ARGS,
    PP_INVALID_ARGS,
    PP_LIBRARY_ERROR,
    PP_UNABLE_TO_OPEN_PARAMS_FILE,
    PP_INVALID_PARAMS_FILE,
    PP_UNABLE_SAVE_STATS,
    PP_OUT_OF_RESOURCES,
    PP_ENERGY_OVERFLOW,
    PARAM_KEYS,
    STAT_FIELDS,
)
