"""gpu_mm_amd: planned pointing operators (map2tod / tod2map) for AMD MI355X.

See gpu_mm_amd.gpu_mm for the data model and API, and gpu_mm_amd.mapmaker for the differentiable
operators and the preconditioned CG map solver.
"""

from .gpu_mm import (  # noqa: F401
    map2tod, tod2map, LocalPixelization, LocalMap, PointingPrePlan, PointingPlan,
    ToyPointing, PointingPlanTester, read_xpointing_npzfile, test_plan_iterator,
    make_random_plan_mt, set_errflag_check, get_errflag_check, mm_dtype, encode_tod16,
    kernel_timer_enable, kernel_timer_read, stream_read, DynamicMap, expand_dynamic_map, cell_broadcast,
    cell_reduce, PlanBatch, BatchConfig, deferred_errflags,
)
from . import gpu_mm  # noqa: F401
from . import mapmaker  # noqa: F401

__doc__ = gpu_mm.__doc__
