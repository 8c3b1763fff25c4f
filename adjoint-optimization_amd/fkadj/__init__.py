"""fkadj -- MI355X Fenton-Karma adjoint-gradient engine (Python side of libfkadj.so)."""
from ._lib import (F32, F64, PULSE_KNOTS, PULSE_PIECEWISE, FkadjError, Model, Problem, lib, last_timing,  # noqa: F401
                   param_string, reference_model, set_batch_groups, set_profiling, set_rolling, strerror)
from .api import (Instance, Plan, Server, batch_forward, batch_grad, fk_forward_Et_mex, fk_grad_Et_mex,  # noqa: F401
                  footprint_bytes, mem_info, snapshot_count)
