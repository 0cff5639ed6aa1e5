"""steps_amd -- MI355X-native direct-summation gravity force engine for StePS.

The product is the HIP library steps_amd/lib/libsteps_amd.so behind the C ABI
in include/steps_force.h (and the C++ drop-in shim in csrc/shim).  This Python
package only binds that ABI (ctypes) and generates the synthetic workloads.
"""
from ._abi import (  # noqa: F401
    EXPORTED, STEPS_ERR_STATE, PREC_FP64, PREC_MIXED, REQ_FORCE, REQ_STEP, TOPO_R3, TOPO_S1R2, TOPO_T3, ForceDesc, Server,
    Sim, StepParams, StepsError,
    device_count, force_eval, load_library, make_desc, radial_table, s1r2_table,
    softening_lengths, t3_table,
)
