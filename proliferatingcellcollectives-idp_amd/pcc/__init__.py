"""Host-side Python binding of libpcc, the MI355X collision-pipeline library.

`pcc.engine.ParticleManager` mirrors the reference's
ParticleManager::step (code/cpp/src/simulation/ParticleManager.cpp:150-166)
over the C ABI declared in include/pcc.h.
"""
from .abi_types import (Bodies, Constraints, Params, StepResult, reference_params,  # noqa: F401
                        MODE_HARD, MODE_SOFT, PREC_FP64, PREC_MIXED)
