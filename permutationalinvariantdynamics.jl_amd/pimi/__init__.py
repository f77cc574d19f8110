"""pimi: MI355X backend for the permutation-invariant Liouvillian hot path.

The numeric work runs in libpimi (HIP kernels for gfx950 behind the C ABI in
include/pimi.h); this package mirrors the reference package's API around it.
"""
from ._native import PIMI_C32, PIMI_C64, NotConverged, PIMIError  # noqa: F401
from .api import (  # noqa: F401
    CollectiveHamiltonian, CollectiveJump, CollectiveOperator, CollectiveProduct, CompiledPIModel,
    DirectPIHamiltonian, DynamicsStreamResult, ExpvAlgorithm, GMRESAlgorithm, LocalHamiltonian, LocalJump, Models,
    PIBasis, PIModel, PIOperator, PIState, SteadyStateResult, collective_operator, compile,
    computational_product_state, expect, iid_state, qubit_ensemble_model, saved_times, solve_dynamics, spin_matrices,
    stationary_state,
)
from .batch import BatchAlgorithm, BatchEngine, batched_stationary_states  # noqa: F401
from .trajectories import (  # noqa: F401
    batched_apply_jumps, batched_channel_intensities, batched_conditional_action, batched_conditional_rk4,
    batched_sample_jumps, batched_trajectory_rngs, trajectory_channels,
)
from .checkpoints import (  # noqa: F401
    PI_CHECKPOINT_VERSION, PIStateCheckpoint, checkpoint_state, load_checkpoint, make_checkpoint, save_checkpoint,
)
from .families import (  # noqa: F401
    CompiledPIModelFamily, ParameterScanPlan, ParameterScanPoint, ParameterScanResult, SpecializedPIModel,
    compile_family, parameter_scan, specialize,
)
from .spectral import evolve, krylov_liouvillian_spectrum, time_evolve  # noqa: F401
