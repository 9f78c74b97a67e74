"""MI355X-native CTMRG hot path of TeNeS (iTPS corner-transfer-matrix
renormalisation) on hand-written gfx950 HIP kernels behind a C ABI."""
from .core import (  # noqa: F401
    Context,
    Environment,
    Tensor,
    TenesError,
    calc_next_ctm,
    calc_next_et,
    calc_projector_left_block,
    calc_projector_updown_blocks,
    create_environment_two_sites,
    peps_parameters,
    simple_update_bond,
    svd_top,
    tensordot,
    tm_parameters,
    transpose,
)

__all__ = [
    "Context", "Environment", "Tensor", "TenesError", "calc_next_ctm", "calc_next_et",
    "calc_projector_left_block", "calc_projector_updown_blocks",
    "create_environment_two_sites", "peps_parameters", "simple_update_bond", "svd_top", "tensordot",
    "tm_parameters", "transpose",
]
