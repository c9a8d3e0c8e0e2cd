"""MI355X-native 1 ms network step of the cerebellar simulator.

Host mirror of the reference's CBMSimCore / CBMState API over the C-ABIs in
include/cbm_hip.h (HIP kernels) and include/cbm_state.h (network state).
"""
from ._abi import FIELDS, ActParams, CbmError, Options, Sizes, hip_lib, state_lib
from .network import Batch, Network, NetState, act_params_from_text, stream_ceiling

__all__ = ["FIELDS", "ActParams", "CbmError", "Options", "Sizes", "hip_lib", "state_lib", "Network",
           "NetState", "act_params_from_text", "Batch", "stream_ceiling"]
