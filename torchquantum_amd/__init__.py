"""torchquantum_amd: an MI355X-native statevector engine behind the
torchquantum API (QuantumDevice, functional gates, Operators,
expval_joint_analytical / expval / MeasureAll).

All state arithmetic runs in hand-written HIP kernels (libtqsv.so, built
from torchquantum_amd/csrc by ``python -m torchquantum_amd.csrc.build``);
there is no CPU fallback.
"""
__version__ = "0.1.0"

from .macro import C_DTYPE, F_DTYPE
from . import gates
from . import functional
from .module import QuantumModule, QuantumModuleList, QuantumModuleDict, static_support
from .device import QuantumDevice, set_engine_options, get_engine_options
from .noisedevice import NoiseDevice
from .operator import *  # noqa: F401,F403
from .operator import op_name_dict, fixed_ops, parameterized_ops, Operator, Operation, Observable
from .measurement import (expval, MeasureAll, expval_joint_analytical, expval_joint_analytical_batch,
                          expval_pauli_sum, MeasureMultiPauliSum, gen_bitstrings, measure, sample_indices,
                          find_observable_groups, expval_joint_sampling, expval_joint_sampling_grouping)
from .layer import (Op1QAllLayer, Op2QAllLayer, EntanglePairwise, LayerTemplate0, RXYZCXLayer0, RandomLayer)
from .encoding import (Encoder, GeneralEncoder, PhaseEncoder, MultiPhaseEncoder, StateEncoder, AmplitudeEncoder,
                       MagnitudeEncoder, encoder_op_list_name_dict)
from .engine import EngineOptions, wait_for_compiled_stages
from .replay import CircuitReplay
from . import device as _device_mod

_device_mod._bind_methods()
