"""Global dtypes (torchquantum/macro.py:41-47): complex64 states, float32 parameters."""
import math

import numpy as np
import torch

C_DTYPE = torch.complex64
F_DTYPE = torch.float32
C_DTYPE_NUMPY = np.complex64
F_DTYPE_NUMPY = np.float32
INV_SQRT2 = 1 / math.sqrt(2)
