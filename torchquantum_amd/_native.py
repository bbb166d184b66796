"""ctypes binding of ``libtqsv.so`` (the C ABI declared in include/tqsv.h).

The library is the only compute path: if it is missing, or a tensor is not
on a ROCm device, calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_int, c_int64, c_uint64, c_void_p, c_float, POINTER
from pathlib import Path

import torch

LIB_PATH = Path(os.environ.get("TQSV_LIB", Path(__file__).resolve().parent / "libtqsv.so"))

_lib = None

c_int_p = POINTER(c_int)
c_u64_p = POINTER(c_uint64)
c_i64_p = POINTER(c_int64)
c_float_p = POINTER(c_float)

# name -> (restype, argtypes); every entry point of include/tqsv.h
SIGNATURES = {
    "tqsv_version": (c_int, []),
    "tqsv_last_error": (ctypes.c_char_p, []),
    "tqsv_num_cu": (c_int, []),
    "tqsv_init_basis": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int64]),
    "tqsv_apply_1q": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_uint64, c_uint64, c_void_p, c_int64]),
    "tqsv_apply_2q": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_int, c_uint64, c_uint64, c_void_p, c_int64]),
    "tqsv_apply_diag": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_uint64, c_uint64, c_uint64, c_void_p, c_int64]),
    "tqsv_apply_swap": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_int, c_uint64, c_uint64]),
    "tqsv_apply_kq": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_int_p, c_uint64, c_uint64, c_void_p, c_int64]),
    "tqsv_gate_words": (c_int, []),
    "tqsv_run_gates": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_i64_p, c_int, c_void_p]),
    "tqsv_stage_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int_p, c_void_p, c_void_p, c_int, c_int]),
    "tqsv_stage_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int_p, c_void_p, c_void_p,
                                   c_float_p, c_int64, c_void_p, c_int64, c_int]),
    "tqsv_stage_adjoint_work_floats": (c_int64, [c_int_p, c_int, c_int64, c_int]),
    "tqsv_adjoint_work_floats": (c_int64, [c_int64, c_int]),
    "tqsv_adjoint_1q": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int, c_uint64, c_uint64, c_void_p,
                                c_int64, c_int, c_float_p, c_int64, c_int, c_void_p, c_int64]),
    "tqsv_adjoint_2q": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int, c_int, c_uint64, c_uint64,
                                c_void_p, c_int64, c_int, c_float_p, c_int64, c_int, c_void_p, c_int64]),
    "tqsv_adjoint_diag": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_uint64, c_uint64, c_uint64,
                                  c_void_p, c_int64, c_int, c_float_p, c_int64, c_int, c_void_p, c_int64]),
    "tqsv_expval_workspace_bytes": (c_int64, [c_int64, c_int, c_int]),
    "tqsv_expval_z_all": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_float_p, c_void_p, c_int64]),
    "tqsv_expval_z_all_noalloc": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_float_p]),
    "tqsv_expval_pauli": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_u64_p, c_u64_p, c_float_p, c_int64,
                                  c_void_p, c_int64]),
    "tqsv_pauli_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int, c_u64_p, c_u64_p, c_void_p,
                                 c_int]),
    "tqsv_pauli_sum_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int, c_u64_p, c_u64_p,
                                    c_void_p, c_float_p, c_float_p, c_int64]),
    "tqsv_pauli_sum_grad_workspace_floats": (c_int64, [c_int64, c_int]),
    "tqsv_batch_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int]),
    "tqsv_batch_queue": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                 c_int, c_int64, c_int]),
    "tqsv_batch_queue_item_bytes": (c_int, []),
    "tqsv_batch_item_bytes": (c_int, []),
    "tqsv_batch_zcirc_bytes": (c_int, []),
    "tqsv_batch_zexpval": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int64, c_float_p, c_float_p]),
    "tqsv_mg_item_bytes": (c_int, []),
    "tqsv_dev_gate_bytes": (c_int, []),
    "tqsv_build_dev_gates": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int64]),
    "tqsv_mg_to_grad": (c_int, [c_void_p, c_float_p, c_int64, c_int64, c_void_p, c_int, c_void_p]),
    "tqsv_sample_workspace_bytes": (c_int64, [c_int64, c_int]),
    "tqsv_sample": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int64, c_uint64, c_void_p, c_void_p, c_int64]),
    "tqsv_jit_compile": (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_void_p, c_int64, c_i64_p, c_void_p, c_int64]),
    "tqsv_jit_load": (c_int, [c_void_p, c_int64, c_int, POINTER(ctypes.c_char_p), POINTER(c_void_p),
                              POINTER(c_void_p)]),
    "tqsv_jit_unload": (c_int, [c_void_p]),
    "tqsv_jit_stage_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int_p, c_void_p,
                                       c_int, c_int]),
    "tqsv_jit_stage_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int,
                                       c_int_p, c_void_p, c_void_p, c_float_p, c_int64, c_void_p, c_int64, c_int]),
    "tqsv_gemm": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                          c_float, c_void_p, c_int, c_void_p, c_int]),
    "tqsv_relu_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "tqsv_colsum": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "tqsv_mse_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tqsv_adam_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float, c_float, c_float,
                               c_float, c_float, c_float]),
}


class NativeError(RuntimeError):
    pass


def lib_available() -> bool:
    return LIB_PATH.exists()


def lib():
    """Load (once) and return the ctypes library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(
                f"{LIB_PATH} is not built; run `python -m torchquantum_amd.csrc.build` "
                "(torchquantum_amd has no CPU fallback)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return {n for n in SIGNATURES if getattr(L, n, None) is not None}


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().tqsv_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise NativeError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args):
    check(getattr(lib(), name)(*args), name)


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def require_gpu(t: torch.Tensor, what: str = "tensor"):
    if t.device.type != "cuda":
        raise NotImplementedError(
            f"torchquantum_amd runs its statevector kernels on MI355X (ROCm) devices only; "
            f"{what} is on '{t.device}'. Create the QuantumDevice with device='cuda'.")


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def fptr(t: torch.Tensor):
    return ctypes.cast(t.data_ptr(), c_float_p)


def u64_array(vals):
    arr = (c_uint64 * max(1, len(vals)))(*[int(v) for v in vals])
    return arr


def int_array(vals):
    return (c_int * max(1, len(vals)))(*[int(v) for v in vals])
