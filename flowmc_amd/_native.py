"""ctypes binding of ``libflowmc_hip.so`` (the C ABI in ``include/flowmc.h``).

The library is the only compute path: there is no Python or CPU fallback.
If the shared library is missing, or a tensor argument is not a contiguous
tensor of the right dtype on a HIP device, the call raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import torch

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libflowmc_hip.so"

ABI_VERSION = 3
FLOW_MAX_HIDDEN = 4  # FMC_FLOW_MAX_HIDDEN of include/flowmc.h
NF_MH_NOALLOC_MAX = 1 << 22  # FMC_NF_MH_NOALLOC_MAX of include/flowmc.h


class FlowMCNativeError(RuntimeError):
    """A C-ABI call returned a non-zero fmc_status."""

    def __init__(self, fn: str, status: int, message: str):
        super().__init__(f"{fn} failed with status {status}: {message}")
        self.fn = fn
        self.status = status
        self.message = message


# --------------------------------------------------------------- ABI types --
class fmc_target(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_dim", C.c_int32),
        ("n_modes", C.c_int32),
        ("reserved", C.c_int32),
        ("params", C.c_void_p),
        ("n_params", C.c_int64),
    ]


class fmc_launch_cfg(C.Structure):
    _fields_ = [
        ("lanes_per_chain", C.c_int32),
        ("vec_per_lane", C.c_int32),
        ("block_size", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fmc_hmc_slot(C.Structure):
    _fields_ = [
        ("target", fmc_target),
        ("x", C.c_void_p),
        ("logp", C.c_void_p),
        ("metric", C.c_void_p),
        ("momentum_chol", C.c_void_p),
        ("noise", C.c_void_p),
        ("log_u", C.c_void_p),
        ("out_pos", C.c_void_p),
        ("out_logp", C.c_void_p),
        ("out_acc", C.c_void_p),
        ("n_chains", C.c_int64),
        ("buf_len", C.c_int64),
        ("cursor", C.c_int64),
        ("key0", C.c_uint32),
        ("key1", C.c_uint32),
        ("lanes_per_chain", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fmc_work_item(C.Structure):
    _fields_ = [
        ("slot", C.c_int32),
        ("chain0", C.c_int32),
        ("n_chains", C.c_int32),
        ("reserved", C.c_int32),
    ]


class fmc_flow(C.Structure):
    _fields_ = [
        ("n_dim", C.c_int32),
        ("n_layers", C.c_int32),
        ("n_bins", C.c_int32),
        ("n_hidden", C.c_int32),
        ("hidden", C.c_int32 * 4),  # FMC_FLOW_MAX_HIDDEN
        ("precision", C.c_int32),
        ("flags", C.c_int32),
        ("range_min", C.c_float),
        ("range_max", C.c_float),
        ("min_bin_size", C.c_float),
        ("min_knot_slope", C.c_float),
        ("theta", C.c_void_p),
        ("packed", C.c_void_p),
        ("packed_bias", C.c_void_p),
    ]


FLOW_DENSE_BASE = 1  # FMC_FLOW_DENSE_BASE


def make_flow(n_dim, n_layers, n_bins, hidden, precision, range_min=-10.0, range_max=10.0, min_bin_size=1e-4,
              min_knot_slope=1e-4, theta=None, packed=None, packed_bias=None, flags=0) -> "fmc_flow":
    """fmc_flow descriptor with the conditioner's hidden widths given as a list."""
    hidden = [int(h) for h in hidden]
    if not 1 <= len(hidden) <= FLOW_MAX_HIDDEN:
        raise ValueError(f"the flow kernels take 1 to {FLOW_MAX_HIDDEN} hidden layers, got {len(hidden)}")
    hw = (C.c_int32 * FLOW_MAX_HIDDEN)(*(hidden + [0] * (FLOW_MAX_HIDDEN - len(hidden))))
    return fmc_flow(n_dim, n_layers, n_bins, len(hidden), hw, precision, flags, range_min, range_max, min_bin_size,
                    min_knot_slope, theta, packed, packed_bias)


class fmc_rnvp(C.Structure):
    _fields_ = [
        ("n_dim", C.c_int32),
        ("n_layers", C.c_int32),
        ("n_hidden", C.c_int32),
        ("dt", C.c_float),
        ("theta", C.c_void_p),
        ("packed", C.c_void_p),
        ("layout", C.c_void_p),
    ]


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
U32 = C.c_uint32
F32 = C.c_float

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGNATURES: dict[str, list] = {
    "fmc_abi_version": [],
    "fmc_last_error": [],
    "fmc_device_query": [C.c_int, P, P, P, P, P],
    "fmc_target_check": [C.POINTER(fmc_target)],
    "fmc_target_eval": [P, C.POINTER(fmc_target), P, I64, P, P, C.POINTER(fmc_launch_cfg)],
    "fmc_mala_run": [P, C.POINTER(fmc_target), P, P, I64, I32, I32, F32, U32, U32, P, P, P, P, P, I64, I64, I32,
                     C.POINTER(fmc_launch_cfg)],
    "fmc_grw_run": [P, C.POINTER(fmc_target), P, P, I64, I32, I32, F32, U32, U32, P, P, P, P, P, I64, I64, I32,
                    C.POINTER(fmc_launch_cfg)],
    "fmc_hmc_run": [P, C.POINTER(fmc_target), P, P, I64, I32, I32, F32, I32, P, P, U32, U32, P, P, P, P, P, I64,
                    I64, I32, C.POINTER(fmc_launch_cfg)],
    "fmc_hmc_run_batched": [P, P, I32, P, I32, I32, I32, I32, F32, I32, I32, I32, I32, I32],
    "fmc_tempered_run": [P, I32, C.POINTER(fmc_target), P, P, I64, I32, P, I32, I32, F32, I32, P, P, U32, U32, P, P,
                         P, I64, C.POINTER(fmc_launch_cfg)],
    "fmc_pt_exchange": [P, P, P, P, P, I64, I32, I32, U32, U32, P, P],
    "fmc_finite_rows": [P, P, I64, I32, P],
    "fmc_rnvp_sizes": [C.POINTER(fmc_rnvp), P, P, P, P],
    "fmc_rnvp_pack": [P, C.POINTER(fmc_rnvp)],
    "fmc_rnvp_pass": [P, C.POINTER(fmc_rnvp), I32, I32, P, I64, U32, U32, I64, P, P, P, P, P],
    "fmc_rnvp_backward_layer": [P, C.POINTER(fmc_rnvp), I32, P, I64, P, P, P, P, I64],
    "fmc_rnvp_weight_grad": [P, C.POINTER(fmc_rnvp), I32, P, I64, I64, P],
    "fmc_pt_adapt": [P, P, P, I64, I32],
    "fmc_mala_propose": [P, P, P, P, I64, I32, F32, U32, U32, I32, P],
    "fmc_mala_accept": [P, P, P, P, P, P, P, I64, I32, F32, U32, U32, I32, P, P, P, P, I64, I64, I32],
    "fmc_grw_propose": [P, P, P, I64, I32, F32, U32, U32, I32, P],
    "fmc_grw_accept": [P, P, P, P, P, I64, I32, U32, U32, I32, P, P, P, P, I64, I64, I32],
    "fmc_hmc_momentum": [P, P, P, I64, I32, P, P, U32, U32, I32, P],
    "fmc_hmc_drift": [P, P, P, P, I64, I32, F32],
    "fmc_hmc_kick": [P, P, P, I64, I32, F32],
    "fmc_hmc_accept": [P, P, P, P, P, P, P, P, I64, I32, U32, U32, I32, P, P, P, P, I64, I64, I32],
    "fmc_nf_mh": [P, P, P, P, P, P, P, I64, I32, I32, U32, U32, P, P, P, P, P, P, P, I64, I64, I32],
    "fmc_nf_mh_noalloc": [P, P, P, P, P, P, P, I64, I32, I32, U32, U32, P, P, P, P, P, P, I64, I64, I32],
    "fmc_philox_blocks": [P, U32, U32, U32, U32, U32, U32, I64, P],
    "fmc_rng_normal": [P, U32, U32, I64, I32, I32, I32, I32, P],
    "fmc_rng_log_uniform": [P, U32, U32, I64, I32, I32, I32, P],
    "fmc_flow_sizes": [C.POINTER(fmc_flow), P, P, P, P, P, P],
    "fmc_flow_pack": [P, C.POINTER(fmc_flow)],
    "fmc_flow_pass": [P, C.POINTER(fmc_flow), I32, I32, P, I64, U32, U32, I64, P, P, P, P, P],
    "fmc_flow_backward_layer": [P, C.POINTER(fmc_flow), I32, P, I64, P, P, P, P, I64, P, I32],
    "fmc_flow_weight_grad": [P, C.POINTER(fmc_flow), I32, P, I64, P, I32, P],
    "fmc_flow_weight_grad_ws_size": [C.POINTER(fmc_flow), I64, C.POINTER(C.c_int64)],
    "fmc_flow_weight_grad_split": [P, C.POINTER(fmc_flow), I32, P, I64, P, I32, P, P, I64],
    "fmc_adamw_clip": [P, P, P, P, P, I64, F32, F32, F32, F32, F32, F32, I32, P, I32, P],
    "fmc_adamw_clip_dev": [P, P, P, P, P, I64, F32, F32, F32, F32, F32, F32, P, P, I32, P],
    "fmc_flow_backward": [P, P, P, P, I64, P, P, P, P, P, I64, P, P, I32, P, I32, P, I64],
}
_RESTYPES = {"fmc_last_error": C.c_char_p}

_lib: Optional[C.CDLL] = None


def exported_names() -> list[str]:
    return list(_SIGNATURES)


def load(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Load (once) and type the shared library. Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"flowmc_amd native library not found at {p}; build it with `python -m flowmc_amd.build`"
        )
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if an export is missing
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, C.c_int)
    v = lib.fmc_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libflowmc_hip ABI version {v} != expected {ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    return load().fmc_last_error().decode(errors="replace")


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise FlowMCNativeError(name, rc, last_error())


# ---------------------------------------------------------- tensor checks --
def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("flowmc_amd kernels need a HIP GPU (torch.cuda.is_available() is False)")


def dptr(t: Optional[torch.Tensor], dtype: torch.dtype = torch.float32, name: str = "tensor",
         numel: Optional[int] = None) -> Optional[int]:
    """Device pointer of a contiguous HIP tensor (None passes through as NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the HIP device, got {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")
    return t.data_ptr()


def stream_ptr(stream: Optional[torch.cuda.Stream] = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
