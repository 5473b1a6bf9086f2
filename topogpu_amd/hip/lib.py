"""ctypes binding of libtopohip.so (the C ABI declared in include/topogpu_hip.h).

The library is built in-tree by ``topogpu_amd.csrc.build``. Loading fails
loudly when it is missing: there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent.parent / "libtopohip.so"

TG_OP_HEX_F32, TG_OP_HEX_BF16, TG_OP_HEX_BF16X3, TG_OP_Q8_F32, TG_OP_STORED_F32, TG_OP_STENCIL_F32 = 1, 2, 3, 4, 5, 6
TG_OP_Q8_BF16, TG_OP_Q8_BF16X3, TG_OP_STENCIL_BF16 = 7, 8, 9
(TG_EP_STORE, TG_EP_DOT, TG_EP_RESID, TG_EP_RESID_NORM, TG_EP_CHEB, TG_EP_CHEB_R, TG_EP_RESID_UPDATE,
 TG_EP_CHEB_FIRST, TG_EP_RESID_FIRST, TG_EP_CHEB_LAST) = range(10)
OC_STATE_LEN = 32   # device doubles of the OC bisection state (TG_OC_STATE_LEN)
OC_MEAN_STATE_LEN = 80   # ... of the mean-volume OC (TG_OC_MEAN_STATE_LEN)

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_f32 = C.c_float
_f64 = C.c_double


class TgGrid(C.Structure):
    _fields_ = [("nelx", _i32), ("nely", _i32), ("nelz", _i32)]


class TgOp32(C.Structure):
    _fields_ = [("kind", _i32), ("grid", TgGrid), ("mask", _vp), ("weights", _vp), ("elem_mats", _vp),
                ("stencil", _vp)]


class TgEp32(C.Structure):
    _fields_ = [("mode", _i32), ("rhs", _vp), ("dinv", _vp), ("d", _vp), ("x", _vp), ("r", _vp),
                ("a1", _f32), ("a2", _f32), ("a3", _f32), ("dot", _vp)]


def ep(mode, rhs=None, dinv=None, d=None, x=None, r=None, a1=0.0, a2=0.0, a3=0.0, dot=None) -> TgEp32:
    """Epilogue descriptor from tensors (None -> NULL)."""
    return TgEp32(mode, ptr(rhs), ptr(dinv), ptr(d), ptr(x), ptr(r), a1, a2, a3, ptr(dot))


_SIGS = {
    "tg_init": [_vp],
    "tg_mapped_pointer": [_vp, _vp],
    "tg_last_error": None,
    "tg_version": [],
    "tg_apply32": [C.POINTER(TgOp32), _vp, _vp, C.POINTER(TgEp32), _vp],
    "tg_graph_begin": [_vp],
    "tg_graph_end": [_vp, C.POINTER(_vp)],
    "tg_graph_launch": [_vp, _vp],
    "tg_graph_destroy": [_vp],
    "tg_apply64": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_apply64_noalloc": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_mfma_probe": [_vp, _vp, _vp, _vp],
    "tg_diag_inv_hex": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp, _vp],
    "tg_diag_inv_stored": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_diag_inv_q8": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_restrict32": [C.POINTER(TgGrid), C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_prolong32": [C.POINTER(TgGrid), C.POINTER(TgGrid), _vp, _vp, _vp, C.c_int, _vp],
    "tg_galerkin_elems": [C.POINTER(TgGrid), C.POINTER(TgGrid), _vp, C.c_int, _vp, _vp, _vp, _vp, _vp],
    "tg_galerkin_from_fine": [C.POINTER(TgGrid), C.c_int, _vp, _vp, _vp, _vp, _vp],
    "tg_assemble_dense": [C.POINTER(TgGrid), _vp, _vp, _i32, _vp, _vp],
    "tg_assemble_stencil": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_assemble_stencil_bf16": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_dense_spd_inverse": [_vp, _i32, _vp],
    "tg_dense_spd_inverse_async": [_vp, _i32, _vp, _vp],
    "tg_coarsest_apply": [_vp, _i32, _vp, _i64, _vp, _vp, _vp],
    "tg_bcr_size": [C.POINTER(TgGrid), _vp],
    "tg_bcr_setup": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp],
    "tg_bcr_apply_b": [C.POINTER(TgGrid), _vp, _i32, _vp, _vp, _vp, _vp],
    "tg_dot32": [_i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_dot64": [_i64, _vp, _vp, _vp, _vp],
    "tg_pcg_xr": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_pcg_dir": [_i64, _vp, _vp, _vp, _vp, _vp],
    "tg_axpy_f32_to_f64": [_i64, _vp, _vp, _vp],
    "tg_axpy64": [_i64, _vp, _vp, _vp, _vp, _f64, _i32, _vp],
    "tg_normalize64": [_i64, _vp, _vp, _vp, _vp, _vp],
    "tg_passive64": [_i64, _vp, _vp, _f64, _vp],
    "tg_cheb_init32": [_i64, _vp, _vp, _f32, _vp, _vp, _vp, _vp],
    "tg_power_step32": [_i64, _vp, _vp, _vp, _vp, _vp],
    "tg_fill32": [_i64, _vp, _f32, _vp],
    "tg_fill64": [_i64, _vp, _f64, _vp],
    "tg_copy32": [_i64, _vp, _vp, _vp],
    "tg_compliance_dd": [C.POINTER(TgGrid), _vp, _vp, _vp, _vp, _vp],
    "tg_oc_mean_prepare": [_vp, _vp],
    "tg_oc_mean_step": [_i64, _vp, _vp, _f64, _f64, _f64, _vp, _vp],
    "tg_oc_mean_finalize": [_i64, _vp, _vp, _f64, _f64, _vp, _vp, _vp, _vp],
    "tg_apply32_b": [C.POINTER(TgOp32), _i32, _vp, _vp, C.POINTER(TgEp32), _vp],
    "tg_apply64_b": [C.POINTER(TgGrid), _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_restrict32_b": [C.POINTER(TgGrid), C.POINTER(TgGrid), _vp, _i32, _vp, _vp, _vp],
    "tg_prolong32_b": [C.POINTER(TgGrid), C.POINTER(TgGrid), _vp, _i32, _vp, _vp, C.c_int, _vp],
    "tg_coarsest_apply_b": [_vp, _i32, _vp, _i64, _i32, _vp, _vp, _vp],
    "tg_dot32_b": [_i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_dot64_b": [_i64, _i32, _vp, _vp, _vp, _vp],
    "tg_pcg_xr_b": [_i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_pcg_dir_b": [_i64, _i32, _vp, _vp, _vp, _vp, _vp],
    "tg_cheb_init32_b": [_i64, _i32, _vp, _vp, _f32, _vp, _vp, _vp, _vp],
    "tg_sensitivity64_b": [C.POINTER(TgGrid), _i32, _vp, _vp, _vp, _vp, _f64, _f64, _vp, _vp, _vp],
    "tg_stencil64": [C.POINTER(TgGrid), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "tg_density_chain": [_i64, _vp, _f64, _f64, _f64, _f64, _f64, _vp, _vp, _vp, _vp, _vp],
    "tg_sensitivity64": [C.POINTER(TgGrid), _vp, _vp, _vp, _f64, _f64, _vp, _vp, _vp],
    "tg_oc_prepare": [_i64, _vp, _vp, _vp, _f64, _vp, _vp, _vp, _vp, _vp],
    "tg_oc_bisect_step": [C.POINTER(TgGrid), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _f64, _f64, _f64, C.c_int,
                          _vp, _vp],
    "tg_oc_finalize": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
}

EXPORTED = tuple(_SIGS)

_lib = None


def load() -> C.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m topogpu_amd.csrc.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            if name == "tg_last_error":
                fn.restype = C.c_char_p
                fn.argtypes = []
            else:
                fn.restype = C.c_int
                fn.argtypes = args
        _lib = lib
    return _lib


class TopoHipError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != 0:
        msg = load().tg_last_error().decode(errors="replace")
        raise TopoHipError(f"libtopohip error {rc}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def grid(dims) -> TgGrid:
    return TgGrid(int(dims[0]), int(dims[1]), int(dims[2]))


def ptr(t) -> int | None:
    """Raw device pointer of a tensor (None for None)."""
    return None if t is None else t.data_ptr()


_initialised: set[int] = set()


def init_device(ke: np.ndarray, device: int) -> None:
    """Upload KE to the device's constant memory once per process."""
    if device in _initialised:
        return
    ke = np.ascontiguousarray(ke, dtype=np.float64)
    assert ke.shape == (24, 24)
    call("tg_init", ke.ctypes.data)
    _initialised.add(device)
