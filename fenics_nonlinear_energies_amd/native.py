"""ctypes binding of libfne_hip.so (C ABI declared in include/fne_hip.h).

The product path has no CPU fallback: if the library is missing or cannot be loaded,
``lib()`` raises.  Pointers passed in are raw device addresses (``tensor.data_ptr()``).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("FNE_HIP_LIB", PKG_DIR / "_lib" / "libfne_hip.so"))

FNE_OK = 0
FNE_EINVAL = -1
FNE_EHIP = -2
FNE_ENOMEM = -3
FNE_EUNSUPPORTED = -4
FNE_KIND_PLAPLACE = 1
FNE_KIND_GINZBURG_LANDAU = 2
FNE_KIND_NEOHOOKEAN = 3
FNE_KIND_MOHR_COULOMB = 4
FNE_FP64 = 0
FNE_FP32 = 1
FNE_FLAG_ABS_DET = 1
FNE_FLAG_ROW_GATHER = 2
FNE_FLAG_COLORED = 4
FNE_FLAG_STAGED = 8
FNE_FLAG_NODE_GEO = 16
FNE_FLAG_GL_PRODUCTS = 32
FNE_PCG_NOALLOC_MAX_BYTES = 16 << 20

_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


class FnePlanDesc(C.Structure):
    """Mirror of ``fne_plan_desc``."""

    _fields_ = [
        ("kind", C.c_int32),
        ("n_elem", C.c_int32),
        ("n_loc", C.c_int32),
        ("n_geo", C.c_int32),
        ("n_free", C.c_int32),
        ("n_total", C.c_int32),
        ("nnz", C.c_int64),
        ("edof", C.c_void_p),
        ("egeo", C.c_void_p),
        ("eweight", C.c_void_p),
        ("indptr", C.c_void_p),
        ("indices", C.c_void_p),
        ("n_patch", C.c_int32),
        ("chunk", C.c_int32),
        ("max_patch_nnz", C.c_int32),
        ("max_row_len", C.c_int32),
        ("prow", C.c_void_p),
        ("pelem_ptr", C.c_void_p),
        ("pelem", C.c_void_p),
        ("inc_ptr", C.c_void_p),
        ("inc", C.c_void_p),
        ("inc_aux", C.c_void_p),
        ("n_pe", C.c_int64),
        ("pedof", C.c_void_p),
        ("pgeo", C.c_void_p),
        ("pown", C.c_void_p),
        ("pitem", C.c_void_p),
        ("pcolors", C.c_void_p),
        ("param", C.c_double * 8),
        ("qdata", C.c_void_p),
        ("nq", C.c_int32),
        ("flags", C.c_int32),
        ("dof_xy", C.c_void_p),
    ]


# (name, restype, argtypes) -- every symbol declared in include/fne_hip.h
_SIGNATURES = [
    ("fne_plan_create", C.c_int, [C.POINTER(FnePlanDesc), C.POINTER(C.c_void_p)]),
    ("fne_plan_destroy", C.c_int, [C.c_void_p]),
    ("fne_plan_set_params", C.c_int, [C.c_void_p, _f64p]),
    ("fne_plan_set_mode", C.c_int, [C.c_void_p, C.c_int32]),
    ("fne_energy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_gradient", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    (
        "fne_hessian",
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
    ),
    (
        "fne_energy_grad_hessian",
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
    ),
    (
        "fne_element_eval",
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    ("fne_hvp", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_hessian_diag", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    (
        "fne_spmv",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    (
        "fne_csr_diag",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p],
    ),
    ("fne_golden_state_len", C.c_int, []),
    ("fne_golden_trial", C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_golden_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    (
        "fne_aggregate_qr",
        C.c_int,
        [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    (
        "fne_vcycle_op",
        C.c_int,
        [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    (
        "fne_spgemm_numeric",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    ("fne_dot", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_dot_work_len", C.c_int32, []),
    (
        "fne_color_patch_elements",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p],
    ),
    ("fne_pcg_workspace_bytes", C.c_int64, [C.c_int32, C.c_int64]),
    (
        "fne_pcg",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
         C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
         C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p],
    ),
    ("fne_color_distance2", C.c_int, [C.c_int32, _i32p, _i32p, _i32p]),
    (
        "fne_vcycle_cheb",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_double, C.c_void_p, C.c_void_p],
    ),
    (
        "fne_pcg_noalloc",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
         C.c_double, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
         C.POINTER(C.c_double), C.c_void_p],
    ),
    (
        "fne_pcg_graph",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_int64,
         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p],
    ),
    (
        "fne_pstcg",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
         C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
         C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p],
    ),
    (
        "fne_pstcg_graph",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_int64,
         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p],
    ),
    ("fne_gmres_workspace_bytes", C.c_int64, [C.c_int32, C.c_int32, C.c_int32]),
    (
        "fne_gmres",
        C.c_int,
        [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
         C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p],
    ),
    ("fne_mlp_forward", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_mlp_workspace_bytes", C.c_int64, [C.c_void_p, C.c_int64]),
    (
        "fne_mlp_backward",
        C.c_int,
        [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
         C.c_void_p],
    ),
    (
        "fne_mlp_adam",
        C.c_int,
        [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float,
         C.c_float, C.c_float, C.c_int32, C.c_void_p],
    ),
    ("fne_mlp_to_bf16", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("fne_last_error", C.c_char_p, []),
    ("fne_version", C.c_char_p, []),
]

_LIB: C.CDLL | None = None


class FneError(RuntimeError):
    pass


def exported_symbols() -> list[str]:
    return [name for name, _, _ in _SIGNATURES]


def lib() -> C.CDLL:
    """Load the HIP library (once).  Raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise FneError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback)"
        )
    handle = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, restype, argtypes in _SIGNATURES:
        fn = getattr(handle, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _LIB = handle
    return handle


def check(rc: int, what: str) -> None:
    if rc != FNE_OK:
        msg = lib().fne_last_error()
        raise FneError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


def ptr(t) -> int | None:
    """Raw device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return int(t.data_ptr())


def stream_ptr(stream=None) -> int | None:
    """Raw HIP stream of ``stream`` or of torch's current stream on the current device (None
    for the null stream).  The current stream is read through torch's C accessors: the public
    ``torch.cuda.current_stream()`` costs about 8 us per call in Python, which adds up over the
    tens of thousands of launches of a Newton load path."""
    import torch

    if stream is not None:
        return int(stream.cuda_stream) or None
    try:
        raw = torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())
    except AttributeError:
        raw = torch.cuda.current_stream().cuda_stream
    return int(raw) or None
