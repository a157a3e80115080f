"""Device vectors with the PETSc.Vec method surface used by the Newton loop.

The reference's ``newton`` (core/petsc/minimizers.py:181-1038) drives PETSc vectors via
``duplicate, copy, scale, norm, dot, axpy, waxpy, set`` and ``array``.  ``DeviceVec`` keeps
an FP64 torch tensor on the GPU and implements the same calls; reductions (``dot``/``norm``)
go through the library's deterministic compensated FP64 kernel.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import native


class NormType:
    NORM_2 = 2


class DeviceVec:
    __slots__ = ("t", "_scalar")

    def __init__(self, t: torch.Tensor, scalar: torch.Tensor | None = None):
        if t.dtype != torch.float64:
            raise TypeError("DeviceVec holds float64 data")
        self.t = t
        self._scalar = scalar

    # -- construction -------------------------------------------------------------------
    @staticmethod
    def zeros(n: int, device="cuda") -> "DeviceVec":
        return DeviceVec(torch.zeros(int(n), dtype=torch.float64, device=device))

    @staticmethod
    def from_numpy(a, device="cuda") -> "DeviceVec":
        return DeviceVec(torch.as_tensor(np.asarray(a, dtype=np.float64)).to(device))

    def duplicate(self) -> "DeviceVec":
        return DeviceVec(torch.empty_like(self.t))

    def destroy(self) -> None:
        self.t = None

    # -- data access --------------------------------------------------------------------
    @property
    def array(self) -> np.ndarray:
        """Host copy of the values (the reference reads ``vec.array[:]`` by copy)."""
        return self.t.detach().cpu().numpy()

    @array.setter
    def array(self, values) -> None:
        self.t.copy_(torch.as_tensor(np.asarray(values, dtype=np.float64)))

    def getSize(self) -> int:
        return int(self.t.numel())

    def assemble(self) -> None:
        return None

    # -- BLAS-1 ---------------------------------------------------------------------------
    def copy(self, other: "DeviceVec | None" = None) -> "DeviceVec":
        if other is None:
            return DeviceVec(self.t.clone())
        other.t.copy_(self.t)
        return other

    def set(self, value: float) -> None:
        self.t.fill_(float(value))

    def scale(self, alpha: float) -> None:
        self.t.mul_(float(alpha))

    def axpy(self, alpha: float, x: "DeviceVec") -> None:
        self.t.add_(x.t, alpha=float(alpha))

    def aypx(self, alpha: float, x: "DeviceVec") -> None:
        self.t.mul_(float(alpha)).add_(x.t)

    def waxpy(self, alpha: float, x: "DeviceVec", y: "DeviceVec") -> None:
        """self = alpha*x + y"""
        torch.add(y.t, x.t, alpha=float(alpha), out=self.t)

    def pointwiseMult(self, x: "DeviceVec", y: "DeviceVec") -> None:
        torch.mul(x.t, y.t, out=self.t)

    def dot(self, other: "DeviceVec") -> float:
        return float(dot_device(self.t, other.t).item())

    def norm(self, norm_type=NormType.NORM_2) -> float:
        if norm_type != NormType.NORM_2:
            raise ValueError("only NORM_2 is supported")
        return math.sqrt(max(self.dot(self), 0.0))


_SCRATCH: dict[int, tuple[torch.Tensor, torch.Tensor]] = {}


def _dot_scratch(device: torch.device) -> tuple[torch.Tensor, torch.Tensor]:
    """Per-device (result, partials) buffers of ``fne_dot``, allocated once by torch."""
    dev = device.index if device.index is not None else torch.cuda.current_device()
    buf = _SCRATCH.get(dev)
    if buf is None:
        n_pairs = int(native.lib().fne_dot_work_len())
        buf = (torch.empty(8, dtype=torch.float64, device=device),
               torch.empty(2 * n_pairs, dtype=torch.float64, device=device))
        _SCRATCH[dev] = buf
    return buf


def dot_device(a: torch.Tensor, b: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Deterministic compensated FP64 dot product, result left on the device.

    Uses one partials buffer per device: calls are ordered on the current stream."""
    res, work = _dot_scratch(a.device)
    if out is None:
        out = res[:1]
    native.check(
        native.lib().fne_dot(int(a.numel()), native.ptr(a), native.ptr(b), native.ptr(out), native.ptr(work),
                             native.stream_ptr()),
        "fne_dot",
    )
    return out
