"""Learned warm start for load and continuation steps (north-star learned part).

A small per-node MLP predicts the correction from the extrapolated guess
u_pred = u_k + (u_k - u_{k-1}) to the converged next state u_{k+1}.  Inputs per node:
reference coordinates, the last two displacements and the next load parameter (scaled);
output: the displacement correction.  The fallback -- and the reference's behaviour
(``hyperelasticity/jax_petsc/solver.py:351-359``: the next step starts from the previous
converged state) -- is used whenever the network is not trained or history is short.

Training and inference run on the device through the HIP library (``fne_mlp_*``: BF16 MFMA
products with FP32 accumulation, FP32 master weights, Adam); torch only holds the buffers.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import native

_MAX_LAYERS = 8


class _MlpDesc(C.Structure):
    _fields_ = [
        ("n_layers", C.c_int32),
        ("dims", C.c_int32 * (_MAX_LAYERS + 1)),
        ("w", C.c_void_p * _MAX_LAYERS),
        ("b", C.c_void_p * _MAX_LAYERS),
    ]


def _lib():
    return native.lib()


class NodeMLP:
    """Fully connected tanh network on the device: dims = [d_in, h, ..., h, d_out]."""

    def __init__(self, dims, device="cuda", seed=0):
        if len(dims) - 1 > _MAX_LAYERS:
            raise ValueError(f"at most {_MAX_LAYERS} layers")
        self.dims = [int(d) for d in dims]
        self.device = torch.device(device)
        rng = np.random.default_rng(seed)
        self.w32, self.b32 = [], []
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            lim = math.sqrt(6.0 / (din + dout))  # Glorot uniform
            self.w32.append(torch.as_tensor(rng.uniform(-lim, lim, (dout, din)).astype(np.float32)).to(self.device))
            self.b32.append(torch.zeros(dout, dtype=torch.float32, device=self.device))
        self.wbf = [torch.empty(w.shape, dtype=torch.int16, device=self.device) for w in self.w32]
        self._sync_bf16()
        self.m = [torch.zeros_like(t) for t in self.w32 + self.b32]
        self.v = [torch.zeros_like(t) for t in self.w32 + self.b32]
        self.step = 0
        self._x = None
        self._acts = None
        self._ws = None

    # ----------------------------------------------------------------------------------
    def _sync_bf16(self) -> None:
        lib = _lib()
        for w, wb in zip(self.w32, self.wbf):
            native.check(lib.fne_mlp_to_bf16(w.numel(), native.ptr(w), native.ptr(wb), native.stream_ptr()),
                         "fne_mlp_to_bf16")

    def _desc(self) -> _MlpDesc:
        d = _MlpDesc()
        d.n_layers = len(self.w32)
        for i, v in enumerate(self.dims):
            d.dims[i] = v
        for l, (wb, b) in enumerate(zip(self.wbf, self.b32)):
            d.w[l] = native.ptr(wb)
            d.b[l] = native.ptr(b)
        return d

    def weights_bf16_values(self):
        """BF16 weights as FP32 tensors (for checks)."""
        return [(wb.to(torch.int32) << 16).view(torch.float32) for wb in self.wbf]

    def forward(self, x: torch.Tensor, keep: bool = False) -> torch.Tensor:
        x = x.to(device=self.device, dtype=torch.float32).contiguous()
        n = x.shape[0]
        y = torch.empty((n, self.dims[-1]), dtype=torch.float32, device=self.device)
        nh = sum(self.dims[1:-1])
        acts = torch.empty(max(1, n * nh), dtype=torch.float32, device=self.device)
        d = self._desc()
        native.check(
            _lib().fne_mlp_forward(C.byref(d), n, native.ptr(x), native.ptr(y), native.ptr(acts), native.stream_ptr()),
            "fne_mlp_forward",
        )
        if keep:
            self._x, self._acts = x, acts
        return y

    def backward(self, dy: torch.Tensor):
        """Gradients (dw list, db list) of sum(dy * y) for the last ``forward(keep=True)``."""
        if self._x is None:
            raise RuntimeError("backward needs forward(..., keep=True) first")
        dy = dy.to(device=self.device, dtype=torch.float32).contiguous()
        n = self._x.shape[0]
        lib = _lib()
        d = self._desc()
        nbytes = int(lib.fne_mlp_workspace_bytes(C.byref(d), n))
        if nbytes < 0:
            native.check(-1, "fne_mlp_workspace_bytes")
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        dw = [torch.empty_like(w) for w in self.w32]
        db = [torch.empty_like(b) for b in self.b32]
        DW = (C.c_void_p * len(dw))(*[native.ptr(t) for t in dw])
        DB = (C.c_void_p * len(db))(*[native.ptr(t) for t in db])
        native.check(
            lib.fne_mlp_backward(C.byref(d), n, native.ptr(self._x), native.ptr(self._acts), native.ptr(dy),
                                 C.cast(DW, C.c_void_p), C.cast(DB, C.c_void_p), native.ptr(self._ws), nbytes,
                                 native.stream_ptr()),
            "fne_mlp_backward",
        )
        return dw, db

    def adam_step(self, dw, db, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0) -> None:
        self.step += 1
        lib = _lib()
        params = self.w32 + self.b32
        grads = list(dw) + list(db)
        bf = self.wbf + [None] * len(self.b32)
        for p, g, m, v, wb in zip(params, grads, self.m, self.v, bf):
            native.check(
                lib.fne_mlp_adam(p.numel(), native.ptr(p), native.ptr(wb), native.ptr(g), native.ptr(m), native.ptr(v),
                                 lr, beta1, beta2, eps, weight_decay, self.step, native.stream_ptr()),
                "fne_mlp_adam",
            )

    def train_mse(self, x: torch.Tensor, target: torch.Tensor, *, epochs=200, batch=65536, lr=2e-3, seed=0,
                  lr_final=None) -> list[float]:
        """Mini-batch Adam on the mean squared error; returns per-epoch mean losses."""
        x = x.to(self.device, torch.float32)
        target = target.to(self.device, torch.float32)
        n = x.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            lr_e = lr if lr_final is None else lr * (lr_final / lr) ** (ep / max(1, epochs - 1))
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = 0.0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                xb, tb = x[idx], target[idx]
                y = self.forward(xb, keep=True)
                r = y - tb
                dy = (2.0 / r.numel()) * r
                dw, db = self.backward(dy)
                self.adam_step(dw, db, lr=lr_e)
                tot += float((r * r).sum().item())
            losses.append(tot / (n * self.dims[-1]))
        return losses


class NodalPredictor:
    """Initial guess for the next load / continuation step from the converged history.

    ``mode``: "previous" (the reference's choice: last converged state), "extrapolate"
    (u_k + (u_k - u_{k-1}) for equally spaced steps), "learned" (extrapolation plus the MLP
    correction).  Works on the free DOFs of the assembler's reordered numbering; per-node
    features use the nodes whose DOFs are all free (``free_node_dofs``).

    ``state_offset`` (free values) is subtracted to form displacements (the HE unknown is the
    deformed position, so the offset is the reference position; zero for slope displacements).
    """

    def __init__(self, free_node_dofs: np.ndarray, coords: np.ndarray, *, state_offset, disp_scale: float,
                 load_scale: float, net: NodeMLP | None = None, mode: str = "learned", first_step: bool = False):
        self.idx = torch.as_tensor(np.asarray(free_node_dofs, dtype=np.int64))  # (n_nodes_free, dim)
        self.coords = np.asarray(coords, dtype=np.float64)
        self.dim = self.idx.shape[1]
        self.offset = state_offset
        self.disp_scale = float(disp_scale)
        self.load_scale = float(load_scale)
        self.net = net
        self.mode = str(mode)
        # first_step: the network also predicts the first step of a path, from the start state
        # alone (features with u_{k-1} = u_k, i.e. no velocity); trained on the start sample of
        # every history (training_pairs)
        self.first_step = bool(first_step)
        self._dev = None
        self.feat_mean = self.feat_std = self.targ_std = None
        self.energy_fn = None  # optional: accept the learned guess only if its energy is lower
        self.accepted = []

    def _to(self, t):
        if self._dev is None:
            self._dev = self.idx.to(t.device)
            self._X = torch.as_tensor(self.coords, dtype=torch.float32, device=t.device)
        return self._dev

    def raw_features(self, u_k: torch.Tensor, u_km1: torch.Tensor, load_next: float) -> torch.Tensor:
        idx = self._to(u_k)
        dk = (u_k - self.offset)[idx].float() / self.disp_scale
        dkm1 = (u_km1 - self.offset)[idx].float() / self.disp_scale
        lp = torch.full((idx.shape[0], 1), float(load_next) / self.load_scale, device=u_k.device)
        return torch.cat((self._X / self.coord_scale, dk, dkm1, lp), dim=1).contiguous()

    def features(self, u_k, u_km1, load_next) -> torch.Tensor:
        f = self.raw_features(u_k, u_km1, load_next)
        if self.feat_mean is not None:
            f = (f - self.feat_mean) / self.feat_std
        return f.contiguous()

    def fit_normalization(self, raw_x: torch.Tensor, raw_t: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """Standardise inputs (per feature) and targets (per component) from training data."""
        self.feat_mean = raw_x.mean(dim=0)
        self.feat_std = raw_x.std(dim=0).clamp_min(1e-12)
        self.targ_std = raw_t.std(dim=0).clamp_min(1e-12)
        return (raw_x - self.feat_mean) / self.feat_std, raw_t / self.targ_std

    @property
    def coord_scale(self) -> float:
        return float(np.abs(self.coords).max()) or 1.0

    @property
    def n_in(self) -> int:
        return 3 * self.dim + 1

    def target(self, u_k, u_km1, u_next) -> torch.Tensor:
        """Training target: (u_{k+1} - extrapolation) per free node, scaled."""
        idx = self._to(u_k)
        pred = 2.0 * u_k - u_km1
        return ((u_next - pred)[idx].float() / self.disp_scale).contiguous()

    def __call__(self, history, load_next):
        if self.mode == "previous" or not history:
            return None  # caller keeps the previous converged state
        if len(history) < 2:
            if not (self.first_step and self.mode == "learned" and self.net is not None):
                return None
            u_k = u_km1 = history[-1][1]  # the start state: the guess is the previous state
        else:
            u_k, u_km1 = history[-1][1], history[-2][1]
        guess = 2.0 * u_k - u_km1
        if self.mode == "extrapolate" or self.net is None:
            return guess
        idx = self._to(u_k)
        corr = self.net.forward(self.features(u_k, u_km1, load_next))
        if self.targ_std is not None:
            corr = corr * self.targ_std
        learned = guess.clone()
        learned[idx] += self.disp_scale * corr.double()
        if self.energy_fn is not None:
            # Newton's own merit function decides: keep the extrapolation if the network's
            # guess has the higher energy
            ok = self.energy_fn(learned) <= self.energy_fn(guess)
            self.accepted.append(bool(ok))
            return learned if ok else guess
        return learned


def training_pairs(predictor: NodalPredictor, history) -> tuple[torch.Tensor, torch.Tensor]:
    """(raw features, raw targets) of every (k-1, k, k+1) triple of consecutive, equally spaced
    loads in a converged history (and, for a predictor with ``first_step``, the start sample
    (0, 0, 1) with u_{-1} = u_0).  The history holds converged steps only, so a step that did
    not converge leaves a gap; a triple across it would pair a two-step increment with the
    one-step extrapolation the features describe, and is skipped."""
    xs, ts = [], []
    if predictor.first_step and len(history) >= 2:
        # start sample: (u_0, u_0) -> u_1, the first step from the start state alone
        (_, u_0), (load_1, u_1) = history[0], history[1]
        xs.append(predictor.raw_features(u_0, u_0, load_1))
        ts.append(predictor.target(u_0, u_0, u_1))
    for j in range(1, len(history) - 1):
        (load_km1, u_km1), (load_k, u_k), (load_next, u_next) = history[j - 1], history[j], history[j + 1]
        h0, h1 = float(load_k) - float(load_km1), float(load_next) - float(load_k)
        if abs(h1 - h0) > 1e-9 * max(abs(h0), abs(h1), 1e-300):
            continue
        xs.append(predictor.raw_features(u_k, u_km1, load_next))
        ts.append(predictor.target(u_k, u_km1, u_next))
    if not xs:
        raise ValueError("the history holds no triple of equally spaced loads")
    return torch.cat(xs), torch.cat(ts)
