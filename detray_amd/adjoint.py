"""Differentiable chi2 of hit residuals through the propagation.

The reference transports covariances with Jacobians (parameter_updater.hpp:
422-572, rk_stepper.hpp:633-956); it has no reverse mode.  Here the
gradient of

    chi2 = sum over matched hits of (loc0 - m0)^2 / s0^2 + (loc1 - m1)^2 / s1^2

with respect to the bound start parameters of each track comes from a
hand-written adjoint kernel:

* ``dtr_adjoint_forward`` propagates every track from its bound start (the
  same stepper, navigator and actor chain as :class:`~.Propagator`) and
  writes a tape: the step sizes, the free state every ``checkpoint_every``
  steps, and every sensitive / material surface event.  It also returns the
  gradient that the chain of the reference's transport Jacobians gives.
* ``dtr_adjoint_backward`` replays each segment from its checkpoint and runs
  the adjoint back through the Runge-Kutta steps and surface transitions.

:func:`chi2` wraps the two passes in a ``torch.autograd.Function``.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace
from typing import Optional

import numpy as np

from . import capi
from .propagator import Detector, MagneticField, PropagationConfig

# toy detector: the beam pipe is surface 0, a concentric cylinder of radius
# 19 mm (build_toy_detector.hpp:1094-1142)
BEAMPIPE_SURFACE = 0
BEAMPIPE_RADIUS = 19.0

# hit resolutions (mm) and start-parameter uncertainties of the
# differentiation workload
SIGMA_LOC0 = 0.010
SIGMA_LOC1 = 0.050
SIGMA_START_LOC = 0.010
SIGMA_START_ANGLE = 1e-3
SIGMA_START_QOP_REL = 0.01


@dataclass
class AdjointConfig:
    """``dtr_adjoint_config``: tape capacity, checkpoint interval K,
    precision ("fp64" / "mixed"), model ("exact": derivative of the
    propagation; "transport": the derivative the reference's Jacobians
    describe), launch (block size, waves per SIMD the kernels are built
    for) and the track order of both passes ("eta_phi": grouped by
    direction, "none": as given)."""

    max_steps: int = 512
    max_events: int = 64
    checkpoint_every: int = 32
    precision: str = "mixed"
    model: str = "exact"
    block_size: int = 64
    waves_per_simd: int = 2
    sort: str = "eta_phi"
    sigma_loc0: float = SIGMA_LOC0
    sigma_loc1: float = SIGMA_LOC1

    def to_c(self, model: Optional[str] = None) -> capi.AdjointConfig:
        a = capi.AdjointConfig()
        a.max_steps = int(self.max_steps)
        a.max_events = int(self.max_events)
        a.checkpoint_every = int(self.checkpoint_every)
        a.precision = {"fp64": capi.PRECISION_FP64, "mixed": capi.PRECISION_MIXED}[self.precision]
        a.model = {"exact": capi.ADJOINT_EXACT, "transport": capi.ADJOINT_TRANSPORT}[model or self.model]
        a.block_size = int(self.block_size)
        a.waves_per_simd = int(self.waves_per_simd)
        a.sigma_loc0 = float(self.sigma_loc0)
        a.sigma_loc1 = float(self.sigma_loc1)
        return a


@dataclass
class Hits:
    """Hits of a batch (device tensors): surface [max_hits, n] int32 in
    crossing order, loc [2, max_hits, n] float64, n_hits [n] int32."""

    surface: "object"
    loc: "object"
    n_hits: "object"

    @property
    def max_hits(self) -> int:
        return int(self.surface.shape[0])

    @property
    def n(self) -> int:
        return int(self.surface.shape[1])

    @classmethod
    def empty(cls, max_hits: int, n: int, device="cuda") -> "Hits":
        import torch

        return cls(torch.full((max_hits, n), -1, dtype=torch.int32, device=device),
                   torch.zeros((2, max_hits, n), dtype=torch.float64, device=device),
                   torch.zeros((n,), dtype=torch.int32, device=device))

    def check(self, n: int) -> None:
        import torch

        if (self.surface.dtype != torch.int32 or self.surface.dim() != 2 or self.surface.shape[1] != n
                or self.loc.dtype != torch.float64 or tuple(self.loc.shape) != (2, self.max_hits, n)
                or self.n_hits.dtype != torch.int32 or tuple(self.n_hits.shape) != (n,)
                or not (self.surface.is_cuda and self.loc.is_cuda and self.n_hits.is_cuda)):
            raise ValueError("hits must be CUDA tensors surface int32 [m, n], loc float64 [2, m, n], "
                             "n_hits int32 [n]")
        if not (self.surface.is_contiguous() and self.loc.is_contiguous() and self.n_hits.is_contiguous()):
            raise ValueError("hits tensors must be contiguous")

    def to_c(self) -> capi.Hits:
        return capi.Hits(self.max_hits, 0, self.surface.data_ptr(), self.loc.data_ptr(), self.n_hits.data_ptr())

    def smeared(self, sigma_loc0: float = SIGMA_LOC0, sigma_loc1: float = SIGMA_LOC1,
                generator=None) -> "Hits":
        """Measurements: the true intersections plus Gaussian noise."""
        import torch

        noise = torch.randn(self.loc.shape, dtype=torch.float64, device=self.loc.device, generator=generator)
        scale = torch.tensor([sigma_loc0, sigma_loc1], dtype=torch.float64, device=self.loc.device)
        return Hits(self.surface.clone(), self.loc + noise * scale[:, None, None], self.n_hits.clone())


@dataclass
class ForwardResult:
    chi2: "object"             # [n] float64
    grad_transport: "object"   # [6, n] float64 or None
    status: "object"           # [n] int32 (dtr_tape_status)
    tape: "object"             # uint8 tensor
    order: "object" = None     # [n] int32 track order of the pass, or None
    # the adjoint configuration the tape was written with (the backward pass
    # reads the tape with the same layout and precision)
    adjoint: "object" = None


def direction_order(bound_params, eta_bins_per_unit: float = 20.0):
    """Permutation (int32, device) grouping bound tracks by direction: by
    pseudorapidity bins of width 1/eta_bins_per_unit, then by phi."""
    import torch

    theta = bound_params[3].clamp(1e-9, math.pi - 1e-9)
    eta = -torch.log(torch.tan(0.5 * theta))
    eta_bin = torch.round(eta * eta_bins_per_unit).to(torch.int64) + 4096
    phi_q = ((bound_params[2] + math.pi) * (2 ** 20 / (2 * math.pi))).to(torch.int64).clamp(0, 2 ** 20 - 1)
    return torch.argsort((eta_bin << 21) | phi_q).to(torch.int32)


def _stream(device):
    import torch

    return torch.cuda.current_stream(device).cuda_stream


class Chi2Propagator:
    """chi2 of hit residuals and its adjoint over batches of bound tracks."""

    def __init__(self, config: Optional[PropagationConfig] = None, adjoint: Optional[AdjointConfig] = None):
        self.config = config or PropagationConfig()
        self.adjoint = adjoint or AdjointConfig()
        self._c_config = self.config.to_c()

    def tape_bytes(self, n: int) -> int:
        return self.tape_bytes_for(self.adjoint, n)

    @staticmethod
    def tape_bytes_for(adjoint: "AdjointConfig", n: int) -> int:
        return int(capi.load().dtr_adjoint_tape_size(C.byref(adjoint.to_c()), int(n)))

    def _inputs(self, detector: Detector, bound_params, start_surface):
        import torch

        if (not isinstance(bound_params, torch.Tensor) or not bound_params.is_cuda
                or bound_params.dtype != torch.float64 or bound_params.dim() != 2 or bound_params.shape[0] != 6):
            raise ValueError("bound_params must be a float64 CUDA tensor [6, n]")
        n = int(bound_params.shape[1])
        if (not isinstance(start_surface, torch.Tensor) or not start_surface.is_cuda
                or start_surface.dtype != torch.int32 or tuple(start_surface.shape) != (n,)):
            raise ValueError("start_surface must be an int32 CUDA tensor [n]")
        if n and (int(start_surface.min().item()) < 0 or int(start_surface.max().item()) >= detector.n_surfaces):
            raise ValueError("start_surface out of range")
        return bound_params.contiguous(), start_surface.contiguous(), n

    def _batch(self, bound, start, n, tape, order) -> capi.AdjointBatch:
        return capi.AdjointBatch(n, 0, bound.data_ptr(), start.data_ptr(), tape.data_ptr(), int(tape.numel()),
                                 order.data_ptr() if order is not None else None)

    def forward(self, detector: Detector, field: MagneticField, bound_params, start_surface,
                hits: Optional[Hits] = None, transport_gradient: bool = False,
                record: Optional[Hits] = None) -> ForwardResult:
        """Recording forward pass: chi2 against ``hits`` (0 without hits),
        the tape, and optionally the Jacobian-chain gradient.  ``record``
        receives the true loc0/loc1 of every sensitive surface crossed."""
        import torch

        bound, start, n = self._inputs(detector, bound_params, start_surface)
        dev = bound.device
        if hits is not None:
            hits.check(n)
        if record is not None:
            record.check(n)
        tape = torch.empty(max(self.tape_bytes(n), 8), dtype=torch.uint8, device=dev)
        chi2 = torch.zeros(n, dtype=torch.float64, device=dev)
        status = torch.zeros(n, dtype=torch.int32, device=dev)
        gt = torch.zeros((6, n), dtype=torch.float64, device=dev) if transport_gradient else None
        order = direction_order(bound) if (self.adjoint.sort == "eta_phi" and n) else None
        lib = capi.load()
        hc = hits.to_c() if hits is not None else None
        rc = record.to_c() if record is not None else None
        capi.check(lib.dtr_adjoint_forward(detector.device(), field.device(), C.byref(self._c_config),
                                           C.byref(self.adjoint.to_c()),
                                           C.byref(self._batch(bound, start, n, tape, order)),
                                           C.byref(hc) if hc is not None else None,
                                           C.byref(rc) if rc is not None else None,
                                           C.c_void_p(chi2.data_ptr()),
                                           C.c_void_p(gt.data_ptr()) if gt is not None else None,
                                           C.c_void_p(status.data_ptr()), C.c_void_p(_stream(dev))),
                   "dtr_adjoint_forward")
        return ForwardResult(chi2, gt, status, tape, order, replace(self.adjoint))

    def backward(self, detector: Detector, field: MagneticField, bound_params, start_surface,
                 hits: Optional[Hits], fwd: ForwardResult, model: Optional[str] = None):
        """d chi2 / d(bound start parameters) [6, n] from the tape of
        ``fwd`` (NaN for tracks whose tape status is not OK)."""
        import torch

        bound, start, n = self._inputs(detector, bound_params, start_surface)
        if hits is not None:
            hits.check(n)
        grad = torch.empty((6, n), dtype=torch.float64, device=bound.device)
        hc = hits.to_c() if hits is not None else None
        # the tape layout and precision are those of the forward pass, even if
        # self.adjoint changed since
        acfg = fwd.adjoint if fwd.adjoint is not None else self.adjoint
        if fwd.tape is None or int(fwd.tape.numel()) < self.tape_bytes_for(acfg, n):
            raise RuntimeError("the forward result holds no tape for this batch")
        capi.check(capi.load().dtr_adjoint_backward(
            detector.device(), field.device(), C.byref(self._c_config),
            C.byref(acfg.to_c(model or self.adjoint.model)),
            C.byref(self._batch(bound, start, n, fwd.tape, fwd.order)), C.byref(hc) if hc is not None else None,
            C.c_void_p(grad.data_ptr()), C.c_void_p(_stream(bound.device))), "dtr_adjoint_backward")
        return grad

    def tape_counters(self, fwd: ForwardResult, n: int):
        """(steps, surface events) per track [n] int32 from the tape of a
        forward pass (the counters that close the tape layout of
        dtr_adjoint_tape_size)."""
        a = self.adjoint
        doubles = (2 * a.max_steps + 7 * (a.max_steps // a.checkpoint_every + 1) + 23 * a.max_events
                   + 7 * a.checkpoint_every + 36)
        import torch

        words = fwd.tape[8 * n * doubles:8 * n * doubles + 4 * n * (5 * a.max_events + 3)].view(torch.int32)
        base = 5 * a.max_events * n
        return words[base:base + n], words[base + n:base + 2 * n]

    def record_hits(self, detector: Detector, field: MagneticField, bound_params, start_surface,
                    max_hits: int = 32) -> Hits:
        """True intersections: loc0/loc1 of every sensitive surface the
        tracks cross."""
        n = int(bound_params.shape[1])
        rec = Hits.empty(max_hits, n, bound_params.device)
        self.forward(detector, field, bound_params, start_surface, None, False, rec)
        return rec

    def chi2(self, detector: Detector, field: MagneticField, bound_params, start_surface, hits: Hits):
        """chi2 per track [n], differentiable with respect to
        ``bound_params`` through the adjoint kernel."""
        return _Chi2Function.apply(bound_params, start_surface, hits.surface, hits.loc, hits.n_hits,
                                   self, detector, field)


def _autograd_function():
    import torch

    class Chi2Function(torch.autograd.Function):
        @staticmethod
        def forward(ctx, bound_params, start_surface, surface, loc, n_hits, prop, detector, field):
            hits = Hits(surface, loc, n_hits)
            fwd = prop.forward(detector, field, bound_params.detach(), start_surface, hits)
            ctx.prop, ctx.detector, ctx.field, ctx.fwd = prop, detector, field, fwd
            ctx.save_for_backward(bound_params, start_surface, surface, loc, n_hits)
            return fwd.chi2

        @staticmethod
        def backward(ctx, grad_chi2):
            bound_params, start_surface, surface, loc, n_hits = ctx.saved_tensors
            # the tape stays with the graph (ctx), so a second backward
            # (retain_graph=True) reads it again
            g = ctx.prop.backward(ctx.detector, ctx.field, bound_params.detach(), start_surface,
                                  Hits(surface, loc, n_hits), ctx.fwd)
            return grad_chi2[None, :] * g, None, None, None, None, None, None, None

    return Chi2Function


class _LazyFunction:
    _fn = None

    def apply(self, *args):
        if _LazyFunction._fn is None:
            _LazyFunction._fn = _autograd_function()
        return _LazyFunction._fn.apply(*args)


_Chi2Function = _LazyFunction()


def chi2(detector: Detector, field: MagneticField, bound_params, start_surface, hits: Hits,
         config: Optional[PropagationConfig] = None, adjoint: Optional[AdjointConfig] = None):
    """chi2 per track as a differentiable torch function of the bound start
    parameters (``torch.autograd.Function`` over the adjoint kernel)."""
    return Chi2Propagator(config, adjoint).chi2(detector, field, bound_params, start_surface, hits)


# ------------------------------------------------------------- workload --
def beampipe_starts(n: int, seed: int = 46, pt_min: float = 0.5, pt_max: float = 50.0, eta_max: float = 2.5,
                    sigma_z: float = 50.0, z_max: float = 200.0):
    """Bound start parameters [6, n] on the toy detector's beam pipe
    (loc0 = phi of the position, loc1 = z): position phi uniform, z
    Gaussian, momentum phi within 10 mrad of the position phi (outgoing),
    eta uniform in [-eta_max, eta_max], pT log-uniform, charge +-1.
    Returns (bound [6, n] float64, start surface [n] int32), numpy."""
    rng = np.random.default_rng(seed)
    b = np.zeros((6, n))
    b[0] = rng.uniform(-math.pi, math.pi, n)
    b[1] = np.clip(rng.normal(0.0, sigma_z, n), -z_max, z_max)
    b[2] = b[0] + rng.normal(0.0, 0.01, n)
    b[2] = (b[2] + math.pi) % (2 * math.pi) - math.pi
    eta = rng.uniform(-eta_max, eta_max, n)
    b[3] = 2.0 * np.arctan(np.exp(-eta))
    pt = np.exp(rng.uniform(math.log(pt_min), math.log(pt_max), n))
    q = rng.choice(np.array([-1.0, 1.0]), n)
    b[4] = q * np.sin(b[3]) / pt
    return b, np.full(n, BEAMPIPE_SURFACE, dtype=np.int32)


def start_sigma(bound: np.ndarray) -> np.ndarray:
    """1-sigma uncertainties [6, n] of beam-pipe start parameters: 10 um in
    position (loc0 is an angle on the cylinder), 1 mrad in angles, 1 % in
    q/p, none in time."""
    n = bound.shape[1]
    return np.stack([np.full(n, SIGMA_START_LOC / BEAMPIPE_RADIUS), np.full(n, SIGMA_START_LOC),
                     np.full(n, SIGMA_START_ANGLE), np.full(n, SIGMA_START_ANGLE),
                     SIGMA_START_QOP_REL * np.abs(bound[4]), np.zeros(n)])


def perturbed(bound: np.ndarray, seed: int = 47) -> np.ndarray:
    """Truth parameters: the start parameters moved by a Gaussian 1-sigma
    perturbation."""
    rng = np.random.default_rng(seed)
    return bound + start_sigma(bound) * rng.normal(size=bound.shape)
