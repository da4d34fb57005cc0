"""Beamlet-weight dose and its adjoint gradient, exposed to PyTorch.

A plan's fields are divided into beamlets (square pencils of the aperture at
the isocentre plane).  Every source record belongs to one beamlet; the dose of
the plan is D = sum_b w_b D_b, where D_b is the dose of beamlet b's histories
at unit weight.  ``BeamletDose`` is a ``torch.autograd.Function``: its forward
runs the batched event-queued transport with the weights applied to the source
records, its backward replays the same histories from their counter-based
Philox streams (gegs_adjoint) and gathers dL/dD at every deposit.  Weights,
dose, dL/dD and the gradient stay in device memory (the library's
``_device`` entry points).  A queue of plans is one BeamletPlan
(``BeamletPlan.from_jobs``): one dose buffer and one objective term per
plan, all plans in one transport and one replay.  The gradient with respect
to the source records' own weights is the same replay with one beamlet per
history (``phase_space_weight_gradient``).  The reference (RicciWoo/GEGS)
has no differentiable path; this is the addition the dose-optimisation use of
the transport needs.
"""
from __future__ import annotations

import numpy as np


def beamlet_index(src, field, n_side, pitch):
    """Beamlet of each source record: the record's straight line is carried to
    the isocentre plane (z = source_to_isocenter in the head frame) and
    binned on an n_side x n_side grid of pitch cm centred on the axis.
    Records that miss the grid get index -1."""
    sid = float(field.source_to_isocenter)
    z = src["z"].astype(np.float64)
    w = src["w"].astype(np.float64)
    t = np.where(w > 0, (sid - z) / np.where(w > 0, w, 1.0), np.nan)
    x = src["x"] + t * src["u"]
    y = src["y"] + t * src["v"]
    half = 0.5 * n_side * pitch
    ix = np.floor((x + half) / pitch)
    iy = np.floor((y + half) / pitch)
    ok = (ix >= 0) & (ix < n_side) & (iy >= 0) & (iy < n_side) & np.isfinite(t)
    idx = np.where(ok, iy * n_side + ix, -1)
    return idx.astype(np.int64)


def _torch():
    import torch
    return torch


class BeamletPlan:
    """Jobs (fields) x beamlets on a loaded context.

    BeamletPlan(ctx, fields, n_per_field, ...): one plan whose field f uses
    source records [f * n_per_field, (f + 1) * n_per_field).
    BeamletPlan.from_jobs(ctx, jobs, n_plans, src, ...): any jobs (dicts with
    field, plan, first, count and optionally history_offset), n_plans dose
    buffers.  Job j owns beamlets [j * n_side^2, (j + 1) * n_side^2); records
    that miss the beamlet grid keep weight 1 through one last beamlet that
    the optimiser does not see.
    """

    def __init__(self, ctx, fields, n_per_field, n_side=20, pitch=0.5, seed=1, n_batches=4, src=None):
        jobs = [dict(field=f, plan=0, first=i * n_per_field, count=n_per_field) for i, f in enumerate(fields)]
        self._setup(ctx, jobs, 1, src, n_side, pitch, seed, n_batches)

    @classmethod
    def from_jobs(cls, ctx, jobs, n_plans, src, n_side=20, pitch=0.5, seed=1, n_batches=4):
        self = cls.__new__(cls)
        self._setup(ctx, jobs, n_plans, src, n_side, pitch, seed, n_batches)
        return self

    def _setup(self, ctx, jobs, n_plans, src, n_side, pitch, seed, n_batches):
        self.ctx = ctx
        self.jobs = [dict(j) for j in jobs]
        self.fields = [j["field"] for j in self.jobs]
        self.n_plans = int(n_plans)
        self.n_side = n_side
        self.seed = seed
        self.n_batches = n_batches
        per_field = n_side * n_side
        self.n_beamlets = per_field * len(self.jobs)
        n_slots = sum(int(j["count"]) for j in self.jobs)
        bl = np.empty(n_slots, dtype=np.int64)
        k = 0
        for i, j in enumerate(self.jobs):
            lo, n = int(j["first"]), int(j["count"])
            b = beamlet_index(src[lo:lo + n], j["field"], n_side, pitch)
            bl[k:k + n] = np.where(b >= 0, b + i * per_field, self.n_beamlets)
            k += n
        self.beamlet = bl.astype(np.uint32)
        self.plan_of_beamlet = np.repeat(np.array([int(j["plan"]) for j in self.jobs], dtype=np.int64), per_field)
        self._n_all = self.n_beamlets + 1

    def plan_beamlets(self, plan):
        """Indices of the beamlets of one plan's fields."""
        return np.flatnonzero(self.plan_of_beamlet == plan)

    def _weights(self, w):
        full = np.ones(self._n_all, dtype=np.float32)
        full[:self.n_beamlets] = np.asarray(w, dtype=np.float32)
        return full

    def _own_map(self):
        # the context holds one beamlet map: (re)load this plan's when another
        # plan's is in place
        if getattr(self.ctx, "_beamlet_owner", None) is not self:
            self.ctx.set_beamlets(self.beamlet, np.ones(self._n_all, dtype=np.float32))
            self.ctx._beamlet_owner = self

    def forward(self, w):
        """Dose for beamlet weights w (host arrays): (nvox,) for one plan,
        (n_plans, nvox) for a queue."""
        self.ctx.set_beamlets(self.beamlet, self._weights(w))
        self.ctx._beamlet_owner = self
        r = self.ctx.run_batched(self.jobs, self.n_plans, self.seed, self.n_batches)
        return r["dose"][0] if self.n_plans == 1 else r["dose"]

    def adjoint(self, dLdD):
        """Gradient with respect to the beamlet weights for dL/dD of shape
        (nvox,) or (n_plans, nvox) (host arrays)."""
        self._own_map()
        g = self.ctx.adjoint(self.jobs, self.n_plans, self.seed, np.asarray(dLdD).reshape(self.n_plans, -1),
                             self._n_all, self.n_batches)
        return g[:self.n_beamlets]

    def forward_device(self, w):
        """Dose (float64 device tensor, (n_plans, nvox)) for the beamlet
        weights w (a 1-D device tensor); nothing leaves the device."""
        torch = _torch()
        self._own_map()
        full = torch.ones(self._n_all, dtype=torch.float32, device=w.device)
        full[:self.n_beamlets] = w.detach().to(torch.float32)
        dose = torch.empty((self.n_plans, self.ctx.nvox), dtype=torch.float64, device=w.device)
        # the library works on its own stream: torch's writes must be done
        torch.cuda.current_stream(w.device).synchronize()
        self.ctx.set_beamlet_weights_device(full.data_ptr(), self._n_all)
        self.ctx.run_batched_device(self.jobs, self.n_plans, self.seed, self.n_batches, dose.data_ptr())
        return dose

    def adjoint_device(self, dLdD):
        """Gradient (float64 device tensor, (n_beamlets,)) for dL/dD given as
        a device tensor of n_plans * nvox values."""
        torch = _torch()
        self._own_map()
        g = dLdD.detach().to(torch.float64).contiguous()
        if g.numel() != self.n_plans * self.ctx.nvox:
            raise ValueError("dL/dD must have n_plans * nvox values")
        grad = torch.empty(self._n_all, dtype=torch.float64, device=g.device)
        torch.cuda.current_stream(g.device).synchronize()
        self.ctx.adjoint_device(self.jobs, self.n_plans, self.seed, g.data_ptr(), grad.data_ptr(), self.n_batches)
        return grad[:self.n_beamlets]


def phase_space_weight_gradient(ctx, jobs, n_plans, src, seed, dLdD, n_batches=4):
    """Gradient of L with respect to the weight of every source record the
    jobs transport, in history-slot order (job after job; a record used by
    two jobs has two slots, and its gradient is their sum).  Each history is
    its own beamlet at unit weight; the replay gathers dL/dD x E over the
    history's deposits, whose energies carry the record's weight, so
    dL/dwt = that sum / wt.  dLdD: (n_plans, nvox) host array."""
    n_slots = sum(int(j["count"]) for j in jobs)
    ctx.set_beamlets(np.arange(n_slots, dtype=np.uint32), np.ones(n_slots, dtype=np.float32))
    ctx._beamlet_owner = None
    g = ctx.adjoint(jobs, n_plans, seed, np.asarray(dLdD).reshape(n_plans, -1), n_slots, n_batches)
    wt = np.concatenate([src["wt"][int(j["first"]):int(j["first"]) + int(j["count"])] for j in jobs])
    return g / wt.astype(np.float64)


class BeamletDose:
    """``BeamletDose.apply(w, plan)``: dose of the plan for the beamlet
    weights w (1-D device tensor), float64 on w's device, (nvox,) for one
    plan or (n_plans, nvox) for a queue; differentiable in w.  Forward and
    backward stay in device memory."""

    _fn = None

    @classmethod
    def apply(cls, w, plan):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, weights, bplan):
                    d = bplan.forward_device(weights)
                    ctx.bplan = bplan
                    ctx.wdtype = weights.dtype
                    return d[0] if bplan.n_plans == 1 else d

                @staticmethod
                def backward(ctx, g):
                    grad = ctx.bplan.adjoint_device(g)
                    return grad.to(ctx.wdtype), None

            cls._fn = _F
        return cls._fn.apply(w, plan)
