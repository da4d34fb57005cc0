"""Beamlet-weight dose and its adjoint gradient, exposed to PyTorch.

A plan's fields are divided into beamlets (square pencils of the aperture at
the isocentre plane).  Every source record belongs to one beamlet; the dose of
the plan is D = sum_b w_b D_b, where D_b is the dose of beamlet b's histories
at unit weight.  ``BeamletDose`` is a ``torch.autograd.Function``: its forward
runs the batched event-queued transport with the weights applied to the source
records, its backward replays the same histories from their counter-based
Philox streams (gegs_adjoint) and gathers dL/dD at every deposit.  The
reference (RicciWoo/GEGS) has no differentiable path; this is the addition
the dose-optimisation use of the transport needs.
"""
from __future__ import annotations

import numpy as np

from . import workloads


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


class BeamletPlan:
    """Fields x beamlets on a loaded context.

    fields:  list of workloads.Field
    src:     source records (already loaded into ctx), one block per field:
             field f uses records [f*n_per_field, (f+1)*n_per_field)
    """

    def __init__(self, ctx, fields, n_per_field, n_side=20, pitch=0.5, seed=1, n_batches=4, src=None):
        self.ctx = ctx
        self.fields = fields
        self.n_side = n_side
        self.seed = seed
        self.n_batches = n_batches
        self.jobs = [dict(field=f, plan=0, first=i * n_per_field, count=n_per_field) for i, f in enumerate(fields)]
        per_field = n_side * n_side
        self.n_beamlets = per_field * len(fields)
        bl = np.zeros(n_per_field * len(fields), dtype=np.int64)
        for i, f in enumerate(fields):
            part = src[i * n_per_field:(i + 1) * n_per_field]
            b = beamlet_index(part, f, n_side, pitch)
            # records outside the beamlet grid keep weight 1 through a
            # dedicated last beamlet that the optimiser does not see
            bl[i * n_per_field:(i + 1) * n_per_field] = np.where(b >= 0, b + i * per_field, self.n_beamlets)
        self.beamlet = bl.astype(np.uint32)
        self._n_all = self.n_beamlets + 1

    def _weights(self, w):
        full = np.ones(self._n_all, dtype=np.float32)
        full[:self.n_beamlets] = np.asarray(w, dtype=np.float32)
        return full

    def forward(self, w):
        self.ctx.set_beamlets(self.beamlet, self._weights(w))
        r = self.ctx.run_batched(self.jobs, 1, self.seed, self.n_batches)
        return r["dose"][0]

    def adjoint(self, dLdD):
        self.ctx.set_beamlets(self.beamlet, self._weights(np.ones(self.n_beamlets)))
        g = self.ctx.adjoint(self.jobs, 1, self.seed, np.asarray(dLdD).reshape(1, -1), self._n_all, self.n_batches)
        return g[:self.n_beamlets]


def _torch():
    import torch
    return torch


class BeamletDose:
    """``BeamletDose.apply(w, plan)``: dose (nvox, FP64) of the plan for the
    beamlet weights w (1-D tensor); differentiable in w."""

    _fn = None

    @classmethod
    def apply(cls, w, plan):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, weights, bplan):
                    d = bplan.forward(weights.detach().to("cpu", torch.float32).numpy())
                    ctx.bplan = bplan
                    ctx.wmeta = (weights.dtype, weights.device)
                    return torch.from_numpy(d).to(weights.device)

                @staticmethod
                def backward(ctx, g):
                    grad = ctx.bplan.adjoint(g.detach().to("cpu", torch.float64).numpy())
                    dtype, device = ctx.wmeta
                    return torch.from_numpy(grad).to(device=device, dtype=dtype), None

            cls._fn = _F
        return cls._fn.apply(w, plan)
