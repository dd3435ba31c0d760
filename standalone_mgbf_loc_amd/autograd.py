"""torch.autograd over the MGBF operator B_w = W A_w W^T.

`localize(op, z, weights)` returns B_w z, where A_w scales generation g of
the multigrid vector by weights[g] (a scalar tensor, or a per-point field of
the generation's shape: the reference's a_diff_f / a_diff_h,
src.mgbf/mg_intstate.f90:333-351). Gradients:

    dJ/dz   = B_w g                          (B_w is symmetric)
    dJ/dw_g = < (W^T g)_g , (W^T z)_g >      scalar weight, FP64 accumulation
    dJ/dw_g = (W^T g)_g * (W^T z)_g          per-point weight field

with g = dJ/dy. The per-generation states W^T z are recomputed in the
backward pass, not stored; W^T g is computed once and also gives
B_w g = W A_w (W^T g). The adjoint halves are the hand-written kernels of
the library (W^T: restriction chain + radial filter; W: radial filter +
prolongation chain + vertical filter).
"""
from __future__ import annotations

import torch

from . import mgbf as M


def _scalar_values(weights):
    """Host values of the scalar weights (None for fields), read from the
    device in one transfer."""
    idx = [i for i, w in enumerate(weights) if w.dim() == 0 or w.numel() == 1]
    vals = [None] * len(weights)
    if idx:
        host = torch.stack([weights[i].detach().reshape(()).double() for i in idx]).cpu().tolist()
        for i, v in zip(idx, host):
            vals[i] = v
    return vals


def _set_weights(op: M.MGBF, weights, values):
    for g, (w, v) in enumerate(zip(weights, values), start=1):
        if v is not None:
            op.set_weights(g, value=v)
        else:
            op.set_weights(g, field=w.detach().to(torch.float32).contiguous())


class _Localize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, op, *weights):
        if op.cfg.mgbf_proc == 7 and op.cfg.l_vertical_filter:
            # filtering_lin2_ens ends with Vf^T (mg_filtering.f90:1303-1305)
            raise ValueError("mgbf_proc=7 ends with the adjoint vertical filter: its W A W^T is "
                             "not symmetric, and the gradients here assume it is")
        if len(weights) != op.dims.gm:
            raise ValueError(f"need {op.dims.gm} generation weights")
        values = _scalar_values(weights)
        _set_weights(op, weights, values)
        zc = z.detach().contiguous()
        y = op.apply(M.OP_WAWT, zc)
        ctx.op = op
        ctx.values = values
        ctx.save_for_backward(zc, *weights)
        return y

    @staticmethod
    def backward(ctx, gy):
        op = ctx.op
        z, *weights = ctx.saved_tensors
        _set_weights(op, weights, ctx.values)
        gy = gy.contiguous()
        need_w = any(ctx.needs_input_grad[2:])
        gz, wg = None, None
        if need_w:
            # W^T g serves both: the weight gradients and B g = W A (W^T g)
            wg = op.apply(M.OP_WT, gy)
            if ctx.needs_input_grad[0]:
                gz = op.apply(M.OP_W, op.apply(M.OP_A, wg))
        elif ctx.needs_input_grad[0]:
            gz = op.apply(M.OP_WAWT, gy)
        gw = [None] * len(weights)
        if need_w:
            # recompute the per-generation states of both vectors
            vg = op.split_mg(wg)
            vz = op.split_mg(op.apply(M.OP_WT, z))
            for g, w in enumerate(weights):
                if not ctx.needs_input_grad[2 + g]:
                    continue
                if g == 0 and not _filters_g1(op):
                    gw[g] = torch.zeros_like(w)
                elif w.dim() == 0 or w.numel() == 1:
                    val = op.dot(vg[g], vz[g])
                    gw[g] = torch.tensor(val, dtype=w.dtype, device=w.device).reshape(w.shape)
                else:
                    gw[g] = (vg[g] * vz[g]).to(w.dtype)
        return (gz, None, *gw)


def _filters_g1(op: M.MGBF) -> bool:
    # l_filt_g1 is read by the ensemble procedures (6-8) only
    return op.cfg.mgbf_proc in (1, 2, 3, 4, 5) or bool(op.cfg.l_filt_g1)


def localize(op: M.MGBF, z: torch.Tensor, weights) -> torch.Tensor:
    """y = W A_w W^T z with differentiable z and generation weights."""
    return _Localize.apply(z, op, *weights)


def loss_and_grads(op: M.MGBF, a, x, y, weights):
    """J = 1/2 ||W A_w W^T (a * x) - y||^2 and its gradients w.r.t. a and w.
    J is accumulated in FP64 on the device (mgbf_dot); dJ/dr = r enters the
    autograd graph of r = W A_w W^T (a * x) - y directly."""
    a = a.detach().requires_grad_(True)
    ws = [w.detach().requires_grad_(True) for w in weights]
    r = localize(op, a * x, ws) - y
    rd = r.detach().contiguous()
    J = 0.5 * op.dot(rd, rd)
    r.backward(rd)
    return J, a.grad, [w.grad for w in ws]
