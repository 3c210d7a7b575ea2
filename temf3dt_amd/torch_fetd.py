"""torch.autograd.Function over the libtemfetd forward loop and its adjoint.

    fields = fetd_fields(log_sigma, fe, inv_mu)      # [rows][R][9] (or [rows][K][R][9])
    loss = ((gate_values(fields, t, gates)[..., 5] - d_obs) ** 2).sum()
    loss.backward()                                  # log_sigma.grad from tf_adjoint

The forward sets sigma = exp(log_sigma) on the context, re-assembles, and
runs the BDF2 loop with its tape recorded (history checkpoints every
`ckpt_every` steps); the backward hands dloss/dfields to tf_adjoint, which
sweeps backward in time reusing the LDL^T factors and returns
dloss/dlog(sigma) per tet.  The controller's keep/change decisions are those
of the forward run, or the `replay` list when given (a fixed schedule makes
the map smooth in sigma).  PyTorch only carries the tensors; all the
arithmetic runs in libtemfetd on the GPU.
"""
import numpy as np
import torch


class FetdForward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_sigma, fe, inv_mu, ckpt_every, replay):
        sig = torch.exp(log_sigma.detach()).to("cpu", torch.float64).numpy()
        fe.set_materials(sig, inv_mu)
        fe.assemble()
        if replay is not None:
            fe.set_replay(replay)
        fe.set_checkpointing(ckpt_every)
        rows = fe.run()
        ctx.fe = fe
        ctx.rows = rows
        out = torch.from_numpy(np.ascontiguousarray(rows["fields"])).to(log_sigma.device)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        w = grad_out.detach().to("cpu", torch.float64).contiguous().numpy()
        g = ctx.fe.adjoint(w)
        return torch.from_numpy(g).to(grad_out.device), None, None, None, None


def fetd_fields(log_sigma, fe, inv_mu, ckpt_every=50, replay=None):
    """Receiver rows (ex ey ez dBx/dt dBy/dt dBz/dt bx by bz) as a function of
    per-tet log-conductivity, differentiable through the adjoint."""
    return FetdForward.apply(log_sigma, fe, np.ascontiguousarray(inv_mu, np.float64), int(ckpt_every), replay)


def gate_values(fields, t, gates):
    """The .field2 values: each gate from the first row with t >= gate,
    linearly interpolated with the row before (src/FETD_FWD.f90:1625-1641).
    Gates before the second row or beyond the last are left out; returns
    (values [ngates_kept, ...], kept gate indices)."""
    t = np.asarray(t, np.float64)
    idx, a0, a1 = [], [], []
    for g, tg in enumerate(gates):
        k = int(np.searchsorted(t, tg, side="left"))
        if k == 0 or k >= len(t):
            continue
        x1, x2 = t[k - 1], t[k]
        # y = a*tg + b with a = (y1-y2)/(x1-x2), b = y1 - a*x1: weights on y1, y2
        w2 = (tg - x1) / (x2 - x1)
        idx.append(g)
        a0.append(1.0 - w2)
        a1.append(w2)
    if not idx:
        return fields[:0], []
    ks = [int(np.searchsorted(t, gates[g], side="left")) for g in idx]
    lo = torch.as_tensor([k - 1 for k in ks], device=fields.device)
    hi = torch.as_tensor(ks, device=fields.device)
    shape = (-1,) + (1,) * (fields.dim() - 1)
    w0 = torch.as_tensor(a0, dtype=fields.dtype, device=fields.device).reshape(shape)
    w1 = torch.as_tensor(a1, dtype=fields.dtype, device=fields.device).reshape(shape)
    return w0 * fields[lo] + w1 * fields[hi], idx
