"""Differentiable transmission: fused forward kernel + hand-written adjoint.

``transmission(...)`` returns T(E_b) for B energies of one system and is a
``torch.autograd.Function``: ``backward`` receives dJ/dT_b and calls
``dpnegf_rgf_transmission_grad``, which returns dJ/dH for every on-site block
``hd_q`` and coupling block ``hu_q`` / ``hl_q`` summed over energies, the
same for the overlap blocks ``sd_q`` / ``su_q`` / ``sl_q``, and dJ/dSigma_L,
dJ/dSigma_R per energy (torch's complex-gradient convention, dJ/dRe + i
dJ/dIm); each only when some input needs it.  The reference has no such
kernel; it relies on autograd through ``recursive_gf``
(``dpnegf/negf/recursive_green_cal.py:89-112``), which is what the tests
compare against, together with central finite differences.

Memory: the plain adjoint keeps every block of the sweep for all B energies,
O(K) blocks per energy.  With ``checkpoint`` it keeps the left-connected g_q
and the first-row products only every s blocks and recomputes each segment of
s blocks during the gradient sweep (``dpnegf_rgf_transmission_grad_ckpt``):
O(K / s + s) blocks per energy for one more forward sweep, s = ceil(sqrt(K))
by default.
"""
import math

import torch

from .. import _native as nat
from ._btd import C128, split_blocks
from .recursive_green_cal import BTDSystems, rgf_transmission


class _Transmission(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta, z, sigL, sigR, *blocks):
        K, precision, checkpoint = meta
        hd = list(blocks[:K])
        hu = list(blocks[K:2 * K - 1])
        hl = list(blocks[2 * K - 1:3 * K - 2])
        sd = list(blocks[3 * K - 2:4 * K - 2])
        su = list(blocks[4 * K - 2:5 * K - 3])
        sl = list(blocks[5 * K - 3:6 * K - 4])
        dev = z.device
        bs = BTDSystems([(hd, hu, hl, sd, su, sl)], dev)
        T = rgf_transmission(bs, z, sigL, sigR, precision=precision)
        ctx.K = K
        ctx.bs = bs
        ctx.checkpoint = checkpoint
        ctx.save_for_backward(z, sigL, sigR)
        return T

    @staticmethod
    def backward(ctx, gT):
        z, sigL, sigR = ctx.saved_tensors
        bs = ctx.bs
        K = ctx.K
        sizes = bs.sizes
        dev = z.device
        B = z.shape[0]
        w = gT.detach().to(device=dev, dtype=torch.float64).contiguous()
        h_bsz = nat.host_i32(sizes)
        stride = checkpoint_stride(ctx.checkpoint, sizes, B, dev)
        want_sig = ctx.needs_input_grad[2] or ctx.needs_input_grad[3]
        gsL = torch.empty_like(sigL) if want_sig else None
        gsR = torch.empty_like(sigR) if want_sig else None
        want_s = any(ctx.needs_input_grad[4 + 3 * K - 2:])
        if stride:
            nwork = nat.load().dpnegf_rgf_grad_ckpt_workspace_elems(K, h_bsz, B, stride)
        else:
            nwork = nat.load().dpnegf_rgf_grad_workspace_elems(K, h_bsz, B)
        work = torch.empty(nwork, dtype=C128, device=dev)
        nsq = sum(n * n for n in sizes)
        nx = sum(sizes[q] * sizes[q + 1] for q in range(K - 1))
        ghd = torch.empty(nsq, dtype=C128, device=dev)
        ghu = torch.empty(max(nx, 1), dtype=C128, device=dev)
        ghl = torch.empty(max(nx, 1), dtype=C128, device=dev)
        gsd = torch.empty(nsq, dtype=C128, device=dev) if want_s else None
        gsu = torch.empty(max(nx, 1), dtype=C128, device=dev) if want_s else None
        gsl = torch.empty(max(nx, 1), dtype=C128, device=dev) if want_s else None
        status = torch.zeros(B, dtype=torch.int32, device=dev)
        args = (nat.stream_ptr(dev), K, h_bsz, nat.ptr(bs.arena), B, nat.ptr(z), nat.ptr(sigL),
                sigL.shape[-1], nat.ptr(sigR), sigR.shape[-1], int(bs.ortho), nat.ptr(w),
                nat.ptr(ghd), nat.ptr(ghu), nat.ptr(ghl), nat.ptr(gsL), nat.ptr(gsR),
                nat.ptr(gsd), nat.ptr(gsu), nat.ptr(gsl))
        if stride:
            nat.call("dpnegf_rgf_transmission_grad_ckpt", *args, stride, nat.ptr(work),
                     nat.ptr(status))
        else:
            nat.call("dpnegf_rgf_transmission_grad", *args, nat.ptr(work), nat.ptr(status))
        g_hd = [t[0] for t in split_blocks(ghd, 1, [(n, n) for n in sizes])]
        up = [(sizes[q], sizes[q + 1]) for q in range(K - 1)]
        g_hu = [t[0] for t in split_blocks(ghu, 1, up)]
        g_hl = [t[0] for t in split_blocks(ghl, 1, [(b, a) for a, b in up])]
        if want_s:
            g_s = ([t[0] for t in split_blocks(gsd, 1, [(n, n) for n in sizes])]
                   + [t[0] for t in split_blocks(gsu, 1, up)]
                   + [t[0] for t in split_blocks(gsl, 1, [(b, a) for a, b in up])])
        else:
            g_s = [None] * (3 * K - 2)
        return (None, None, gsL if ctx.needs_input_grad[2] else None,
                gsR if ctx.needs_input_grad[3] else None, *g_hd, *g_hu, *g_hl, *g_s)


def checkpoint_stride(checkpoint, sizes, B, dev, mem_fraction=0.5):
    """Checkpoint stride of the adjoint (0: keep every block).

    ``checkpoint``: None picks 0 when the plain adjoint's work buffer fits in
    ``mem_fraction`` of the free device memory, else ceil(sqrt(K)); "sqrt"
    always ceil(sqrt(K)); an int >= 1 that stride; 0 keeps every block."""
    K = len(sizes)
    sqrt_k = max(1, math.isqrt(K - 1) + 1) if K > 1 else 1
    if checkpoint is None:
        full = nat.load().dpnegf_rgf_grad_workspace_elems(K, nat.host_i32(sizes), B) * 16
        free, _ = torch.cuda.mem_get_info(dev)
        return 0 if full <= mem_fraction * free else sqrt_k
    if checkpoint == "sqrt":
        return sqrt_k
    stride = int(checkpoint)
    if stride < 0:
        raise ValueError(f"checkpoint stride must be >= 0, got {checkpoint}")
    return stride


def transmission(energy, hd, hu, hl, sd, su, sl, sigL, sigR, eta, E_ref=0.0, precision="fp64",
                 checkpoint=None):
    """Differentiable T(E) [B] of one system.

    ``energy`` [B] relative to ``E_ref``; ``sigL`` [B, nL, nL] / ``sigR``
    [B, nR, nR] the lead self-energies trimmed to their corners.  Gradients
    flow to ``hd``, ``hu``, ``hl`` and ``sd``, ``su``, ``sl`` (summed over
    the energies) and to ``sigL``, ``sigR`` (per energy).  ``checkpoint``:
    memory mode of the backward pass (``checkpoint_stride``).
    """
    dev = hd[0].device
    nat.require_device(hd[0])
    K = len(hd)
    z = (torch.as_tensor(energy).to(device=dev, dtype=C128).reshape(-1) + E_ref + 1j * eta)
    z = z.contiguous()
    meta = (K, precision, checkpoint)
    return _Transmission.apply(meta, z, sigL.contiguous(), sigR.contiguous(), *hd, *hu, *hl,
                               *sd, *su, *sl)
