"""Differentiable Ozaki density: dn/dH by a hand-written adjoint.

``density_blocks(...)`` returns the density matrix of every diagonal block,
DM_q = DM_eq,q + DM_neq,q with (``density.py:164-195``)

  DM_eq  = sum_b Re(c_b G_qq(z_b))                       (R term and poles)
  DM_neq = sum_i a_i Re[G_{q,K-1} Gamma_R G_{q,K-1}^dagger](x_i)   (bias window)

and is a ``torch.autograd.Function``: given W_q = dJ/dDM_q it returns dJ/dH
for every on-site block hd_q and coupling block hu_q, hl_q (torch's complex
convention, dJ/dRe + i dJ/dIm; lead self-energies held fixed).  The
reference has no such kernel: it would rely on autograd through its Python
loop over poles.  With G = (zS - H - Sigma)^-1, dG = G dH G, so

* equilibrium: dJ = Re Tr[Y dH] with Y = sum_b c_b G W^T G, computed block by
  block by the library's bilinear recursion (``dpnegf_rgf_full`` with a
  block-diagonal source S_b = c_b W^T), then dJ/dH_ab = (sum_b Y_ba)^H
  (``dpnegf_batch_sum_conjt``);
* bias window: with W~ = (W + W^T) / 2 and A_R = G Gamma_R G^dagger
  Hermitian, dJ = 2 Re Tr[A_R W~ G dH] a_i.  Since A_R = c Gamma_R c^dagger
  with c = G_{., K-1}, the BTD blocks of A_R W~ G are c_b Gamma_R y_a with the
  block row y = c^dagger W~ G, which one sweep forms from the left-connected
  g_q and the diagonal blocks:
      l_0 = 0,  l_{a+1} = (l_a + r_a) g_a U_a            (r_a = c_a^dagger W~_a)
      t_{K-1} = 0,  t_a = (r_{a+1} G_{a+1,a+1} + t_{a+1}) L_a g_a
      y_a = (l_a + r_a) G_aa + t_a.
The backward pass recomputes the sweeps (nothing but the inputs is kept
from the forward pass).
"""
import torch

from .. import _native as nat
from ._btd import C128, block_sizes, is_orthogonal, pack_arena, pack_batched_blocks, split_blocks
from ._ops import mm, sum_conjt
from .density import eq_density, neq_density, ozaki_points
from .recursive_green_cal import _run_full, full_points_per_call


def _grad_eq(arena, sizes, ortho, eq, W, g_hd, g_hu, g_hl):
    z, sL, nL, sR, nR, coef = eq
    P = z.shape[0]
    K = len(sizes)
    WT = [w.T.to(C128) for w in W]
    chunk = full_points_per_call(sizes, False, P, z.device, mem_fraction=0.25)
    sq = [(n, n) for n in sizes]
    lo = [(sizes[q + 1], sizes[q]) for q in range(K - 1)]
    up = [(sizes[q], sizes[q + 1]) for q in range(K - 1)]
    for c0 in range(0, P, chunk):
        c = slice(c0, min(P, c0 + chunk))
        B = z[c].shape[0]
        cc = coef[c].view(B, 1, 1)
        S = pack_batched_blocks([cc * wt for wt in WT], B, z.device, sq, "bilinear source")
        raw = _run_full(z[c].contiguous(), arena, sizes, ortho, sL[c].contiguous(), nL,
                        sR[c].contiguous(), nR, False, s_bil=S)
        _, _, xd, xl, xu = raw["bilinear"]
        acc = c0 > 0
        for q, blk in enumerate(split_blocks(xd, B, sq)):
            sum_conjt(g_hd[q], blk, acc)
        for q, blk in enumerate(split_blocks(xl, B, lo)):   # Y_{q+1,q} -> dJ/dH_{q,q+1}
            sum_conjt(g_hu[q], blk, acc)
        for q, blk in enumerate(split_blocks(xu, B, up)):   # Y_{q,q+1} -> dJ/dH_{q+1,q}
            sum_conjt(g_hl[q], blk, acc)
        del raw, S


def _grad_neq(arena, sizes, ortho, neq, W, hu, hl, su, sl, g_hd, g_hu, g_hl):
    zx, sLx, nL, sRx, nR, gamR, a = neq
    dev = zx.device
    G = zx.shape[0]
    K = len(sizes)
    nK = sizes[-1]
    Ws = [(0.5 * (w + w.T)).to(C128) for w in W]
    chunk = full_points_per_call(sizes, True, G, dev, mem_fraction=0.25)
    for c0 in range(0, G, chunk):
        c = slice(c0, min(G, c0 + chunk))
        B = zx[c].shape[0]
        zc = zx[c].view(B, 1, 1)
        raw = _run_full(zx[c].contiguous(), arena, sizes, ortho, sLx[c].contiguous(), nL,
                        sRx[c].contiguous(), nR, True)
        gl = split_blocks(raw["gl"], B, [(n, n) for n in sizes])
        grd = split_blocks(raw["grd"], B, [(n, n) for n in sizes])
        lc = split_blocks(raw["lc"], B, [(n, nK) for n in sizes])
        # U_q = hu_q - z su_q and L_q = hl_q - z sl_q (shared when S is orthogonal)
        U = [hu[q] if ortho else hu[q] - zc * su[q] for q in range(K - 1)]
        L = [hl[q] if ortho else hl[q] - zc * sl[q] for q in range(K - 1)]
        r = [mm(lc[q], Ws[q], 1, 0) for q in range(K)]            # [B, nK, n_q]
        ell = [None] * K
        for q in range(K - 1):
            x = r[q] if ell[q] is None else ell[q] + r[q]
            ell[q + 1] = mm(mm(x, gl[q]), U[q])
        tau = [None] * K
        for q in range(K - 2, -1, -1):
            x = mm(r[q + 1], grd[q + 1])
            if tau[q + 1] is not None:
                x = x + tau[q + 1]
            tau[q] = mm(mm(x, L[q]), gl[q])
        y = []
        for q in range(K):
            x = r[q] if ell[q] is None else ell[q] + r[q]
            v = mm(x, grd[q])
            y.append(v if tau[q] is None else v + tau[q])
        # P_q = 2 a_i c_q Gamma_R ; blocks of 2 a A_R W~ G are P_b y_a
        ac = (2.0 * a[c]).to(C128).view(B, 1, 1)
        Pm = [mm(lc[q], gamR[c]) * ac for q in range(K)]
        for q in range(K):
            sum_conjt(g_hd[q], mm(Pm[q], y[q]), True)
        for q in range(K - 1):
            sum_conjt(g_hu[q], mm(Pm[q + 1], y[q]), True)   # X_{q+1,q} -> dJ/dH_{q,q+1}
            sum_conjt(g_hl[q], mm(Pm[q], y[q + 1]), True)   # X_{q,q+1} -> dJ/dH_{q+1,q}
        del raw


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, meta, *blocks):
        K, sd, su, sl, eq, neq = meta
        hd = list(blocks[:K])
        hu = list(blocks[K:2 * K - 1])
        hl = list(blocks[2 * K - 1:3 * K - 2])
        dev = hd[0].device
        sizes = block_sizes(hd)
        arena = pack_arena(hd, hu, hl, sd, su, sl, dev)
        ortho = is_orthogonal(sd, su, sl)
        dm = eq_density(arena, sizes, ortho, eq)
        if neq is not None:
            dn = neq_density(arena, sizes, ortho, neq)
            dm = [x + y for x, y in zip(dm, dn)]
        ctx.meta = meta
        ctx.save_for_backward(*blocks)
        return tuple(dm)

    @staticmethod
    def backward(ctx, *gdm):
        K, sd, su, sl, eq, neq = ctx.meta
        blocks = [b.detach() for b in ctx.saved_tensors]
        hd, hu, hl = blocks[:K], blocks[K:2 * K - 1], blocks[2 * K - 1:3 * K - 2]
        dev = hd[0].device
        sizes = block_sizes(hd)
        arena = pack_arena(hd, hu, hl, sd, su, sl, dev)
        ortho = is_orthogonal(sd, su, sl)
        W = [(g if g is not None else torch.zeros(n, n, dtype=torch.float64, device=dev))
             .to(torch.float64).contiguous() for g, n in zip(gdm, sizes)]
        g_hd = [torch.zeros(n, n, dtype=C128, device=dev) for n in sizes]
        g_hu = [torch.zeros(sizes[q], sizes[q + 1], dtype=C128, device=dev) for q in range(K - 1)]
        g_hl = [torch.zeros(sizes[q + 1], sizes[q], dtype=C128, device=dev) for q in range(K - 1)]
        _grad_eq(arena, sizes, ortho, eq, W, g_hd, g_hu, g_hl)
        if neq is not None:
            _grad_neq(arena, sizes, ortho, neq, W, hu, hl, su, sl, g_hd, g_hu, g_hl)
        return (None, *g_hd, *g_hu, *g_hl)


def density_blocks(hd, hu, hl, sd, su, sl, eq, neq=None):
    """Differentiable DM_eq + DM_neq of every diagonal block (real
    [n_q, n_q] tensors); ``eq`` / ``neq`` are the point sets of
    :func:`~dpnegf_amd.negf.density.ozaki_points` (their self-energies are
    constants).  Gradients flow to ``hd``, ``hu``, ``hl``."""
    nat.require_device(hd[0])
    K = len(hd)
    meta = (K, [s.detach() for s in sd], [s.detach() for s in su], [s.detach() for s in sl],
            eq, neq)
    return list(_Density.apply(meta, *hd, *hu, *hl))


def ozaki_density_differentiable(blocks, lead_L, lead_R, kBT, mu_L, mu, R, poles, residues,
                                 E_ref, eta_lead, eta_device, vL=0.0, vR=0.0, n_gauss=10,
                                 fermi_L=None, fermi_R=None, eta_lead0=1e-5, eta_device0=0.0):
    """:func:`~dpnegf_amd.negf.density.ozaki_density` (same arguments) as one
    differentiable total: returns DM_eq + DM_neq per diagonal block, with
    dn/dH for every device block through the hand-written adjoint."""
    hd, hu, hl, sd, su, sl = blocks
    dev = hd[0].device
    sizes = block_sizes(hd)
    eq, neq = ozaki_points(sizes, lead_L, lead_R, kBT, mu_L, mu, R, poles, residues, E_ref,
                           eta_lead, eta_device, vL, vR, n_gauss, fermi_L, fermi_R, eta_lead0,
                           eta_device0, dev)
    return density_blocks(hd, hu, hl, sd, su, sl, eq, neq)
