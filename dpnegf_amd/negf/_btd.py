"""Device-side packing of block-tridiagonal inputs into the library's arenas.

The C ABI takes one contiguous complex128 "arena" per system with the blocks
in the order hd_q, sd_q, hu_q, su_q, hl_q, sl_q (q = 0..K-1, the last block
without couplings).  These helpers build it with torch ops on the GPU; they
carry no NEGF arithmetic.
"""
import numpy as np
import torch

C128 = torch.complex128


def _flat(t, device):
    return t.to(device=device, dtype=C128).contiguous().reshape(-1)


def block_sizes(hd):
    sizes = [int(h.shape[-1]) for h in hd]
    for h in hd:
        assert h.ndim == 2 and h.shape[0] == h.shape[1], "diagonal blocks must be square 2-D"
    return sizes


def is_orthogonal(sd, su, sl):
    """True when every S_qq is the identity and every off-diagonal S block is zero."""
    for s in sd:
        if not torch.equal(s, torch.eye(s.shape[0], dtype=s.dtype, device=s.device)):
            return False
    for s in list(su) + list(sl):
        if torch.count_nonzero(s).item() != 0:
            return False
    return True


def pack_arena(hd, hu, hl, sd, su, sl, device):
    K = len(hd)
    assert len(sd) == K and len(hu) == K - 1 and len(hl) == K - 1
    assert len(su) == K - 1 and len(sl) == K - 1
    parts = []
    for q in range(K):
        parts += [_flat(hd[q], device), _flat(sd[q], device)]
        if q < K - 1:
            parts += [_flat(hu[q], device), _flat(su[q], device), _flat(hl[q], device),
                      _flat(sl[q], device)]
    return torch.cat(parts)


def batch_of(t, B, what="tensor"):
    """``t`` as [B, m, n]: a 2-D matrix or a [1, m, n] batch is broadcast over
    the B energies (as torch broadcasting does in the reference,
    ``recursive_green_cal.py:397-428``); any other batch size is an error, so
    no caller can hand the library fewer than B matrices."""
    if t.ndim == 2:
        return t.unsqueeze(0).expand(B, -1, -1)
    if t.ndim != 3:
        raise ValueError(f"{what} must be 2-D or [B, m, n] (got shape {tuple(t.shape)})")
    if t.shape[0] == B:
        return t
    if t.shape[0] == 1:
        return t.expand(B, -1, -1)
    raise ValueError(f"{what} has batch size {t.shape[0]}, expected {B} energies")


def pack_batched_blocks(blocks, B, device, shapes=None, what="block"):
    """List of [B, m, n] (or 2-D / [1, m, n], broadcast) tensors -> one
    contiguous arena; ``shapes`` (optional) are the required (m, n) per block."""
    out = []
    for q, t in enumerate(blocks):
        t = batch_of(t.to(device=device, dtype=C128), B, f"{what} {q}")
        if shapes is not None and tuple(t.shape[-2:]) != tuple(shapes[q]):
            raise ValueError(f"{what} {q} has shape {tuple(t.shape[-2:])}, expected "
                             f"{tuple(shapes[q])}")
        out.append(t.contiguous().reshape(-1))
    return torch.cat(out)


def split_blocks(arena, B, shapes):
    """Inverse of pack_batched_blocks: views [B, m, n] into one arena."""
    out, off = [], 0
    for (m, n) in shapes:
        sz = B * m * n
        out.append(arena[off:off + sz].view(B, m, n))
        off += sz
    return out


def corner_self_energy(se, n_block, B, at_end, device):
    """Trim a self-energy to its corner block the way recursive_gf does
    (dpnegf/negf/recursive_green_cal.py:416-428); returns ([B, r, r], r)."""
    se = batch_of(se, B, "self-energy")
    r = min(n_block, se.shape[-2])
    c = min(n_block, se.shape[-1])
    if r != c:
        raise ValueError("self-energy corner must be square")
    se = se[:, -r:, -c:] if at_end else se[:, :r, :c]
    return se.to(device=device, dtype=C128).contiguous(), r


def kvec(kpoint):
    """A k-point as a host float64 tensor (a list goes through numpy, since
    torch would make float32 of Python floats under its default dtype)."""
    if torch.is_tensor(kpoint):
        return kpoint.detach().cpu().to(torch.float64)
    return torch.as_tensor(np.asarray(kpoint, dtype=float))
