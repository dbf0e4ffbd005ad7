"""Batched complex products through the library's kernels (``dpnegf_zgemm_batched``,
``dpnegf_diag_product``, ``dpnegf_batch_sum_conjt``) for the host-side glue
of the observables and the adjoints; no arithmetic of their own."""
import torch

from .. import _native as nat
from ._btd import C128


def batched(x):
    x = x.to(C128)
    return (x if x.ndim == 3 else x.unsqueeze(0)).contiguous()


def mm(A, B, opA=0, opB=0):
    """op(A) op(B) through the library's batched product
    (``dpnegf_zgemm_batched``); 2-D operands are shared over the batch, a 2-D
    result is returned when both are 2-D.  op 0 = N, 1 = conjugate transpose."""
    assert opA in (0, 1) and opB in (0, 1)
    squeeze = A.ndim == 2 and B.ndim == 2
    Ab, Bb = batched(A), batched(B)
    nb = max(Ab.shape[0], Bb.shape[0])
    m, k = (Ab.shape[1], Ab.shape[2]) if opA == 0 else (Ab.shape[2], Ab.shape[1])
    k2, n = (Bb.shape[1], Bb.shape[2]) if opB == 0 else (Bb.shape[2], Bb.shape[1])
    assert k == k2, "inner dimensions differ"
    C = torch.empty(nb, m, n, dtype=C128, device=Ab.device)
    sa = 0 if Ab.shape[0] == 1 else Ab.shape[1] * Ab.shape[2]
    sb = 0 if Bb.shape[0] == 1 else Bb.shape[1] * Bb.shape[2]
    st = nat.stream_ptr(Ab.device)
    for b0 in range(0, nb, 65535):
        nbc = min(65535, nb - b0)
        nat.call("dpnegf_zgemm_batched", st, nbc, m, n, k, nat.ptr(Ab[b0 * (sa > 0):]), Ab.shape[2],
                 sa, opA, nat.ptr(Bb[b0 * (sb > 0):]), Bb.shape[2], sb, opB, 1.0, 0.0, 0.0, 0.0,
                 nat.ptr(C[b0:]), n, m * n)
    return C[0] if squeeze else C


def diag_of_product(pairs, batch, device):
    """sum over (A, B) pairs of diag(A B) through ``dpnegf_diag_product``:
    A [B, m, n] (or 2-D, shared), B [n, m] (or batched); returns [B, m]."""
    out = None
    st = nat.stream_ptr(device)
    for i, (A, Bm) in enumerate(pairs):
        Ab, Bb = batched(A), batched(Bm)
        m, n = Ab.shape[1], Ab.shape[2]
        assert Bb.shape[1] == n and Bb.shape[2] == m
        if out is None:
            out = torch.empty(batch, m, dtype=C128, device=device)
        sa = 0 if Ab.shape[0] == 1 else m * n
        sb = 0 if Bb.shape[0] == 1 else m * n
        nat.call("dpnegf_diag_product", st, nat.ptr(out), nat.ptr(Ab), sa, nat.ptr(Bb), sb, m, n,
                 batch, int(i > 0))
    return out


def sum_conjt(out, x, accumulate):
    """out (+)= (sum_b x[b])^H for x [B, m, n] contiguous; out [n, m]."""
    B, m, n = x.shape
    nat.call("dpnegf_batch_sum_conjt", nat.stream_ptr(x.device), nat.ptr(out), nat.ptr(x), m, n,
             B, int(accumulate))
    return out
