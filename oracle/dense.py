"""Dense brute-force Green's function (oracle pin for ``oracle.rgf``).

Assembles the full matrix ``A = z S - H - Sigma_L - Sigma_R`` of a
block-tridiagonal device and inverts it with LAPACK; the RGF blocks must
equal the corresponding blocks of ``A^{-1}``.
"""
import numpy as np


def assemble(hd, hu, hl, sd, su, sl, sigL=None, sigR=None):
    """Dense H, S (numpy complex128) and block offsets."""
    sizes = [h.shape[0] for h in hd]
    off = np.concatenate([[0], np.cumsum(sizes)])
    N = int(off[-1])
    H = np.zeros((N, N), complex)
    S = np.zeros((N, N), complex)
    for q, n in enumerate(sizes):
        a, b = off[q], off[q + 1]
        H[a:b, a:b] = hd[q]
        S[a:b, a:b] = sd[q]
        if q + 1 < len(sizes):
            c = off[q + 2]
            H[a:b, b:c] = hu[q]
            S[a:b, b:c] = su[q]
            H[b:c, a:b] = hl[q]
            S[b:c, a:b] = sl[q]
    Sig = np.zeros((N, N), complex)
    if sigL is not None:
        r = min(sigL.shape[0], sizes[0])
        Sig[:r, :r] += sigL[:r, :r]
    if sigR is not None:
        r = min(sigR.shape[0], sizes[-1])
        Sig[N - r:, N - r:] += sigR[-r:, -r:]
    return H, S, Sig, off


def dense_green(z, hd, hu, hl, sd, su, sl, sigL=None, sigR=None):
    """Full retarded G = (z S - H - Sigma)^{-1} and the block offsets."""
    H, S, Sig, off = assemble(hd, hu, hl, sd, su, sl, sigL, sigR)
    return np.linalg.inv(z * S - H - Sig), off
