"""Ozaki continued-fraction poles and residues (``dpnegf/negf/ozaki_res_cal.py:8-37``).

A one-off host computation (a 2M x 2M symmetric tridiagonal eigenproblem),
as in the reference; the poles then drive the batched GPU density sum.
"""
import numpy as np
import torch
from scipy.linalg import eigh_tridiagonal


def ozaki_residues(M_cut: int = 1000):
    """Positive-half poles (ascending, in units of kT) and their residues.

    The matrix has zero diagonal and off-diagonal 1/(2 sqrt((2n-1)(2n+1)));
    poles are 1/lambda of its positive eigenvalues and residues
    |v_0|^2 / (4 lambda^2) (Karrasch, Meden, Schoenhammer, PRB 82, 125114).
    """
    M_cut = int(M_cut)
    N = 2 * M_cut
    n = np.arange(1, N, dtype=np.float64)
    off = 0.5 / np.sqrt((2.0 * n - 1.0) * (2.0 * n + 1.0))
    lam, vec = eigh_tridiagonal(np.zeros(N), off, select='i', select_range=(N // 2, N - 1))
    poles = (1.0 / lam)[::-1].copy()
    res = (np.abs(vec[0, :]) ** 2 / (4.0 * lam ** 2))[::-1].copy()
    return torch.from_numpy(poles), torch.from_numpy(res)
