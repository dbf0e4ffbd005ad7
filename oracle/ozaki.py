"""Ozaki pole-sum equilibrium density (oracle only).

* ``ozaki_residues``  -> ``dpnegf/negf/ozaki_res_cal.py:8-37``
* ``gauss_xw``        -> ``dpnegf/negf/negf_utils.py:136-159``
* ``ozaki_density``   -> ``Ozaki.integrate`` ``dpnegf/negf/density.py:143-199``,
  with the ``.mu`` attributes that do not exist at HEAD (``density.py:166``)
  mapped to the chemical potentials they stand for, and generalised from
  block 0 to every diagonal block.
"""
import numpy as np
import torch
from scipy.linalg import eigh_tridiagonal


def ozaki_residues(M_cut=1000):
    """Poles (ascending, units of kT) and residues of the Ozaki continued
    fraction: 2M x 2M tridiagonal, zero diagonal, off-diagonal
    1/(2 sqrt((2n-1)(2n+1))); the positive half of its spectrum."""
    M_cut = int(M_cut)
    N = 2 * M_cut
    d = np.zeros(N)
    e = np.array([0.5 / np.sqrt((2.0 * n - 1) * (2.0 * n + 1)) for n in range(1, N)])
    lam, vec = eigh_tridiagonal(d=d, e=e, select='i', select_range=(N // 2, N - 1))
    poles = np.flip(1.0 / lam)
    res = np.flip(np.abs(vec[0, :]) ** 2 / (4.0 * lam ** 2))
    return torch.from_numpy(poles.copy()), torch.from_numpy(res.copy())


def gauss_xw(xl, xu, n=100):
    """Gauss-Legendre nodes/weights on [xl, xu]."""
    ndim = len(xu.shape)
    x, w = np.polynomial.legendre.leggauss(n)
    x = torch.tensor(x, dtype=xu.dtype, device=xu.device)[(...,) + (None,) * ndim]
    w = torch.tensor(w, dtype=xu.dtype, device=xu.device)[(...,) + (None,) * ndim]
    w = w * (0.5 * (xu - xl))
    xs = x * (0.5 * (xu - xl)) + 0.5 * (xu + xl)
    return xs, w


def ozaki_density(grd_of, gammaR_of, fR, fL, poles, residues, kBT, R, mu_shift,
                  vL, vR, n_gauss):
    """Density matrices of every diagonal block, Ozaki form of ``density.py:164-199``.

    ``grd_of(E)`` returns the list of diagonal retarded blocks at (complex)
    energy ``E`` (relative to ``E_ref``); ``gammaR_of(E)`` the right-lead
    broadening padded into the last block's shape, or into every block when
    ``grd_of`` is full-column.  Returns ``(DM_eq blocks, DM_neq blocks or 0)``.
    ``mu_shift`` = mu_L - mu (the pole offset of ``density.py:166``).
    """
    z = 1j * poles * kBT + mu_shift
    g0 = grd_of(1j * R - 0.0)
    dm = [1j * R * g for g in g0]
    for i in range(len(z)):
        g = grd_of(z[i])
        for q in range(len(dm)):
            dm[q] = dm[q] - ((-4 * 1j * kBT) * g[q] * residues[i]).imag
    dm = [d.real for d in dm]
    if abs(vL - vR) <= 1e-14:
        return dm, 0.0
    xl, xu = min(vL, vR) - 8 * kBT, max(vL, vR) + 8 * kBT
    xs, wl = gauss_xw(torch.scalar_tensor(xl, dtype=torch.float64),
                      torch.scalar_tensor(xu, dtype=torch.float64), n=n_gauss)
    neq = None
    for i in range(len(xs)):
        e = xs[i]
        blocks = gammaR_of(e)
        f = fR(e) - fL(e)
        term = [b.real * f * wl[i] for b in blocks]
        neq = term if neq is None else [a + b for a, b in zip(neq, term)]
    return dm, neq
