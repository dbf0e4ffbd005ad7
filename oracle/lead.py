"""Lead-side helpers (oracle only), restating ``dpnegf/negf/lead_property.py``.

* ``KB_EV``           -> ``Boltzmann / eV2J`` (``lead_property.py:82``,
  ``utils/constants.py:3,108``)
* ``HDL_reduced``     -> ``:386-434``
* ``sigma2gamma``     -> ``sigmaLR2Gamma`` ``:437-453``
* ``fermi_dirac``     -> ``:455-456`` with ``chemiPot_lead = efermi - voltage`` ``:89``
* ``bloch_unfold``    -> ``Bloch.unfold_points`` (``dpnegf/negf/bloch.py:25-58``)
* ``bloch_self_energy`` -> the ``useBloch`` branch of ``self_energy_cal`` (``:330-379``)
"""
import numpy as np
import torch
from scipy.constants import Boltzmann

EV2J = 1.6021766208e-19
KB_EV = Boltzmann / EV2J


def kBT(e_T=300.0):
    return Boltzmann * e_T / EV2J


def HDL_reduced(HDL, SDL, subblocks):
    """Trim device-lead coupling rows to the first/last BTD block (``:386-434``)."""
    assert HDL.ndim == 2 and SDL.ndim == 2 and HDL.shape == SDL.shape
    nz = HDL.nonzero()
    lo, hi = nz.min(dim=0).values, nz.max(dim=0).values
    if subblocks is None:
        cut = (lo, hi)
    else:
        cut = ((subblocks[-1], subblocks[-1]), (subblocks[0], subblocks[0]))
    if HDL.shape[0] == 1:
        return HDL, SDL
    if lo[0] > 0:                      # right lead
        r = cut[0][0] if subblocks is None else -1 * cut[0][0]
        return HDL[r:, :], SDL[r:, :]
    r = cut[1][0] + 1 if subblocks is None else cut[1][0]
    return HDL[:r, :], SDL[:r, :]


def sigma2gamma(se):
    """Gamma = i (Sigma - Sigma^dagger)."""
    return 1j * (se - se.mH)


def fermi_dirac(x, chemiPot, kbt):
    return 1 / (1 + torch.exp((x - chemiPot) / kbt))


def bloch_unfold(bloch_factor, k):
    """The B0 B1 B2 points k' = (k + j) / B that fold onto k, x fastest
    (``bloch.py:50-58``)."""
    B = np.asarray(bloch_factor)
    out = np.empty([B[2], B[1], B[0], 3])
    out[:, :, :, 0] = (np.arange(B[0]).reshape(1, 1, -1) + k[0]) / B[0]
    out[:, :, :, 1] = (np.arange(B[1]).reshape(1, -1, 1) + k[1]) / B[1]
    out[:, :, :, 2] = (np.arange(B[2]).reshape(-1, 1, 1) + k[2]) / B[2]
    return out.reshape(-1, 3)


def bloch_self_energy(gs_list, k_bloch, R_list, sorted_idx, hDL, sDL, ee):
    """Self-energy of a lead whose principal layer is m = B0 B1 copies of a
    smaller cell (``lead_property.py:336-378``): g_s of the small cell at each
    unfolded k' (``gs_list``, from the decimation without a coupling block),
    g = mean_k' kron(P(k'), g_s(k')) with P_ij = exp(2 pi i (R_j - R_i) . k')
    and P_ii = 1, rows and columns taken in ``sorted_idx`` order, then
    Sigma = (e S_DL - H_DL) g[:b, :b] (e S_DL^H - H_DL^H) with e = E + E_ref
    real (no eta), b = H_DL's column count."""
    m = len(R_list)
    acc = []
    for gs, kb in zip(gs_list, k_bloch):
        kb = torch.as_tensor(np.asarray(kb), dtype=torch.float64)
        P = torch.zeros(m, m, dtype=torch.complex128)
        for i in range(m):
            for j in range(m):
                P[i, j] = 1.0 if i == j else torch.exp(
                    1j * 2 * torch.pi * torch.dot(torch.as_tensor(R_list[j], dtype=torch.float64)
                                                  - torch.as_tensor(R_list[i], dtype=torch.float64),
                                                  kb))
        acc.append(torch.kron(P, gs))
    g = torch.stack(acc).sum(0) / len(acc)
    g = g[sorted_idx, :][:, sorted_idx]
    b = hDL.shape[1]
    e = torch.as_tensor(ee, dtype=torch.complex128)
    return (e * sDL - hDL) @ g[:b, :b] @ (e * sDL.conj().T - hDL.conj().T)
