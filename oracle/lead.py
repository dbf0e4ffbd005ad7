"""Lead-side helpers (oracle only), restating ``dpnegf/negf/lead_property.py``.

* ``KB_EV``           -> ``Boltzmann / eV2J`` (``lead_property.py:82``,
  ``utils/constants.py:3,108``)
* ``HDL_reduced``     -> ``:386-434``
* ``sigma2gamma``     -> ``sigmaLR2Gamma`` ``:437-453``
* ``fermi_dirac``     -> ``:455-456`` with ``chemiPot_lead = efermi - voltage`` ``:89``
"""
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
