"""Fiori real-axis free charge (oracle only), restating
``dpnegf/negf/density.py:234-393`` (class ``Fiori``):

* ``integration_points`` -> ``Fiori.__init__`` ``:236-265`` (``"gauss"``:
  Gauss-Legendre nodes and weights on [e_grid[0], e_grid[-1]]; ``"direct"``:
  the grid with weight dE = e_grid[1] - e_grid[0]);
* ``free_charge``        -> ``density_integrate_Fiori`` ``:268-360``, one energy
  at a time as the reference loops: A_R = G_{q,K-1} Gamma_R G_{q,K-1}^dagger
  (or G_qq Gamma_R G_qq^dagger for one block), A_L = i (G_qq - G_qq^dagger)
  - A_R, G^n = A_L f_L(e + E_ref) + A_R f_R(e + E_ref), G^p = A_L + A_R - G^n;
  each atom gains pre_factor * (-1 / pi) Tr G^n over its orbitals where
  e >= -potential_at_atom (electrons) and pre_factor * (1 / pi) Tr G^p
  otherwise (holes);
* ``subblock_index``     -> ``get_subblock_index`` ``:362-393``.

The reference's ``gammaR[-x1:, -x1:]`` with ``gr_lc`` requires the right
self-energy to fill the last block; here Gamma_R is padded into the last
block, which is the same product when it does.  Every self-energy is the
caller's (the reference recomputes or loads it per energy).
"""
import math

import numpy as np
import torch

from . import ozaki as ooz
from . import rgf as orgf


def integration_points(e_grid, integrate_way, n_gauss=None):
    e_grid = torch.as_tensor(e_grid, dtype=torch.float64)
    if integrate_way == "gauss":
        xs, w = ooz.gauss_xw(torch.scalar_tensor(float(e_grid[0]), dtype=torch.float64),
                             torch.scalar_tensor(float(e_grid[-1]), dtype=torch.float64),
                             n=n_gauss)
        return xs, w
    if integrate_way == "direct":
        dE = e_grid[1] - e_grid[0]
        return e_grid, dE * torch.ones(len(e_grid), dtype=torch.float64)
    raise ValueError("integrate_way only supports 'gauss' and 'direct' in this version")


def subblock_index(subblocks, atom_index, device_atom_norbs):
    """Blocks an atom's orbitals occupy, with its start in the first and its
    end in the last (``density.py:362-393``)."""
    cum = [0] + list(np.cumsum(subblocks))
    pre = sum(device_atom_norbs[:atom_index])
    last = pre + device_atom_norbs[atom_index]
    block_index = []
    orb_start = orb_end = None
    for i in range(len(cum) - 1):
        if pre >= cum[i] and last <= cum[i + 1]:
            block_index.append(i)
            orb_start, orb_end = pre - cum[i], last - cum[i]
            break
        elif pre >= cum[i] and pre < cum[i + 1] and last > cum[i + 1]:
            block_index.append(i)
            orb_start = pre - cum[i]
            for j in range(i + 1, len(cum) - 1):
                block_index.append(j)
                if last <= cum[j + 1]:
                    orb_end = last - cum[j]
                    break
    return block_index, orb_start, orb_end


def free_charge(points, weights, sigma_L, sigma_R, blocks, fermi_L, fermi_R, E_ref, eta_device,
                potential_at_atom, device_atom_norbs, subblocks, block_tridiagonal=True):
    """Per-atom free charge [natom] summed over the integration points.

    ``sigma_X(e)`` -> [a, a] self-energy at energy e (relative to E_ref);
    ``blocks`` = (hd, hu, hl, sd, su, sl) of the device (2-D tensors)."""
    hd, hu, hl, sd, su, sl = blocks
    natom = len(device_atom_norbs)
    charge = torch.zeros(natom, dtype=torch.complex128)
    for eidx, e in enumerate(points):
        e = float(e)
        sL, sR = sigma_L(e), sigma_R(e)
        ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), hl, hd, hu, sd, su, sl,
                                sL, sR, eta=eta_device, E_ref=E_ref, need_gr_lc=True)
        grd, gr_lc = ans[2], ans[1]
        tx, ty = ans[0].shape
        x1 = min(sR.shape[0], ty)
        gammaR = torch.zeros(ty, ty, dtype=torch.complex128)
        gammaR[-x1:, -x1:] += (1j * (sR - sR.mH))[-x1:, -x1:]
        if not block_tridiagonal:
            A_R = [g @ gammaR @ g.conj().T for g in grd]
        else:
            A_R = [c @ gammaR @ c.conj().T for c in gr_lc]
        A_L = [1j * (grd[i] - grd[i].conj().T) - A_R[i] for i in range(len(A_R))]
        ee = torch.tensor(e + E_ref, dtype=torch.float64)
        fl, fr = fermi_L(ee), fermi_R(ee)
        gnd = [A_L[i] * fl + A_R[i] * fr for i in range(len(A_L))]
        gpd = [A_L[i] + A_R[i] - gnd[i] for i in range(len(A_L))]
        for atom, Ei in enumerate(-1 * torch.as_tensor(potential_at_atom, dtype=torch.float64)):
            mats, sign = (gnd, -1.0) if e >= float(Ei) else (gpd, 1.0)
            pref = float(weights[eidx]) * 2 * sign / 2 / math.pi
            if not block_tridiagonal:
                pre = sum(device_atom_norbs[:atom])
                last = pre + device_atom_norbs[atom]
                charge[atom] += pref * torch.trace(mats[0][pre:last, pre:last])
                continue
            idx, o0, o1 = subblock_index(subblocks, atom, device_atom_norbs)
            if len(idx) == 1:
                charge[atom] += pref * torch.trace(mats[idx[0]][o0:o1, o0:o1])
            else:
                for b in idx:
                    if b == idx[0]:
                        charge[atom] += pref * torch.trace(mats[b][o0:, o0:])
                    elif b == idx[-1]:
                        charge[atom] += pref * torch.trace(mats[b][:o1, :o1])
                    else:
                        charge[atom] += pref * torch.trace(mats[b])
    return charge
