"""Device-side observables (oracle only), restating
``dpnegf/negf/device_property.py``:

* ``build_s_in``    -> ``_build_s_in_batched`` ``:21-31`` and ``:259-272``
* ``transmission``  -> ``_cal_tc_`` ``:374-406``
* ``dos`` / ``ldos``-> ``_cal_dos_`` ``:408-439`` / ``_cal_ldos_`` ``:441-470``
* ``local_current`` -> ``_cal_local_current_`` ``:472-503``
* ``current_nscf``  -> ``_cal_current_nscf_`` ``:333-372``
"""
import math

import numpy as np
import torch


def _corner_pad(full, shape, at_end):
    """Zero matrix of ``shape`` with ``full``'s leading (or trailing) corner."""
    out = torch.zeros(shape, dtype=torch.complex128, device=full.device)
    r = min(full.shape[-2], shape[-2])
    if at_end:
        out[..., -r:, -r:] = full[..., -r:, -r:]
    else:
        out[..., :r, :r] = full[..., :r, :r]
    return out


def transmission(g_trans, gammaL, gammaR):
    """T = Re Tr[Gamma_L G_{0,K-1} Gamma_R G_{0,K-1}^dagger] (``:383-404``)."""
    tx, ty = g_trans.shape[-2], g_trans.shape[-1]
    lead = g_trans.shape[:-2]
    gl = _corner_pad(gammaL, lead + (tx, tx), at_end=False)
    gr = _corner_pad(gammaR, lead + (ty, ty), at_end=True)
    return (gl @ g_trans @ gr @ g_trans.mH).diagonal(dim1=-2, dim2=-1).real.sum(-1)


def _block_spectral(grd, grl, gru, sd, su, sl, block_tridiagonal=True):
    out = []
    K = len(grd)
    for j in range(K):
        if not block_tridiagonal or len(gru) == 0:
            t = grd[j] @ sd[j]
        elif j == 0:
            t = grd[j] @ sd[j] + gru[j] @ sl[j]
        elif j == K - 1:
            t = grd[j] @ sd[j] + grl[j - 1] @ su[j - 1]
        else:
            t = grd[j] @ sd[j] + grl[j - 1] @ su[j - 1] + gru[j] @ sl[j]
        out.append(t)
    return out


def dos(grd, grl, gru, sd, su, sl, block_tridiagonal=True):
    """Total DOS with spin factor 2 (``:408-439``)."""
    val = 0
    for t in _block_spectral(grd, grl, gru, sd, su, sl, block_tridiagonal):
        val = val - t.imag.diagonal(dim1=-2, dim2=-1).sum(-1) / math.pi
    return val * 2


def ldos(grd, grl, gru, sd, su, sl, norbs_per_atom, block_tridiagonal=True):
    """Per-atom LDOS with spin factor 2 (``:441-470``)."""
    parts = [-t.imag.diagonal(dim1=-2, dim2=-1) / math.pi
             for t in _block_spectral(grd, grl, gru, sd, su, sl, block_tridiagonal)]
    orb = torch.cat(parts, dim=-1)
    acc = np.cumsum([0] + list(norbs_per_atom))
    return torch.stack([orb[..., acc[i]:acc[i + 1]].sum(-1)
                        for i in range(len(acc) - 1)], dim=-1) * 2


def build_s_in(hd_shapes, seinL, seinR):
    """Per-block in-scattering with the lead terms in the corners (``:21-31``)."""
    B = seinL.shape[0]
    s_in = [torch.zeros((B,) + tuple(s), dtype=torch.complex128) for s in hd_shapes]
    r0 = min(hd_shapes[0][0], seinL.shape[-1])
    r1 = min(hd_shapes[-1][0], seinR.shape[-1])
    s_in[0][:, :r0, :r0] += seinL[:, :r0, :r0]
    s_in[-1][:, -r1:, -r1:] += seinR[:, -r1:, -r1:]
    return s_in


def local_current(hd, gnd, norbs_per_atom):
    """Bond currents between atoms from the whole-device H and lesser G^n
    (``device_property.py:472-503``), atom by atom as the reference loops:
    J_ij = Re Tr[H_ij (i G^n_ji)] - Re Tr[H_ji (i G^n_ij)], J_ii = 0."""
    na = len(norbs_per_atom)
    acc = np.cumsum([0] + list(norbs_per_atom))
    J = torch.zeros(na, na, dtype=torch.float64)
    for i in range(na):
        for j in range(na):
            if i != j:
                a0, a1, b0, b1 = acc[i], acc[i + 1], acc[j], acc[j + 1]
                ki = hd[a0:a1, b0:b1] @ (1j * gnd[b0:b1, a0:a1])
                kj = hd[b0:b1, a0:a1] @ (1j * gnd[a0:a1, b0:b1])
                J[i, j] = ki.real.diag().sum() - kj.real.diag().sum()
    return J


def current_nscf(energy_grid, tc, efermi, kbt):
    """Non-self-consistent I-V from the zero-bias T(E) (``:333-372``): biases
    vv = 2 dv on a 0.1 V step up to the window the grid allows, each current
    Simpson's rule over (f(E + E_F; E_F - vm + dv) - f(E + E_F; E_F - vm - dv)) T."""
    from scipy.integrate import simpson
    f = lambda x, mu: 1 / (1 + torch.exp((x - mu) / kbt))
    emin, emax = energy_grid.min(), energy_grid.max()
    vmin, vmax = emin + 4 * kbt, emax - 4 * kbt
    vm = 0.5 * (vmin + vmax)
    vmid = vm - vmin
    vv = torch.linspace(start=0., end=float(vmid), steps=int(vmid / 0.1) + 1) * 2
    cc = []
    for dv in vv * 0.5:
        y = (f(energy_grid + efermi, efermi - vm + dv) - f(energy_grid + efermi, efermi - vm - dv)) * tc
        cc.append(simpson(y=y.numpy(), x=energy_grid.numpy()))
    return vv, cc
