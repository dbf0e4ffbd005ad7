"""Ozaki pole-sum density on the GPU (``dpnegf.negf.density.Ozaki`` interface).

``Ozaki.integrate`` follows ``dpnegf/negf/density.py:143-199``:

  DM_eq  = Re[ iR G(iR - mu) ] + sum_p 4 kT R_p Re G(z_p),   z_p = i zeta_p kT + mu_L - mu
  DM_neq = sum_i w_i Re[G Gamma_R G^dagger](x_i) (f_R(x_i + mu) - f_L(x_i + mu))

with x_i Gauss-Legendre nodes on [min V - 8kT, max V + 8kT].  All poles (and
all quadrature nodes) are one batch: one Lopez-Sancho launch per lead, one
block-RGF call, one weighted reduction per block (``dpnegf_weighted_re_sum``)
-- instead of the reference's Python loop with a B=1 RGF per pole.

The reference computes block 0 only and reads ``lead_L.mu`` / ``mu``
attributes that do not exist at HEAD; here every diagonal block is computed
(``all_blocks=True`` returns them all), the lead chemical potential is
``chemiPot_lead`` (``lead_property.py:89``) and the device one defaults to
efermi - (V_L + V_R)/2 (the expression of ``device_property.py:134``).  The
non-equilibrium term uses G_{q,K-1} Gamma_R G_{q,K-1}^dagger, which is the
reference's G_00 Gamma_R G_00^dagger for a single-block device.
"""
import math

import numpy as np
import torch

from .. import _native as nat
from ._btd import C128, block_sizes, corner_self_energy, is_orthogonal, pack_arena, split_blocks
from ._ops import diag_of_product, mm
from .ozaki_res_cal import ozaki_residues
from .recursive_green_cal import _run_full, full_points_per_call
from .surface_green import self_energy_batched


def gauss_xw(xl, xu, n=100):
    """Gauss-Legendre nodes and weights on [xl, xu] (``negf_utils.py:136-159``)."""
    ndim = len(xu.shape)
    x, w = np.polynomial.legendre.leggauss(n)
    x = torch.tensor(x, dtype=xu.dtype, device=xu.device)[(...,) + (None,) * ndim]
    w = torch.tensor(w, dtype=xu.dtype, device=xu.device)[(...,) + (None,) * ndim]
    w = w * (0.5 * (xu - xl))
    return x * (0.5 * (xu - xl)) + 0.5 * (xu + xl), w


def _lead_sigma(lead, e, eta, E_ref):
    """Self-energies [B, a, a] at the energies ``e`` (relative to ``E_ref``):
    ``lead`` is either a tuple (hL, hLL, hDL, sL, sLL, sDL) or a callable
    ``lead(e, eta)`` (a lead object's own route, which applies Bloch
    unfolding when the lead uses it)."""
    if callable(lead):
        return lead(e, eta)
    return self_energy_batched([lead], e, eta, E_ref=E_ref).sigma


def _sigmas(lead_L, lead_R, e, eta, E_ref, nblk0, nblkK):
    sL = _lead_sigma(lead_L, e, eta, E_ref)
    sR = _lead_sigma(lead_R, e, eta, E_ref)
    B = e.shape[0]
    sL, nL = corner_self_energy(sL, nblk0, B, False, e.device)
    sR, nR = corner_self_energy(sR, nblkK, B, True, e.device)
    return sL, nL, sR, nR


def _weighted_blocks(arena, B, shapes, coef, acc=None):
    """[sum_b Re(c_b X_b)] for every block of a block-major arena, added to
    the blocks of ``acc`` when given."""
    out = []
    off = 0
    for q, (m, n) in enumerate(shapes):
        blk = arena[off * B:(off + m * n) * B]
        res = acc[q] if acc is not None else torch.empty(m, n, dtype=torch.float64,
                                                         device=arena.device)
        nat.call("dpnegf_weighted_re_sum", nat.stream_ptr(arena.device), nat.ptr(res),
                 nat.ptr(blk), nat.ptr(coef), m * n, B, int(acc is not None))
        out.append(res)
        off += m * n
    return out


def ozaki_points(sizes, lead_L, lead_R, kBT, mu_L, mu, R, poles, residues, E_ref, eta_lead,
                 eta_device, vL=0.0, vR=0.0, n_gauss=10, fermi_L=None, fermi_R=None,
                 eta_lead0=1e-5, eta_device0=0.0, device=None):
    """The energy points of the Ozaki density (``density.py:164-195``) with
    their self-energies and weights.  Returns ``(eq, neq)``:

    * ``eq`` = (z [M+1], sigL, nL, sigR, nR, coef [M+1]): the zeroth moment
      at iR - mu (broadening eta_lead0 / eta_device0) and the poles, with
      DM_eq = sum_b Re(coef_b G_qq(z_b));
    * ``neq`` = (z [n_gauss], sigL, nL, sigR, nR, gammaR [n_gauss, nK, nK],
      a [n_gauss]) for a bias window (None at zero bias), with
      DM_neq = sum_i a_i Re[G_{q,K-1} Gamma_R G_{q,K-1}^dagger](x_i)."""
    dev = device
    poles = torch.as_tensor(poles, dtype=torch.float64).to(dev)
    residues = torch.as_tensor(residues, dtype=torch.float64).to(dev)
    e0 = torch.tensor([1j * R - mu], dtype=C128, device=dev)
    ep = (1j * poles * kBT + (mu_L - mu)).to(C128)
    s0 = _sigmas(lead_L, lead_R, e0, eta_lead0, E_ref, sizes[0], sizes[-1])
    sp = _sigmas(lead_L, lead_R, ep, eta_lead, E_ref, sizes[0], sizes[-1])
    sL = torch.cat([s0[0], sp[0]]).contiguous()
    sR = torch.cat([s0[2], sp[2]]).contiguous()
    z = torch.cat([e0 + E_ref + 1j * eta_device0, ep + E_ref + 1j * eta_device]).contiguous()
    coef = torch.cat([torch.tensor([1j * R], dtype=C128, device=dev),
                      (4.0 * kBT * residues).to(C128)]).contiguous()
    eq = (z, sL, s0[1], sR, s0[3], coef)
    if abs(vL - vR) <= 1e-14:
        return eq, None
    xl, xu = min(vL, vR) - 8 * kBT, max(vL, vR) + 8 * kBT
    xs, wl = gauss_xw(torch.scalar_tensor(xl, dtype=torch.float64),
                      torch.scalar_tensor(xu, dtype=torch.float64), n=n_gauss)
    ex = xs.to(dev).to(C128)
    G = ex.shape[0]
    sLx, nL, sRx, nR = _sigmas(lead_L, lead_R, ex, eta_lead, E_ref, sizes[0], sizes[-1])
    zx = (ex + E_ref + 1j * eta_device).contiguous()
    nK = sizes[-1]
    gam = 1j * (sRx - sRx.mH)
    gamR = torch.zeros(G, nK, nK, dtype=C128, device=dev)
    gamR[:, nK - nR:, nK - nR:] = gam
    xsd = xs.to(dev)
    f = fermi_R(xsd + mu) - fermi_L(xsd + mu)
    a = (wl.to(dev) * f).to(torch.float64).contiguous()
    return eq, (zx, sLx, nL, sRx, nR, gamR.contiguous(), a)


def eq_density(arena, sizes, ortho, eq):
    """DM_eq blocks of the equilibrium point set (see :func:`ozaki_points`);
    all points in one call unless their outputs would not fit (blocks of
    several hundred orbitals): then in chunks, summed in order."""
    z, sL, nL, sR, nR, coef = eq
    P = z.shape[0]
    dm = None
    chunk = full_points_per_call(sizes, False, P, z.device)
    for c0 in range(0, P, chunk):
        c = slice(c0, min(P, c0 + chunk))
        raw = _run_full(z[c].contiguous(), arena, sizes, ortho, sL[c].contiguous(), nL,
                        sR[c].contiguous(), nR, False)
        dm = _weighted_blocks(raw["grd"], z[c].shape[0], [(n, n) for n in sizes],
                              coef[c].contiguous(), dm)
        del raw
    return dm


def neq_density(arena, sizes, ortho, neq):
    """DM_neq blocks of the bias-window point set (see :func:`ozaki_points`)."""
    zx, sLx, nL, sRx, nR, gamR, a = neq
    dev = zx.device
    G = zx.shape[0]
    nK = sizes[-1]
    coef = a.to(C128).contiguous()
    dm = [None] * len(sizes)
    chunk = full_points_per_call(sizes, True, G, dev)
    for c0 in range(0, G, chunk):
        c = slice(c0, min(G, c0 + chunk))
        Gc = zx[c].shape[0]
        raw = _run_full(zx[c].contiguous(), arena, sizes, ortho, sLx[c].contiguous(), nL,
                        sRx[c].contiguous(), nR, True)
        lc = split_blocks(raw["lc"], Gc, [(n, nK) for n in sizes])
        gR = gamR[c].contiguous()
        cc = coef[c].contiguous()
        for q, n in enumerate(sizes):
            X = torch.empty(Gc, n, nK, dtype=C128, device=dev)
            Y = torch.empty(Gc, n, n, dtype=C128, device=dev)
            lcq = lc[q].contiguous()
            st = nat.stream_ptr(dev)
            nat.call("dpnegf_zgemm_batched", st, Gc, n, nK, nK, nat.ptr(lcq), nK, n * nK, 0,
                     nat.ptr(gR), nK, nK * nK, 0, 1.0, 0.0, 0.0, 0.0, nat.ptr(X), nK, n * nK)
            nat.call("dpnegf_zgemm_batched", st, Gc, n, n, nK, nat.ptr(X), nK, n * nK, 0,
                     nat.ptr(lcq), nK, n * nK, 1, 1.0, 0.0, 0.0, 0.0, nat.ptr(Y), n, n * n)
            acc = [dm[q]] if dm[q] is not None else None
            dm[q] = _weighted_blocks(Y.reshape(-1), Gc, [(n, n)], cc, acc)[0]
        del raw
    return dm


def ozaki_density(blocks, lead_L, lead_R, kBT, mu_L, mu, R, poles, residues, E_ref, eta_lead,
                  eta_device, vL=0.0, vR=0.0, n_gauss=10, fermi_L=None, fermi_R=None,
                  eta_lead0=1e-5, eta_device0=0.0):
    """Density matrices of every diagonal block of one k-point.

    ``blocks`` = (hd, hu, hl, sd, su, sl) CUDA tensors; ``lead_X`` = lead tuples
    (hL, hLL, hDL, sL, sLL, sDL), or callables ``lead_X(e, eta)`` returning the
    self-energies [B, a, a] at energies ``e``.  Energies are relative to ``E_ref``.
    Returns (DM_eq list, DM_neq list or 0.)."""
    hd, hu, hl, sd, su, sl = blocks
    dev = hd[0].device
    nat.require_device(hd[0])
    sizes = block_sizes(hd)
    arena = pack_arena(hd, hu, hl, sd, su, sl, dev)
    ortho = is_orthogonal(sd, su, sl)
    eq, neq = ozaki_points(sizes, lead_L, lead_R, kBT, mu_L, mu, R, poles, residues, E_ref,
                           eta_lead, eta_device, vL, vR, n_gauss, fermi_L, fermi_R, eta_lead0,
                           eta_device0, dev)
    dm_eq = eq_density(arena, sizes, ortho, eq)
    if neq is None:
        return dm_eq, 0.
    return dm_eq, neq_density(arena, sizes, ortho, neq)


class Ozaki(object):
    """``dpnegf/negf/density.py:107-199`` on the GPU."""

    def __init__(self, R, M_cut, n_gauss):
        self.poles, self.residues = ozaki_residues(M_cut)
        self.R = R
        self.n_gauss = n_gauss

    def integrate(self, deviceprop, kpoint, eta_lead=1e-5, eta_device=0., Vbias=None,
                  block_tridiagonal=False, all_blocks=False):
        kBT = deviceprop.kBT
        lL, lR = deviceprop.lead_L, deviceprop.lead_R
        mu_L = getattr(lL, "mu", lL.chemiPot_lead)
        mu = getattr(deviceprop, "mu", None)
        if mu is None:
            mu = deviceprop.efermi - 0.5 * (lL.voltage + lR.voltage)
        V = Vbias if Vbias is not None else torch.tensor(0.)
        hs = deviceprop.hamiltonian.get_hs_device(kpoint, V, block_tridiagonal)
        dev = deviceprop.rgf_device
        hd, sd, hl, su, sl, hu = [[b.to(dev) for b in blk] for blk in hs]
        # the leads' own batched route (Bloch-unfolded when they use it,
        # lead_property.py:330-379), as the reference reaches them through
        # lead.self_energy (density.py:167-177)
        dm_eq, dm_neq = ozaki_density(
            (hd, hu, hl, sd, su, sl), lL.sigma_fn(kpoint), lR.sigma_fn(kpoint), kBT, mu_L,
            mu, self.R, self.poles, self.residues, deviceprop.E_ref, eta_lead, eta_device,
            lL.voltage, lR.voltage, self.n_gauss, lL.fermi_dirac, lR.fermi_dirac)
        if all_blocks:
            return dm_eq, dm_neq
        return dm_eq[0], (dm_neq[0] if isinstance(dm_neq, list) else dm_neq)

    def get_density_onsite(self, deviceprop, DM):
        """``density.py:201-230``."""
        if len(DM.shape) == 2:
            DM = DM.diag()
        norbs = np.cumsum([0] + list(deviceprop.norbs_per_atom))
        onsite = torch.stack([DM[norbs[i]:norbs[i + 1]].sum() for i in range(len(norbs) - 1)])
        pos = torch.as_tensor(np.asarray(deviceprop.structure.positions)).to(onsite)
        return torch.cat([pos, onsite.unsqueeze(-1)], dim=-1)


def _lead_sigmas(lead, kpoint, E, eta):
    """Self-energies [B, a, a] of ``lead`` at the energies ``E``: values stored
    by ``compute_all_self_energy`` when present (the reference's Dirichlet-lead
    branch loads them), else one batched decimation."""
    stored = lead._stored(kpoint, E) if hasattr(lead, "_stored") else None
    if stored is not None:
        return stored
    return lead.sigma_fn(kpoint)(E, eta)


def fiori_free_charge(blocks, sigL, sigR, E, pre_factor, fL, fR, E_ref, eta_device,
                      potential_at_atom, device_atom_norbs):
    """Per-atom free charge [natom] (float64) of ``density.py:268-360`` over
    the integration points ``E`` [B] (relative to ``E_ref``) with weights
    ``pre_factor`` [B] and lead occupations ``fL``, ``fR`` [B] at E + E_ref.

    Per point and orbital o: a_D = diag i(G - G^dagger) = -2 Im G_oo,
    a_R = diag(G_{q,K-1} Gamma_R G_{q,K-1}^dagger) (G_qq for one block),
    G^n_oo = (a_D - a_R) f_L + a_R f_R, G^p_oo = a_D - G^n_oo; atom j gains
    -pre/pi sum_o G^n_oo where E >= -potential_at_atom[j] (electrons) and
    +pre/pi sum_o G^p_oo otherwise (holes).  Orbitals are counted in device
    order, so an atom straddling block edges sums over its blocks as the
    reference's ``get_subblock_index`` does."""
    hd, hu, hl, sd, su, sl = blocks
    dev = hd[0].device
    sizes = block_sizes(hd)
    K = len(sizes)
    nK = sizes[-1]
    arena = pack_arena(hd, hu, hl, sd, su, sl, dev)
    ortho = is_orthogonal(sd, su, sl)
    E = torch.as_tensor(E).to(device=dev, dtype=torch.float64).reshape(-1)
    B = E.shape[0]
    sL, nL = corner_self_energy(sigL, sizes[0], B, False, dev)
    sR, nR = corner_self_energy(sigR, nK, B, True, dev)
    gamR = torch.zeros(B, nK, nK, dtype=C128, device=dev)
    gamR[:, nK - nR:, nK - nR:] = 1j * (sR - sR.mH)
    z = (E.to(C128) + E_ref + 1j * eta_device).contiguous()
    aD, aR = [], []
    chunk = full_points_per_call(sizes, True, B, dev)
    for c0 in range(0, B, chunk):
        c = slice(c0, min(B, c0 + chunk))
        Bc = z[c].shape[0]
        raw = _run_full(z[c].contiguous(), arena, sizes, ortho, sL[c].contiguous(), nL,
                        sR[c].contiguous(), nR, True)
        grd = split_blocks(raw["grd"], Bc, [(n, n) for n in sizes])
        lc = split_blocks(raw["lc"], Bc, [(n, nK) for n in sizes])
        gR = gamR[c].contiguous()
        dq, rq = [], []
        for q in range(K):
            dq.append(-2.0 * torch.diagonal(grd[q], dim1=-2, dim2=-1).imag)
            X = mm(lc[q], gR)
            rq.append(diag_of_product([(X, lc[q].mH.contiguous())], Bc, dev).real)
        aD.append(torch.cat(dq, dim=-1))
        aR.append(torch.cat(rq, dim=-1))
        del raw
    aD, aR = torch.cat(aD), torch.cat(aR)
    fL = torch.as_tensor(fL).to(device=dev, dtype=torch.float64).reshape(B, 1)
    fR = torch.as_tensor(fR).to(device=dev, dtype=torch.float64).reshape(B, 1)
    gn = (aD - aR) * fL + aR * fR
    gp = aD - gn
    norbs = [int(x) for x in device_atom_norbs]
    atom = torch.repeat_interleave(torch.arange(len(norbs), device=dev),
                                   torch.as_tensor(norbs, device=dev))
    onehot = torch.nn.functional.one_hot(atom, len(norbs)).to(torch.float64)
    Qn, Qp = gn @ onehot, gp @ onehot
    Ei = -torch.as_tensor(potential_at_atom).to(device=dev, dtype=torch.float64).reshape(1, -1)
    electron = E.reshape(B, 1) >= Ei
    pre = torch.as_tensor(pre_factor).to(device=dev, dtype=torch.float64).reshape(B, 1)
    return (torch.where(electron, -Qn, Qp) * pre).sum(0) / math.pi


class Fiori(object):
    """``dpnegf/negf/density.py:234-393`` on the GPU: the free charge of the
    Poisson-NEGF loop from a real-axis integral, every integration energy in
    one batch (one decimation launch per lead, full sweeps with the last block
    column, per-orbital spectral diagonals) instead of the reference's Python
    loop with one RGF per energy."""

    def __init__(self, n_gauss=None, integrate_way="direct", e_grid=None):
        self.n_gauss = n_gauss
        self.e_grid = e_grid
        self.integrate_way = integrate_way
        self.xs = None
        self.wlg = None
        self.e_grid_Fiori = None
        if integrate_way == "gauss":
            assert self.n_gauss is not None, \
                "n_gauss must be set in the Fiori class with gauss integration"
            self.xs, self.wlg = gauss_xw(torch.scalar_tensor(float(e_grid[0]), dtype=torch.float64),
                                         torch.scalar_tensor(float(e_grid[-1]),
                                                             dtype=torch.float64),
                                         n=self.n_gauss)
            self.e_grid_Fiori = e_grid
            integrate_range, pre_factor = self.xs, self.wlg
        elif integrate_way == "direct":
            e_grid = torch.as_tensor(e_grid, dtype=torch.float64)
            dE = e_grid[1] - e_grid[0]
            integrate_range = e_grid
            pre_factor = dE * torch.ones(len(e_grid), dtype=torch.float64)
        else:
            raise ValueError("integrate_way only supports 'gauss' and 'direct' in this version")
        self.integrate_range = integrate_range
        self.pre_factor = pre_factor

    def density_integrate_Fiori(self, e_grid, kpoint, Vbias, block_tridiagonal, subblocks,
                                integrate_way, deviceprop, device_atom_norbs, potential_at_atom,
                                with_Dirichlet_leads, free_charge, E_ref, eta_lead=1e-5,
                                eta_device=1e-5, self_energy_save_path=None):
        """Adds each atom's free charge to ``free_charge[str(kpoint)]`` as the
        reference does (``density.py:268-360``).  Self-energies stored by
        ``compute_all_self_energy`` are used when present (the reference loads
        them for Dirichlet leads), otherwise computed in one batch."""
        E = torch.as_tensor(self.integrate_range, dtype=torch.float64)
        dev = deviceprop.rgf_device
        V = Vbias if Vbias is not None else torch.tensor(0.)
        hs = deviceprop.hamiltonian.get_hs_device(kpoint, V, block_tridiagonal)
        hd, sd, hl, su, sl, hu = [[b.to(dev) for b in blk] for blk in hs]
        lL, lR = deviceprop.lead_L, deviceprop.lead_R
        sigL = _lead_sigmas(lL, kpoint, E, eta_lead)
        sigR = _lead_sigmas(lR, kpoint, E, eta_lead)
        Ed = E.to(dev)
        charge = fiori_free_charge((hd, hu, hl, sd, su, sl), sigL, sigR, Ed, self.pre_factor,
                                   lL.fermi_dirac(Ed + E_ref), lR.fermi_dirac(Ed + E_ref), E_ref,
                                   eta_device, potential_at_atom, device_atom_norbs)
        key = str(kpoint)
        cur = free_charge[key]
        free_charge[key] = cur + charge.to(device=cur.device, dtype=cur.dtype) \
            if torch.is_tensor(cur) else cur + charge.cpu().numpy()
        return free_charge

    def get_subblock_index(self, subblocks, atom_index, device_atom_norbs):
        """Blocks an atom's orbitals occupy, with its start in the first block
        and its end in the last (``density.py:362-393``)."""
        cum = [0] + list(np.cumsum(subblocks))
        pre = sum(device_atom_norbs[:atom_index])
        last = pre + device_atom_norbs[atom_index]
        block_index, orb_start, orb_end = [], None, None
        for i in range(len(cum) - 1):
            if cum[i] <= pre and last <= cum[i + 1]:
                return [i], pre - cum[i], last - cum[i]
            if cum[i] <= pre < cum[i + 1] < last:
                block_index.append(i)
                orb_start = pre - cum[i]
                for j in range(i + 1, len(cum) - 1):
                    block_index.append(j)
                    if last <= cum[j + 1]:
                        orb_end = last - cum[j]
                        break
        return block_index, orb_start, orb_end
