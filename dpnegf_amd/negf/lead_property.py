"""Lead self-energies on the GPU (``dpnegf.negf.lead_property`` interface).

:class:`LeadProperty` keeps the reference's constructor, attributes and
methods (``dpnegf/negf/lead_property.py:30-460``): ``self_energy`` /
``self_energy_cal`` compute Sigma with the batched Lopez-Sancho kernel (a 1-D
energy tensor gives [B, n, n] in one launch instead of the joblib pool of
``compute_all_self_energy``, ``:599-683``), ``HDL_reduced``,
``sigmaLR2Gamma``, ``fermi_dirac`` and ``gamma`` are unchanged.  Bloch-unfolded
leads (``useBloch``, ``:330-379``) take the small cell's surface Green's
functions at every unfolded point from one decimation launch
(:func:`bloch_self_energy`).  The HDF5 self-energy cache (``:932-984``) is not
part of this package.
"""
import numpy as np
import torch

from .. import _native as nat
from ._btd import C128, kvec as _kvec
from .bloch import Bloch
from .surface_green import self_energy_batched

_EV2J = 1.6021766208e-19
_BOLTZMANN = 1.380649e-23


class LeadProperty(object):
    def __init__(self, tab, hamiltonian, structure, results_path, voltage,
                 structure_leads_fold=None, bloch_sorted_indice=None, useBloch=False,
                 bloch_factor=[1, 1, 1], bloch_R_list=None, e_T=300, efermi=0.0, E_ref=None,
                 device="cuda"):
        self.hamiltonian = hamiltonian
        self.structure = structure
        self.tab = tab
        self.voltage = voltage
        self.results_path = results_path
        self.kBT = _BOLTZMANN * e_T / _EV2J
        self.e_T = e_T
        self.efermi = efermi
        self.E_ref = efermi if E_ref is None else E_ref
        self.chemiPot_lead = efermi - voltage
        self.kpoint = None
        self.voltage_old = None
        self.useBloch = useBloch
        self.bloch_factor = bloch_factor
        self.bloch_sorted_indice = bloch_sorted_indice
        self.bloch_R_list = bloch_R_list
        if self.useBloch:
            assert self.bloch_sorted_indice is not None
            assert self.bloch_R_list is not None
            assert self.bloch_factor is not None
        self.device = torch.device(device)

    def _lead_blocks(self, kpoint):
        kp = _kvec(kpoint)
        if (not hasattr(self, "HLk")) or self.voltage_old is None \
                or abs(self.voltage_old - self.voltage) > 1e-6 \
                or max(abs(self.kpoint - kp)) > 1e-6:
            blocks = self.hamiltonian.get_hs_lead(kpoint, tab=self.tab, v=self.voltage)
            self.HLk, self.HLLk, self.HDLk, self.SLk, self.SLLk, self.SDLk = [
                b.to(device=self.device, dtype=C128) for b in blocks]
            self.voltage_old = self.voltage
            self.kpoint = kp
        hdl, sdl = self.HDL_reduced(self.HDLk, self.SDLk, self._subblocks(kpoint))
        return (self.HLk, self.HLLk, hdl, self.SLk, self.SLLk, sdl)

    def self_energy(self, kpoint, energy, eta_lead=1e-5, method="Lopez-Sancho", save_path=None,
                    save_format="h5", se_info_display=False, HS_inmem=True):
        """Sets ``self.se``: [a, a] for a scalar energy, [B, a, a] for a 1-D one.

        Load-or-compute as the reference (``lead_property.py:105-158``): a
        self-energy stored by :func:`compute_all_self_energy` for this k-point
        and energy (the key of the reference's files, ``E_{e:.8f}`` and the
        k-point) is returned without a launch; otherwise it is computed."""
        assert len(np.array(kpoint).reshape(-1)) == 3
        hit = self._stored(kpoint, energy)
        if hit is not None:
            self.se = hit
            return
        self.se = self.self_energy_cal(kpoint, energy, eta_lead=eta_lead, method=method,
                                       HS_inmem=HS_inmem)

    def store_self_energies(self, kpoints, energies, eta_lead):
        """Self-energies of every (k, E) in one batched launch per k-point,
        kept on the device for :meth:`self_energy` to read (in place of the
        reference's per-(k, E) HDF5 records, ``lead_property.py:932-951``)."""
        E = torch.as_tensor(energies).reshape(-1)
        keys = [f"{float(e.real):.8f}" for e in E]
        if not hasattr(self, "_se_store"):
            self._se_store = {}
        for k in kpoints:
            sig = self.sigma_fn(k)(E, eta_lead)
            self._se_store[_k_key(k)] = ({key: i for i, key in enumerate(keys)}, sig)

    def _stored(self, kpoint, energy):
        store = getattr(self, "_se_store", None)
        if not store:
            return None
        ent = store.get(_k_key(kpoint))
        if ent is None:
            return None
        e = torch.as_tensor(energy)
        if e.ndim == 0:
            i = ent[0].get(f"{float(e.real):.8f}")
            return None if i is None else ent[1][i]
        idx = [ent[0].get(f"{float(x.real):.8f}") for x in e.reshape(-1)]
        if any(i is None for i in idx):
            return None
        return ent[1][torch.as_tensor(idx, device=ent[1].device)]

    def self_energy_cal(self, kpoint, energy, eta_lead=1e-5, method="Lopez-Sancho",
                        se_numba_jit=None, HS_inmem=True):
        e = torch.as_tensor(energy)
        if self.useBloch:
            sig = self._bloch_self_energy(kpoint, e.reshape(-1), eta_lead)
        else:
            lead = self._lead_blocks(kpoint)
            sig = self_energy_batched([lead], e.reshape(-1), eta_lead, E_ref=self.E_ref).sigma
        self.se = sig[0] if e.ndim == 0 else sig
        if not HS_inmem:
            del self.HLk, self.HLLk, self.HDLk, self.SLk, self.SLLk, self.SDLk
        return self.se

    def sigma_fn(self, kpoint):
        """``f(e, eta)`` -> self-energies [B, a, a] at the energies ``e`` ([B],
        complex allowed, relative to ``E_ref``) with decimation broadening
        ``eta``, through the same route as ``self_energy_cal`` (Bloch-unfolded
        when ``useBloch``).  Leaves ``self.se`` unchanged."""
        def f(e, eta):
            e = torch.as_tensor(e).reshape(-1)
            if self.useBloch:
                return self._bloch_self_energy(kpoint, e, eta)
            lead = self._lead_blocks(kpoint)
            return self_energy_batched([lead], e, eta, E_ref=self.E_ref).sigma
        return f

    def _subblocks(self, kpoint):
        if hasattr(self.hamiltonian, "get_hs_device"):
            try:
                return self.hamiltonian.get_hs_device(kpoint, only_subblocks=True)
            except TypeError:
                return None
        return None

    def _bloch_self_energy(self, kpoint, e, eta_lead):
        """``lead_property.py:330-378``: the small cell's blocks at every point
        that folds onto ``kpoint``; the device-lead coupling is the one
        ``get_hs_lead`` returns with them (the supercell's, at ``kpoint``)."""
        self.kpoint = _kvec(kpoint)
        self.voltage_old = self.voltage
        kb = Bloch(self.bloch_factor).unfold_points(list(self.kpoint.numpy()))
        small = []
        for k in kb:
            blocks = self.hamiltonian.get_hs_lead(torch.tensor(k), tab=self.tab, v=self.voltage)
            self.HLk, self.HLLk, self.HDLk, self.SLk, self.SLLk, self.SDLk = [
                b.to(device=self.device, dtype=C128) for b in blocks]
            small.append((self.HLk, self.HLLk, self.SLk, self.SLLk))
        hdl, sdl = self.HDL_reduced(self.HDLk, self.SDLk, self._subblocks(kpoint))
        return bloch_self_energy(small, kb, self.bloch_R_list, self.bloch_sorted_indice, hdl, sdl,
                                 e, eta_lead, self.E_ref)

    @staticmethod
    def HDL_reduced(HDL, SDL, subblocks):
        """``lead_property.py:386-434``."""
        assert len(HDL.shape) == 2 and len(SDL.shape) == 2 and HDL.shape == SDL.shape
        nz = HDL.nonzero()
        lo, hi = nz.min(dim=0).values, nz.max(dim=0).values
        cut = (lo, hi) if subblocks is None else ((subblocks[-1], subblocks[-1]),
                                                   (subblocks[0], subblocks[0]))
        if HDL.shape[0] == 1:
            return HDL, SDL
        if lo[0] > 0:
            r = int(cut[0][0]) if subblocks is None else -1 * int(cut[0][0])
            return HDL[r:, :], SDL[r:, :]
        r = int(cut[1][0]) + 1 if subblocks is None else int(cut[1][0])
        return HDL[:r, :], SDL[:r, :]

    def sigmaLR2Gamma(self, se):
        return 1j * (se - se.mH)

    def fermi_dirac(self, x):
        return 1 / (1 + torch.exp((x - self.chemiPot_lead) / self.kBT))

    @property
    def gamma(self):
        return self.sigmaLR2Gamma(self.se)


def _k_key(k):
    """A k-point as a hashable key (the reference keys its records the same way)."""
    a = np.asarray(torch.as_tensor(k).detach().cpu().numpy() if torch.is_tensor(k) else k,
                   dtype=float).reshape(3)
    return (float(a[0]), float(a[1]), float(a[2]))


def compute_all_self_energy(eta, lead_L, lead_R, kpoints_grid, energy_grid,
                            self_energy_save_path=None, n_jobs=-1, batch_size=200, n_cpus=None,
                            se_numba_jit=None):
    """``lead_property.py:599-684`` on the GPU: the self-energies of both
    leads at every (k, E) of the grid, one batched decimation launch per lead
    and k-point instead of a loky pool over (k, E) tasks, kept in device
    memory on the lead objects instead of HDF5 files.  The runner's loop then
    reads them through ``LeadProperty.self_energy`` as it reads the files
    (``NEGF.py:762-774``).  ``self_energy_save_path``, ``n_jobs``,
    ``batch_size``, ``n_cpus`` and ``se_numba_jit`` are accepted for the
    reference's signature and not needed."""
    for lead in (lead_L, lead_R):
        lead.store_self_energies(kpoints_grid, energy_grid, eta)


def bloch_self_energy(small, k_bloch, R_list, sorted_idx, hDL, sDL, energies, eta_lead,
                      E_ref=0.0):
    """Self-energies [P, a, a] of a lead whose principal layer is m copies of a
    small cell (``lead_property.py:336-378``).

    ``small``: per unfolded point k' (``k_bloch`` [nk, 3]) the small cell's
    ``(hL, hLL, sL, sLL)`` on the device; ``R_list``: the m cell offsets;
    ``sorted_idx``: supercell orbital order in terms of kron(cell, orbital);
    ``hDL``/``sDL`` [a, b]: device-lead coupling (rows already reduced).
    g_s of every (E, k') comes from one decimation launch; then
    g = mean_k' kron(P(k'), g_s(k')) with P_ij = exp(2 pi i (R_j - R_i) . k'),
    and Sigma = (e S_DL - H_DL) g[:b, :b] (e S_DL^H - H_DL^H), e = E + E_ref
    real-axis (no eta) as in the reference."""
    dev = hDL.device
    e = torch.as_tensor(energies).to(device=dev, dtype=C128).reshape(-1)
    P, nk = e.shape[0], len(small)
    e_all = e.repeat(nk).contiguous()  # point index k' * P + p
    sys_idx = torch.arange(nk, dtype=torch.int32, device=dev).repeat_interleave(P).contiguous()
    leads = [(hL, hLL, None, sL, sLL, None) for (hL, hLL, sL, sLL) in small]
    gs = self_energy_batched(leads, e_all, eta_lead, sys_idx=sys_idx if nk > 1 else None,
                             E_ref=E_ref).gs
    n = gs.shape[-1]
    m = len(R_list)
    R = torch.stack([torch.as_tensor(r, dtype=torch.float64) for r in R_list]).to(dev)
    kb = torch.as_tensor(np.asarray(k_bloch, dtype=float), dtype=torch.float64, device=dev)
    # phase[k', i, j] = exp(2 pi i (R_j - R_i) . k')
    arg = (R[None, None, :, :] - R[None, :, None, :]).mul(kb[:, None, None, :]).sum(-1)
    phase = torch.polar(torch.ones_like(arg), 2 * np.pi * arg)
    gs = gs.view(nk, P, n, n)
    g = torch.zeros(P, m, n, m, n, dtype=C128, device=dev)
    for q in range(nk):
        g += phase[q][None, :, None, :, None] * gs[q][:, None, :, None, :]
    g = (g / nk).reshape(P, m * n, m * n)
    idx = torch.as_tensor(sorted_idx, device=dev)
    b = hDL.shape[1]
    gb = g[:, idx[:b]][:, :, idx[:b]].contiguous()
    a = hDL.shape[0]
    ee = (e + E_ref).view(P, 1, 1)
    A = (ee * sDL - hDL).contiguous()                       # [P, a, b]
    C = (ee * sDL.conj().T - hDL.conj().T).contiguous()     # [P, b, a]
    X = torch.empty(P, a, b, dtype=C128, device=dev)
    sig = torch.empty(P, a, a, dtype=C128, device=dev)
    st = nat.stream_ptr(dev)
    nat.call("dpnegf_zgemm_batched", st, P, a, b, b, nat.ptr(A), b, a * b, 0, nat.ptr(gb), b,
             b * b, 0, 1.0, 0.0, 0.0, 0.0, nat.ptr(X), b, a * b)
    nat.call("dpnegf_zgemm_batched", st, P, a, a, b, nat.ptr(X), b, a * b, 0, nat.ptr(C), a,
             b * a, 0, 1.0, 0.0, 0.0, 0.0, nat.ptr(sig), a, a * a)
    return sig
