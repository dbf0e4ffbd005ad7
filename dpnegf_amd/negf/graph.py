"""T(E, k) over a fixed grid as one captured HIP graph.

An SCF loop re-evaluates the same NEGF work every iteration with only the
Hamiltonian values changed (the potential enters through ``get_hs_device`` /
``get_hs_lead``, ``negf_hamiltonian_init.py:857-974``).  ``TransmissionGraph``
packs the systems, leads and energy grid once, captures the three launches of
the transmission path -- Sigma_L and Sigma_R (Lopez-Sancho, one launch per
lead over all points) and the fused RGF transmission sweep -- and replays them
with no host work.  ``update`` rewrites the packed blocks in place between
replays.  Replays and updates run on the caller's current stream.

Any block and lead size the library takes can be captured.  Blocks above 64
use the block-operation sweep (one call per k-point, the points of each
k-point gathered and scattered by precomputed index lists); leads above 48
orbitals use the batched decimation, which under capture launches its full
101 iterations with converged points frozen instead of asking the host when
to stop (at most 65535 points per graph for either).
"""
import torch

from .. import _native as nat
from ._btd import is_orthogonal, pack_arena
from .recursive_green_cal import BTDSystems
from .surface_green import LS_STOP, _orthogonal, pack_lead

C128 = torch.complex128


class TransmissionGraph:
    def __init__(self, systems, leads_L, leads_R, energies, eta_lead=1e-5, eta_device=0.0,
                 sys_idx=None, E_ref=0.0):
        """``systems``: list of ``(hd, hu, hl, sd, su, sl)`` (one per k-point, equal block
        sizes); ``leads_L`` / ``leads_R``: lists of ``(hL, hLL, hDL, sL, sLL, sDL)``, one per
        k-point; ``energies``: [P] real or complex; ``sys_idx``: [P] int32 k-point of each
        point (required with more than one system)."""
        self.bs = BTDSystems(systems)
        dev = self.bs.device
        self.blocked = max(self.bs.sizes) > nat.TRANSMISSION_NMAX
        if len(systems) != len(leads_L) or len(systems) != len(leads_R):
            raise ValueError("one left and one right lead per system")
        if len(systems) > 1 and sys_idx is None:
            raise ValueError("sys_idx is required with more than one system")
        e = torch.as_tensor(energies).to(device=dev, dtype=C128).reshape(-1) + E_ref
        self.P = e.shape[0]
        if self.P > 65535 and (self.blocked or max(
                side[0][0].shape[-1] for side in (leads_L, leads_R)) > nat.LEAD_NMAX):
            raise ValueError("at most 65535 points per graph with blocks above 64 or leads "
                             "above 48 orbitals")
        self.e = e.contiguous()
        self.z = (self.e + 1j * float(eta_device)).contiguous()
        self.sys_idx = (None if sys_idx is None
                        else sys_idx.to(device=dev, dtype=torch.int32).contiguous())
        self.eta_lead = float(eta_lead)
        self.leads = []
        for side in (leads_L, leads_R):
            recs = [pack_lead(l[0], l[1], l[3], l[4], l[2], l[5], dev)[0] for l in side]
            if side[0][2] is None:
                raise ValueError("graph capture needs the device-lead couplings hDL")
            n = side[0][0].shape[-1]
            a = side[0][2].shape[0]
            if a > min(self.bs.sizes[0], self.bs.sizes[-1]):
                raise ValueError("coupling rows exceed the end blocks")
            ortho = all(_orthogonal(l[3].to(dev), l[4].to(dev)) for l in side)
            ws = torch.empty(nat.load().dpnegf_lead_self_energy_workspace_elems(n, a, self.P),
                             dtype=C128, device=dev)
            self.leads.append(dict(
                arena=torch.cat(recs), stride=recs[0].numel(), nsys=len(side), n=n, a=a,
                ortho=int(ortho), ws=ws,
                gs=torch.empty(self.P, n, n, dtype=C128, device=dev),
                sigma=torch.empty(self.P, a, a, dtype=C128, device=dev),
                iters=torch.zeros(self.P, dtype=torch.int32, device=dev),
                status=torch.zeros(self.P, dtype=torch.int32, device=dev),
                resid=torch.zeros(self.P, dtype=torch.float64, device=dev)))
        self.T = torch.empty(self.P, dtype=torch.float64, device=dev)
        self.status = torch.zeros(self.P, dtype=torch.int32, device=dev)
        if self.blocked:
            self._prepare_blocked()
        else:
            self.work = torch.empty(nat.load().dpnegf_rgf_transmission_workspace_elems(
                len(self.bs.sizes), self.bs._h_bsz, self.P), dtype=C128, device=dev)
        self._launch()  # eager run: validates arguments before capture
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):  # torch's capture stream is current inside
            self._launch()
        torch.cuda.synchronize(dev)

    def _prepare_blocked(self):
        """Per k-point: its point indices and the gathered z, Sigma_L, Sigma_R,
        T, status and work buffers of the block-operation sweep."""
        dev = self.bs.device
        K = len(self.bs.sizes)
        idx = (self.sys_idx if self.sys_idx is not None
               else torch.zeros(self.P, dtype=torch.int32, device=dev)).cpu()
        nL = self.leads[0]["a"]
        nR = self.leads[1]["a"]
        self.parts = []
        for s in range(self.bs.nsys):
            sel = (idx == s).nonzero().reshape(-1)
            B = int(sel.numel())
            if B == 0:
                continue
            nwork = nat.load().dpnegf_rgf_transmission_blocked_workspace_elems(
                K, self.bs._h_bsz, B)
            self.parts.append(dict(
                s=s, sel=sel.to(dev), B=B,
                z=torch.empty(B, dtype=C128, device=dev),
                sL=torch.empty(B, nL, nL, dtype=C128, device=dev),
                sR=torch.empty(B, nR, nR, dtype=C128, device=dev),
                T=torch.empty(B, dtype=torch.float64, device=dev),
                st=torch.zeros(B, dtype=torch.int32, device=dev),
                work=torch.empty(nwork, dtype=C128, device=dev)))

    def _launch_blocked(self, st):
        K = len(self.bs.sizes)
        sL, sR = self.leads[0]["sigma"], self.leads[1]["sigma"]
        for p in self.parts:
            torch.index_select(self.z, 0, p["sel"], out=p["z"])
            torch.index_select(sL, 0, p["sel"], out=p["sL"])
            torch.index_select(sR, 0, p["sel"], out=p["sR"])
            arena = self.bs.arena[p["s"] * self.bs.stride:(p["s"] + 1) * self.bs.stride]
            nat.call("dpnegf_rgf_transmission_blocked", st, K, self.bs._h_bsz, nat.ptr(arena),
                     p["B"], nat.ptr(p["z"]), nat.ptr(p["sL"]), p["sL"].shape[-1],
                     nat.ptr(p["sR"]), p["sR"].shape[-1], int(self.bs.ortho), nat.ptr(p["T"]),
                     nat.ptr(p["work"]), nat.ptr(p["st"]))
            self.T.index_copy_(0, p["sel"], p["T"])
            self.status.index_copy_(0, p["sel"], p["st"])

    def _launch(self):
        st = nat.stream_ptr(self.bs.device)
        for L in self.leads:
            nat.call("dpnegf_lead_self_energy", st, L["n"], L["a"], 0, nat.ptr(L["arena"]),
                     L["stride"], L["nsys"], nat.ptr(self.sys_idx), self.P, nat.ptr(self.e),
                     self.eta_lead, float(LS_STOP), nat.ptr(L["ws"]), nat.ptr(L["gs"]),
                     nat.ptr(L["sigma"]), L["ortho"], nat.ptr(L["iters"]), nat.ptr(L["status"]),
                     nat.ptr(L["resid"]))
        if self.blocked:
            self._launch_blocked(st)
            return
        sL, sR = self.leads[0]["sigma"], self.leads[1]["sigma"]
        nat.call("dpnegf_rgf_transmission", st, nat.DPNEGF_FP64, len(self.bs.sizes),
                 self.bs._h_bsz, nat.ptr(self.bs.bsz), nat.ptr(self.bs.arena), self.bs.stride,
                 self.bs.nsys, nat.ptr(self.sys_idx), self.P, nat.ptr(self.z), nat.ptr(sL),
                 sL.shape[-1], nat.ptr(sR), sR.shape[-1], int(self.bs.ortho), nat.ptr(self.T),
                 nat.ptr(self.work), nat.ptr(self.status))

    def replay(self):
        """Runs the captured launches; returns T [P] (the graph's own buffer)."""
        self.graph.replay()
        return self.T

    @property
    def lead_status(self):
        return self.leads[0]["status"], self.leads[1]["status"]

    def update(self, systems=None, leads_L=None, leads_R=None):
        """Rewrites the packed blocks in place (same shapes, same overlap pattern) before
        the next ``replay``; the copies run on the current stream, ahead of it."""
        if systems is not None:
            if len(systems) != self.bs.nsys:
                raise ValueError("number of systems changed")
            if self.bs.ortho and not all(is_orthogonal(b[3], b[4], b[5]) for b in systems):
                raise ValueError("overlap pattern changed (captured as orthogonal)")
            for s, sysb in enumerate(systems):
                a = pack_arena(*sysb, self.bs.device)
                if a.numel() != self.bs.stride:
                    raise ValueError("block shapes changed")
                self.bs.arena[s * self.bs.stride:(s + 1) * self.bs.stride].copy_(a)
        for L, side in zip(self.leads, (leads_L, leads_R)):
            if side is None:
                continue
            if L["ortho"] and not all(_orthogonal(l[3].to(self.bs.device), l[4].to(self.bs.device))
                                      for l in side):
                raise ValueError("lead overlap pattern changed (captured as orthogonal)")
            recs = [pack_lead(l[0], l[1], l[3], l[4], l[2], l[5], self.bs.device)[0] for l in side]
            new = torch.cat(recs)
            if new.numel() != L["arena"].numel():
                raise ValueError("lead shapes changed")
            L["arena"].copy_(new)
