"""Lead surface Green's function and self-energy on the GPU
(``dpnegf.negf.surface_green`` interface).

* :func:`selfEnergy` / :func:`surface_green` -- the reference signatures
  (``dpnegf/negf/surface_green.py:9-10, 229-231``) for one energy.
* :func:`self_energy_batched` -- new: every (E, k) point in one launch of the
  Lopez-Sancho kernel (``dpnegf_lead_self_energy``), replacing the joblib
  process pool of ``lead_property.py:599-683``.

The decimation keeps the reference's stopping rule (max |alpha| + |beta| <
1e-40, ``surface_green.py:208``), 100-iteration limit and residual checks;
per-point status words are turned into the reference's exceptions
(``RuntimeError`` / ``ArithmeticError``) and warnings here.
"""
import logging
import weakref

import numpy as np
import torch

from .. import _native as nat
from . import autotune
from ._btd import C128

log = logging.getLogger(__name__)

LS_STOP = 1.0e-40


class SelfEnergyResult:
    """Outputs of :func:`self_energy_batched`."""

    def __init__(self, sigma, gs, iters, status, resid):
        self.sigma, self.gs, self.iters, self.status, self.resid = sigma, gs, iters, status, resid


def _dev(x, device):
    if not torch.is_tensor(x):
        x = torch.as_tensor(np.asarray(x))
    return x.to(device=device, dtype=C128)


def pack_lead(hL, hLL, sL, sLL, hDL=None, sDL=None, device=None):
    """One lead's blocks as the library's per-system record."""
    n = hL.shape[-1]
    parts = [_dev(hL, device), _dev(hLL, device), _dev(sL, device), _dev(sLL, device)]
    if hDL is not None:
        sDL = sDL if sDL is not None else torch.zeros_like(hDL)
        parts += [_dev(hDL, device), _dev(sDL, device)]
    return torch.cat([p.contiguous().reshape(-1) for p in parts]), n


# id(sL) -> (weak reference to sL, its version, weak reference to sLL, its
# version, answer); an entry leaves with its sL
_ortho_seen = {}


def _forget(key, r):
    if _ortho_seen.get(key, (None,))[0] is r:
        del _ortho_seen[key]


def _orthogonal(sL, sLL):
    """S_L == I and S_LL == 0 (the orthogonal kernels apply).  The test reads
    the answer back from the device, which waits for every launch queued
    before it; the answer is kept per (sL, sLL) pair of tensors with their
    version counters (bumped by every in-place change), so a lead passed
    again unchanged, as in every energy chunk and SCF step, costs no wait."""
    key = id(sL)
    seen = _ortho_seen.get(key)
    if (seen is not None and seen[0]() is sL and seen[1] == sL._version
            and seen[2]() is sLL and seen[3] == sLL._version):
        return seen[4]
    ans = bool(torch.equal(sL, torch.eye(sL.shape[0], dtype=sL.dtype, device=sL.device))
               and torch.count_nonzero(sLL).item() == 0)
    _ortho_seen[key] = (weakref.ref(sL, lambda r, k=key: _forget(k, r)), sL._version,
                        weakref.ref(sLL), sLL._version, ans)
    return ans


def points_per_call(n, a, P, device, mem_fraction=0.5):
    """Points per ``dpnegf_lead_self_energy`` call: all ``P`` for the fused
    kernel (leads up to 48 orbitals); for larger leads at most 65535, and no
    more than fit, with their outputs, in ``mem_fraction`` of the free device
    memory (the batched decimation's workspace grows by about 12 n^2 complex
    elements per point)."""
    if P < 1 or (n <= nat.LEAD_NMAX and a <= nat.LEAD_NMAX):
        return max(P, 1)
    lib = nat.load()
    w1 = lib.dpnegf_lead_self_energy_workspace_elems(n, a, 1)
    w2 = lib.dpnegf_lead_self_energy_workspace_elems(n, a, 2)
    per = 16 * ((w2 - w1) + n * n + a * a) + 16 + 8 + 4 + 4
    free, _ = torch.cuda.mem_get_info(device)
    fit = (int(free * mem_fraction) - 16 * (2 * w1 - w2)) // per
    if fit < 1:
        raise torch.cuda.OutOfMemoryError(
            f"one point of the {n}-orbital lead decimation does not fit in free device memory")
    return min(P, fit, 65535)


def self_energy_batched(leads, energies, eta_lead, sys_idx=None, E_ref=0.0, Bulk=False,
                        raise_on_error=True, stop=LS_STOP):
    """Self-energies of P points.

    ``leads``: list (one per k-point) of tuples ``(hL, hLL, hDL, sL, sLL, sDL)``
    in ``get_hs_lead`` order (``hDL``/``sDL`` may be ``None``: then
    Sigma = (eS - H) - g_s^{-1}, or g_s^{-1} with ``Bulk``), all CUDA tensors.
    ``energies``: [P] (complex allowed) relative to ``E_ref``.
    Returns a :class:`SelfEnergyResult` with sigma [P, a, a] and g_s [P, n, n].
    """
    device = leads[0][0].device
    nat.require_device(leads[0][0])
    e = torch.as_tensor(energies).to(device=device, dtype=C128).reshape(-1) + E_ref
    e = e.contiguous()
    P = e.shape[0]
    recs = []
    has_dl = leads[0][2] is not None
    for (hL, hLL, hDL, sL, sLL, sDL) in leads:
        if (hDL is not None) != has_dl:
            raise ValueError("all leads of a batch must agree on hDL")
        rec, n = pack_lead(hL, hLL, sL, sLL, hDL, sDL, device)
        recs.append(rec)
    n = leads[0][0].shape[-1]
    a = leads[0][2].shape[0] if has_dl else n
    mode = 0 if has_dl else (2 if Bulk else 1)
    stride = recs[0].numel()
    arena = torch.cat(recs)
    ortho = all(_orthogonal(l[3], l[4]) for l in leads)
    if len(leads) > 1 and sys_idx is None:
        raise ValueError("sys_idx is required with more than one lead system")
    if sys_idx is not None:
        sys_idx = sys_idx.to(device=device, dtype=torch.int32).contiguous()
    gs = torch.empty(P, n, n, dtype=C128, device=device)
    sig = torch.empty(P, a, a, dtype=C128, device=device)
    iters = torch.zeros(P, dtype=torch.int32, device=device)
    status = torch.zeros(P, dtype=torch.int32, device=device)
    resid = torch.zeros(P, dtype=torch.float64, device=device)
    chunk = points_per_call(n, a, P, device)
    if n > nat.LEAD_NMAX and autotune.enabled():
        _tune_decimation(arena, stride, n, a, mode, e, eta_lead, stop, ortho, min(chunk, P), device)
    eps = torch.empty(nat.load().dpnegf_lead_self_energy_workspace_elems(n, a, chunk),
                      dtype=C128, device=device)
    for c0 in range(0, P, chunk):
        c1 = min(P, c0 + chunk)
        nat.call("dpnegf_lead_self_energy", nat.stream_ptr(device), n, a, mode, nat.ptr(arena),
                 stride, len(leads), nat.ptr(sys_idx[c0:c1] if sys_idx is not None else None),
                 c1 - c0, nat.ptr(e[c0:c1]), float(eta_lead), float(stop), nat.ptr(eps),
                 nat.ptr(gs[c0:c1]), nat.ptr(sig[c0:c1]), int(ortho), nat.ptr(iters[c0:c1]),
                 nat.ptr(status[c0:c1]), nat.ptr(resid[c0:c1]))
    res = SelfEnergyResult(sig, gs, iters, status, resid)
    if raise_on_error:
        check_status(res, e)
    return res


def _tune_decimation(arena, stride, n, a, mode, e, eta_lead, stop, ortho, B, device, sample=4):
    """Kernel shapes of the batched decimation (``autotune.tune_lead``),
    recorded on the first ``sample`` points of system 0 into scratch outputs."""
    m = min(sample, e.shape[0])

    def run_sample():
        ws = torch.empty(nat.load().dpnegf_lead_self_energy_workspace_elems(n, a, m), dtype=C128,
                         device=device)
        gs = torch.empty(m, n, n, dtype=C128, device=device)
        sig = torch.empty(m, a, a, dtype=C128, device=device)
        it = torch.zeros(m, dtype=torch.int32, device=device)
        st = torch.zeros(m, dtype=torch.int32, device=device)
        rs = torch.zeros(m, dtype=torch.float64, device=device)
        nat.call("dpnegf_lead_self_energy", nat.stream_ptr(device), n, a, mode, nat.ptr(arena),
                 stride, 1, None, m, nat.ptr(e[:m].contiguous()), float(eta_lead), float(stop),
                 nat.ptr(ws), nat.ptr(gs), nat.ptr(sig), int(ortho), nat.ptr(it), nat.ptr(st),
                 nat.ptr(rs))

    autotune.tune_lead(run_sample, n, B, device)


def check_status(res, e):
    """Map per-point status words onto the reference's exceptions/warnings
    (``surface_green.py:214-224``)."""
    st = res.status
    if int(st.max().item()) == 0:
        return
    st = st.cpu()
    bad = (st & nat.STATUS_BAD_INDEX) != 0
    if bad.any():
        raise ValueError(f"sys_idx outside the lead systems at {int(bad.sum())} points")
    bad = (st & nat.STATUS_LS_NOCONV) != 0
    if bad.any():
        raise RuntimeError("Lopez-scheme not converged.")
    bad = (st & (nat.STATUS_LS_RESID_FAIL | nat.STATUS_SINGULAR)) != 0
    if bad.any():
        raise ArithmeticError("Criteria not met. Please check output...")
    warn = ((st & nat.STATUS_LS_RESID_WARN) != 0).nonzero().reshape(-1)
    if len(warn):
        ee = e.cpu()
        r = res.resid.cpu()
        for i in warn[:8].tolist():
            log.warning("Lopez-scheme not-so-well converged at E = %.4f eV: %g", ee[i].real, r[i])


def selfEnergy(hL, hLL, sL, sLL, ee, hDL=None, sDL=None, etaLead=1e-8, Bulk=False, E_ref=0.0,
               dtype=np.complex128, device='cuda', method='Lopez-Sancho', numba_jit=None):
    """Reference signature (``surface_green.py:9-10``); returns ``(Sigma, g_s)``
    on the GPU.  ``method='GEP'`` is served by the decimation kernel as well
    (both give the same surface Green's function)."""
    dev = torch.device(device) if not torch.is_tensor(hL) else hL.device
    if dev.type != "cuda":
        dev = torch.device("cuda")
    lead = tuple(_dev(x, dev) if x is not None else None for x in (hL, hLL, hDL, sL, sLL, sDL))
    if lead[2] is not None and lead[5] is None:
        lead = lead[:5] + (torch.zeros_like(lead[2]),)
    ee = torch.as_tensor(ee).reshape(-1)
    if ee.numel() != 1:
        raise ValueError("selfEnergy takes one energy; use self_energy_batched for many")
    res = self_energy_batched([lead], ee, etaLead, E_ref=E_ref, Bulk=Bulk)
    return res.sigma[0], res.gs[0]


def surface_green(H, h01, S, s01, ee, method='Lopez-Sancho', numba_jit=None):
    """Surface Green's function at complex energy ``ee`` (already including
    i*eta), ``surface_green.py:229-279``."""
    dev = H.device if torch.is_tensor(H) else torch.device("cuda")
    n = H.shape[-1]
    zero = torch.zeros(n, n, dtype=C128, device=dev)
    lead = (_dev(H, dev), _dev(h01, dev), _dev(zero, dev), _dev(S, dev), _dev(s01, dev), zero)
    ee = complex(ee)
    res = self_energy_batched([lead], torch.tensor([ee.real], dtype=C128), ee.imag)
    return res.gs[0]
