"""dpnegf_lead_self_energy_noalloc (decimation workspace in the library's
file-scope device scratch) against dpnegf_lead_self_energy (caller workspace):
same kernels, so the same self-energies, surface Green's functions,
iteration counts, status words and residuals, on two calls in a row (the
second smaller, with the scratch holding the first call's contents), for the
fused kernel (48-orbital ribbon lead, 4096 energies: two chunks of the
scratch) and the batched decimation (52 orbitals: two chunks, then one)."""
import ctypes

import numpy as np
import pytest
import torch

from dpnegf_amd import _native as nat

C = torch.complex128


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _outputs(P, n, a, dev):
    return dict(gs=torch.full((P, n, n), float("nan"), dtype=C, device=dev),
                sig=torch.full((P, a, a), float("nan"), dtype=C, device=dev),
                it=torch.full((P,), -7, dtype=torch.int32, device=dev),
                st=torch.full((P,), -7, dtype=torch.int32, device=dev),
                rs=torch.full((P,), float("nan"), dtype=torch.float64, device=dev))


def _run(noalloc, rec, n, a, mode, e, eta, ortho, dev):
    P = e.shape[0]
    o = _outputs(P, n, a, dev)
    st = nat.stream_ptr(dev)
    common = (n, a, mode, nat.ptr(rec), rec.numel(), 1, nat.ptr(None), P, nat.ptr(e), eta, 1e-40)
    tail = (nat.ptr(o["gs"]), nat.ptr(o["sig"]), ortho, nat.ptr(o["it"]), nat.ptr(o["st"]),
            nat.ptr(o["rs"]))
    if noalloc:
        nat.call("dpnegf_lead_self_energy_noalloc", st, *common, *tail)
    else:
        ws = torch.empty(nat.load().dpnegf_lead_self_energy_workspace_elems(n, a, P), dtype=C,
                         device=dev)
        nat.call("dpnegf_lead_self_energy", st, *common, nat.ptr(ws), *tail)
    torch.cuda.synchronize(dev)
    return {k: v.cpu() for k, v in o.items()}


def _same(x, y, what):
    assert torch.equal(x["it"], y["it"]), what
    assert torch.equal(x["st"], y["st"]), what
    for k in ("gs", "sig", "rs"):
        a, b = x[k], y[k]
        assert torch.isfinite(a).all(), (what, k)
        scale = b.abs().max().item()
        err = (a - b).abs().max().item() / scale
        assert err < 1e-12, (what, k, err)


def _ribbon_lead(dev):
    from dpnegf_amd import workloads as wl
    from dpnegf_amd.negf.surface_green import pack_lead
    s = wl.ribbon(W=48, K=2, seed=0)
    hL, hLL, hDL, sL, sLL, sDL = s.lead_L
    rec, n = pack_lead(hL, hLL, sL, sLL, hDL, sDL, dev)
    return rec, n, hDL.shape[0]


def _random_lead(n, dev, seed):
    from dpnegf_amd.negf.surface_green import pack_lead
    g = torch.Generator().manual_seed(seed)
    rc = lambda: (torch.randn(n, n, generator=g, dtype=torch.float64)
                  + 1j * torch.randn(n, n, generator=g, dtype=torch.float64))
    h0 = 0.15 * rc()
    h0 = h0 + h0.mH
    h1 = 0.12 * rc()
    rec, _ = pack_lead(h0, h1, torch.eye(n, dtype=C), torch.zeros(n, n, dtype=C), device=dev)
    return rec


@pytest.mark.gpu
def test_noalloc_fused_lead_matches_the_workspace_entry():
    dev = _dev()
    print("case noalloc fused n=48, P=4096 then P=300", flush=True)
    rec, n, a = _ribbon_lead(dev)
    for P in (4096, 300):
        E = torch.linspace(-3.0, 3.0, P, dtype=torch.float64).to(C).to(dev)
        ref = _run(False, rec, n, a, 0, E, 1e-6, 1, dev)
        got = _run(True, rec, n, a, 0, E, 1e-6, 1, dev)
        _same(got, ref, f"fused P={P}")


@pytest.mark.gpu
def test_noalloc_batched_lead_matches_the_workspace_entry():
    dev = _dev()
    n = 52
    per = nat.load().dpnegf_lead_self_energy_workspace_elems(n, n, 1)
    cap = nat.NOALLOC_SCRATCH_ELEMS // per
    print(f"case noalloc batched n={n}, {cap} points per chunk", flush=True)
    rec = _random_lead(n, dev, 5)
    for P in (cap + 40, 64):  # two chunks, then one
        E = (torch.linspace(-1.0, 1.0, P, dtype=torch.float64) + 0.02j).to(dev)
        ref = _run(False, rec, n, n, 1, E, 1e-5, 1, dev)
        got = _run(True, rec, n, n, 1, E, 1e-5, 1, dev)
        _same(got, ref, f"batched P={P}")


def test_noalloc_rejects_a_lead_larger_than_the_scratch():
    """One point of a 900-orbital lead needs more than the scratch holds: the
    call fails on the host before any device work (no device needed)."""
    lib = nat.load()
    n = 900
    need = lib.dpnegf_lead_self_energy_workspace_elems(n, n, 1)
    assert need > nat.NOALLOC_SCRATCH_ELEMS
    dummy = ctypes.c_void_p(16)
    rc = lib.dpnegf_lead_self_energy_noalloc(None, n, n, 1, dummy, 4 * n * n, 1, None, 3, dummy,
                                             1e-5, 1e-40, dummy, dummy, 1, None, None, None)
    assert rc == -2
    assert b"no-allocation scratch" in lib.dpnegf_last_error()
