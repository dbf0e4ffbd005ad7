"""Captured transmission path (TransmissionGraph) against the eager pipeline:
several replays with eager work and synchronisation between them, then an
SCF-like update of the Hamiltonian (a potential on the device layers) and a
replay against the eager result for the new Hamiltonian.  Bar 1e-12 of max T."""
import pytest
import torch

from dpnegf_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _blocks(s):
    return (s.hd, s.hu, s.hl, s.sd, s.su, s.sl)


def test_transmission_graph_replay_and_update():
    from dpnegf_amd.negf.device_property import transmission_pipeline
    from dpnegf_amd.negf.graph import TransmissionGraph
    dev = _dev()
    print("case graph ribbon16 K12", flush=True)
    s = wl.ribbon(W=16, K=12, seed=2).to(dev)
    E = torch.linspace(-2.5, 2.5, 300, dtype=torch.float64, device=dev)
    eta = 1e-4
    ref = transmission_pipeline(s, E, eta_lead=eta, eta_device=eta)
    g = TransmissionGraph([_blocks(s)], [s.lead_L], [s.lead_R], E, eta_lead=eta, eta_device=eta)
    scale = float(ref.abs().max())
    other = wl.ribbon(W=8, K=6, seed=5).to(dev)
    for r in range(3):
        T = g.replay().clone()
        torch.cuda.synchronize()
        assert float((T - ref).abs().max()) <= 1e-12 * scale, r
        transmission_pipeline(other, E[:50], eta_lead=eta, eta_device=eta)  # eager work between
        torch.cuda.synchronize()
    # potential step on the device layers (SCF-like), leads unchanged
    V = torch.linspace(0.0, 0.3, len(s.hd), dtype=torch.float64)
    s2 = wl.BTDSystem([h - float(v) * torch.eye(h.shape[0], dtype=h.dtype, device=dev)
                       for h, v in zip(s.hd, V)], s.hu, s.hl, s.sd, s.su, s.sl, s.lead_L, s.lead_R)
    ref2 = transmission_pipeline(s2, E, eta_lead=eta, eta_device=eta)
    g.update(systems=[_blocks(s2)])
    T2 = g.replay().clone()
    torch.cuda.synchronize()
    assert float((T2 - ref2).abs().max()) <= 1e-12 * float(ref2.abs().max())
    assert float((ref2 - ref).abs().max()) > 1e-3  # the update changed the answer
    assert int(g.status.count_nonzero()) == 0


def test_transmission_graph_large_blocks_two_kpoints():
    """Blocks 72/90/81/72 (block-operation sweep) with 72-orbital leads
    (batched decimation, launched with all its iterations under capture and
    converged points frozen) for 2 k-points: replays with eager work between
    them and after an update, against the eager pipeline at 1e-12."""
    from dpnegf_amd.negf.device_property import transmission_pipeline
    from dpnegf_amd.negf.graph import TransmissionGraph
    dev = _dev()
    print("case graph multiorbital 72-90 blocks, 2 k-points", flush=True)
    systems = [wl.multiorbital(widths=(24, 30, 27), K=4, norb=3, kpoint=k / 32, seed=4).to(dev)
               for k in (0, 7)]
    nE = 40
    E = torch.linspace(-1.2, 1.2, nE, dtype=torch.float64, device=dev)
    eta = 1e-4
    ref = transmission_pipeline(systems, E, eta_lead=eta, eta_device=eta).reshape(-1)
    sys_idx = torch.arange(2, dtype=torch.int32, device=dev).repeat_interleave(nE)
    g = TransmissionGraph([_blocks(s) for s in systems], [s.lead_L for s in systems],
                          [s.lead_R for s in systems], E.repeat(2), eta_lead=eta, eta_device=eta,
                          sys_idx=sys_idx)
    assert g.blocked
    scale = float(ref.abs().max())
    other = wl.ribbon(W=8, K=6, seed=5).to(dev)
    for r in range(3):
        T = g.replay().clone()
        torch.cuda.synchronize()
        assert float((T - ref).abs().max()) <= 1e-12 * scale, r
        transmission_pipeline(other, E[:20], eta_lead=eta, eta_device=eta)
        torch.cuda.synchronize()
    V = 0.2
    shifted = [wl.BTDSystem([h - V * torch.eye(h.shape[0], dtype=h.dtype, device=dev)
                             for h in s.hd], s.hu, s.hl, s.sd, s.su, s.sl, s.lead_L, s.lead_R)
               for s in systems]
    ref2 = transmission_pipeline(shifted, E, eta_lead=eta, eta_device=eta).reshape(-1)
    g.update(systems=[_blocks(s) for s in shifted])
    T2 = g.replay().clone()
    torch.cuda.synchronize()
    assert float((T2 - ref2).abs().max()) <= 1e-12 * float(ref2.abs().max())
    assert float((ref2 - ref).abs().max()) > 1e-3
    assert int(g.status.count_nonzero()) == 0
    assert all(int(st.count_nonzero()) == 0 for st in g.lead_status)
