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
