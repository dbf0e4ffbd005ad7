"""Large-block paths vs the oracle: batched Lopez-Sancho (lead size > 48) and the
block-operation transmission sweep (block size > 64, uneven).  Bar 1e-10 (FP64)."""
import numpy as np
import pytest
import torch

from oracle import device as odev, lead as olead, rgf as orgf, surface_green as osg
from dpnegf_amd import workloads as wl


def _f64(x):
    """Energies as float64 (a Python float would give a float32 tensor)."""
    return torch.tensor(x, dtype=torch.float64)


pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


@pytest.mark.parametrize("W", [64, 104])
def test_batched_lopez_sancho_matches_oracle(W):
    """Leads above 48 orbitals (batched decimation); at 104 the decimation's
    inverse takes the Schur-complement form."""
    from dpnegf_amd.negf.surface_green import self_energy_batched
    from tests.test_gpu_rgf import _leads_sigmas_oracle, check_surface
    dev = _dev()
    s = wl.ribbon(W=W, K=2, disorder=0.0)
    E = np.array([-2.9, -0.7, 0.0, 1.3])
    eta = 1e-6
    print(f"case ribbon{W} leads", flush=True)
    ref = _leads_sigmas_oracle(s, E, eta)
    for lead, r in zip((s.lead_L, s.lead_R), ref):
        res = self_energy_batched([tuple(x.to(dev) for x in lead)], _f64(E), eta,
                                  raise_on_error=False)
        check_surface(res, r)


def test_blocked_transmission_uneven_large_blocks():
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    s = wl.multiorbital(widths=(24, 30, 27), K=4, norb=3, seed=4)  # blocks 72, 90, 81, 72
    E = np.array([-1.1, 0.35])
    eta = 1e-5
    print("case multiorbital 72/90/81", flush=True)
    ref = []
    for e in E:
        sig = [osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e), hDL=l[2], sDL=l[5],
                              etaLead=eta)[0] for l in (s.lead_L, s.lead_R)]
        ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), s.hl, s.hd, s.hu, s.sd,
                                s.su, s.sl, sig[0], sig[1], eta=eta)
        ref.append(float(odev.transmission(ans[0], olead.sigma2gamma(sig[0]),
                                           olead.sigma2gamma(sig[1]))))
    T = transmission_pipeline(s.to(dev), _f64(E), eta_lead=eta, eta_device=eta).cpu().numpy()
    assert np.abs(T - np.array(ref)).max() < 1e-10 * max(1.0, np.abs(ref).max()), (T, ref)


@pytest.mark.parametrize("sizes", [[256, 192, 240], [144, 168, 136], [100, 128, 97]])
def test_recursive_gf_blocks_up_to_256(sizes):
    """Blocks above 96 (the Schur-complement inverse: two halves and a
    Newton step): every output block of recursive_gf for uneven blocks with
    non-orthogonal S."""
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    from tests.test_gpu_rgf import _cmp_tuple, _rand_se
    dev = _dev()
    print(f"case sizes={sizes}", flush=True)
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=7, overlap=True)
    B = 2
    sL, sR = _rand_se(B, sizes[0], 8), _rand_se(B, sizes[-1], 9)
    E = torch.tensor([-0.6, 0.45], dtype=torch.float64) + 1e-3j
    ref = orgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=1e-4, need_gr_lc=True)
    to = lambda xs: [x.to(dev) for x in xs]
    out = recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                       sR.to(dev), eta=1e-4, need_gr_lc=True)
    _cmp_tuple(out, ref)


@pytest.mark.parametrize("sizes", [[300, 264], [520, 300]])
def test_recursive_gf_blocks_above_256(sizes):
    """Blocks above 256: the inverse's halves are themselves inverted through
    the Schur complement (300 -> 150 -> 75, 520 -> 260 -> 130 -> 65); every
    output block of recursive_gf, non-orthogonal S, against the oracle."""
    test_recursive_gf_blocks_up_to_256(sizes)


def test_transmission_and_leads_of_300_orbitals():
    """10 orbitals per site, layers of 30 and 33 sites (blocks 300 / 330)
    with leads of the same widths: batched Lopez-Sancho and the blocked
    sweep above 256 orbitals, T against the oracle; the self-energies also
    in chunks of 2 points, as a device without room for all points runs them."""
    from dpnegf_amd.negf import surface_green as sg
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    s = wl.multiorbital(widths=(30, 33), K=2, norb=10, seed=12)
    E = np.array([-0.8, 0.3, 0.6])
    eta = 1e-4
    print("case multiorbital 300/330", flush=True)
    ref = []
    for e in E:
        sig = [osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e),
                              hDL=l[2], sDL=l[5], etaLead=eta)[0] for l in (s.lead_L, s.lead_R)]
        ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), s.hl, s.hd, s.hu, s.sd,
                                s.su, s.sl, sig[0], sig[1], eta=eta)
        ref.append(float(odev.transmission(ans[0], olead.sigma2gamma(sig[0]),
                                           olead.sigma2gamma(sig[1]))))
    sd = s.to(dev)
    T = transmission_pipeline(sd, _f64(E), eta_lead=eta, eta_device=eta).cpu().numpy()
    assert np.abs(T - np.array(ref)).max() < 1e-10 * max(1.0, np.abs(ref).max()), (T, ref)
    lead = tuple(x.to(dev) if x is not None else None for x in s.lead_R)
    whole = sg.self_energy_batched([lead], _f64(E), eta)
    orig = sg.points_per_call
    try:
        sg.points_per_call = lambda n, a, P, device: 2
        parts = sg.self_energy_batched([lead], _f64(E), eta)
    finally:
        sg.points_per_call = orig
    scale = whole.sigma.abs().max().item()
    assert (parts.sigma - whole.sigma).abs().max().item() < 1e-13 * scale
    assert torch.equal(parts.iters, whole.iters)


def test_transmission_multiorbital_256_blocks():
    """8 orbitals per site, layers of 32 / 24 / 30 sites (blocks 256 / 192 /
    240): transmission through the blocked sweep with leads of 256 and 240
    orbitals through the batched Lopez-Sancho path, against the oracle."""
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    s = wl.multiorbital(widths=(32, 24, 30), K=3, norb=8, seed=11)
    E = np.array([-0.9, 0.25])
    eta = 1e-4
    print("case multiorbital 256/192/240", flush=True)
    ref = []
    for e in E:
        sig = [osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e),
                              hDL=l[2], sDL=l[5], etaLead=eta)[0] for l in (s.lead_L, s.lead_R)]
        ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), s.hl, s.hd, s.hu, s.sd,
                                s.su, s.sl, sig[0], sig[1], eta=eta)
        ref.append(float(odev.transmission(ans[0], olead.sigma2gamma(sig[0]),
                                           olead.sigma2gamma(sig[1]))))
    T = transmission_pipeline(s.to(dev), _f64(E), eta_lead=eta,
                              eta_device=eta).cpu().numpy()
    assert np.abs(T - np.array(ref)).max() < 1e-10 * max(1.0, np.abs(ref).max()), (T, ref)


def test_autotuned_chunk_gives_identical_transmission():
    """The block-operation sweep computes every point independently, so the
    chunk the autotuner picks changes nothing in T: bit-identical to fixed
    chunks, and the pick is a candidate that fits the memory budget."""
    from dpnegf_amd.negf import autotune
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems, rgf_transmission
    dev = _dev()
    s = wl.random_btd([72, 80, 72], seed=5)
    hd, sd, hl, hu, sl, su = [[t.to(dev) for t in blk] for blk in s]
    bs = BTDSystems([(hd, hu, hl, sd, su, sl)], dev)
    P = 600
    print(f"case autotune blocks {bs.sizes} P={P}", flush=True)
    z = (torch.linspace(-2, 2, P, dtype=torch.float64, device=dev) + 1e-4j).to(torch.complex128)
    g = torch.Generator().manual_seed(3)
    sig = lambda n: (0.1 * torch.randn(P, n, n, dtype=torch.complex128, generator=g)
                     - 0.05j * torch.eye(n, dtype=torch.complex128)).to(dev)
    sL, sR = sig(8), sig(6)
    chosen = autotune.blocked_chunk(bs, P)
    free, _ = torch.cuda.mem_get_info(dev)
    assert chosen in autotune.CANDIDATES and chosen <= autotune.max_batch(bs.sizes, free)
    T_auto = rgf_transmission(bs, z, sL, sR)
    for c in (128, 4096):
        assert torch.equal(rgf_transmission(bs, z, sL, sR, chunk=c), T_auto), c
    assert torch.isfinite(T_auto).all()
