"""GPU parity of the RGF kernels against the CPU oracle (complex128).

Bar: relative 1e-10 on every output block / T(E) (FP64 path).  The oracle is
the reference restatement in ``oracle/`` (pinned by tests/test_oracle_*.py).
"""
import numpy as np
import pytest
import torch

from oracle import device as odev, lead as olead, rgf as orgf, surface_green as osg
from dpnegf_amd import workloads as wl


def _f64(x):
    """Energies as float64 (a Python float would give a float32 tensor)."""
    return torch.tensor(x, dtype=torch.float64)

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _rand_se(B, n, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(B, n, n, generator=g, dtype=torch.float64) + 1j * torch.randn(
        B, n, n, generator=g, dtype=torch.float64)
    return 0.3 * a - 0.5j * torch.eye(n, dtype=torch.complex128)


def _cmp(a, b, scale, name):
    a = a.detach().cpu()
    b = b.detach().cpu()
    err = (a - b).abs().max().item() / scale
    assert err < RTOL, f"{name}: rel err {err:.3e}"


def _cmp_tuple(gpu, ref):
    names = ["g_trans", "gr_lc", "grd", "grl", "gru", "gr_left", "gnd", "gnl", "gnu", "gin_left",
             "gpd", "gpl", "gpu", "gip_left"]
    for name, g, r in zip(names, gpu, ref):
        if r is None:
            assert g is None, name
            continue
        if torch.is_tensor(r):
            _cmp(g, r, r.abs().max().item() + 1e-300, name)
        else:
            assert isinstance(g, list) and len(g) == len(r), name
            scale = max([t.abs().max().item() for t in r] + [1e-300])
            for q, (gq, rq) in enumerate(zip(g, r)):
                assert gq.shape == rq.shape, (name, q)
                _cmp(gq, rq, scale, f"{name}[{q}]")


@pytest.mark.parametrize("sizes,overlap", [([8, 6, 8], False), ([5, 9, 3, 7], True), ([12], False),
                                           ([16, 16, 16, 16, 16], False), ([40, 64, 33], True)])
def test_recursive_gf_retarded_matches_oracle(sizes, overlap):
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    dev = _dev()
    print(f"case sizes={sizes} overlap={overlap}", flush=True)
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=1, overlap=overlap)
    B = 5
    sL, sR = _rand_se(B, sizes[0], 2), _rand_se(B, sizes[-1], 3)
    E = torch.linspace(-1, 1, B, dtype=torch.float64) + 1e-3j
    ref = orgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=1e-4, need_gr_lc=True)
    to = lambda xs: [x.to(dev) for x in xs]
    out = recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                       sR.to(dev), eta=1e-4, need_gr_lc=True)
    _cmp_tuple(out, ref)


def test_recursive_gf_lesser_greater_and_scalar_squeeze():
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    dev = _dev()
    sizes = [6, 4, 6]
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=4)
    B = 3
    sL, sR = _rand_se(B, 6, 5), _rand_se(B, 6, 6)
    E = torch.tensor([-0.4, 0.1, 0.9], dtype=torch.complex128)
    s_in = odev.build_s_in([(n, n) for n in sizes], olead.sigma2gamma(sL) * 0.7,
                           olead.sigma2gamma(sR) * 0.2)
    s_out = odev.build_s_in([(n, n) for n in sizes], olead.sigma2gamma(sL) * 0.3,
                            olead.sigma2gamma(sR) * 0.8)
    ref = orgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=1e-5, s_in=s_in, s_out=s_out,
                            need_lesser=True, need_greater=True, need_gr_lc=True)
    to = lambda xs: [x.to(dev) for x in xs]
    out = recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                       sR.to(dev), eta=1e-5, s_in=to(s_in), s_out=to(s_out), need_lesser=True,
                       need_greater=True, need_gr_lc=True)
    _cmp_tuple(out, ref)
    # scalar energy + 2-D self-energies squeeze the batch dim
    ref1 = orgf.recursive_gf(E[1], hl, hd, hu, sd, su, sl, sL[1], sR[1], eta=1e-5)
    out1 = recursive_gf(E[1].to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl),
                        sL[1].to(dev), sR[1].to(dev), eta=1e-5)
    assert out1[0].ndim == 2
    _cmp_tuple(out1, ref1)


def test_single_block_outputs_are_empty_lists():
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    dev = _dev()
    hd, sd, hl, hu, sl, su = wl.random_btd([7], seed=2)
    sL = _rand_se(1, 7, 1)[0]
    out = recursive_gf(torch.tensor(0.3, dtype=torch.complex128), [], [hd[0].to(dev)], [],
                       [sd[0].to(dev)], [], [], sL.to(dev), sL.to(dev), eta=1e-5)
    ref = orgf.recursive_gf(torch.tensor(0.3, dtype=torch.complex128), [], hd, [], sd, [], [], sL,
                            sL, eta=1e-5)
    assert out[3] == [] and out[4] == []
    _cmp_tuple(out, ref)


def test_cpu_tensors_are_rejected():
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    hd, sd, hl, hu, sl, su = wl.random_btd([3, 3], seed=2)
    with pytest.raises(RuntimeError, match="GPU only"):
        recursive_gf(torch.tensor(0.0), hl, hd, hu, sd, su, sl, hd[0], hd[1])


def _leads_sigmas_oracle(s, E, eta_lead):
    """Oracle g_s and Sigma per energy, plus the spread between two stable
    orderings of the decimation (LU solve vs refined inverse): the agreement an
    ill-conditioned point (small eta near a band edge) allows at all."""
    out = []
    for lead in (s.lead_L, s.lead_R):
        hL, hLL, hDL, sL, sLL, sDL = lead
        sig, gs, it, spread = [], [], [], []
        for e in E:
            a = osg.lopez_sancho(hL.numpy(), hLL.numpy(), sL.numpy(), sLL.numpy(),
                                 complex(e) + 1j * eta_lead)
            b = osg.lopez_sancho(hL.numpy(), hLL.numpy(), sL.numpy(), sLL.numpy(),
                                 complex(e) + 1j * eta_lead, variant="refined")
            it.append(a[1])
            gs.append(torch.tensor(a[0]))
            spread.append(np.abs(a[0] - b[0]).max() / np.abs(a[0]).max())
            sig.append(osg.selfEnergy(hL, hLL, sL, sLL, _f64(e), hDL=hDL, sDL=sDL,
                                      etaLead=eta_lead)[0])
        out.append((torch.stack(sig), torch.stack(gs), it, spread))
    return out


def check_surface(res, ref):
    """Per energy: |g_s - g_s^oracle| and |Sigma - Sigma^oracle| within
    max(1e-10, 100 x ordering spread) of the largest element; iterations +-1."""
    rsig, rgs, rit, spread = ref
    for b in range(len(rit)):
        tol = max(RTOL, 100 * spread[b])
        g = res.gs[b].cpu()
        sg = res.sigma[b].cpu()
        assert (g - rgs[b]).abs().max().item() <= tol * rgs[b].abs().max().item(), (b, spread[b])
        assert (sg - rsig[b]).abs().max().item() <= tol * rsig[b].abs().max().item(), (b, spread[b])
        assert abs(int(res.iters[b]) - rit[b]) <= 1, (b, int(res.iters[b]), rit[b])


@pytest.mark.parametrize("name", ["chain2", "ribbon16", "ribbon48"])
def test_lead_self_energy_matches_oracle(name):
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    if name == "chain2":
        s, eta, E = wl.two_orbital_chain(4), 1e-5, np.linspace(-3, 3, 13)
    elif name == "ribbon16":
        s, eta, E = wl.ribbon(W=16, K=2, disorder=0.0), 1e-6, np.linspace(-3, 3, 9)
    else:
        s, eta, E = wl.ribbon(W=48, K=2, disorder=0.0), 1e-6, np.linspace(-3, 3, 5)
    print(f"case {name}", flush=True)
    ref = _leads_sigmas_oracle(s, E, eta)
    for lead, r in zip((s.lead_L, s.lead_R), ref):
        res = self_energy_batched([tuple(x.to(dev) for x in lead)], _f64(E), eta,
                                  raise_on_error=False)
        check_surface(res, r)


def _random_lead(n, nr, overlap, seed):
    """Lead tuple (hL, hLL, hDL, sL, sLL, sDL) with dense Hermitian hL, dense
    hLL and hDL (nr x n), and for `overlap` a Hermitian positive-definite sL
    with small sLL, sDL; else the orthogonal pattern (I, 0, 0)."""
    rng = np.random.default_rng(seed)

    def rc(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    a = rc(n, n)
    hL = 0.25 * (a + a.conj().T)
    hLL = 0.3 * rc(n, n)
    hDL = 0.3 * rc(nr, n)
    if overlap:
        b = rc(n, n)
        sL = np.eye(n) + 0.02 * (b + b.conj().T)
        sLL, sDL = 0.02 * rc(n, n), 0.02 * rc(nr, n)
    else:
        sL, sLL, sDL = np.eye(n), np.zeros((n, n)), np.zeros((nr, n))
    return tuple(torch.tensor(x, dtype=torch.complex128) for x in (hL, hLL, hDL, sL, sLL, sDL))


@pytest.mark.parametrize("n,nr,overlap", [(24, 24, True), (32, 20, False), (40, 40, False),
                                          (44, 30, True), (48, 48, True)])
def test_lead_self_energy_sizes_and_overlap(n, nr, overlap):
    """Every instantiation of the fused decimation (padded to 32 or 48, with
    and without an overlap matrix) against the oracle."""
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    lead = _random_lead(n, nr, overlap, seed=n + nr)

    class _S:  # the oracle helper takes an object with lead_L / lead_R
        lead_L = lead_R = lead

    E = np.linspace(-1.5, 1.5, 5)
    print(f"case lead n={n} nr={nr} overlap={overlap}", flush=True)
    ref = _leads_sigmas_oracle(_S, E, 1e-4)[0]
    res = self_energy_batched([tuple(x.to(dev) for x in lead)], _f64(E), 1e-4,
                              raise_on_error=False)
    assert int((res.status & 7).sum().item()) == 0, res.status
    check_surface(res, ref)


# Points of the 4096-energy grid linspace(-3, 3, 4096) where the W=48 ribbon
# lead's decimation is worst conditioned at eta = 1e-6 (Dyson residual of an
# explicit-inverse decimation above 1e-3 there).
EDGE_IDX = (137, 452, 1036, 1352, 1550, 2545, 2743, 3643)


def test_lead_self_energy_band_edges_eta_1e6():
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    s = wl.ribbon(W=48, K=2, disorder=0.0)
    E = np.linspace(-3, 3, 4096)[list(EDGE_IDX)]
    print("case ribbon48 band edges", flush=True)
    ref = _leads_sigmas_oracle(s, E, 1e-6)
    for lead, r in zip((s.lead_L, s.lead_R), ref):
        res = self_energy_batched([tuple(x.to(dev) for x in lead)], _f64(E), 1e-6,
                                  raise_on_error=False)
        assert int((res.status & 4).sum().item()) == 0, res.resid
        assert float(res.resid.max().item()) < 1e-6, res.resid
        check_surface(res, r)


def _T_oracle(s, E, eta_lead, eta_dev):
    out = []
    for e in E:
        sig = [osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e), hDL=l[2], sDL=l[5],
                              etaLead=eta_lead)[0] for l in (s.lead_L, s.lead_R)]
        ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), s.hl, s.hd, s.hu, s.sd,
                                s.su, s.sl, sig[0], sig[1], eta=eta_dev)
        out.append(float(odev.transmission(ans[0], olead.sigma2gamma(sig[0]),
                                           olead.sigma2gamma(sig[1]))))
    return np.array(out)


@pytest.mark.parametrize("name", ["chain2", "ribbon16", "ribbon48", "uneven"])
def test_transmission_pipeline_matches_oracle(name):
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    if name == "chain2":
        s, eta, E = wl.two_orbital_chain(16), 1e-5, np.linspace(-3, 3, 21)
    elif name == "ribbon16":
        s, eta, E = wl.ribbon(W=16, K=12, seed=3), 1e-6, np.linspace(-3, 3, 11)
    elif name == "ribbon48":
        s, eta, E = wl.ribbon(W=48, K=6, seed=3), 1e-6, np.linspace(-3, 3, 5)
    else:
        s = wl.multiorbital(widths=(4, 6, 5), K=5, norb=3, seed=2)
        eta, E = 1e-5, np.linspace(-2, 2, 7)
    print(f"case {name}", flush=True)
    ref = _T_oracle(s, E, eta, eta)
    T = transmission_pipeline(s.to(dev), _f64(E), eta_lead=eta, eta_device=eta).cpu().numpy()
    err = np.abs(T - ref).max() / max(1.0, np.abs(ref).max())
    assert err < RTOL, (err, T, ref)


def test_mixed_precision_transmission_close_to_fp64():
    """FP32 fused sweep (precision="mixed") against the FP64 path on a
    well-conditioned ribbon (W=16, K=16, eta=1e-3): FP32 rounding only, bounded
    by 1e-3 absolute on T <= 16 (a wrong matrix-core layout gives O(1))."""
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems, rgf_transmission
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    s = wl.ribbon(W=16, K=16, seed=3).to(dev)
    E = torch.linspace(-2.5, 2.5, 64, dtype=torch.float64, device=dev)
    eta = 1e-3
    print("case mixed ribbon16 K16", flush=True)
    sL = self_energy_batched([s.lead_L], E, eta, raise_on_error=False).sigma
    sR = self_energy_batched([s.lead_R], E, eta, raise_on_error=False).sigma
    bs = BTDSystems([(s.hd, s.hu, s.hl, s.sd, s.su, s.sl)], dev)
    z = (E.to(torch.complex128) + 1j * eta).contiguous()
    T64 = rgf_transmission(bs, z, sL, sR, precision="fp64")
    T32 = rgf_transmission(bs, z, sL, sR, precision="mixed")
    err = float((T32 - T64).abs().max().item())
    print(f"mixed max |dT| = {err:.3e}", flush=True)
    assert err < 1e-3, err


def test_recursive_gf_energy_dependent_seP():
    """A 3-D (per-energy) probe self-energy seP (``recursive_green_cal.py:411-414``)
    next to a 2-D one and a [1, n, n] one broadcast over the batch, with the
    lesser outputs, against the oracle."""
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    dev = _dev()
    sizes = [5, 7, 4]
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=8, overlap=True)
    B = 4
    sL, sR = _rand_se(B, 5, 9), _rand_se(B, 4, 10)
    seP = [_rand_se(B, 5, 11) * 0.2, _rand_se(1, 7, 12)[0] * 0.1, _rand_se(1, 4, 13) * 0.3]
    E = torch.linspace(-0.8, 0.6, B, dtype=torch.float64).to(torch.complex128)
    s_in = odev.build_s_in([(n, n) for n in sizes], olead.sigma2gamma(sL) * 0.6,
                           olead.sigma2gamma(sR) * 0.4)
    ref = orgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, seP=seP, eta=1e-5, s_in=s_in,
                            need_lesser=True, need_gr_lc=True)
    to = lambda xs: [x.to(dev) for x in xs]
    out = recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                       sR.to(dev), seP=to(seP), eta=1e-5, s_in=to(s_in), need_lesser=True,
                       need_gr_lc=True)
    _cmp_tuple(out, ref)
    # without the energy-dependent part the result differs (the addend is used)
    plain = recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                         sR.to(dev), seP=[0, seP[1].to(dev), seP[2].to(dev)], eta=1e-5)
    assert (plain[2][0].cpu() - ref[2][0]).abs().max().item() > 1e-6
    with pytest.raises(ValueError, match="batch size 2"):
        recursive_gf(E.to(dev), to(hl), to(hd), to(hu), to(sd), to(su), to(sl), sL.to(dev),
                     sR.to(dev), seP=[seP[0][:2].to(dev), 0, 0], eta=1e-5)
