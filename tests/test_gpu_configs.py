"""GPU parity at the shapes of the benchmark workloads (multi-orbital uneven
blocks, the adjoint ribbon, the Ozaki density device) and of the kernel
instantiations those workloads select, against the oracle.  Bars: 1e-10
relative for T and densities (FP64), 1e-9 for gradients against autograd
through the oracle, 1e-5 against central finite differences (delta 1e-6 eV)."""
import numpy as np
import pytest
import torch

from oracle import device as odev, lead as olead, ozaki as ooz, rgf as orgf, surface_green as osg
from dpnegf_amd import workloads as wl

pytestmark = pytest.mark.gpu
C = torch.complex128


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _f64(x):
    return torch.tensor(x, dtype=torch.float64)


def _sig(s, e, eta):
    return [osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e), hDL=l[2], sDL=l[5], etaLead=eta)[0]
            for l in (s.lead_L, s.lead_R)]


def _oracle_T(s, E, eta_lead, eta_dev):
    out = []
    for e in E:
        sL, sR = _sig(s, e, eta_lead)
        ans = orgf.recursive_gf(torch.tensor(e, dtype=C), s.hl, s.hd, s.hu, s.sd, s.su, s.sl,
                                sL, sR, eta=eta_dev, keep_gr_left=False)
        out.append(float(odev.transmission(ans[0], olead.sigma2gamma(sL), olead.sigma2gamma(sR))))
    return np.array(out)


def _check_T(systems, E, eta, rtol=1e-10):
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    T, st = transmission_pipeline([s.to(dev) for s in systems], _f64(E), eta_lead=eta,
                                  eta_device=eta, return_status=True)
    T = T.cpu().numpy()
    assert int(st.abs().sum()) == 0
    for k, s in enumerate(systems):
        ref = _oracle_T(s, E, eta, eta)
        err = np.abs(T[k] - ref).max() / max(1.0, np.abs(ref).max())
        print(f"  k-point {k}: T = {np.round(T[k], 6)}, max rel err {err:.2e}", flush=True)
        assert err < rtol, (k, T[k], ref)


def _golden():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                "config_shapes.npz"))


def test_multiorbital_96_to_168_blocks_two_kpoints():
    """The multi-orbital workload at its own shape: 3 orbitals per site, layer
    widths 32/40/48/56 (blocks 96/120/144/168), 40 layers, leads of 96 and 168
    orbitals; 2 transverse k-points x 4 energies in one call (batched
    decimation over both k-points with sys_idx, block-operation sweep),
    against oracle T stored by tests/golden/make_config_shapes.py."""
    from dpnegf_amd.negf.device_property import transmission_pipeline
    from tests.golden import make_config_shapes as mk
    dev = _dev()
    print("case multiorbital blocks 96-168 K=40, 2 k x 4 E", flush=True)
    systems = [wl.multiorbital(kpoint=k / 32, seed=0) for k in mk.MO_KPTS]
    assert sorted(set(systems[0].sizes)) == [96, 120, 144, 168] and len(systems[0].sizes) == 40
    T, st = transmission_pipeline([s.to(dev) for s in systems], _f64(mk.MO_E), mk.MO_ETA,
                                  mk.MO_ETA, return_status=True)
    assert int(st.abs().sum()) == 0
    g = _golden()
    for i, k in enumerate(mk.MO_KPTS):
        ref = g[f"mo_T_k{k}"]
        err = np.abs(T[i].cpu().numpy() - ref).max() / max(1.0, np.abs(ref).max())
        print(f"  k = {k}/32: T = {np.round(ref, 6)}, max rel err {err:.2e}", flush=True)
        assert err < 1e-10


@pytest.mark.parametrize("widths,K,kpts", [((8, 10, 12), 6, (0, 8, 13)),
                                           ((16,), 5, (0, 8, 13)),
                                           ((17, 21, 19), 5, (0, 11))])
def test_fused_kernel_several_kpoints(widths, K, kpts):
    """One fused launch over several k-point systems (sys_idx, nsys > 1):
    uneven blocks 24/30/36 (generic kernel, NMAX 48), uniform 48 (the
    compile-time-shaped kernel) and uneven 51/63/57 (NMAX 64; leads of 51 and
    57 orbitals go through the batched decimation)."""
    print(f"case fused multi-k widths={widths} K={K} k={kpts}", flush=True)
    systems = [wl.multiorbital(widths=widths, K=K, norb=3, kpoint=k / 32, seed=2) for k in kpts]
    _check_T(systems, np.array([-0.9, 0.2, 0.75]), 1e-5)


@pytest.mark.parametrize("W,K", [(32, 6), (64, 64)])
def test_fixed_kernels_32_and_64(W, K):
    """The compile-time-shaped kernel at NMAX 32 and at 64 (the adjoint
    workload's ribbon, W = 64, K = 64, eta = 1e-6)."""
    print(f"case ribbon W={W} K={K}", flush=True)
    s = wl.ribbon(W=W, K=K, seed=0)
    _check_T([s], np.array([-2.2, -0.75, 0.05, 1.4]), 1e-6)


def test_mixed_precision_at_48():
    """The FP32 instantiation at NMAX 48 (ribbon W = 48, 16 layers) stays within
    1e-3 of the FP64 path (FP32 rounding, no refinement)."""
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    s = wl.ribbon(W=48, K=16, seed=1).to(dev)
    E = _f64(np.linspace(-2.5, 2.5, 64))
    T64 = transmission_pipeline(s, E, 1e-4, 1e-4)
    T32 = transmission_pipeline(s, E, 1e-4, 1e-4, precision="mixed")
    err = (T64 - T32).abs().max().item()
    print(f"  FP32 vs FP64 max |dT| {err:.2e} (T up to {T64.max().item():.1f})", flush=True)
    assert err < 1e-3


def test_adjoint_ribbon64_K64_bias_window():
    """The adjoint workload's device (ribbon W = 64, K = 64, eta = 1e-6) with
    the (f_L - f_R) weights of a 0.2 V bias at 300 K on 4 energies inside the
    window: dJ/dH for every block against autograd through the oracle (1e-9)
    and central finite differences (1e-5); the sqrt(K)-checkpointed adjoint
    equals the full-storage one."""
    from dpnegf_amd.negf.adjoint import transmission
    dev = _dev()
    print("case adjoint ribbon64 K64, 4 energies", flush=True)
    s = wl.ribbon(W=64, K=64, seed=0)
    eta, kT, V = 1e-6, 0.025852, 0.2
    E = np.array([-0.11, -0.035, 0.02, 0.095])
    w = torch.tensor(1 / (1 + np.exp((E - V / 2) / kT)) - 1 / (1 + np.exp((E + V / 2) / kT)))
    sig = [_sig(s, e, eta) for e in E]
    sL = torch.stack([x[0] for x in sig])
    sR = torch.stack([x[1] for x in sig])
    K = len(s.hd)
    # autograd through the oracle
    hd = [h.clone().requires_grad_(True) for h in s.hd]
    hu = [h.clone().requires_grad_(True) for h in s.hu]
    hl = [h.clone().requires_grad_(True) for h in s.hl]
    J = 0
    for b, e in enumerate(E):
        ans = orgf.recursive_gf(torch.tensor(e, dtype=C), hl, hd, hu, s.sd, s.su, s.sl, sL[b],
                                sR[b], eta=eta, keep_gr_left=False)
        J = J + w[b] * odev.transmission(ans[0], olead.sigma2gamma(sL[b]),
                                         olead.sigma2gamma(sR[b]))
    ref = torch.autograd.grad(J, hd + hu + hl)
    scale = max(r.abs().max().item() for r in ref)
    dv = lambda xs: [x.to(dev) for x in xs]
    grads = {}
    for mode in (0, "sqrt"):
        params = [x.to(dev).requires_grad_(True) for x in s.hd + s.hu + s.hl]
        T = transmission(_f64(E), params[:K], params[K:2 * K - 1], params[2 * K - 1:], dv(s.sd),
                         dv(s.su), dv(s.sl), sL.to(dev), sR.to(dev), eta, checkpoint=mode)
        Jg = (T * w.to(dev)).sum()
        assert abs(float(Jg.detach()) - float(J.detach())) < 1e-10 * max(1.0, abs(float(J.detach())))
        grads[mode] = [g.cpu() for g in torch.autograd.grad(Jg, params)]
        err = max((r - g).abs().max().item() for r, g in zip(ref, grads[mode])) / scale
        print(f"  mode {mode}: J = {float(Jg):.6f}, max rel grad err vs autograd {err:.2e}",
              flush=True)
        assert err < 1e-9
    diff = max((a - b).abs().max().item() for a, b in zip(grads[0], grads["sqrt"])) / scale
    assert diff < 1e-12, diff
    # central finite differences on the GPU forward, real and imaginary directions
    delta = 1e-6
    rng = np.random.default_rng(1)
    base = [x.to(dev) for x in s.hd + s.hu + s.hl]
    for trial in range(4):
        p = int(rng.integers(len(base)))
        i, j = int(rng.integers(64)), int(rng.integers(64))
        for direction in (1.0, 1j):
            vals = []
            for sgn in (1, -1):
                blocks = [x.clone() for x in base]
                blocks[p][i, j] += sgn * delta * direction
                Tp = transmission(_f64(E), blocks[:K], blocks[K:2 * K - 1], blocks[2 * K - 1:],
                                  dv(s.sd), dv(s.su), dv(s.sl), sL.to(dev), sR.to(dev), eta)
                vals.append(float((Tp * w.to(dev)).sum()))
            fd = (vals[0] - vals[1]) / (2 * delta)
            g = grads[0][p][i, j]
            an = g.real.item() if direction == 1.0 else g.imag.item()
            assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-3 * scale), (p, i, j, direction, fd, an)


def _density_device():
    return wl.multiorbital(widths=(32, 24, 28, 30), K=24, norb=8, kpoint=0.0, seed=0)


def test_ozaki_density_blocks_256_pole_subset():
    """The density workload's device (8 orbitals per site, widths 32/24/28/30:
    blocks 256/192/224/240, 24 layers, leads of 256 and 240 orbitals): the R
    term, 3 Ozaki poles and a 16-node bias window (0.1 V, 300 K), every
    diagonal block (its diagonal, first row and last column), against the
    oracle values stored by tests/golden/make_config_shapes.py, at 1e-10."""
    from dpnegf_amd.negf.density import ozaki_density
    from tests.golden import make_config_shapes as mk
    dev = _dev()
    print("case density blocks 256/192/224/240 K=24, 3 poles + 16 GL", flush=True)
    s = _density_device()
    assert s.sizes[:4] == [256, 192, 224, 240] and len(s.sizes) == 24
    d = mk.DENS
    poles, res = ooz.ozaki_residues(d["M"])
    idx = list(d["pole_idx"])
    fL = lambda x: 1 / (1 + torch.exp((x - (d["mu"] - d["vL"])) / d["kBT"]))
    fR = lambda x: 1 / (1 + torch.exp((x - (d["mu"] - d["vR"])) / d["kBT"]))
    sd = s.to(dev)
    eq, neq = ozaki_density((sd.hd, sd.hu, sd.hl, sd.sd, sd.su, sd.sl), sd.lead_L, sd.lead_R,
                            d["kBT"], d["mu_L"], d["mu"], d["R"], poles[idx], res[idx], 0.0,
                            d["eta_lead"], d["eta_dev"], d["vL"], d["vR"], d["n_gauss"], fL, fR)
    g = _golden()
    for name, dm in (("eq", eq), ("neq", neq)):
        refs = [g[f"dens_{name}_{q}"] for q in range(len(dm))]
        got = [mk.entries(x.cpu().numpy()) for x in dm]
        scale = max(np.abs(r).max() for r in refs)
        err = max(np.abs(a - r).max() for a, r in zip(got, refs)) / scale
        print(f"  {name}: max rel err {err:.2e}", flush=True)
        assert err < 1e-10


def test_ozaki_density_blocks_256_all_poles_properties():
    """The full 160-pole equilibrium density of the same device (real
    symmetric H at k = 0): every diagonal block is real symmetric, each
    orbital's occupation 2 f lies in [0, 2], and 160 poles agree with 240."""
    from dpnegf_amd.negf.density import ozaki_density
    dev = _dev()
    print("case density blocks 256 K=24, 160 and 240 poles", flush=True)
    sd = _density_device().to(dev)
    kBT, R = 0.025852, 1e6
    dms = {}
    for M in (160, 240):
        poles, res = ooz.ozaki_residues(M)
        eq, _ = ozaki_density((sd.hd, sd.hu, sd.hl, sd.sd, sd.su, sd.sl), sd.lead_L, sd.lead_R,
                              kBT, 0.0, 0.0, R, poles, res, 0.0, 1e-5, 1e-6)
        dms[M] = [d.cpu() for d in eq]
    scale = max(d.abs().max().item() for d in dms[160])
    for d in dms[160]:
        assert (d - d.T).abs().max().item() < 1e-12 * scale
        diag = d.diagonal()
        assert diag.min().item() > -1e-9 and diag.max().item() < 2 + 1e-9
    conv = max((a - b).abs().max().item() for a, b in zip(dms[160], dms[240])) / scale
    n_el = sum(float(d.diagonal().sum()) for d in dms[160])
    print(f"  electrons {n_el:.6f} of {2 * sum(sd.sizes)} states x 2; 160 vs 240 poles "
          f"{conv:.2e}", flush=True)
    assert conv < 1e-9


def test_mixed_refined_meets_1e6_on_the_headline_device():
    """FP32 block inverses refined by one FP64 Newton step (FP64 products):
    on the headline device (ribbon W = 48, 128 layers, eta = 1e-6) at 512 of
    the 4096 grid energies, max |T_refined - T_fp64| <= 1e-6 (T up to about
    30), including the band-edge energies the lead decimation finds hardest."""
    from dpnegf_amd import _native as nat
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems, rgf_transmission
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    print("case mixed_refined ribbon48 K128, 512 energies", flush=True)
    s = wl.ribbon(W=48, K=128, seed=0).to(dev)
    grid = np.linspace(-3.0, 3.0, 4096)
    idx = np.unique(np.concatenate([np.arange(0, 4096, 8),
                                    [137, 452, 1036, 1352, 1550, 2545, 2743, 3643]]))
    E = torch.tensor(grid[idx], dtype=C, device=dev)
    eta = 1e-6
    sL = self_energy_batched([s.lead_L], E, eta, raise_on_error=False).sigma
    sR = self_energy_batched([s.lead_R], E, eta, raise_on_error=False).sigma
    bs = BTDSystems([(s.hd, s.hu, s.hl, s.sd, s.su, s.sl)], dev)
    z = (E + 1j * eta).contiguous()
    T64 = rgf_transmission(bs, z, sL, sR, precision="fp64")
    st = torch.zeros(E.shape[0], dtype=torch.int32, device=dev)
    Tr = rgf_transmission(bs, z, sL, sR, precision="mixed_refined", status=st)
    d = (Tr - T64).abs()
    nfb = int(((st & nat.STATUS_REFINE_FALLBACK) != 0).sum())
    print(f"  max |dT| {d.max().item():.2e}, FP64 fallback at {nfb} points", flush=True)
    assert d.max().item() <= 1e-6
    assert int((st & nat.STATUS_SINGULAR).sum()) == 0


def test_mixed_refined_generic_kernel_uneven_blocks():
    """The refined variant on the generic (uneven-block) kernel, blocks
    24/30/36, against the FP64 path at 1e-8 relative; blocks above 48 are
    refused."""
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    s = wl.multiorbital(widths=(8, 10, 12), K=6, norb=3, kpoint=0.1, seed=2).to(dev)
    E = _f64(np.linspace(-1.5, 1.5, 24))
    T64 = transmission_pipeline(s, E, 1e-5, 1e-5)
    Tr = transmission_pipeline(s, E, 1e-5, 1e-5, precision="mixed_refined")
    assert (Tr - T64).abs().max().item() < 1e-8 * max(1.0, T64.abs().max().item())
    big = wl.ribbon(W=52, K=3, seed=0).to(dev)
    with pytest.raises(Exception, match="refined mixed precision needs blocks"):
        transmission_pipeline(big, E[:2], 1e-5, 1e-5, precision="mixed_refined")
