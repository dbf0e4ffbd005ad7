"""Hand-written adjoint of the transmission vs torch autograd through the oracle
restatement of ``recursive_gf`` and vs central finite differences (delta = 1e-6 eV).

Objective J = sum_b w_b T(E_b) (a current-like weighted sum).  Bar: gradient
relative error <= 1e-5 against finite differences, <= 1e-9 against autograd.
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


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _sigmas(s, E, eta):
    out = []
    for l in (s.lead_L, s.lead_R):
        out.append(torch.stack([osg.selfEnergy(l[0], l[1], l[3], l[4], _f64(e), hDL=l[2],
                                               sDL=l[5], etaLead=eta)[0] for e in E]))
    return out


def _oracle_J(s, E, sL, sR, eta, w, hd, hu, hl):
    J = 0
    for b, e in enumerate(E):
        z = torch.tensor(e, dtype=torch.complex128)
        ans = orgf.recursive_gf(z, hl, hd, hu, s.sd, s.su, s.sl, sL[b], sR[b], eta=eta,
                                keep_gr_left=False)
        J = J + w[b] * odev.transmission(ans[0], olead.sigma2gamma(sL[b]), olead.sigma2gamma(sR[b]))
    return J


@pytest.mark.parametrize("case", ["ribbon", "uneven"])
def test_adjoint_matches_autograd_and_fd(case):
    from dpnegf_amd.negf.adjoint import transmission
    dev = _dev()
    if case == "ribbon":
        s, eta = wl.ribbon(W=6, K=5, seed=2), 1e-4
        E = np.array([-1.3, 0.2, 0.9])
    else:
        s, eta = wl.multiorbital(widths=(2, 3, 2), K=4, norb=2, seed=5), 1e-3
        E = np.array([-0.8, 0.1, 0.5])
    print(f"case {case}", flush=True)
    sL, sR = _sigmas(s, E, eta)
    w = torch.tensor([0.7, -0.4, 1.1], dtype=torch.float64)

    # autograd through the oracle
    hd = [h.clone().requires_grad_(True) for h in s.hd]
    hu = [h.clone().requires_grad_(True) for h in s.hu]
    hl = [h.clone().requires_grad_(True) for h in s.hl]
    J = _oracle_J(s, E, sL, sR, eta, w, hd, hu, hl)
    ref = torch.autograd.grad(J, hd + hu + hl)

    # GPU forward + adjoint
    g_hd = [h.to(dev).requires_grad_(True) for h in s.hd]
    g_hu = [h.to(dev).requires_grad_(True) for h in s.hu]
    g_hl = [h.to(dev).requires_grad_(True) for h in s.hl]
    sdv = lambda xs: [x.to(dev) for x in xs]
    T = transmission(_f64(E), g_hd, g_hu, g_hl, sdv(s.sd), sdv(s.su), sdv(s.sl),
                     sL.to(dev), sR.to(dev), eta)
    Jg = (T * w.to(dev)).sum()
    Jv, Jgv = float(J.detach()), float(Jg.detach())
    assert abs(Jgv - Jv) < 1e-10 * max(1, abs(Jv))
    got = torch.autograd.grad(Jg, g_hd + g_hu + g_hl)
    scale = max(r.abs().max().item() for r in ref)
    for r, g in zip(ref, got):
        err = (r - g.cpu()).abs().max().item() / scale
        assert err < 1e-9, err

    # central finite differences on a few entries, real and imaginary directions
    delta = 1e-6
    rng = np.random.default_rng(0)
    params = g_hd + g_hu + g_hl
    for trial in range(6):
        p = int(rng.integers(len(params)))
        i = int(rng.integers(params[p].shape[0]))
        j = int(rng.integers(params[p].shape[1]))
        for direction in (1.0, 1j):
            vals = []
            for sgn in (1, -1):
                blocks = [x.detach().clone() for x in params]
                blocks[p][i, j] += sgn * delta * direction
                K = len(g_hd)
                Tp = transmission(_f64(E), blocks[:K], blocks[K:2 * K - 1],
                                  blocks[2 * K - 1:], sdv(s.sd), sdv(s.su), sdv(s.sl), sL.to(dev),
                                  sR.to(dev), eta)
                vals.append(float((Tp * w.to(dev)).sum()))
            fd = (vals[0] - vals[1]) / (2 * delta)
            an = got[p][i, j].real.item() if direction == 1.0 else got[p][i, j].imag.item()
            assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-3 * scale), (p, i, j, direction, fd, an)


@pytest.mark.parametrize("stride", [1, 2, 3, 5, "sqrt", 12])
def test_checkpointed_adjoint_matches_full_storage(stride):
    """Checkpointing every s blocks recomputes the sweep in segments; the
    gradients are those of the adjoint that keeps every block (uneven blocks,
    non-orthogonal S, 12 blocks, 4 energies), to 1e-12 relative."""
    from dpnegf_amd.negf.adjoint import transmission
    from tests.test_gpu_rgf import _rand_se
    dev = _dev()
    sizes = [6, 9, 7, 8, 6, 10, 7, 9, 8, 6, 7, 9]
    print(f"case checkpoint stride={stride} sizes={sizes}", flush=True)
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=3, overlap=True)
    B = 4
    sL, sR = _rand_se(B, sizes[0], 8).to(dev), _rand_se(B, sizes[-1], 9).to(dev)
    E = _f64([-0.9, -0.2, 0.3, 1.1])
    w = torch.tensor([0.5, -1.2, 0.8, 0.3], dtype=torch.float64, device=dev)
    to = lambda xs: [x.to(dev) for x in xs]
    grads, Ts = [], []
    for ck in (0, stride):
        params = [x.to(dev).requires_grad_(True) for x in hd + hu + hl]
        K = len(sizes)
        T = transmission(E, params[:K], params[K:2 * K - 1], params[2 * K - 1:], to(sd), to(su),
                         to(sl), sL, sR, 1e-4, checkpoint=ck)
        Ts.append(T.detach())
        grads.append(torch.autograd.grad((T * w).sum(), params))
    assert torch.equal(Ts[0], Ts[1])
    scale = max(g.abs().max().item() for g in grads[0])
    for a, b in zip(*grads):
        assert (a - b).abs().max().item() < 1e-12 * scale


def test_checkpoint_auto_keeps_every_block_when_it_fits():
    from dpnegf_amd.negf.adjoint import checkpoint_stride
    dev = _dev()
    sizes = [8] * 9
    assert checkpoint_stride(None, sizes, 4, dev) == 0
    assert checkpoint_stride(None, sizes, 4, dev, mem_fraction=0.0) == 3


@pytest.mark.parametrize("sizes", [[140, 96, 130], [300, 264]])
def test_adjoint_blocks_above_128_matches_autograd(sizes):
    """Blocks above 128 (the Schur-complement inverse, recursive above 256)
    in the adjoint sweep: gradients of sum_b w_b T(E_b) against autograd
    through the oracle, full storage and checkpointed."""
    from dpnegf_amd.negf.adjoint import transmission
    from tests.test_gpu_rgf import _rand_se
    dev = _dev()
    print(f"case adjoint sizes={sizes}", flush=True)
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=6, overlap=True)
    B, eta = 2, 1e-3
    sL, sR = _rand_se(B, sizes[0], 4), _rand_se(B, sizes[-1], 5)
    E = [-0.4, 0.35]
    w = torch.tensor([0.9, -0.6], dtype=torch.float64)
    p_hd = [h.clone().requires_grad_(True) for h in hd]
    p_hu = [h.clone().requires_grad_(True) for h in hu]
    p_hl = [h.clone().requires_grad_(True) for h in hl]
    J = 0
    for b, e in enumerate(E):
        z = torch.tensor(e, dtype=torch.complex128)
        ans = orgf.recursive_gf(z, p_hl, p_hd, p_hu, sd, su, sl, sL[b], sR[b], eta=eta,
                                keep_gr_left=False)
        J = J + w[b] * odev.transmission(ans[0], olead.sigma2gamma(sL[b]),
                                         olead.sigma2gamma(sR[b]))
    ref = torch.autograd.grad(J, p_hd + p_hu + p_hl)
    scale = max(r.abs().max().item() for r in ref)
    to = lambda xs: [x.to(dev) for x in xs]
    for ck in (0, 1):
        params = [x.to(dev).requires_grad_(True) for x in hd + hu + hl]
        K = len(sizes)
        T = transmission(_f64(E), params[:K], params[K:2 * K - 1], params[2 * K - 1:], to(sd),
                         to(su), to(sl), sL.to(dev), sR.to(dev), eta, checkpoint=ck)
        Jg = (T * w.to(dev)).sum()
        Jv = float(J.detach())
        assert abs(float(Jg.detach()) - Jv) < 1e-10 * max(1.0, abs(Jv))
        got = torch.autograd.grad(Jg, params)
        for r, g in zip(ref, got):
            err = (r - g.cpu()).abs().max().item() / scale
            assert err < 1e-9, (ck, err)


@pytest.mark.parametrize("sizes,nL,nR", [([6, 9, 7, 8, 10], 4, 7), ([8], 5, 6)])
def test_adjoint_self_energy_gradients(sizes, nL, nR):
    """dJ/dSigma_L and dJ/dSigma_R per energy (Sigma trimmed to its corner,
    smaller than the corner block; K = 1 puts both corners in one block) and
    dJ/dS for the overlap blocks from the adjoint, full storage and
    checkpointed, against autograd through the oracle (for Sigma the corner of
    the gradient of the zero-padded self-energy), to 1e-9 relative, and
    Sigma's against central finite differences (delta 1e-6) on a few entries.
    Sigma enters both G and Gamma = i (Sigma - Sigma^H); S only M = z S - H."""
    import torch.nn.functional as F
    from dpnegf_amd.negf.adjoint import transmission
    from tests.test_gpu_rgf import _rand_se
    dev = _dev()
    print(f"case sigma gradients sizes={sizes} nL={nL} nR={nR}", flush=True)
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=7, overlap=True)
    B, eta = 3, 1e-3
    E = [-0.7, 0.15, 0.6]
    w = torch.tensor([0.8, -0.5, 1.2], dtype=torch.float64)
    sL0, sR0 = _rand_se(B, nL, 11), _rand_se(B, nR, 12)
    n0, nK = sizes[0], sizes[-1]
    # autograd through the oracle, on the zero-padded corner self-energies
    pL, pR = sL0.clone().requires_grad_(True), sR0.clone().requires_grad_(True)
    fL = F.pad(pL, (0, n0 - nL, 0, n0 - nL))
    fR = F.pad(pR, (nK - nR, 0, nK - nR, 0))
    p_s = [x.clone().requires_grad_(True) for x in sd + su + sl]
    K = len(sizes)
    o_sd, o_su, o_sl = p_s[:K], p_s[K:2 * K - 1], p_s[2 * K - 1:]
    J = 0
    for b, e in enumerate(E):
        z = torch.tensor(e, dtype=torch.complex128)
        ans = orgf.recursive_gf(z, hl, hd, hu, o_sd, o_su, o_sl, fL[b], fR[b], eta=eta,
                                keep_gr_left=False)
        J = J + w[b] * odev.transmission(ans[0], olead.sigma2gamma(fL[b]),
                                         olead.sigma2gamma(fR[b]))
    rL, rR, *r_s = torch.autograd.grad(J, [pL, pR] + p_s)
    scale = max(rL.abs().max().item(), rR.abs().max().item())
    scale_s = max(r.abs().max().item() for r in r_s)
    to = lambda xs: [x.to(dev) for x in xs]
    for ck in (0, 1, 3):
        gL = sL0.to(dev).requires_grad_(True)
        gR = sR0.to(dev).requires_grad_(True)
        g_s = [x.to(dev).requires_grad_(True) for x in sd + su + sl]
        T = transmission(_f64(E), to(hd), to(hu), to(hl), g_s[:K], g_s[K:2 * K - 1],
                         g_s[2 * K - 1:], gL, gR, eta, checkpoint=ck)
        Jg = (T * w.to(dev)).sum()
        assert abs(float(Jg.detach()) - float(J.detach())) < 1e-10 * max(1.0, abs(float(J)))
        aL, aR, *a_s = torch.autograd.grad(Jg, [gL, gR] + g_s)
        for name, r, a in (("sigL", rL, aL), ("sigR", rR, aR)):
            err = (r - a.cpu()).abs().max().item() / scale
            assert err < 1e-9, (ck, name, err)
        for q, (r, a) in enumerate(zip(r_s, a_s)):
            err = (r - a.cpu()).abs().max().item() / scale_s
            assert err < 1e-9, (ck, "S block", q, err)
    # central finite differences on a few entries
    delta = 1e-6
    rng = np.random.default_rng(1)
    for trial in range(4):
        which = trial % 2
        base = (sL0, sR0)[which]
        b, i, j = (int(rng.integers(s)) for s in base.shape)
        for direction in (1.0, 1j):
            vals = []
            for sgn in (1, -1):
                pert = [sL0.clone(), sR0.clone()]
                pert[which][b, i, j] += sgn * delta * direction
                Tp = transmission(_f64(E), to(hd), to(hu), to(hl), to(sd), to(su), to(sl),
                                  pert[0].to(dev), pert[1].to(dev), eta)
                vals.append(float((Tp * w.to(dev)).sum()))
            fd = (vals[0] - vals[1]) / (2 * delta)
            g = (aL, aR)[which][b, i, j]
            an = g.real.item() if direction == 1.0 else g.imag.item()
            assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-3 * scale), (which, b, i, j, fd, an)
