"""Pins the oracle's Lopez-Sancho, self-energy, transmission and Ozaki restatements.

* closed-form surface Green's function of the 1-orbital chain;
* the independent GEP (transfer-matrix) route of ``surface_green.py:284-347``;
* a long finite chain (brute force) for a multi-orbital lead;
* transmission = number of open channels for clean leads/devices
  (two-orbital chain via root counting of det(H(k) - E), ribbon via modes);
* Ozaki poles/residues: first pole ~ pi, the Fermi-function expansion identity;
* planted mistakes (wrong lead orientation, missing Sigma) must fail.
"""
import numpy as np
import pytest
import torch

from oracle import device as odev, lead as olead, ozaki as ooz, rgf, surface_green as osg
from dpnegf_amd import workloads as wl


def _chain_gs_closed(E, eta, eps, t):
    """Surface g of a 1-orbital semi-infinite chain, branch with Im g < 0."""
    w = E + 1j * eta - eps
    r = np.sqrt(w - 2 * t + 0j) * np.sqrt(w + 2 * t + 0j)
    g1 = (w - r) / (2 * t * t)
    g2 = (w + r) / (2 * t * t)
    return g1 if g1.imag < 0 else g2


@pytest.mark.parametrize("E", [-2.1, -0.5, 0.0, 0.77, 1.9, 3.5])
def test_lopez_sancho_closed_form(E):
    eps, t, eta = 0.1, -1.0, 1e-5
    H = np.array([[eps]], complex)
    h01 = np.array([[t]], complex)
    S = np.eye(1, dtype=complex)
    s01 = np.zeros((1, 1), complex)
    gs, it, res = osg.lopez_sancho(H, h01, S, s01, E + 1j * eta)
    ref = _chain_gs_closed(E, eta, eps, t)
    assert abs(gs[0, 0] - ref) < 1e-9 * max(1.0, abs(ref))
    assert 1 < it <= 100 and res < 1e-8


def _lead_blocks(seed, n):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = 0.5 * (a + a.conj().T)
    h01 = 0.6 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    return H, h01


@pytest.mark.parametrize("E", [-1.3, 0.2, 2.4])
def test_lopez_sancho_vs_gep_and_finite_chain(E):
    n = 4
    H, h01 = _lead_blocks(1, n)
    S = np.eye(n, dtype=complex)
    s01 = np.zeros((n, n), complex)
    eta = 1e-3
    z = E + 1j * eta
    gs, it, res = osg.lopez_sancho(H, h01, S, s01, z)
    gep = osg.calcg0(z, H, S, h01, s01)
    assert np.abs(gs - gep).max() < 1e-8 * np.abs(gs).max()
    # finite chain of 4000 layers: the surface block converges to g_s for eta > 0
    # (decay length ~ v/eta, so use a larger eta for this brute-force leg)
    z2 = E + 2e-2j
    gs2 = osg.lopez_sancho(H, h01, S, s01, z2)[0]
    g = np.linalg.inv(z2 * S - H)
    for _ in range(4000):
        g = np.linalg.inv(z2 * S - H - h01 @ g @ h01.conj().T)
    assert np.abs(gs2 - g).max() < 1e-8 * np.abs(gs2).max()


def _sigmas(s, E, eta_lead):
    out = []
    for lead in (s.lead_L, s.lead_R):
        hL, hLL, hDL, sL, sLL, sDL = lead
        out.append(osg.selfEnergy(hL, hLL, sL, sLL, E, hDL=hDL, sDL=sDL, etaLead=eta_lead)[0])
    return out


def _T(s, E, eta_lead, eta_dev):
    sL, sR = _sigmas(s, E, eta_lead)
    ans = rgf.recursive_gf(torch.tensor(E, dtype=torch.complex128), s.hl, s.hd, s.hu, s.sd,
                           s.su, s.sl, sL, sR, eta=eta_dev)
    return float(odev.transmission(ans[0], olead.sigma2gamma(sL), olead.sigma2gamma(sR)))


def _chain2_channels(E):
    H00 = np.array(wl.CHAIN2_H00)
    H01 = np.array(wl.CHAIN2_H01)
    ks = np.linspace(1e-7, np.pi - 1e-7, 40001)
    ev = np.array([np.linalg.eigvalsh(H00 + H01 * np.exp(1j * k) + H01.T * np.exp(-1j * k))
                   for k in ks])
    return int(sum(np.sum(np.diff(np.sign(ev[:, b] - E)) != 0) for b in range(2)))


@pytest.mark.parametrize("E", [-2.5, -1.0, 0.0, 0.3, 1.7, 2.9])
def test_two_orbital_chain_channel_count(E):
    s = wl.two_orbital_chain(16)
    assert abs(_T(s, E, 1e-5, 1e-5) - _chain2_channels(E)) < 2e-3


@pytest.mark.parametrize("E", [-3.0, -1.1, 0.4, 2.2])
def test_clean_ribbon_mode_count(E):
    s = wl.ribbon(W=8, K=6, disorder=0.0)
    assert abs(_T(s, E, 1e-6, 1e-6) - wl.ribbon_open_channels(E, 8)) < 1e-3


def test_planted_wrong_lead_orientation_fails():
    s = wl.ribbon(W=6, K=4, disorder=0.0)
    E = 0.4
    good = _T(s, E, 1e-6, 1e-6)
    # swap the left lead's coupling orientation: hLL -> hLL^dagger is harmless for a
    # symmetric ribbon, so plant the error on the 2-orbital chain instead
    c = wl.two_orbital_chain(6)
    hL, hLL, hDL, sL, sLL, sDL = c.lead_L
    c.lead_L = (hL, hLL.mH.clone() * 0.5, hDL, sL, sLL, sDL)
    bad = _T(c, 0.3, 1e-5, 1e-5)
    assert abs(good - wl.ribbon_open_channels(E, 6)) < 1e-3
    assert abs(bad - _chain2_channels(0.3)) > 1e-2


# Energies of the 4096-point grid linspace(-3, 3, 4096) where the W=48 ribbon lead's
# decimation is ill-conditioned at eta = 1e-6 (grid indices 1550 and 1352).
EDGE_E = [float(np.linspace(-3, 3, 4096)[i]) for i in (1550, 1352)]


@pytest.mark.parametrize("E", EDGE_E)
def test_lopez_sancho_association_near_band_edge(E):
    """The reference solves W X = [alpha|beta] (surface_green.py:189-194).  That
    association, and a refined inverse, keep the Dyson residual near 1e-7 here,
    while W^-1 times [alpha|beta] fails the reference's own 3e-5 check: the
    kernel must solve, not invert and multiply."""
    s = wl.ribbon(W=48, K=2, disorder=0.0)
    H, h01, _, S, s01, _ = [x.numpy() for x in s.lead_L]
    z = E + 1e-6j
    g_solve, it_s, r_solve = osg.lopez_sancho(H, h01, S, s01, z, variant="solve")
    g_ref, it_r, r_ref = osg.lopez_sancho(H, h01, S, s01, z, variant="refined")
    assert r_solve < 1e-6 and r_ref < 1e-6 and it_s == it_r == 29
    assert np.abs(g_solve - g_ref).max() < 1e-4 * np.abs(g_solve).max()
    with pytest.raises(ArithmeticError):
        osg.lopez_sancho(H, h01, S, s01, z, variant="inverse")


def test_ozaki_poles_and_fermi_identity():
    poles, res = ooz.ozaki_residues(M_cut=100)
    poles, res = poles.numpy(), res.numpy()
    assert np.all(np.diff(poles) > 0)
    assert abs(poles[0] - np.pi) < 1e-10          # first Matsubara pole
    assert abs(poles[1] - 3 * np.pi) < 1e-8
    assert abs(res[0] - 1.0) < 1e-10
    # 1/(1+e^x) = 1/2 - sum_p R_p 2x/(x^2 + zeta_p^2) for |x| well inside the cut
    for x in [-20.0, -3.0, -0.5, 0.0, 0.7, 4.0, 25.0]:
        approx = 0.5 - np.sum(res * 2 * x / (x * x + poles ** 2))
        assert abs(approx - 1.0 / (1.0 + np.exp(x))) < 1e-12


def test_gauss_xw_integrates_polynomials():
    xs, w = ooz.gauss_xw(torch.scalar_tensor(-0.3, dtype=torch.float64),
                         torch.scalar_tensor(1.7, dtype=torch.float64), n=10)
    for p in range(0, 19):
        exact = (1.7 ** (p + 1) - (-0.3) ** (p + 1)) / (p + 1)
        assert abs(float((w * xs ** p).sum()) - exact) < 1e-12 * max(1, abs(exact))


@pytest.mark.parametrize("m,K,overlap", [(2, 0.0, False), (3, 0.3, True), (4, 0.71, False)])
def test_bloch_unfolded_self_energy_equals_the_supercell(m, K, overlap):
    """Bloch's theorem pins the restated useBloch branch
    (``lead_property.py:336-378``): for a lead whose layer is m copies of a
    small cell, the average of kron(P(k'), g_s(k')) over the m unfolded points
    k' = (K + j)/m is the supercell's own surface Green's function at K, so its
    self-energy equals the direct decimation of the supercell.  The supercell is
    stored orbital-major, which also exercises ``bloch_sorted_indice``."""
    lead = wl.BlochLead(m=m, n=3, seed=10 + m, overlap=overlap, a_dev=5)
    kpt = [0.0, K, 0.0]
    kb = olead.bloch_unfold(lead.bloch_factor, kpt)
    assert kb.shape == (m, 3) and np.allclose(kb[:, 1], (K + np.arange(m)) / m)
    eta = 1e-4
    big = lead.supercell(kpt)
    for E in (-0.7, 0.15, 0.9):
        e64 = torch.tensor(E, dtype=torch.float64)
        gs = [osg.selfEnergy(*lead.small(k), e64, etaLead=eta)[1] for k in kb]
        sig = olead.bloch_self_energy(gs, kb, lead.R_list, lead.sorted_idx,
                                      torch.as_tensor(lead.hdl), torch.as_tensor(lead.sdl), E)
        ref = osg.selfEnergy(*big, e64, hDL=lead.hdl, sDL=lead.sdl, etaLead=eta)[0]
        err = (sig - ref).abs().max().item() / ref.abs().max().item()
        assert err < 1e-10, (E, err)
        # a plausible slip, the phase with R_i - R_j, breaks the identity
        bad = olead.bloch_self_energy(gs, -kb, lead.R_list, lead.sorted_idx,
                                      torch.as_tensor(lead.hdl), torch.as_tensor(lead.sdl), E)
        if K != 0.0 or m > 2:
            assert (bad - ref).abs().max().item() > 1e-4 * ref.abs().max().item()
