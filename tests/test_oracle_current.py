"""Bond currents and the non-self-consistent I-V (``device_property.py:333-372,
472-503``): the oracle restatements pinned by current conservation, the
Landauer identity and a closed-form integral; the package's host-side I-V
against the oracle (no device work)."""
import numpy as np
import pytest
import torch

from oracle import device as odev, rgf as orgf


def _device(seed=0, norbs=(2, 2, 3, 2), E=0.3):
    """One-block device of four atoms; the left lead couples to atom 0 and the
    right lead to the last atom (complex symmetric self-energies with a
    negative-definite imaginary part).  Returns H, G, G^n with f_L = 1,
    f_R = 0, and T."""
    rng = np.random.default_rng(seed)
    n = sum(norbs)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = torch.tensor(0.5 * (a + a.conj().T) * 0.4)
    nL, nR = norbs[0], norbs[-1]

    def sig(m):
        b = rng.normal(size=(m, m))
        return torch.tensor(0.2 * (b + b.T) - 0.3j * (np.eye(m) + 0.2 * np.ones((m, m))))
    sL, sR = sig(nL), sig(nR)
    gam = lambda s: 1j * (s - s.mH)
    s_in = torch.zeros(n, n, dtype=torch.complex128)
    s_in[:nL, :nL] = gam(sL)  # f_L = 1, f_R = 0
    ans = orgf.recursive_gf(torch.tensor(E, dtype=torch.complex128), [], [H],
                            [], [torch.eye(n, dtype=torch.complex128)], [], [], sL, sR, eta=0.0,
                            s_in=[s_in], need_lesser=True)
    G, Gn = ans[2][0], ans[6][0]
    GL = torch.zeros(n, n, dtype=torch.complex128)
    GR = torch.zeros(n, n, dtype=torch.complex128)
    GL[:nL, :nL], GR[-nR:, -nR:] = gam(sL), gam(sR)
    T = torch.trace(GL @ G @ GR @ G.mH).real.item()
    return H, G, Gn, T


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_local_current_conserved_and_equal_to_landauer(seed):
    norbs = (2, 2, 3, 2)
    H, G, Gn, T = _device(seed, norbs)
    J = odev.local_current(H, Gn, norbs)
    assert (J + J.T).abs().max() < 1e-14 and J.diagonal().abs().max() == 0
    out = J.sum(1)
    # interior atoms: what flows in flows out
    assert out[1:-1].abs().max() < 1e-12 * T, out
    # the current leaving the left contact atom and entering the right one is T (f_L - f_R)
    assert abs(abs(out[0].item()) - T) < 1e-12 * T and abs(out[0] + out[-1]) < 1e-12 * T
    # a plausible slip, keeping only the first trace term, halves every bond
    # current (for Hermitian H and G^n the two terms are equal): the Landauer
    # identity catches it
    na = len(norbs)
    acc = np.cumsum((0,) + norbs)
    half = torch.zeros(na, na, dtype=torch.float64)
    for i in range(na):
        for j in range(na):
            if i != j:
                blk = H[acc[i]:acc[i + 1], acc[j]:acc[j + 1]] @ (1j * Gn[acc[j]:acc[j + 1],
                                                                       acc[i]:acc[i + 1]])
                half[i, j] = blk.real.diag().sum()
    assert abs(abs(half.sum(1)[0].item()) - T) > 0.4 * T


def test_current_nscf_closed_form():
    """T(E) = 1 on the whole grid: each current is the window width, mu_1 -
    mu_2 = 2 dv = vv, up to the Fermi tails the grid cuts off."""
    kbt = 0.025852
    eg = torch.linspace(-2.0, 2.0, 8001, dtype=torch.float64)
    vv, cc = odev.current_nscf(eg, torch.ones_like(eg), 0.3, kbt)
    assert len(vv) == len(cc) == int((2.0 - 4 * kbt) / 0.1) + 1  # vmid = half the grid - 4 kT
    for v, c in zip(vv, cc):
        if v / 2 < 2.0 - 4 * kbt - 0.4:  # both Fermi steps at least 0.4 eV inside the grid
            assert abs(c - float(v)) < 1e-6, (float(v), c)


def test_package_current_nscf_matches_oracle():
    from dpnegf_amd.negf.device_property import DeviceProperty

    class _Lead:
        efermi = 0.3
    dp = DeviceProperty(None, None, None, e_T=300, efermi=0.3, rgf_device="cpu")
    dp.set_leadLR(_Lead(), _Lead())
    eg = torch.linspace(-1.5, 1.7, 1601, dtype=torch.float64)
    tc = 1.0 + 0.5 * torch.sin(3 * eg) ** 2
    vv, cc = dp._cal_current_nscf_(eg, tc)
    rv, rc = odev.current_nscf(eg, tc, 0.3, dp.kBT)
    assert torch.equal(vv, rv) and np.allclose(cc, rc, rtol=0, atol=1e-14)
