"""``DeviceProperty`` end to end on the GPU (``device_property.py:137-503``):
cal_green_function with the lesser function on a one-block device of four
atoms, then T, DOS and the bond currents against the oracle (1e-10), for one
energy and for a batch of two."""
import numpy as np
import pytest
import torch

from oracle import device as odev, rgf as orgf
from tests.test_oracle_current import _device

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


class _Ham:
    def __init__(self, H, norbs):
        self.H, self.device_norbs = H, list(norbs)

    def get_hs_device(self, kpoint, V, block_tridiagonal=False, **kw):
        n = self.H.shape[0]
        eye = torch.eye(n, dtype=torch.complex128)
        return [self.H - V.cpu() * eye], [eye], [], [], [], []


class _Lead:
    """Fixed self-energy; f = 1 (left) or 0 (right) over the whole window."""

    def __init__(self, se, mu, efermi=0.0):
        self.se, self.mu, self.efermi, self.kBT = se, mu, efermi, 0.025852

    def fermi_dirac(self, x):
        return 1 / (1 + torch.exp((x.real - self.mu) / self.kBT))

    @property
    def gamma(self):
        return 1j * (self.se - self.se.mH)


@pytest.mark.parametrize("energies", [[0.3], [0.3, -0.45]])
def test_device_property_currents_dos_and_t(energies):
    from dpnegf_amd.negf.device_property import DeviceProperty
    dev = _dev()
    norbs = (2, 2, 3, 2)
    print(f"case DeviceProperty energies={energies}", flush=True)
    refs = []
    for E in energies:
        H, G, Gn, T = _device(0, norbs, E)
        refs.append((odev.local_current(H, Gn, norbs), T, G))
    # the self-energies _device draws (same seed): rebuild them on the device
    rng = np.random.default_rng(0)
    n = sum(norbs)
    rng.normal(size=(n, n)), rng.normal(size=(n, n))

    def sig(m):
        b = rng.normal(size=(m, m))
        return torch.tensor(0.2 * (b + b.T) - 0.3j * (np.eye(m) + 0.2 * np.ones((m, m))))
    sL, sR = sig(norbs[0]), sig(norbs[-1])
    B = len(energies)
    batched = B > 1
    lead_L = _Lead((sL.expand(B, -1, -1) if batched else sL).to(dev), mu=50.0)
    lead_R = _Lead((sR.expand(B, -1, -1) if batched else sR).to(dev), mu=-50.0)
    dp = DeviceProperty(_Ham(H, norbs), None, None, e_T=300, efermi=0.0, rgf_device=dev)
    dp.set_leadLR(lead_L, lead_R)
    e = torch.tensor(energies if batched else energies[0], dtype=torch.float64)
    dp.cal_green_function(e, [0, 0, 0], eta_device=0.0, block_tridiagonal=False,
                          Vbias=torch.tensor(0.0), need_lesser=True)
    J = dp.lcurrent.cpu()
    tc = dp.tc.cpu().reshape(-1)
    dos = torch.as_tensor(dp.dos).cpu().reshape(-1)
    assert J.shape == ((B, 4, 4) if batched else (4, 4))
    for b, (Jr, T, G) in enumerate(refs):
        Jb = J[b] if batched else J
        assert (Jb - Jr).abs().max().item() < 1e-10 * T, (b, (Jb - Jr).abs().max().item())
        assert abs(tc[b].item() - T) < 1e-10 * T
        dos_ref = -torch.diagonal(G).imag.sum().item() / np.pi * 2
        assert abs(dos[b].item() - dos_ref) < 1e-10 * abs(dos_ref)
    assert dp.get_index(2) == [4, 7]


class _BTDHam:
    """Uneven blocks with overlap (S_offdiag != 0, so every term of the
    spectral diagonal enters), Hamiltonian-object interface."""

    def __init__(self, sizes, norbs, seed):
        from dpnegf_amd import workloads as wl
        self.blocks = wl.random_btd(sizes, seed=seed, overlap=True)
        self.device_norbs = list(norbs)

    def get_hs_device(self, kpoint, V, block_tridiagonal=True, **kw):
        hd, sd, hl, hu, sl, su = self.blocks
        return [list(hd), list(sd), list(hl), list(su), list(sl), list(hu)]


@pytest.mark.parametrize("energies", [[0.2], [-0.6, 0.15, 0.7]])
def test_block_tridiagonal_tc_dos_and_ldos(energies):
    """``cal_green_function`` on uneven blocks 5/7/4 with overlap, then
    ``_cal_tc_``, ``_cal_dos_`` and ``_cal_ldos_`` (``device_property.py:374-470``,
    through the library's product and diagonal kernels) against the oracle's
    recursive_gf blocks and formulas, 1e-10; atoms straddle block edges."""
    from dpnegf_amd.negf.device_property import DeviceProperty
    from tests.test_gpu_rgf import _rand_se
    dev = _dev()
    sizes, norbs = [5, 7, 4], (2, 4, 3, 3, 2, 2)
    print(f"case DeviceProperty BTD energies={energies}", flush=True)
    ham = _BTDHam(sizes, norbs, seed=21)
    B = len(energies)
    batched = B > 1
    sL, sR = _rand_se(B, 5, 22), _rand_se(B, 3, 23)  # right self-energy smaller than its block
    lead_L = _Lead((sL if batched else sL[0]).to(dev), mu=0.0)
    lead_R = _Lead((sR if batched else sR[0]).to(dev), mu=0.0)
    dp = DeviceProperty(ham, None, None, e_T=300, efermi=0.0, E_ref=0.05, rgf_device=dev)
    dp.set_leadLR(lead_L, lead_R)
    e = torch.tensor(energies if batched else energies[0], dtype=torch.float64)
    eta = 1e-4
    dp.cal_green_function(e, [0, 0, 0], eta_device=eta, block_tridiagonal=True,
                          Vbias=torch.tensor(0.0))
    hd, sd, hl, hu, sl, su = ham.blocks
    E = torch.tensor(energies, dtype=torch.complex128)
    ans = orgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=eta, E_ref=0.05)
    grd, grl, gru = ans[2], ans[3], ans[4]
    T_ref = odev.transmission(ans[0], 1j * (sL - sL.mH), 1j * (sR - sR.mH))
    dos_ref = odev.dos(grd, grl, gru, sd, su, sl)
    ldos_ref = odev.ldos(grd, grl, gru, sd, su, sl, norbs)
    tc, dos, ldos = dp.tc.cpu(), dp.dos.cpu(), dp.ldos.cpu()
    if not batched:
        assert tc.ndim == 0 and dos.ndim == 0 and ldos.shape == (len(norbs),)
        tc, dos, ldos = tc.reshape(1), dos.reshape(1), ldos.reshape(1, -1)
    assert ldos.shape == (B, len(norbs))
    assert (tc - T_ref).abs().max().item() < 1e-10 * T_ref.abs().max().item()
    assert (dos - dos_ref).abs().max().item() < 1e-10 * dos_ref.abs().max().item()
    assert (ldos - ldos_ref).abs().max().item() < 1e-10 * ldos_ref.abs().max().item()
    assert abs(ldos.sum(-1) - dos).max().item() < 1e-10 * dos.abs().max().item()
