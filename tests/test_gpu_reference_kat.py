"""The reference's own known-answer values for the 4-site chain at E = 0
(dpnegf/tests/test_negf_device_property.py:103-186) through the HIP path,
checked directly against the fixture at its printed precision (1e-5):
Sigma_L / Sigma_R from the decimation kernel, g_trans, grd, gr_left and the
lesser gnd from the full-output RGF, and T from the fused transmission sweep.
The same inputs pin the oracle in tests/test_oracle_rgf.py."""
import numpy as np
import pytest
import torch

from tests.test_oracle_rgf import KAT_GND_ROW0, KAT_GTRANS_ROW0, KAT_SIGMA, _kat_chain

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def test_chain_golden_values_on_gpu():
    from dpnegf_amd.negf.device_property import transmission_pipeline
    from dpnegf_amd.negf.recursive_green_cal import recursive_gf
    from dpnegf_amd.negf.surface_green import selfEnergy
    dev = _dev()
    print("case reference 4-site chain KAT", flush=True)
    s = _kat_chain()
    E = torch.tensor(0.0, dtype=torch.complex128)
    seL, seR = [selfEnergy(l[0], l[1], l[3], l[4], E, hDL=l[2], sDL=l[5], etaLead=1e-5)[0]
                for l in (s.lead_L, s.lead_R)]
    assert seL.is_cuda and seR.is_cuda
    assert abs(complex(seL[0, 0]) - KAT_SIGMA) < 1e-5 and abs(complex(seR[3, 3]) - KAT_SIGMA) < 1e-5
    assert seL.abs()[1:, :].max() < 1e-12 and seR.abs()[:3, :].max() < 1e-12

    gam = lambda x: 1j * (x - x.conj().T)
    # the fixture's lesser input: s_in = (Gamma_L + Gamma_R) f with f(E = mu) = 1/2
    s_in = [(gam(seL) + gam(seR)) * 0.5]
    to = lambda xs: [x.to(dev) for x in xs]
    ans = recursive_gf(E.to(dev), to(s.hl), to(s.hd), to(s.hu), to(s.sd), to(s.su), to(s.sl),
                       seL, seR, eta=0.0, s_in=s_in, need_lesser=True)
    g_trans, grd, grl, gru, gr_left, gnd = ans[0], ans[2], ans[3], ans[4], ans[5], ans[6]
    assert (g_trans[0].cpu() - torch.tensor(KAT_GTRANS_ROW0)).abs().max() < 1e-5
    assert (grd[0] - g_trans).abs().max() < 1e-12 and grl == [] and gru == []
    assert (gr_left[0] - g_trans).abs().max() < 1e-12
    assert (gnd[0][0].cpu() - torch.tensor(KAT_GND_ROW0)).abs().max() < 1e-5
    # T = Re Tr[Gamma_L G Gamma_R G^H] = 1 in the band: directly, and from the fused sweep
    T_direct = torch.trace(gam(seL) @ g_trans @ gam(seR) @ g_trans.conj().T).real
    assert abs(float(T_direct) - 1.0) < 1e-5
    T = transmission_pipeline(s.to(dev), torch.tensor(np.array([0.0])), eta_lead=1e-5,
                              eta_device=0.0)
    assert abs(float(T[0]) - 1.0) < 1e-5
