"""Pins the CPU oracle's RGF restatement before it is trusted as a checker.

* dense brute-force inverse of the assembled device matrix (every retarded
  output block, uneven blocks, non-orthogonal S, complex energies);
* lesser G^< = G Sigma^< G^dagger for Hermitian H, S_offdiag = 0;
* the reference's batched == scalar contract
  (``dpnegf/tests/test_recursive_gf_batched.py:104-197``);
* the reference's 4-site carbon chain golden values
  (``dpnegf/tests/test_negf_device_property.py:103-186``);
* planted-mistake checks: a dropped term / transposed operand in a copy of the
  recursion must fail the dense pin.
"""
import math

import numpy as np
import pytest
import torch

from oracle import dense, device as odev, lead as olead, rgf, surface_green as osg
from dpnegf_amd import workloads as wl


def _rand_se(B, n, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(B, n, n, generator=g, dtype=torch.float64) + 1j * torch.randn(
        B, n, n, generator=g, dtype=torch.float64)
    # retarded-like: negative-definite anti-Hermitian part keeps the problem sane
    return 0.3 * a - 0.5j * torch.eye(n, dtype=torch.complex128)


@pytest.mark.parametrize("sizes,overlap", [([3, 5, 4, 2], False), ([4, 4, 4], True),
                                           ([6], False), ([2, 7], True)])
def test_rgf_matches_dense_inverse(sizes, overlap):
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=3, overlap=overlap)
    B = 3
    sL = _rand_se(B, sizes[0], 1)
    sR = _rand_se(B, sizes[-1], 2)
    E = torch.tensor([-0.7 + 0.01j, 0.2 + 0.3j, 1.1 + 0.0j], dtype=torch.complex128)
    eta = 1e-3
    ans = rgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=eta, need_gr_lc=True)
    g_trans, gr_lc, grd, grl, gru, gr_left = ans[:6]
    K = len(sizes)
    for b in range(B):
        G, off = dense.dense_green(complex(E[b]) + 1j * eta, [h.numpy() for h in hd],
                                   [h.numpy() for h in hu], [h.numpy() for h in hl],
                                   [s.numpy() for s in sd], [s.numpy() for s in su],
                                   [s.numpy() for s in sl], sL[b].numpy(), sR[b].numpy())
        blk = lambda i, j: G[off[i]:off[i + 1], off[j]:off[j + 1]]
        scale = np.abs(G).max()
        tol = 1e-12 * scale
        assert np.abs(g_trans[b].numpy() - blk(0, K - 1)).max() < tol
        for q in range(K):
            assert np.abs(grd[q][b].numpy() - blk(q, q)).max() < tol
            assert np.abs(gr_lc[q][b].numpy() - blk(q, K - 1)).max() < tol
        for q in range(K - 1):
            assert np.abs(grl[q][b].numpy() - blk(q + 1, q)).max() < tol
            assert np.abs(gru[q][b].numpy() - blk(q, q + 1)).max() < tol
    if K == 1:
        assert grl == [] and gru == []


def test_rgf_lesser_matches_dense():
    sizes = [3, 4, 3]
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=5)
    B = 2
    sL = _rand_se(B, 3, 7)
    sR = _rand_se(B, 3, 8)
    E = torch.tensor([-0.3, 0.6], dtype=torch.complex128)
    f = torch.tensor([0.8, 0.3], dtype=torch.float64).view(B, 1, 1)
    s_in = odev.build_s_in([(n, n) for n in sizes], olead.sigma2gamma(sL) * f,
                           olead.sigma2gamma(sR) * 0.5)
    eta = 1e-4
    ans = rgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=eta, s_in=s_in,
                           need_lesser=True)
    gnd, gnl, gnu = ans[6], ans[7], ans[8]
    for b in range(B):
        G, off = dense.dense_green(complex(E[b]) + 1j * eta, [h.numpy() for h in hd],
                                   [h.numpy() for h in hu], [h.numpy() for h in hl],
                                   [s.numpy() for s in sd], [s.numpy() for s in su],
                                   [s.numpy() for s in sl], sL[b].numpy(), sR[b].numpy())
        N = G.shape[0]
        Sin = np.zeros((N, N), complex)
        Sin[:3, :3] = s_in[0][b].numpy()
        Sin[-3:, -3:] = s_in[-1][b].numpy()
        Gn = G @ Sin @ G.conj().T
        blk = lambda i, j: Gn[off[i]:off[i + 1], off[j]:off[j + 1]]
        tol = 1e-11 * np.abs(Gn).max()
        for q in range(3):
            assert np.abs(gnd[q][b].numpy() - blk(q, q)).max() < tol
        for q in range(2):
            assert np.abs(gnl[q][b].numpy() - blk(q + 1, q)).max() < tol
            assert np.abs(gnu[q][b].resolve_conj().numpy() - blk(q, q + 1)).max() < tol


def test_batched_equals_scalar():
    sizes = [8, 6, 8]
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=0)
    B = 6
    sL = _rand_se(B, 8, 11)
    sR = _rand_se(B, 8, 12)
    E = torch.linspace(-1, 1, B, dtype=torch.float64) + 1e-4j
    batched = rgf.recursive_gf(E, hl, hd, hu, sd, su, sl, sL, sR, eta=1e-5, need_gr_lc=True)
    for b in range(B):
        one = rgf.recursive_gf(E[b], hl, hd, hu, sd, su, sl, sL[b], sR[b], eta=1e-5,
                               need_gr_lc=True)
        assert one[0].ndim == 2
        assert torch.allclose(one[0], batched[0][b], atol=1e-10)
        for q in range(3):
            assert torch.allclose(one[2][q], batched[2][q][b], atol=1e-10)


# ---- the reference's 4-site chain KAT -------------------------------------------
# Values from dpnegf/tests/test_negf_device_property.py:103-186.  The model there
# (DeePTB nnsk carbon s-chain) is a uniform 1-orbital chain evaluated at its band
# centre; |t| follows from the fixture's own Sigma_L = -i|t| and 1/(2|t|) = 0.82022,
# the sign of t from the phase of g_trans[0,1].
KAT_T = 0.609597  # inside the window (0.609595, 0.609600) that satisfies both fixtures at 1e-5
KAT_SIGMA = -3.3171e-07 - 0.6096j
KAT_GTRANS_ROW0 = [1.0983e-11 - 8.2022e-01j, -8.2022e-01 + 4.4634e-07j,
                   8.9264e-07 + 8.2022e-01j, 8.2022e-01 - 1.3390e-06j]
KAT_GND_ROW0 = [8.2022e-01 + 0.0j, -4.4634e-07 + 2.2204e-16j, -8.2022e-01 - 3.1764e-22j,
                1.3390e-06 - 5.5511e-17j]


def _kat_chain():
    t = KAT_T
    h = np.zeros((4, 4))
    for i in range(3):
        h[i, i + 1] = h[i + 1, i] = t
    c = np.zeros((4, 4))
    c[3, 0] = t          # cell i -> cell i+1 : last site couples to first
    return wl.uniform_chain(h, c, 1)


def test_chain_golden_values():
    s = _kat_chain()
    E = torch.tensor(0.0, dtype=torch.complex128)
    ses = []
    for lead in (s.lead_L, s.lead_R):
        hL, hLL, hDL, sL_, sLL, sDL = lead
        se, _ = osg.selfEnergy(hL, hLL, sL_, sLL, E, hDL=hDL, sDL=sDL, etaLead=1e-5)
        ses.append(se)
    seL, seR = ses
    assert abs(seL[0, 0] - KAT_SIGMA) < 1e-5 and abs(seR[3, 3] - KAT_SIGMA) < 1e-5
    assert seL.abs()[1:, :].max() < 1e-12 and seR.abs()[:3, :].max() < 1e-12
    gamL, gamR = olead.sigma2gamma(seL), olead.sigma2gamma(seR)
    f = 1.0 / (1.0 + math.exp(0.0))
    s_in = [(gamL * f + gamR * f)]
    ans = rgf.recursive_gf(E, s.hl, s.hd, s.hu, s.sd, s.su, s.sl, seL, seR, eta=0.0,
                           s_in=s_in, need_lesser=True)
    g_trans, grd, grl, gru, gr_left = ans[0], ans[2], ans[3], ans[4], ans[5]
    assert (g_trans[0] - torch.tensor(KAT_GTRANS_ROW0)).abs().max() < 1e-5
    assert (grd[0] - g_trans).abs().max() < 1e-12 and grl == [] and gru == []
    assert (gr_left[0] - g_trans).abs().max() < 1e-12
    # gnd golden row; the reference built s_in with f(E) at 300 K, E = mu -> f = 1/2
    gnd = ans[6]
    assert (gnd[0][0] - torch.tensor(KAT_GND_ROW0)).abs().max() < 1e-5
    assert ans[7] == [] and ans[8] == []
    assert ans[10] is None and ans[13] is None
    T = odev.transmission(g_trans, gamL, gamR)
    assert abs(float(T) - 1.0) < 1e-5
    dos = odev.dos(grd, grl, gru, s.sd, s.su, s.sl)
    assert abs(float(dos) - 2.0887) < 1e-4
    ldos = odev.ldos(grd, grl, gru, s.sd, s.su, s.sl, [1, 1, 1, 1])
    assert (ldos - 0.2611 * 2).abs().max() < 1e-4


# ---- planted mistakes must fail the dense pin ------------------------------------
def _mutant_grd(E, hd, hu, hl, sd, su, sl, sL, sR, eta, mutation):
    z = E + 1j * eta
    M = [hd[q] - z * sd[q] for q in range(len(hd))]
    M[0] = M[0] + sL
    M[-1] = M[-1] + sR
    L = [hl[q] - z * sl[q] for q in range(len(hl))]
    U = [hu[q] - z * su[q] for q in range(len(hu))]
    gl = [torch.linalg.inv(-M[0])]
    for q in range(len(hd) - 1):
        coup = L[q] @ gl[q] @ U[q]
        if mutation == "transpose":
            coup = U[q].T @ gl[q] @ L[q].T
        if mutation == "drop":
            coup = 0 * coup
        if mutation == "sign":
            coup = -coup
        gl.append(torch.linalg.inv(-M[q + 1] - coup))
    return gl[-1]


@pytest.mark.parametrize("mutation", ["transpose", "drop", "sign"])
def test_planted_mistakes_fail_dense(mutation):
    sizes = [3, 4, 3]
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=9, hermitian_offdiag=False)
    sL = _rand_se(1, 3, 1)[0]
    sR = _rand_se(1, 3, 2)[0]
    E, eta = 0.3, 1e-3
    G, off = dense.dense_green(E + 1j * eta, [h.numpy() for h in hd], [h.numpy() for h in hu],
                               [h.numpy() for h in hl], [s.numpy() for s in sd],
                               [s.numpy() for s in su], [s.numpy() for s in sl],
                               sL.numpy(), sR.numpy())
    good = _mutant_grd(E, hd, hu, hl, sd, su, sl, sL, sR, eta, None)
    bad = _mutant_grd(E, hd, hu, hl, sd, su, sl, sL, sR, eta, mutation)
    ref = G[off[2]:off[3], off[2]:off[3]]
    assert np.abs(good.numpy() - ref).max() < 1e-11
    assert np.abs(bad.numpy() - ref).max() > 1e-4
