"""Pins the oracle's Fiori free charge (``oracle/fiori.py``,
``density.py:234-393``) before it is trusted as a checker.

* Dense identity: at eta_device = 0, i (G - G^dagger) = G (Gamma_L + Gamma_R)
  G^dagger, so the reference's A_L = i (G - G^dagger) - A_R is G Gamma_L
  G^dagger; the per-atom charge from the dense inverse, G (Gamma_L f_L +
  Gamma_R f_R) G^dagger for electrons and G (Gamma_L (1 - f_L) + Gamma_R
  (1 - f_R)) G^dagger for holes, summed over each atom's global orbital range,
  must equal the oracle's block-by-block sum (atoms straddling block edges,
  both branches, non-orthogonal S).
* ``subblock_index`` against the global orbital range on random partitions.
* Planted mistake: dropping the "- A_R" fails the dense identity.
"""
import math

import numpy as np
import pytest
import torch

from oracle import dense, fiori, lead as olead
from dpnegf_amd import workloads as wl

C = torch.complex128


def _se(n, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(n, n, generator=g, dtype=torch.float64) + 1j * torch.randn(
        n, n, generator=g, dtype=torch.float64)
    return 0.2 * a - 0.4j * torch.eye(n, dtype=C)


def _dense_charge(points, weights, sL, sR, blocks, fL, fR, E_ref, pot, norbs):
    hd, hu, hl, sd, su, sl = blocks
    natom = len(norbs)
    acc = np.cumsum([0] + list(norbs))
    out = np.zeros(natom, complex)
    for w, e in zip(weights, points):
        e = float(e)
        z = e + E_ref
        G, off = dense.dense_green(z, [h.numpy() for h in hd], [h.numpy() for h in hu],
                                   [h.numpy() for h in hl], [s.numpy() for s in sd],
                                   [s.numpy() for s in su], [s.numpy() for s in sl],
                                   sL(e).numpy(), sR(e).numpy())
        N = G.shape[0]
        gL = np.zeros((N, N), complex)
        gR = np.zeros((N, N), complex)
        a, b = sL(e).shape[0], sR(e).shape[0]
        gL[:a, :a] = olead.sigma2gamma(sL(e)).numpy()
        gR[N - b:, N - b:] = olead.sigma2gamma(sR(e)).numpy()
        fl = float(fL(torch.tensor(e + E_ref, dtype=torch.float64)))
        fr = float(fR(torch.tensor(e + E_ref, dtype=torch.float64)))
        gn = G @ (gL * fl + gR * fr) @ G.conj().T
        gp = G @ (gL * (1 - fl) + gR * (1 - fr)) @ G.conj().T
        for atom in range(natom):
            sl_ = slice(acc[atom], acc[atom + 1])
            if e >= -pot[atom]:
                out[atom] += float(w) * (-1 / math.pi) * np.trace(gn[sl_, sl_])
            else:
                out[atom] += float(w) * (1 / math.pi) * np.trace(gp[sl_, sl_])
    return out


def _setup(sizes, overlap, seed):
    hd, sd, hl, hu, sl, su = wl.random_btd(sizes, seed=seed, overlap=overlap)
    blocks = (hd, hu, hl, sd, su, sl)
    SL, SR = _se(sizes[0], seed + 1), _se(sizes[-1], seed + 2)
    kT = 0.025852
    fL = lambda x: 1 / (1 + torch.exp((x - 0.1) / kT))
    fR = lambda x: 1 / (1 + torch.exp((x + 0.1) / kT))
    return blocks, (lambda e: SL * (1 + 0.1 * e)), (lambda e: SR * (1 - 0.05 * e)), fL, fR


@pytest.mark.parametrize("sizes,norbs,bt", [([4, 6, 5], (2, 3, 3, 2, 3, 2), True),
                                            ([4, 6, 5], (4, 7, 4), True),
                                            ([9], (3, 2, 4), False)])
@pytest.mark.parametrize("way", ["direct", "gauss"])
def test_fiori_matches_dense_inverse(sizes, norbs, bt, way):
    blocks, sL, sR, fL, fR = _setup(sizes, overlap=True, seed=sum(sizes))
    pts, w = fiori.integration_points(np.linspace(-0.6, 0.5, 12), way, n_gauss=9)
    natom = len(norbs)
    pot = np.linspace(-0.3, 0.4, natom)  # some atoms electrons, some holes at each energy
    got = fiori.free_charge(pts, w, sL, sR, blocks, fL, fR, 0.05, 0.0, pot, list(norbs), sizes,
                            block_tridiagonal=bt)
    ref = _dense_charge(pts, w, sL, sR, blocks, fL, fR, 0.05, pot, norbs)
    assert np.abs(got.numpy() - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(got.imag.numpy()).max() < 1e-12 * np.abs(ref).max()


def test_planted_missing_right_spectral_term_fails():
    """With A_L = i (G - G^dagger) (the "- A_R" dropped) the electron charge
    becomes A_L f_L + A_R (f_L + f_R): the oracle run with f_R replaced by
    f_L + f_R gives exactly that, and it misses the dense charge."""
    sizes, norbs = [4, 6, 5], (4, 7, 4)
    blocks, sL, sR, fL, fR = _setup(sizes, overlap=False, seed=3)
    pts, w = fiori.integration_points(np.linspace(-0.5, 0.5, 8), "direct")
    pot = np.full(3, 5.0)  # every atom in the electron branch
    ref = _dense_charge(pts, w, sL, sR, blocks, fL, fR, 0.0, pot, norbs)
    good = fiori.free_charge(pts, w, sL, sR, blocks, fL, fR, 0.0, 0.0, pot, list(norbs), sizes)
    assert np.abs(good.numpy() - ref).max() < 1e-11 * np.abs(ref).max()
    planted = fiori.free_charge(pts, w, sL, sR, blocks, fL, lambda x: fL(x) + fR(x), 0.0, 0.0,
                                pot, list(norbs), sizes)
    assert np.abs(planted.numpy() - ref).max() > 1e-3 * np.abs(ref).max()


def test_subblock_index_matches_global_ranges():
    rng = np.random.default_rng(0)
    for trial in range(200):
        nb = int(rng.integers(1, 6))
        sub = [int(x) for x in rng.integers(1, 7, nb)]
        total = sum(sub)
        # random atom partition of the same orbitals
        cuts = sorted(rng.choice(np.arange(1, total), size=min(total - 1, int(rng.integers(0, 6))),
                                 replace=False).tolist()) if total > 1 else []
        bounds = [0] + cuts + [total]
        norbs = [bounds[i + 1] - bounds[i] for i in range(len(bounds) - 1)]
        cum = np.cumsum([0] + sub)
        for atom in range(len(norbs)):
            idx, o0, o1 = fiori.subblock_index(sub, atom, norbs)
            lo, hi = bounds[atom], bounds[atom + 1]
            covered = []
            for b in idx:
                s = o0 if b == idx[0] else 0
                e = o1 if b == idx[-1] else sub[b]
                covered += list(range(cum[b] + s, cum[b] + e))
            assert covered == list(range(lo, hi)), (sub, norbs, atom, idx, o0, o1)
