"""Seeded synthetic tight-binding devices for the five benchmark workloads.

Every builder returns a :class:`BTDSystem`: the device blocks in the layout
``recursive_gf`` takes (``hd, hu, hl, sd, su, sl`` lists of 2-D complex128
tensors, the layout of ``NEGFHamiltonianInit.get_hs_device``,
``dpnegf/negf/negf_hamiltonian_init.py:745-880``) plus, per lead, the tuple
``(hL, hLL, hDL, sL, sLL, sDL)`` in the order of ``get_hs_lead``
(``negf_hamiltonian_init.py:882-974``), already reduced to the rows of the
adjacent device block (``HDL_reduced``, ``lead_property.py:386-434``).

Lead orientation follows the reference: ``hLL`` couples the principal layer
next to the device to the next layer further out; ``hDL`` couples device
rows to the lead's surface layer.  For a chain of cells joined by ``H01``
(cell i -> cell i+1, left to right) that gives ``hLL = hDL = H01^dagger`` on
the left and ``hLL = hDL = H01`` on the right.

All energies are in eV; S = I and S_offdiag = 0 unless stated.
"""
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np
import torch

C128 = torch.complex128


@dataclass
class BTDSystem:
    hd: List[torch.Tensor]
    hu: List[torch.Tensor]
    hl: List[torch.Tensor]
    sd: List[torch.Tensor]
    su: List[torch.Tensor]
    sl: List[torch.Tensor]
    lead_L: Tuple[torch.Tensor, ...]
    lead_R: Tuple[torch.Tensor, ...]
    meta: dict = field(default_factory=dict)

    @property
    def sizes(self):
        return [int(h.shape[0]) for h in self.hd]

    def to(self, device):
        mv = lambda xs: [x.to(device) for x in xs]
        return BTDSystem(mv(self.hd), mv(self.hu), mv(self.hl), mv(self.sd),
                         mv(self.su), mv(self.sl), tuple(mv(self.lead_L)),
                         tuple(mv(self.lead_R)), dict(self.meta))


def _t(a):
    return torch.as_tensor(np.asarray(a), dtype=C128)


def _eye(n):
    return torch.eye(n, dtype=C128)


def _zeros(a, b):
    return torch.zeros(a, b, dtype=C128)


def uniform_chain(H00, H01, K, H00_lead=None, diag_blocks=None):
    """Device of K identical cells (optionally per-cell diagonal blocks) with
    identical semi-infinite leads built from ``H00_lead`` / ``H01``."""
    H00 = _t(H00)
    H01 = _t(H01)
    n = H00.shape[0]
    lead00 = H00 if H00_lead is None else _t(H00_lead)
    hd = [(_t(b) if diag_blocks is not None else H00.clone())
          for b in (diag_blocks if diag_blocks is not None else range(K))]
    hu = [H01.clone() for _ in range(K - 1)]
    hl = [H01.mH.clone() for _ in range(K - 1)]
    sd = [_eye(n) for _ in range(K)]
    su = [_zeros(n, n) for _ in range(K - 1)]
    sl = [_zeros(n, n) for _ in range(K - 1)]
    lead_L = (lead00.clone(), H01.mH.clone(), H01.mH.clone(), _eye(n), _zeros(n, n), _zeros(n, n))
    lead_R = (lead00.clone(), H01.clone(), H01.clone(), _eye(n), _zeros(n, n), _zeros(n, n))
    return BTDSystem(hd, hu, hl, sd, su, sl, lead_L, lead_R)


# -- workload 1: two-orbital chain ------------------------------------------------
CHAIN2_H00 = [[0.0, 0.5], [0.5, 1.0]]
CHAIN2_H01 = [[-1.0, 0.2], [0.2, -0.8]]


def two_orbital_chain(K=16):
    """On-site diag [0, 1.0] eV, intra-cell hopping 0.5 eV, inter-cell
    hopping [[-1.0, 0.2], [0.2, -0.8]] eV; K principal layers of size 2."""
    s = uniform_chain(CHAIN2_H00, CHAIN2_H01, K)
    s.meta.update(name="two_orbital_chain", K=K, n=2)
    return s


def chain2_energies(nE=201):
    return torch.linspace(-3.0, 3.0, nE, dtype=torch.float64)


# -- workloads 2 and 4: square-lattice ribbon ---------------------------------------
def ribbon_layer(W, t=-2.7, onsite=None):
    h = np.zeros((W, W))
    i = np.arange(W - 1)
    h[i, i + 1] = t
    h[i + 1, i] = t
    if onsite is not None:
        h[np.arange(W), np.arange(W)] = onsite
    return h


def ribbon(W=48, K=128, t=-2.7, disorder=0.25, seed=0):
    """Square-lattice ribbon, one orbital per site, hopping ``t``, width W
    (block size W), K layers with Anderson disorder U[-disorder, disorder]
    from ``seed``; clean leads."""
    rng = np.random.default_rng(seed)
    diag = [ribbon_layer(W, t, rng.uniform(-disorder, disorder, W)) for _ in range(K)]
    s = uniform_chain(ribbon_layer(W, t), t * np.eye(W), K,
                      H00_lead=ribbon_layer(W, t), diag_blocks=diag)
    s.meta.update(name="ribbon", W=W, K=K, t=t, disorder=disorder, seed=seed)
    return s


def ribbon_open_channels(E, W, t=-2.7):
    """Clean-ribbon transmission: number of transverse modes with
    |E - eps_m| < 2|t|, eps_m = 2 t cos(m pi/(W+1))."""
    m = np.arange(1, W + 1)
    eps = 2 * t * np.cos(m * np.pi / (W + 1))
    E = np.asarray(E, dtype=float)[..., None]
    return (np.abs(E - eps) < 2 * abs(t)).sum(-1)


# -- workloads 3 and 5: uneven multi-orbital layers with transverse k ---------------
def _herm(rng, m, sigma):
    a = rng.normal(0.0, sigma, (m, m))
    return 0.5 * (a + a.T)


def multiorbital(widths=(32, 40, 48, 56), K=40, norb=3, kpoint=0.0, seed=0,
                 onsite_sigma=0.5, hop_sigma=0.3, shift=None):
    """``norb`` orbitals per site; layer q has ``widths[q % len(widths)]``
    sites on a transverse ring closed with Bloch phase exp(2 pi i k).

    * on-site ``norb x norb`` blocks N(0, onsite_sigma^2), symmetrised, one
      per site (fixed by ``seed``);
    * one nearest-neighbour hopping matrix along the ring (``Ty``) and one
      between layers (``Tx``), entries N(0, hop_sigma^2);
    * layer q couples site i to site i of layer q+1 for i < min(W_q, W_q+1);
    * leads repeat the first / last layer, coupled through ``I (x) Tx``.
    """
    rng = np.random.default_rng(seed)
    Ty = rng.normal(0.0, hop_sigma, (norb, norb))
    Tx = rng.normal(0.0, hop_sigma, (norb, norb))
    Ws = [widths[q % len(widths)] for q in range(K)]
    phase = np.exp(2j * np.pi * kpoint)

    def layer(W, onsite):
        n = W * norb
        h = np.zeros((n, n), complex)
        for i in range(W):
            h[i * norb:(i + 1) * norb, i * norb:(i + 1) * norb] = onsite[i]
            j = (i + 1) % W
            ph = phase if j == 0 else 1.0
            h[i * norb:(i + 1) * norb, j * norb:(j + 1) * norb] += Ty * ph
            h[j * norb:(j + 1) * norb, i * norb:(i + 1) * norb] += (Ty * ph).conj().T
        return h

    def couple(Wa, Wb):
        c = np.zeros((Wa * norb, Wb * norb), complex)
        for i in range(min(Wa, Wb)):
            c[i * norb:(i + 1) * norb, i * norb:(i + 1) * norb] = Tx
        return c

    onsites = [[_herm(rng, norb, onsite_sigma) for _ in range(W)] for W in Ws]
    if shift is not None:
        onsites = [[o + shift * np.eye(norb) for o in layer_] for layer_ in onsites]
    hd = [_t(layer(W, o)) for W, o in zip(Ws, onsites)]
    hu = [_t(couple(Ws[q], Ws[q + 1])) for q in range(K - 1)]
    hl = [h.mH.clone() for h in hu]
    sd = [_eye(h.shape[0]) for h in hd]
    su = [_zeros(*h.shape) for h in hu]
    sl = [_zeros(*h.shape) for h in hl]
    cL = _t(couple(Ws[0], Ws[0]))
    cR = _t(couple(Ws[-1], Ws[-1]))
    nL, nR = hd[0].shape[0], hd[-1].shape[0]
    lead_L = (hd[0].clone(), cL.mH.clone(), cL.mH.clone(), _eye(nL), _zeros(nL, nL), _zeros(nL, nL))
    lead_R = (hd[-1].clone(), cR.clone(), cR.clone(), _eye(nR), _zeros(nR, nR), _zeros(nR, nR))
    s = BTDSystem(hd, hu, hl, sd, su, sl, lead_L, lead_R)
    s.meta.update(name="multiorbital", widths=tuple(widths), K=K, norb=norb,
                  kpoint=kpoint, seed=seed)
    return s


def random_btd(block_sizes, seed=0, hermitian_offdiag=True, overlap=False):
    """Random BTD device (Hermitian diagonal blocks, ``hl = hu^dagger``),
    in the style of ``dpnegf/tests/test_recursive_gf_batched.py:18-54``."""
    g = torch.Generator().manual_seed(seed)

    def rc(*shape):
        return (torch.randn(*shape, generator=g, dtype=torch.float64)
                + 1j * torch.randn(*shape, generator=g, dtype=torch.float64))

    hd, sd, hu, hl, su, sl = [], [], [], [], [], []
    for n in block_sizes:
        h = rc(n, n)
        hd.append(0.5 * (h + h.mH))
        if overlap:
            a = 0.05 * rc(n, n)
            sd.append(torch.eye(n, dtype=C128) + 0.5 * (a + a.mH))
        else:
            sd.append(torch.eye(n, dtype=C128))
    for q in range(len(block_sizes) - 1):
        a, b = block_sizes[q], block_sizes[q + 1]
        u = rc(a, b)
        hu.append(u)
        hl.append(u.mH.clone() if hermitian_offdiag else rc(b, a))
        if overlap:
            s_ = 0.02 * rc(a, b)
            su.append(s_)
            sl.append(s_.mH.clone())
        else:
            su.append(torch.zeros(a, b, dtype=C128))
            sl.append(torch.zeros(b, a, dtype=C128))
    return hd, sd, hl, hu, sl, su


class BlochLead:
    """A lead whose principal layer is ``m`` copies of a small cell of ``n``
    orbitals along a periodic transverse direction y, for Bloch-unfolded
    self-energies (``bloch_factor = [1, m, 1]``, ``bloch_R_list`` the cell
    offsets (0, j, 0)).  Terms, per small cell R and layer (fixed by ``seed``):
    on-site ``h0`` (Hermitian), ``ty`` from cell R to R + 1 in y,
    ``tx`` to cell R of the next layer, ``txy`` to cell R + 1 of the next layer.
    The supercell is stored orbital-major (orbital a of cell i at a m + i), so
    ``sorted_idx`` maps it onto the cell-major order of kron(P, g).

    ``get_hs_lead(k)`` follows ``negf_hamiltonian_init.py:882-975``: the small
    cell's blocks at the unfolded point ``k`` and the supercell's device-lead
    coupling at the point ``k`` folds onto; ``supercell(k)`` gives the whole
    supercell lead at ``k`` for a direct decimation."""

    def __init__(self, m=2, n=3, seed=0, a_dev=None, overlap=False):
        rng = np.random.default_rng(seed)
        self.m, self.n = m, n
        self.h0 = _herm(rng, n, 0.5)
        self.ty = rng.normal(0, 0.3, (n, n)) + 1j * rng.normal(0, 0.1, (n, n))
        self.tx = rng.normal(0, 0.4, (n, n))
        self.txy = rng.normal(0, 0.15, (n, n))
        self.s_on = np.eye(n) + (0.05 * _herm(rng, n, 1.0) if overlap else 0)
        self.s_x = 0.03 * rng.normal(0, 1, (n, n)) if overlap else np.zeros((n, n))
        a = a_dev if a_dev is not None else m * n
        self.hdl = rng.normal(0, 0.3, (a, m * n)) + 1j * rng.normal(0, 0.05, (a, m * n))
        self.sdl = 0.02 * rng.normal(0, 1, (a, m * n)) if overlap else np.zeros((a, m * n))
        self.sorted_idx = torch.tensor([i * n + o for o in range(n) for i in range(m)])
        self.R_list = [torch.tensor([0.0, float(j), 0.0]) for j in range(m)]
        self.bloch_factor = [1, m, 1]
        self.tab = "lead_L"

    def small(self, k):
        """Small-cell (hL, hLL, sL, sLL) at reduced wave vector k (3-vector),
        H(k) = sum_d H(0 -> d) exp(-2 pi i k d): the sign that the unfolding's
        phase exp(2 pi i (R_j - R_i) . k) assumes."""
        ph = np.exp(-2j * np.pi * float(k[1]))
        hL = self.h0 + self.ty * ph + self.ty.conj().T / ph
        hLL = self.tx + self.txy * ph
        return hL, hLL, self.s_on.astype(complex), self.s_x.astype(complex)

    def supercell(self, k):
        """Supercell (hL, hLL, sL, sLL) at k, orbital-major."""
        m, n = self.m, self.n
        ph = np.exp(-2j * np.pi * float(k[1]))
        big = [np.zeros((m * n, m * n), complex) for _ in range(4)]
        terms = [(0, 0, self.h0), (0, 1, self.ty), (0, -1, self.ty.conj().T),  # in-layer
                 (1, 0, self.tx), (1, 1, self.txy),  # to the next layer
                 (2, 0, self.s_on), (3, 0, self.s_x)]
        for which, d, t in terms:
            for i in range(m):
                j, L = (i + d) % m, (i + d) // m
                blk = t * ph ** L
                for o1 in range(n):
                    for o2 in range(n):
                        big[which][o1 * m + i, o2 * m + j] += blk[o1, o2]
        return tuple(big)

    def get_hs_lead(self, kpoint, tab=None, v=0.0):
        k = np.asarray(kpoint, dtype=float).reshape(3)
        hL, hLL, sL, sLL = self.small(k)
        return (_t(hL - v * sL), _t(hLL - v * sLL), _t(self.hdl), _t(sL), _t(sLL), _t(self.sdl))
