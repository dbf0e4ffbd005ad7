"""Writes config_shapes.npz: oracle results at the shapes of two benchmark
workloads, too slow to recompute inside a test run.

* multi-orbital device, 3 orbitals per site, widths 32/40/48/56 (blocks
  96/120/144/168), 40 layers, seed 0, k = 0 and 5/32, energies
  -1.7, -0.45, 0.3, 1.25 eV, eta = 1e-5: T per (k, E);
* density device, 8 orbitals per site, widths 32/24/28/30 (blocks
  256/192/224/240), 24 layers, seed 0, k = 0: the Ozaki density with the R term
  (R = 1e6) and poles 0, 7, 60 of the 160-pole set, kT = 25.852 meV, mu = 0,
  plus a 16-node Gauss-Legendre bias window (V_L = 0.05, V_R = -0.05,
  eta_lead = 1e-5, eta_device = 1e-6); per diagonal block its diagonal, first
  row and last column, equilibrium and non-equilibrium.

Every number comes from the oracle (oracle/).  Run from the repository root:
    python tests/golden/make_config_shapes.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from dpnegf_amd import workloads as wl  # noqa: E402
from oracle import ozaki as ooz  # noqa: E402
from tests.test_gpu_configs import _oracle_T  # noqa: E402
from tests.test_gpu_density import _oracle_density  # noqa: E402

MO_KPTS = (0, 5)
MO_E = np.array([-1.7, -0.45, 0.3, 1.25])
MO_ETA = 1e-5
DENS = dict(kBT=0.025852, mu=0.0, mu_L=0.0, R=1e6, n_gauss=16, vL=0.05, vR=-0.05,
            eta_lead=1e-5, eta_dev=1e-6, pole_idx=(0, 7, 60), M=160)


def entries(m):
    """Diagonal, first row and last column of a square block."""
    m = np.asarray(m)
    return np.concatenate([np.diag(m), m[0], m[:, -1]])


def main():
    out = {}
    for k in MO_KPTS:
        s = wl.multiorbital(kpoint=k / 32, seed=0)
        out[f"mo_T_k{k}"] = _oracle_T(s, MO_E, MO_ETA, MO_ETA)
        print("multiorbital k", k, out[f"mo_T_k{k}"], flush=True)
    d = DENS
    s = wl.multiorbital(widths=(32, 24, 28, 30), K=24, norb=8, kpoint=0.0, seed=0)
    poles, res = ooz.ozaki_residues(d["M"])
    idx = list(d["pole_idx"])
    fL = lambda x: 1 / (1 + torch.exp((x - (d["mu"] - d["vL"])) / d["kBT"]))
    fR = lambda x: 1 / (1 + torch.exp((x - (d["mu"] - d["vR"])) / d["kBT"]))
    eq, neq = _oracle_density(s, d["kBT"], d["mu_L"], d["mu"], d["R"], None, 0.0, d["eta_lead"],
                              d["eta_dev"], d["vL"], d["vR"], d["n_gauss"], fL, fR,
                              poles=poles[idx], res=res[idx])
    for q in range(len(eq)):
        n = s.sizes[q]
        out[f"dens_eq_{q}"] = entries(eq[q].reshape(n, n).numpy())
        out[f"dens_neq_{q}"] = entries(np.asarray(neq[q]).reshape(n, n))
    np.savez_compressed(os.path.join(HERE, "config_shapes.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
