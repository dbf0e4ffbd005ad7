"""Writes ribbon48_K128_T.json: oracle T(E) of the W=48, K=128 ribbon (Anderson
U[-0.25, 0.25] seed 0, clean leads, eta_lead = eta_device = 1e-6) at 24 of the
4096 energies linspace(-3, 3, 4096): 16 evenly spaced plus the 8 points where
the lead decimation is worst conditioned.  For each energy T is computed with
the decimation solved by LU (the reference's association) and by a refined
inverse; their difference is the conditioning spread a GPU result is held to.

Run from the repository root:  python tests/golden/make_ribbon48_K128.py
"""
import json
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from dpnegf_amd import workloads as wl  # noqa: E402
from oracle import device as odev, lead as olead, rgf as orgf, surface_green as osg  # noqa: E402

EDGE_IDX = (137, 452, 1036, 1352, 1550, 2545, 2743, 3643)
ETA = 1e-6


def sigma(lead, e, variant):
    H, h01, hDL, S, s01, sDL = [x.numpy() for x in lead]
    gs = osg.lopez_sancho(H, h01, S, s01, e + 1j * ETA, variant=variant)[0]
    b = hDL.shape[1]
    return torch.tensor((e * sDL - hDL) @ gs[:b, :b] @ (e * sDL.conj().T - hDL.conj().T))


def main():
    torch.set_num_threads(min(8, os.cpu_count() or 1))
    s = wl.ribbon(W=48, K=128, seed=0)
    E = np.linspace(-3.0, 3.0, 4096)
    idx = sorted(set(np.linspace(0, 4095, 16).round().astype(int).tolist()) | set(EDGE_IDX))
    rows = []
    for i in idx:
        e = float(E[i])
        Ts = []
        for variant in ("solve", "refined"):
            sL, sR = sigma(s.lead_L, e, variant), sigma(s.lead_R, e, variant)
            ans = orgf.recursive_gf(torch.tensor(e, dtype=torch.complex128), s.hl, s.hd, s.hu,
                                    s.sd, s.su, s.sl, sL, sR, eta=ETA, keep_gr_left=False)
            Ts.append(float(odev.transmission(ans[0], olead.sigma2gamma(sL),
                                              olead.sigma2gamma(sR))))
        rows.append({"index": i, "E": e, "T": Ts[0], "T_refined": Ts[1]})
        print(i, e, Ts[0], Ts[1] - Ts[0], flush=True)
    out = {"workload": "ribbon W=48 K=128 t=-2.7 Anderson U[-0.25,0.25] seed 0, clean leads, "
                       "eta_lead=eta_device=1e-6, E=linspace(-3,3,4096)",
           "points": rows}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ribbon48_K128_T.json"),
              "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
