"""Writes reference_rgf.npz + reference_rgf.json: inputs and all 14 outputs of
the reference's own ``recursive_gf`` (``dpnegf/negf/recursive_green_cal.py``)
on seeded cases that exercise what no closed form pins: the greater branch
with its ``.conj()`` quirk (``:142-161, 266-289``), the lesser branch with
S_offdiag != 0 and complex energies (``:120-140, 240-264``), a 3-D
(energy-dependent) ``seP`` (``:411-414``), the stacked uniform-block path
(``:72-168``), K = 1 and the scalar squeeze (``:386-395, 474-475``).

The reference file imports only torch, so it is loaded by file path; the
package itself is not imported.  This script runs only where the reference
tree exists (the build container); the fixtures it writes are what the tests
read.

Run from the repository root:  python tests/golden/make_reference_rgf.py
"""
import importlib.util
import json
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/dpnegf/negf/recursive_green_cal.py"
NAMES = ("g_trans", "gr_lc", "grd", "grl", "gru", "gr_left", "gnd", "gnl", "gnu", "gin_left",
         "gpd", "gpl", "gpu", "gip_left")
C = torch.complex128


def load_reference():
    spec = importlib.util.spec_from_file_location("ref_recursive_green_cal", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rc(rng, *shape):
    return torch.tensor(rng.normal(size=shape) + 1j * rng.normal(size=shape), dtype=C)


def herm(rng, n, s):
    a = rc(rng, n, n) * s
    return a + a.mH


def blocks(rng, sizes, overlap):
    """Hermitian H (and S) in block-tridiagonal form, hl = hu^H."""
    K = len(sizes)
    hd = [herm(rng, n, 0.4) for n in sizes]
    hu = [rc(rng, sizes[q], sizes[q + 1]) * 0.3 for q in range(K - 1)]
    hl = [h.mH.clone() for h in hu]
    if overlap:
        sd = [torch.eye(n, dtype=C) + herm(rng, n, 0.02) for n in sizes]
        su = [rc(rng, sizes[q], sizes[q + 1]) * 0.03 for q in range(K - 1)]
        sl = [s.mH.clone() for s in su]
    else:
        sd = [torch.eye(n, dtype=C) for n in sizes]
        su = [torch.zeros(sizes[q], sizes[q + 1], dtype=C) for q in range(K - 1)]
        sl = [s.mH.clone() for s in su]
    return hd, hu, hl, sd, su, sl


def retarded_se(rng, B, n):
    a = rc(rng, B, n, n) * 0.2
    return a - 0.4j * torch.eye(n, dtype=C)


def psd(rng, B, n, s):
    a = rc(rng, B, n, n) * s
    return a @ a.mH


CASES = {
    # uneven blocks, S_offdiag != 0, complex energies, lesser + greater + gr_lc,
    # a right self-energy smaller than its corner block, a 3-D seP on block 1
    "uneven": dict(sizes=[5, 7, 4], overlap=True, B=3, energy=[-0.6 + 0.02j, 0.15, 0.7 + 0.05j],
                   E_ref=0.1, eta=1e-4, nL=5, nR=3, seP=(1,), lesser=True, greater=True,
                   gr_lc=True, keep=True),
    # uniform blocks: the reference's stacked path
    "uniform": dict(sizes=[4, 4, 4, 4], overlap=True, B=2, energy=[0.3 + 0.01j, -0.45],
                    E_ref=0.0, eta=1e-5, nL=4, nR=4, seP=(), lesser=True, greater=True,
                    gr_lc=True, keep=True),
    # one block, scalar (0-d) energy with 2-D self-energies: squeezed outputs
    "single": dict(sizes=[6], overlap=False, B=None, energy=0.25 + 0.03j, E_ref=-0.05,
                   eta=1e-5, nL=6, nR=6, seP=(), lesser=True, greater=True, gr_lc=True,
                   keep=True),
    # two uneven blocks, scalar energy, retarded only, gr_left not kept
    "scalar": dict(sizes=[3, 5], overlap=True, B=None, energy=-0.35, E_ref=0.0, eta=1e-5,
                   nL=3, nR=4, seP=(), lesser=False, greater=False, gr_lc=False, keep=False),
}


def run_case(ref, name, c, seed):
    rng = np.random.default_rng(seed)
    sizes = c["sizes"]
    K = len(sizes)
    hd, hu, hl, sd, su, sl = blocks(rng, sizes, c["overlap"])
    B = c["B"]
    Bx = 1 if B is None else B
    energy = torch.tensor(c["energy"], dtype=C)  # 0-d for a scalar caller
    sL = retarded_se(rng, Bx, c["nL"])
    sR = retarded_se(rng, Bx, c["nR"])
    s_in = [psd(rng, Bx, n, 0.3) for n in sizes]
    s_out = [psd(rng, Bx, n, 0.3) for n in sizes]
    seP = None
    if c["seP"]:
        seP = [retarded_se(rng, Bx, n) * 0.5 if q in c["seP"] else 0 for q, n in enumerate(sizes)]
    if B is None:  # scalar caller: 2-D energy-dependent inputs
        sL, sR = sL[0], sR[0]
        s_in = [t[0] for t in s_in]
        s_out = [t[0] for t in s_out]
    inputs = dict(hd=hd, hu=hu, hl=hl, sd=sd, su=su, sl=sl, left_se=sL, right_se=sR,
                  s_in=s_in, s_out=s_out, energy=energy)
    if seP is not None:
        inputs["seP"] = [t if torch.is_tensor(t) else torch.zeros(sizes[q], sizes[q], dtype=C)
                         for q, t in enumerate(seP)]
    # the reference's wrapper mutates nothing it is given (it copies hd), but
    # hand it clones anyway so the stored inputs are exactly what went in
    cl = lambda xs: [x.clone() if torch.is_tensor(x) else x for x in xs]
    ans = ref.recursive_gf(energy.clone(), cl(hl), cl(hd), cl(hu), cl(sd), cl(su), cl(sl),
                           sL.clone(), sR.clone(), seP=cl(seP) if seP is not None else None,
                           E_ref=c["E_ref"], s_in=cl(s_in) if c["lesser"] else 0,
                           s_out=cl(s_out) if c["greater"] else 0, eta=c["eta"],
                           need_lesser=c["lesser"], need_greater=c["greater"],
                           need_gr_lc=c["gr_lc"], keep_gr_left=c["keep"])
    arrays, meta = {}, {"K": K, "sizes": sizes, "scalar": B is None, "E_ref": c["E_ref"],
                        "eta": c["eta"], "need_lesser": c["lesser"],
                        "need_greater": c["greater"], "need_gr_lc": c["gr_lc"],
                        "keep_gr_left": c["keep"], "seP": seP is not None, "outputs": {}}
    for key, val in inputs.items():
        if isinstance(val, list):
            for q, t in enumerate(val):
                arrays[f"{name}/in/{key}/{q}"] = t.resolve_conj().resolve_neg().numpy()
            meta[f"len_{key}"] = len(val)
        else:
            arrays[f"{name}/in/{key}"] = val.resolve_conj().resolve_neg().numpy()
    for key, val in zip(NAMES, ans):
        if val is None:
            meta["outputs"][key] = None
        elif isinstance(val, list):
            meta["outputs"][key] = len(val)
            for q, t in enumerate(val):
                arrays[f"{name}/out/{key}/{q}"] = t.resolve_conj().resolve_neg().numpy()
        else:
            meta["outputs"][key] = "tensor"
            arrays[f"{name}/out/{key}"] = val.resolve_conj().resolve_neg().numpy()
    return arrays, meta


def main():
    ref = load_reference()
    arrays, meta = {}, {"source": "dpnegf/negf/recursive_green_cal.py (recursive_gf)",
                        "cases": {}}
    for i, (name, c) in enumerate(CASES.items()):
        a, m = run_case(ref, name, c, seed=100 + i)
        arrays.update(a)
        meta["cases"][name] = m
    np.savez_compressed(os.path.join(HERE, "reference_rgf.npz"), **arrays)
    with open(os.path.join(HERE, "reference_rgf.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {len(arrays)} arrays for {len(CASES)} cases")


if __name__ == "__main__":
    main()
