"""Reader of tests/golden/reference_rgf.{npz,json}: inputs and outputs of the
reference's own ``recursive_gf`` (written by tests/golden/make_reference_rgf.py)."""
import json
import os

import numpy as np
import torch

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("g_trans", "gr_lc", "grd", "grl", "gru", "gr_left", "gnd", "gnl", "gnu", "gin_left",
         "gpd", "gpl", "gpu", "gip_left")


def load():
    meta = json.load(open(os.path.join(HERE, "reference_rgf.json")))
    arr = np.load(os.path.join(HERE, "reference_rgf.npz"))  # allow_pickle stays False
    return meta, arr


def case(name, meta=None, arr=None):
    """(kwargs for recursive_gf, expected 14-tuple, case metadata)."""
    if meta is None:
        meta, arr = load()
    m = meta["cases"][name]
    t = lambda key: torch.from_numpy(arr[f"{name}/{key}"].copy())
    lst = lambda key: [t(f"in/{key}/{q}") for q in range(m[f"len_{key}"])]
    kw = dict(energy=t("in/energy"), hl=lst("hl"), hd=lst("hd"), hu=lst("hu"), sd=lst("sd"),
              su=lst("su"), sl=lst("sl"), left_se=t("in/left_se"), right_se=t("in/right_se"),
              seP=lst("seP") if m["seP"] else None, E_ref=m["E_ref"],
              s_in=lst("s_in") if m["need_lesser"] else 0,
              s_out=lst("s_out") if m["need_greater"] else 0, eta=m["eta"],
              need_lesser=m["need_lesser"], need_greater=m["need_greater"],
              need_gr_lc=m["need_gr_lc"], keep_gr_left=m["keep_gr_left"])
    expect = []
    for key in NAMES:
        kind = m["outputs"][key]
        if kind is None:
            expect.append(None)
        elif kind == "tensor":
            expect.append(t(f"out/{key}"))
        else:
            expect.append([t(f"out/{key}/{q}") for q in range(kind)])
    return kw, tuple(expect), m


def compare(got, expect, rtol):
    """Max relative error of every output (scale: the largest entry of that
    output); None must stay None and block lists keep their lengths."""
    worst = 0.0
    for key, g, r in zip(NAMES, got, expect):
        if r is None:
            assert g is None, key
            continue
        if torch.is_tensor(r):
            g, r = [g], [r]
        assert isinstance(g, list) and len(g) == len(r), key
        scale = max([x.abs().max().item() for x in r] + [1e-300])
        for q, (gq, rq) in enumerate(zip(g, r)):
            gq = gq.detach().cpu()
            assert gq.shape == rq.shape, (key, q, gq.shape, rq.shape)
            err = (gq - rq).abs().max().item() / scale
            assert err < rtol, f"{key}[{q}]: rel err {err:.3e}"
            worst = max(worst, err)
    return worst
