"""Kernel-shape tuning table and autotuner of the block-operation path.

* every inverse form the table can select (register-resident and LDS
  Gauss-Jordan, including the 33-64 LDS form only the table reaches) against
  numpy at 1e-10, with row pivoting;
* every product tile shape against the built-in shape (same k order, so equal
  to 1e-14 of the largest element);
* the recorded keys of a blocked sweep, and the sweep and the batched
  decimation with tuned and with forced alternative shapes against the
  built-in choice (1e-10 of max T / of max |Sigma|).
The table is process-wide: every test leaves it empty."""
import numpy as np
import pytest
import torch

from dpnegf_amd import _native as nat
from dpnegf_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


@pytest.fixture(autouse=True)
def _empty_table():
    nat.load().dpnegf_tune_clear()
    yield
    nat.load().dpnegf_tune_clear()


def _inv(A, dev):
    t = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    st = torch.zeros(A.shape[0], dtype=torch.int32, device=dev)
    nat.call("dpnegf_zinv_batched", nat.stream_ptr(dev), nat.ptr(t), A.shape[1], A.shape[0],
             nat.ptr(st))
    torch.cuda.synchronize()
    return t.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("n", [20, 40, 64, 72, 90, 96])
@pytest.mark.parametrize("form", ["register", "lds"])
def test_inverse_forms_match_numpy(n, form):
    dev = _dev()
    print(f"case tuned inverse n={n} form={form}", flush=True)
    nat.call("dpnegf_tune_set", nat.TUNE_INVERSE, n, 0, 0,
             nat.INV_REGISTER if form == "register" else nat.INV_LDS)
    rng = np.random.default_rng(300 + n)
    B = 4
    A = np.empty((B, n, n), complex)
    for b in range(B):
        D = 4.0 * np.eye(n) + 0.1 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        R = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + np.sqrt(n) * np.eye(n)
        A[b] = D[rng.permutation(n)] if b % 2 else R
    X, st = _inv(A, dev)
    assert (st == 0).all()
    ref = np.linalg.inv(A)
    for b in range(B):
        rel = np.abs(X[b] - ref[b]).max() / np.abs(ref[b]).max()
        res = np.abs(A[b] @ X[b] - np.eye(n)).max()
        assert rel < 1e-10 and res < 1e-11, (b, rel, res)


def test_lds_form_is_refused_above_96():
    assert nat.load().dpnegf_tune_set(nat.TUNE_INVERSE, 100, 0, 0, nat.INV_LDS) != 0
    assert nat.load().dpnegf_tune_get(nat.TUNE_INVERSE, 100, 0, 0) == 0


@pytest.mark.parametrize("mnk", [(96, 168, 120), (40, 72, 33), (130, 17, 64)])
def test_product_tiles_agree(mnk):
    dev = _dev()
    m, n, k = mnk
    print(f"case tuned product {mnk}", flush=True)
    g = torch.Generator().manual_seed(m + n + k)
    B = 5

    def rnd(*s):
        return torch.complex(torch.randn(*s, generator=g, dtype=torch.float64),
                             torch.randn(*s, generator=g, dtype=torch.float64)).to(dev)

    A, Bm = rnd(B, m, k), rnd(B, k, n)

    def run():
        C = torch.empty(B, m, n, dtype=torch.complex128, device=dev)
        nat.call("dpnegf_zgemm_batched", nat.stream_ptr(dev), B, m, n, k, nat.ptr(A), k, m * k, 0,
                 nat.ptr(Bm), n, k * n, 0, 1.0, 0.0, 0.0, 0.0, nat.ptr(C), n, m * n)
        torch.cuda.synchronize()
        return C

    ref = run()
    assert float((ref - A @ Bm).abs().max()) <= 1e-12 * float(ref.abs().max())
    for c in (nat.TILE_32x32, nat.TILE_32x64, nat.TILE_64x32, nat.TILE_64x64):
        nat.call("dpnegf_tune_set", nat.TUNE_GEMM_TILE, m, n, k, c)
        C = run()
        assert float((C - ref).abs().max()) <= 1e-14 * float(ref.abs().max()), c


def _mo_system(dev):
    return wl.multiorbital(widths=(24, 30, 27), K=4, norb=3, kpoint=0.1, seed=6).to(dev)


def test_recorded_shapes_and_tuned_sweep(monkeypatch):
    from dpnegf_amd.negf import autotune
    from dpnegf_amd.negf.device_property import transmission_pipeline
    dev = _dev()
    print("case tuned blocked sweep 72/90/81/72", flush=True)
    s = _mo_system(dev)
    E = torch.linspace(-1.0, 1.0, 24, dtype=torch.float64, device=dev)
    eta = 1e-4
    ref = transmission_pipeline(s, E, eta_lead=eta, eta_device=eta)
    scale = float(ref.abs().max())
    # every inverse with its other form, every product with 32x32 tiles
    keys = autotune.record_shapes(lambda: transmission_pipeline(s, E, eta_lead=eta,
                                                                eta_device=eta))
    inv = {k[1] for k in keys if k[0] == nat.TUNE_INVERSE}
    assert {72, 81, 90} <= inv
    assert any(k[0] == nat.TUNE_GEMM_TILE for k in keys)
    for k in keys:
        if k[0] == nat.TUNE_INVERSE:
            other = nat.INV_REGISTER if 64 < k[1] <= 96 else nat.INV_LDS
            if k[1] <= 96:
                nat.call("dpnegf_tune_set", *k, other)
        else:
            nat.call("dpnegf_tune_set", *k, nat.TILE_32x32)
    forced = transmission_pipeline(s, E, eta_lead=eta, eta_device=eta)
    assert float((forced - ref).abs().max()) <= 1e-10 * scale
    nat.load().dpnegf_tune_clear()
    # the autotuner's own picks, through the environment switch
    monkeypatch.setenv("DPNEGF_AUTOTUNE", "1")
    tuned = transmission_pipeline(s, E, eta_lead=eta, eta_device=eta)
    assert float((tuned - ref).abs().max()) <= 1e-10 * scale
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems
    res = autotune.tune_blocked(BTDSystems([(s.hd, s.hu, s.hl, s.sd, s.su, s.sl)], dev), 24)
    assert res and all(c in (0,) + autotune._candidates(k) for k, (c, _) in res.items())


def test_tuned_lead_decimation(monkeypatch):
    from dpnegf_amd.negf.surface_green import self_energy_batched
    dev = _dev()
    print("case tuned decimation lead 72", flush=True)
    s = _mo_system(dev)
    E = torch.linspace(-1.0, 1.0, 16, dtype=torch.float64, device=dev)
    lead = tuple(x.to(dev) if x is not None else None for x in s.lead_L)
    ref = self_energy_batched([lead], E, 1e-4)
    monkeypatch.setenv("DPNEGF_AUTOTUNE", "1")
    tuned = self_energy_batched([lead], E, 1e-4)
    scale = float(ref.sigma.abs().max())
    assert float((tuned.sigma - ref.sigma).abs().max()) <= 1e-10 * scale
    assert int(tuned.status.count_nonzero()) == 0
