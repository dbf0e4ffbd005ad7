"""Building blocks on the GPU: batched complex GEMM (matrix-core tiles, every
op mode, shared and strided operands), diag-of-product and weighted sums, vs
plain CPU products (complex128)."""
import pytest
import torch

from dpnegf_amd import _native as nat

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _rc(g, *shape):
    return torch.randn(*shape, generator=g, dtype=torch.float64) + 1j * torch.randn(
        *shape, generator=g, dtype=torch.float64)


def _op(x, op):
    return {0: x, 1: x.mH, 2: x.mT, 3: x.conj()}[op]


@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (1, 1, 1), (48, 48, 48), (37, 21, 13), (96, 48, 48),
                                   (70, 130, 65)])
@pytest.mark.parametrize("opA,opB", [(0, 0), (1, 0), (0, 1), (2, 3)])
def test_zgemm_batched_matches_cpu(m, n, k, opA, opB):
    dev = _dev()
    g = torch.Generator().manual_seed(m * 131 + n * 7 + k + 11 * opA + opB)
    B = 3
    A = _rc(g, B, *((m, k) if opA in (0, 3) else (k, m)))
    Bm = _rc(g, B, *((k, n) if opB in (0, 3) else (n, k)))
    C0 = _rc(g, B, m, n)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
    Ad, Bd, Cd = A.to(dev), Bm.to(dev), C0.clone().to(dev)
    nat.call("dpnegf_zgemm_batched", nat.stream_ptr(dev), B, m, n, k, nat.ptr(Ad), A.shape[-1],
             A[0].numel(), opA, nat.ptr(Bd), Bm.shape[-1], Bm[0].numel(), opB, alpha.real,
             alpha.imag, beta.real, beta.imag, nat.ptr(Cd), n, m * n)
    ref = alpha * (_op(A, opA) @ _op(Bm, opB)) + beta * C0
    assert (Cd.cpu() - ref).abs().max().item() < 1e-12 * max(1.0, ref.abs().max().item())


def test_shared_operand_and_reductions():
    dev = _dev()
    g = torch.Generator().manual_seed(5)
    B, m, n = 4, 9, 7
    A = _rc(g, B, m, n).to(dev)
    S = _rc(g, n, m).to(dev)  # shared across the batch (stride 0)
    out = torch.empty(B, m, dtype=torch.complex128, device=dev)
    nat.call("dpnegf_diag_product", nat.stream_ptr(dev), nat.ptr(out), nat.ptr(A), m * n,
             nat.ptr(S), 0, m, n, B, 0)
    ref = torch.diagonal(A.cpu() @ S.cpu(), dim1=-2, dim2=-1)
    assert (out.cpu() - ref).abs().max() < 1e-12
    c = _rc(g, B).to(dev)
    w = torch.empty(m * n, dtype=torch.float64, device=dev)
    nat.call("dpnegf_weighted_re_sum", nat.stream_ptr(dev), nat.ptr(w), nat.ptr(A), nat.ptr(c),
             m * n, B, 0)
    ref = (c.cpu().view(B, 1, 1) * A.cpu()).real.sum(0).reshape(-1)
    assert (w.cpu() - ref).abs().max() < 1e-12


def test_nan_inputs_give_nan_not_a_fault():
    """Regression: a NaN column used to leave the Gauss-Jordan pivot unset and
    its permutation partly uninitialised; the selection now always picks an
    unused row, so NaN inputs propagate as NaN results."""
    from dpnegf_amd import workloads as wl
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems, recursive_gf, rgf_transmission
    dev = _dev()
    hd, sd, hl, hu, sl, su = wl.random_btd([6, 6, 6], seed=1)
    to = lambda xs: [x.to(dev) for x in xs]
    E = torch.tensor([0.1, float("nan")], dtype=torch.complex128, device=dev)
    se = torch.zeros(2, 6, 6, dtype=torch.complex128, device=dev) - 0.3j * torch.eye(6, device=dev)
    out = recursive_gf(E, to(hl), to(hd), to(hu), to(sd), to(su), to(sl), se, se, eta=1e-3)
    torch.cuda.synchronize()
    assert torch.isfinite(out[2][0][0]).all() and not torch.isfinite(out[2][0][1]).all()
    bs = BTDSystems([(to(hd), to(hu), to(hl), to(sd), to(su), to(sl))], dev)
    T = rgf_transmission(bs, (E + 1e-3j).contiguous(), se, se)
    torch.cuda.synchronize()
    assert torch.isfinite(T[0]) and torch.isnan(T[1])


def _call_lead(dev, lead_rec, n, nsys, sys_idx, e, eta):
    P = e.shape[0]
    ws = torch.empty(nat.load().dpnegf_lead_self_energy_workspace_elems(n, n, P),
                     dtype=torch.complex128, device=dev)
    gs = torch.empty(P, n, n, dtype=torch.complex128, device=dev)
    sig = torch.empty(P, n, n, dtype=torch.complex128, device=dev)
    it = torch.zeros(P, dtype=torch.int32, device=dev)
    st = torch.zeros(P, dtype=torch.int32, device=dev)
    rs = torch.zeros(P, dtype=torch.float64, device=dev)
    nat.call("dpnegf_lead_self_energy", nat.stream_ptr(dev), n, n, 1, nat.ptr(lead_rec),
             lead_rec.numel() // nsys, nsys, nat.ptr(sys_idx), P, nat.ptr(e), eta, 1e-40,
             nat.ptr(ws), nat.ptr(gs), nat.ptr(sig), 0, nat.ptr(it), nat.ptr(st), nat.ptr(rs))
    torch.cuda.synchronize()
    return sig.cpu(), st.cpu()


@pytest.mark.parametrize("n", [6, 52])
def test_out_of_range_sys_idx_is_flagged_not_read(n):
    """A sys_idx entry outside [0, nsys) is reported per point
    (DPNEGF_STATUS_BAD_INDEX) by the fused kernels and the batched decimation
    (n > 48), which read no record outside the arena for it; the other points
    are unchanged.  Every index the kernels use is checked before any load
    through it, so the test reads nothing out of bounds."""
    from dpnegf_amd.negf.surface_green import pack_lead
    dev = _dev()
    print(f"case bad sys_idx n={n}", flush=True)
    g = torch.Generator().manual_seed(n)
    h0 = _rc(g, n, n) * 0.2
    h0 = h0 + h0.mH
    h1 = _rc(g, n, n) * 0.15
    eye = torch.eye(n, dtype=torch.complex128)
    rec = pack_lead(h0, h1, eye, torch.zeros(n, n, dtype=torch.complex128), device=dev)[0]
    e = torch.tensor([0.3 + 0.05j, -0.2 + 0.05j, 0.1 + 0.05j, 0.4 + 0.05j],
                     dtype=torch.complex128, device=dev)
    good = torch.zeros(4, dtype=torch.int32, device=dev)
    bad = torch.tensor([0, 1, -3, 0], dtype=torch.int32, device=dev)
    s_ok, st_ok = _call_lead(dev, rec, n, 1, good, e, 1e-5)
    s_bad, st_bad = _call_lead(dev, rec, n, 1, bad, e, 1e-5)
    assert (st_ok & nat.STATUS_BAD_INDEX).sum() == 0
    assert [(int(x) & nat.STATUS_BAD_INDEX) != 0 for x in st_bad] == [False, True, True, False]
    for p in (0, 3):
        assert torch.equal(s_ok[p], s_bad[p]), p
    if n > nat.TRANSMISSION_NMAX:
        return
    # fused transmission sweep: T = NaN and the flag for the bad points
    from dpnegf_amd import workloads as wl
    from dpnegf_amd.negf.recursive_green_cal import BTDSystems
    hd, sd, hl, hu, sl, su = wl.random_btd([n, n, n], seed=3)
    to = lambda xs: [x.to(dev) for x in xs]
    bs = BTDSystems([(to(hd), to(hu), to(hl), to(sd), to(su), to(sl))], dev)
    sig = s_ok.to(dev).contiguous()
    T = torch.empty(4, dtype=torch.float64, device=dev)
    st = torch.zeros(4, dtype=torch.int32, device=dev)
    work = torch.empty(nat.load().dpnegf_rgf_transmission_workspace_elems(3, bs._h_bsz, 4),
                       dtype=torch.complex128, device=dev)
    for idx in (good, bad):
        nat.call("dpnegf_rgf_transmission", nat.stream_ptr(dev), nat.DPNEGF_FP64, 3, bs._h_bsz,
                 nat.ptr(bs.bsz), nat.ptr(bs.arena), bs.stride, 1, nat.ptr(idx), 4,
                 nat.ptr((e + 1e-4j).contiguous()), nat.ptr(sig), n, nat.ptr(sig), n, int(bs.ortho),
                 nat.ptr(T), nat.ptr(work), nat.ptr(st))
        torch.cuda.synchronize()
        if idx is good:
            T_ok = T.cpu().clone()
            assert st.sum().item() == 0
    Tb, stb = T.cpu(), st.cpu()
    assert torch.isnan(Tb[1]) and torch.isnan(Tb[2])
    assert stb.tolist() == [0, nat.STATUS_BAD_INDEX, nat.STATUS_BAD_INDEX, 0]
    assert Tb[0] == T_ok[0] and Tb[3] == T_ok[3]
