"""Host-side argument checks that keep every library call inside its buffers
(no device needed): batch broadcasting of per-energy inputs and the per-point
inputs of the fused transmission entry."""
import types

import pytest
import torch

from dpnegf_amd.negf._btd import batch_of, corner_self_energy, pack_batched_blocks
from dpnegf_amd.negf.recursive_green_cal import _check_points

C = torch.complex128


def _m(*shape):
    return torch.randn(*shape, dtype=C)


def test_batch_of_broadcasts_2d_and_single_batch_and_rejects_others():
    B = 4
    assert batch_of(_m(3, 3), B).shape == (B, 3, 3)
    one = _m(1, 3, 3)
    out = batch_of(one, B)
    assert out.shape == (B, 3, 3) and torch.equal(out[2], one[0])
    full = _m(B, 3, 3)
    assert batch_of(full, B) is full
    with pytest.raises(ValueError, match="batch size 2"):
        batch_of(_m(2, 3, 3), B)
    with pytest.raises(ValueError, match="2-D or"):
        batch_of(_m(3), B)


def test_pack_batched_blocks_checks_batch_and_shape():
    B = 3
    arena = pack_batched_blocks([_m(2, 2), _m(1, 3, 3)], B, "cpu")
    assert arena.numel() == B * (4 + 9)
    with pytest.raises(ValueError, match="batch size 2"):
        pack_batched_blocks([_m(2, 2, 2)], B, "cpu")
    with pytest.raises(ValueError, match="expected"):
        pack_batched_blocks([_m(2, 2)], B, "cpu", shapes=[(3, 3)])


def test_corner_self_energy_broadcasts_and_rejects_short_batches():
    B = 5
    se, r = corner_self_energy(_m(1, 4, 4), 6, B, False, "cpu")
    assert se.shape == (B, 4, 4) and r == 4
    se, r = corner_self_energy(_m(B, 8, 8), 6, B, True, "cpu")  # trimmed to the corner
    assert se.shape == (B, 6, 6) and r == 6
    with pytest.raises(ValueError, match="batch size 3"):
        corner_self_energy(_m(3, 4, 4), 6, B, False, "cpu")


def _systems(nsys=2, sizes=(4, 5)):
    return types.SimpleNamespace(sizes=list(sizes), nsys=nsys, device=torch.device("cpu"))


def test_check_points_converts_int64_sys_idx_and_checks_its_range():
    P = 6
    bs = _systems()
    z = _m(P)
    sL, sR = _m(P, 3, 3), _m(P, 5, 5)
    idx = torch.arange(P) % 2  # int64, as torch.arange gives
    z2, sL2, sR2, idx2 = _check_points(bs, z, sL, sR, idx)
    assert idx2.dtype == torch.int32 and idx2.is_contiguous()
    assert idx2.tolist() == idx.tolist()
    with pytest.raises(ValueError, match=r"\[0, 2\)"):
        _check_points(bs, z, sL, sR, torch.tensor([0, 1, 2, 0, 1, 0]))
    with pytest.raises(ValueError, match=r"\[0, 2\)"):
        _check_points(bs, z, sL, sR, torch.tensor([0, -1, 1, 0, 1, 0]))
    with pytest.raises(ValueError, match="sys_idx is required"):
        _check_points(bs, z, sL, sR, None)
    with pytest.raises(ValueError, match=r"sys_idx must be \[P=6\]"):
        _check_points(bs, z, sL, sR, idx[:4])
    with pytest.raises(TypeError, match="integer"):
        _check_points(bs, z, sL, sR, idx.double())


def test_check_points_rejects_short_wrong_dtype_and_oversized_inputs():
    P = 4
    bs = _systems(nsys=1)
    z, sL, sR = _m(P), _m(P, 4, 4), _m(P, 5, 5)
    _check_points(bs, z, sL, sR, None)
    with pytest.raises(ValueError, match="sigL must be"):
        _check_points(bs, z, sL[:3], sR, None)
    with pytest.raises(ValueError, match="sigR must be"):
        _check_points(bs, z, sL, _m(P, 6, 6), None)  # larger than the last block
    with pytest.raises(TypeError, match="z must be"):
        _check_points(bs, z.to(torch.complex64), sL, sR, None)
    with pytest.raises(ValueError, match="z must be 1-D"):
        _check_points(bs, z.reshape(2, 2), sL, sR, None)
    # non-contiguous inputs come back contiguous
    sLt = _m(P, 4, 4).transpose(1, 2)
    assert not sLt.is_contiguous()
    assert _check_points(bs, z, sLt, sR, None)[1].is_contiguous()
