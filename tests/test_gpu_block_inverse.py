"""Batched block inverse through the C ABI (dpnegf_zinv_batched) against
numpy.linalg.inv: every size class of the kernels (up to 16, 32, 64 and 128),
blocks that need row pivoting, and the singular flag of one block in a
batch.  Bar: relative error 1e-10 against numpy and
residual |A X - I| below 1e-11 (FP64)."""
import numpy as np
import pytest
import torch

from dpnegf_amd import _native as nat

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda")


def _inv_gpu(A):
    dev = _dev()
    t = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    st = torch.zeros(A.shape[0], dtype=torch.int32, device=dev)
    print(f"case zinv n={A.shape[1]} B={A.shape[0]}", flush=True)
    nat.call("dpnegf_zinv_batched", nat.stream_ptr(dev), nat.ptr(t), A.shape[1], A.shape[0],
             nat.ptr(st))
    torch.cuda.synchronize()
    return t.cpu().numpy(), st.cpu().numpy()


def _check(A, X):
    ref = np.linalg.inv(A)
    n = A.shape[-1]
    for b in range(A.shape[0]):
        rel = np.abs(X[b] - ref[b]).max() / np.abs(ref[b]).max()
        res = np.abs(A[b] @ X[b] - np.eye(n)).max()
        assert rel < 1e-10 and res < 1e-11, (b, rel, res)


SIZES = [1, 7, 16, 33, 64, 65, 72, 80, 84, 90, 96, 100, 113, 120, 127, 128]


@pytest.mark.parametrize("n", SIZES)
def test_block_inverse_matches_numpy(n):
    rng = np.random.default_rng(100 + n)
    B = 5
    A = rng.normal(size=(B, n, n)) + 1j * rng.normal(size=(B, n, n))
    A += np.sqrt(n) * np.eye(n)
    X, st = _inv_gpu(A)
    assert (st == 0).all()
    _check(A, X)


@pytest.mark.parametrize("n", [20, 64, 84, 128])
def test_block_inverse_pivots_rows(n):
    """Rows of a diagonally dominant block in a random order: the largest
    entry of every column sits off the diagonal, and the diagonal is small."""
    rng = np.random.default_rng(7 + n)
    B = 3
    A = np.empty((B, n, n), complex)
    for b in range(B):
        D = 4.0 * np.eye(n) + 0.1 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        A[b] = D[rng.permutation(n)]
    X, st = _inv_gpu(A)
    assert (st == 0).all()
    _check(A, X)


@pytest.mark.parametrize("n", [72, 128])
def test_block_inverse_flags_a_singular_block(n):
    """A zero column makes block 1 exactly singular: its status bit is set,
    and the other blocks of the batch are inverted as usual."""
    rng = np.random.default_rng(n)
    B = 3
    A = rng.normal(size=(B, n, n)) + 1j * rng.normal(size=(B, n, n)) + np.sqrt(n) * np.eye(n)
    A[1][:, n // 3] = 0.0
    X, st = _inv_gpu(A)
    assert st[1] & 1 and st[0] == 0 and st[2] == 0
    _check(A[[0, 2]], X[[0, 2]])
