"""The C-ABI library loads without a GPU and exports every symbol the public
header declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from dpnegf_amd import _native as nat

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "dpnegf_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dpnegf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_bound_symbols():
    assert set(_declared()) == set(nat.EXPORTED)


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nat.lib_path())
    for name in _declared():
        assert hasattr(lib, name), name
    nat.load()
    declared = int(re.search(r"#define DPNEGF_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert nat.load().dpnegf_abi_version() == declared == nat.ABI_VERSION


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_argument_errors_are_reported_without_a_device():
    # invalid sizes are rejected on the host before anything touches the GPU
    lib = nat.load()
    bsz = nat.host_i32([80])
    rc = lib.dpnegf_rgf_transmission(None, 0, 1, bsz, None, None, 0, 1, None, 1, None, None, 1,
                                     None, 1, 1, None, None, None)
    assert rc == -2 and b"exceeds" in lib.dpnegf_last_error()
    # a missing work buffer is refused like any other null device pointer
    ok = nat.host_i32([48, 48])
    rc = lib.dpnegf_rgf_transmission(None, 0, 2, ok, 8, 8, 0, 1, None, 1, 8, 8, 48, 8, 48, 1, 8,
                                     None, None)
    assert rc == -1 and b"null" in lib.dpnegf_last_error()
    assert lib.dpnegf_rgf_transmission_workspace_elems(2, ok, 10) == 10 * 48 * 48
    assert lib.dpnegf_rgf_transmission_workspace_elems(2, nat.host_i32([40, 48]), 3) == 3 * 40 * 48
    assert lib.dpnegf_rgf_transmission_workspace_elems(1, bsz, 3) == -1
    rc = lib.dpnegf_lead_self_energy(None, 0, 4, 0, None, 0, 1, None, 1, None, 1e-5, 1e-40,
                                     None, None, None, 1, None, None, None)
    assert rc == -1
    with pytest.raises(ValueError):
        nat.call("dpnegf_rgf_full", None, 0, None, None, 1, None, None, 0, None, 0, 1,
                 *([None] * 23))
    for n in (0, 129):
        assert lib.dpnegf_zinv_batched(None, None, n, 1, None) < 0
        assert b"outside" in lib.dpnegf_last_error()
    assert lib.dpnegf_zinv_batched(None, None, 16, 1, None) < 0  # null pointer
    assert lib.dpnegf_zinv_batched(None, None, 16, 0, None) == 0  # empty batch


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_autotune_query_fits_the_workspace():
    """The largest batch the query returns fits the budget and the next one
    does not (workspace plus z, both self-energies, T and status per point)."""
    lib = nat.load()
    for sizes in ([72, 80, 96], [168, 120, 144, 96], [256] * 4):
        h = nat.host_i32(sizes)
        K = len(sizes)

        def need(B):
            w = lib.dpnegf_rgf_transmission_blocked_workspace_elems(K, h, B) * 16
            return w + B * ((1 + sizes[0] ** 2 + sizes[-1] ** 2) * 16 + 8 + 4)

        for mem in (10 ** 7, 3 * 10 ** 8, 5 * 10 ** 9):
            b = lib.dpnegf_autotune_query(K, h, mem)
            assert 0 <= b <= 65535
            assert need(b) <= mem
            if b < 65535:
                assert need(b + 1) > mem
    assert lib.dpnegf_autotune_query(0, nat.host_i32([1]), 10 ** 6) < 0


def test_every_package_module_imports():
    """Import errors surface on the CPU, not first on a GPU host."""
    import importlib
    import pkgutil

    import dpnegf_amd
    for m in pkgutil.walk_packages(dpnegf_amd.__path__, "dpnegf_amd."):
        importlib.import_module(m.name)


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_checkpointed_adjoint_workspace():
    """The checkpointed adjoint's work buffer follows its stated formula and is
    several times smaller than the full-storage one at 64 blocks of 64."""
    lib = nat.load()
    for sizes in ([64] * 64, [6, 9, 7, 8, 6, 10, 7, 9, 8, 6, 7, 9], [200, 90, 256, 120]):
        K, h = len(sizes), nat.host_i32(sizes)
        B, nmax, n0 = 3, max(sizes), sizes[0]
        inv = lib.dpnegf_rgf_grad_ckpt_workspace_elems(K, h, B, K) - B * (
            sum(sizes[c] ** 2 + n0 * sizes[c] for c in range(0, K, K)) - n0 * n0
            + (2 * K + 18) * nmax ** 2)
        assert inv >= 0  # the large-block inverse's workspace, independent of the stride
        for s in (1, 2, 3, 8, K):
            ck = sum(sizes[c] ** 2 + n0 * sizes[c] for c in range(0, K, s)) - n0 * n0
            want = B * (ck + (2 * s + 18) * nmax ** 2) + inv
            assert lib.dpnegf_rgf_grad_ckpt_workspace_elems(K, h, B, s) == want, (sizes, s)
    h = nat.host_i32([64] * 64)
    full = lib.dpnegf_rgf_grad_workspace_elems(64, h, 16)
    ck = lib.dpnegf_rgf_grad_ckpt_workspace_elems(64, h, 16, 8)
    assert ck * 4 < full
    assert lib.dpnegf_rgf_grad_ckpt_workspace_elems(64, h, 16, 0) == -1
    rc = lib.dpnegf_rgf_transmission_grad_ckpt(None, 1, nat.host_i32([4]), None, 1, None, None, 1,
                                               None, 1, 1, None, None, None, None, None, None,
                                               None, None, None, 0, None, None)
    assert rc < 0 and b"stride" in lib.dpnegf_last_error()


def test_checkpoint_stride_choices():
    from dpnegf_amd.negf.adjoint import checkpoint_stride
    assert checkpoint_stride("sqrt", [8] * 10, 2, None) == 4
    assert checkpoint_stride("sqrt", [8] * 9, 2, None) == 3
    assert checkpoint_stride("sqrt", [8], 2, None) == 1
    assert checkpoint_stride(5, [8] * 10, 2, None) == 5
    assert checkpoint_stride(0, [8] * 10, 2, None) == 0
    with pytest.raises(ValueError):
        checkpoint_stride(-1, [8] * 10, 2, None)


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_full_output_workspace_covers_the_library_size(monkeypatch):
    """The work buffer recursive_gf (and the density) hand to dpnegf_rgf_full
    is at least what the library declares; above 128 orbitals that includes
    the two-level inverse's 3 B n^2 elements (a 5 B n^2 buffer was written
    past its end there).  The library call itself is intercepted."""
    import torch
    from dpnegf_amd.negf import recursive_green_cal as rgc
    passed = []
    seen = {}
    monkeypatch.setattr(nat, "ptr", lambda t: passed.append(t) or ctypes.c_void_p(0))
    monkeypatch.setattr(nat, "stream_ptr", lambda device=None: ctypes.c_void_p(0))
    monkeypatch.setattr(nat, "call", lambda name, *args: seen.__setitem__(name, list(passed)))
    lib = nat.load()
    for sizes in ([256, 192, 240], [144, 168, 136], [8, 12], [130]):
        B = 3
        passed.clear()
        z = torch.zeros(B, dtype=torch.complex128)
        sig = torch.zeros(B, 4, 4, dtype=torch.complex128)
        rgc._run_full(z, torch.zeros(1, dtype=torch.complex128), sizes, True, sig, 4, sig, 4, True)
        work, status = seen["dpnegf_rgf_full"][-2:]
        need = lib.dpnegf_rgf_full_workspace_elems(len(sizes), nat.host_i32(sizes), B)
        assert status.numel() == B
        assert work.numel() >= need, (sizes, work.numel(), need)


def test_entry_module_helpers_run_without_a_gpu():
    """smoke()'s host-side helpers resolve their imports (the GPU part runs at
    round end on the device)."""
    import importlib
    import torch
    entry = importlib.import_module("__graft_entry__")
    t = entry._f64([0.5, -1.0])
    assert t.dtype == torch.float64 and t.tolist() == [0.5, -1.0]
    assert callable(entry.build) and callable(entry.smoke)


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_block_limit_is_1024_orbitals():
    """Blocks up to 1024 orbitals pass the size check (the inverse recurses
    over diagonal halves above 128) and 1025 is refused before any device
    call; the inverse's workspace stays 3 B n^2 at every depth."""
    lib = nat.load()
    assert nat.FULL_NMAX == 1024
    rc = lib.dpnegf_rgf_transmission_blocked(None, 2, nat.host_i32([64, 1025]), None, 1, None,
                                             None, 1, None, 1, 1, None, None, None)
    assert rc < 0 and b"1025" in lib.dpnegf_last_error()
    rc = lib.dpnegf_lead_self_energy(None, 1025, 1025, 0, None, 0, 1, None, 1, None, 1e-5,
                                     1e-40, None, None, None, 1, None, None, None)
    assert rc < 0 and b"exceed" in lib.dpnegf_last_error()
    for n in (300, 520, 1024):
        h = nat.host_i32([n])
        assert lib.dpnegf_rgf_full_workspace_elems(1, h, 2) == 2 * 8 * n * n
        assert lib.dpnegf_rgf_transmission_blocked_workspace_elems(1, h, 2) == 2 * 11 * n * n


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_points_per_call_fit_the_free_memory(monkeypatch):
    """The chunks of the large-lead decimation, of the full-output sweep (the
    density) and of the blocked transmission fit, with their outputs, in half
    the free memory, and one more point would not (or the chunk is P or the
    per-call limit)."""
    import torch
    from dpnegf_amd.negf import autotune
    from dpnegf_amd.negf.recursive_green_cal import full_points_per_call
    from dpnegf_amd.negf.surface_green import points_per_call
    lib = nat.load()
    free = {"v": 0}
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (free["v"], 2 * free["v"]))
    assert points_per_call(48, 48, 5000, None) == 5000  # fused kernel: no chunking
    for n, a in ((300, 300), (520, 300), (1024, 1024)):
        def need(c):
            return 16 * (lib.dpnegf_lead_self_energy_workspace_elems(n, a, c)
                         + c * (n * n + a * a)) + 32 * c
        for budget in (10 ** 9, 7 * 10 ** 10, 2 * 10 ** 11):
            free["v"] = budget
            c = points_per_call(n, a, 4096, None)
            assert need(c) <= budget // 2, (n, budget, c)
            assert c in (4096, 65535) or need(c + 1) > budget // 2
    free["v"] = 10 ** 6
    with pytest.raises(torch.cuda.OutOfMemoryError):
        points_per_call(1024, 1024, 10, None)
    for sizes, lc in (([520, 300, 520], True), ([1024] * 4, False), ([64, 96], True)):
        K = len(sizes)
        nsq = sum(n * n for n in sizes)
        nx = sum(sizes[q] * sizes[q + 1] for q in range(K - 1))
        nlc = sum(n * sizes[-1] for n in sizes) if lc else 0

        def need_full(c):
            w = lib.dpnegf_rgf_full_workspace_elems(K, nat.host_i32(sizes), c)
            return 16 * (w + c * (2 * nsq + 2 * nx + nlc + sizes[0] * sizes[-1])) + 4 * c
        for budget in (10 ** 10, 2 * 10 ** 11):
            free["v"] = budget
            c = full_points_per_call(sizes, lc, 161, None)
            assert need_full(c) <= budget // 2, (sizes, budget, c)
            assert c == 161 or need_full(c + 1) > budget // 2

    class _Sys:
        device = None
        sizes = [1024, 1024]
    free["v"] = 10 ** 11
    c = autotune.default_chunk(_Sys)
    assert 1 <= c < autotune.DEFAULT_CHUNK
    free["v"] = 10 ** 13
    assert autotune.default_chunk(_Sys) == autotune.DEFAULT_CHUNK


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_lead_records_need_coupling_blocks_only_in_mode_0():
    """With several lead systems, a record of 4 n^2 elements (no H_DL, S_DL)
    is too short when Sigma needs the coupling (mode 0); the stride check
    refuses it before any device work."""
    lib = nat.load()
    fake = ctypes.c_void_p(16)  # never dereferenced: the call fails at the stride check
    rc = lib.dpnegf_lead_self_energy(None, 4, 4, 0, fake, 4 * 16, 2, fake, 1, fake, 1e-5,
                                     1e-40, fake, fake, fake, 1, None, None, None)
    assert rc < 0 and b"lead_stride" in lib.dpnegf_last_error()


def test_lead_kpoints_stay_float64():
    """Regression: a k-point list went through torch's default float32, which
    moved the unfolded Bloch points by ~1e-8 and the self-energy by ~1e-7."""
    import torch
    from dpnegf_amd.negf.lead_property import _kvec
    assert _kvec([0, 0.37, 0]).dtype == torch.float64 and _kvec([0, 0.37, 0])[1].item() == 0.37
    t = torch.tensor([0.1, 0.37, 0.0], dtype=torch.float64)
    assert torch.equal(_kvec(t), t)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says why,
    before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_tuning_table_without_a_device():
    """The kernel-shape table is host state: entries are validated, read back,
    removed by choice 0 and cleared; nothing is recorded while recording is off."""
    lib = nat.load()
    lib.dpnegf_tune_clear()
    try:
        assert lib.dpnegf_tune_set(nat.TUNE_GEMM_TILE, 96, 168, 120, nat.TILE_32x64) == 0
        assert lib.dpnegf_tune_get(nat.TUNE_GEMM_TILE, 96, 168, 120) == nat.TILE_32x64
        assert lib.dpnegf_tune_get(nat.TUNE_GEMM_TILE, 96, 168, 121) == 0
        assert lib.dpnegf_tune_set(nat.TUNE_GEMM_TILE, 96, 168, 120, 5) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_GEMM_TILE, 0, 168, 120, 1) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 90, 0, 0, nat.INV_LDS) == 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 97, 0, 0, nat.INV_LDS) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 129, 0, 0, nat.INV_REGISTER) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 96, 0, 0, nat.INV_SCHUR) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 100, 0, 0, nat.INV_SCHUR) == 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 100, 0, 0, nat.INV_REGISTER) == 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 100, 0, 0, 4) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 90, 1, 0, nat.INV_LDS) != 0
        assert lib.dpnegf_tune_set(2, 1, 1, 1, 1) != 0
        assert lib.dpnegf_tune_set(nat.TUNE_INVERSE, 90, 0, 0, 0) == 0
        assert lib.dpnegf_tune_get(nat.TUNE_INVERSE, 90, 0, 0) == 0
        lib.dpnegf_tune_clear()
        assert lib.dpnegf_tune_get(nat.TUNE_GEMM_TILE, 96, 168, 120) == 0
        lib.dpnegf_tune_record(1)
        lib.dpnegf_tune_record(0)
        assert lib.dpnegf_tune_recorded(None, 0) == 0
    finally:
        lib.dpnegf_tune_clear()


def test_tuning_builtin_choices():
    """The built-in choice the autotuner compares against: 32-wide tiles in a
    dimension below 256 whose last 64-wide tile is at most half full, the LDS
    inverse for 17..96 orbitals, the Schur-complement inverse for 97..128."""
    lib = nat.load()
    assert lib.dpnegf_tune_builtin(nat.TUNE_GEMM_TILE, 96, 96, 96) == nat.TILE_32x32
    assert lib.dpnegf_tune_builtin(nat.TUNE_GEMM_TILE, 96, 168, 96) == nat.TILE_32x64
    assert lib.dpnegf_tune_builtin(nat.TUNE_GEMM_TILE, 120, 144, 8) == nat.TILE_64x32
    assert lib.dpnegf_tune_builtin(nat.TUNE_GEMM_TILE, 480, 128, 64) == nat.TILE_64x64
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 16, 0, 0) == nat.INV_REGISTER
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 17, 0, 0) == nat.INV_LDS
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 64, 0, 0) == nat.INV_LDS
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 96, 0, 0) == nat.INV_LDS
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 97, 0, 0) == nat.INV_SCHUR
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 128, 0, 0) == nat.INV_SCHUR
    assert lib.dpnegf_tune_builtin(nat.TUNE_INVERSE, 129, 0, 0) == 0
