"""Autotuner: batch shape of the block-operation transmission sweep, and the
kernel shapes (product tiles, inverse form) of the block-operation path.

Blocks above 64 orbitals go through ``dpnegf_rgf_transmission_blocked``, one
launch sequence per chunk of points.  Larger chunks fill more of the GPU per
batched product and inverse, but their workspace grows linearly with the chunk
(``dpnegf_autotune_query`` gives the largest chunk that fits a memory budget),
and past the point where one chunk fills the chip nothing is gained.  The
reference leaves the chunk to the user (``runner/NEGF.py:602-631``).

``blocked_chunk`` times the candidate chunk sizes that fit on a three-block
prefix of the caller's own system (the per-block cost is what the chunk size
changes), keeps the fastest per (device, block sizes, overlap pattern) for the
life of the process, and returns it.  Timing runs on the caller's current
stream.  The sweep uses it only when ``DPNEGF_AUTOTUNE=1`` (``enabled()``);
otherwise the chunk is ``default_chunk``: DEFAULT_CHUNK, or the largest chunk
that fits in half the free memory when that is smaller.

``tune_kernels`` picks kernel shapes.  The library looks each batched product
(m, n, k) and batched inverse n up in a process-wide table before it launches
it (``dpnegf_tune_set``); keys without an entry keep the built-in choice (64-
or 32-wide output tiles by how full the last tile is; the LDS blocked inverse
for 17-96 orbitals, the register-resident one elsewhere).  ``tune_kernels``
runs the caller's path once with the library recording the keys it looks up,
times every candidate for each key at the caller's batch size on random
operands, and installs the fastest where it beats the built-in choice by more
than ``MIN_GAIN``.  Every candidate gives the same result to rounding.  With
``DPNEGF_AUTOTUNE=1`` the blocked transmission sweep and the batched lead
decimation tune their shapes on first use (``tune_blocked``, ``tune_lead``).
"""
import ctypes
import os

import torch

from .. import _native as nat

C128 = torch.complex128
CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)
DEFAULT_CHUNK = 4096
MIN_GAIN = 0.03
_cache = {}
_kernel_cache = {}


def enabled():
    return os.environ.get("DPNEGF_AUTOTUNE", "0") == "1"


def max_batch(sizes, mem_bytes):
    """Largest points per blocked call whose buffers fit in ``mem_bytes``."""
    r = int(nat.load().dpnegf_autotune_query(len(sizes), nat.host_i32(list(sizes)),
                                             int(mem_bytes)))
    if r < 0:
        raise ValueError(nat.load().dpnegf_last_error().decode())
    return r


def default_chunk(systems, mem_fraction=0.5):
    """DEFAULT_CHUNK points per blocked call, or fewer when the buffers of
    that many would not fit in ``mem_fraction`` of the free device memory
    (blocks of several hundred orbitals)."""
    free, _ = torch.cuda.mem_get_info(systems.device)
    bmax = max_batch(systems.sizes, int(free * mem_fraction))
    if bmax < 1:
        raise torch.cuda.OutOfMemoryError("no batch of the blocked sweep fits in free memory")
    return min(DEFAULT_CHUNK, bmax)


def _time_chunk(systems, s, B, trial_blocks):
    """Seconds per point of one blocked launch of B points over the first
    ``trial_blocks`` blocks of system ``s`` (the arena prefix is a valid arena
    for those blocks)."""
    dev = systems.device
    sizes = systems.sizes[:trial_blocks]
    K = len(sizes)
    h_bsz = nat.host_i32(list(sizes))
    arena = systems.arena[s * systems.stride:(s + 1) * systems.stride]
    nL, nR = min(4, sizes[0]), min(4, sizes[-1])
    z = torch.linspace(-1.0, 1.0, B, dtype=torch.float64, device=dev).to(C128) + 1e-3j
    sigL = torch.zeros(B, nL, nL, dtype=C128, device=dev)
    sigR = torch.zeros(B, nR, nR, dtype=C128, device=dev)
    T = torch.empty(B, dtype=torch.float64, device=dev)
    st = torch.zeros(B, dtype=torch.int32, device=dev)
    work = torch.empty(nat.load().dpnegf_rgf_transmission_blocked_workspace_elems(K, h_bsz, B),
                       dtype=C128, device=dev)

    def launch():
        nat.call("dpnegf_rgf_transmission_blocked", nat.stream_ptr(dev), K, h_bsz,
                 nat.ptr(arena), B, nat.ptr(z), nat.ptr(sigL), nL, nat.ptr(sigR), nR,
                 int(systems.ortho), nat.ptr(T), nat.ptr(work), nat.ptr(st))

    launch()  # warm-up (first-use code loading)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    launch()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / B


def blocked_chunk(systems, P, mem_fraction=0.5, candidates=CANDIDATES, trial_blocks=3):
    """Points per ``dpnegf_rgf_transmission_blocked`` call for ``systems``
    (a BTDSystems) and ``P`` points in total."""
    dev = systems.device
    free, _ = torch.cuda.mem_get_info(dev)
    bmax = max_batch(systems.sizes, int(free * mem_fraction))
    if bmax < 1:
        raise torch.cuda.OutOfMemoryError("no batch of the blocked sweep fits in free memory")
    fits = [c for c in candidates if c <= bmax] or [bmax]
    # no candidate beyond the first that covers all points: larger ones only add padding time
    cands = tuple([c for c in fits if c < P] + [c for c in fits if c >= P][:1])
    key = (torch.cuda.get_device_name(dev), tuple(systems.sizes), bool(systems.ortho), cands)
    if key in _cache:
        return _cache[key]
    trial = min(trial_blocks, len(systems.sizes))
    best = min(cands, key=lambda B: _time_chunk(systems, 0, B, trial))
    _cache[key] = best
    return best


# -- kernel shapes -------------------------------------------------------------------
def record_shapes(run):
    """Keys (kind, m, n, k) of the products and inverses ``run()`` launches
    through the library, in first-seen order."""
    lib = nat.load()
    lib.dpnegf_tune_record(1)
    try:
        run()
        torch.cuda.synchronize()
    finally:
        lib.dpnegf_tune_record(0)
    cnt = int(lib.dpnegf_tune_recorded(None, 0))
    buf = (ctypes.c_int32 * max(4 * cnt, 1))()
    lib.dpnegf_tune_recorded(ctypes.cast(buf, ctypes.c_void_p), cnt)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(cnt)]


def _candidates(key):
    kind, m, _, _ = key
    if kind == nat.TUNE_GEMM_TILE:
        return (nat.TILE_32x32, nat.TILE_32x64, nat.TILE_64x32, nat.TILE_64x64)
    # the Schur form (97-128) is not timed: dpnegf_zinv_batched has no workspace
    # for it, and its key keeps the built-in choice
    return (nat.INV_REGISTER, nat.INV_LDS) if m <= 96 else (nat.INV_REGISTER,)


def _operands(key, B, dev, seed=0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    kind, m, n, k = key

    def rnd(*shape):
        return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                             torch.randn(*shape, generator=g, dtype=torch.float64)).to(dev)

    if kind == nat.TUNE_GEMM_TILE:
        return rnd(B, m, k), rnd(B, k, n), torch.empty(B, m, n, dtype=C128, device=dev)
    # diagonally dominant: the inverse is again well conditioned, so repeated
    # in-place inversion stays finite
    eye = torch.eye(m, dtype=C128, device=dev)
    return (rnd(B, m, m) * (0.5 / m) + 2.0 * eye,)


def _launcher(key, ops, B, dev):
    kind, m, n, k = key
    st = nat.stream_ptr(dev)
    if kind == nat.TUNE_GEMM_TILE:
        A, Bm, C = ops
        return lambda: nat.call("dpnegf_zgemm_batched", st, B, m, n, k, nat.ptr(A), k, m * k, 0,
                                nat.ptr(Bm), n, k * n, 0, 1.0, 0.0, 0.0, 0.0, nat.ptr(C), n, m * n)
    (A,) = ops
    return lambda: nat.call("dpnegf_zinv_batched", st, nat.ptr(A), m, B, None)


def time_choice(key, choice, B, dev, reps=5, ops=None):
    """Median milliseconds per launch of the product or inverse ``key`` at
    batch B with ``choice`` installed (0: the built-in choice), over ``reps``
    separately timed launches.  Leaves the table entry of ``key`` as it found it."""
    lib = nat.load()
    old = int(lib.dpnegf_tune_get(*key))
    ops = ops if ops is not None else _operands(key, B, dev)
    launch = _launcher(key, ops, B, dev)
    nat.call("dpnegf_tune_set", *key, int(choice))
    try:
        launch()  # warm-up
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for e0, e1 in ev:
            e0.record()
            launch()
            e1.record()
        ev[-1][1].synchronize()
    finally:
        nat.call("dpnegf_tune_set", *key, old)
    return float(sorted(e0.elapsed_time(e1) for e0, e1 in ev)[reps // 2])


def tune_keys(keys, B, dev, reps=5, min_gain=MIN_GAIN, install=True):
    """Times every candidate of each key at batch B; returns
    {key: (choice, {choice: ms})}, choice 0 where the built-in choice is kept
    (no candidate beats it by more than ``min_gain``), and installs the
    non-zero choices when ``install``."""
    lib = nat.load()
    out = {}
    for key in keys:
        if key[0] == nat.TUNE_INVERSE and key[1] > 128:
            continue
        builtin = int(lib.dpnegf_tune_builtin(*key))
        ops = _operands(key, B, dev)
        times = {c: time_choice(key, c, B, dev, reps, ops) for c in _candidates(key)}
        best = min(times, key=times.get)
        choice = (best if best != builtin and builtin in times
                  and times[best] < (1.0 - min_gain) * times[builtin] else 0)
        out[key] = (choice, times)
        del ops
        if install:
            nat.call("dpnegf_tune_set", *key, int(choice))
    return out


def tune_kernels(run, B, dev, reps=5, min_gain=MIN_GAIN, cache_key=None):
    """Records the keys ``run()`` launches and tunes them at batch B (see
    ``tune_keys``); with ``cache_key`` a second call for the same key and B
    returns the first result without running anything."""
    ck = None
    if cache_key is not None:
        ck = (torch.cuda.get_device_name(dev), cache_key, int(B))
        if ck in _kernel_cache:
            return _kernel_cache[ck]
    res = tune_keys(record_shapes(run), B, dev, reps, min_gain)
    if ck is not None:
        _kernel_cache[ck] = res
    return res


def tune_blocked(systems, B, sample=8):
    """Kernel shapes of the blocked transmission sweep of ``systems`` at B
    points per call (recorded on ``sample`` points of system 0)."""
    dev = systems.device
    K = len(systems.sizes)
    h_bsz = nat.host_i32(list(systems.sizes))
    arena = systems.arena[:systems.stride]
    n0, nK = systems.sizes[0], systems.sizes[-1]

    def run():
        z = torch.linspace(-1.0, 1.0, sample, dtype=torch.float64, device=dev).to(C128) + 1e-3j
        sigL = torch.zeros(sample, n0, n0, dtype=C128, device=dev)
        sigR = torch.zeros(sample, nK, nK, dtype=C128, device=dev)
        T = torch.empty(sample, dtype=torch.float64, device=dev)
        st = torch.zeros(sample, dtype=torch.int32, device=dev)
        work = torch.empty(nat.load().dpnegf_rgf_transmission_blocked_workspace_elems(
            K, h_bsz, sample), dtype=C128, device=dev)
        nat.call("dpnegf_rgf_transmission_blocked", nat.stream_ptr(dev), K, h_bsz, nat.ptr(arena),
                 sample, nat.ptr(z), nat.ptr(sigL), n0, nat.ptr(sigR), nK, int(systems.ortho),
                 nat.ptr(T), nat.ptr(work), nat.ptr(st))

    return tune_kernels(run, B, dev, cache_key=("blocked", tuple(systems.sizes),
                                                bool(systems.ortho)))


def tune_lead(run_sample, n, B, dev):
    """Kernel shapes of the batched decimation of an n-orbital lead at B points
    per call; ``run_sample()`` runs that decimation on a few points."""
    return tune_kernels(run_sample, B, dev, cache_key=("lead", int(n)))
