"""Batch-shape autotuner for the block-operation transmission sweep.

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
"""
import os

import torch

from .. import _native as nat

C128 = torch.complex128
CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)
DEFAULT_CHUNK = 4096
_cache = {}


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
