"""Recursive Green's function on the GPU, ``dpnegf.negf.recursive_green_cal`` interface.

* :func:`recursive_gf` -- same signature, defaults and 14-tuple as the
  reference wrapper (``dpnegf/negf/recursive_green_cal.py:324-477``); every
  block is computed by ``dpnegf_rgf_full`` (batched block operations, all
  quirks of ``recursive_gf_cal`` kept).
* :func:`rgf_transmission` -- new entry: T(E, k) only, one fused kernel
  launch over every (E, k) point (``dpnegf_rgf_transmission``).

Inputs must be CUDA (HIP) tensors; there is no CPU path.
"""
import torch

from .. import _native as nat
from . import autotune
from ._btd import (C128, block_sizes, corner_self_energy, is_orthogonal, pack_arena,
                   pack_batched_blocks, split_blocks)

_MAX_B = 65535


def _squeeze_ans(ans):
    """``_squeeze_ans`` (``recursive_green_cal.py:480-491``)."""
    def sq(x):
        if x is None:
            return None
        if torch.is_tensor(x):
            return x.squeeze(0)
        if isinstance(x, list):
            return [t.squeeze(0) if torch.is_tensor(t) else t for t in x]
        return x
    return tuple(sq(x) for x in ans)


def _energies(energy, E_ref, left_se, right_se, device):
    e = energy + E_ref
    if not torch.is_tensor(e):
        e = torch.as_tensor(e, dtype=C128)
    probe = left_se if isinstance(left_se, torch.Tensor) else right_se
    squeezed = isinstance(probe, torch.Tensor) and probe.ndim == 2
    if e.ndim == 0:
        e = e.reshape(1)
    return e.to(device=device, dtype=C128), squeezed


def recursive_gf(energy, hl, hd, hu, sd, su, sl, left_se, right_se, seP=None, E_ref=0.0, s_in=0,
                 s_out=0, eta=1e-5, need_lesser=False, need_greater=False, need_gr_lc=False,
                 keep_gr_left=True):
    """Retarded (and optionally lesser / greater) Green's function blocks.

    Returns ``(g_trans, gr_lc, grd, grl, gru, gr_left, gnd, gnl, gnu, gin_left,
    gpd, gpl, gpu, gip_left)`` exactly as ``recursive_green_cal.py:299-321``;
    the leading energy dimension is squeezed for scalar callers (2-D
    self-energies), ``grl == gru == []`` for a single block.
    """
    device = hd[0].device
    nat.require_device(*hd)
    e, squeezed = _energies(energy, E_ref, left_se, right_se, device)
    B = e.shape[0]
    z = (e + 1j * eta).contiguous()
    if need_lesser:
        assert isinstance(s_in, list), \
            "need_lesser=True requires s_in to be a list of coupling matrices"
    if need_greater:
        assert isinstance(s_out, list), \
            "need_greater=True requires s_out to be a list of coupling matrices"

    hd_eff = list(hd)
    seP_blocks = None  # energy-dependent probe self-energies ([B, n, n] per block)
    if seP is not None:
        if len(seP) != len(hd):
            raise ValueError(f"seP has {len(seP)} blocks, the device {len(hd)}")
        for i in range(len(hd_eff)):
            sp = seP[i]
            if torch.is_tensor(sp) and sp.ndim == 3 and sp.shape[0] == 1:
                sp = sp[0]  # one matrix broadcast over the batch
            if torch.is_tensor(sp) and sp.ndim == 3:
                if seP_blocks is None:
                    seP_blocks = [torch.zeros_like(h) for h in hd]
                seP_blocks[i] = sp
            elif torch.is_tensor(sp):
                hd_eff[i] = hd_eff[i] + sp.to(device)
            elif sp != 0:
                hd_eff[i] = hd_eff[i] + sp
    sizes = block_sizes(hd_eff)
    K = len(sizes)
    arena = pack_arena(hd_eff, hu, hl, sd, su, sl, device)
    ortho = is_orthogonal(sd, su, sl)
    sigL = sigR = None
    nL = nR = 0
    if isinstance(left_se, torch.Tensor):
        sigL, nL = corner_self_energy(left_se, sizes[0], B, False, device)
    if isinstance(right_se, torch.Tensor):
        sigR, nR = corner_self_energy(right_se, sizes[-1], B, True, device)

    sq = [(n, n) for n in sizes]
    raw = _run_full(z, arena, sizes, ortho, sigL, nL, sigR, nR, need_gr_lc,
                    pack_batched_blocks(s_in, B, device, sq, "s_in") if need_lesser else None,
                    pack_batched_blocks(s_out, B, device, sq, "s_out") if need_greater else None,
                    seP=(pack_batched_blocks(seP_blocks, B, device, sq, "seP")
                         if seP_blocks is not None else None))
    sq_shapes = [(n, n) for n in sizes]
    lo_shapes = [(sizes[q + 1], sizes[q]) for q in range(K - 1)]
    up_shapes = [(sizes[q], sizes[q + 1]) for q in range(K - 1)]
    lc_shapes = [(n, sizes[-1]) for n in sizes]
    gl, grd, grl, gru, lc, g_trans = (raw[k] for k in ("gl", "grd", "grl", "gru", "lc", "g_trans"))
    lesser, greater = raw["lesser"], raw["greater"]
    grd_l = split_blocks(grd, B, sq_shapes)
    grl_l = split_blocks(grl, B, lo_shapes)
    gru_l = split_blocks(gru, B, up_shapes)
    gl_l = split_blocks(gl, B, sq_shapes) if keep_gr_left else None
    lc_l = split_blocks(lc, B, lc_shapes) if need_gr_lc else None
    if need_gr_lc:
        g_trans = lc_l[0]
    ans = [g_trans, lc_l, grd_l, grl_l, gru_l, gl_l]
    for part in (lesser, greater):
        if part is None:
            ans += [None] * 4
        else:
            _, gi, gd, gL, gU = part
            ans += [split_blocks(gd, B, sq_shapes), split_blocks(gL, B, lo_shapes),
                    split_blocks(gU, B, up_shapes), split_blocks(gi, B, sq_shapes)]
    ans = tuple(ans)
    if squeezed:
        ans = _squeeze_ans(ans)
    return ans


def full_points_per_call(sizes, need_lc, P, device, mem_fraction=0.5):
    """Points per ``_run_full`` call (retarded outputs only) whose outputs and
    work buffer fit in ``mem_fraction`` of the free device memory, at most
    ``P`` and ``_MAX_B``."""
    K = len(sizes)
    h_bsz = nat.host_i32(list(sizes))
    lib = nat.load()
    w1 = lib.dpnegf_rgf_full_workspace_elems(K, h_bsz, 1)
    w2 = lib.dpnegf_rgf_full_workspace_elems(K, h_bsz, 2)
    nsq = sum(n * n for n in sizes)
    nx = sum(sizes[q] * sizes[q + 1] for q in range(K - 1))
    nlc = sum(n * sizes[-1] for n in sizes) if need_lc else 0
    per = 16 * ((w2 - w1) + 2 * nsq + 2 * nx + nlc + sizes[0] * sizes[-1]) + 4
    free, _ = torch.cuda.mem_get_info(device)
    fit = (int(free * mem_fraction) - 16 * (2 * w1 - w2)) // per
    if fit < 1:
        raise torch.cuda.OutOfMemoryError("one point of the block sweep does not fit in free "
                                          "device memory")
    return max(1, min(P, fit, _MAX_B))


def _run_full(z, arena, sizes, ortho, sigL, nL, sigR, nR, need_lc, s_in=None, s_out=None,
              seP=None, s_bil=None):
    """One ``dpnegf_rgf_full`` call; returns the raw block-major arenas.
    ``seP``: optional per-energy diagonal addend, a block-major [B, n_q, n_q] arena.
    ``s_bil``: optional block-diagonal source of the bilinear product G S G
    (``out["bilinear"]`` = (s_bil, x_left, X_qq, X_{q+1,q}, X_{q,q+1}))."""
    device = z.device
    B = z.shape[0]
    K = len(sizes)
    if B > _MAX_B:
        raise ValueError(f"at most {_MAX_B} energies per call (got {B}); "
                         "split the energy grid into chunks as the runner does")
    nsq = sum(n * n for n in sizes)
    nx = sum(sizes[q] * sizes[q + 1] for q in range(K - 1))
    nlc = sum(n * sizes[-1] for n in sizes)

    def alloc(count):
        return torch.empty(B * max(count, 1), dtype=C128, device=device)

    out = {"gl": alloc(nsq), "grd": alloc(nsq), "grl": alloc(nx), "gru": alloc(nx),
           "lc": alloc(nlc) if need_lc else None,
           "g_trans": torch.empty(B, sizes[0], sizes[-1], dtype=C128, device=device)}
    out["lesser"] = ((s_in, alloc(nsq), alloc(nsq), alloc(nx), alloc(nx))
                     if s_in is not None else None)
    out["greater"] = ((s_out, alloc(nsq), alloc(nsq), alloc(nx), alloc(nx))
                      if s_out is not None else None)
    out["bilinear"] = ((s_bil, alloc(nsq), alloc(nsq), alloc(nx), alloc(nx))
                       if s_bil is not None else None)
    status = torch.zeros(B, dtype=torch.int32, device=device)
    # the library's own size: it includes the block inverse's workspace for blocks above 128
    nwork = nat.load().dpnegf_rgf_full_workspace_elems(K, nat.host_i32(sizes), B)
    if nwork < 0:
        raise ValueError(nat.load().dpnegf_last_error().decode())
    work = torch.empty(nwork, dtype=C128, device=device)
    args = [nat.stream_ptr(device), K, nat.host_i32(sizes), nat.ptr(arena), B, nat.ptr(z),
            nat.ptr(sigL), nL, nat.ptr(sigR), nR, int(ortho), nat.ptr(seP),
            nat.ptr(out["gl"]), nat.ptr(out["grd"]), nat.ptr(out["grl"]), nat.ptr(out["gru"]),
            nat.ptr(out["lc"]), nat.ptr(out["g_trans"])]
    for part in (out["lesser"], out["greater"], out["bilinear"]):
        args += [nat.ptr(t) for t in part] if part else [nat.ptr(None)] * 5
    args += [nat.ptr(work), nat.ptr(status)]
    nat.call("dpnegf_rgf_full", *args)
    if int(status.max().item()) & nat.STATUS_SINGULAR:
        raise torch.linalg.LinAlgError("singular block in the recursive Green's function sweep")
    return out


class BTDSystems:
    """Device-resident Hamiltonians of one or more k-points with identical
    block sizes, packed once for repeated transmission sweeps."""

    def __init__(self, systems, device=None):
        """``systems``: list of ``(hd, hu, hl, sd, su, sl)`` block lists."""
        if device is None:
            device = systems[0][0][0].device
        nat.require_device(systems[0][0][0].to(device))
        self.device = torch.device(device)
        self.sizes = block_sizes(systems[0][0])
        for s in systems:
            if block_sizes(s[0]) != self.sizes:
                raise ValueError("all systems must share the block sizes")
        arenas = [pack_arena(*s, self.device) for s in systems]
        self.stride = arenas[0].numel()
        self.arena = torch.cat(arenas)
        self.nsys = len(systems)
        self.ortho = all(is_orthogonal(s[3], s[4], s[5]) for s in systems)
        self.bsz = torch.tensor(self.sizes, dtype=torch.int32, device=self.device)
        self._h_bsz = nat.host_i32(self.sizes)


def _check_points(systems, z, sigL, sigR, sys_idx):
    """The per-point inputs of :func:`rgf_transmission` as the library reads
    them: ``z`` [P] complex128, ``sigL`` / ``sigR`` [P, r, r] complex128 with r
    at most the corner block, contiguous, on the systems' device, and
    ``sys_idx`` [P] int32 with every entry naming one of the systems."""
    dev = systems.device
    for name, t in (("z", z), ("sigL", sigL), ("sigR", sigR)):
        if not torch.is_tensor(t) or t.dtype != C128:
            raise TypeError(f"{name} must be a complex128 tensor")
        if t.device.type != dev.type or (dev.index is not None and t.device.index is not None
                                         and t.device.index != dev.index):
            raise ValueError(f"{name} is on {t.device}, the systems on {dev}")
    if z.ndim != 1:
        raise ValueError(f"z must be 1-D [P] (got shape {tuple(z.shape)})")
    P = z.shape[0]
    for name, t, nb in (("sigL", sigL, systems.sizes[0]), ("sigR", sigR, systems.sizes[-1])):
        if t.ndim != 3 or t.shape[0] != P or t.shape[1] != t.shape[2] or not 1 <= t.shape[1] <= nb:
            raise ValueError(f"{name} must be [P={P}, r, r] with r <= {nb} "
                             f"(got shape {tuple(t.shape)})")
    if sys_idx is not None:
        if sys_idx.ndim != 1 or sys_idx.shape[0] != P:
            raise ValueError(f"sys_idx must be [P={P}] (got shape {tuple(sys_idx.shape)})")
        if sys_idx.dtype.is_floating_point or sys_idx.dtype.is_complex:
            raise TypeError("sys_idx must be an integer tensor")
        sys_idx = sys_idx.to(device=dev, dtype=torch.int32).contiguous()
        if P and (int(sys_idx.min()) < 0 or int(sys_idx.max()) >= systems.nsys):
            raise ValueError(f"sys_idx must lie in [0, {systems.nsys})")
    elif systems.nsys > 1:
        raise ValueError("sys_idx is required with more than one system")
    return z.contiguous(), sigL.contiguous(), sigR.contiguous(), sys_idx


def rgf_transmission(systems, z, sigL, sigR, sys_idx=None, precision="fp64", status=None, out=None,
                     chunk=None):
    """T for every point: ``z`` [P] complex (E + E_ref + i eta), ``sigL`` [P, nL, nL],
    ``sigR`` [P, nR, nR] (already trimmed to their corners), ``sys_idx`` [P] int32
    (k-point of each point).  Returns T [P] float64 (and fills ``status``).
    ``precision``: "fp64"; "mixed" (FP32 throughout, about 1e-4 on T of order
    30); or "mixed_refined" (blocks <= 48: FP32 block inverses refined by one
    FP64 Newton step, FP64 products; ``status`` gets STATUS_REFINE_FALLBACK
    where a block was inverted in FP64 instead).
    ``chunk``: points per call of the block-operation sweep (blocks above 64);
    by default ``autotune.default_chunk`` (DEFAULT_CHUNK, fewer when that many
    would not fit in half the free device memory), or
    ``autotune.blocked_chunk``'s pick when ``DPNEGF_AUTOTUNE=1`` (which also
    tunes the sweep's kernel shapes, ``autotune.tune_blocked``)."""
    if not isinstance(systems, BTDSystems):
        raise TypeError("systems must be a BTDSystems")
    dev = systems.device
    z, sigL, sigR, sys_idx = _check_points(systems, z, sigL, sigR, sys_idx)
    P = z.shape[0]
    T = out if out is not None else torch.empty(P, dtype=torch.float64, device=dev)
    if max(systems.sizes) <= nat.TRANSMISSION_NMAX:
        prec = {"fp64": nat.DPNEGF_FP64, "mixed": nat.DPNEGF_MIXED,
                "mixed_refined": nat.DPNEGF_MIXED_REFINED}[precision]
        work = torch.empty(nat.load().dpnegf_rgf_transmission_workspace_elems(
            len(systems.sizes), systems._h_bsz, P), dtype=C128, device=dev)
        nat.call("dpnegf_rgf_transmission", nat.stream_ptr(dev), prec, len(systems.sizes),
                 systems._h_bsz, nat.ptr(systems.bsz), nat.ptr(systems.arena), systems.stride,
                 systems.nsys, nat.ptr(sys_idx), P, nat.ptr(z), nat.ptr(sigL), sigL.shape[-1],
                 nat.ptr(sigR), sigR.shape[-1], int(systems.ortho), nat.ptr(T), nat.ptr(work),
                 nat.ptr(status))
        return T
    # larger blocks: block-operation sweep, one call per system and chunk of points
    if precision != "fp64":
        raise NotImplementedError("mixed precision needs block sizes <= 64")
    K = len(systems.sizes)
    idx = (sys_idx.cpu() if sys_idx is not None else torch.zeros(P, dtype=torch.int32))
    if chunk is None:
        chunk = (autotune.blocked_chunk(systems, P) if autotune.enabled()
                 else autotune.default_chunk(systems))
    if autotune.enabled():
        autotune.tune_blocked(systems, min(chunk, P))
    for s in range(systems.nsys):
        pts = (idx == s).nonzero().reshape(-1)
        if pts.numel() == 0:
            continue
        pts = pts.to(dev)
        arena = systems.arena[s * systems.stride:(s + 1) * systems.stride]
        for c0 in range(0, pts.numel(), chunk):
            sel = pts[c0:c0 + chunk]
            B = sel.numel()
            zz = z[sel].contiguous()
            sl_ = sigL[sel].contiguous()
            sr_ = sigR[sel].contiguous()
            tt = torch.empty(B, dtype=torch.float64, device=dev)
            st = torch.zeros(B, dtype=torch.int32, device=dev)
            nwork = nat.load().dpnegf_rgf_transmission_blocked_workspace_elems(K, systems._h_bsz, B)
            work = torch.empty(nwork, dtype=C128, device=dev)
            nat.call("dpnegf_rgf_transmission_blocked", nat.stream_ptr(dev), K, systems._h_bsz,
                     nat.ptr(arena), B, nat.ptr(zz), nat.ptr(sl_), sl_.shape[-1], nat.ptr(sr_),
                     sr_.shape[-1], int(systems.ortho), nat.ptr(tt), nat.ptr(work), nat.ptr(st))
            T[sel] = tt
            if status is not None:
                status[sel] = st
    return T
