"""Recursive Green's function, complex128 torch restatement (oracle only).

Follows ``dpnegf/negf/recursive_green_cal.py`` of the reference:

* ``recursive_gf``      -> wrapper ``:324-477`` (energy shift ``:378``, batch
  broadcast ``:397-409``, ``seP`` ``:411-414``, corner self-energies
  ``:416-428``, scalar squeeze ``:386-395, 474-475``).
* ``_rgf_core``         -> ``recursive_gf_cal`` list path ``:173-296``.  The
  stacked path ``:72-168`` performs the same arithmetic on a ``[K,B,n,n]``
  tensor, so one restatement covers both.
* ``_pack``             -> ``_pack_ans`` ``:299-321``.
* ``_squeeze``          -> ``_squeeze_ans`` ``:480-491``.

Everything is plain differentiable torch, so autograd through it is the
"reference torch autograd gradient" used to check the adjoint kernels.
"""
import torch


def _bcast(t, B):
    """2-D k-block -> [B, ...] view (``recursive_green_cal.py:397-400``)."""
    if t.ndim == 2:
        return t.unsqueeze(0).expand(B, -1, -1)
    return t


def _rgf_core(z, D, L, U, SD, SL, SU, s_in, s_out, need_lesser, need_greater,
              need_gr_lc, keep_gr_left):
    """Forward/backward sweeps of ``recursive_green_cal.py:173-296``.

    ``z`` is ``[B]`` complex (E + E_ref + i*eta).  ``D`` holds the diagonal
    blocks with the self-energies already added.  Returns the 14-tuple.
    """
    K = len(D)
    zz = z.view(-1, 1, 1)
    # M_q = H_q - z S_q, L_q/U_q likewise (``:175-182``)
    Md = [D[q] - zz * SD[q] for q in range(K)]
    Ml = [L[q] - zz * SL[q] for q in range(K - 1)]
    Mu = [U[q] - zz * SU[q] for q in range(K - 1)]
    B = z.shape[0]

    def eye(n):
        return torch.eye(n, dtype=z.dtype).expand(B, n, n)

    # forward: left-connected g_q (``:198-207``)
    gl = [None] * K
    gl[0] = torch.linalg.solve(-Md[0], eye(Md[0].shape[-1]))
    for q in range(K - 1):
        a = -Md[q + 1] - Ml[q] @ gl[q] @ Mu[q]
        gl[q + 1] = torch.linalg.solve(a, eye(a.shape[-1]))

    # backward: full retarded blocks (``:209-238``)
    grd = [None] * K
    grl = [None] * (K - 1)
    gru = [None] * (K - 1)
    grd[K - 1] = gl[K - 1].clone()
    col = gl[K - 1].clone()          # G_{q,K-1}
    lc = [col] if need_gr_lc else None
    for q in range(K - 2, -1, -1):
        gu = gl[q] @ Mu[q]
        grl[q] = grd[q + 1] @ Ml[q] @ gl[q]
        gru[q] = gu @ grd[q + 1]
        grd[q] = gl[q] + gu @ grl[q]
        col = gu @ col
        if need_gr_lc:
            lc.append(col)
    if need_gr_lc:
        lc.reverse()

    gnd = gnl = gnu = gin = None
    if need_lesser:
        # forward ``:243-249``; backward ``:256-264`` (keeps the reference's
        # use of L_q where the textbook has U_q^dagger)
        gin = [None] * K
        gin[0] = gl[0] @ s_in[0] @ gl[0].mH
        for q in range(K - 1):
            src = s_in[q + 1] + Ml[q] @ gin[q] @ Mu[q]
            gin[q + 1] = gl[q + 1] @ src @ gl[q + 1].mH
        gnd = [None] * K
        gnl = [None] * (K - 1)
        gnu = [None] * (K - 1)
        gnd[K - 1] = gin[K - 1].clone()
        for q in range(K - 2, -1, -1):
            lga = Ml[q] @ gl[q].mH
            gnl[q] = grd[q + 1] @ Ml[q] @ gin[q] + gnd[q + 1] @ lga
            gnd[q] = (gin[q] + gl[q] @ Mu[q] @ gnd[q + 1] @ lga
                      + gin[q] @ Mu[q] @ gru[q].mH
                      + gru[q] @ Ml[q] @ gin[q])
            gnu[q] = gnl[q].mH

    gpd = gpl = gpu = gip = None
    if need_greater:
        # ``:269-289`` -- note ``.conj()`` (not ``.mH``) as in the reference
        gip = [None] * K
        gip[0] = gl[0] @ s_out[0] @ gl[0].conj()
        for q in range(K - 1):
            src = s_out[q + 1] + Ml[q] @ gip[q] @ Mu[q].conj()
            gip[q + 1] = gl[q + 1] @ src @ gl[q + 1].conj()
        gpd = [None] * K
        gpl = [None] * (K - 1)
        gpu = [None] * (K - 1)
        gpd[K - 1] = gip[K - 1].clone()
        for q in range(K - 2, -1, -1):
            lcgc = Ml[q].conj() @ gl[q].conj()
            gpl[q] = grd[q + 1] @ Ml[q] @ gip[q] + gpd[q + 1] @ lcgc
            gpd[q] = (gip[q] + gl[q] @ Mu[q] @ gpd[q + 1] @ lcgc
                      + gip[q] @ Mu[q].conj() @ grl[q].conj()
                      + gru[q] @ Ml[q] @ gip[q])
            gpu[q] = gpl[q].mH

    if not keep_gr_left:
        gl = None
    return _pack(col, lc, grd, grl, gru, gl, gnd, gnl, gnu, gin,
                 gpd, gpl, gpu, gip, need_lesser, need_greater)


def _pack(g_trans, gr_lc, grd, grl, gru, gr_left, gnd, gnl, gnu, gin_left,
          gpd, gpl, gpu, gip_left, need_lesser, need_greater):
    """Output contract of ``_pack_ans`` (``recursive_green_cal.py:299-321``)."""
    less = (gnd, gnl, gnu, gin_left) if need_lesser else (None,) * 4
    great = (gpd, gpl, gpu, gip_left) if need_greater else (None,) * 4
    return (g_trans, gr_lc, grd, grl, gru, gr_left) + less + great


def _squeeze(ans):
    """``_squeeze_ans`` (``recursive_green_cal.py:480-491``)."""
    out = []
    for x in ans:
        if x is None:
            out.append(None)
        elif torch.is_tensor(x):
            out.append(x.squeeze(0))
        elif isinstance(x, list):
            out.append([t.squeeze(0) if torch.is_tensor(t) else t for t in x])
        else:
            out.append(x)
    return tuple(out)


def recursive_gf(energy, hl, hd, hu, sd, su, sl, left_se, right_se, seP=None,
                 E_ref=0.0, s_in=0, s_out=0, eta=1e-5, need_lesser=False,
                 need_greater=False, need_gr_lc=False, keep_gr_left=True):
    """Same signature, defaults and 14-tuple as the reference wrapper
    (``recursive_green_cal.py:324-326``)."""
    e = energy + E_ref
    if not torch.is_tensor(e):
        e = torch.as_tensor(e, dtype=torch.complex128)
    probe = left_se if isinstance(left_se, torch.Tensor) else right_se
    squeezed = isinstance(probe, torch.Tensor) and probe.ndim == 2
    if e.ndim == 0:
        e = e.reshape(1)
    B = e.shape[0]
    z = e.to(torch.complex128) + 1j * eta

    D = [_bcast(h, B) * 1.0 for h in hd]
    if seP is not None:
        for i in range(len(D)):
            sp = _bcast(seP[i], B) if torch.is_tensor(seP[i]) else seP[i]
            D[i] = D[i] + sp
    if isinstance(left_se, torch.Tensor):
        se = _bcast(left_se, B)
        r = min(D[0].shape[-2], se.shape[-2])
        c = min(D[0].shape[-1], se.shape[-1])
        D[0] = D[0].clone()
        D[0][:, :r, :c] = D[0][:, :r, :c] + se[:, :r, :c]
    if isinstance(right_se, torch.Tensor):
        se = _bcast(right_se, B)
        r = min(D[-1].shape[-2], se.shape[-2])
        c = min(D[-1].shape[-1], se.shape[-1])
        D[-1] = D[-1].clone()
        D[-1][:, -r:, -c:] = D[-1][:, -r:, -c:] + se[:, -r:, -c:]
    L = [_bcast(t, B) for t in hl]
    U = [_bcast(t, B) for t in hu]
    SD = [_bcast(t, B) for t in sd]
    SL = [_bcast(t, B) for t in sl]
    SU = [_bcast(t, B) for t in su]
    if isinstance(s_in, list):
        s_in = [_bcast(t, B) if torch.is_tensor(t) else t for t in s_in]
    if isinstance(s_out, list):
        s_out = [_bcast(t, B) if torch.is_tensor(t) else t for t in s_out]
    if need_lesser:
        assert isinstance(s_in, list)
    if need_greater:
        assert isinstance(s_out, list)

    ans = _rgf_core(z, D, L, U, SD, SL, SU, s_in, s_out, need_lesser,
                    need_greater, need_gr_lc, keep_gr_left)
    if squeezed:
        ans = _squeeze(ans)
    return ans
