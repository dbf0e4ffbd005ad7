"""CPU oracle for the NEGF hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain complex128 torch / numpy / scipy, the
algorithms of the reference implementation (deepmodeling/dpnegf) that the
HIP library in ``dpnegf_amd`` replaces.  It exists to *check* the GPU code
and to be timed as the CPU baseline; nothing in ``dpnegf_amd`` imports it,
and only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it.

Module map (reference file:line each restatement follows):

* ``oracle.rgf``            -- ``dpnegf/negf/recursive_green_cal.py:4-491``
* ``oracle.surface_green``  -- ``dpnegf/negf/surface_green.py:9-91, 169-226, 284-365``
* ``oracle.lead``           -- ``dpnegf/negf/lead_property.py:82-89, 386-460``
* ``oracle.device``         -- ``dpnegf/negf/device_property.py:21-31, 137-296, 374-470``
* ``oracle.ozaki``          -- ``dpnegf/negf/ozaki_res_cal.py:8-37``,
                               ``dpnegf/negf/density.py:143-199``,
                               ``dpnegf/negf/negf_utils.py:136-159``
* ``oracle.dense``          -- dense brute-force inverse used to pin ``oracle.rgf``

Pinning status (what each function is checked against in ``tests/``):

* ``rgf.recursive_gf``, all 14 outputs: the reference's own ``recursive_gf``
  outputs on seeded cases (``tests/golden/reference_rgf.npz``, written by
  ``tests/golden/make_reference_rgf.py``, which loads
  ``dpnegf/negf/recursive_green_cal.py`` by file path) -- uneven blocks with
  S_offdiag != 0 and complex energies, the stacked uniform path, a 3-D seP,
  K = 1 and the scalar squeeze, greater branch (``.conj()`` quirk,
  ``recursive_green_cal.py:145-160``) and lesser branch (``L_q`` quirk,
  ``:134-139``) included; agreement is bit-exact.
* ``rgf.recursive_gf`` retarded outputs: also the dense brute-force inverse,
  the reference's batched==scalar self-consistency, and the reference's 4-site
  chain golden values (``dpnegf/tests/test_negf_device_property.py:103-186``).
* ``rgf.recursive_gf`` lesser / greater outputs: also the dense
  ``G Sigma G^dagger`` where the quirks are exact (Hermitian H, S_offdiag = 0;
  for greater, real symmetric H and complex-symmetric Sigma).
* ``lead.HDL_reduced``: trimming keeps the coupling's self-energy (the trimmed
  one is the corner of the untrimmed one, zero elsewhere), both branches, with
  and without ``subblocks``.
* ``surface_green``: closed-form 1-orbital chain, the Dyson fixed-point
  residual, the GEP route and a long finite chain.
* ``ozaki``: Matsubara limit / Fermi-function identity, brute-force real-axis
  density.
"""
