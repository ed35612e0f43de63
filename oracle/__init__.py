"""Test infrastructure: the CPU oracle of the RCWA layer-eigensolve + S-matrix path.

Nothing here is part of the product. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import, call or load anything under ``oracle/``, and only as the checker.

Contents
--------
* ``rcwa1d.py``   -- numpy restatement of the planar TE/TM solve (rcwa1d.hpp).
* ``conical.py``  -- numpy restatement of the conical solve (rcwa1d_conical.hpp).
* ``rcwa2d.py``   -- numpy restatement of the crossed-grating solve (rcwa2d.hpp, rcwa2d_operators.hpp).
* ``ref.py``      -- ctypes loader for ``oracle/_ref/libsrcwa_ref.so``, the reference's own header-only
                     C++ built from its sources by ``oracle/Makefile`` (absent where the reference is absent),
                     including its planar near fields (``planar_fields``: solve_with_fields, field_at, field_map),
                     pinned by flux conservation, tangential continuity and a planted anchor swap in
                     ``tests/test_oracle_fields.py``.
* ``flat.py``     -- decoder of the flat stack records that cross the C-ABI.
"""
