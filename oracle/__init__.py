"""CPU oracle for the block-sparse DMRG hot path (test infrastructure only).

This package restates, in numpy, the algorithms of QuantiT v0.2.1 that the
MI355X library replaces: conserved-charge bookkeeping, ``btensor::tensordot``,
``reshape``/``permute``, the sector-wise block SVD/eigh, ``truncate_impl`` and
the two-site DMRG update.  Every function cites the reference file:line it
follows.

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / the CPU baseline.  The
product (``quantit_amd``) never imports it and has no CPU fallback.
"""
