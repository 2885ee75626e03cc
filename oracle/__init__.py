"""CPU oracle for the MI355X Dagger backend -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the timed CPU baseline.  The product
(dagger.jl_amd/) never imports it.

Parity pinning (what makes each oracle function trustworthy):
  dagger_tiled.cholesky   vs scipy dense dpotrf (U and L, ragged tiles), pure
                          Python Cholesky on tiny inputs, closed forms
                          (identity, diagonal, constructed U^T U), LAPACK info
                          on non-PD inputs, committed golden fixtures, and a
                          planted-mistake test that must fail.
  dagger_tiled.trsm/ldiv  vs numpy.linalg.solve on the assembled triangle and
                          the residual of the solved system.
  dagger_tiled.datadeps_edges  vs hand-derived edge sets of the reference's
                          rules (R->R none, R->W, W->R, W->W).
  gen.philox4x32_10       vs the published Random123 known-answer vectors.
  gen.randn               vs its moments and the Box-Muller identity.
"""
