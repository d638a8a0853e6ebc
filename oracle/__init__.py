"""CPU oracle for the chromatix FFT-propagation path -- TEST INFRASTRUCTURE.

NumPy/SciPy restatements of the reference's arithmetic, used only as the
checker: by ``tests/``, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``.  Nothing in ``chromatix_amd`` imports
it; the product path has no CPU fallback.

Two precisions per function:

* ``fp64``  -- the reference's formulas evaluated in float64 with exact
  inputs: the accuracy yardstick.
* ``ref32`` -- the reference's float32 operation order for every quantity it
  computes in float32 (frequency grid, propagator phase), with FFTs in
  complex64.  It stands in for chromatix's own jnp.fft path, which cannot run
  here (no JAX; the reference needs Python >= 3.12).

Parity pins (see tests/test_oracle_*.py): the reference tests' own constants
and closed forms (square-aperture Fresnel diffraction, tests/test_propagate.py:
10-33; thin-sample knowns, tests/test_samples.py; CZT vs DFT / scipy,
tests/test_czt.py), physical plane-wave eigenfunctions, z/-z round trips,
Parseval, and brute-force DFTs on tiny grids.
"""
