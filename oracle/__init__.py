"""CPU oracle for the sdepy path-integration hot path -- TEST INFRASTRUCTURE.

This package restates, in NumPy, the algorithm of sdepy's Euler path
integration (sdepy/integration.py, sdepy/infrastructure.py), so that the HIP
path of ``sdepy_amd`` can be checked against it on identical inputs:

* ``oracle.philox``   -- Philox4x32-10 counter-based generator, uniforms,
  Box-Muller normals and Poisson inversion, bit-for-bit the mapping used by
  the kernels (host twin of sdepy_amd/csrc/philox.hpp);
* ``oracle.euler``    -- paths_generator / integrator.euler_next / SDE.A /
  SDE.dZ / store / let / exit for the preset processes, fed either recorded
  increments (the reference's own replay hook: any source-protocol callable)
  or Philox draws;
* ``oracle.montecarlo`` -- sdepy's montecarlo accumulator (moments 1-4 and
  histograms over cumulated samples, antithetic folding); the process
  reductions pmean / pvar / psum / pstd are ndarray.mean / var / sum / std
  over the paths axis, which the tests call on the oracle's arrays;
* ``oracle.closed_forms`` -- the textbook closed forms used to pin the above
  (lognormal, discrete-Euler OU moments, Black-Scholes, Merton series, Heston
  log-mean);
* ``oracle.greeks``   -- FP64 pathwise adjoint and common-random-number finite
  differences for the Greeks kernels.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed
CPU baseline: the product path (``sdepy_amd``) never imports it and fails
loudly when its HIP library is missing.

Pinning (which independent check fixes each function) is listed in the
docstring of each module.
"""
