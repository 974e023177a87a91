"""CPU oracle for the NUTS/HMC hot path -- TEST INFRASTRUCTURE ONLY.

This package is a float64 numpy restatement of the reference's
``potential_fn``-driven HMC/NUTS path (numpyro 0.21.0):

* ``jax_random``   -- threefry2x32-20 and the jax.random key/bit/float
                      derivations used by the sampler,
* ``distributions``-- the log-densities and bijections of the five workloads,
* ``models``       -- potentials U(z) and their analytic gradients,
* ``hmc_util``     -- velocity Verlet, iterative NUTS tree, dual averaging,
                      Welford, window schedule, warmup adapter,
* ``hmc``          -- the functional ``hmc()`` init/sample kernels,
* ``diagnostics``  -- ESS / split R-hat,
* ``cpu_nuts.c``   -- a C/OpenMP restatement of the logistic NUTS path used
                      as the timed CPU baseline.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import or execute anything here, and only as the checker.
The product (``numpyro_amd``) never imports this package and has no CPU
fallback.

Each function cites the reference file:line it restates.  Every function is
pinned by at least one of: a reference known-answer test re-expressed in
``tests/test_oracle_*.py``, a published known-answer vector, a closed form,
scipy, or central finite differences.  Functions without such a pin say
"parity unpinned" in their docstring.

Parity unpinned against jax itself (jax is absent and was never imported):
``jax_random.split`` / ``random_bits`` in the partitionable mode (the
original mode is pinned by the documented ``split(PRNGKey(0))`` value), and
the float32 bits of ``jax_random.normal`` (pinned only to the mathematical
sqrt(2) * erfinv(u), which XLA's float32 erf_inv approximates).
"""
