"""CPU complex128 oracle for the dynamiqs ``mesolve`` hot path.

TEST INFRASTRUCTURE ONLY. This package is the correctness checker for the HIP
path in ``dynamiqs_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check or to time
a CPU restatement beside the GPU. The product path never imports it and never
falls back to it.

It restates, in NumPy/SciPy complex128, the algorithm of the reference's
``dq.mesolve`` for ``method=Tsit5()`` and ``method=Rouchon1()``:

* ``operators``  -- Fock-space operators and states built as the reference builds
  them (``dynamiqs/utils/operators.py``, ``dynamiqs/utils/states.py``).
* ``dia``        -- the sparse-DIA storage convention and its products
  (``dynamiqs/qarrays/sparsedia_primitives.py``).
* ``lindblad``   -- the three Lindblad vector-field forms
  (``dynamiqs/integrators/core/diffrax_integrator.py:312-327``).
* ``tsit5``      -- Tsit5 tableau, dense output, PID step controller and initial
  step (third-party ``diffrax>=0.6.1``, not vendored in the reference; restated
  from Tsitouras 2011 and Hairer-Norsett-Wanner II.4).
* ``rouchon``    -- Rouchon1 Kraus step with Cholesky normalisation
  (``dynamiqs/integrators/core/rouchon_integrator.py:226-269, 552-582``).
* ``mesolve``    -- the solver driver with PWC and modulated time dependence,
  saving and expectation values (``dynamiqs/integrators/apis/mesolve.py``,
  ``dynamiqs/time_qarray.py``, ``dynamiqs/integrators/core/save_mixin.py``), also
  for Dopri5 (``dopri5``), Rouchon2 / Rouchon3 and Euler.
* ``sesolve``    -- the Schrodinger-equation driver
  (``dynamiqs/integrators/apis/sesolve.py``).
* ``propagators`` -- ``sepropagator`` / ``mepropagator`` column by column.

How each function is pinned (closed forms, library routines, reference KATs,
order conditions) is stated in its module header and tested in
``tests/test_oracle.py``.
"""
