"""CPU oracle for the flux-reconstruction right-hand side.

THIS PACKAGE IS TEST INFRASTRUCTURE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product (``pyfr_amd``) never imports,
links or executes anything under ``oracle/``.

It restates, in plain NumPy FP64, the reference algorithm of PyFR 3.1 for the
hot path:

* ``oracle.basis``    -- orthonormal bases, nodal bases and the ``clean``
  decorator (``pyfr/polys.py``, ``pyfr/nputil.py:23-57``) and the hex/quad/line
  FR operator matrices M0..M6 (``pyfr/shapes.py:79-135``, ``:231-259``).
* ``oracle.geom``     -- metric terms (``pyfr/solvers/base/elements.py:309-440``,
  ``pyfr/solvers/baseadvec/kernels/smats.mako:13-25``).
* ``oracle.mesh``     -- periodic box meshes and a brute-force interface pairing
  (geometric, independent of the product's sort-based pairing).
* ``oracle.physics``  -- Euler/NS fluxes and Riemann solvers
  (``pyfr/solvers/euler/kernels/*.mako``, ``pyfr/solvers/navstokes/kernels/*.mako``).
* ``oracle.rhs``      -- the RHS in the reference's own kernel decomposition
  (``pyfr/solvers/baseadvec/system.py:57-136``,
  ``pyfr/solvers/baseadvecdiff/system.py:53-231``) and the RK4/RK45 steppers
  (``pyfr/integrators/explicit/steppers.py:60-243``).
* ``oracle/frcpu.c``  -- a C/OpenMP restatement of the Euler/NS RHS used as the
  timed CPU baseline (the reference OpenMP backend cannot run here).

Each function cites the reference file:line it follows.  How each part is
pinned (golden fixture, closed forms, invariants, brute force) is recorded in
the docstrings of ``tests/test_oracle_*.py``.
"""
