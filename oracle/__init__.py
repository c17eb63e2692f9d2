"""CPU oracle for the fastesp ESP hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``fastesp_amd``) imports, links or runs anything in
this directory.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker / the
CPU baseline, never as the thing measured on the GPU.

Contents (each function cites the reference file:line it restates):

* ``ints``     -- Boys (scipy ``hyp1f1``, as ``fastesp/ints.py:35-40``), the
                  a-table, 1-D/3-D Hermite E coefficients, (d/dR)^t F and the
                  bare primitive integral (``fastesp/ints.py:46-158``), the
                  per-L contraction map (``fastesp/esp.py:54-105``).
* ``engine``   -- the point-independent pair-table build with the NumPy
                  screen, monomial schedule and flattening
                  (``fastesp/esp.py:168-257``, ``fastesp/table.py:176-259``).
* ``evaluate`` -- the per-point evaluator (``fastesp/table.py:73-104``) and
                  its analytic adjoint (field dV/dC, dL/dG, dL/dP), which the
                  reference does not have.
* ``cpu_port.cpp`` -- a C++/OpenMP restatement of the builder
                  (``fastesp/native/build_pairs.cpp:162-425``) and evaluator
                  (``fastesp/native/esp_gpu.cu:36-66,161-229``), used as the
                  timed CPU baseline and the large-N cross-check.
* ``Makefile`` -- builds ``cpu_port`` and, from the reference's own source
                  where it lies, ``_ref/build_pairs`` (the reference C++ table
                  builder) used to pin ``engine``.

Parity status of every function is recorded in the module docstrings: each is
pinned by a closed form, a quadrature, the reference's own known-answer tests
(``tests/test_integrals.py``, ``tests/test_pipeline.py``) or the compiled
reference builder.
"""
