"""CPU oracle for the AeroJAX collocated projection step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``aerojax_amd``)
imports this package.  It may be imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``, and
there only as the checker or the timed CPU baseline, never as the thing that
is measured or shipped.

Modules
-------
reference_step  NumPy restatement (FP32 and FP64) of the reference hot path:
                geometry, IC, Smagorinsky LES, ``rk_step_unified``,
                ``divergence_nonperiodic``, ``poisson_multigrid`` (reference
                mode) and the collocated branch of ``step_pure``.
textbook_mg     NumPy restatement of this project's own textbook multigrid
                (red-black Gauss-Seidel, cell-centred transfers, coarse
                direct solve) and the FP64 defect-correction driver, plus the
                consistent outlet-Dirichlet operator used to define the
                ``||div||`` metric.
"""
