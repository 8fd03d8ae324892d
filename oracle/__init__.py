"""CPU oracle for the FDTD time-stepping path — TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the algorithm of rfx's uniform and
graded-mesh Yee time loop (rfx/simulation.py, rfx/nonuniform.py,
rfx/core/yee.py, rfx/boundaries/cpml.py, rfx/boundaries/pec.py) so the HIP
path can be checked against it.  It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``;
the product package ``rfx_amd`` never imports, links or executes anything
here.

Modules
-------
fdtd_np   numpy restatement (float64 or float32 arithmetic): grid, materials,
          CPML profiles, sources, the step, probes, DFT planes, lumped-port
          V/I DFTs and S11, the graded-mesh variant.
cstep     ctypes binding of ``cstep.c``: the same step written as plain C
          loops (double and float), used for larger cases and as the timed
          CPU baseline.
"""
