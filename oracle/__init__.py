"""CPU oracle for the FlowPM particle-mesh path.

TEST INFRASTRUCTURE ONLY: a restatement of the reference's algorithm used to
check the HIP path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the ``flowpm_amd``
package never does, and never falls back to it.

* ``oracle.background`` -- growth ODE + interpolators (flowpm/tfbackground.py)
* ``oracle.power``      -- Eisenstein-Hu P(k), sigma8 norm, interp_tf
                           (flowpm/tfpower.py, flowpm/scipy/*)
* ``oracle.pm``         -- CIC, FFT, kernels, LPT, force / kick / drift / nbody
                           (flowpm/utils.py, flowpm/kernels.py, flowpm/tfpm.py)

Every function is pinned by a check that does not restate its formula:
closed forms, library routines, brute force on tiny inputs, invariants and the
reference's stored P(k) tables.  Which test pins which function is listed in
each module header.
"""
