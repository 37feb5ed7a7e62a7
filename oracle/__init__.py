"""Parity oracle for the MGPCG Poisson solver -- TEST INFRASTRUCTURE ONLY.

This package is the checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (the HIP library behind ``mgpcg.MGPCGPoissonSolver``)
never calls into it and fails loudly when its own extension is missing.

Contents
--------
core        ctypes bindings of the C restatement of mgpcg.py
            (oracle/mgpcg_oracle.c): coarsen, smooth, restrict, prolongate,
            compute_Ap, reduce, v_cycle, solve in PLAIN / MIXED / F64 modes.
strategies  numpy restatements of the operator-assembly strategies
            (pressure_project.py, volume_control.py, magnetic_force.py,
            potential_function.py, surface_tension.py) and of utils.sample,
            LevelSet.theta/delta.
fp64        scipy.sparse assembly of the level-0 operator and an FP64 CG
            reference solve.

Every function cites the reference file:line it restates.  How each part is
pinned (closed forms, brute force, invariants) is listed in the test files
``tests/test_oracle_*.py``; the reference ships no golden vectors and Taichi
is not installable here, so the only anchors are those checks.
"""
