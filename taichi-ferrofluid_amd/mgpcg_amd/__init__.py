"""mgpcg_amd: the MGPCG Poisson solver of taichi-ferrofluid (mgpcg.py) and its
operator-assembly strategies, on one MI355X through libmgpcg.so (HIP, gfx950).

Modules mirror the reference's: ``mgpcg`` (MGPCGPoissonSolver),
``pressure_project``, ``volume_control``, ``magnetic_force``,
``potential_function``, ``surface_tension``, ``level_set``, ``utils``; plus
``batch`` and ``serve`` (batched and continuously batched solves),
``projection`` and ``cells`` (the steps before and after a pressure solve),
``scenes`` and ``systems`` (seeded workloads), ``autotune`` and ``build``.
Put ``taichi-ferrofluid_amd/`` on ``sys.path`` to import it.
"""
