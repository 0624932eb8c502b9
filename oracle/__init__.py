"""CPU oracle for the guarded-floor GMG-FGMRES solver path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy/SciPy restatement of the
reference algorithm (``src/gpu_fem`` + ``experiments/phase5`` of
solver-admissibility-and-guarded-floor-selection).  It exists to check the HIP
implementation and to serve as the timed CPU baseline.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker; the product package never imports it.

Modules
-------
fem      hex8 stiffness, element/DOF tables, presets, boundary conditions,
         matrix-free K_ff(E) u, diag(K_ff), node-block Jacobi.
gmg      trilinear transfers, injection coarse free sets, level-1 element-wise
         Galerkin assembly, R(KP) coarse operators, spectral estimates,
         Chebyshev/Jacobi smoothers, V/W cycles, safeguarded coarse
         correction, root-local correction, inner-Krylov wrapper.
krylov   flexible GMRES (MGS + Givens), PCG, fixed-count coarse PCG.
policy   density states, connectivity, probe classification, floor ladder,
         recomputed-residual guard, CSV row schema.

Parity status per function is listed in each module header.
"""
