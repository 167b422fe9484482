"""MI355X-native SUCEM:FEM hot path.

Mirrors the reference package's problem API (``sucemfem.ProblemConfigurations``,
``SystemMatrices``, ``Forms``, ``BoundaryConditions``, ``Materials``,
``Sources``, ``Utilities.LinalgSolvers``) on top of libsucem_hip, whose HIP
kernels do the numbering, assembly, CSR operators and Krylov/eigen solves.
"""
__version__ = "1.0.0"
