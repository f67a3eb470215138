"""fvm_amd: MI355X-native algebraic multigrid for finite-volume CRMatrix systems.

Host-side mirror of the c-PRIMED/fvm ``LinearSolver``/``AMG`` interface over the
C ABI of ``libfvm_amd.so`` (hand-written HIP kernels for gfx950).
"""
from .linear_system import LinearSystem, read_matrix_market
from .amg import AMG, BCGStab, CG, FGMRES, BatchAMG
from .learned import LearnedCorrection

__all__ = ["AMG", "BCGStab", "CG", "FGMRES", "LinearSystem", "read_matrix_market"]
