"""fmm-gen for AMD Instinct MI355X (gfx950).

A generator of Strassen-like fast-matrix-multiplication (FMM) kernels with the
surface of flame/fmm-gen: the ``control.py`` CLI, the [U,V,W] coefficient
tables with Kronecker multi-level composition, and the generated
``bl_dgemm_strassen_<pack>`` C-ABI plus the ``test_<job>`` driver.  The
kernels are hand-written HIP for gfx950 (FP64 MFMA), reached from Python via
ctypes on the C-ABI library built from ``csrc/``.

Import name: ``fmm_gen_amd`` (the directory name has a hyphen, so callers load
it through ``fmm_gen_amd_loader`` style importlib helpers; see ``load()``).
"""
from .coeffs import (Algorithm, CoeffError, brent_residual, compose, grid_tables,
                     load_algorithm, parse_coeff, read_coeffs)
from .alg_dic import alg_dic, alg_set, lookup
from .plan import (MR, NR, PACK_TYPES, Partition, build_algorithm, fmm_flops,
                   jobname, partition)

__all__ = [
    "Algorithm", "CoeffError", "brent_residual", "compose", "grid_tables",
    "load_algorithm", "parse_coeff", "read_coeffs", "alg_dic", "alg_set",
    "lookup", "MR", "NR", "PACK_TYPES", "Partition", "build_algorithm",
    "fmm_flops", "jobname", "partition",
]
