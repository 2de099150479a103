"""MI355X (gfx950) implementation of 2pcf's pair-counting hot path.

Binned scalar-scalar (kappa-kappa) and shear-scalar (kappa . gamma_t/gamma_x)
correlations with per-lens marks and per-patch-pair sums, behind the C-ABI of
``include/tpcf_gpu.h``.  The package is importable as ``2pcf_amd`` through
``importlib.import_module`` (its name starts with a digit).
"""
from . import _native  # noqa: F401
from ._native import TpcfError  # noqa: F401
from .tpcf import (IHS, KAPPA, LIN, LOG, SHEAR, SQRT, Batch, Binning, Context, KappaMap, LensCatalog,  # noqa: F401
                   LensingSurveyField, Shear2pcfResults, ShapeCatalog, Tape, adjoint, adjoint_record, adjoint_replay,
                   block_sums, bootstrap, correlate, create_bins, default_context, find_bin_index, kappa_2pcf, kmeans,
                   rect_blocks, shear_2pcf, stack_fields)

__all__ = [
    "TpcfError", "Batch", "Binning", "Context", "KappaMap", "LensCatalog", "LensingSurveyField", "Shear2pcfResults",
    "ShapeCatalog", "Tape", "adjoint", "adjoint_record", "adjoint_replay", "block_sums", "bootstrap", "correlate", "create_bins", "default_context",
    "find_bin_index", "kappa_2pcf", "kmeans", "rect_blocks", "shear_2pcf", "stack_fields", "LIN", "LOG", "IHS", "SQRT", "SHEAR", "KAPPA",
]
