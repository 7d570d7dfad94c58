"""trajsimimeasures_amd -- MI355X (gfx950) engine for TrajSimiMeasures'
batched dynamic-programming trajectory-similarity path: DTW, EDR, LCSS, ERP,
discrete Frechet and Hausdorff, behind the reference's ``core`` call
surface (``from trajsimimeasures_amd import core; core.dtw(...)``)."""
from . import core, data  # noqa: F401
from ._lib import DTW, EDR, ERP, FRECHET, HAUSDORFF, LCSS, TsmError, lib  # noqa: F401

__all__ = ["core", "data", "lib", "TsmError", "DTW", "EDR", "LCSS", "ERP", "FRECHET", "HAUSDORFF"]
