"""Drop-in for the reference's ``core`` package (core/__init__.py:1-10) for
the six DP / Hausdorff measures and the spatio-temporal STEDR / STLCSS / CDDS /
SAR:
``core.dtw(t1, len1, t2, len2, ...)`` etc., same signatures and float64
[batch] results, computed by libtsm.so."""
from .cdds import *  # noqa: F401,F403
from .discret_frechet import *  # noqa: F401,F403
from .dtw import *  # noqa: F401,F403
from .edr import *  # noqa: F401,F403
from .erp import *  # noqa: F401,F403
from .hausdorff import *  # noqa: F401,F403
from .lcss import *  # noqa: F401,F403
from .sar import *  # noqa: F401,F403
from .stedr import *  # noqa: F401,F403
from .stlcss import *  # noqa: F401,F403
from .utils import pad_lists_to_array  # noqa: F401

__all__ = ['cdds', 'dfrechet', 'dtw', 'edr', 'erp', 'hausdorff', 'lcss', 'sar', 'stedr', 'stlcss',
           'pad_lists_to_array']
