"""L.A.Cosmic cosmic-ray cleaning on MI355X (gfx950).

A from-scratch HIP implementation of the cusmic cleaning loop behind the cusmic
C ABI (``include/cusmic.h``, ``libcusmic.so``), with this Python surface on
PyTorch-ROCm tensors in place of the reference's CuPy package
(mod/cusmic/__init__.py).
"""

__version__ = "0.1.0"

from .api import remove_cosmics, remove_cosmics_ragged
from .cleaner import Cleaner
from .image import Image

__all__ = ["Image", "Cleaner", "remove_cosmics", "remove_cosmics_ragged"]
