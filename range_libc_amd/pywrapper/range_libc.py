"""The reference's Python module `range_libc` (RangeLibc.pyx), compiled from
RangeLibc.pyx in this directory as `_range_libc` over the C ABI.

PyTorch ships its own HIP runtime and links it by the same unversioned names
as the system one.  Importing torch first (when it is installed) makes the
library bind to that runtime; otherwise a later `import torch` in the same
process would load a second runtime, which then finds no device.
"""
try:
    import torch  # noqa: F401
except ImportError:
    pass

from _range_libc import *  # noqa: E402,F401,F403
from _range_libc import __doc__  # noqa: E402,F401
