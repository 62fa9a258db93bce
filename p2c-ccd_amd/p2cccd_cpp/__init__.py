"""p2cccd_cpp -- the reference's native module name (src/cpp/bindings/
py_module.cpp:1803-2015) over the gfx950 library.

The compiled bindings live in _p2cccd_cpp and link libp2c_hip.so, which
needs the HIP runtime.  A process must hold one HIP runtime only: PyTorch
ships its own, and a second one loaded beside it loses the device.  So torch,
when it is installed, is imported before the bindings, and the library binds
to the runtime torch brought in.
"""
try:
    import torch  # noqa: F401  (selects the process's HIP runtime)
except ImportError:  # the bindings then load the system runtime
    pass

from ._p2cccd_cpp import *  # noqa: E402,F401,F403
from ._p2cccd_cpp import __doc__  # noqa: E402,F401
from ._p2cccd_cpp import _dense_group_columns, _schedule_runtime_columns  # noqa: E402,F401  (p2cccd.policy)
