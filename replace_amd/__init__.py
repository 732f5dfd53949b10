"""MI355X-native (gfx950) RePlAce Nesterov global placement.

The compute path is the HIP library ``libreplace_amd.so`` behind the C ABI of
include/replace/rp_capi.h; this package holds its ctypes binding
(``capi``), the host SoA netlist (``netlist``), the LEF/DEF reader
(``lefdef``) and the synthetic netlist generator (``netgen``).
"""
from .netlist import Netlist, params_default  # noqa: F401
