"""spmvdare: MI355X-native FP64 SpMV layer with the plugin contract of
cslab-ntua/spmv-dare (Matrix_Format / csr_to_format).

Submodules:
    gen  -- the artificial matrix generator (native, bit-exact) and the
            benchmark's value / x definitions
    hip  -- the gfx950 formats behind the C ABI (include/spmvhip.h)
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PACKAGE_DIR)  # spmv-dare_amd/
