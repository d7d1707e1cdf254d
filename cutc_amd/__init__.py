"""MI355X-native (gfx950) transcorrelated-integral path of CuTC.

The product is the C-ABI library ``cutc_amd/lib/libcutcint.so`` (declared in
``include/cutc_int.h``).  This package holds its sources (``csrc/``), a ctypes
host-side mirror of the reference interface (:mod:`cutc_amd.cutc`) and the
seeded synthetic-input generator (:mod:`cutc_amd.synth`).
"""
