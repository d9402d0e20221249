"""tisphi_amd: MI355X-native engine for the tiSPHi particle hot path.

Host code (this package) mirrors the reference's ``eng/`` interface and calls
hand-written HIP kernels for gfx950 through the C ABI in include/tisphi.h
(libtisphi_hip.so, loaded by ``tisphi_amd._lib``).
"""
__version__ = "0.1.0"
