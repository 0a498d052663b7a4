"""Python harness for the MI355X sliding-window VI Schur backend.

`abi` binds the C-ABI of `libvib.so` (include/vib.h) through ctypes; `synth` generates seeded
synthetic windows. The compute path is the native library; nothing here computes residuals,
normal equations or solves.
"""
