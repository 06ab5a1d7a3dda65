"""math_amd: MI355X-native n-space path tracer (the painter path of Math Viewer).

The product is libnsp.so (HIP kernels for gfx950 behind the C ABI of
include/nsp.h) and the C++17 painter adapter in csrc/painter/.  This Python
package only binds the C ABI for tests and benchmarks (native.py) and holds the
workload configurations (configs.py).
"""
