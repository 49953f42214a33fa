"""prism3d_amd: PRISM-3D's per-cell chemistry and thermal step on MI355X (gfx950).

Hand-written HIP kernels behind a C ABI (include/prism3d_amd.h, built into
prism3d_amd/libp3d_hip.so); Python mirrors the upstream call surfaces:

  prism3d_amd._prism3d_core   drop-in for the upstream native module
  prism3d_amd._rust_backend   has_rust()/get_rust() loader twin
  prism3d_amd.chemistry       ChemistrySolver.solve_steady_state + batched SoA solve
  prism3d_amd.grid            device-resident RT -> chemistry -> thermal loop
"""

__version__ = "0.1.0"
