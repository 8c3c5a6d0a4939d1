"""drops_amd -- MI355X-native trace-FEM assembly and Krylov core (drop-in for the
DROPS P1 trace-FEM path).  The compute path is the HIP library libdhip.so behind
the C ABI declared in include/dhip.h; drops_amd.dhip is its Python binding."""
