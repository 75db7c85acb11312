"""floodca_amd -- MI355X (gfx950) backend for the FloodCA weighted-cellular-automata
step: hand-written HIP kernels behind a C-ABI (include/floodca_hip.h), with this
package as the ctypes binding for the Python prep and analysis side."""
