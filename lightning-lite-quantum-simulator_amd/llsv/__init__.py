"""llsv -- host-side package of the MI355X complex128 state-vector engine.

  circuits   seeded workload builders (op tuples)
  _native    ctypes binding of libllsv.so (include/llsv.h)

The StateVector surface itself is the pybind11 module quantum_sim_v3 (and its v2
alias) next to this package; the fusion pre-pass is gate_fusion.
"""
__all__ = ["circuits"]
