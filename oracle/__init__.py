"""CPU oracle for Muscle.jl's contraction hot path — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, as the checker and as the timed CPU baseline; the product
(muscle.jl_amd/) never does.

  backend_base  numpy restatement of BackendBase (permutedims + BLAS gemm),
                the front-end index algebra, the hyperindex extension, the
                unary (OMEinsum) rule and hadamard.
  brute         long-double brute-force einsum in C (oracle/brute.c), the
                independent check that pins backend_base.

The reference (Julia) cannot be built or run here: no Julia toolchain is
installed, so there is no oracle/_ref build.  Every oracle function is pinned
by the reference's own known-answer tests plus brute force; see the header of
backend_base.py for the list.
"""
