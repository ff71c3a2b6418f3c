"""Algorithmic work units used for the roofline of each kernel.

ssa_kernel is bound by VALU issue (integer Philox + scalar FP64 per lane; it
reads and writes almost nothing).  Its algorithmic unit is one SSA step (one
Philox4x32-10 block spent on an event proposal, a boundary crossing or 128
division coin flips).  Lane-operations per event step, counted from the
source of the step (each IEEE op, integer op, conversion or compare = 1):

    Philox4x32-10: 10 rounds x (2 mul_lo + 2 mul_hi + 4 xor + 2 key adds)  100
    counter update and packing                                                4
    uniform for the waiting time (cvt, add, mul)                              3
    log (exponent/mantissa split 12, s=(m-1)/(m+1) 3, z 1, Horner 18,
         reconstruction 10)                                                  44
    waiting time (neg, div), t + tau, boundary compare                        4
    propensities (3 selects, 3 int adds, 2 cvt, 2 mul, 4 add)                14
    channel uniform and selection compares                                    6
    state update                                                              4
                                                                      ------
                                                                        179

The peak is one VALU lane-op per lane per clock: 256 CUs x 64 lanes x
2.4 GHz = 39.3 Top/s (half the published 78.6 TFLOP/s FP64 vector figure,
which counts an FMA as two flops).
"""
SSA_OPS_PER_STEP = 179
VALU_PEAK_TOPS = 39.3
