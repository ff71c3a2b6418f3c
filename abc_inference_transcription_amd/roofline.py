"""Algorithmic work units used for the roofline of each kernel.

ssa_kernel is bound by VALU issue (integer Philox + scalar FP64 per lane; it
reads and writes almost nothing).  Its algorithmic unit is one SSA step (one
Philox4x32-10 block spent on an event proposal, a boundary crossing or 128
division coin flips).  Lane-operations per event step, counted from the
source of the step (each IEEE op or fused multiply-add, integer op, 32x32->64
product, three-input xor, conversion, compare or table load = 1; gfx950 does a
three-input xor in one v_bitop3_b32):

    Philox4x32-10: 10 rounds x (2 products + 2 three-input xor + 2 key adds) 60
    counter update and packing                                                4
    uniform for the waiting time (cvt, add, mul)                              3
    log (index/exponent/reduction bits 8, table load 1, r and r^2 2,
         log1p polynomial 7, reconstruction 7)                               25
    waiting time (neg, div, t + tau), two boundary compares                   5
    propensities (select, 4 int adds, 2 cvt, 2 mul, 2 add, test)             12
    channel uniform and selection compares                                    6
    state update                                                              4
                                                                      ------
                                                                        119

The peak is one VALU lane-op per lane per clock: 256 CUs x 64 lanes x
2.4 GHz = 39.3 Top/s (half the published 78.6 TFLOP/s FP64 vector figure,
which counts an FMA as two flops).
"""
SSA_OPS_PER_STEP = 119
VALU_PEAK_TOPS = 39.3

# Moment model (moment_transient_kernel, moment_traj_kernel): FP64 operations
# per Dormand-Prince step, counted from the source of the step (the library
# is built without contraction, so every add, multiply and divide is one
# instruction):
#
#     stage combinations  9 x (3 + 5 + 7 + 9 + 11) + yn 9 x 11         414
#     stage times t + c h                                                9
#     size scaling of alpha at 6 stage times (4 each)                   24
#     right-hand side, 6 evaluations x 67                              402
#     error norm 9 x 15 (6 products, 5 sums, h, scale, divide, ...)    135
#     step-size control (own log and exp)                               60
#                                                                   ------
#                                                                     1044
#
# The FP64 peak for such unfused operations is one per lane per clock:
# 39.3 TFLOP/s (the published 78.6 counts a fused multiply-add as two).
MOMENT_FLOPS_PER_STEP = 1044

# Method 1 (the gene block in closed form, Dormand-Prince over the other five
# moments), counted the same way per step:
#
#     stage combinations  5 x (3 + 5 + 7 + 9 + 11) + yn 5 x 11         230
#     stage times t + c h and offsets from the piece start              14
#     size scaling of alpha at 5 stage times (4 each)                   20
#     right-hand side of the five moments, 6 evaluations x 41          246
#     gene block at 5 stage times x 102 (two own exp of 21 and two
#       products, three (E1, E2) pairs of 10 and gamma - K, the
#       source sum J of 13, the four moments of 12)                    510
#     error norm 5 x 15                                                 75
#     step-size control (own log and exp)                               60
#                                                                   ------
#                                                                     1155
MOMENT_FLOPS_PER_STEP_BY_METHOD = {0: MOMENT_FLOPS_PER_STEP, 1: 1155}
FP64_PEAK_TFLOPS = 39.3
