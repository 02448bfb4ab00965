"""Known answers for the reference's FP64 rounding order (hand-derived).

Each value is IEEE-754 round-to-nearest-even applied by hand to the
reference's own expression, with inputs chosen so that the plausible
alternative form rounds differently.  tests/test_oracle.py checks the oracle
against them (and a copy of the oracle with the alternative planted, which
must fail); tests/test_gpu_parity.py checks the engine against them.
"""
import numpy as np

# x: fl(fl(0.1 + 0.2) + 0.4) = 0x1.6666666666667p-1 (0.7000000000000001), and
# fl(1/3) = 0x1.5555555555555p-2; their product rounds to 0x1.ddddddddddddep-3,
# while the quotient 0x1.6666666666667p-1 / 3 rounds to 0x1.ddddddddddddfp-3.
# y: fl(fl(0.3 + 0.3) + 0.3) = 0x1.cccccccccccccp-1 (0.8999999999999999); times fl(1/3)
#    the exact product 0.29999999999999995707... lies below the midpoint
#    0.29999999999999996115... of its two neighbours, so it rounds to
#    0x1.3333333333332p-2 (the quotient by 3 rounds to 0x1.3333333333333p-2).
# z: 1 + 2 + 4 = 7 (exact) * fl(1/3) = 0x1.2aaaaaaaaaaaap+1 (7/3 = ...aaabp+1).
# size: max(|p| - c) over all components = 4 - 0x1.2aaaaaaaaaaaap+1.
MEAN_KAT_POS = np.array([[0.1, 0.3, 1.0], [0.2, 0.3, 2.0], [0.4, 0.3, 4.0]])
MEAN_KAT_C = [float.fromhex("0x1.ddddddddddddep-3"), float.fromhex("0x1.3333333333332p-2"),
              float.fromhex("0x1.2aaaaaaaaaaaap+1")]


# kick: v = 0.782, a = 1.915, dt = 1e-3: fl(a dt) = 0x1.f601797cc3a00p-10, halved
# exactly, plus v rounds to 0x1.90dfce3150daep-1;
# a fused multiply-add of the exact a * dt/2 onto v rounds to ...dafp-1.
# drift: x = -0.002, v = -1.512, dt = 1e-3: x + fl(v dt) = -0x1.cc532a497fa5ep-9;
# fused, -0x1.cc532a497fa5dp-9.
# subnormal: a = 2^-1074, dt = 2.75, v = 0: fl(a dt) = 3 * 2^-1074 (2.75 rounds up),
# * 0.5 = 1.5 * 2^-1074 rounds to even, 2 * 2^-1074; a * fl(dt * 0.5) = 1.375 * 2^-1074
# rounds to 1 * 2^-1074.  Then the drift x = 0 + fl(2 * 2^-1074 * 2.75) = 6 * 2^-1074
# (5.5 rounds to even).
TINY = 2.0 ** -1074


KICK_KAT = dict(v=[[0.782, 0.0, 0.0]], a=[[1.915, 0.0, 0.0]], want=float.fromhex("0x1.90dfce3150daep-1"))
DRIFT_KAT = dict(x=[[-0.002, 0.0, 0.0]], v=[[-1.512, 0.0, 0.0]], want=float.fromhex("-0x1.cc532a497fa5ep-9"))


# theta = 0.6 (fl(0.36) = 0x1.70a3d70a3d70ap-2).  Point at the origin.
# A: baryCentre (2.2, 0.25, 0.2), edge 0x1.557abc473c9c1p+0: fl(s^2 / d^2) is not
#    below fl(theta^2) (the node opens), but s^2 < fl(theta^2 d^2) holds.
# B: baryCentre (0.2, 0.6, 0), edge 0x1.8494a75f057abp-2: fl(s^2 / d^2) < fl(theta^2)
#    (accepted), but s^2 < fl(theta^2 d^2) fails.
MAC_KATS = [((2.2, 0.25, 0.2), "0x1.557abc473c9c1p+0", False),
            ((0.2, 0.6, 0.0), "0x1.8494a75f057abp-2", True)]
