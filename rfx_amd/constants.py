"""Physical constants used by the reference solver.

C0 is rfx/grid.py:8; EPS_0 and MU_0 are rfx/core/yee.py:117-118 (post-2019
SI values; 1/sqrt(MU_0 EPS_0) equals C0 to rounding).
"""
C0 = 299_792_458.0
EPS_0 = 8.8541878128e-12
MU_0 = 1.25663706212e-6
