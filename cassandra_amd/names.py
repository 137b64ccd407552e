"""Pattern masks and pattern sets (Pattern/Names.h:23-211).

Bit b of a mask is the square (x, y) with b = x + 8y; the masks are the values
the reference's "..."_pattern literals evaluate to.
"""

L0 = 0x00000000000000FF
L1 = 0x000000000000FF00
L2 = 0x0000000000FF0000
L3 = 0x00000000FF000000
D4 = 0x0000000001020408
D5 = 0x0000000102040810
D6 = 0x0000010204081020
D7 = 0x0001020408102040
D8 = 0x0102040810204080
L02X = 0x00000000000042FF
B4 = 0x0000000000000F0F
B5 = 0x0000000000001F1F
B6 = 0x0000000000003F3F
Q3 = 0x0000000000070707
C3 = 0x0000000000010307
C3p1 = 0x000000000101030F
C3p2 = 0x000000010101031F
Ep = 0x0000000000003CBD
Comet = 0x8040201008040303

PATTERNS = dict(L0=L0, L1=L1, L2=L2, L3=L3, D4=D4, D5=D5, D6=D6, D7=D7, D8=D8, L02X=L02X, B4=B4, B5=B5, B6=B6,
                Q3=Q3, C3=C3, C3p1=C3p1, C3p2=C3p2, Ep=Ep, Comet=Comet)

# Pattern/Names.h:195-211
logistello = [D4, D5, D6, D7, D8, L02X, L1, L2, L3, Q3, B5]
edax = [D4, D5, D6, D7, D8, L02X, L1, L2, L3, Q3, C3p2, Ep]
cassandra = [D4, D5, D6, D7, Comet, L02X, L1, L2, L3, Q3, C3p2, B5]

# Workload pattern sets: edge+2X, 2x5 corner block, 3x3 corner, main diagonal
# (k = 18, 102,222 columns); plus two edge rows, the second edge variant and
# the 7-square anti-diagonal (k = 34, 139,644 columns); plus the corner
# triangle, the 2x4 block, the third edge row and the 5/6-square diagonals
# (k = 54, 150,282 columns).
FIT4 = [L02X, B5, Q3, D8]
FIT8 = [L02X, B5, Q3, D8, L1, L2, Ep, D7]
FIT12 = [L02X, B5, Q3, D8, C3p2, B4, L1, L2, L3, D5, D6, D7]

# The reference's pattern test masks (PatternTest/Pattern.h:4-82): one per
# symmetry class.
test_a = 0x0000000000000003
test_c = 0x0000000000008040
test_d = 0x0000000000000102
test_h = 0x0000000000000081
test_v = 0x0100000000000001
test_vh = 0x4200000000000042
test_dc = 0x4080000000000102
test_vhdc = 0x8100000000000081
test_mix = [test_a, test_c, test_d, test_h, test_v, test_vh, test_dc, test_vhdc]
