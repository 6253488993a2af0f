"""Algorithmic work of one energy + force evaluation, for roofline figures.

Operation counts are read off the reference's loops and multiplied by the
exact per-structure counts that hdnnp_work_counts returns.  Convention:
FP64 add/sub/mul/div/sqrt = 1 flop, FMA = 2, transcendentals (exp, cos,
pow) are counted separately and not as flops.

Forward part (G values; SymGrpExpAngn.cpp:248-297, SymGrpExpAngw.cpp
:221-267, SymGrpExpRad.cpp:190-200):
  narrow pair test  8   dr_jk (3 sub) + |dr_jk|^2 (3 mul, 2 add)
  wide pair         3   dr_jk (3 sub)
  narrow triplet   22   sqrt 1, f_c(r_jk) 5, 1/r_ij/r_ik 2, cos 6, f_c
                        product 2, r^2 sum 5, exp argument 1  (+cos, exp)
  wide triplet     13   1/r_ij/r_ik 2, cos 6, f_c product 1, r^2 sum 3,
                        exp argument 1                        (+exp)
  angular member    6   1 + lambda cos 2, fg 1, fg*plambda*fc 2, sum 1,
                        plus the pow_int multiplies (counted exactly)
  radial member     6   (r - rs)^2 eta 4, pexp*fc + sum 2     (+exp)
Derivative part (SymGrpExpAngn.cpp:298-385, SymGrpExpAngw.cpp:268-358,
SymGrpExpRad.cpp:201-230, Mode.cpp:1849-1913, Atom.cpp:371-401):
  narrow triplet   15   pr1, pr2, pr3 9; f_c'(r_jk) 6
  wide triplet      6   pr1, pr2
  narrow member    49   fg*norm, pfczl, p2etapl 3; p1 7; p2 7; p3 5;
                        p*dr 9; three dG/dr vector updates 18
  wide member      46   as narrow with p3 = fg pfczl / r_ij r_ik (2)
  radial member    17   p1 8; p1*dr 3; two dG/dr updates 6
  contraction       6   per dE/dG_s * dG_s/dr term (3 multiply-adds)
Networks (NeuralNetwork.cpp:761-896 and :396-442): 2 n_in n_out per layer
forward, the same again for dE/dG.
"""

FWD = {"pair_n": 8, "pair_w": 3, "trip_n": 22, "trip_w": 13, "mem": 6,
       "rad": 6}
BWD = {"trip_n": 15, "trip_w": 6, "mem_n": 49, "mem_w": 46, "rad": 17,
       "contract": 6}


def forward_flops(c):
    return (FWD["pair_n"] * c[2] + FWD["pair_w"] * c[3]
            + FWD["trip_n"] * c[4] + FWD["trip_w"] * c[5]
            + FWD["mem"] * (c[6] + c[7]) + c[8] + FWD["rad"] * c[1])


def backward_flops(c):
    return (BWD["trip_n"] * c[4] + BWD["trip_w"] * c[5]
            + BWD["mem_n"] * c[6] + BWD["mem_w"] * c[7]
            + BWD["rad"] * c[1] + BWD["contract"] * c[9])


def transcendentals(c):
    return 2 * c[4] + c[5] + c[1]


def network_flops(widths_per_atom):
    """widths_per_atom: list of layer-width lists, one per atom."""
    tot = 0
    for w in widths_per_atom:
        f = sum(2 * a * b for a, b in zip(w[:-1], w[1:]))
        tot += 2 * f
    return tot


def neighbor_bytes(n_atoms, n_entries):
    """Algorithmic HBM bytes of the neighbour build: positions read (24 B
    per atom) and per entry written: index 4, image key 8, dr and d 32,
    partner slot 4."""
    return 24 * n_atoms + 48 * n_entries
