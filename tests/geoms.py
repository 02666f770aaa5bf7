"""Small geometries for parity tests (grid, planes, rings, size bins)."""
from barbie_amd import Geometry, RINGS_REFERENCE, RINGS_UNIFORM

ORACLE_KEYS = ("NN", "Nmax", "Nr", "n_mtbin", "UL_k1", "LL_k1", "BW_k1", "LL", "ff", "beta")


def tiny():      # NN=16: brute-force triangle sums are cheap
    return Geometry(NN=16, UL_k1=6, LL_k1=1, BW_k1=1, Nmax=4, Nr=4, ring_rule=RINGS_UNIFORM,
                    ring_width=1.0, last_ring_to_nmax=1)


def small():     # NN=24, radix-3 FFT length
    return Geometry(NN=24, UL_k1=8, LL_k1=1, BW_k1=1, Nmax=6, Nr=6, ring_rule=RINGS_UNIFORM,
                    ring_width=1.0, last_ring_to_nmax=0)


def medium():    # NN=48, reference ring rule truncated at Nr=12
    return Geometry(NN=48, UL_k1=14, LL_k1=1, BW_k1=1, Nmax=12, Nr=12,
                    ring_rule=RINGS_REFERENCE)


def radix5():    # NN=40: radix-5 FFT length
    return Geometry(NN=40, UL_k1=11, LL_k1=1, BW_k1=2, Nmax=10, Nr=8, ring_rule=RINGS_UNIFORM,
                    ring_width=1.2, last_ring_to_nmax=1)


def config1():   # NN=128, Nmax=32, Nr=12 rings of width 1 (strict), UL=21
    return Geometry(NN=128, UL_k1=21, LL_k1=1, BW_k1=1, Nmax=32, Nr=12,
                    ring_rule=RINGS_UNIFORM, ring_width=1.0, last_ring_to_nmax=0)


def oracle_params(g):
    return {k: getattr(g, k) for k in ORACLE_KEYS}
