"""The seeded synthetic graphs of the measured workloads, defined once.

Every builder returns the host CSR (`CSRGraph`, OpenMP generators) or, with
``gpu=True``, the same graph built in HBM (`DeviceGraph`, kz_gpu_gen_*): the
two are bit for bit the same CSR (same digest).  Clean-up follows
datasets/live_journal/standardize.cc:66-114 (symmetrise, dedupe, drop
self-loops and isolated vertices, relabel ascending) and orientation
toolkits/to_oriented_dataset_converter.cc:27-124 (degree order).

    er_2p16_2p20      Erdos-Renyi G(n = 2^16, m = 2^20 draws), seed 1, symmetric
    rmat(scale, ef)   RMAT a/b/c/d = .57/.19/.19/.05, seed 1, degree-oriented
    chunglu_census    Chung-Lu n = 4.85 M, 43 M draws, exponent 2.3, weight cap
                      20,000, seed 1, symmetric
    pool_graph(i)     graph i of the 4096-graph pool (seed 1000 + i): even i
                      Erdos-Renyi, odd i RMAT; n ~ U[2^10, 2^18], average
                      degree ~ U[4, 64]
"""

from __future__ import annotations

import math

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)
CHUNGLU_CENSUS = (4850000, 43000000, 2.3, 20000, 1)
POOL_SIZE = 4096


def er_2p16_2p20(kz, gpu: bool = False, orient: bool = False):
    if gpu:
        return kz.gpu_gen_erdos_renyi(1 << 16, 1 << 20, 1, orient)
    g = kz.gen_erdos_renyi(1 << 16, 1 << 20, 1)
    return g.oriented() if orient else g


def rmat(kz, scale: int, edge_factor: int = 16, seed: int = 1, gpu: bool = False, orient: bool = True):
    if gpu:
        return kz.gpu_gen_rmat(scale, edge_factor, *RMAT_ABC, seed, orient)
    g = kz.gen_rmat(scale, edge_factor, *RMAT_ABC, seed)
    if not orient:
        return g
    o = g.oriented()
    del g
    return o


def chunglu_census(kz, gpu: bool = False):
    if gpu:
        return kz.gpu_gen_chung_lu(*CHUNGLU_CENSUS)
    return kz.gen_chung_lu(*CHUNGLU_CENSUS)


def pool_params(i: int) -> tuple:
    """("er", n, m, seed) or ("rmat", scale, edge_factor, seed) of pool graph i."""
    rng = np.random.default_rng(1000 + i)
    n = int(rng.integers(2**10, 2**18 + 1))
    deg = float(rng.uniform(4, 64))
    if i % 2 == 0:
        return ("er", n, int(n * deg / 2), 1000 + i)
    return ("rmat", max(10, round(math.log2(n))), max(2, round(deg / 2)), 1000 + i)


def pool_graph(kz, i: int, gpu: bool = False, orient: bool = False):
    kind, a, b, seed = pool_params(i)
    if gpu:
        if kind == "er":
            return kz.gpu_gen_erdos_renyi(a, b, seed, orient)
        return kz.gpu_gen_rmat(a, b, *RMAT_ABC, seed, orient)
    g = kz.gen_erdos_renyi(a, b, seed) if kind == "er" else kz.gen_rmat(a, b, *RMAT_ABC, seed)
    return g.oriented() if orient else g
