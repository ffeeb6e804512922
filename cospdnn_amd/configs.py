"""The benchmark networks (synthetic, deterministic).

Bias by N follows the reference's bias map (src/cospdnn.cpp:97-102).  Every
network is generated from MASTER_SEED + its index: weights of 0.0625 with 32
nonzeros per output column (random distinct rows, or the banded pattern
(pi_l(j) + k*N/32) mod N), and a binary input whose row density is
0.15 + 0.30*U.
"""
from dataclasses import dataclass

MASTER_SEED = 20261203


@dataclass(frozen=True)
class NetSpec:
    name: str
    N: int
    L: int
    S: int
    bias: float
    kind: int  # 0 random rows, 1 banded
    seed: int
    K: int = 32
    wval: float = 0.0625

    @property
    def edges(self):
        return self.S * self.N * self.K * self.L


def _c5():
    shapes = [(1024, 120, 60000, -0.3), (4096, 120, 30000, -0.35), (4096, 480, 10000, -0.35),
              (16384, 120, 45000, -0.4), (16384, 480, 20000, -0.4), (1024, 480, 50000, -0.3)]
    return tuple(NetSpec(f"c5.{i}", N, L, S, b, 0, MASTER_SEED + 4 + 16 * (i + 1)) for i, (N, L, S, b) in
                 enumerate(shapes))


CONFIGS = {
    "c1": (NetSpec("c1", 1024, 120, 60000, -0.3, 0, MASTER_SEED + 0),),
    "c2": (NetSpec("c2", 4096, 120, 60000, -0.35, 0, MASTER_SEED + 1),),
    "c3": (NetSpec("c3", 16384, 120, 60000, -0.4, 1, MASTER_SEED + 2),),
    "c4": (NetSpec("c4", 65536, 480, 60000, -0.45, 0, MASTER_SEED + 3),),
    "c5": _c5(),
}


def config_nets(name):
    return CONFIGS[name]
