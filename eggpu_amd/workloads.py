"""Seeded synthetic workloads for the batched and served forms.

``hetero_specs`` describes the heterogeneous graph set solved in shared
launches: node counts log-uniform in [n_lo, n_hi], average out-degree
uniform in [deg_lo, deg_hi], alternating Erdos-Renyi G(n, M) and R-MAT
(a=0.57, b=c=0.19) draws, unweighted.  ``serving_pool_specs`` describes the
device-resident pool of R-MAT graphs (1e4 to 5e7 draws, mixed weighted and
unweighted) that the serving loop answers requests against.  Every value is
a function of the seed, so the CPU oracle can rebuild any graph of a set.
"""
from dataclasses import dataclass

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)


@dataclass(frozen=True)
class GraphSpec:
    kind: str         # "er" or "rmat"
    n: int
    nedges: int       # draws before self-loop / duplicate removal
    seed: int
    weighted: bool = False
    wseed: int = 0

    def build(self, native, alpha=0.85, keep_csr=False):
        """The graph on the device (eggpu_amd.native.DeviceGraph)."""
        if self.kind == "er":
            return native.DeviceGraph.er(self.n, self.nedges, seed=self.seed, weighted=self.weighted,
                                         wseed=self.wseed, alpha=alpha, keep_csr=keep_csr)
        a, b, c = RMAT_ABC
        return native.DeviceGraph.rmat_n(self.n, self.nedges, a=a, b=b, c=c, seed=self.seed,
                                         weighted=self.weighted, wseed=self.wseed, alpha=alpha, keep_csr=keep_csr)


def hetero_specs(count=512, seed=2024, n_lo=1e3, n_hi=1e6, deg_lo=4.0, deg_hi=32.0):
    rng = np.random.default_rng(seed)
    logn = rng.uniform(np.log(n_lo), np.log(n_hi), size=count)
    deg = rng.uniform(deg_lo, deg_hi, size=count)
    specs = []
    for i in range(count):
        n = max(int(round(float(np.exp(logn[i])))), 2)
        specs.append(GraphSpec("er" if i % 2 == 0 else "rmat", n, int(round(n * float(deg[i]))),
                               seed=seed * 1000003 + i))
    return specs


def serving_pool_specs(count=32, seed=77, m_lo=1e4, m_hi=5e7, edge_factor=16):
    """R-MAT graphs whose draw counts are log-uniform in [m_lo, m_hi]; every
    other graph carries U(0.1, 1.0) FP32 weights."""
    rng = np.random.default_rng(seed)
    logm = rng.uniform(np.log(m_lo), np.log(m_hi), size=count)
    specs = []
    for i in range(count):
        m = int(round(float(np.exp(logm[i]))))
        n = max(m // edge_factor, 2)
        specs.append(GraphSpec("rmat", n, m, seed=seed * 1000003 + i, weighted=(i % 2 == 1),
                               wseed=seed * 7 + i))
    return specs
