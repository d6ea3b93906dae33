"""Synthetic operands of the benchmark and autotuner workloads (seeded, host).

Workload names:
  er1024           C = A*A, A = Erdos-Renyi n=1024, 8 columns/row, seed 0
  rmat:S           C = A*A, A = R-MAT scale S, edge factor 16, seed 1
  stage:S:k        SUMMA stage k of R-MAT scale S (seed 2) cut into a 4x4
                   tile grid: the 16 products A_ik * A_kj
  galerkin:n       C = P^T (A P), A = 27-point Laplacian on n^3, P the
                   smoothed-aggregation prolongator (3^3 aggregates)
"""
from __future__ import annotations

from typing import Dict, Tuple

from . import gen


def product_operand(name: str) -> gen.Csr:
    """A of the plain product C = A*A named by `name` (er1024 or rmat:S)."""
    if name == "er1024":
        return gen.erdos_renyi(1024, 8, seed=0)
    if name.startswith("rmat:"):
        return gen.rmat(int(name.split(":")[1]), 16, seed=1)
    raise ValueError(f"not a plain product workload: {name}")


def stage_tiles(scale: int, k: int, g: int = 4, seed: int = 2) -> Dict[Tuple[int, int], gen.Csr]:
    """Tiles (i, k) and (k, j) of R-MAT scale `scale` cut into a g x g grid."""
    A = gen.rmat(scale, 16, seed=seed)
    n = A.nrows // g
    tiles = {}
    for i in range(g):
        tiles[(i, k)] = gen.submatrix(A, i * n, (i + 1) * n, k * n, (k + 1) * n)
        tiles[(k, i)] = gen.submatrix(A, k * n, (k + 1) * n, i * n, (i + 1) * n)
    return tiles


def galerkin_operators(n: int) -> Tuple[gen.Csr, gen.Csr]:
    """(A, P): 27-point Laplacian on n^3 (26 / -1) and its SA prolongator."""
    return gen.laplace27(n), gen.sa_prolongator(n)
