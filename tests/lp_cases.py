"""Hand-built LPs for edge cases of the PDHG path, with known answers.

degenerate_lp: the closed-form LP of the random generator (x >= 0, Ax = b,
optimal pair x*, y*) extended by every structural special case the engine's
set-up and kernels must carry: an empty ranged row, an empty free row, a
free row with entries, an empty column with positive cost, an empty free
column with zero cost and a fixed column whose entries lie only in the free
row.  None of them moves the optimum: the new rows take y* = 0, the empty
column sits at its lower bound 0 (reduced cost 1), the free empty column
has zero reduced cost, and the fixed column adds c_f * value to the
objective.

infeasible_lp: x1 + x2 <= -1 with x >= 0, which PDHG cannot solve; a solve
runs to its iteration limit.
"""
import numpy as np
import scipy.sparse as sp

import pdhg_amd as P

FIXED_COST, FIXED_VALUE = 0.7, 2.0


def degenerate_lp(seed=5, m=200, n=400):
    base = P.gen_random_lp(seed, m, n, 4.0, 0.0)
    A0 = base.scipy_csc().tocoo()
    rng = np.random.default_rng(seed)
    free_cols = rng.choice(n, 30, replace=False)
    rows = np.concatenate([A0.row, np.full(31, m + 2)])
    cols = np.concatenate([A0.col, free_cols, [n + 2]])
    vals = np.concatenate([A0.data, rng.standard_normal(31)])
    A = sp.csc_matrix((vals, (rows, cols)), shape=(m + 3, n + 3))
    A.sort_indices()
    inf = np.inf
    lp = P.LP(m + 3, n + 3, A.indptr, A.indices, A.data,
              c=np.concatenate([base.c, [1.0, 0.0, FIXED_COST]]),
              l=np.concatenate([base.l, [0.0, -inf, FIXED_VALUE]]),
              u=np.concatenate([base.u, [5.0, inf, FIXED_VALUE]]),
              L=np.concatenate([base.L, [-1.0, -inf, -inf]]),
              U=np.concatenate([base.U, [1.0, inf, inf]]),
              offset=base.offset)
    lp.xstar = np.concatenate([base.xstar, [0.0, 0.0, FIXED_VALUE]])
    lp.ystar = np.concatenate([base.ystar, [0.0, 0.0, 0.0]])
    obj = float(base.c @ base.xstar) + FIXED_COST * FIXED_VALUE + base.offset
    return lp, obj


def degenerate_costs(lp, k, seed=3):
    """k cost vectors of degenerate_lp that keep every case bounded: non-negative
    increments on the x >= 0 columns, the free empty column's cost kept at 0."""
    rng = np.random.default_rng(seed)
    costs = np.repeat(lp.c[None, :], k, axis=0)
    costs[:, :lp.n - 3] += rng.uniform(0.0, 0.2, (k, lp.n - 3))
    costs[:, lp.n - 3] += rng.uniform(0.0, 1.0, k)  # empty column, lower bound 0
    costs[:, lp.n - 1] += rng.uniform(-1.0, 1.0, k)  # fixed column: any cost
    return costs


def infeasible_lp():
    inf = np.inf
    return P.LP(1, 2, np.array([0, 1, 2]), np.array([0, 0]), np.array([1.0, 1.0]),
                c=np.array([1.0, 1.0]), l=np.zeros(2), u=np.full(2, inf),
                L=np.array([-inf]), U=np.array([-1.0]))
