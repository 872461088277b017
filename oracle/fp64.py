"""Float64 closed form of the node-level force of infection.  TEST INFRASTRUCTURE ONLY.

Restates Transmission_ABM.step Part 2b (src/laser_polio/model.py:1327-1363) in
float64 throughout:
    beta'   = beta + N^T beta - beta * rowsum(N)        (network mix, model.py:1334-1338)
    beta''  = beta' * seasonality * r0_scalar           (model.py:1341)
    rate    = beta'' / max(pop_t, 1)                    (model.py:1344-1347)
    p       = 1 - exp(-rate)  (evaluated as -expm1(-rate))   (model.py:1348-1351)
    expected= sum(risk of S) * sum_s p                  (model.py:1362-1363)
with N[i, j] the migration fraction from node i to node j.
"""

import numpy as np

Q24 = 2.0**-24


def foi_fp64(IB, SW, network, r0s, season, pop_t):
    """IB: int64 [N, K] fixed-point infectious weight; SW: int64 [N] fixed-point S risk."""
    beta = np.asarray(IB, dtype=np.float64) * Q24
    net = np.asarray(network, dtype=np.float64)
    mixed = beta + net.T @ beta - beta * net.sum(axis=1)[:, None]
    b2 = mixed * float(season) * np.asarray(r0s, dtype=np.float64)[:, None]
    rate = b2 / np.maximum(np.asarray(pop_t, dtype=np.float64), 1.0)[:, None]
    p = np.maximum(-np.expm1(-rate), 0.0)
    expected = np.asarray(SW, dtype=np.float64) * Q24 * p.sum(axis=1)
    return p, expected, mixed
