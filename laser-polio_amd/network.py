"""Migration networks between nodes (init time, host, float64).

``gravity`` and ``row_normalizer`` restate laser_core.migration as the
reference calls them (model.py:1243-1258); their formulas are written out in
the reference's scripts/sandbox/debug_negative_network.py:46-61 (k * p_i^a *
p_j^b / d_ij^c, zero diagonal) and scripts/sandbox/debug_row_normalizer.py:28-33
(rows whose sum exceeds the cap are scaled down to the cap).  ``radiation``
restates the radiation model of Simini et al. (Nature 2012) with the origin
flux k * p_i; ``distance`` is the haversine great-circle distance in km.
laser_core is not vendored by the reference, so radiation/distance parity with
it is unpinned.
"""

import numpy as np

EARTH_RADIUS_KM = 6371.0


def gravity(pops, distances, k, a, b, c):
    p = np.asarray(pops, dtype=np.float64)
    d = np.asarray(distances, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        net = k * np.power(p[:, None], a) * np.power(p[None, :], b) / np.power(d, c)
    np.fill_diagonal(net, 0.0)
    net[~np.isfinite(net)] = 0.0
    return net


def row_normalizer(network, max_rowsum):
    net = np.array(network, dtype=np.float64, copy=True)
    rs = net.sum(axis=1)
    over = rs > max_rowsum
    net[over] = net[over] * (max_rowsum / rs[over])[:, None]
    return net


def radiation(pops, distances, k, include_home=False):
    p = np.asarray(pops, dtype=np.float64)
    d = np.asarray(distances, dtype=np.float64)
    n = len(p)
    net = np.zeros((n, n))
    for i in range(n):
        order = np.argsort(d[i], kind="stable")
        s = np.zeros(n)
        cum = 0.0
        for j in order:
            if j == i:
                continue
            s[j] = cum  # population strictly closer than j (excluding origin and j)
            cum += p[j]
        for j in range(n):
            if j == i:
                continue
            mi, nj = p[i], p[j]
            denom = (mi + s[j]) * (mi + nj + s[j])
            net[i, j] = k * mi * (mi * nj) / denom if denom > 0 else 0.0
        if include_home:
            net[i, i] = k * p[i]
    return net


def distance(lat1, lon1, lat2, lon2):
    lat1, lon1, lat2, lon2 = map(np.radians, (lat1, lon1, lat2, lon2))
    h = np.sin((lat2 - lat1) / 2) ** 2 + np.cos(lat1) * np.cos(lat2) * np.sin((lon2 - lon1) / 2) ** 2
    return 2 * EARTH_RADIUS_KM * np.arcsin(np.sqrt(h))
