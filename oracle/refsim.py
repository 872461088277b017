"""A second CPU restatement of the reference's daily step, with the reference's own random
streams and number formats.  TEST INFRASTRUCTURE ONLY: used by tests/test_ensemble.py;
nothing in the product imports it.

The oracle (lp_oracle.cpp) replays the GPU's draws bit for bit, so it shares the GPU's
substitutions: Philox counter streams, fixed-point selection weights, uint16 timers, the
fixed-order float32 node map and no strain multinomial.  This module keeps the reference's
own choices instead, so that an ensemble of runs of the two can show that the
substitutions leave the distribution of outcomes unchanged:

  disease_state_step_kernel  int8 timers tested with <= 0 and decremented (model.py:393-452),
                             numpy uniforms for paralysis
  tx_step_prep               float32 sums of risk and infectious weight per node
                             (model.py:932-1007)
  Transmission_ABM.step      node map in float64 numpy: (beta * network.T).T transfer,
                             seasonality x r0_scalars, 1 - exp(-rate), Poisson or
                             zero-inflated negative binomial counts, multinomial strain
                             split (model.py:1327-1407)
  tx_infect_nb               float32 cumsum of risk x FOI over the node's susceptibles,
                             searchsorted of uniform probes, unique, zeroed weights for
                             further rounds, strain drawn per exposed agent
                             (model.py:1010-1149)
  count_SEIRP                per-node S / E / I / R counts (model.py:869-929)

It covers DiseaseState, Transmission and the log; births, deaths, RI and SIA are not
restated here (the ensembles that use it run without them, VitalDynamics only carrying
the population forward).  The random stream is numpy's default_rng(seed), in sequence.
"""

import numpy as np


class ReferenceSim:
    """One simulation from the product's exported SimTables (its initial agents), stepped
    with the reference's algorithm and random stream."""

    def __init__(self, tables, pop0, seed):
        tb = tables
        live = np.concatenate([np.arange(b, e) for b, e in zip(tb.seg_begin, tb.live_end)])
        c = tb.cols
        self.N, self.K, self.nt = tb.n_nodes, tb.n_strains, tb.nt
        self.state = c["disease_state"][live].astype(np.int8)
        self.strain = c["strain"][live].astype(np.int8)
        self.node = np.repeat(np.arange(self.N), (tb.live_end - tb.seg_begin).astype(np.int64))
        clip8 = lambda x: np.minimum(x.astype(np.int64), 127).astype(np.int8)  # noqa: E731
        self.et, self.it, self.pt = clip8(c["exposure_timer"][live]), clip8(c["infection_timer"][live]), clip8(c["paralysis_timer"][live])
        self.pp, self.par, self.ipv = c["potentially_paralyzed"][live].copy(), c["paralyzed"][live].copy(), c["ipv_protected"][live].copy()
        self.risk = c["acq_risk_multiplier"][live].astype(np.float32)
        self.inf = c["daily_infectivity"][live].astype(np.float32)
        nt_ = np.asarray(tb.network_t, dtype=np.float64).reshape(self.N, self.N)
        self.network = nt_.T.copy()                  # network[i, j]: fraction migrating i -> j
        self.r0s = np.asarray(tb.r0s, dtype=np.float64)
        self.strain_r0 = np.asarray(tb.strain_r0[: self.K], dtype=np.float32)
        self.season = np.asarray(tb.season, dtype=np.float64)
        self.p_par = float(tb.thr_paralysis) / 2.0**32
        self.zi, self.disp_r = float(tb.zi), int(tb.disp_r)
        self.pop = np.asarray(pop0, dtype=np.int64)
        self.rng = np.random.default_rng(seed)
        z = lambda: np.zeros((self.nt, self.N), np.int64)  # noqa: E731
        self.res = {"S": z(), "E": z(), "I": z(), "R": z(), "new_exposed": z()}

    # -- DiseaseState_ABM.step --------------------------------------------------------------
    def _disease(self):
        st = self.state
        E = st == 1
        st[E & (self.et <= 0)] = 2
        self.et[E] -= 1
        I = st == 2
        st[I & (self.it <= 0)] = 3
        self.it[I] -= 1
        s0 = I & (self.strain == 0)
        cand = s0 & (self.pt <= 0) & (self.pp == -1)
        hit = cand & (self.ipv == 0)
        self.pp[hit] = 1
        idx = np.nonzero(hit)[0]
        self.par[idx[self.rng.random(len(idx)) < self.p_par]] = 1
        self.pp[cand & (self.ipv != 0)] = 0
        self.pt[s0] -= 1

    # -- Transmission_ABM.step --------------------------------------------------------------
    def _transmission(self, t):
        N, K, rng = self.N, self.K, self.rng
        st, node = self.state, self.node
        S = st == 0
        exposure = np.bincount(node[S], weights=self.risk[S], minlength=N).astype(np.float32)
        sus = np.bincount(node[S], minlength=N)
        I = st == 2
        beta = np.zeros((N, K), np.float32)
        for k in range(K):
            m = I & (self.strain == k)
            beta[:, k] = np.bincount(node[m], weights=self.inf[m] * self.strain_r0[k], minlength=N).astype(np.float32)
        pre = beta.copy()
        beta = beta.astype(np.float64)
        for k in range(K):
            transfer = (beta[:, k] * self.network.T).T
            beta[:, k] += transfer.sum(axis=0) - transfer.sum(axis=1)
        beta = beta * self.season[t] * self.r0s[:, None]
        rate = beta / np.maximum(self.pop[:, None], 1)
        prob = np.maximum(1 - np.exp(-rate), 0)
        tot_prob = prob.sum(axis=1)
        expected = exposure * tot_prob
        counts = np.zeros((N, K), np.int64)
        for n in range(N):
            x = expected[n]
            if x <= 0:
                continue
            if pre[n].sum() == 0:
                if self.zi >= 1.0:
                    total = 0
                else:
                    mean = x / (1 - self.zi)
                    r = self.disp_r
                    total = 0 if rng.random() < self.zi else int(rng.negative_binomial(r, r / (r + mean)))
            else:
                total = int(rng.poisson(x))
            if total > 0:
                sp = prob[n] / tot_prob[n] if tot_prob[n] > 0 else np.zeros(K)
                counts[n] = rng.multinomial(total, sp / sp.sum())
        new = np.zeros((N, K), np.int64)
        for n in range(N):
            need = int(counts[n].sum())
            if need <= 0 or sus[n] == 0 or tot_prob[n] <= 0:
                continue
            idx = np.nonzero(S & (node == n))[0]
            p = self.risk[idx] * np.float32(tot_prob[n])
            size = min(need, len(idx))
            chosen = []
            while len(chosen) < size:
                cdf = np.cumsum(p, dtype=np.float32)
                if cdf[-1] <= 0:
                    break
                x = rng.random(size - len(chosen))
                pick = np.unique(np.searchsorted(cdf, x * cdf[-1], side="right"))
                pick = pick[pick < len(idx)]
                take = pick[: size - len(chosen)]
                chosen.extend(take.tolist())
                p[pick] = 0.0
            sp = prob[n] / tot_prob[n]
            csp = np.cumsum(sp)
            for j in chosen:
                a = idx[j]
                if st[a] != 0:
                    continue
                k = int(np.searchsorted(csp, rng.random(), side="right"))
                k = min(k, K - 1)
                st[a] = 1
                self.strain[a] = k
                new[n, k] += 1
        self.res["new_exposed"][t] = new.sum(axis=1)

    def _log(self, t):
        for v, name in ((0, "S"), (1, "E"), (2, "I"), (3, "R")):
            self.res[name][t] = np.bincount(self.node[self.state == v], minlength=self.N)

    def run(self):
        self._log(0)
        for t in range(1, self.nt):
            self._disease()
            self._transmission(t)
            self._log(t)
        return self
