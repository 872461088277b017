#pragma once
// Population initialisation on the device (lp_init_population, include/lp_step.h):
// the per-agent columns that SEIR_ABM and its components draw at construction
// (model.py:561-713 ages, immunity, initial infections; 816-866 heterogeneity;
// 568-587 disease timers; 1548-1669 death dates; 1887-1894 RI timers), for every
// slot of every node segment, from counter-based Philox streams keyed by (init seed,
// sim-local slot).  Draws use float32 samplers of the named distributions; they are
// the same distributions as the host generator's, not the same numbers.
#include "lp_kernels.hip"

namespace lp {

constexpr uint32_t ST_INIT = 9;   // Philox stream of the initial draws (sub-stream << 8)
enum : uint32_t { SI_AGE = 0, SI_AGE2, SI_EXP, SI_INF, SI_PAR, SI_HET, SI_STATE, SI_RI, SI_MISSED };

// Uniform in (0, 1) from the top 24 bits.
__device__ __forceinline__ float u01o(uint32_t r) { return ((float)(r >> 8) + 0.5f) * 5.9604644775390625e-08f; }

// Standard normal by Box-Muller from two 32-bit draws.
__device__ __forceinline__ float normal_bm(uint32_t r1, uint32_t r2) {
  const float u1 = u01o(r1), u2 = u01o(r2);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853071795865f * u2);
}

// Draw j of sub-stream `sub` for agent `slot` (the day word holds -1 - j: no step
// of a run uses negative days).
__device__ __forceinline__ uint4 init_draw(uint64_t seed, uint64_t slot, uint32_t sub, uint32_t j) {
  return philox((uint32_t)slot, (uint32_t)(slot >> 32), (uint32_t)(-1 - (int32_t)j), ST_INIT | (sub << 8), seed);
}

// Gamma(shape k, scale theta) by Marsaglia & Tsang (ACM TOMS 26(3), 2000): for k >= 1
// d = k - 1/3, c = 1/sqrt(9d), x ~ N(0,1), v = (1 + c x)^3, accept d v when
// log u < x^2/2 + d - d v + d log v; for k < 1, Gamma(k + 1) * u^(1/k).
__device__ float gamma_mt(float k, float theta, uint64_t seed, uint64_t slot, uint32_t sub) {
  const bool boost = k < 1.0f;
  const float kk = boost ? k + 1.0f : k;
  const float dd = kk - 1.0f / 3.0f, c = 1.0f / sqrtf(9.0f * dd);
  float g = dd;
  for (uint32_t j = 0; j < 64; ++j) {
    const uint4 o = init_draw(seed, slot, sub, j);
    const float x = normal_bm(o.x, o.y);
    const float t = 1.0f + c * x;
    if (t <= 0.0f) continue;
    const float v = t * t * t;
    const float u = u01o(o.z);
    if (logf(u) < 0.5f * x * x + dd - dd * v + dd * logf(v)) { g = dd * v; break; }
  }
  if (boost) {
    const uint4 o = init_draw(seed, slot, sub, 64);
    g *= powf(u01o(o.x), 1.0f / k);
  }
  return g * theta;
}

// A sample of duration distribution `ds` (lp_dist) for agent `slot`.
__device__ float dist_sample(const lp_dist& ds, uint64_t seed, uint64_t slot, uint32_t sub) {
  switch (ds.kind) {
    case LP_DIST_CONSTANT: return ds.a;
    case LP_DIST_NORMAL: { const uint4 o = init_draw(seed, slot, sub, 0); return ds.a + ds.b * normal_bm(o.x, o.y); }
    case LP_DIST_GAMMA: return gamma_mt(ds.a, ds.b, seed, slot, sub);
    case LP_DIST_LOGNORMAL: {
      if (!(ds.a > 0.0f)) return 0.0f;
      const float m = ds.a, s = ds.b;
      const float mu = logf((m * m) / sqrtf(s * s + m * m)), sg = sqrtf(logf((s * s) / (m * m) + 1.0f));
      const uint4 o = init_draw(seed, slot, sub, 0);
      return expf(mu + sg * normal_bm(o.x, o.y));
    }
    case LP_DIST_POISSON: {                                // inversion (small means: durations)
      const uint4 o = init_draw(seed, slot, sub, 0);
      const float u = u01o(o.x);
      float p = expf(-ds.a), F = p;
      int k = 0;
      while (u > F && k < 1000) { k += 1; p = p * ds.a / (float)k; F += p; }
      return (float)k;
    }
    case LP_DIST_EXPONENTIAL: { const uint4 o = init_draw(seed, slot, sub, 0); return -ds.a * logf(u01o(o.x)); }
    case LP_DIST_UNIFORM_INT: {
      const uint4 o = init_draw(seed, slot, sub, 0);
      return floorf(ds.a + u01o(o.x) * (ds.b - ds.a));
    }
    default: return 0.0f;
  }
}

// Float duration -> timer: truncated toward zero, clipped to [0, 127] (model.py:575-579).
__device__ __forceinline__ uint32_t to_timer(float x) {
  const float tr = truncf(x);
  return tr <= 0.0f ? 0u : tr >= 127.0f ? 127u : (uint32_t)tr;
}

// Age in days of an initial agent (exponential capped, or a pyramid bin then a day in it).
__device__ int32_t init_age(const lp_init_sim& sp, uint64_t seed, uint64_t slot) {
  const uint4 o = init_draw(seed, slot, SI_AGE, 0);
  int32_t age;
  if (sp.age_model == 0) {
    float a = -sp.age_mean_days * logf(u01o(o.x));
    if (a > sp.age_cap_days) a = sp.age_cap_days;
    age = (int32_t)a;
  } else {
    const float u = u01o(o.x);
    int b = 0;
    while (b < sp.n_pyr - 1 && u >= sp.pyr_cdf[b]) ++b;
    const int32_t lo = sp.pyr_lo[b], hi = sp.pyr_hi[b];
    age = lo + (int32_t)floorf(u01o(o.y) * (float)(hi - lo));
  }
  return age == 0 ? 1 : age;
}

// One workgroup per span: every slot's timers and multipliers; the initial agents'
// ages, death dates, immunity, RI timers and node ids; unborn slots' defaults.
__global__ void __launch_bounds__(WG) k_init(Dev d, const lp_init_sim* specs) {
  const int4 span = d.spans[blockIdx.x];
  const int node = span.x, sim = span.w;
  const lp_sim_params& simp = d.sims[sim];
  const lp_init_sim& sp = specs[sim];
  const uint64_t seed = sp.seed;
  const int64_t live_hi = d.live_end[node];
  const int16_t lnode = (int16_t)(node - simp.node_begin);
  const int64_t* km = d.km_cd + (size_t)sim * LP_KM_LEN;
  for (int u = 0; u < span.z; ++u) {
    const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)threadIdx.x * 4;
    uint32_t st = 0, mi = 0, ppw = 0, nid_lo = 0, nid_hi = 0, et_lo = 0, et_hi = 0, it_lo = 0, it_hi = 0,
             pt_lo = 0, pt_hi = 0, ri_lo = 0, ri_hi = 0;
    int4 dob4, dod4;
    float4 risk4, inf4;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = i0 + a;
      const uint64_t slot = (uint64_t)(i - simp.slot_base);
      // disease timers and heterogeneity: every slot (newborns inherit theirs)
      const uint32_t et = to_timer(dist_sample(sp.dur_exp, seed, slot, SI_EXP));
      const uint32_t it = to_timer(dist_sample(sp.dur_inf, seed, slot, SI_INF));
      float pr = dist_sample(sp.t_to_paralysis, seed, slot, SI_PAR) - (float)et;
      const float pmax = (float)it;
      pr = pr < 0.0f ? 0.0f : pr > pmax ? pmax : pr;
      const uint32_t pt = (uint32_t)pr;
      float risk = 1.0f, inf = sp.inf_mean;
      if (sp.het_model != 0) {
        const uint4 o = init_draw(seed, slot, SI_HET, 0);
        const float z0 = normal_bm(o.x, o.y), z1 = normal_bm(o.z, o.w);
        const float y1 = sp.rho * z0 + sqrtf(1.0f - sp.rho * sp.rho) * z1;   // Cholesky of [[1, rho], [rho, 1]]
        risk = expf(sp.risk_mu + sp.risk_sigma * z0);
        if (sp.het_model == 1) {
          inf = sp.inf_mean * expf(sp.risk_mu + sp.risk_sigma * y1);
        } else {
          // Gamma(shape 1) quantile of Phi(y1): -mean * log(1 - Phi(y1)) = -mean * log(Phi(-y1))
          inf = -sp.inf_mean * logf(0.5f * erfcf(y1 * 0.70710678118654752f));
        }
      }
      int32_t dob = -1, dod = sp.vd ? 0 : INT_MAX;
      int16_t nid = -1, ri = -1;
      int8_t s = -1, missed = 0;
      if (i < live_hi) {                                   // an initial agent
        const int32_t age = init_age(sp, seed, slot);
        dob = -age;
        s = 0;
        const uint4 o = init_draw(seed, slot, SI_STATE, 0);
        if (sp.init_immun > 0.0f && u01o(o.x) < sp.init_immun) s = 3;
        const float ub = u01o(o.y);                        // one draw for all age bands
        for (int b = 0; b < sp.n_bands; ++b)
          if (age >= sp.band_lo[b] && age < sp.band_hi[b] && ub < sp.band_frac[b]) s = 3;
        if (sp.vd) dod = km_age_at_death(km, seed, slot, -1, age) - age;
        if (sp.ri) {
          const float days = sp.ri_hi > sp.ri_lo ? sp.ri_lo + u01o(init_draw(seed, slot, SI_RI, 0).x) * (sp.ri_hi - sp.ri_lo)
                                                 : sp.ri_lo;
          ri = (int16_t)(int32_t)((float)dob + days);
        }
        if (sp.missed_frac > 0.0f && u01o(o.z) < sp.missed_frac) missed = 1;
        nid = lnode;
      }
      st |= (uint32_t)(uint8_t)s << (8 * a);
      mi |= (uint32_t)(uint8_t)missed << (8 * a);
      ppw |= 0xFFu << (8 * a);                             // potentially_paralyzed -1
      const int sh = 16 * (a & 1);
      if (a < 2) {
        nid_lo |= (uint32_t)(uint16_t)nid << sh; et_lo |= et << sh; it_lo |= it << sh; pt_lo |= pt << sh;
        ri_lo |= (uint32_t)(uint16_t)ri << sh;
      } else {
        nid_hi |= (uint32_t)(uint16_t)nid << sh; et_hi |= et << sh; it_hi |= it << sh; pt_hi |= pt << sh;
        ri_hi |= (uint32_t)(uint16_t)ri << sh;
      }
      if (a == 0) { dob4.x = dob; dod4.x = dod; risk4.x = risk; inf4.x = inf; }
      else if (a == 1) { dob4.y = dob; dod4.y = dod; risk4.y = risk; inf4.y = inf; }
      else if (a == 2) { dob4.z = dob; dod4.z = dod; risk4.z = risk; inf4.z = inf; }
      else { dob4.w = dob; dod4.w = dod; risk4.w = risk; inf4.w = inf; }
    }
    reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = st;
    reinterpret_cast<uint32_t*>(d.strain)[i0 >> 2] = 0u;
    reinterpret_cast<uint32_t*>(d.pp)[i0 >> 2] = ppw;
    reinterpret_cast<uint32_t*>(d.paralyzed)[i0 >> 2] = 0u;
    reinterpret_cast<uint32_t*>(d.ipv)[i0 >> 2] = 0u;
    reinterpret_cast<uint32_t*>(d.missed)[i0 >> 2] = mi;
    *reinterpret_cast<uint2*>(d.node_id + i0) = make_uint2(nid_lo, nid_hi);
    *reinterpret_cast<uint2*>(d.exp_timer + i0) = make_uint2(et_lo, et_hi);
    *reinterpret_cast<uint2*>(d.inf_timer + i0) = make_uint2(it_lo, it_hi);
    *reinterpret_cast<uint2*>(d.par_timer + i0) = make_uint2(pt_lo, pt_hi);
    *reinterpret_cast<uint2*>(d.ri_timer + i0) = make_uint2(ri_lo, ri_hi);
    *reinterpret_cast<int4*>(d.dob + i0) = dob4;
    *reinterpret_cast<int4*>(d.dod + i0) = dod4;
    *reinterpret_cast<float4*>(d.risk + i0) = risk4;
    *reinterpret_cast<float4*>(d.infect + i0) = inf4;
  }
}

// The initial infections chosen by the host (exact counts per node): state I.
__global__ void k_init_infected(Dev d, const int64_t* slots, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = slots[j];
    if (i >= 0 && i < d.capacity && d.state[i] >= 0) d.state[i] = 2;
  }
}

}  // namespace lp
