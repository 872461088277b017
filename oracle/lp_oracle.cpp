// CPU oracle for the laser-polio daily agent step.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, function by function, the numba kernels of the reference
// (src/laser_polio/model.py) that make up one simulated day, keeping their pass
// structure: one parallel pass per kernel with thread-local [threads, nodes]
// bins reduced afterwards, a serial scan that builds per-node susceptible
// lists, per-node cumulative-weight sampling rounds, and a per-node host loop
// for the count draws.  It is compiled with g++ -O3 -fopenmp and used
//   * as the checker in tests/ (replay parity with the HIP path), and
//   * as the host-core CPU baseline in bench.py ("port").
// Nothing in the product links or loads it.
//
// Documented substitutions relative to the reference (each one is applied
// identically by the HIP path, so replay is bit-exact):
//   R1  random numbers: Philox4x32-10 (Salmon et al., SC'11) keyed by the sim
//       seed, counter = (index lo, index hi, day, stream | sub<<8), instead of
//       numpy/numba MT19937 streams.  Bernoulli draws compare the 32-bit output
//       r with an integer threshold ceil(p * 2^32)  (<=>  r*2^-32 < p).
//   R2  selection weights: acq_risk_multiplier quantised to int64 fixed point
//       W = rint(risk * 2^24); per-node sums and the sampling CDF are exact
//       integers (reference: float32 cumsum, model.py:1108).
//   R3  timers: uint16 with saturating decrement and ==0 tests instead of int8
//       with <=0 tests (model.py:419-447); equivalent on every reachable path.
//   R4  node-level FOI math in float32 with a fixed summation order and the
//       functions orc_expf/orc_expm1f/orc_log1pf below (Cody-Waite reduction +
//       Taylor polynomials); prob = -expm1(-rate) instead of 1-exp(-rate).
//   R5  the per-node Multinomial strain split (model.py:1400-1407) is not
//       drawn: tx_infect_nb only uses its total (model.py:1068) and re-draws
//       each exposed agent's strain (model.py:1134-1141).
//
// Parity status per function is listed in the header of oracle/oracle.py.

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr uint32_t ST_PARALYSIS = 1, ST_RI = 2, ST_SIA = 3, ST_STRAIN = 4, ST_KM = 5,
                   ST_BIRTH = 6, ST_COUNT = 7, ST_PROBE = 8;

inline void philox(uint32_t c[4], uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0; c[1] = (uint32_t)p1; c[2] = n2; c[3] = (uint32_t)p0;
  }
}

inline void draw(uint32_t out[4], uint64_t seed, uint64_t idx, int32_t t, uint32_t stream) {
  out[0] = (uint32_t)idx; out[1] = (uint32_t)(idx >> 32); out[2] = (uint32_t)t; out[3] = stream;
  philox(out, seed);
}

inline uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

// ---- float32 elementary functions (R4) -------------------------------------
constexpr float LN2_HI = 0.693145751953125f;      // 0x3f317200, 15 significant bits
constexpr float LN2_LO = 1.42860676533018672e-06f;
constexpr float LOG2E = 1.44269502162933349609f;

inline float em1_poly(float r) {  // expm1(r), |r| <= ln2/2, Taylor to r^8/8!
  float p = 2.48015873015873e-05f;            // 1/40320
  p = p * r + 1.98412698412698e-04f;          // 1/5040
  p = p * r + 1.38888888888889e-03f;          // 1/720
  p = p * r + 8.33333333333333e-03f;          // 1/120
  p = p * r + 4.16666666666667e-02f;          // 1/24
  p = p * r + 1.66666666666667e-01f;          // 1/6
  p = p * r + 0.5f;
  return r + (r * r) * p;
}

inline float orc_expm1f(float x) {
  if (x < -17.5f) return -1.0f;
  if (x > 88.0f) return INFINITY;
  if (std::fabs(x) <= 0.34657359f) return em1_poly(x);
  const float kf = std::nearbyint(x * LOG2E);
  const float r = (x - kf * LN2_HI) - kf * LN2_LO;
  const float e = em1_poly(r);
  const int k = (int)kf;
  const float s = std::ldexp(1.0f, k);
  return std::fma(s, e, s - 1.0f);
}

inline float orc_expf(float x) {
  if (x < -87.0f) return 0.0f;
  if (x > 88.0f) return INFINITY;
  const float kf = std::nearbyint(x * LOG2E);
  const float r = (x - kf * LN2_HI) - kf * LN2_LO;
  return std::ldexp(1.0f + em1_poly(r), (int)kf);
}

inline float orc_logf(float y) {  // y > 0, normal
  uint32_t b; std::memcpy(&b, &y, 4);
  int e = (int)((b >> 23) & 0xffu) - 127;
  b = (b & 0x7fffffu) | 0x3f800000u;
  float f; std::memcpy(&f, &b, 4);
  if (f > 1.41421356f) { f = f * 0.5f; e += 1; }
  const float s = (f - 1.0f) / (f + 1.0f);
  const float z = s * s;
  float p = 0.0769230769231f;                // 1/13
  p = p * z + 0.0909090909091f;               // 1/11
  p = p * z + 0.111111111111f;                // 1/9
  p = p * z + 0.142857142857f;                // 1/7
  p = p * z + 0.2f;
  p = p * z + 0.333333333333f;
  const float lf = (s + s) + (s + s) * (z * p);
  const float ef = (float)e;
  return ef * LN2_HI + (ef * LN2_LO + lf);
}

inline float orc_log1pf(float x) {  // x >= 0
  const float y = 1.0f + x;
  if (y == 1.0f) return x;
  const float c = (y - 1.0f) - x;
  return orc_logf(y) - c / y;
}

inline uint64_t thr_from_prob(float p) {  // ceil(p * 2^32), clamped to [0, 2^32]
  if (!(p > 0.0f)) return 0;
  const float s = p * 4294967296.0f;
  if (s >= 4294967296.0f) return 4294967296ull;
  return (uint64_t)std::ceil(s);
}

inline float u24(uint32_t r) { return (float)(r >> 8) * 5.9604644775390625e-08f; }  // [0,1) on 2^-24 grid

// Poisson(lam) as a sum of independent Poisson chunks: floor(lam/16) chunks of
// mean 16 and one chunk of the remainder, each by sequential inversion.
int64_t poisson_chunked(float lam, uint64_t seed, uint32_t node, int32_t t) {
  const float nf = std::floor(lam / 16.0f);
  const int64_t n_full = (int64_t)nf;
  const float last = lam - 16.0f * nf;
  const int64_t nch = n_full + (last > 0.0f ? 1 : 0);
  int64_t total = 0;
  for (int64_t j = 0; j < nch; ++j) {
    const float lc = j < n_full ? 16.0f : last;
    uint32_t o[4];
    o[0] = (uint32_t)j; o[1] = node; o[2] = (uint32_t)t; o[3] = ST_COUNT; philox(o, seed);
    const float u = u24(o[0]);
    float p = orc_expf(-lc), F = p;
    int k = 0;
    while (u >= F) {
      k += 1;
      p = (p * lc) / (float)k;
      const float Fn = F + p;
      if (Fn == F) break;
      F = Fn;
    }
    total += k;
  }
  return total;
}

// NegBin(r, p) with p = 1/(1+q), q = mean/r, as a sum of NB(r_c, p) chunks by inversion.
int64_t negbin_chunked(int32_t r, float q, uint64_t seed, uint32_t node, int32_t t) {
  const float lq = orc_log1pf(q);          // -log p
  if (!(lq > 0.0f)) return 0;
  const float omp = q / (1.0f + q);        // 1 - p
  int32_t rc_max = (int32_t)(16.0f / lq);
  if (rc_max < 1) rc_max = 1;
  int64_t total = 0;
  int32_t rem = r;
  uint32_t j = 1;
  while (rem > 0) {
    const int32_t rc = rem < rc_max ? rem : rc_max;
    rem -= rc;
    uint32_t o[4];
    o[0] = j; o[1] = node; o[2] = (uint32_t)t; o[3] = ST_COUNT; philox(o, seed);
    const float u = u24(o[0]);
    float P = orc_expf(-((float)rc * lq)), F = P;
    int64_t k = 0;
    while (u >= F) {
      P = (P * ((float)(k + rc) * omp)) / (float)(k + 1);
      k += 1;
      const float Fn = F + P;
      if (Fn == F) break;
      F = Fn;
    }
    total += k;
    j += 1;
  }
  return total;
}

inline int32_t km_age_at_death(const int64_t* cd, int32_t km_len, uint64_t seed, uint64_t slot,
                               int32_t t, int32_t age) {
  uint32_t o[4];
  draw(o, seed, slot, t, ST_KM);
  const int32_t max_year = km_len - 1;
  int32_t ay = age / 365;
  if (ay > max_year) ay = max_year;
  const int64_t lo = ay == 0 ? 0 : cd[ay - 1];
  const int64_t hi = cd[max_year];
  int32_t yod = ay;
  if (hi > lo) {
    const uint64_t u = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
    const int64_t v = lo + (int64_t)mulhi64(u, (uint64_t)(hi - lo));
    yod = (int32_t)(std::upper_bound(cd, cd + km_len, v) - cd);
  }
  int32_t lo_d = yod * 365;
  if (lo_d < age + 1) lo_d = age + 1;
  const int32_t hi_d = (yod + 1) * 365;
  if (hi_d <= lo_d) return lo_d;
  return lo_d + (int32_t)(((uint64_t)o[2] * (uint64_t)(hi_d - lo_d)) >> 32);
}

// expected_births = step_size * birth_rate * pop[t-1] in float64 (model.py:1712; brs is
// step_size * birth_rate, the product's first factor): returns its integer part
// (astype(int32), model.py:1713) and sets *thr to the Bernoulli threshold of the
// fractional part (model.py:1714-1715), ceil(frac * 2^32) (R1).
int32_t birth_expected(double brs, int32_t pop, uint64_t* thr) {
  *thr = 0;
  const double e = brs * (double)pop;
  if (!(e > 0.0)) return 0;
  const int32_t bi = (int32_t)e;
  const double frac = e - (double)bi;
  if (frac > 0.0) *thr = (uint64_t)std::ceil(frac * 4294967296.0);
  return bi;
}

}  // namespace

extern "C" {

// Everything one single-sim oracle run needs.  Layout mirrored by oracle.py (ctypes).
struct orc_model {
  int64_t capacity;
  int32_t n_nodes, n_strains, n_threads, pad0;
  uint64_t seed;
  int8_t* state; int8_t* strain; int16_t* node_id;
  uint16_t* exp_timer; uint16_t* inf_timer; uint16_t* par_timer;
  int8_t* pp; int8_t* paralyzed; int8_t* ipv; uint8_t* missed;
  float* risk; float* infect; int32_t* dob; int32_t* dod; int16_t* ri_timer;
  int64_t* seg_begin; int64_t* live_end; int64_t* seg_end;
  float* r0s; float* keep; double* birth_rate_step; float* NT;
  uint64_t* thr_ri; uint64_t* thr_ipv;
  int64_t* km_cd; int32_t km_len; int32_t newborn_ri_timer;
  float strain_r0[4];
  uint64_t thr_par; uint64_t thr_zi;
  int32_t disp_r; int32_t ri_strain; int32_t ri_step; float zi;
};

int64_t orc_struct_size(void) { return (int64_t)sizeof(orc_model); }
int32_t orc_birth_expected(double brs, int32_t pop, uint64_t* thr) { return birth_expected(brs, pop, thr); }
int32_t orc_max_threads(void) { return omp_get_max_threads(); }

void orc_philox(const uint32_t* ctr, uint64_t seed, uint32_t* out) {
  for (int i = 0; i < 4; ++i) out[i] = ctr[i];
  philox(out, seed);
}
float orc_expm1f_c(float x) { return orc_expm1f(x); }
float orc_expf_c(float x) { return orc_expf(x); }
float orc_log1pf_c(float x) { return orc_log1pf(x); }
int64_t orc_poisson(float lam, uint64_t seed, uint32_t node, int32_t t) { return poisson_chunked(lam, seed, node, t); }
int64_t orc_negbin(int32_t r, float q, uint64_t seed, uint32_t node, int32_t t) { return negbin_chunked(r, q, seed, node, t); }
int32_t orc_km_age_at_death(const int64_t* cd, int32_t km_len, uint64_t seed, uint64_t slot, int32_t t, int32_t age) {
  return km_age_at_death(cd, km_len, seed, slot, t, age);
}

static inline int nthreads(const orc_model* m) { return m->n_threads > 0 ? m->n_threads : omp_get_max_threads(); }

// get_deaths (model.py:1767-1781): state>=0 && dod<=t -> dead; per-node counts via TLS bins.
void orc_get_deaths(const orc_model* m, int32_t t, int32_t* deaths) {
  const int T = nthreads(m), N = m->n_nodes;
  std::vector<int32_t> tls((size_t)T * N, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    if (m->state[i] >= 0 && m->dod[i] <= t) {
      m->state[i] = -1;
      tls[(size_t)omp_get_thread_num() * N + m->node_id[i]] += 1;
    }
  }
  for (int n = 0; n < N; ++n) { int32_t s = 0; for (int k = 0; k < T; ++k) s += tls[(size_t)k * N + n]; deaths[n] = s; }
}

// VitalDynamics_ABM births (model.py:1711-1734): floor(step*rate*pop[t-1]) + Bernoulli(frac),
// newborns appended to the node's tail.  Returns the number of nodes whose capacity ran out.
int32_t orc_births(const orc_model* m, int32_t t, const int32_t* pop_prev, int32_t* births) {
  const int N = m->n_nodes;
  int32_t overflow = 0;
  std::vector<int64_t> nb_begin(N);
  for (int n = 0; n < N; ++n) {
    uint64_t thr = 0;
    int32_t b = birth_expected(m->birth_rate_step[n], pop_prev[n], &thr);
    if (thr) {
      uint32_t o[4];
      o[0] = 0; o[1] = (uint32_t)n; o[2] = (uint32_t)t; o[3] = ST_BIRTH; philox(o, m->seed);
      b += (uint64_t)o[0] < thr ? 1 : 0;
    }
    const int64_t room = m->seg_end[n] - m->live_end[n];
    if (b > room) { b = (int32_t)room; overflow += 1; }
    births[n] = b;
    nb_begin[n] = m->live_end[n];
    m->live_end[n] += b;
  }
  for (int n = 0; n < N; ++n) {
    const int64_t lo = nb_begin[n], hi = m->live_end[n];
#pragma omp parallel for num_threads(nthreads(m)) schedule(static)
    for (int64_t i = lo; i < hi; ++i) {
      m->dob[i] = t;
      m->dod[i] = t + km_age_at_death(m->km_cd, m->km_len, m->seed, (uint64_t)i, t, 0);
      m->state[i] = 0;
      m->node_id[i] = (int16_t)n;
      if (m->newborn_ri_timer >= 0) m->ri_timer[i] = (int16_t)m->newborn_ri_timer;
    }
  }
  return overflow;
}

// disease_state_step_kernel (model.py:393-454) + wrapper reduction (model.py:366-390).
void orc_disease_state_step(const orc_model* m, int32_t t, int32_t* new_pp, int32_t* new_p) {
  const int T = nthreads(m), N = m->n_nodes;
  std::vector<int32_t> tpp((size_t)T * N, 0), tp((size_t)T * N, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    int8_t s = m->state[i];
    if (s == 1) {
      if (m->exp_timer[i] == 0) { s = 2; m->state[i] = 2; }
      if (m->exp_timer[i] > 0) m->exp_timer[i] -= 1;
    }
    if (s == 2) {
      if (m->inf_timer[i] == 0) m->state[i] = 3;
      if (m->inf_timer[i] > 0) m->inf_timer[i] -= 1;
      if (m->strain[i] == 0) {
        if (m->par_timer[i] == 0 && m->pp[i] == -1) {
          const int tid = omp_get_thread_num();
          const int n = m->node_id[i];
          if (m->ipv[i] == 0) {
            m->pp[i] = 1;
            tpp[(size_t)tid * N + n] += 1;
            uint32_t o[4];
            draw(o, m->seed, (uint64_t)i, t, ST_PARALYSIS);
            if ((uint64_t)o[0] < m->thr_par) { m->paralyzed[i] = 1; tp[(size_t)tid * N + n] += 1; }
          } else {
            m->pp[i] = 0;
          }
        }
        if (m->par_timer[i] > 0) m->par_timer[i] -= 1;
      }
    }
  }
  for (int n = 0; n < N; ++n) {
    int32_t a = 0, b = 0;
    for (int k = 0; k < T; ++k) { a += tpp[(size_t)k * N + n]; b += tp[(size_t)k * N + n]; }
    new_pp[n] = a; new_p[n] = b;
  }
}

// fast_ri (model.py:1784-1855).
void orc_fast_ri(const orc_model* m, int32_t t, int32_t* ri_vacc, int32_t* ri_prot, int32_t* ipv_vacc) {
  const int T = nthreads(m), N = m->n_nodes;
  const int32_t step = m->ri_step;
  std::vector<int32_t> a((size_t)T * N, 0), b((size_t)T * N, 0), c((size_t)T * N, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    const int8_t s = m->state[i];
    if (s < 0 || m->missed[i] == 1) continue;
    const int n = m->node_id[i];
    const int16_t timer = (int16_t)(m->ri_timer[i] - step);
    m->ri_timer[i] = timer;
    bool eligible = false;
    if (t == step) eligible = timer <= 0 && timer >= -step;
    else if (t > step) eligible = timer <= 0 && timer > -step;
    if (!eligible) continue;
    uint32_t o[4];
    draw(o, m->seed, (uint64_t)i, t, ST_RI);
    const int tid = omp_get_thread_num();
    if ((uint64_t)o[0] < m->thr_ri[n]) {
      a[(size_t)tid * N + n] += 1;
      if (s == 0) { m->state[i] = 1; m->strain[i] = (int8_t)m->ri_strain; b[(size_t)tid * N + n] += 1; }
    }
    if ((uint64_t)o[1] < m->thr_ipv[n]) { c[(size_t)tid * N + n] += 1; m->ipv[i] = 1; }
  }
  for (int n = 0; n < N; ++n) {
    int32_t x = 0, y = 0, z = 0;
    for (int k = 0; k < T; ++k) { x += a[(size_t)k * N + n]; y += b[(size_t)k * N + n]; z += c[(size_t)k * N + n]; }
    ri_vacc[n] = x; ri_prot[n] = y; ipv_vacc[n] = z;
  }
}

// fast_sia (model.py:1994-2060) for one campaign event; thr1/thr2 are 0 for untargeted nodes.
void orc_fast_sia(const orc_model* m, int32_t t, int32_t ev, int32_t sia_strain, int32_t min_age, int32_t max_age,
                  const uint64_t* thr1, const uint64_t* thr2, const uint8_t* targeted,
                  int32_t* vacc, int32_t* prot) {
  const int T = nthreads(m), N = m->n_nodes;
  std::vector<int32_t> a((size_t)T * N, 0), b((size_t)T * N, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    if (m->state[i] < 0 || m->missed[i] == 1) continue;
    const int32_t age = t - m->dob[i];
    if (!(min_age <= age && age <= max_age)) continue;
    const int n = m->node_id[i];
    if (targeted[n] == 0) continue;
    uint32_t o[4];
    draw(o, m->seed, (uint64_t)i, t, ST_SIA | ((uint32_t)ev << 8));
    if ((uint64_t)o[0] < thr1[n]) {
      const int tid = omp_get_thread_num();
      a[(size_t)tid * N + n] += 1;
      if (m->state[i] == 0 && (uint64_t)o[0] < thr2[n]) {
        m->state[i] = 1; m->strain[i] = (int8_t)sia_strain; b[(size_t)tid * N + n] += 1;
      }
    }
  }
  for (int n = 0; n < N; ++n) {
    int32_t x = 0, y = 0;
    for (int k = 0; k < T; ++k) { x += a[(size_t)k * N + n]; y += b[(size_t)k * N + n]; }
    vacc[n] = x; prot[n] = y;
  }
}

static inline int64_t risk_weight(float r) {
  if (!(r > 0.0f)) return 0;
  return (int64_t)std::nearbyint(r * 16777216.0f);
}

// tx_step_prep_kernel (model.py:980-1007): per node #S, sum of S weights, and per node x strain
// sum of daily_infectivity * strain_r0 over I agents (R2 fixed point).
void orc_tx_step_prep(const orc_model* m, int32_t* S_cnt, int32_t* S_pos, int64_t* SW, int64_t* IB) {
  const int T = nthreads(m), N = m->n_nodes, K = m->n_strains;
  std::vector<int32_t> sc((size_t)T * N, 0), sp((size_t)T * N, 0);
  std::vector<int64_t> sw((size_t)T * N, 0), ib((size_t)T * N * K, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    const int8_t s = m->state[i];
    const int tid = omp_get_thread_num();
    if (s == 0) {
      const int n = m->node_id[i];
      const int64_t w = risk_weight(m->risk[i]);
      sc[(size_t)tid * N + n] += 1;
      sp[(size_t)tid * N + n] += (w > 0);
      sw[(size_t)tid * N + n] += w;
    } else if (s == 2) {
      const int n = m->node_id[i];
      const int k = m->strain[i];
      const float v = m->infect[i] * m->strain_r0[k];
      ib[((size_t)tid * N + n) * K + k] += risk_weight(v);
    }
  }
  for (int n = 0; n < N; ++n) {
    int32_t a = 0, b = 0; int64_t c = 0;
    for (int q = 0; q < T; ++q) { a += sc[(size_t)q * N + n]; b += sp[(size_t)q * N + n]; c += sw[(size_t)q * N + n]; }
    S_cnt[n] = a; S_pos[n] = b; SW[n] = c;
    for (int k = 0; k < K; ++k) {
      int64_t d = 0;
      for (int q = 0; q < T; ++q) d += ib[((size_t)q * N + n) * K + k];
      IB[(size_t)n * K + k] = d;
    }
  }
}

// Transmission_ABM.step Part 2b (model.py:1327-1407): network mix, seasonality, r0 scalars,
// rate -> probability, expected exposures, per-node Poisson / zero-inflated NB count.
// Outputs per node: p[n*K+k], expected[n], total[n], size[n], take_all[n], thr_strain[n*K+k].
void orc_node_map(const orc_model* m, int32_t t, float season, const int32_t* pop_t,
                  const int32_t* S_cnt, const int32_t* S_pos, const int64_t* SW, const int64_t* IB,
                  float* p, float* expected, int64_t* total, int32_t* size, int32_t* take_all,
                  uint64_t* thr_strain, float* beta_mixed) {
  const int N = m->n_nodes, K = m->n_strains;
  std::vector<float> beta((size_t)N * K);
  for (int j = 0; j < N; ++j)
    for (int k = 0; k < K; ++k) beta[(size_t)j * K + k] = (float)IB[(size_t)j * K + k] * 5.9604644775390625e-08f;
  float part[256];
  for (int i = 0; i < N; ++i) {
    float bm[4];
    for (int k = 0; k < K; ++k) {
      for (int l = 0; l < 256; ++l) {
        float acc = 0.0f;
        for (int j = l; j < N; j += 256) acc = std::fma(m->NT[(size_t)i * N + j], beta[(size_t)j * K + k], acc);
        part[l] = acc;
      }
      for (int w = 128; w >= 1; w >>= 1)
        for (int l = 0; l < w; ++l) part[l] = part[l] + part[l + w];
      bm[k] = std::fma(beta[(size_t)i * K + k], m->keep[i], part[0]);
    }
    const int32_t popc = pop_t[i] > 1 ? pop_t[i] : 1;
    float ptot = 0.0f;
    float pk[4];
    bool local_zero = true;
    for (int k = 0; k < K; ++k) {
      if (beta_mixed) beta_mixed[(size_t)i * K + k] = bm[k];
      const float b2 = (bm[k] * season) * m->r0s[i];
      const float rate = b2 / (float)popc;
      pk[k] = -orc_expm1f(-rate);
      if (!(pk[k] > 0.0f)) pk[k] = 0.0f;
      p[(size_t)i * K + k] = pk[k];
      ptot = ptot + pk[k];
      if (IB[(size_t)i * K + k] != 0) local_zero = false;
    }
    const float expo = (float)SW[i] * 5.9604644775390625e-08f;
    const float ex = expo * ptot;
    expected[i] = ex;
    int64_t tot = 0;
    if (ex > 0.0f) {
      if (local_zero) {                      // zero-inflated NB (model.py:1381-1393)
        if (m->zi < 1.0f) {
          uint32_t o[4];
          o[0] = 0xFFFFFFFFu; o[1] = (uint32_t)i; o[2] = (uint32_t)t; o[3] = ST_COUNT; philox(o, m->seed);
          if ((uint64_t)o[0] >= m->thr_zi) {
            const float dm = ex / (1.0f - m->zi);
            tot = negbin_chunked(m->disp_r, dm / (float)m->disp_r, m->seed, (uint32_t)i, t);
          }
        }
      } else {                               // Poisson (model.py:1395-1397)
        tot = poisson_chunked(ex, m->seed, (uint32_t)i, t);
      }
    }
    total[i] = tot;
    size[i] = (int32_t)(tot < (int64_t)S_cnt[i] ? tot : (int64_t)S_cnt[i]);
    take_all[i] = (size[i] > 0 && size[i] >= S_pos[i]) ? 1 : 0;
    float cum = 0.0f;
    for (int k = 0; k < K; ++k) {
      cum = cum + (ptot > 0.0f ? pk[k] / ptot : 0.0f);
      thr_strain[(size_t)i * K + k] = thr_from_prob(cum);
    }
  }
}

// tx_infect_nb (model.py:1010-1149): serial scan builds per-node susceptible lists, then per node
// (parallel) weighted sampling without replacement by rounds of with-replacement probes + unique,
// then one strain draw per selected agent.
void orc_tx_infect(const orc_model* m, int32_t t, const int32_t* S_cnt, const int32_t* size,
                   const int32_t* take_all, const uint64_t* thr_strain, int32_t* new_exp) {
  const int N = m->n_nodes, K = m->n_strains;
  std::vector<int64_t> off(N + 1, 0);
  for (int n = 0; n < N; ++n) off[n + 1] = off[n] + S_cnt[n];
  std::vector<int64_t> sus_idx(off[N]);
  std::vector<int64_t> sus_w(off[N]);
  std::vector<int64_t> next(off.begin(), off.end() - 1);
  for (int64_t i = 0; i < m->capacity; ++i) {          // serial scan (model.py:1055-1061)
    if (m->state[i] != 0) continue;
    const int n = m->node_id[i];
    if (size[n] <= 0) continue;
    const int64_t idx = next[n]++;
    sus_idx[idx] = i;
    sus_w[idx] = risk_weight(m->risk[i]);
  }
  std::vector<uint8_t> already(m->capacity, 0);         // model.py:1063
  for (int n = 0; n < N * K; ++n) new_exp[n] = 0;
#pragma omp parallel for num_threads(nthreads(m)) schedule(dynamic, 1)
  for (int n = 0; n < N; ++n) {
    const int32_t sz = size[n];
    if (sz <= 0 || S_cnt[n] == 0) continue;
    const int64_t o0 = off[n], cnt = S_cnt[n];
    std::vector<int64_t> selected;
    if (take_all[n]) {
      for (int64_t q = 0; q < cnt; ++q) if (sus_w[o0 + q] > 0) selected.push_back(q);
    } else {
      std::vector<int64_t> w(sus_w.begin() + o0, sus_w.begin() + o0 + cnt);
      std::vector<int64_t> cdf(cnt);
      int64_t n_uniq = 0;
      uint32_t round = 1;
      while (n_uniq < sz) {
        int64_t acc = 0;
        for (int64_t q = 0; q < cnt; ++q) { acc += w[q]; cdf[q] = acc; }
        if (acc <= 0) break;
        const int64_t mr = sz - n_uniq;
        std::vector<int64_t> hits(mr);
        for (int64_t j = 0; j < mr; ++j) {
          uint32_t o[4];
          o[0] = (uint32_t)j; o[1] = (uint32_t)n; o[2] = (uint32_t)t; o[3] = ST_PROBE | (round << 8);
          philox(o, m->seed);
          const uint64_t u = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
          const int64_t T = (int64_t)mulhi64(u, (uint64_t)acc);
          hits[j] = std::upper_bound(cdf.begin(), cdf.end(), T) - cdf.begin();
        }
        std::sort(hits.begin(), hits.end());
        hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
        for (int64_t q : hits) { selected.push_back(q); w[q] = 0; }
        n_uniq += (int64_t)hits.size();
        round += 1;
      }
    }
    std::vector<int32_t> local(K, 0);
    for (int64_t q : selected) {
      const int64_t i = sus_idx[o0 + q];
      if (already[i] || m->state[i] != 0) continue;
      uint32_t o[4];
      draw(o, m->seed, (uint64_t)i, t, ST_STRAIN);
      int k = 0;
      for (int s = 0; s < K; ++s) { if ((uint64_t)o[0] < thr_strain[(size_t)n * K + s]) { k = s; break; } }
      m->state[i] = 1; m->strain[i] = (int8_t)k; already[i] = 1; local[k] += 1;
    }
    for (int k = 0; k < K; ++k) new_exp[(size_t)n * K + k] = local[k];
  }
}

// count_SEIRP_kernel (model.py:902-929) + wrapper reduction (model.py:876-899).
void orc_count_seirp(const orc_model* m, int32_t* S, int32_t* Ebs, int32_t* Ibs, int32_t* R,
                     int32_t* POTP, int32_t* P) {
  const int T = nthreads(m), N = m->n_nodes, K = m->n_strains;
  std::vector<int32_t> s((size_t)T * N, 0), r((size_t)T * N, 0), pp((size_t)T * N, 0), pa((size_t)T * N, 0);
  std::vector<int32_t> e((size_t)T * N * K, 0), in((size_t)T * N * K, 0);
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < m->capacity; ++i) {
    const int8_t ds = m->state[i];
    if (ds < 0) continue;
    const int tid = omp_get_thread_num();
    const int n = m->node_id[i];
    const int k = m->strain[i];
    if (ds == 0) s[(size_t)tid * N + n] += 1;
    else if (ds == 1) e[((size_t)tid * N + n) * K + k] += 1;
    else if (ds == 2) in[((size_t)tid * N + n) * K + k] += 1;
    else if (ds == 3) r[(size_t)tid * N + n] += 1;
    if (m->pp[i] == 1) pp[(size_t)tid * N + n] += 1;
    if (m->paralyzed[i] == 1) pa[(size_t)tid * N + n] += 1;
  }
  for (int n = 0; n < N; ++n) {
    int32_t a = 0, b = 0, c = 0, d = 0;
    for (int q = 0; q < T; ++q) {
      a += s[(size_t)q * N + n]; b += r[(size_t)q * N + n]; c += pp[(size_t)q * N + n]; d += pa[(size_t)q * N + n];
    }
    S[n] = a; R[n] = b; POTP[n] = c; P[n] = d;
    for (int k = 0; k < K; ++k) {
      int32_t x = 0, y = 0;
      for (int q = 0; q < T; ++q) { x += e[((size_t)q * N + n) * K + k]; y += in[((size_t)q * N + n) * K + k]; }
      Ebs[(size_t)n * K + k] = x; Ibs[(size_t)n * K + k] = y;
    }
  }
}

}  // extern "C"
