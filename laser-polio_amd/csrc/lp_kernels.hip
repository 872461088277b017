#pragma once
// Kernels of the laser-polio daily agent step on MI355X (gfx950, wave64).
//
// One simulated day is this kernel sequence (all stages fused, reference order
// VitalDynamics -> DiseaseState -> RI -> SIA -> Transmission -> log):
//   k_day_begin    per node: births (count + newborn slot range), per-day resets,
//                  early-stop rule                     model.py:1711-1719, 788-795
//   agent pass     k_ei: DiseaseState and the E / I tally over the per-node lists
//                  of E and I agents (every day the lists are complete; plain days
//                  need nothing else); k_full<VD|RI|SIA> on the other days: births,
//                  deaths (only chunks with a death due), RI and SIA over the
//                  node-segmented agent slots, one workgroup per span of tiles;
//                  k_plain / k_full<F, true> run DiseaseState inline instead when
//                  no list is available, k_agents when every carried statistic is
//                  recounted               model.py:1767-1781, 393-454, 1784-1855,
//                                              1994-2060, 980-1007, 902-929
//   k_node_reduce  per node: reduce span partials (and k_ei's accumulators), carried
//                  S / R counts, exclusive prefix of the span weights, result rows
//                  of VD/DS/RI/SIA
//   k_mix_foi      per node: network mix (fixed-order tree), seasonality, r0 scalars,
//                  rate -> probability, expected exposures, Poisson / zero-inflated NB
//                  count draw (chunks in parallel across the wave)  model.py:1327-1407
//   k_plan, k_sample, k_finish
//                  weighted sampling without replacement by rounds of with-
//                  replacement probes + unique (model.py:1089-1122): every probe
//                  searches its node's span prefix, then a span's chunks, then one
//                  chunk, one wavefront per probe; round 1 grid-wide (planned
//                  inside k_sample, by k_plan above 8191 nodes; k_sample claims
//                  its agents, k_finish makes them E), later rounds and the
//                  take-all nodes per node in k_finish; each exposed agent's strain is drawn
//                  as it is exposed (S -> E) and it joins the next day's E / I list
//                                                               model.py:1125-1147
//   k_day_end      per node: log rows + transmission incidence; its last
//                  workgroup plans the next day's list pieces and advances the day
//                  counter                             model.py:1436-1441, 1461-1483
// A day is replayed as one hipGraph (one graph per day kind).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>
#include <type_traits>
#include "lp_device.h"
#include "../../include/lp_step.h"

namespace lp {

constexpr int WG = 256;
constexpr int MAX_TPW = 16;         // tiles per span at most
constexpr int NPART = 32;           // fields of a tally partial
constexpr int8_t SEL_MARK = 0x7F;   // state byte of an agent selected for exposure today
constexpr uint64_t SEL_INVALID = ~0ull;
constexpr uint64_t SEL_SLOT = (1ull << 40) - 1;   // slot field of a selected-list entry; all ones: no agent
constexpr int MAX_ROUNDS = 4096;    // bound on the finisher's rounds (each round claims >= 1 agent)

// Tally fields.
enum : int {
  P_S = 0, P_SPOS, P_R, P_POTP, P_P, P_DEATHS, P_DPP, P_DP, P_NEWPP, P_NEWP,
  P_RIV, P_RIP, P_IPV, P_SIAV, P_SIAP,
  P_E0 = 15, P_I0 = 19, P_SIAPK0 = 23, P_SW = 27, P_IB0 = 28
};

enum : int { ERR_CAPACITY = 1, ERR_OVERFLOW = 2, ERR_SAMPLING = 4, ERR_KIND = 8, ERR_LIST = 16 };

// Per-node constants the list and sampling kernels need, gathered in one 64-byte
// record (rebuilt by k_node_info after every RECOUNT pass).
struct NodeInfo {
  uint64_t seed, thr_paralysis;
  int64_t slot_base, seg_begin;
  float r0k[LP_MAX_STRAINS];
  int32_t sim, lnode;          // sim index; node index inside the sim
  int32_t span_begin, n_spans;
  // the simulation's fields the network mix and count draw read (one load level, not two)
  int64_t net_off;
  uint64_t thr_zi;
  float zi;
  int32_t disp_r, sim_nodes, sim_node_begin;
};

// Per node, lane l < 64: span l of the node's (at most 64) spans, its exclusive prefix of
// the carried selection weights and its tiles, written by k_node_reduce so that the
// sampler loads them with the node's other data (prefix INT64_MAX past the last span).
struct SpanLane {
  int64_t prefix;
  int32_t tile0, ntiles;
};

struct Dev {
  // agent columns
  int8_t* state; int8_t* strain; int16_t* node_id;
  uint16_t* exp_timer; uint16_t* inf_timer; uint16_t* par_timer;
  int8_t* pp; int8_t* paralyzed; int8_t* ipv; uint8_t* missed;
  float* risk; float* infect; int32_t* dob; int32_t* dod; int16_t* ri_timer;
  uint4* rec;   // [capacity] agent records: the DiseaseState / tally fields of each slot in 16 B
  // node tables
  const int64_t* seg_begin; int64_t* live_end; const int64_t* seg_end; const int32_t* node_sim;
  const float* r0s; const float* keep;
  const uint64_t* birth_rate_step;   // step * birth rate per node, float64 bit patterns
  const uint64_t* thr_ri; const uint64_t* thr_ipv;
  // sim tables
  const lp_sim_params* sims; const float* network_t; const int64_t* km_cd;
  // schedule
  const int32_t* day_flags; const int32_t* day_sia_begin; const int32_t* day_sia_count;
  const float* season;
  const int32_t* sia_strain; const int32_t* sia_min_age; const int32_t* sia_max_age; const uint64_t* sia_thr;
  // results
  int32_t *r_S, *r_E, *r_I, *r_R, *r_Ebs, *r_Ibs, *r_potp, *r_p, *r_newpp, *r_newp, *r_pop,
          *r_births, *r_deaths, *r_newexp, *r_newexp_bs, *r_riv, *r_rip, *r_ri_newexp_bs, *r_ipv,
          *r_siav, *r_siap, *r_sia_newexp_bs;
  // scratch
  int32_t* day;   // [8] tick, end of the lp_run_days call, unused x2, the day record (today())
  int32_t* stop_day; int32_t* err;
  const int4* spans; const int32_t* node_span_begin;
  int64_t* partials; int64_t* chunk_sum;
  int64_t* nb_begin; int32_t* births;
  int64_t* node_tally; int32_t* potp_run; int32_t* p_run;
  int64_t* node_carry;   // [nodes][3] carried counts: S, S with positive weight, R
  int32_t* chunk_min_dod;  // [chunks] earliest date_of_death among the chunk's live agents
  uint4* ei_list;          // [2][capacity] per-node lists of E / I agents (16-byte entries)
  int32_t* ei_cnt;         // [2][nodes] list lengths
  int32_t* ei_piece_off;   // [nodes + 1] prefix over nodes of the day's list pieces
  int64_t* node_acc;       // [nodes][32] tallies accumulated by k_ei (zeroed each day)
  float* beta_f;           // [nodes][4] infectious weight per strain (node tally / 2^24)
  float* mix_buf;          // [nodes][4] network mix sum_j N_ji beta_j on MFMA (mix_mode 1), else null
  int32_t* sim_cases;      // [sims] 1 when a node of the sim logged E or I today
  NodeInfo* ninfo;         // [nodes]
  const int32_t* ri_cnt;   // [nt + 1] ri_cnt[t + 1] = RI days among ticks 0..t
  int32_t* ri_ref;         // [2] ri_timer reference tick; latest tick whose RI step is applied
  uint32_t* ri_list;       // [2][capacity] per-node lists of RI candidates (slots)
  int32_t* ri_lcnt;        // [2][nodes] their lengths
  float* p_dbg; float* expected_dbg; int64_t* total_dbg; float* beta_mixed;
  int64_t* birth_dbg;      // [nodes][2] integer part and fraction threshold of the last VD step's expected births
  int32_t* size; int32_t* take_all; int32_t* n_uniq; int32_t* m_round; int64_t* W_round;
  uint64_t* thr_strain; int32_t* newexp;
  int64_t* probe_off; int64_t* round_base; uint64_t* sel_list; int64_t* sel_count; int64_t sel_cap;
  uint32_t* sel_word;      // [sel_cap] round 1: a claimed agent's selection weight (all ones: recompute)
  int64_t* fin_prefix; int64_t* span_prefix;
  SpanLane* span_lane;     // [nodes][64]
  int32_t* resv;           // [nodes] first list position of today's round-1 exposures (-1: appended)
  uint32_t* ticket;   // [4] last-block-done counters
  int64_t capacity; int32_t n_nodes, n_sims, K, nt, n_spans, tpw;
  int32_t tpw_sh;                  // tpw = 1 << tpw_sh (1, 2, 4, 8 or 16 tiles per span)
};

__device__ __forceinline__ bool sim_stopped(const Dev& d, int sim, int t) { return d.stop_day[sim] < t; }

// The current day's record at day + 4: {t, day_flags[t], ri_cnt[t], ri_cnt[t + 1]}, so
// a kernel reads its day, flags and RI counts with one 16-byte load instead of a
// load that depends on the day counter.  Written where the day is set (k_set_day)
// and where it advances (k_day_end).
__device__ __forceinline__ int4 today(const Dev& d) { return *reinterpret_cast<const int4*>(d.day + 4); }
__device__ __forceinline__ int4 day_record(const Dev& d, int32_t t) {
  const int32_t f = t < d.nt ? d.day_flags[t] : 0;
  const int32_t r0 = t <= d.nt ? d.ri_cnt[t] : 0;          // ri_cnt has nt + 1 entries
  const int32_t r1 = t + 1 <= d.nt ? d.ri_cnt[t + 1] : r0;
  return make_int4(t, f, r0, r1);
}
__device__ __forceinline__ void set_today(const Dev& d, const int4& rec) {
  d.day[0] = rec.x;
  *reinterpret_cast<int4*>(d.day + 4) = rec;
}

__device__ __forceinline__ int32_t stage_day_flags(int32_t flags, int mask) {
  int32_t f = 0;
  if (mask & LP_STAGE_VITAL) f |= flags & (LP_DAY_VD_STEP | LP_DAY_VD_COPY);
  if (mask & LP_STAGE_DISEASE) f |= flags & (LP_DAY_DS | LP_DAY_STOPCHK);
  if (mask & LP_STAGE_RI) f |= flags & LP_DAY_RI;
  if (mask & LP_STAGE_SIA) f |= flags & LP_DAY_SIA;
  if (mask & (LP_STAGE_TRANSMISSION | LP_STAGE_LOG)) f |= flags & LP_DAY_TX;
  return f;
}

// ---------------------------------------------------------------------------
// day[0] = current tick; day[1] = end of the current lp_run_days call (0 outside one).
__global__ void k_set_day(Dev d, int32_t t, int32_t until) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    set_today(d, day_record(d, t));
    d.day[1] = until;
  }
}

// Kaplan-Meier style age at death (days) for an agent of `age` days, from a
// cumulative deaths table by year of age: year of death drawn in proportion to
// deaths among years >= current year, day uniform within that year (> age).
__device__ int32_t km_age_at_death(const int64_t* cd, uint64_t seed, uint64_t slot, int32_t t, int32_t age) {
  const uint4 o = draw(seed, slot, t, ST_KM);
  const int32_t max_year = LP_KM_LEN - 1;
  int32_t ay = age / 365;
  if (ay > max_year) ay = max_year;
  const int64_t lo = ay == 0 ? 0 : cd[ay - 1];
  const int64_t hi = cd[max_year];
  int32_t yod = ay;
  if (hi > lo) {
    const uint64_t u = (uint64_t)o.x | ((uint64_t)o.y << 32);
    const int64_t v = lo + (int64_t)mulhi64(u, (uint64_t)(hi - lo));
    int32_t a = 0, b = LP_KM_LEN;          // first k with cd[k] > v
    while (a < b) { const int32_t m = (a + b) >> 1; if (cd[m] > v) b = m; else a = m + 1; }
    yod = a;
  }
  int32_t lo_d = yod * 365;
  if (lo_d < age + 1) lo_d = age + 1;
  const int32_t hi_d = (yod + 1) * 365;
  if (hi_d <= lo_d) return lo_d;
  return lo_d + (int32_t)(((uint64_t)o.z * (uint64_t)(hi_d - lo_d)) >> 32);
}

// ---------------------------------------------------------------------------
// Per node: births, per-day resets (model.py:1711-1719), in two parts so that a
// caller can issue the loads (day_begin_load) ahead of unrelated stores of its own:
// the compiler cannot move them past stores through other pointers of Dev.
struct DayBeginIn {
  int32_t flags, stopped, pop_prev, lnode;
  uint64_t rate, seed;
  int64_t live_end, seg_end;
};
__device__ __forceinline__ DayBeginIn day_begin_load(const Dev& d, int n, int32_t t, int mask) {
  DayBeginIn in;
  const int sim = d.node_sim[n];
  const lp_sim_params& sp = d.sims[sim];
  in.flags = stage_day_flags(d.day_flags[t], mask);
  in.stopped = sim_stopped(d, sim, t) ? 1 : 0;
  in.pop_prev = t >= 1 ? d.r_pop[(size_t)(t - 1) * d.n_nodes + n] : 0;
  in.lnode = n - sp.node_begin;
  in.rate = d.birth_rate_step[n];
  in.seed = sp.seed;
  in.live_end = d.live_end[n];
  in.seg_end = d.seg_end[n];
  return in;
}
__device__ __forceinline__ void day_begin_apply(const Dev& d, int n, int32_t t, const DayBeginIn& in) {
  for (int k = 0; k < LP_MAX_STRAINS; ++k) d.newexp[n * LP_MAX_STRAINS + k] = 0;
  d.n_uniq[n] = 0; d.size[n] = 0; d.take_all[n] = 0; d.m_round[n] = 0;
  d.births[n] = 0;
  d.nb_begin[n] = in.live_end;
  d.ei_cnt[(size_t)((t + 1) & 1) * d.n_nodes + n] = 0;
  if (in.stopped) return;
  if ((in.flags & LP_DAY_VD_STEP) && t >= 1) {
    // expected = step * rate * pop[t-1] in float64 as the reference computes it
    // (model.py:1712-1716), reproduced exactly in integer arithmetic
    uint64_t thr = 0;
    int64_t b = in.pop_prev > 0 ? f64_mul_floor(in.rate, (uint32_t)in.pop_prev, &thr) : 0;
    d.birth_dbg[2 * n] = b;
    d.birth_dbg[2 * n + 1] = (int64_t)thr;
    if (thr) {
      const uint4 o = philox(0u, (uint32_t)in.lnode, (uint32_t)t, ST_BIRTH, in.seed);
      b += (uint64_t)o.x < thr ? 1 : 0;
    }
    const int64_t room = in.seg_end - in.live_end;
    if (b > room) { b = room; atomicOr(d.err, ERR_CAPACITY); }
    d.births[n] = (int32_t)b;
    d.live_end[n] = in.live_end + b;
  }
}
__device__ void day_begin_node(const Dev& d, int n, int32_t t, int mask) {
  day_begin_apply(d, n, t, day_begin_load(d, n, t, mask));
}

// sim_cases[sim]: 1 when a node of the sim logged E or I on the previous day (set by
// k_day_end), 0 when none did; CASES_CHECKED when today's early-stop rule has run
// already (k_day_begin, which sums the rows).  k_node_reduce runs the rule for
// days begun inside lp_run_days and clears the flag for today's log.
constexpr int32_t CASES_CHECKED = 2;

// Per sim: the early-stop rule (model.py:788-795) on the previous day's E and I;
// `cases` < 0: sum the previous day's rows, else whether any was positive.
__device__ void day_stop_check(const Dev& d, int sim, int32_t t, int mask, int cases) {
  const lp_sim_params& sp = d.sims[sim];
  const int32_t flags = stage_day_flags(d.day_flags[t], mask);
  if ((flags & LP_DAY_STOPCHK) && t >= 1 && d.stop_day[sim] == INT_MAX) {
    int64_t any = cases;
    if (cases < 0) {
      any = 0;
      const size_t row = (size_t)(t - 1) * d.n_nodes;
      for (int j = sp.node_begin; j < sp.node_begin + sp.n_nodes; ++j) any += d.r_E[row + j] + d.r_I[row + j];
    }
    if (any == 0 && !(sp.last_seed_day > t)) d.stop_day[sim] = t;
  }
}

// Per node: births, per-day resets, early-stop rule.  Inside lp_run_days only the
// first day launches it: k_day_end of each day begins the next one.
__global__ void k_day_begin(Dev d, int mask) {
  const int32_t t = d.day[0];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n == 0) d.sel_count[0] = 0;
  if (n >= d.n_nodes) return;
  day_begin_node(d, n, t, mask);
  const int sim = d.node_sim[n];
  if (n == d.sims[sim].node_begin) {
    day_stop_check(d, sim, t, mask, -1);
    d.sim_cases[sim] = CASES_CHECKED;
  }
}

// ---------------------------------------------------------------------------
// The agent pass.  One workgroup per span (up to U tiles of 1024 slots of one
// node).  Thread tid owns slots [tile*1024 + 4*tid, +4) of each tile, so each
// wave covers one 256-slot sampling chunk per tile.
//   phase A  the state words of all U tiles, then the day's campaign / death
//            columns (dod on VD step days; ri_timer + chronically_missed on RI
//            days; dob on SIA days; risk only when recounting)
//   phase B  births and deaths                          model.py:1719-1732, 1775-1777
//   phase C  one vector load per E/I column for every quad holding an E or I agent
//   phase D  disease progression, RI, SIA and the tallies
// The susceptible statistics the transmission step needs (per node: #S, #S with
// positive weight, #R; per span and per chunk: the fixed-point weight sum of
// S agents) are exact integers carried from day to day.  Every change of an
// agent into or out of S (birth, death, RI, SIA, transmission) or into or out
// of R applies its exact delta, so only a RECOUNT pass reads every agent's risk.
// Rare tallies go through LDS atomics; one partial per span, no global atomics.
enum : int { L_E0 = 0, L_I0 = 4, L_SIAPK0 = 8, L_POTP = 12, L_P, L_DEATHS, L_DPP, L_DP, L_NEWPP, L_NEWP,
             L_RIV, L_RIP, L_IPV, L_SIAV, L_SIAP, L_N };

__device__ __forceinline__ int8_t byte_of(uint32_t w, int a) { return (int8_t)((w >> (8 * a)) & 0xffu); }
__device__ __forceinline__ int32_t lane_of(const int4& v, int a) { return a == 0 ? v.x : a == 1 ? v.y : a == 2 ? v.z : v.w; }
__device__ __forceinline__ float lane_of(const float4& v, int a) { return a == 0 ? v.x : a == 1 ? v.y : a == 2 ? v.z : v.w; }

// Calls f(std::integral_constant<int, u>) for u = 0..U-1 with u a compile-time constant.
template <int U, int u = 0, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (u < U) {
    f(std::integral_constant<int, u>{});
    static_for<U, u + 1>(f);
  }
}

// Day kinds of the agent pass:
//   PLAIN    no births, deaths, RI or SIA (about 86% of days): k_plain reads the
//            state column and the E/I columns of quads holding an E or I agent
//   FULL     VD step / RI / SIA days: k_agents adds the campaign and death columns
//   RECOUNT  the first pass and any pass after the host changed bound buffers:
//            k_agents recounts every carried statistic from the agent columns
// FULL days use kinds 1..7 (FULL_* bits); KIND_PLAIN_LIST: a plain day from the E / I list (k_ei);
// KIND_FULL_LIST + F - 1: FULL day F with DiseaseState from the list (k_ei, then k_full<F, false>)
enum : int { KIND_PLAIN = 0, KIND_FULL = 1, KIND_RECOUNT = 8, KIND_PLAIN_LIST = 9, KIND_FULL_LIST = 10,
             N_KINDS = 17 };

__device__ __forceinline__ bool is_plain_day(int32_t flags) {
  return !(flags & (LP_DAY_VD_STEP | LP_DAY_RI | LP_DAY_SIA));
}

// True if any byte of the packed state word is 1 (E) or 2 (I): the classic
// zero-byte test (v - 0x01..) & ~v & 0x80.. on w ^ 0x01.. and w ^ 0x02.. (exact
// for "any byte": a borrow only propagates out of a byte that is itself zero).
__device__ __forceinline__ bool has_ei_byte(uint32_t w) {
  const uint32_t x1 = w ^ 0x01010101u, x2 = w ^ 0x02020202u;
  return ((((x1 - 0x01010101u) & ~x1) | ((x2 - 0x01010101u) & ~x2)) & 0x80808080u) != 0;
}

// True if any byte of the packed state word equals v (same zero-byte test).
__device__ __forceinline__ bool has_byte(uint32_t w, uint32_t v) {
  const uint32_t x = w ^ (v * 0x01010101u);
  return (((x - 0x01010101u) & ~x) & 0x80808080u) != 0;
}

// Position of the j-th (0-based) set bit of m; m has more than j set bits.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int j) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int c = __popcll(m & ((1ull << w) - 1));
    if (j >= c) { j -= c; m >>= w; pos += w; }
  }
  return pos;
}

// Disease progression of one E or I agent (disease_state_step_kernel,
// model.py:419-452) on the quad's packed columns: a = agent in the quad, s its
// state (1 or 2, updated), e2 / i2 / p2 the quad's exposure / infection /
// paralysis timers, ppw its potentially_paralyzed bytes.  Returns true on I -> R.
__device__ __forceinline__ bool progress_agent(int a, int8_t& s, int strain, uint2& e2, uint2& i2, uint2& p2,
                                               uint32_t& ppw, uint32_t ipvw, uint64_t seed, uint64_t slot, int32_t t,
                                               uint64_t thr_paralysis, int8_t* paralyzed_i, int32_t* lc_newpp,
                                               int32_t* lc_newp) {
  const int sh16 = 16 * (a & 1);
  const uint32_t m16 = 0xffffu << sh16;
  bool to_r = false;
  if (s == 1) {
    const uint32_t ew = a < 2 ? e2.x : e2.y;
    const uint16_t et = (uint16_t)((ew >> sh16) & 0xffffu);
    if (et == 0) s = 2;
    else {
      const uint32_t nw = (ew & ~m16) | ((uint32_t)(et - 1) << sh16);
      if (a < 2) e2.x = nw; else e2.y = nw;
    }
  }
  if (s == 2) {
    const uint32_t iw = a < 2 ? i2.x : i2.y;
    const uint16_t it = (uint16_t)((iw >> sh16) & 0xffffu);
    if (it == 0) { s = 3; to_r = true; }
    else {
      const uint32_t nw = (iw & ~m16) | ((uint32_t)(it - 1) << sh16);
      if (a < 2) i2.x = nw; else i2.y = nw;
    }
    if (strain == 0) {
      const uint32_t pw = a < 2 ? p2.x : p2.y;
      const uint16_t pt = (uint16_t)((pw >> sh16) & 0xffffu);
      if (pt == 0 && byte_of(ppw, a) == -1) {
        const uint32_t bm = 0xffu << (8 * a);
        if (byte_of(ipvw, a) == 0) {
          ppw = (ppw & ~bm) | (1u << (8 * a));
          atomicAdd(lc_newpp, 1);
          const uint4 o = draw(seed, slot, t, ST_PARALYSIS);
          if ((uint64_t)o.x < thr_paralysis) { *paralyzed_i = 1; atomicAdd(lc_newp, 1); }
        } else {
          ppw = ppw & ~bm;
        }
      }
      if (pt > 0) {
        const uint32_t nw = (pw & ~m16) | ((uint32_t)(pt - 1) << sh16);
        if (a < 2) p2.x = nw; else p2.y = nw;
      }
    }
  }
  return to_r;
}

// Agent record (Dev::rec, 16 B per slot): the mutable fields of DiseaseState
// and the E / I tally in one aligned line segment, so that an E or I agent costs
// one 16-byte access instead of one cache line from each of seven columns.
//   x: exposure_timer (bits 0-15) | infection_timer (16-31)
//   y: paralysis_timer (0-15) | strain (16-23) | potentially_paralyzed (24-31)
//   z: daily_infectivity (float bits)
//   w: ipv_protected (0-7) | paralyzed (8-15) | unused (16-31; the disease state
//      lives in its column and, for E and I agents, in their list entries)
// k_pack fills the records from the bound columns after a RECOUNT pass and
// k_unpack writes them back (lp_sync); in between the records are the state of
// these fields.  Single fields are written with byte stores (RB_*), so writers
// of different fields of one record never overwrite each other.
enum : int { RB_STRAIN = 6, RB_PP = 7, RB_IPV = 12, RB_PARALYZED = 13 };

__device__ __forceinline__ int8_t* rec_byte(const Dev& d, int64_t i, int b) {
  return reinterpret_cast<int8_t*>(d.rec + i) + b;
}

// DiseaseState step of one E or I agent on its record (model.py:419-452): s its
// state (1 or 2, updated), r its record (x and y updated).  Sets `to_pp` when the
// agent becomes potentially paralysed today and `to_par` when it is paralysed
// (the caller counts both and stores the paralyzed byte).  Returns true on I -> R.
__device__ __forceinline__ bool progress_rec(int8_t& s, uint4& r, uint64_t seed, uint64_t slot, int32_t t,
                                             uint64_t thr_paralysis, bool& to_pp, bool& to_par) {
  bool to_r = false;
  if (s == 1) {
    const uint32_t et = r.x & 0xffffu;
    if (et == 0) s = 2;
    else r.x = (r.x & 0xffff0000u) | (et - 1);
  }
  if (s == 2) {
    const uint32_t it = r.x >> 16;
    if (it == 0) { s = 3; to_r = true; }
    else r.x = (r.x & 0xffffu) | ((it - 1) << 16);
    if (((r.y >> 16) & 0xffu) == 0) {                  // strain 0
      const uint32_t pt = r.y & 0xffffu;
      if (pt == 0 && (int8_t)(r.y >> 24) == -1) {
        if ((r.w & 0xffu) == 0) {                      // not ipv_protected
          r.y = (r.y & 0x00ffffffu) | (1u << 24);
          to_pp = true;
          const uint4 o = draw(seed, slot, t, ST_PARALYSIS);
          if ((uint64_t)o.x < thr_paralysis) to_par = true;
        } else {
          r.y = r.y & 0x00ffffffu;
        }
      }
      if (pt > 0) r.y = (r.y & 0xffff0000u) | (pt - 1);
    }
  }
  return to_r;
}

// Per-simulation constants of the E / I stage, copied into registers once per
// workgroup (read through `sp` they would be reloaded after every record store).
struct EIConst {
  uint64_t seed, thr_paralysis;
  int64_t slot_base;
  float r0k[LP_MAX_STRAINS];
};

__device__ __forceinline__ EIConst ei_const(const lp_sim_params& sp) {
  EIConst c;
  c.seed = sp.seed;
  c.thr_paralysis = sp.thr_paralysis;
  c.slot_base = sp.slot_base;
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) c.r0k[k] = sp.strain_r0[k];
  return c;
}

// ri_timer <- actual values: for live agents that are not chronically missed,
// minus RI's decrements from the reference tick through the latest applied RI
// step (one workgroup per span); k_set_ri_ref then moves the reference there.
__global__ void __launch_bounds__(WG) k_ri_sync(Dev d) {
  const int4 span = d.spans[blockIdx.x];
  const int32_t dc = d.ri_cnt[d.ri_ref[1] + 1] - d.ri_cnt[d.ri_ref[0] + 1];
  const int32_t shift = d.sims[span.w].ri_step * dc;
  if (shift == 0) return;
  for (int u = 0; u < span.z; ++u) {
    const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)threadIdx.x * 4;
    const uint32_t w = reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2];
    if ((w & 0x80808080u) == 0x80808080u) continue;   // no live agent in the quad
    const uint32_t mw = reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2];
    const uint2 r = *reinterpret_cast<const uint2*>(d.ri_timer + i0);
    uint2 n = r;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (byte_of(w, a) < 0 || byte_of(mw, a) == 1) continue;
      const int sh16 = 16 * (a & 1);
      uint32_t& h = a < 2 ? n.x : n.y;
      const int16_t v = (int16_t)((int16_t)((h >> sh16) & 0xffffu) - shift);
      h = (h & ~(0xffffu << sh16)) | ((uint32_t)(uint16_t)v << sh16);
    }
    if (n.x != r.x || n.y != r.y) *reinterpret_cast<uint2*>(d.ri_timer + i0) = n;
  }
}

// through >= -1: ri_timer is actual as of tick `through` (after a RECOUNT pass);
// through < -1: as of the latest applied RI step (after k_ri_sync).
__global__ void k_set_ri_ref(Dev d, int32_t through) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int32_t v = through >= -1 ? through : d.ri_ref[1];
    d.ri_ref[0] = v;
    d.ri_ref[1] = v;
  }
}

// ---------------------------------------------------------------------------
// E/I lists.  ei_list[b] holds, per node n, entries of the node's E and I
// agents at ei_list + b * capacity + seg_begin[n] (room for the whole segment),
// ei_cnt[b * n_nodes + n] of them, in no particular order.  Day t reads list
// t & 1 (the agents that are E or I when the day starts) and every pass of the
// day appends to list (t + 1) & 1: the agent passes their agents that are E or I
// afterwards (survivors of the progression, RI / SIA exposures), the sampler the
// agents it exposes.  Plain days then process the list alone (k_ei) instead of
// scanning every agent; the order of a list does not affect any result.

// E/I list entry (16 bytes).  While the lists carry the disease timers, an E or
// I agent's countdowns are kept as absolute days (the day its timer reads 0 at the
// start of DiseaseState) and its record keeps the timers of when it entered the
// list (an I agent's exposure timer is actual: 0 after E -> I, as drawn for one
// infected from the start); ei_actual_timers writes the actual timers back.  An E
// entry also holds what E -> I needs (none of it changes while S or E), so the
// list pass reads a record only for the paralysis decision (ipv_protected).
//   x: slot
//   y: bits 0-1 state (1 E, 2 I); 2 strain 0 (paralysis timer tracked); 3 paralysis
//      decision pending (potentially_paralyzed -1); 4-5 strain; 8-31 E: day the
//      exposure timer reads 0, I: day the paralysis timer reads 0 (strain 0)
//   z: E: infection timer | paralysis timer << 16; I: day the infection timer
//      reads 0
//   w: infectious weight rint(infectivity * strain r0 * 2^24); W_REC when it does
//      not fit 32 bits (then recomputed from the record)
constexpr uint32_t W_REC = 0xFFFFFFFFu;
constexpr uint32_t DAY_NONE = 0xFFFFFFu;

__device__ __forceinline__ int ent_state(const uint4& e) { return (int)(e.y & 3u); }
__device__ __forceinline__ int ent_strain(const uint4& e) { return (int)((e.y >> 4) & 3u); }
__device__ __forceinline__ bool ent_s0(const uint4& e) { return (e.y >> 2) & 1u; }
__device__ __forceinline__ bool ent_pend(const uint4& e) { return (e.y >> 3) & 1u; }
__device__ __forceinline__ int32_t ent_day(const uint4& e) { return (int32_t)(e.y >> 8); }
// The next day the entry's agent changes: E -> I, the paralysis decision, I -> R.
__device__ __forceinline__ int32_t ent_event(const uint4& e) {
  const int32_t d8 = ent_day(e);
  if (ent_state(e) == 1) return d8;
  return ent_pend(e) && d8 < (int32_t)e.z ? d8 : (int32_t)e.z;
}
__device__ __forceinline__ uint32_t ent_flags(int s, int k, bool pend, int32_t day) {
  return (uint32_t)s | (k == 0 ? 4u : 0u) | (pend ? 8u : 0u) | ((uint32_t)k << 4) | ((uint32_t)day << 8);
}
__device__ __forceinline__ uint32_t weight_entry(int64_t w) { return w >= (int64_t)W_REC ? W_REC : (uint32_t)w; }
__device__ __forceinline__ float r0_sel(const float* r0k, int k) {
  return k == 0 ? r0k[0] : k == 1 ? r0k[1] : k == 2 ? r0k[2] : r0k[3];
}

// Entry of an E agent with timers `et`, `it`, `pt`, potentially_paralyzed `pp` at the
// start of day t_next and infectious weight `weight` once I.
__device__ __forceinline__ uint4 entry_e(uint32_t slot, int k, uint32_t et, uint32_t it, uint32_t pt, int8_t pp,
                                         int32_t t_next, int64_t weight) {
  return make_uint4(slot, ent_flags(1, k, k == 0 && pp == -1, t_next + (int32_t)et), it | (pt << 16),
                    weight_entry(weight));
}

// Entry of an I agent with timers `it`, `pt` and potentially_paralyzed `pp` at the start
// of day t_next.
__device__ __forceinline__ uint4 entry_i(uint32_t slot, int k, uint32_t it, uint32_t pt, int8_t pp, int32_t t_next,
                                         int64_t weight) {
  const int32_t p_day = k == 0 ? t_next + (int32_t)pt : (int32_t)DAY_NONE;
  return make_uint4(slot, ent_flags(2, k, k == 0 && pp == -1, p_day), (uint32_t)t_next + it, weight_entry(weight));
}

// Entry of an E or I agent from its (actual) record at the start of day t_next.
__device__ __forceinline__ uint4 entry_from_rec(uint32_t slot, int s, const uint4& r, int32_t t_next, const float* r0k) {
  const int k = (int)((r.y >> 16) & (LP_MAX_STRAINS - 1));
  const int64_t w = weight_q24(__uint_as_float(r.z) * r0_sel(r0k, k));
  if (s == 1) return entry_e(slot, k, r.x & 0xFFFFu, r.x >> 16, r.y & 0xFFFFu, (int8_t)(r.y >> 24), t_next, w);
  return entry_i(slot, k, r.x >> 16, r.y & 0xFFFFu, (int8_t)(r.y >> 24), t_next, w);
}

// A timer that reads 0 on day `end`, as of the start of day t_next (it stays at 0).
__device__ __forceinline__ uint32_t timer_left(int32_t end, int32_t t_next) {
  const int32_t v = end - t_next;
  return v < 0 ? 0u : (uint32_t)v;
}

// The record's actual timers of an E or I agent at the start of day t_next.
__device__ __forceinline__ void ei_actual_timers(const Dev& d, const uint4& e, int32_t t_next) {
  uint16_t* hw = reinterpret_cast<uint16_t*>(d.rec + e.x);
  if (ent_state(e) == 1) {
    hw[0] = (uint16_t)timer_left(ent_day(e), t_next);           // exposure timer
  } else {
    hw[1] = (uint16_t)timer_left((int32_t)e.z, t_next);         // infection timer
    if (ent_s0(e)) hw[2] = (uint16_t)timer_left(ent_day(e), t_next);   // paralysis timer
  }
}

// Pieces of EI_PIECE list entries: one wavefront of k_ei takes one piece at a time.
constexpr int EI_PIECE = 256;
constexpr int EI_LDS_NODES = 4095;   // k_ei stages the piece offsets in LDS up to this many nodes

__device__ __forceinline__ uint4* ei_list_of(const Dev& d, int b, int node) {
  return d.ei_list + (size_t)b * (size_t)d.capacity + d.seg_begin[node];
}

// Appends one entry (one atomic, returning); for the rare paths (inline DiseaseState,
// RECOUNT) that append from divergent code.
__device__ __forceinline__ void append_entry(const Dev& d, int b, int node, const uint4& e) {
  const int pos = atomicAdd(d.ei_cnt + (size_t)b * d.n_nodes + node, 1);
  ei_list_of(d, b, node)[pos] = e;
}

// Appends slots i0 + a (bit a of `bits`) of every lane to node `node`'s list b:
// one atomic per wave, ranks from per-position ballots.  Called by all 64 lanes
// of the wave together.
__device__ __forceinline__ void wave_append_quad_to(uint32_t* L, int32_t* cnt, int64_t i0, uint32_t bits, int lane) {
  const uint64_t lt = (1ull << lane) - 1ull;
  uint64_t mk[4];
  int tot = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) { mk[a] = __ballot((bits >> a) & 1u); tot += __popcll(mk[a]); }
  if (tot == 0) return;
  int base = 0;
  if (lane == 0) base = atomicAdd(cnt, tot);
  base = __shfl(base, 0, 64);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if ((bits >> a) & 1u) L[base + __popcll(mk[a] & lt)] = (uint32_t)(i0 + a);
    base += __popcll(mk[a]);
  }
}


// One compacted quad holding E or I agents: loads the quad's four records (one
// 64-byte segment, four loads in flight together), runs the disease progression
// of its E and I agents (model.py:419-452) and tallies them as E / I.  `wei`:
// state bytes that identify today's E and I agents; `w`: the quad's state word
// to update and write back if it differs from `w_orig`.  The agents still E or I
// afterwards join node `node`'s list of tomorrow; returns them (bit a).
__device__ __forceinline__ uint32_t ei_quad(const Dev& d, int64_t i0, int64_t live_hi, uint32_t wei, uint32_t w,
                                        uint32_t w_orig, bool do_ds, bool do_tally, const EIConst& ec,
                                        int32_t t, int32_t* lc, unsigned long long* lib, int32_t& c_r, int node) {
  const uint4* rq = d.rec + i0;
  uint4 r[4] = {rq[0], rq[1], rq[2], rq[3]};
  uint32_t surv = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + a;
    int8_t sa = byte_of(wei, a);
    if ((sa != 1 && sa != 2) || i >= live_hi) continue;
    const uint4 r0 = r[a];
    const int k = (r0.y >> 16) & (LP_MAX_STRAINS - 1);
    bool to_pp = false, to_par = false;
    if (do_ds && progress_rec(sa, r[a], ec.seed, (uint64_t)(i - ec.slot_base), t, ec.thr_paralysis, to_pp, to_par))
      c_r += 1;
    if (to_pp) atomicAdd(&lc[L_NEWPP], 1);
    if (to_par) atomicAdd(&lc[L_NEWP], 1);
    if (do_tally) {                                     // tx_step_prep / count_SEIRP tally of E and I
      if (sa == 1) {
        atomicAdd(&lc[L_E0 + k], 1);
      } else if (sa == 2) {
        atomicAdd(&lc[L_I0 + k], 1);
        const float r0k = k == 0 ? ec.r0k[0] : k == 1 ? ec.r0k[1] : k == 2 ? ec.r0k[2] : ec.r0k[3];
        atomicAdd(&lib[k], (unsigned long long)weight_q24(__uint_as_float(r0.z) * r0k));
      }
    }
    uint32_t* rw = reinterpret_cast<uint32_t*>(d.rec + i);
    if (r[a].x != r0.x) rw[0] = r[a].x;
    if (r[a].y != r0.y) rw[1] = r[a].y;
    if (to_par) *rec_byte(d, i, RB_PARALYZED) = 1;
    w = (w & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)sa << (8 * a));
    if (sa == 1 || sa == 2) {                            // into tomorrow's list, timers as of tomorrow
      surv |= 1u << a;
      append_entry(d, (t + 1) & 1, node, entry_from_rec((uint32_t)i, sa, r[a], t + 1, ec.r0k));
    }
  }
  if (w != w_orig) reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = w;
  return surv;
}

// Remove an exposed agent's weight w from its chunk and from span s.
__device__ __forceinline__ void unweight_in(const Dev& d, int64_t s, int64_t slot, int64_t w) {
  const unsigned long long dw = (unsigned long long)(-w);
  atomicAdd(reinterpret_cast<unsigned long long*>(d.chunk_sum + (slot >> 8)), dw);
  atomicAdd(reinterpret_cast<unsigned long long*>(d.partials + (size_t)s * NPART + P_SW), dw);
}

// The span of node `node` holding `slot`.
__device__ __forceinline__ int64_t span_of(const Dev& d, int node, int64_t slot) {
  const int64_t tile = slot / LP_TILE, first = d.seg_begin[node] / LP_TILE;
  return d.node_span_begin[node] + ((tile - first) >> d.tpw_sh);   // tile >= first
}

// RI candidate lists: ri_list[b] holds, per node, the slots of live agents that are
// not chronically missed and whose RI window can still open (at ri_list + b *
// capacity + seg_begin[n], ri_lcnt[b * n_nodes + n] of them).  Buffer b = (RI days
// before the next RI day) & 1: births append to it, the next RI day reads it and
// keeps its candidates in the other buffer; k_node_reduce empties the buffer read.
__device__ __forceinline__ uint32_t* ri_list_of(const Dev& d, int b, int node) {
  return d.ri_list + (size_t)b * (size_t)d.capacity + d.seg_begin[node];
}
__device__ __forceinline__ int32_t* ri_cnt_of(const Dev& d, int b, int node) {
  return d.ri_lcnt + (size_t)b * d.n_nodes + node;
}

// Appends entry `e` of each active lane with `flag` to node `node`'s list b (every
// active lane passes the same node): one atomic per call.
__device__ __forceinline__ void wave_append_entry(const Dev& d, int b, int node, const uint4& e, bool flag, int lane) {
  const uint64_t mk = __ballot(flag);
  if (mk == 0) return;
  const int first = __ffsll((unsigned long long)mk) - 1;
  const int rank = __popcll(mk & ((1ull << lane) - 1ull));
  int base = 0;
  if (lane == first) base = atomicAdd(d.ei_cnt + (size_t)b * d.n_nodes + node, __popcll(mk));
  base = __shfl(base, first, 64);
  if (flag) ei_list_of(d, b, node)[base + rank] = e;
}

// The entry of an agent exposed on day t with strain k, whose r0 scalar is rk (its
// timers, potentially_paralyzed and infectivity are the record's, untouched while
// it was S).
__device__ __forceinline__ uint4 exposed_entry_rec(int64_t slot, const uint4& r, int k, int32_t t, float rk) {
  return entry_e((uint32_t)slot, k, r.x & 0xFFFFu, r.x >> 16, r.y & 0xFFFFu, (int8_t)(r.y >> 24), t + 1,
                 weight_q24(__uint_as_float(r.z) * rk));
}
__device__ __forceinline__ uint4 exposed_entry(const Dev& d, int64_t slot, int k, int32_t t, float rk) {
  return exposed_entry_rec(slot, d.rec[slot], k, t, rk);
}

// NodeInfo of every node (one thread per node).
__global__ void k_node_info(Dev d) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.n_nodes) return;
  const int sim = d.node_sim[n];
  const lp_sim_params& sp = d.sims[sim];
  NodeInfo ni;
  ni.seed = sp.seed;
  ni.thr_paralysis = sp.thr_paralysis;
  ni.slot_base = sp.slot_base;
  ni.seg_begin = d.seg_begin[n];
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) ni.r0k[k] = sp.strain_r0[k];
  ni.sim = sim;
  ni.lnode = n - sp.node_begin;
  ni.span_begin = d.node_span_begin[n];
  ni.n_spans = d.node_span_begin[n + 1] - ni.span_begin;
  ni.net_off = sp.net_off;
  ni.thr_zi = sp.thr_zi;
  ni.zi = sp.zi;
  ni.disp_r = sp.disp_r;
  ni.sim_nodes = sp.n_nodes;
  ni.sim_node_begin = sp.node_begin;
  d.ninfo[n] = ni;
}

// Agent records <- bound columns, slots [0, capacity) (one quad per thread).
__global__ void __launch_bounds__(WG) k_pack(Dev d) {
  const int64_t nq = (d.capacity + 3) >> 2;
  for (int64_t q = (int64_t)blockIdx.x * WG + threadIdx.x; q < nq; q += (int64_t)gridDim.x * WG) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = q * 4 + a;
      if (i >= d.capacity) break;
      d.rec[i] = make_uint4((uint32_t)d.exp_timer[i] | ((uint32_t)d.inf_timer[i] << 16),
                            (uint32_t)d.par_timer[i] | ((uint32_t)(uint8_t)d.strain[i] << 16) |
                                ((uint32_t)(uint8_t)d.pp[i] << 24),
                            __float_as_uint(d.infect[i]),
                            (uint32_t)(uint8_t)d.ipv[i] | ((uint32_t)(uint8_t)d.paralyzed[i] << 8));
    }
  }
}

// Bound columns <- agent records (daily_infectivity is never written, so not copied back).
__global__ void __launch_bounds__(WG) k_unpack(Dev d) {
  const int64_t nq = (d.capacity + 3) >> 2;
  for (int64_t q = (int64_t)blockIdx.x * WG + threadIdx.x; q < nq; q += (int64_t)gridDim.x * WG) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = q * 4 + a;
      if (i >= d.capacity) break;
      const uint4 r = d.rec[i];
      d.exp_timer[i] = (uint16_t)(r.x & 0xffffu);
      d.inf_timer[i] = (uint16_t)(r.x >> 16);
      d.par_timer[i] = (uint16_t)(r.y & 0xffffu);
      d.strain[i] = (int8_t)(r.y >> 16);
      d.pp[i] = (int8_t)(r.y >> 24);
      d.ipv[i] = (int8_t)(r.w & 0xffu);
      d.paralyzed[i] = (int8_t)(r.w >> 8);
    }
  }
}

// Writes the span partial: carried-count deltas from red[][], tallies from LDS.
__device__ __forceinline__ void write_partial(const Dev& d, int tid, const int64_t (*red)[4], const int32_t* lc,
                                              const unsigned long long* lib, bool recount, bool sw_delta) {
  if (tid >= NPART) return;
  int64_t v = 0;
  if (tid == P_S || tid == P_SPOS || tid == P_R || tid == P_SW) {
    const int q = tid == P_S ? 0 : tid == P_SPOS ? 1 : tid == P_R ? 2 : 3;
#pragma unroll
    for (int w = 0; w < WG / 64; ++w) v += red[w][q];
  } else if (tid >= P_E0 && tid < P_E0 + 4) v = lc[L_E0 + tid - P_E0];
  else if (tid >= P_I0 && tid < P_I0 + 4) v = lc[L_I0 + tid - P_I0];
  else if (tid >= P_SIAPK0 && tid < P_SIAPK0 + 4) v = lc[L_SIAPK0 + tid - P_SIAPK0];
  else if (tid >= P_IB0 && tid < P_IB0 + 4) v = (int64_t)lib[tid - P_IB0];
  else {
    switch (tid) {
      case P_POTP: v = lc[L_POTP]; break;
      case P_P: v = lc[L_P]; break;
      case P_DEATHS: v = lc[L_DEATHS]; break;
      case P_DPP: v = lc[L_DPP]; break;
      case P_DP: v = lc[L_DP]; break;
      case P_NEWPP: v = lc[L_NEWPP]; break;
      case P_NEWP: v = lc[L_NEWP]; break;
      case P_RIV: v = lc[L_RIV]; break;
      case P_RIP: v = lc[L_RIP]; break;
      case P_IPV: v = lc[L_IPV]; break;
      case P_SIAV: v = lc[L_SIAV]; break;
      case P_SIAP: v = lc[L_SIAP]; break;
      default: break;
    }
  }
  int64_t* slot = d.partials + (size_t)blockIdx.x * NPART + tid;
  if (tid != P_SW) *slot = v;
  else if (recount) *slot = v;
  else if (sw_delta && v != 0)   // the carried weight: a delta, added without a dependent load
    atomicAdd(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)v);
}

// ---------------------------------------------------------------------------
// PLAIN days.  The workgroup loads the state words of all U tiles of its span at
// once, then every wave compacts its quads holding an E or I agent (ballot +
// prefix of popcounts) so that each lane takes one such quad per round: one
// round of E/I column loads serves up to 64 quads of any of the U tiles.
// S statistics do not change on these days; I -> R adds to the R delta.
template <int U>
__global__ void __launch_bounds__(WG) k_plain(Dev d, int mask) {
  const int4 span = d.spans[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ uint32_t s_st[U][WG];
  __shared__ int32_t lc[L_N];
  __shared__ unsigned long long lib[LP_MAX_STRAINS];
  __shared__ int64_t red[WG / 64][4];
  uint32_t st[U];
  static_for<U>([&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
    st[u] = (u < span.z) ? reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2] : 0xFFFFFFFFu;
  });
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int node = span.x, sim = span.w;
  const int64_t live_hi = d.live_end[node];
  const int32_t stop_day = d.stop_day[sim];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (stop_day < t) return;
  if (!is_plain_day(flags)) {
    if (tid < NPART && tid != P_SW) d.partials[(size_t)blockIdx.x * NPART + tid] = 0;
    if (tid == 0) atomicOr(d.err, ERR_KIND);
    return;
  }
  const bool do_ds = flags & LP_DAY_DS;
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const bool need_ei = do_ds || do_tally;
  const lp_sim_params& sp = d.sims[sim];
  if (tid < L_N) lc[tid] = 0;
  if (tid < LP_MAX_STRAINS) lib[tid] = 0;

  // quads holding an E or I agent, per tile, as wave masks (slots past the
  // node's live range hold -1, so no range test is needed)
  uint64_t m[U];
  int n_ei = 0;
  static_for<U>([&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const bool q = has_ei_byte(st[u]);
    m[u] = __ballot(q && need_ei);
    n_ei += __popcll(m[u]);
    s_st[u][tid] = st[u];
  });
  __syncthreads();

  int32_t c_r = 0;
  const EIConst ec = ei_const(sp);
  for (int base = 0; base < n_ei; base += 64) {
    int j = base + lane, su = -1, src = 0;
    static_for<U>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int c = __popcll(m[u]);
      if (su < 0) {
        if (j < c) { su = u; src = nth_set_bit(m[u], j); }
        else j -= c;
      }
    });
    if (su < 0) continue;
    const int qt = wave * 64 + src;
    const int64_t i0 = ((int64_t)span.y + su) * LP_TILE + (int64_t)qt * 4;
    const uint32_t w0 = s_st[su][qt];
    ei_quad(d, i0, live_hi, w0, w0, w0, do_ds, do_tally, ec, t, lc, lib, c_r, node);
  }
  {
    const int32_t a2 = wave_sum_i32(c_r);
    if (lane == 0) { red[wave][0] = 0; red[wave][1] = 0; red[wave][2] = a2; red[wave][3] = 0; }
  }
  __syncthreads();
  write_partial(d, tid, red, lc, lib, false, false);
}

// ---------------------------------------------------------------------------
// DiseaseState and the E / I tally from the E / I list: the agents in list
// t & 1 only (no other agent is E or I), survivors appended to list (t + 1) & 1.
// On plain days this is the whole agent pass; on FULL days it runs first and
// k_full<F, false> then does births, deaths, RI and SIA.  On VD step days the
// agents due to die today are skipped: the reference removes them before
// DiseaseState (model.py:226-241), so neither step sees them.  Each wavefront takes
// pieces of EI_PIECE entries of one node (ei_piece_off) in turn and adds the
// piece's tallies to node_acc with one atomic per field; k_node_reduce reads
// them in place of the span partials.  Cost: proportional to the E and I
// agents, not to the population.

// Exclusive prefix over n in [0, N) of cnt(n), by every thread of one workgroup:
// emit(n, prefix) for each n; returns the total.  Each pass takes 4 * WG nodes, 4
// consecutive ones per thread with their loads in flight together.
template <typename Cnt, typename Emit>
__device__ int64_t block_scan_nodes(int N, Cnt cnt, Emit emit) {
  __shared__ int64_t s_w[WG / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int64_t carry = 0;
  for (int base = 0; base < N; base += 4 * WG) {
    int64_t v[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int n = base + 4 * tid + a;
      v[a] = n < N ? cnt(n) : 0;
    }
    const int64_t sum = v[0] + v[1] + v[2] + v[3];
    const int64_t incl = wave_incl_scan_i64(sum, lane);
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    int64_t off = carry + incl - sum, tot = 0;
#pragma unroll
    for (int j = 0; j < WG / 64; ++j) {
      const int64_t x = s_w[j];
      off += j < w ? x : 0;
      tot += x;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int n = base + 4 * tid + a;
      if (n < N) emit(n, off);
      off += v[a];
    }
    carry += tot;
    __syncthreads();
  }
  return carry;
}

// Largest n in [0, N) with off[n] <= p (off[0] = 0 <= p): 64-ary search, all lanes.
__device__ __forceinline__ int find_piece_node(const int32_t* off, int N, int32_t p, int lane) {
  int lo = 0, hi = N;
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) / 64;
    const int c = lo + lane * step;
    const int32_t v = c < hi ? off[c] : INT_MAX;
    const uint64_t ok = __ballot(v <= p);
    lo += (63 - __clzll((unsigned long long)ok)) * step;
    hi = lo + step < hi ? lo + step : hi;
  }
  return lo;
}

__global__ void __launch_bounds__(WG) k_ei(Dev d, int mask) {
  // the piece offsets in LDS (when they fit) for the node lookup of each piece
  __shared__ int32_t s_off[EI_LDS_NODES + 1];
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int cur = t & 1, nxt = cur ^ 1;
  const int N = d.n_nodes, K = d.K;
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * (WG / 64) + (threadIdx.x >> 6), nw = gridDim.x * (WG / 64);
  const int32_t flags = stage_day_flags(dy.y, mask);
  const bool do_ds = flags & LP_DAY_DS;
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const bool vd = flags & LP_DAY_VD_STEP;   // agents due to die today leave before DiseaseState
  const bool lds = N <= EI_LDS_NODES;
  int32_t total;
  if (lds) {
    // every workgroup scans the nodes' piece counts of today's list itself
    const int32_t* cnt = d.ei_cnt + (size_t)cur * N;
    total = (int32_t)block_scan_nodes(
        N, [&](int n) { return (int64_t)((cnt[n] + EI_PIECE - 1) / EI_PIECE); },
        [&](int n, int64_t off) { s_off[n] = (int32_t)off; });
    __syncthreads();
  } else {
    total = d.ei_piece_off[N];   // planned by k_ei_plan
  }
  for (int32_t p = gw; p < total; p += nw) {
    int n;
    if (lds) {
      int a = 0, b = N;   // largest n with s_off[n] <= p
      while (b - a > 1) { const int m = (a + b) >> 1; if (s_off[m] <= p) a = m; else b = m; }
      n = a;
    } else {
      n = find_piece_node(d.ei_piece_off, N, p, lane);
    }
    n = __builtin_amdgcn_readfirstlane(n);   // the same in every lane: scalar loads of the node's data
    const NodeInfo ni = d.ninfo[n];
    const int32_t cnt = d.ei_cnt[(size_t)cur * N + n];
    const int32_t j0 = (p - (lds ? s_off[n] : d.ei_piece_off[n])) * EI_PIECE;
    const bool stopped = sim_stopped(d, ni.sim, t);
    EIConst ec;
    ec.seed = ni.seed;
    ec.thr_paralysis = ni.thr_paralysis;
    ec.slot_base = ni.slot_base;
#pragma unroll
    for (int k = 0; k < LP_MAX_STRAINS; ++k) ec.r0k[k] = ni.r0k[k];
    const uint4* L = d.ei_list + (size_t)cur * (size_t)d.capacity + ni.seg_begin;
    // four entries per lane; lanes past the end repeat entry j0 (unconditional loads)
    uint4 ent[4];
    bool v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t j = j0 + q * 64 + lane;
      v[q] = j < cnt;
      ent[q] = L[v[q] ? j : j0];
      if (v[q] && (int64_t)ent[q].x >= d.capacity) { atomicOr(d.err, ERR_LIST); v[q] = false; }
    }
    int32_t dd[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
    if (vd && !stopped) {
#pragma unroll
      for (int q = 0; q < 4; ++q) if (v[q]) dd[q] = d.dod[ent[q].x];
    }
    // tallies: E and I counts per strain in 16-bit fields, I->R / new potentially
    // paralysed / new paralysed in 16-bit fields, infectious weight per strain
    int64_t e_pk = 0, i_pk = 0, misc = 0, ib0 = 0, ib1 = 0, ib2 = 0, ib3 = 0;
    uint32_t surv = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint4 e = ent[q];
      int sa = ent_state(e);
      if (!v[q] || (sa != 1 && sa != 2)) continue;
      // a stopped simulation keeps its agents as they are: actual timers into the records
      if (stopped) { ei_actual_timers(d, e, t); continue; }
      const uint32_t slot = e.x;
      if (dd[q] <= t) { ei_actual_timers(d, e, t); continue; }   // dies today: the VD pass takes it
      const int k = ent_strain(e);
      if (do_ds && ent_event(e) <= t) {
        // today's transitions (model.py:419-452)
        uint16_t* hw = reinterpret_cast<uint16_t*>(d.rec + slot);
        int32_t i_end = (int32_t)e.z, p_day = ent_day(e);
        bool pend = ent_pend(e), to_par = false;
        const bool s0 = k == 0;
        const int s_in = sa;
        if (sa == 1) {                                    // E -> I: the exposure timer reads 0
          sa = 2;
          i_end = t + (int32_t)(e.z & 0xFFFFu);
          p_day = s0 ? t + (int32_t)(e.z >> 16) : (int32_t)DAY_NONE;
          hw[0] = 0;
        }
        if (pend && p_day <= t) {                         // the paralysis timer reads 0 with pp -1
          pend = false;
          int8_t pp = 0;
          if (*rec_byte(d, slot, RB_IPV) == 0) {          // not ipv_protected
            pp = 1;
            misc += 1ll << 16;
            const uint4 o = draw(ec.seed, (uint64_t)((int64_t)slot - ec.slot_base), t, ST_PARALYSIS);
            if ((uint64_t)o.x < ec.thr_paralysis) { to_par = true; misc += 1ll << 32; }
          }
          *rec_byte(d, slot, RB_PP) = pp;
        }
        if (i_end <= t) {                                 // I -> R: the infection timer reads 0
          sa = 3;
          misc += 1;
          hw[1] = 0;
          if (s0) hw[2] = (uint16_t)timer_left(p_day, t + 1);   // stops after today's count-down
        }
        if (to_par) *rec_byte(d, slot, RB_PARALYZED) = 1;
        if (sa != s_in) d.state[slot] = (int8_t)sa;
        if (sa == 2) e = make_uint4(slot, ent_flags(2, k, pend, p_day), (uint32_t)i_end, e.w);
      }
      if (sa != 1 && sa != 2) continue;
      surv |= 1u << q;
      ent[q] = e;
      if (do_tally) {                                     // tx_step_prep / count_SEIRP tally of E and I
        if (sa == 1) {
          e_pk += 1ll << (16 * k);
        } else {
          i_pk += 1ll << (16 * k);
          int64_t wv = (int64_t)e.w;
          if (e.w == W_REC) {
            const float rk = k == 0 ? ec.r0k[0] : k == 1 ? ec.r0k[1] : k == 2 ? ec.r0k[2] : ec.r0k[3];
            wv = weight_q24(__uint_as_float(d.rec[slot].z) * rk);
          }
          ib0 += k == 0 ? wv : 0; ib1 += k == 1 ? wv : 0; ib2 += k == 2 ? wv : 0; ib3 += k == 3 ? wv : 0;
        }
      }
    }
    // survivors -> next day's list: one atomic per piece
    {
      uint64_t mk[4];
      int rank[4], tot = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        mk[q] = __ballot((surv >> q) & 1u);
        rank[q] = tot + __popcll(mk[q] & ((1ull << lane) - 1ull));
        tot += __popcll(mk[q]);
      }
      if (tot) {
        int base = 0;
        if (lane == 0) base = atomicAdd(d.ei_cnt + (size_t)nxt * N + n, tot);
        base = __shfl(base, 0, 64);
        uint4* Ln = d.ei_list + (size_t)nxt * (size_t)d.capacity + ni.seg_begin;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if ((surv >> q) & 1u) Ln[base + rank[q]] = ent[q];
      }
    }
    if (stopped) continue;
    // the piece's tallies -> node_acc (lanes 0..14 one field each)
    e_pk = wave_sum_i64(e_pk); i_pk = wave_sum_i64(i_pk); misc = wave_sum_i64(misc);
    ib0 = wave_sum_i64(ib0); ib1 = wave_sum_i64(ib1); ib2 = wave_sum_i64(ib2); ib3 = wave_sum_i64(ib3);
    int f = -1;
    int64_t val = 0;
    if (lane < LP_MAX_STRAINS) {
      if (lane < K) { f = P_E0 + lane; val = (e_pk >> (16 * lane)) & 0xffff; }
    } else if (lane < 2 * LP_MAX_STRAINS) {
      const int k = lane - LP_MAX_STRAINS;
      if (k < K) { f = P_I0 + k; val = (i_pk >> (16 * k)) & 0xffff; }
    } else if (lane < 3 * LP_MAX_STRAINS) {
      const int k = lane - 2 * LP_MAX_STRAINS;
      if (k < K) { f = P_IB0 + k; val = k == 0 ? ib0 : k == 1 ? ib1 : k == 2 ? ib2 : ib3; }
    } else if (lane == 12) {
      f = P_R; val = misc & 0xffff;
    } else if (lane == 13) {
      f = P_NEWPP; val = (misc >> 16) & 0xffff;
    } else if (lane == 14) {
      f = P_NEWP; val = (misc >> 32) & 0xffff;
    }
    if (f >= 0 && val != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.node_acc + (size_t)n * NPART + f), (unsigned long long)val);
  }
}

// Actual timers into the records of every agent in E / I list b, whose entries are
// as of day t_next (one workgroup per node): run before anything reads the records'
// timers after days that processed the list (column sync, stage calls, a day
// without DiseaseState).
__global__ void __launch_bounds__(WG) k_ei_actual(Dev d, int b, int32_t t_next) {
  const int n = blockIdx.x;
  const int32_t cnt = d.ei_cnt[(size_t)b * d.n_nodes + n];
  const uint4* L = ei_list_of(d, b, n);
  for (int32_t j = threadIdx.x; j < cnt; j += WG) {
    const uint4 e = L[j];
    const int s = ent_state(e);
    if ((int64_t)e.x < d.capacity && (s == 1 || s == 2)) ei_actual_timers(d, e, t_next);   // not a hole
  }
}

// ---------------------------------------------------------------------------
// RI from the candidate lists (fast_ri, model.py:1826-1854), gridDim.y workgroups
// per node, four candidates per thread at a time.  ri_timer is a reference-tick value
// (timer today = col - ri_step * RI days since the reference), so only agents whose
// window can open are visited: each candidate is checked against today's window,
// drawn if eligible (vaccinated; susceptibles exposed with the RI strain, IPV
// doses), and kept for later RI days while its window can still open.  Deltas go
// to node_acc (k_node_reduce adds them to the node tallies) and to the chunk /
// span weights.  Runs after k_ei (DiseaseState reads ipv_protected before RI).
__global__ void __launch_bounds__(WG) k_ri_list(Dev d, int mask) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int n = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int sim = d.node_sim[n];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (sim_stopped(d, sim, t) || !(flags & LP_DAY_RI)) return;
  if (n == 0 && blockIdx.y == 0 && tid == 0) d.ri_ref[1] = t;
  const lp_sim_params& sp = d.sims[sim];
  const int32_t step = sp.ri_step;
  const int32_t sh_today = step * (dy.w - d.ri_cnt[d.ri_ref[0] + 1]);
  const uint64_t thr_ri = d.thr_ri[n], thr_ipv = d.thr_ipv[n];
  const uint64_t seed = sp.seed;
  const int64_t sbase = sp.slot_base;
  const int rk = sp.ri_strain & (LP_MAX_STRAINS - 1);
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const int A = dy.z & 1, B = A ^ 1, nxt = (t + 1) & 1;
  const int32_t cnt = *ri_cnt_of(d, A, n);
  const uint32_t* L = ri_list_of(d, A, n);
  uint32_t* LB = ri_list_of(d, B, n);
  __shared__ int32_t s_c[8];   // vaccinated, exposed, ipv, S, S with positive weight
  if (tid < 8) s_c[tid] = 0;
  __syncthreads();
  int32_t c_riv = 0, c_rip = 0, c_ipv = 0, c_spos = 0;
  // blockIdx.y of gridDim.y workgroups of the node takes every gridDim.y-th batch
  for (int32_t j0 = (int32_t)blockIdx.y * 4 * WG; j0 < cnt; j0 += (int32_t)gridDim.y * 4 * WG) {
    uint32_t sl[4];
    bool v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t j = j0 + q * WG + tid;
      v[q] = j < cnt;
      const uint32_t raw = v[q] ? L[j] : 0u;
      if (v[q] && (int64_t)raw >= d.capacity) { atomicOr(d.err, ERR_LIST); v[q] = false; }
      sl[q] = v[q] ? raw : 0u;
    }
    int8_t st[4];
    uint8_t ms[4];
    int16_t col[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      st[q] = v[q] ? d.state[sl[q]] : (int8_t)-1;
      ms[q] = v[q] ? d.missed[sl[q]] : (uint8_t)1;
      col[q] = v[q] ? d.ri_timer[sl[q]] : (int16_t)0;
    }
    uint32_t keep = 0, expo = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int8_t s = st[q];
      if (s < 0 || ms[q] == 1) continue;                // dead or never reached by RI
      const int16_t timer = (int16_t)(col[q] - sh_today);
      if (timer > -step) keep |= 1u << q;
      bool eligible = false;
      if (t == step) eligible = timer <= 0 && timer >= -step;
      else if (t > step) eligible = timer <= 0 && timer > -step;
      if (!eligible) continue;
      const uint32_t slot = sl[q];
      const uint4 o = draw(seed, (uint64_t)((int64_t)slot - sbase), t, ST_RI);
      if ((uint64_t)o.x < thr_ri) {
        c_riv += 1;
        if (s == 0) {                                     // exposed with the RI strain, counted as E today
          const int64_t w = weight_q24(d.risk[slot]);
          unweight_in(d, span_of(d, n, slot), slot, w);
          d.state[slot] = 1;
          *rec_byte(d, slot, RB_STRAIN) = (int8_t)rk;
          c_rip += 1;
          c_spos += (w > 0);
          expo |= 1u << q;
        }
      }
      if ((uint64_t)o.y < thr_ipv) { c_ipv += 1; *rec_byte(d, slot, RB_IPV) = 1; }
    }
    // survivors -> the other candidate buffer (one atomic per wave and batch),
    // exposures -> tomorrow's E / I list
    {
      uint64_t mk[4];
      int rank[4], tot = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        mk[q] = __ballot((keep >> q) & 1u);
        rank[q] = tot + __popcll(mk[q] & ((1ull << lane) - 1ull));
        tot += __popcll(mk[q]);
      }
      if (tot) {
        int base = 0;
        if (lane == 0) base = atomicAdd(ri_cnt_of(d, B, n), tot);
        base = __shfl(base, 0, 64);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if ((keep >> q) & 1u) LB[base + rank[q]] = sl[q];
      }
    }
    if (__ballot(expo != 0)) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool f = (expo >> q) & 1u;
        wave_append_entry(d, nxt, n, f ? exposed_entry(d, sl[q], rk, t, sp.strain_r0[rk]) : make_uint4(0u, 0u, 0u, 0u),
                          f, lane);
      }
    }
  }
  const int32_t a0 = wave_sum_i32(c_riv), a1 = wave_sum_i32(c_rip), a2 = wave_sum_i32(c_ipv),
                a3 = wave_sum_i32(c_spos);
  if (lane == 0) {
    if (a0) atomicAdd(&s_c[0], a0);
    if (a1) atomicAdd(&s_c[1], a1);
    if (a2) atomicAdd(&s_c[2], a2);
    if (a3) atomicAdd(&s_c[3], a3);
  }
  __syncthreads();
  if (tid < 8) {
    int f = -1;
    int64_t val = 0;
    switch (tid) {
      case 0: f = P_RIV; val = s_c[0]; break;
      case 1: f = P_RIP; val = s_c[1]; break;
      case 2: f = P_IPV; val = s_c[2]; break;
      case 3: f = P_S; val = -(int64_t)s_c[1]; break;
      case 4: f = P_SPOS; val = -(int64_t)s_c[3]; break;
      case 5: if (do_tally) { f = P_E0 + rk; val = s_c[1]; } break;
      default: break;
    }
    if (f >= 0 && val != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.node_acc + (size_t)n * NPART + f), (unsigned long long)val);
  }
}

// ---------------------------------------------------------------------------
// FULL days (VD step, RI and / or SIA; F = OR of the FULL_* bits).  EI: the
// pass also runs DiseaseState and the E / I tally of its E and I agents (when no
// E / I list is available); otherwise k_ei has done that already and, on VD-only
// days, only chunks with a death due or newborn slots are visited.  The span's
// tiles go through in groups of G: one round of dense loads (state and the
// day's campaign / death columns), births, deaths, RI and SIA per agent in
// registers, then the group's quads holding an agent that was E or I after
// births and deaths are compacted per wave for disease progression and the
// E / I tally, as on plain days.  Agents exposed by RI or SIA today are
// tallied as E directly and are not progressed (DiseaseState runs before RI and
// SIA in the reference day, model.py:226-241).
enum : int { FULL_VD = 1, FULL_RI = 2, FULL_SIA = 4 };

template <int F, bool EI>
__global__ void __launch_bounds__(WG, 5) k_full(Dev d, int mask) {
  // tiles per group: RI days carry ri_timer and chronically_missed per tile, SIA days date_of_birth too
  constexpr int G = (F & FULL_SIA) && (F & FULL_RI) ? 1 : (F & (FULL_SIA | FULL_RI)) ? 2 : 8;
  constexpr bool VD = F & FULL_VD, RI = F & FULL_RI, SIA = F & FULL_SIA;
  constexpr bool SPARSE_VD = F == FULL_VD && !EI;
  const int4 span0 = d.spans[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);   // the same value, known uniform
  // earliest death date of each of the span's chunks (<= 64): lane l holds chunk l
  const int32_t cm_all = (VD && lane < span0.z * (LP_TILE / LP_CHUNK))
                             ? d.chunk_min_dod[(int64_t)span0.y * (LP_TILE / LP_CHUNK) + lane] : INT_MAX;
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int node = span0.x, sim = span0.w;
  const int64_t live_hi = d.live_end[node];
  const int32_t stop_day = d.stop_day[sim];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (stop_day < t) return;
  const int want = ((flags & LP_DAY_VD_STEP) ? FULL_VD : 0) | ((flags & LP_DAY_RI) ? FULL_RI : 0) |
                   ((flags & LP_DAY_SIA) ? FULL_SIA : 0);
  if (want != F) {
    if (tid < NPART && tid != P_SW) d.partials[(size_t)blockIdx.x * NPART + tid] = 0;
    if (tid == 0) atomicOr(d.err, ERR_KIND);
    return;
  }
  const bool do_ds = flags & LP_DAY_DS;
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const bool need_ei = EI && (do_ds || do_tally);
  const int32_t sia_b = SIA ? d.day_sia_begin[t] : 0;
  const int32_t sia_n = SIA ? d.day_sia_count[t] : 0;

  // state words for the inline E / I stage: original, after VD, after RI / SIA (the
  // last only on campaign days); without it (EI false) nothing is shared between
  // threads until the partial, so the groups need no barrier
  constexpr int GW = EI ? G : 1;
  constexpr int GW2 = EI && (F & (FULL_RI | FULL_SIA)) ? G : 1;
  __shared__ uint32_t s_w0[GW][WG], s_w1[GW][WG], s_w2[GW2][WG];
  // per quad of the span: its agents exposed by RI / SIA today (bit a), appended to
  // the next day's list once at the end (the inline E / I stage appends its own)
  __shared__ uint8_t s_sv[MAX_TPW][WG];
  __shared__ int32_t lc[L_N];
  __shared__ unsigned long long lib[LP_MAX_STRAINS];
  __shared__ int64_t red[WG / 64][4];
  // VD days: the simulation's Kaplan-Meier table in LDS for the newborns' death dates
  // (a binary search of dependent loads per newborn)
  __shared__ int64_t s_km[VD ? LP_KM_LEN : 1];

  const lp_sim_params& sp = d.sims[sim];
  const int64_t nb_lo = VD ? d.nb_begin[node] : live_hi;
  const uint64_t seed = sp.seed;
  const int64_t sbase = sp.slot_base;
  const EIConst ec = ei_const(sp);
  const uint64_t thr_ri = RI ? d.thr_ri[node] : 0, thr_ipv = RI ? d.thr_ipv[node] : 0;
  const int nxt = (t + 1) & 1;
  // ri_timer holds, for live agents not chronically missed, the timer as of tick
  // ri_ref[0]: RI's decrements since then are applied by arithmetic (sh_prev: RI days
  // before today, for births and deaths; sh_today: through today, for RI)
  const int32_t cnt_ref = d.ri_cnt[d.ri_ref[0] + 1];
  const int32_t sh_prev = sp.ri_step * (dy.z - cnt_ref);
  const int32_t sh_today = sp.ri_step * (dy.w - cnt_ref);
  if (RI && blockIdx.x == 0 && tid == 0) d.ri_ref[1] = t;
  // does any of today's campaigns target this node?  (uniform over the workgroup)
  bool sia_here = false;
  if (SIA)
    for (int e = 0; e < sia_n; ++e) sia_here = sia_here || d.sia_thr[((size_t)(sia_b + e) * d.n_nodes + node) * 2] != 0;
  const bool camp = RI || sia_here;
  if (tid < L_N) lc[tid] = 0;
  if (tid < LP_MAX_STRAINS) lib[tid] = 0;
  if (VD && tid < LP_KM_LEN) s_km[tid] = d.km_cd[(size_t)sim * LP_KM_LEN + tid];
  __syncthreads();

  int32_t c_s = 0, c_spos = 0, c_r = 0;
  int64_t c_sw = 0;

  for (int g0 = 0; g0 < span0.z; g0 += G) {
    const int tile0 = span0.y + g0;
    const int ng = span0.z - g0 < G ? span0.z - g0 : G;
    // ---- dense loads ----
    // date_of_death is read only in chunks whose earliest death date among live agents is due
    uint32_t st[G], mw[G];
    int4 db[G];
    uint2 rt[G];
    int32_t cmin[G];
    bool actv[G];   // SPARSE_VD: this wave's chunk of tile u has work today (wave-uniform)
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)tile0 + u) * LP_TILE + (int64_t)tid * 4;
      const bool v = (u < ng) && (i0 < live_hi);
      cmin[u] = (VD && u < ng) ? __builtin_amdgcn_readlane(cm_all, (g0 + u) * (LP_TILE / LP_CHUNK) + wave_u) : INT_MAX;
      // VD-only days after k_ei: a chunk with no death due and no newborn slot has nothing
      // to do (wave-uniform test); its state is not read
      bool act = u < ng;
      if (SPARSE_VD) {
        const int64_t c0 = ((int64_t)tile0 + u) * LP_TILE + (int64_t)wave_u * LP_CHUNK;
        act = act && (cmin[u] <= t || (c0 + LP_CHUNK > nb_lo && c0 < live_hi));
      }
      actv[u] = act;
      st[u] = act ? reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2] : 0xFFFFFFFFu;
      if (camp) mw[u] = v ? reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2] : 0x01010101u;
      if (RI) rt[u] = v ? *reinterpret_cast<const uint2*>(d.ri_timer + i0) : make_uint2(0u, 0u);
      if (SIA) db[u] = (v && sia_here) ? *reinterpret_cast<const int4*>(d.dob + i0) : make_int4(0, 0, 0, 0);
    });

    // ---- births, deaths, RI, SIA per agent ----
    uint64_t m[G];
    uint32_t ipv_new = 0;  // today's IPV doses (bit 4u + a), written after the group's disease progression
    int n_ei = 0;
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)tile0 + u) * LP_TILE + (int64_t)tid * 4;
      if (SPARSE_VD && !actv[u]) {                        // nothing due in this chunk today
        m[u] = 0;
        if (u < ng) s_sv[g0 + u][tid] = 0;
        return;
      }
      const uint32_t w0 = st[u];
      uint32_t w1 = w0, w2 = w0;
      // death dates only in the (rare) chunks with a death due; the branch is uniform per wave
      int4 ddu = make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
      if (VD && cmin[u] <= t && u < ng && i0 < live_hi) ddu = *reinterpret_cast<const int4*>(d.dod + i0);
      uint2 r2 = RI ? rt[u] : make_uint2(0u, 0u);
      int64_t cw = 0;
      int32_t dmin = INT_MAX;   // earliest death date among the quad's agents alive after today
      uint32_t cnew = 0;        // agents exposed by RI / SIA today (bit a)
      uint32_t rcand = 0;       // RI: agents whose window can still open after today (bit a)
      bool q = false;
      // Fast path for the common quad: all four slots in the live range, no newborn, no
      // death due, no SIA here today and no agent entering its RI window; the only change
      // is then RI's timer decrement.  Anything else takes the per-agent path below.
      bool fast = false;
      if (u < ng && i0 + 3 < live_hi && !(VD && i0 + 3 >= nb_lo) && !(SIA && sia_here)) {
        const uint32_t alive = ~w0 & 0x80808080u;        // bit 8a+7 set <=> agent a has state >= 0
        bool ok = true;
        int32_t qmin = INT_MAX;
        if (VD && cmin[u] <= t) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const int32_t dod = lane_of(ddu, a);
            if ((alive >> (8 * a + 7)) & 1u) { ok = ok && dod > t; qmin = dod < qmin ? dod : qmin; }
          }
        }
        if (RI && ok) {
          const int32_t step = sp.ri_step;
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            if (!((alive >> (8 * a + 7)) & 1u) || byte_of(mw[u], a) == 1) continue;
            const uint32_t rw = a < 2 ? r2.x : r2.y;
            const int16_t timer = (int16_t)((int16_t)((rw >> (16 * (a & 1))) & 0xffffu) - sh_today);
            const bool eligible = (t == step) ? (timer <= 0 && timer >= -step) : (t > step && timer <= 0 && timer > -step);
            ok = ok && !eligible;
            if (timer > -step) rcand |= 1u << a;
          }
        }
        if (ok) {
          fast = true;
          dmin = qmin;
          q = has_ei_byte(w0);
        }
      }
      if (!fast && u < ng && i0 < live_hi) {
        rcand = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int64_t i = i0 + a;
          if (i >= live_hi) continue;
          int8_t s = byte_of(w0, a);
          const uint64_t slot = (uint64_t)(i - sbase);
          const int sh16 = 16 * (a & 1);
          const uint32_t m16 = 0xffffu << sh16;
          if (VD) {
            if (i >= nb_lo) {                             // newborn (model.py:1719-1732)
              const int32_t dod = t + km_age_at_death(s_km, seed, slot, t, 0);
              d.dob[i] = t;
              d.dod[i] = dod;
              dmin = dod < dmin ? dod : dmin;
              d.node_id[i] = (int16_t)(node - sp.node_begin);   // sim-local node id
              {                                           // its ri_timer, as of the reference tick
                const int16_t base = sp.newborn_ri_timer >= 0
                                         ? (int16_t)sp.newborn_ri_timer
                                         : (RI ? (int16_t)(((a < 2 ? r2.x : r2.y) >> sh16) & 0xffffu) : d.ri_timer[i]);
                const int16_t col = (int16_t)(base + sh_prev);
                if (sp.newborn_ri_timer >= 0 || col != base) d.ri_timer[i] = col;
                if (RI) {
                  const uint32_t hv = (uint32_t)(uint16_t)col << sh16;
                  if (a < 2) r2.x = (r2.x & ~m16) | hv; else r2.y = (r2.y & ~m16) | hv;
                }
              }
              if (SIA) { if (a == 0) db[u].x = t; else if (a == 1) db[u].y = t; else if (a == 2) db[u].z = t; else db[u].w = t; }
              const int64_t w = weight_q24(d.risk[i]);     // a newborn enters S
              c_s += 1; c_spos += (w > 0); c_sw += w; cw += w;
              s = 0;
            } else if (s >= 0 && lane_of(ddu, a) <= t) { // get_deaths (model.py:1775-1777)
              atomicAdd(&lc[L_DEATHS], 1);
              const uint4 rd = d.rec[i];
              if ((int8_t)(rd.y >> 24) == 1) atomicAdd(&lc[L_DPP], 1);
              if ((int8_t)(rd.w >> 8) == 1) atomicAdd(&lc[L_DP], 1);
              // its ri_timer stops here: store the actual value (chronically missed agents
              // are never decremented)
              if (sh_prev != 0 && (camp ? byte_of(mw[u], a) : (int8_t)d.missed[i]) != 1) {
                const int16_t c = RI ? (int16_t)(((a < 2 ? r2.x : r2.y) >> sh16) & 0xffffu) : d.ri_timer[i];
                d.ri_timer[i] = (int16_t)(c - sh_prev);
              }
              if (s == 0) {
                const int64_t w = weight_q24(d.risk[i]);
                c_s -= 1; c_spos -= (w > 0); c_sw -= w; cw -= w;
              } else if (s == 3) {
                c_r -= 1;
              }
              s = -1;
            } else if (s >= 0) {
              const int32_t dod = lane_of(ddu, a);
              dmin = dod < dmin ? dod : dmin;
            }
            w1 = (w1 & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)s << (8 * a));
          }
          q = q || s == 1 || s == 2;
          if (camp && s >= 0 && byte_of(mw[u], a) != 1) {
            const int8_t s_before = s;
            int k = 0;
            if (RI) {                                     // fast_ri (model.py:1826-1854)
              const int32_t step = sp.ri_step;
              const uint32_t rw = a < 2 ? r2.x : r2.y;
              const int16_t timer = (int16_t)((int16_t)((rw >> sh16) & 0xffffu) - sh_today);
              bool eligible = false;
              if (t == step) eligible = timer <= 0 && timer >= -step;
              else if (t > step) eligible = timer <= 0 && timer > -step;
              if (eligible) {
                const uint4 o = draw(seed, slot, t, ST_RI);
                if ((uint64_t)o.x < thr_ri) {
                  atomicAdd(&lc[L_RIV], 1);
                  if (s == 0) {
                    s = 1; k = sp.ri_strain;
                    *rec_byte(d, i, RB_STRAIN) = (int8_t)k;
                    atomicAdd(&lc[L_RIP], 1);
                  }
                }
                if ((uint64_t)o.y < thr_ipv) { atomicAdd(&lc[L_IPV], 1); ipv_new |= 1u << (4 * u + a); }
              }
            }
            if (SIA && sia_here) {                        // fast_sia (model.py:2033-2059), per event
              const int32_t age = t - lane_of(db[u], a);
              for (int e = 0; e < sia_n; ++e) {
                const int ev = sia_b + e;
                if (!(d.sia_min_age[ev] <= age && age <= d.sia_max_age[ev])) continue;
                const uint64_t t1 = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2];
                if (t1 == 0) continue;
                const uint4 o = draw(seed, slot, t, ST_SIA | ((uint32_t)e << 8));
                if ((uint64_t)o.x < t1) {
                  const bool last = (e == sia_n - 1);
                  if (last) atomicAdd(&lc[L_SIAV], 1);
                  if (s == 0 && (uint64_t)o.x < d.sia_thr[((size_t)ev * d.n_nodes + node) * 2 + 1]) {
                    s = 1;
                    k = d.sia_strain[ev] & (LP_MAX_STRAINS - 1);
                    *rec_byte(d, i, RB_STRAIN) = (int8_t)k;
                    atomicAdd(&lc[L_SIAPK0 + k], 1);
                    if (last) atomicAdd(&lc[L_SIAP], 1);
                  }
                }
              }
            }
            if (s_before == 0 && s != 0) {                // vaccinated out of S, counted as E today
              cnew |= 1u << a;
              const int64_t wt = weight_q24(d.risk[i]);
              c_s -= 1; c_spos -= (wt > 0); c_sw -= wt; cw -= wt;
              if (do_tally) atomicAdd(&lc[L_E0 + k], 1);
            }
          }
          if (RI && s >= 0 && byte_of(mw[u], a) != 1) {
            const int16_t tm = (int16_t)((int16_t)(((a < 2 ? r2.x : r2.y) >> sh16) & 0xffffu) - sh_today);
            if (tm > -sp.ri_step) rcand |= 1u << a;
          }
          w2 = (w2 & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)s << (8 * a));
        }
      }
      if (!VD) w1 = w0;
      if (u < ng) s_sv[g0 + u][tid] = (uint8_t)cnew;
      if (RI) {                                           // candidates for later RI days
        const int b = (dy.w) & 1;
        wave_append_quad_to(ri_list_of(d, b, node), ri_cnt_of(d, b, node), i0, u < ng ? rcand : 0u, lane);
      }
      m[u] = __ballot(q && need_ei);
      n_ei += __popcll(m[u]);
      const bool in_ei = (m[u] >> lane) & 1ull;
      if (u < ng) {
        if (!in_ei && w2 != w0) reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = w2;
      }
      if constexpr (EI) {
        s_w0[u][tid] = w0; s_w1[u][tid] = w1;
        if constexpr (GW2 == G) s_w2[u][tid] = w2;
      }
      if (VD && u < ng) {
        // the chunk's earliest death date: exact when its dates were read, else lowered by newborns
        const int32_t wm = wave_min_i32(dmin);
        const int32_t nm = cmin[u] <= t ? wm : (wm < cmin[u] ? wm : cmin[u]);
        if (lane == 0 && nm != cmin[u]) d.chunk_min_dod[((int64_t)tile0 + u) * (LP_TILE / LP_CHUNK) + wave_u] = nm;
      }
      // chunk weight delta (uniform condition: wave_sum needs every lane)
      if (u < ng && ((int64_t)tile0 + u) * LP_TILE < live_hi) {
        const int64_t cs = wave_sum_i64(cw);
        if (lane == 0 && cs != 0) d.chunk_sum[(((int64_t)tile0 + u) * LP_TILE >> 8) + wave] += cs;
      }
    });
    if constexpr (EI) __syncthreads();

    // ---- disease progression and E / I tally of the compacted quads ----
    for (int base = 0; base < n_ei; base += 64) {
      int j = base + lane, su = -1, src = 0;
      static_for<G>([&](auto uc) {
        constexpr int u = decltype(uc)::value;
        const int c = __popcll(m[u]);
        if (su < 0) {
          if (j < c) { su = u; src = nth_set_bit(m[u], j); }
          else j -= c;
        }
      });
      if (su < 0) continue;
      const int qt = wave * 64 + src;
      const int64_t i0 = ((int64_t)tile0 + su) * LP_TILE + (int64_t)qt * 4;
      const uint32_t wb = GW2 == G ? s_w2[GW2 == G ? su : 0][qt] : s_w1[GW == G ? su : 0][qt];
      ei_quad(d, i0, live_hi, s_w1[GW == G ? su : 0][qt], wb, s_w0[GW == G ? su : 0][qt], do_ds, do_tally, ec, t,
              lc, lib, c_r, node);
    }
    // RI's IPV doses land after DiseaseState has read ipv_protected (model.py:226-241)
    if (RI && ipv_new) {
      static_for<G>([&](auto uc) {
        constexpr int u = decltype(uc)::value;
        const int64_t i0 = ((int64_t)tile0 + u) * LP_TILE + (int64_t)tid * 4;
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if ((ipv_new >> (4 * u + a)) & 1u) *rec_byte(d, i0 + a, RB_IPV) = 1;
      });
    }
    if constexpr (EI) __syncthreads();
  }

  {
    // the span's RI / SIA exposures -> next day's list: one atomic per wave
    int c = 0;
    for (int u = 0; u < span0.z; ++u) c += __popc(s_sv[u][tid]);
    const int incl = (int)wave_incl_scan_i64(c, lane);
    const int tot = __shfl(incl, 63, 64);
    if (tot) {
      int base = 0;
      if (lane == 0) base = atomicAdd(d.ei_cnt + (size_t)nxt * d.n_nodes + node, tot);
      int pos = __shfl(base, 0, 64) + incl - c;
      uint4* L = ei_list_of(d, nxt, node);
      for (int u = 0; u < span0.z; ++u) {
        const uint32_t bits = s_sv[u][tid];
        const int64_t i0 = ((int64_t)span0.y + u) * LP_TILE + (int64_t)tid * 4;
#pragma unroll
        for (int a = 0; a < 4; ++a)
          if ((bits >> a) & 1u) {
            const int k = *rec_byte(d, i0 + a, RB_STRAIN) & (LP_MAX_STRAINS - 1);
            L[pos++] = exposed_entry(d, i0 + a, k, t, r0_sel(ec.r0k, k));
          }
      }
    }
  }
  if (VD) {
    // today's newborns in this span join the RI candidates of the next RI day (today's, if
    // today is one: births precede RI); a dense RI pass today lists its own candidates
    const int64_t s_lo = (int64_t)span0.y * LP_TILE, s_hi = s_lo + (int64_t)span0.z * LP_TILE;
    const int64_t lo = nb_lo > s_lo ? nb_lo : s_lo, hi = live_hi < s_hi ? live_hi : s_hi;
    if (!RI && hi > lo) {
      __shared__ int32_t s_base;
      const int b = dy.z & 1;
      if (tid == 0) s_base = atomicAdd(ri_cnt_of(d, b, node), (int32_t)(hi - lo));
      __syncthreads();
      uint32_t* L = ri_list_of(d, b, node);
      for (int64_t x = lo + tid; x < hi; x += WG) L[s_base + (x - lo)] = (uint32_t)x;
    }
  }
  {
    const int32_t a0 = wave_sum_i32(c_s), a1 = wave_sum_i32(c_spos), a2 = wave_sum_i32(c_r);
    const int64_t a3 = wave_sum_i64(c_sw);
    if (lane == 0) { red[wave][0] = a0; red[wave][1] = a1; red[wave][2] = a2; red[wave][3] = a3; }
  }
  __syncthreads();
  write_partial(d, tid, red, lc, lib, false, true);
}

// ---------------------------------------------------------------------------
// VD step days after k_ei, without RI or SIA in the pass (get_deaths and the
// newborns of VitalDynamics_ABM.step, model.py:1711-1734, 1767-1781).  Only chunks
// with a death due today (chunk_min_dod <= t) or newborn slots have work, a few
// thousand of the 1.2 million chunks of the national population, so the pass is
// chunk-granular instead of one workgroup per span:
//   workgroups [0, n_nodes): node n's newborn chunks (those overlapping
//     [nb_begin, live_end)), births and the deaths due in them, the Kaplan-Meier
//     table of the node's simulation in LDS; the newborn slots join the RI candidates;
//   the others: one wavefront per span, lane l checks chunk l's earliest death date,
//     the wave then takes the span's chunks with a death due (and no newborn slot).
// Each chunk belongs to one wavefront, so its state word, chunk weight and earliest
// death date are updated without atomics; span weights (partials[P_SW]) by atomic
// deltas; the per-node tallies (deaths, their paralysis flags, S / S with positive
// weight / R deltas) by one atomic per field and wavefront into node_acc, which
// k_node_reduce reads as on list-only days.
enum : int { VA_DEATHS = 0, VA_DPP, VA_DP, VA_S, VA_SPOS, VA_R, VA_N };

// One chunk of node `node` (one wavefront, lane l slots c*256 + 4l .. +3): deaths
// due today, and births in [nb_lo, live_hi) when BIRTHS.  Adds to the lane tallies
// acc[]; returns nothing (the chunk's words, weights and death date are stored).
template <bool BIRTHS>
__device__ __forceinline__ void vd_chunk(const Dev& d, int64_t c, int node, int64_t nb_lo, int64_t live_hi, int32_t cm,
                                         int32_t t, int32_t sh_prev, const lp_sim_params& sp, const int64_t* km,
                                         int32_t* acc, int lane) {
  const int64_t i0 = c * LP_CHUNK + (int64_t)lane * 4;
  const bool live_q = i0 < live_hi;
  const uint32_t w0 = live_q ? reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2] : 0xFFFFFFFFu;
  int4 ddu = make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
  const bool die_q = cm <= t && live_q;                       // a death may be due in the quad
  if (die_q) ddu = *reinterpret_cast<const int4*>(d.dod + i0);
  // the quad's other inputs load with its state and death dates: the stores below go
  // through other pointers, which the compiler cannot move these loads past
  const bool nb_q = BIRTHS && live_q && i0 + 3 >= nb_lo;     // the quad holds a newborn
  const float4 rq = (nb_q || die_q) ? *reinterpret_cast<const float4*>(d.risk + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
  const uint32_t mq = die_q ? reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2] : 0u;
  const uint2 tq = (die_q || (nb_q && sp.newborn_ri_timer < 0)) ? *reinterpret_cast<const uint2*>(d.ri_timer + i0)
                                                                : make_uint2(0u, 0u);
  auto timer_of = [&](int a) { return (int16_t)(((a < 2 ? tq.x : tq.y) >> (16 * (a & 1))) & 0xffffu); };
  uint32_t w1 = w0;
  int32_t dmin = INT_MAX;
  int64_t cw = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + a;
    if (i >= live_hi) continue;
    int8_t s = byte_of(w0, a);
    if (BIRTHS && i >= nb_lo) {                           // newborn (model.py:1719-1732)
      const uint64_t slot = (uint64_t)(i - sp.slot_base);
      const int32_t dod = t + km_age_at_death(km, sp.seed, slot, t, 0);
      d.dob[i] = t;
      d.dod[i] = dod;
      dmin = dod < dmin ? dod : dmin;
      d.node_id[i] = (int16_t)(node - sp.node_begin);   // sim-local node id
      const int16_t base = sp.newborn_ri_timer >= 0 ? (int16_t)sp.newborn_ri_timer : timer_of(a);
      const int16_t col = (int16_t)(base + sh_prev);     // as of the reference tick
      if (sp.newborn_ri_timer >= 0 || col != base) d.ri_timer[i] = col;
      const int64_t w = weight_q24(lane_of(rq, a));      // a newborn enters S
      acc[VA_S] += 1; acc[VA_SPOS] += (w > 0); cw += w;
      s = 0;
    } else if (s >= 0 && lane_of(ddu, a) <= t) {         // get_deaths (model.py:1775-1777)
      acc[VA_DEATHS] += 1;
      const uint4 rd = d.rec[i];
      if ((int8_t)(rd.y >> 24) == 1) acc[VA_DPP] += 1;
      if ((int8_t)(rd.w >> 8) == 1) acc[VA_DP] += 1;
      // its ri_timer stops here: the actual value (chronically missed agents are never decremented)
      if (sh_prev != 0 && byte_of(mq, a) != 1) d.ri_timer[i] = (int16_t)(timer_of(a) - sh_prev);
      if (s == 0) {
        const int64_t w = weight_q24(lane_of(rq, a));
        acc[VA_S] -= 1; acc[VA_SPOS] -= (w > 0); cw -= w;
      } else if (s == 3) {
        acc[VA_R] -= 1;
      }
      s = -1;
    } else if (s >= 0) {
      const int32_t dod = lane_of(ddu, a);
      dmin = dod < dmin ? dod : dmin;
    }
    w1 = (w1 & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)s << (8 * a));
  }
  if (live_q && w1 != w0) reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = w1;
  // the chunk's earliest death date: exact when its dates were read, else lowered by newborns
  const int32_t wm = wave_min_i32(dmin);
  const int32_t nm = cm <= t ? wm : (wm < cm ? wm : cm);
  const int64_t cs = wave_sum_i64(cw);
  if (lane == 0) {
    if (nm != cm) d.chunk_min_dod[c] = nm;
    if (cs != 0) {   // (no value returned: nothing waits for these)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.chunk_sum + c), (unsigned long long)cs);
      atomicAdd(reinterpret_cast<unsigned long long*>(d.partials + (size_t)span_of(d, node, c * LP_CHUNK) * NPART + P_SW),
                (unsigned long long)cs);
    }
  }
}

// The wave's tallies of node `node` -> node_acc (one atomic per nonzero field).
__device__ __forceinline__ void vd_flush(const Dev& d, int node, int32_t* acc, int lane) {
  int32_t v[VA_N];
#pragma unroll
  for (int f = 0; f < VA_N; ++f) v[f] = wave_sum_i32(acc[f]);
  if (lane < VA_N) {
    const int f = lane;
    const int32_t x = f == 0 ? v[0] : f == 1 ? v[1] : f == 2 ? v[2] : f == 3 ? v[3] : f == 4 ? v[4] : v[5];
    const int pf = f == VA_DEATHS ? P_DEATHS : f == VA_DPP ? P_DPP : f == VA_DP ? P_DP : f == VA_S ? P_S
                 : f == VA_SPOS ? P_SPOS : P_R;
    if (x != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.node_acc + (size_t)node * NPART + pf),
                (unsigned long long)(int64_t)x);
  }
}

__global__ void __launch_bounds__(WG) k_vd(Dev d, int mask) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int N = d.n_nodes;
  const int32_t flags = stage_day_flags(dy.y, mask);
  const int32_t cnt_ref = d.ri_cnt[d.ri_ref[0] + 1];
  int32_t acc[VA_N] = {0, 0, 0, 0, 0, 0};
  if ((int)blockIdx.x < N) {
    // ---- node n's newborn chunks ----
    const int n = blockIdx.x;
    const int sim = d.node_sim[n];
    const int64_t nb_lo = d.nb_begin[n], live_hi = d.live_end[n];
    __shared__ int64_t s_km[LP_KM_LEN];
    if (nb_lo >= live_hi || sim_stopped(d, sim, t) || !(flags & LP_DAY_VD_STEP)) return;
    const lp_sim_params& sp = d.sims[sim];
    if (tid < LP_KM_LEN) s_km[tid] = d.km_cd[(size_t)sim * LP_KM_LEN + tid];
    const int32_t sh_prev = sp.ri_step * (dy.z - cnt_ref);
    {
      // the newborn slots join the candidates of the next RI day (today's, if today is one:
      // births precede RI)
      __shared__ int32_t s_base;
      const int b = dy.z & 1;
      if (tid == 0) s_base = atomicAdd(ri_cnt_of(d, b, n), (int32_t)(live_hi - nb_lo));
      __syncthreads();
      uint32_t* L = ri_list_of(d, b, n);
      for (int64_t x = nb_lo + tid; x < live_hi; x += WG) L[s_base + (x - nb_lo)] = (uint32_t)x;
    }
    const int64_t c_lo = nb_lo / LP_CHUNK, c_hi = (live_hi + LP_CHUNK - 1) / LP_CHUNK;
    for (int64_t c = c_lo + (tid >> 6); c < c_hi; c += WG / 64) {
      const int32_t cm = d.chunk_min_dod[c];
      vd_chunk<true>(d, c, n, nb_lo, live_hi, cm, t, sh_prev, sp, s_km, acc, lane);
    }
    vd_flush(d, n, acc, lane);
    return;
  }
  // ---- one wavefront per span: chunks with a death due ----
  const int64_t s = ((int64_t)blockIdx.x - N) * (WG / 64) + (tid >> 6);
  if (s >= d.n_spans) return;
  const int4 span = d.spans[s];
  const int node = span.x, sim = span.w;
  const int64_t c0 = (int64_t)span.y * (LP_TILE / LP_CHUNK);
  const int nc = span.z * (LP_TILE / LP_CHUNK);
  const int32_t cm = lane < nc ? d.chunk_min_dod[c0 + lane] : INT_MAX;
  const int64_t nb_lo = d.nb_begin[node], live_hi = d.live_end[node];
  if (sim_stopped(d, sim, t)) return;
  if (!(flags & LP_DAY_VD_STEP)) { if (lane == 0) atomicOr(d.err, ERR_KIND); return; }
  // chunks overlapping today's newborn slots belong to the node's workgroup above
  const int64_t c = c0 + lane;
  const bool nb_chunk = nb_lo < live_hi && c * LP_CHUNK < live_hi && (c + 1) * LP_CHUNK > nb_lo;
  uint64_t due = __ballot(lane < nc && cm <= t && !nb_chunk);
  if (due == 0) return;
  const lp_sim_params& sp = d.sims[sim];
  const int32_t sh_prev = sp.ri_step * (dy.z - cnt_ref);
  while (due) {
    const int l = __ffsll((unsigned long long)due) - 1;
    due &= due - 1;
    vd_chunk<false>(d, c0 + l, node, nb_lo, live_hi, __shfl(cm, l, 64), t, sh_prev, sp, nullptr, acc, lane);
  }
  vd_flush(d, node, acc, lane);
}

// ---------------------------------------------------------------------------
// SIA campaign days after k_ei (and k_vd / k_ri_list when VitalDynamics and RI
// step the same day: births, deaths and RI precede SIA, model.py:226-241): fast_sia
// (model.py:1994-2060) per event, over the agents of the nodes an event targets.
// One workgroup per span; a span of a node no event targets returns at once.  Each
// lane takes its quad of each of the span's tiles in turn, with the state,
// chronically_missed and date_of_birth words of the next two tiles already loading
// (a rolled loop: one copy of the per-agent code, whose draws are the bulk of the
// instructions); live, reached agents of an event's age band get one draw per event
// (the day's events in LDS): vaccinated counts, susceptibles exposed with the
// event's vaccine strain.  Tallies per workgroup into node_acc (as k_vd), chunk
// weights per wavefront, the span weight by one atomic; exposed agents join
// tomorrow's E / I list.
constexpr int SIA_LDS_EV = 64;   // events of one day held in LDS (more: read from global memory)
constexpr int SIA_XG = 4;        // tiles whose exposures are queued in LDS and finished together

__global__ void __launch_bounds__(WG) k_sia(Dev d, int mask) {
  const int4 span = d.spans[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int node = span.x, sim = span.w;
  const int64_t live_hi = d.live_end[node];
  const int32_t stop_day = d.stop_day[sim];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (stop_day < t) return;
  if (!(flags & LP_DAY_SIA)) { if (tid == 0) atomicOr(d.err, ERR_KIND); return; }
  const int32_t sia_b = d.day_sia_begin[t], sia_n = d.day_sia_count[t];
  __shared__ uint64_t e_thr[SIA_LDS_EV][2];
  __shared__ int32_t e_lo[SIA_LDS_EV], e_hi[SIA_LDS_EV], e_k[SIA_LDS_EV];
  __shared__ int s_here;
  if (tid == 0) s_here = 0;
  __syncthreads();
  for (int e = tid; e < sia_n; e += WG) {
    const int ev = sia_b + e;
    const uint64_t t1 = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2];
    if (t1 != 0) s_here = 1;
    if (e < SIA_LDS_EV) {
      e_thr[e][0] = t1;
      e_thr[e][1] = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2 + 1];
      e_lo[e] = d.sia_min_age[ev];
      e_hi[e] = d.sia_max_age[ev];
      e_k[e] = d.sia_strain[ev] & (LP_MAX_STRAINS - 1);
    }
  }
  __syncthreads();
  if (!s_here || (int64_t)span.y * LP_TILE >= live_hi) return;
  const lp_sim_params& sp = d.sims[sim];
  const uint64_t seed = sp.seed;
  const int64_t sbase = sp.slot_base;
  float r0k[LP_MAX_STRAINS];
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) r0k[k] = sp.strain_r0[k];
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const int nxt = (t + 1) & 1;
  // lane tallies: vaccinated / protected by the day's last event (the reference's result
  // rows keep the last event's counts), exposures by strain, S and S-with-weight deltas
  int32_t c_v = 0, c_p = 0, c_s = 0, c_spos = 0;
  int32_t c_k[LP_MAX_STRAINS] = {0, 0, 0, 0};
  int64_t c_sw = 0;
  // the words of tiles u (cur), u + 1 and u + 2 (in flight)
  auto load = [&](int u, uint32_t& w, uint32_t& m, int4& b) {
    const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
    const bool v = u < span.z && i0 < live_hi;
    w = v ? reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2] : 0xFFFFFFFFu;
    m = v ? reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2] : 0x01010101u;
    b = v ? *reinterpret_cast<const int4*>(d.dob + i0) : make_int4(0, 0, 0, 0);
  };
  // the widest age band of the events targeting the node: agents outside it are skipped
  int32_t band_lo = INT_MAX, band_hi = INT_MIN;
  for (int e = 0; e < sia_n; ++e) {
    const bool on = e < SIA_LDS_EV ? e_thr[e][0] != 0
                                   : d.sia_thr[((size_t)(sia_b + e) * d.n_nodes + node) * 2] != 0;
    if (!on) continue;
    const int32_t lo = e < SIA_LDS_EV ? e_lo[e] : d.sia_min_age[sia_b + e];
    const int32_t hi = e < SIA_LDS_EV ? e_hi[e] : d.sia_max_age[sia_b + e];
    band_lo = lo < band_lo ? lo : band_lo;
    band_hi = hi > band_hi ? hi : band_hi;
  }
  // per wave: the tile's state words and birth dates, and the compacted list of its
  // agents in the band (quad << 2 | agent), so that each round of draws serves 64 of them
  __shared__ uint32_t c_w[WG / 64][64];
  __shared__ int4 c_db[WG / 64][64];
  __shared__ uint16_t c_list[WG / 64][256];
  // the exposures of up to SIA_XG tiles ((slot - first slot of the span) << 2 | strain),
  // finished together by the workgroup: their loads in flight at once, one reservation
  // in tomorrow's list; the span's chunk weight deltas accumulate in LDS
  __shared__ uint32_t x_buf[SIA_XG * LP_TILE];
  __shared__ int32_t x_cnt, x_base;
  __shared__ int64_t x_cw[MAX_TPW * (LP_TILE / LP_CHUNK)];
  const int64_t span_lo = (int64_t)span.y * LP_TILE;
  if (tid == 0) x_cnt = 0;
  if (tid < MAX_TPW * (LP_TILE / LP_CHUNK)) x_cw[tid] = 0;
  __syncthreads();
  auto finish = [&]() {
    __syncthreads();
    const int n = x_cnt;
    if (n > 0) {
      if (tid == 0) x_base = atomicAdd(d.ei_cnt + (size_t)nxt * d.n_nodes + node, n);
      __syncthreads();
      uint4* L = ei_list_of(d, nxt, node) + x_base;
      for (int j = tid; j < n; j += WG) {
        const uint32_t code = x_buf[j];
        const int64_t i = span_lo + (int64_t)(code >> 2);
        const int k = (int)(code & 3u);
        const float rk = d.risk[i];
        const uint4 rec = d.rec[i];
        d.state[i] = 1;                                   // vaccinated out of S, counted as E today
        *rec_byte(d, i, RB_STRAIN) = (int8_t)k;
        const int64_t w = weight_q24(rk);
        c_s -= 1; c_spos -= (w > 0); c_sw -= w;
        if (w != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&x_cw[(i - span_lo) >> 8]), (unsigned long long)(-w));
        L[j] = exposed_entry_rec(i, rec, k, t, r0_sel(r0k, k));
      }
    }
    __syncthreads();
    if (tid == 0) x_cnt = 0;
    __syncthreads();
  };
  uint32_t w_a, m_a, w_b, m_b, w_c, m_c;
  int4 b_a, b_b, b_c;
  load(0, w_a, m_a, b_a);
  load(1, w_b, m_b, b_b);
#pragma unroll 1
  for (int u = 0; u < span.z; ++u) {
    load(u + 2, w_c, m_c, b_c);
    const int64_t tile = (int64_t)span.y + u;
    const int64_t i0 = tile * LP_TILE + (int64_t)tid * 4;
    uint32_t elig = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int32_t age = t - lane_of(b_a, a);
      if (i0 + a < live_hi && byte_of(w_a, a) >= 0 && byte_of(m_a, a) != 1 && age >= band_lo && age <= band_hi)
        elig |= 1u << a;
    }
    int tot = 0;
    {
      const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const uint64_t mk = __ballot((elig >> a) & 1u);
        if ((elig >> a) & 1u) c_list[wave][tot + __popcll(mk & lt)] = (uint16_t)((lane << 2) | a);
        tot += __popcll(mk);
      }
    }
    c_w[wave][lane] = w_a;
    c_db[wave][lane] = b_a;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int base = 0; base < tot; base += 64) {
      const int j = base + lane;
      bool ex = false;
      int k_ex = 0;
      uint32_t off = 0;
      if (j < tot) {
        const int code = c_list[wave][j];
        const int l = code >> 2, a = code & 3;
        off = (uint32_t)(u * LP_TILE + (wave * 64 + l) * 4 + a);   // lane l of this wave owns quad wave * 64 + l
        int8_t s = byte_of(c_w[wave][l], a);
        const int32_t age = t - lane_of(c_db[wave][l], a);
        const uint64_t slot = (uint64_t)(span_lo + off - sbase);
        for (int e = 0; e < sia_n; ++e) {                 // in the day's order
          uint64_t t1, t2;
          int32_t lo, hi, k;
          if (e < SIA_LDS_EV) {
            lo = e_lo[e]; hi = e_hi[e];
            if (!(lo <= age && age <= hi)) continue;
            t1 = e_thr[e][0]; t2 = e_thr[e][1]; k = e_k[e];
          } else {
            const int ev = sia_b + e;
            lo = d.sia_min_age[ev]; hi = d.sia_max_age[ev];
            if (!(lo <= age && age <= hi)) continue;
            t1 = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2];
            t2 = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2 + 1];
            k = d.sia_strain[ev] & (LP_MAX_STRAINS - 1);
          }
          if (t1 == 0) continue;
          const uint4 o = draw(seed, slot, t, ST_SIA | ((uint32_t)e << 8));
          if ((uint64_t)o.x < t1) {
            const bool last = e == sia_n - 1;
            if (last) c_v += 1;
            if (s == 0 && (uint64_t)o.x < t2) {
              s = 1;
              ex = true;
              k_ex = k;
              c_k[0] += k == 0; c_k[1] += k == 1; c_k[2] += k == 2; c_k[3] += k == 3;
              if (last) c_p += 1;
            }
          }
        }
      }
      // queue the round's exposures: one LDS reservation per wave
      const uint64_t mk = __ballot(ex);
      if (mk) {
        int b0 = 0;
        if (lane == 0) b0 = atomicAdd(&x_cnt, __popcll(mk));
        b0 = __shfl(b0, 0, 64);
        if (ex) x_buf[b0 + __popcll(mk & ((1ull << lane) - 1ull))] = (off << 2) | (uint32_t)k_ex;
      }
    }
    __builtin_amdgcn_wave_barrier();   // the wave's LDS rows are reused for the next tile
    if ((u + 1) % SIA_XG == 0 || u + 1 == span.z) finish();   // uniform
    w_a = w_b; m_a = m_b; b_a = b_b;
    w_b = w_c; m_b = m_c; b_b = b_c;
  }
  // the span's chunk weights
  if (tid < span.z * (LP_TILE / LP_CHUNK) && x_cw[tid] != 0) d.chunk_sum[(int64_t)span.y * (LP_TILE / LP_CHUNK) + tid] += x_cw[tid];
  // the workgroup's tallies -> node_acc, the span's weight delta -> its partial
  __shared__ int64_t red[WG / 64][9];
  {
    int64_t v[9] = {c_v, c_p, c_s, c_spos, c_sw, c_k[0], c_k[1], c_k[2], c_k[3]};
#pragma unroll
    for (int f = 0; f < 9; ++f) v[f] = wave_sum_i64(v[f]);
    if (lane == 0) {
#pragma unroll
      for (int f = 0; f < 9; ++f) red[wave][f] = v[f];
    }
  }
  __syncthreads();
  if (tid < 13) {
    int64_t x = 0;
    int f = -1;
    auto sum = [&](int q) { int64_t s = 0; for (int w = 0; w < WG / 64; ++w) s += red[w][q]; return s; };
    if (tid == 0) { f = P_SIAV; x = sum(0); }
    else if (tid == 1) { f = P_SIAP; x = sum(1); }
    else if (tid == 2) { f = P_S; x = sum(2); }
    else if (tid == 3) { f = P_SPOS; x = sum(3); }
    else if (tid < 8) { const int k = tid - 4; if (k < d.K) { f = P_SIAPK0 + k; x = sum(5 + k); } }
    else if (tid < 12) { const int k = tid - 8; if (k < d.K && do_tally) { f = P_E0 + k; x = sum(5 + k); } }
    else { x = sum(4); if (x != 0) atomicAdd(reinterpret_cast<unsigned long long*>(d.partials + (size_t)blockIdx.x * NPART + P_SW),
                                              (unsigned long long)x); }
    if (f >= 0 && x != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.node_acc + (size_t)node * NPART + f), (unsigned long long)x);
  }
}

// ---------------------------------------------------------------------------
// RECOUNT passes: the span's tiles in groups of G, each group through phases
// A-D (every agent's risk, paralysis flags and state recounted); tallies
// accumulate over the groups, one partial per span.
template <int G, int KIND>
__global__ void __launch_bounds__(WG) k_agents(Dev d, int mask) {
  constexpr bool RECOUNT = KIND == KIND_RECOUNT;
  const int4 span0 = d.spans[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int node = span0.x, sim = span0.w;
  const int64_t live_hi = d.live_end[node];
  const int32_t stop_day = d.stop_day[sim];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (stop_day < t) return;
  if (!RECOUNT && is_plain_day(flags)) {
    if (tid < NPART && tid != P_SW) d.partials[(size_t)blockIdx.x * NPART + tid] = 0;
    if (tid == 0) atomicOr(d.err, ERR_KIND);
    return;
  }
  const bool do_deaths = flags & LP_DAY_VD_STEP;
  const bool do_ds = flags & LP_DAY_DS;
  const bool do_ri = flags & LP_DAY_RI;
  const bool do_sia = flags & LP_DAY_SIA;
  const bool do_tally = (flags & LP_DAY_TX) != 0;
  const bool need_ei = do_ds || do_tally;
  const int32_t sia_b = do_sia ? d.day_sia_begin[t] : 0;
  const int32_t sia_n = do_sia ? d.day_sia_count[t] : 0;

  __shared__ int32_t lc[L_N];
  __shared__ unsigned long long lib[LP_MAX_STRAINS];
  __shared__ int64_t red[WG / 64][4];

  const lp_sim_params& sp = d.sims[sim];
  const int64_t nb_lo = d.nb_begin[node];
  const uint64_t seed = sp.seed;
  const int64_t sbase = sp.slot_base;
  const uint64_t thr_ri = do_ri ? d.thr_ri[node] : 0, thr_ipv = do_ri ? d.thr_ipv[node] : 0;
  if (tid < L_N) lc[tid] = 0;
  if (tid < LP_MAX_STRAINS) lib[tid] = 0;
  __syncthreads();

  // carried-statistic deltas (absolute counts when recounting)
  int32_t c_s = 0, c_spos = 0, c_r = 0;
  int64_t c_sw = 0;

  for (int g0 = 0; g0 < span0.z; g0 += G) {
    const int4 span = make_int4(node, span0.y + g0, span0.z - g0 < G ? span0.z - g0 : G, sim);
    // ---- phase A ----
    uint32_t st[G], st0[G];
    float4 rk[G];
    int4 dd[G], db[G];
    uint2 rt[G];
    uint32_t mw[G];
    bool tv[G];
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
      const bool v = (u < span.z) && (i0 < live_hi);
      tv[u] = v;
      st[u] = (u < span.z) ? reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2] : 0xFFFFFFFFu;
      st0[u] = st[u];
      if (RECOUNT) rk[u] = v ? *reinterpret_cast<const float4*>(d.risk + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
      if (do_deaths || RECOUNT) dd[u] = v ? *reinterpret_cast<const int4*>(d.dod + i0) : make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
      if (do_ri || do_sia) mw[u] = v ? reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2] : 0x01010101u;
      if (do_ri) rt[u] = v ? *reinterpret_cast<const uint2*>(d.ri_timer + i0) : make_uint2(0u, 0u);
      if (do_sia) db[u] = v ? *reinterpret_cast<const int4*>(d.dob + i0) : make_int4(0, 0, 0, 0);
    });
    int64_t cw[G];
    static_for<G>([&](auto uc) { cw[decltype(uc)::value] = 0; });

    // ---- phase B: births and deaths ----
    const bool vd_work = do_deaths || nb_lo < live_hi;
    bool ei[G];
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
      bool any = false;
      if (tv[u]) {
        if (vd_work) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const int64_t i = i0 + a;
            if (i >= live_hi) continue;
            int8_t s = byte_of(st[u], a);
            if (i >= nb_lo) {
              const uint64_t slot = (uint64_t)(i - sbase);
              d.dob[i] = t;
              d.dod[i] = t + km_age_at_death(d.km_cd + (size_t)sim * LP_KM_LEN, seed, slot, t, 0);
              d.node_id[i] = (int16_t)(node - sp.node_begin);   // sim-local node id
              if (sp.newborn_ri_timer >= 0) {
                d.ri_timer[i] = (int16_t)sp.newborn_ri_timer;
                if (do_ri) {
                  const uint32_t hv = (uint32_t)(uint16_t)sp.newborn_ri_timer << (16 * (a & 1));
                  const uint32_t hm = 0xffffu << (16 * (a & 1));
                  if (a < 2) rt[u].x = (rt[u].x & ~hm) | hv; else rt[u].y = (rt[u].y & ~hm) | hv;
                }
              }
              if (do_sia) { if (a == 0) db[u].x = t; else if (a == 1) db[u].y = t; else if (a == 2) db[u].z = t; else db[u].w = t; }
              if (!RECOUNT) {                             // a newborn enters S
                const int64_t w = weight_q24(d.risk[i]);
                c_s += 1; c_spos += (w > 0); c_sw += w; cw[u] += w;
              }
              s = 0;
            } else if (do_deaths && s >= 0 && lane_of(dd[u], a) <= t) {
              atomicAdd(&lc[L_DEATHS], 1);
              if (d.pp[i] == 1) atomicAdd(&lc[L_DPP], 1);
              if (d.paralyzed[i] == 1) atomicAdd(&lc[L_DP], 1);
              if (!RECOUNT) {
                if (s == 0) {                             // a susceptible leaves S
                  const int64_t w = weight_q24(d.risk[i]);
                  c_s -= 1; c_spos -= (w > 0); c_sw -= w; cw[u] -= w;
                } else if (s == 3) {
                  c_r -= 1;
                }
              }
              s = -1;
            }
            st[u] = (st[u] & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)s << (8 * a));
          }
        }
        const uint32_t w = st[u];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int8_t s = byte_of(w, a);
          any = any || ((s == 1 || s == 2) && i0 + a < live_hi);
        }
      }
      ei[u] = any && need_ei;
    });
    if (RECOUNT) {
      // earliest death date among each chunk's live agents (date_of_death of newborns just written)
      static_for<G>([&](auto uc) {
        constexpr int u = decltype(uc)::value;
        const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
        int32_t dmin = INT_MAX;
        if (tv[u]) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const int64_t i = i0 + a;
            if (i >= live_hi || byte_of(st[u], a) < 0) continue;
            const int32_t dod = i >= nb_lo ? d.dod[i] : lane_of(dd[u], a);
            dmin = dod < dmin ? dod : dmin;
          }
        }
        if (u < span.z) {
          const int32_t wm = wave_min_i32(dmin);
          if (lane == 0) d.chunk_min_dod[((int64_t)span.y + u) * (LP_TILE / LP_CHUNK) + wave] = wm;
        }
      });
    }

    // ---- phase C: E/I columns of quads holding an E or I agent ----
    uint2 xe[G], xi[G], xp[G];
    uint32_t xs[G], xpp[G], xipv[G];
    float4 xin[G];
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
      if (ei[u]) {
        xe[u] = *reinterpret_cast<const uint2*>(d.exp_timer + i0);
        xi[u] = *reinterpret_cast<const uint2*>(d.inf_timer + i0);
        xp[u] = *reinterpret_cast<const uint2*>(d.par_timer + i0);
        xs[u] = reinterpret_cast<const uint32_t*>(d.strain)[i0 >> 2];
        xpp[u] = reinterpret_cast<const uint32_t*>(d.pp)[i0 >> 2];
        xipv[u] = reinterpret_cast<const uint32_t*>(d.ipv)[i0 >> 2];
        xin[u] = do_tally ? *reinterpret_cast<const float4*>(d.infect + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        xe[u] = xi[u] = xp[u] = make_uint2(0u, 0u);
        xs[u] = xpp[u] = xipv[u] = 0u;
        xin[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    });

    // ---- phase D: progression, RI, SIA, tallies ----
    static_for<G>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int64_t i0 = ((int64_t)span.y + u) * LP_TILE + (int64_t)tid * 4;
      uint32_t wf = st[u];   // the quad's state word after the pass
      if (tv[u] && !ei[u] && !do_ri && !do_sia) {
        // no E/I in the quad and no campaign today: only a recount looks at it
        if (RECOUNT) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const int8_t s = byte_of(st[u], a);
            const bool live = i0 + a < live_hi;
            if (live && s == 0) {
              const int64_t wt = weight_q24(lane_of(rk[u], a));
              c_s += 1; c_spos += (wt > 0); c_sw += wt; cw[u] += wt;
            }
            c_r += (live && s == 3);
            if (live && s >= 0 && d.paralyzed[i0 + a] == 1) atomicAdd(&lc[L_P], 1);
            if (live && s >= 0 && d.pp[i0 + a] == 1) atomicAdd(&lc[L_POTP], 1);
          }
        }
        if (vd_work && st[u] != st0[u]) reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = st[u];
      } else if (tv[u]) {
        uint32_t w = st[u];
        uint2 e2 = xe[u], i2 = xi[u], p2 = xp[u], r2 = do_ri ? rt[u] : make_uint2(0u, 0u);
        uint32_t ppw = xpp[u];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int64_t i = i0 + a;
          if (i >= live_hi) continue;
          int8_t s = byte_of(w, a);
          const uint64_t slot = (uint64_t)(i - sbase);
          int k = byte_of(xs[u], a) & (LP_MAX_STRAINS - 1);
          const int sh16 = 16 * (a & 1);
          const uint32_t m16 = 0xffffu << sh16;
          if (do_ds && (s == 1 || s == 2)) {
            if (progress_agent(a, s, byte_of(xs[u], a), e2, i2, p2, ppw, xipv[u], seed, slot, t, sp.thr_paralysis,
                               d.paralyzed + i, &lc[L_NEWPP], &lc[L_NEWP]) && !RECOUNT)
              c_r += 1;                                   // I -> R
          }
          if ((do_ri || do_sia) && s >= 0 && byte_of(mw[u], a) != 1) {
            const int8_t s_before = s;
            if (do_ri) {                                  // fast_ri (model.py:1826-1854)
              const int32_t step = sp.ri_step;
              const uint32_t rw = a < 2 ? r2.x : r2.y;
              const int16_t timer = (int16_t)((int16_t)((rw >> sh16) & 0xffffu) - step);
              const uint32_t nw = (rw & ~m16) | ((uint32_t)(uint16_t)timer << sh16);
              if (a < 2) r2.x = nw; else r2.y = nw;
              bool eligible = false;
              if (t == step) eligible = timer <= 0 && timer >= -step;
              else if (t > step) eligible = timer <= 0 && timer > -step;
              if (eligible) {
                const uint4 o = draw(seed, slot, t, ST_RI);
                if ((uint64_t)o.x < thr_ri) {
                  atomicAdd(&lc[L_RIV], 1);
                  if (s == 0) { s = 1; k = sp.ri_strain; d.strain[i] = (int8_t)k; atomicAdd(&lc[L_RIP], 1); }
                }
                if ((uint64_t)o.y < thr_ipv) { atomicAdd(&lc[L_IPV], 1); d.ipv[i] = 1; }
              }
            }
            if (do_sia) {                                 // fast_sia (model.py:2033-2059), per event
              const int32_t age = t - lane_of(db[u], a);
              for (int e = 0; e < sia_n; ++e) {
                const int ev = sia_b + e;
                if (!(d.sia_min_age[ev] <= age && age <= d.sia_max_age[ev])) continue;
                const uint64_t t1 = d.sia_thr[((size_t)ev * d.n_nodes + node) * 2];
                if (t1 == 0) continue;
                const uint4 o = draw(seed, slot, t, ST_SIA | ((uint32_t)e << 8));
                if ((uint64_t)o.x < t1) {
                  const bool last = (e == sia_n - 1);
                  if (last) atomicAdd(&lc[L_SIAV], 1);
                  if (s == 0 && (uint64_t)o.x < d.sia_thr[((size_t)ev * d.n_nodes + node) * 2 + 1]) {
                    s = 1;
                    k = d.sia_strain[ev] & (LP_MAX_STRAINS - 1);
                    d.strain[i] = (int8_t)k;
                    atomicAdd(&lc[L_SIAPK0 + k], 1);
                    if (last) atomicAdd(&lc[L_SIAP], 1);
                  }
                }
              }
            }
            if (!RECOUNT && s_before == 0 && s != 0) {    // vaccinated out of S
              const int64_t wt = weight_q24(d.risk[i]);
              c_s -= 1; c_spos -= (wt > 0); c_sw -= wt; cw[u] -= wt;
            }
          }
          if (s >= 0) {
            if (RECOUNT) {
              if (s == 0) {
                const int64_t wt = weight_q24(lane_of(rk[u], a));
                c_s += 1; c_spos += (wt > 0); c_sw += wt; cw[u] += wt;
              } else if (s == 3) {
                c_r += 1;
              }
              if (d.pp[i] == 1 || byte_of(ppw, a) == 1) atomicAdd(&lc[L_POTP], 1);
              if (d.paralyzed[i] == 1) atomicAdd(&lc[L_P], 1);
            }
            if (do_tally) {                               // tx_step_prep / count_SEIRP tally of E and I
              if (s == 1) {
                atomicAdd(&lc[L_E0 + k], 1);
              } else if (s == 2) {
                atomicAdd(&lc[L_I0 + k], 1);
                const int64_t wv = weight_q24(lane_of(xin[u], a) * sp.strain_r0[k]);
                atomicAdd(&lib[k], (unsigned long long)wv);
              }
            }
          }
          w = (w & ~(0xffu << (8 * a))) | ((uint32_t)(uint8_t)s << (8 * a));
        }
        if (w != st0[u]) reinterpret_cast<uint32_t*>(d.state)[i0 >> 2] = w;
        wf = w;
        if (ei[u]) {
          if (e2.x != xe[u].x || e2.y != xe[u].y) *reinterpret_cast<uint2*>(d.exp_timer + i0) = e2;
          if (i2.x != xi[u].x || i2.y != xi[u].y) *reinterpret_cast<uint2*>(d.inf_timer + i0) = i2;
          if (p2.x != xp[u].x || p2.y != xp[u].y) *reinterpret_cast<uint2*>(d.par_timer + i0) = p2;
          if (ppw != xpp[u]) reinterpret_cast<uint32_t*>(d.pp)[i0 >> 2] = ppw;
        }
        if (do_ri && (r2.x != rt[u].x || r2.y != rt[u].y)) *reinterpret_cast<uint2*>(d.ri_timer + i0) = r2;
      }
      if (u < span.z) {                                   // E / I list of the next day (uniform condition)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int8_t s = byte_of(wf, a);
          if (!(tv[u] && i0 + a < live_hi && (s == 1 || s == 2))) continue;
          const int64_t i = i0 + a;   // timers as written by this pass: as of tomorrow
          const int k = d.strain[i] & (LP_MAX_STRAINS - 1);
          const int64_t w = weight_q24(d.infect[i] * sp.strain_r0[k]);
          const uint4 e = s == 1 ? entry_e((uint32_t)i, k, d.exp_timer[i], d.inf_timer[i], d.par_timer[i], d.pp[i],
                                           t + 1, w)
                                 : entry_i((uint32_t)i, k, d.inf_timer[i], d.par_timer[i], d.pp[i], t + 1, w);
          append_entry(d, (t + 1) & 1, node, e);
        }
        // RI candidates: live agents not chronically missed whose window can still open
        // (ri_timer is actual here; re-read after this pass's own RI step)
        uint32_t rb = 0;
        if (tv[u]) {
          const uint2 rr = *reinterpret_cast<const uint2*>(d.ri_timer + i0);
          const uint32_t mm = reinterpret_cast<const uint32_t*>(d.missed)[i0 >> 2];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const int16_t tm = (int16_t)(((a < 2 ? rr.x : rr.y) >> (16 * (a & 1))) & 0xffffu);
            if (i0 + a < live_hi && byte_of(wf, a) >= 0 && byte_of(mm, a) != 1 && tm > -sp.ri_step) rb |= 1u << a;
          }
        }
        const int rbuf = dy.w & 1;
        wave_append_quad_to(ri_list_of(d, rbuf, node), ri_cnt_of(d, rbuf, node), i0, rb, lane);
      }
      // chunk weight sums: rewritten when recounting, adjusted by their delta otherwise
      // (the condition is uniform over the workgroup: wave_sum needs every lane)
      if (u < span.z && (RECOUNT || ((int64_t)span.y + u) * LP_TILE < live_hi)) {
        const int64_t cs = wave_sum_i64(cw[u]);
        const int64_t c = (((int64_t)span.y + u) * LP_TILE >> 8) + wave;
        if (lane == 0) {
          if (RECOUNT) d.chunk_sum[c] = cs;
          else if (cs != 0) d.chunk_sum[c] += cs;
        }
      }
    });
  }

  // ---- span partial ----
  {
    const int32_t a0 = wave_sum_i32(c_s), a1 = wave_sum_i32(c_spos), a2 = wave_sum_i32(c_r);
    const int64_t a3 = wave_sum_i64(c_sw);
    if (lane == 0) { red[wave][0] = a0; red[wave][1] = a1; red[wave][2] = a2; red[wave][3] = a3; }
  }
  __syncthreads();
  write_partial(d, tid, red, lc, lib, RECOUNT, true);
}

// ---------------------------------------------------------------------------
// Exclusive scan of a node's span weights (partial field P_SW of spans [s0, s1))
// into prefix[s]; returns the total.  One wavefront scans 64 spans per step.
// ATOMIC_IO reads the weights / writes the prefix with agent-scope atomics.
template <bool ATOMIC_IO>
__device__ int64_t node_span_scan(int64_t* partials, int64_t* prefix, int s0, int s1) {
  __shared__ int64_t total_sh;
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid < 64) {
    int64_t carry = 0;
    for (int base = s0; base < s1; base += 64) {
      const int s = base + lane;
      int64_t v = 0;
      if (s < s1) {
        int64_t* p = partials + (size_t)s * NPART + P_SW;
        v = ATOMIC_IO ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
      }
      const int64_t incl = wave_incl_scan_i64(v, lane);
      if (s < s1) {
        const int64_t ex = carry + incl - v;
        if (ATOMIC_IO) __hip_atomic_store(prefix + s, ex, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else prefix[s] = ex;
      }
      carry += __shfl(incl, 63, 64);
    }
    if (tid == 0) total_sh = carry;
  }
  __syncthreads();
  const int64_t total = total_sh;
  __syncthreads();
  return total;
}

// Per node: reduce span partials, span prefix, results rows of VD / DS / RI / SIA.
__global__ void __launch_bounds__(WG) k_node_reduce(Dev d, int mask, int recount, int list) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int n = blockIdx.x;
  const int tid = threadIdx.x;
  const size_t N = d.n_nodes, row = (size_t)t * N;
  // loads that do not depend on the span partials, issued with them
  const int sim = d.node_sim[n];
  const int s0 = d.node_span_begin[n], s1 = d.node_span_begin[n + 1];
  const int64_t ea = (list && tid < NPART) ? d.node_acc[(size_t)n * NPART + tid] : 0;
  const int64_t carry0 = tid < 3 ? d.node_carry[(size_t)n * 3 + tid] : 0;
  const int32_t pop_prev = (tid == 0 && t >= 1) ? d.r_pop[row - N + n] : 0;
  const int32_t births_n = tid == 0 ? d.births[n] : 0;
  // up to 64 spans: wave 0 holds the carried span weights for the span prefix (lane = span)
  const bool lane_spans = s1 - s0 <= 64;
  const int64_t psw = (lane_spans && tid < 64 && s0 + tid < s1) ? d.partials[(size_t)(s0 + tid) * NPART + P_SW] : 0;
  const int4 spn = (lane_spans && tid < 64 && s0 + tid < s1) ? d.spans[s0 + tid] : make_int4(0, 0, 0, 0);
  // lane tid's SpanLane record from its span's exclusive prefix
  auto span_lane = [&](int64_t prefix) {
    d.span_lane[(size_t)n * 64 + tid] = s0 + tid < s1 ? SpanLane{prefix, spn.y, spn.z} : SpanLane{INT64_MAX, 0, 0};
  };
  if (sim_stopped(d, sim, t)) return;
  const lp_sim_params& sp = d.sims[sim];
  const int32_t flags = stage_day_flags(dy.y, mask);
  if (tid == 0 && n == sp.node_begin) {
    // today's early-stop rule on yesterday's log (unless k_day_begin ran it); stopping on
    // day t leaves day t whole, so it can run at any point of the day
    const int32_t cases = d.sim_cases[sim];
    if (cases != CASES_CHECKED) day_stop_check(d, sim, t, mask, cases);
    d.sim_cases[sim] = 0;
  }
  __shared__ int64_t tally[NPART];
  __shared__ int64_t grp[WG / NPART][NPART];
  __shared__ int64_t s_w;
  // list days: the day's tallies are the accumulators of the list-day passes (k_ei,
  // k_vd, k_ri_list, k_sia), cleared here for the next day; the span partials hold only
  // the carried span weights, whose exclusive prefix (for the sampler) and total are
  // taken here.  Other days: every span's partial is reduced.
  if (list) {
    if (lane_spans) {
      if (tid < 64) {
        const int64_t incl = wave_incl_scan_i64(psw, tid);
        if (s0 + tid < s1) d.span_prefix[s0 + tid] = incl - psw;
        span_lane(incl - psw);
        if (tid == 63) s_w = incl;
      }
    } else {
      const int64_t W = node_span_scan<false>(d.partials, d.span_prefix, s0, s1);
      if (tid == 0) s_w = W;
    }
    __syncthreads();
    if (tid < NPART) {
      d.node_acc[(size_t)n * NPART + tid] = 0;
      tally[tid] = tid == P_SW ? s_w : ea;
    }
  } else {
    {
      // thread (f = tid % 32, g = tid / 32) sums field f of spans g, g+8, ...: coalesced
      const int f = tid % NPART, g = tid / NPART;
      int64_t acc = 0;
      constexpr int G = WG / NPART;
      for (int s = s0 + g; s < s1; s += 4 * G) {
        const int64_t a0 = d.partials[(size_t)s * NPART + f];
        const int64_t a1 = s + G < s1 ? d.partials[(size_t)(s + G) * NPART + f] : 0;
        const int64_t a2 = s + 2 * G < s1 ? d.partials[(size_t)(s + 2 * G) * NPART + f] : 0;
        const int64_t a3 = s + 3 * G < s1 ? d.partials[(size_t)(s + 3 * G) * NPART + f] : 0;
        acc += (a0 + a1) + (a2 + a3);
      }
      grp[g][f] = acc;
    }
    __syncthreads();
    if (tid < NPART) {
      int64_t acc = 0;
#pragma unroll
      for (int g = 0; g < WG / NPART; ++g) acc += grp[g][tid];
      tally[tid] = acc;
    }
  }
  __syncthreads();
  if (tid < 3) {
    // carried per-node counts: S, S with positive weight, R (the spans hold deltas)
    const int f = tid == 0 ? P_S : tid == 1 ? P_SPOS : P_R;
    const int64_t v = recount ? tally[f] : carry0 + tally[f];
    d.node_carry[(size_t)n * 3 + tid] = v;
    tally[f] = v;
  }
  __syncthreads();
  if (tid < NPART) d.node_tally[(size_t)n * NPART + tid] = tally[tid];
  if (tid == 0 && (flags & LP_DAY_RI)) *ri_cnt_of(d, dy.z & 1, n) = 0;   // consumed by today's RI
  if (tid < LP_MAX_STRAINS)   // infectious weight per strain as the network mix reads it
    d.beta_f[(size_t)n * LP_MAX_STRAINS + tid] = tid < d.K ? (float)tally[P_IB0 + tid] * 5.9604644775390625e-08f : 0.0f;
  if (flags & LP_DAY_TX) {
    if (list) {
      if (tid == 0) d.W_round[n] = s_w;            // the prefix was written above
    } else if (lane_spans) {
      if (tid < 64) {
        const int64_t incl = wave_incl_scan_i64(psw, tid);
        if (s0 + tid < s1) d.span_prefix[s0 + tid] = incl - psw;
        span_lane(incl - psw);
        const int64_t W = __shfl(incl, 63, 64);
        if (tid == 0) d.W_round[n] = W;
      }
    } else {
      const int64_t W = node_span_scan<false>(d.partials, d.span_prefix, s0, s1);
      if (tid == 0) d.W_round[n] = W;
    }
  }
  if (tid != 0) return;
  const size_t K4 = d.K;
  if (flags & LP_DAY_VD_STEP) {
    d.r_births[row + n] = births_n;
    d.r_deaths[row + n] = (int32_t)tally[P_DEATHS];
    d.r_pop[row + n] = pop_prev + births_n - (int32_t)tally[P_DEATHS];
  } else if (flags & LP_DAY_VD_COPY) {
    d.r_pop[row + n] = pop_prev;
  }
  if (flags & LP_DAY_DS) {
    d.r_newpp[row + n] = (int32_t)tally[P_NEWPP];
    d.r_newp[row + n] = (int32_t)tally[P_NEWP];
  }
  if (flags & LP_DAY_RI) {
    const int32_t rip = (int32_t)tally[P_RIP];
    d.r_riv[row + n] = (int32_t)tally[P_RIV];
    d.r_rip[row + n] = rip;
    d.r_ipv[row + n] = (int32_t)tally[P_IPV];
    d.r_newexp[row + n] += rip;
    d.r_newexp_bs[(row + n) * K4 + sp.ri_strain] += rip;
    d.r_ri_newexp_bs[(row + n) * K4 + sp.ri_strain] = rip;
  }
  if (flags & LP_DAY_SIA) {
    d.r_siav[row + n] = (int32_t)tally[P_SIAV];
    d.r_siap[row + n] = (int32_t)tally[P_SIAP];
    int32_t tot = 0;
    for (int k = 0; k < d.K; ++k) {
      const int32_t v = (int32_t)tally[P_SIAPK0 + k];
      d.r_newexp_bs[(row + n) * K4 + k] += v;
      d.r_sia_newexp_bs[(row + n) * K4 + k] += v;
      tot += v;
    }
    d.r_newexp[row + n] += tot;
  }
  if (recount) {
    d.potp_run[n] = (int32_t)tally[P_POTP];
    d.p_run[n] = (int32_t)tally[P_P];
  } else {
    d.potp_run[n] += (int32_t)(tally[P_NEWPP] - tally[P_DPP]);
    d.p_run[n] += (int32_t)(tally[P_NEWP] - tally[P_DP]);
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of the probes of sampling round `round` over nodes (probe_off),
// and reservation of their space in the selected list; run by one workgroup.
__device__ void plan_scan(const Dev& d, int round) {
  const int N = d.n_nodes;
  const int64_t total = block_scan_nodes(
      N, [&](int n) { return (int64_t)d.m_round[n]; },   // written by k_mix_foi (an earlier launch)
      [&](int n, int64_t off) { d.probe_off[n] = off; });
  if (threadIdx.x == 0) {
    int64_t M = total;
    const int64_t base = __hip_atomic_load(d.sel_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (base + M > d.sel_cap) { atomicOr(d.err, ERR_OVERFLOW); M = 0; }
    d.probe_off[N] = M;
    d.round_base[round] = base;
    d.sel_count[0] = base + M;
  }
}

// ---------------------------------------------------------------------------
// One wavefront per node.  The network mix sum_j N_ji beta_j is added in the
// order of a 256-lane strided sum followed by a halving tree over the 256
// partials: lane l holds the partials of lanes l, l + 64, l + 128, l + 192, folds
// them in-lane for the first two tree levels and across lanes for the rest.
// `off`: no transmission today (the node's sampling sizes are zeroed instead); a
// stopped simulation is tested after the mix's loads are issued.
// t / s, correctly rounded, for a step number s <= 64 that is the same on every lane:
// t * y with y = RN(1/s) (lane s - 1 of ystep), refined by its exact residual
// r = t - s * (t * y) (Markstein's correction).  Over every significand of t and
// s = 1 .. 64 this equals the correctly rounded t / s (tests/test_step_quotient.py); it
// holds at any exponent where the residual stays normal, which t >= 2^-90 ensures.
__device__ __forceinline__ float step_quot(float t, int s, float ystep) {
  const float y = __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ystep), s - 1));
  const float q0 = t * y;
  const float r = __builtin_fmaf(-(float)s, q0, t);
  return __builtin_fmaf(r, y, q0);
}

__device__ void mix_foi_node(const Dev& d, int i, const NodeInfo& ni, int32_t t, int lane, bool off) {
  const int K = d.K;
  const int sim = ni.sim;
  const int ns = ni.sim_nodes, li = ni.lnode;
  const float* NT = d.network_t + ni.net_off + (size_t)li * ns;
  const float4* B = reinterpret_cast<const float4*>(d.beta_f) + ni.sim_node_begin;
  // the node's own inputs, loaded ahead of the mix
  const int64_t* tl = d.node_tally + (size_t)i * NPART;
  int32_t* lcnt = d.ei_cnt + (size_t)((t + 1) & 1) * d.n_nodes + i;   // tomorrow's list (its passes ran)
  const int32_t lbase = *lcnt;
  const int64_t room = d.seg_end[i] - ni.seg_begin;
  const int32_t pop_t = d.r_pop[(size_t)t * d.n_nodes + i];
  const float season = d.season[(size_t)t * d.n_sims + sim];
  const float keep = d.keep[i], r0s = d.r0s[i];
  int64_t ibk[LP_MAX_STRAINS];
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) ibk[k] = k < K ? tl[P_IB0 + k] : 0;
  const int64_t sw = tl[P_SW], scnt = tl[P_S], spos = tl[P_SPOS];
  // lane l's partial of group v sums sources j = l + 64 v + 256 r for r = 0, 1, ...;
  // the sources below MIX_PRE * 64 are loaded together first
  constexpr int MIX_PRE = 16;
  float ntv[MIX_PRE];
  float4 bv[MIX_PRE];
#pragma unroll
  for (int m = 0; m < MIX_PRE; ++m) {
    const int j = lane + 64 * m;
    ntv[m] = j < ns ? NT[j] : 0.0f;
    bv[m] = j < ns ? B[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  float acc[4][LP_MAX_STRAINS];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
#pragma unroll
    for (int k = 0; k < LP_MAX_STRAINS; ++k) acc[v][k] = 0.0f;
  }
#pragma unroll
  for (int m = 0; m < MIX_PRE; ++m) {
    const int v = m & 3;
    if (lane + 64 * m < ns) {
      acc[v][0] = __builtin_fmaf(ntv[m], bv[m].x, acc[v][0]);
      acc[v][1] = __builtin_fmaf(ntv[m], bv[m].y, acc[v][1]);
      acc[v][2] = __builtin_fmaf(ntv[m], bv[m].z, acc[v][2]);
      acc[v][3] = __builtin_fmaf(ntv[m], bv[m].w, acc[v][3]);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    for (int j = lane + 64 * v + 64 * MIX_PRE; j < ns; j += 256) {
      const float nt = NT[j];
      const float4 b = B[j];
      acc[v][0] = __builtin_fmaf(nt, b.x, acc[v][0]);
      acc[v][1] = __builtin_fmaf(nt, b.y, acc[v][1]);
      acc[v][2] = __builtin_fmaf(nt, b.z, acc[v][2]);
      acc[v][3] = __builtin_fmaf(nt, b.w, acc[v][3]);
    }
  }
  float part[LP_MAX_STRAINS];
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) {
    float x = (acc[0][k] + acc[2][k]) + (acc[1][k] + acc[3][k]);   // tree levels 128 and 64
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) x = x + __shfl_down(x, w, 64);
    part[k] = __shfl(x, 0, 64);
  }
  if (d.mix_buf) {                 // mix_mode 1: the contraction came from k_mix_mfma
#pragma unroll
    for (int k = 0; k < LP_MAX_STRAINS; ++k) part[k] = d.mix_buf[(size_t)i * LP_MAX_STRAINS + k];
  }
  if (off || sim_stopped(d, sim, t)) {
    if (lane == 0) { d.m_round[i] = 0; d.size[i] = 0; d.take_all[i] = 0; }
    return;
  }
  // node-level map, then the count draw with chunks spread over lanes
  const float popc = (float)(pop_t > 1 ? pop_t : 1);
  float pk[LP_MAX_STRAINS] = {0.0f, 0.0f, 0.0f, 0.0f};
  float ptot = 0.0f;
  bool local_zero = true;
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) {
    if (k >= K) break;
    const float bi = (float)ibk[k] * 5.9604644775390625e-08f;
    const float bm = __builtin_fmaf(bi, keep, part[k]);
    const float b2 = (bm * season) * r0s;
    const float rate = b2 / popc;
    float p = -expm1f_lp(-rate);
    if (!(p > 0.0f)) p = 0.0f;
    pk[k] = p;
    ptot = ptot + p;
    if (ibk[k] != 0) local_zero = false;
    if (lane == 0) { d.p_dbg[(size_t)i * K + k] = p; d.beta_mixed[(size_t)i * K + k] = bm; }
  }
  const float expo = (float)sw * 5.9604644775390625e-08f;
  const float ex = expo * ptot;
  // mode 0: none, 1: Poisson(ex), 2: NB
  int mode = 0;
  if (ex > 0.0f) {
    if (!local_zero) mode = 1;
    else if (ni.zi < 1.0f) {
      const uint4 o = philox(0xFFFFFFFFu, (uint32_t)li, (uint32_t)t, ST_COUNT, ni.seed);
      if ((uint64_t)o.x >= ni.thr_zi) mode = 2;
    }
  }
  int64_t mine = 0;
  // Each pass inverts one chunk per lane.  Lanes step together, so step s divides by the
  // same s on every lane still stepping: steps 1 .. 64 take RN(1/s) from lane s - 1 of
  // ystep (step_quot), later steps divide.  A stepping lane's numerator stays >= 2^-90:
  // the last term added was >= ulp(F) / 2 >= 2^-48 (F >= e^-16), times lc >= 2^-25 (a
  // smaller chunk has F = e^-lc = 1 and never steps; the quick path is kept off for a
  // node whose last chunk is that small) or times (k + rc) omp >= omp >= 2^-40.
  if (mode == 1) {
    const float ystep = 1.0f / (float)(lane + 1);
    const float nf = floorf(ex / 16.0f);
    const int64_t n_full = (int64_t)nf;
    const float last = ex - 16.0f * nf;
    const int64_t nch = n_full + (last > 0.0f ? 1 : 0);
    const bool quick = !(last > 0.0f && last < 0x1p-25f);
    for (int64_t j0 = 0; j0 < nch; j0 += 64) {
      const int64_t j = j0 + lane;
      float lc = 0.0f, u = 0.0f, p = 0.0f, F = 0.0f;
      bool go = false;
      if (j < nch) {
        lc = j < n_full ? 16.0f : last;
        const uint4 o = philox((uint32_t)j, (uint32_t)li, (uint32_t)t, ST_COUNT, ni.seed);
        u = u24(o.x);
        p = expf_lp(-lc);
        F = p;
        go = u >= F;
      }
      int k = 0;
      if (quick) {
        for (int st = 1; st <= 64 && __builtin_amdgcn_ballot_w64(go) != 0; ++st) {
          const float q = step_quot(p * lc, st, ystep);
          const float Fn = F + q;
          k = go ? st : k;
          p = q;
          const bool more = go && Fn != F;
          F = more ? Fn : F;
          go = more && u >= Fn;
        }
      }
      while (go) {
        k += 1;
        p = (p * lc) / (float)k;
        const float Fn = F + p;
        if (Fn == F) break;
        F = Fn;
        go = u >= F;
      }
      mine += k;
    }
  } else if (mode == 2) {
    const float dm = ex / (1.0f - ni.zi);
    const float q = dm / (float)ni.disp_r;
    const float lq = log1pf_lp(q);
    if (lq > 0.0f) {
      const float omp = q / (1.0f + q);
      int32_t rc_max = (int32_t)(16.0f / lq);
      if (rc_max < 1) rc_max = 1;
      const int32_t n_full = ni.disp_r / rc_max, rest = ni.disp_r % rc_max;
      const int32_t nch = n_full + (rest > 0 ? 1 : 0);
      // (float)(k + rc) in 32 bits equals the 64-bit conversion while k + rc < 2^31
      const bool quick = omp >= 0x1p-40f && rc_max <= (1 << 24);
      const float ystep = 1.0f / (float)(lane + 1);
      for (int32_t j0 = 0; j0 < nch; j0 += 64) {
        const int32_t j = j0 + lane;
        int32_t rc = 0;
        float u = 0.0f, P = 0.0f, F = 0.0f;
        bool go = false;
        if (j < nch) {
          rc = j < n_full ? rc_max : rest;
          const uint4 o = philox((uint32_t)(j + 1), (uint32_t)li, (uint32_t)t, ST_COUNT, ni.seed);
          u = u24(o.x);
          P = expf_lp(-((float)rc * lq));
          F = P;
          go = u >= F;
        }
        int64_t k = 0;
        if (quick) {
          for (int st = 1; st <= 64 && __builtin_amdgcn_ballot_w64(go) != 0; ++st) {
            const float q1 = step_quot(P * ((float)(st - 1 + rc) * omp), st, ystep);
            const float Fn = F + q1;
            k = go ? st : k;
            P = q1;
            const bool more = go && Fn != F;
            F = more ? Fn : F;
            go = more && u >= Fn;
          }
        }
        while (go) {
          P = (P * ((float)(k + rc) * omp)) / (float)(k + 1);
          k += 1;
          const float Fn = F + P;
          if (Fn == F) break;
          F = Fn;
          go = u >= F;
        }
        mine += k;
      }
    }
  }
  const int64_t total = wave_sum_i64(mine);
  if (lane != 0) return;
  d.expected_dbg[i] = ex;
  d.total_dbg[i] = total;
  const int32_t sz = (int32_t)(total < scnt ? total : scnt);
  const int32_t ta = (sz > 0 && (int64_t)sz >= spos) ? 1 : 0;
  d.size[i] = sz;
  d.take_all[i] = ta;
  const int32_t m = ta ? 0 : sz;
  d.m_round[i] = m;
  // positions for the round's probes in tomorrow's list (k_sample), when the segment
  // also holds every later exposure of the node (at most sz more, appended)
  if ((int64_t)lbase + m + sz <= room) {
    *lcnt = lbase + m;
    d.resv[i] = lbase;
  } else {
    d.resv[i] = -1;
  }
  float cum = 0.0f;
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) {
    if (k >= K) break;
    cum = cum + (ptot > 0.0f ? pk[k] / ptot : 0.0f);
    d.thr_strain[(size_t)i * LP_MAX_STRAINS + k] = thr_from_prob(cum);
  }
}

// mix_mode 1: every simulation's network mix as a matrix product on MFMA,
// mix[i][k] = sum_j NT[i][j] beta[j][k] with NT the simulation's network_t block (row i:
// the weights into node i) and beta the nodes' infectious weight per strain, in float32.
// One wavefront per (simulation, 16 rows): v_mfma_f32_16x16x4_f32 over k-steps of 4
// sources, A = NT[i0 .. i0+15][j0 .. j0+3] (lane i + 16 k holds A[i][k]), B = beta[j0 ..
// j0+3][0 .. 15] with strains in columns 0-3 and zeros beyond (lane j + 16 k holds
// B[k][j]), D[i][j] in lane j + 16 (i / 4), element i % 4.  The accumulation order is the
// matrix unit's, not the oracle's fixed order.
__global__ void __launch_bounds__(64) k_mix_mfma(Dev d) {
  const int sim = blockIdx.x, tile = blockIdx.y, lane = threadIdx.x;
  const lp_sim_params& sp = d.sims[sim];
  const int n = sp.n_nodes, i0 = tile * 16;
  if (i0 >= n) return;
  const float* NT = d.network_t + sp.net_off;
  const float* B = d.beta_f + (size_t)sp.node_begin * LP_MAX_STRAINS;
  const int ai = i0 + (lane & 15), kk = lane >> 4, bj = lane & 15;
  typedef float v4f __attribute__((ext_vector_type(4)));
  // eight accumulators over interleaved k-steps, summed as a tree at the end: the
  // sequential depth of each sum is n / 32, not n / 4 (float32 error ~ sqrt of it)
  constexpr int NACC = 8;
  v4f acc[NACC];
#pragma unroll
  for (int q = 0; q < NACC; ++q) acc[q] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  for (int j0 = 0; j0 < n; j0 += 4 * NACC) {
#pragma unroll
    for (int q = 0; q < NACC; ++q) {
      const int j = j0 + 4 * q + kk;
      const float a = (ai < n && j < n) ? NT[(size_t)ai * n + j] : 0.0f;
      const float b = (bj < LP_MAX_STRAINS && j < n) ? B[(size_t)j * LP_MAX_STRAINS + bj] : 0.0f;
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
    }
  }
  const v4f s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  if (bj < LP_MAX_STRAINS) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = i0 + 4 * kk + r;
      if (row < n) d.mix_buf[(size_t)(sp.node_begin + row) * LP_MAX_STRAINS + bj] = s[r];
    }
  }
}

// Per node: network mix, FOI, count draw.  One wavefront per node.  MINW = 4 keeps at
// least 4 waves per SIMD (<= 128 VGPRs; one 8-byte value spills, stored at entry and
// reloaded once), for more nodes than are resident at once, where the network reads need
// the occupancy; MINW = 1 leaves the kernel unspilled for fewer nodes.
template <int MINW>
__global__ void __launch_bounds__(WG, MINW) k_mix_foi(Dev d, int mask) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int i = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i < d.n_nodes) {
    const int32_t flags = stage_day_flags(dy.y, mask);
    mix_foi_node(d, i, d.ninfo[i], t, lane, !(flags & LP_DAY_TX));
  }
}

// One workgroup: plan sampling round `round` (probe offsets over nodes, list space).
__global__ void __launch_bounds__(WG) k_plan(Dev d, int round) { plan_scan(d, round); }


// ---------------------------------------------------------------------------
// Infection sampling (model.py:1089-1122): rounds of `need` with-replacement
// probes proportional to the remaining selection weights; the distinct agents
// hit are exposed.  Round 1 runs grid-wide (k_sample, with k_plan first above
// LDS_OFF_MAX - 1 nodes); k_finish then takes each node on its own: round 1's
// agents become E, then its later rounds, and the take-all nodes.

__device__ __forceinline__ int64_t find_node(const int64_t* off, int N, int64_t p) {
  int a = 0, b = N;   // largest n in [0, N) with off[n] <= p < off[n+1]
  while (b - a > 1) { const int m = (a + b) >> 1; if (off[m] <= p) a = m; else b = m; }
  return a;
}

template <bool ATOMIC_IO>
__device__ __forceinline__ int64_t ld64(const int64_t* p) {
  return ATOMIC_IO ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
}

// The agent hit by target T inside the span of `ntiles` tiles from tile `tile0`,
// whose prefix (weight before it) is sp_base: the chunk by a wave scan of the span's
// chunk weights, the agent by a wave scan of the chunk's agent weights (S agents,
// state 0, weigh).  ATOMIC_IO: the weights and states are read with atomic loads
// (other wavefronts change them during the kernel).  MARKED_AS_S: an agent claimed
// earlier in the round (SEL_MARK) keeps its weight, so that every probe of a round
// sees the weights of the round's start.  Returns the slot, or -1 if the weights were
// inconsistent; *word (when given) receives the hit agent's state word as read here.
template <bool ATOMIC_IO, bool MARKED_AS_S>
__device__ int64_t resolve_in_tiles(const Dev& d, int64_t tile0, int ntiles, int64_t sp_base, int64_t T, int lane,
                                    uint32_t* word = nullptr) {
  const int64_t c0 = tile0 * (LP_TILE / LP_CHUNK);
  const int nc = ntiles * (LP_TILE / LP_CHUNK);
  const int64_t cw = lane < nc ? ld64<ATOMIC_IO>(d.chunk_sum + c0 + lane) : 0;
  const int64_t cincl = wave_incl_scan_i64(cw, lane);
  const uint64_t cb = __ballot(lane < nc && sp_base + cincl > T);
  if (cb == 0) return -1;
  const int Lc = __ffsll((unsigned long long)cb) - 1;
  const int64_t cp = sp_base + __shfl(cincl, Lc, 64) - __shfl(cw, Lc, 64);
  const int64_t c = c0 + Lc;
  const int64_t i0 = c * LP_CHUNK + (int64_t)lane * 4;
  uint32_t st4;
  if (ATOMIC_IO || MARKED_AS_S)
    st4 = __hip_atomic_load(reinterpret_cast<const uint32_t*>(d.state) + (i0 >> 2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    st4 = reinterpret_cast<const uint32_t*>(d.state)[i0 >> 2];
  const float4 rk = *reinterpret_cast<const float4*>(d.risk + i0);
  int64_t w[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const uint32_t b = (st4 >> (8 * a)) & 0xffu;
    const bool s_like = b == 0 || (MARKED_AS_S && b == (uint32_t)(uint8_t)SEL_MARK);
    w[a] = s_like ? weight_q24(lane_of(rk, a)) : 0;
  }
  const int64_t lsum = w[0] + w[1] + w[2] + w[3];
  const int64_t incl = wave_incl_scan_i64(lsum, lane);
  int64_t cum = cp + (incl - lsum);
  int hit = -1;
  for (int a = 0; a < 4; ++a) { cum += w[a]; if (hit < 0 && cum > T) hit = a; }
  const uint64_t hb = __ballot(hit >= 0);
  if (hb == 0) return -1;
  const int L = __ffsll((unsigned long long)hb) - 1;
  const int ha = __shfl(hit, L, 64);
  if (word) *word = __shfl(st4, L, 64);
  return c * LP_CHUNK + (int64_t)L * 4 + ha;
}

// resolve_in_tiles over span s.
template <bool ATOMIC_IO, bool MARKED_AS_S>
__device__ __forceinline__ int64_t resolve_in_span(const Dev& d, int64_t s, int64_t sp_base, int64_t T, int lane,
                                                   uint32_t* word = nullptr) {
  const int4 span = d.spans[s];
  return resolve_in_tiles<ATOMIC_IO, MARKED_AS_S>(d, span.y, span.z, sp_base, T, lane, word);
}


// One wavefront per probe: target T in [0, W); the span with prefix <= T by a
// 64-ary search over the node's span prefix, the chunk inside the span by a wave
// scan of its chunk weights, the agent inside the chunk by a wave scan of the
// agents' weights (ATOMIC_IO, MARKED_AS_S as for resolve_in_tiles).  Returns the
// slot, or -1 if the weights were inconsistent.
template <bool ATOMIC_IO, bool MARKED_AS_S>
__device__ int64_t resolve_probe(const Dev& d, int node, int64_t T, const int64_t* prefix, int lane,
                                 uint32_t* word = nullptr) {
  int64_t lo = d.node_span_begin[node];
  int64_t hi = d.node_span_begin[node + 1];   // search spans [lo, hi)
  while (hi - lo > 1) {
    const int64_t range = hi - lo;
    const int64_t step = (range + 63) / 64;
    const int64_t c = lo + (int64_t)lane * step;
    const int64_t v = (c < hi) ? ld64<ATOMIC_IO>(prefix + c) : INT64_MAX;
    const uint64_t ok = __ballot(v <= T);        // monotone: lane 0 always true
    const int L = 63 - __clzll(ok);
    lo = lo + (int64_t)L * step;
    const int64_t nhi = lo + step;
    hi = nhi < hi ? nhi : hi;
  }
  const int64_t s = lo;
  return resolve_in_span<ATOMIC_IO, MARKED_AS_S>(d, s, ld64<ATOMIC_IO>(prefix + s), T, lane, word);
}

// Claim a hit: state byte 0 -> SEL_MARK with a word CAS; false if the agent was
// already claimed (a duplicate probe) or is not susceptible.
__device__ __forceinline__ bool claim(int8_t* state, int64_t slot) {
  uint32_t* w = reinterpret_cast<uint32_t*>(state) + (slot >> 2);
  const int sh = 8 * (int)(slot & 3);
  uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    if (((old >> sh) & 0xffu) != 0) return false;
    const uint32_t nw = (old & ~(0xffu << sh)) | ((uint32_t)(uint8_t)SEL_MARK << sh);
    const uint32_t prev = atomicCAS(w, old, nw);
    if (prev == old) return true;
    old = prev;
  }
}

// claim() starting from the state word `seen` (the word read when the agent was
// resolved, its byte then S): the first compare-and-swap needs no load before it,
// and a word changed since costs one more attempt with the value it returns.
__device__ __forceinline__ bool claim_from(int8_t* state, int64_t slot, uint32_t seen) {
  uint32_t* w = reinterpret_cast<uint32_t*>(state) + (slot >> 2);
  const int sh = 8 * (int)(slot & 3);
  uint32_t old = seen & ~(0xffu << sh);
  while (true) {
    const uint32_t nw = old | ((uint32_t)(uint8_t)SEL_MARK << sh);
    const uint32_t prev = atomicCAS(w, old, nw);
    if (prev == old) return true;
    if (((prev >> sh) & 0xffu) != 0) return false;
    old = prev;
  }
}

// Remove an exposed agent's weight from its chunk and its span.
__device__ __forceinline__ void unweight(const Dev& d, int node, int64_t slot) {
  unweight_in(d, span_of(d, node, slot), slot, weight_q24(d.risk[slot]));
}

// The strain of a selected agent, drawn by cumulative strain probability
// (model.py:1125-1147) against the node's thresholds thr[0..K) (already loaded).
__device__ __forceinline__ int strain_draw(const Dev& d, const uint64_t* thr, uint64_t seed, int64_t slot_base,
                                           int64_t slot, int32_t t) {
  const uint4 o = draw(seed, (uint64_t)(slot - slot_base), t, ST_STRAIN);
  int k = 0;
#pragma unroll
  for (int s = LP_MAX_STRAINS - 1; s >= 0; --s)
    if (s < d.K && (uint64_t)o.x < thr[s]) k = s;   // the first strain whose threshold exceeds the draw
  return k;
}

// S -> E of a selected agent with its drawn strain; returns the strain (the caller
// counts it).
__device__ __forceinline__ int expose_agent_thr(const Dev& d, const uint64_t* thr, uint64_t seed, int64_t slot_base,
                                                int64_t slot, int32_t t) {
  const int k = strain_draw(d, thr, seed, slot_base, slot, t);
  d.state[slot] = 1;
  *rec_byte(d, slot, RB_STRAIN) = (int8_t)k;
  return k;
}

constexpr uint64_t SEL_WON = 1ull << 63;   // round-1 entry whose probe claimed its agent

// Sampling round 1, grid-wide: one wavefront per probe resolves it to an agent, and
// the first probe of each agent hit claims it (the round's distinct agents, counted
// per node in n_uniq).  Every probe of the round must see the weights of the round's
// start, so a claimed agent keeps its weight and reads as S (state SEL_MARK) until
// k_finish makes it E and removes its weight; its strain, record byte and entry for
// tomorrow's list are written here.  Probe j of node n owns position resv[n] + j of
// the node's list for tomorrow (reserved by k_mix_foi): the agent's entry, or a hole
// (state 0, skipped by every reader of the lists) when the probe lost the claim; a
// node without a reservation (resv -1) appends with atomics instead.  When the probe
// offsets fit in LDS (dynamic shared memory of (n_nodes + 1) * 8 bytes) every
// workgroup scans the nodes' round-1 probe counts itself and workgroup 0 publishes the
// plan (round 1 is the day's first reservation in the selected list: base 0);
// otherwise k_plan planned the round.
constexpr int LDS_OFF_MAX = 8192;
__global__ void __launch_bounds__(WG) k_sample(Dev d) {
  extern __shared__ int64_t s_off[];
  const int N = d.n_nodes;
  const bool lds = N + 1 <= LDS_OFF_MAX;
  int64_t M;
  if (lds) {
    const bool pub = blockIdx.x == 0;
    const int64_t total = block_scan_nodes(
        N, [&](int n) { return (int64_t)d.m_round[n]; },   // written by k_mix_foi (an earlier launch)
        [&](int n, int64_t off) {
          s_off[n] = off;
          if (pub) d.probe_off[n] = off;
        });
    M = total > d.sel_cap ? 0 : total;
    if (threadIdx.x == 0) s_off[N] = M;
    if (pub && threadIdx.x == 0) {
      if (total > d.sel_cap) atomicOr(d.err, ERR_OVERFLOW);
      d.probe_off[N] = M;
      d.round_base[1] = 0;
      d.sel_count[0] = M;
    }
    __syncthreads();
  } else {
    M = d.probe_off[N];
  }
  const int64_t nw = (int64_t)gridDim.x * (WG / 64);
  const int64_t first = (int64_t)blockIdx.x * (WG / 64);
  if (first >= M) return;
  const int64_t* off = lds ? s_off : d.probe_off;
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int nxt = (t + 1) & 1;
  const int lane = threadIdx.x & 63;
  const int64_t base = lds ? 0 : d.round_base[1];
  // each wave takes a contiguous range of probes (probes are ordered by node); the
  // node's seed, weight and span prefix (one span per lane, up to 64 spans) are
  // loaded once per node within a group of 64 probes, so a probe costs the
  // chunk-weight and the agent loads
  const int64_t per = (M + nw - 1) / nw;
  const int64_t p0 = (first + (threadIdx.x >> 6)) * per;
  const int64_t p1 = p0 + per < M ? p0 + per : M;
  // groups of up to 64 probes: resolved one at a time by the whole wave into the wave's
  // LDS row (entry q: probe g0 + q's agent and node, its state word and index in the
  // node), then every lane claims its own at once.  The node's data is reloaded per
  // group and the row is in LDS, so neither is held in registers through the other
  // phase (more wavefronts resident)
  __shared__ uint4 s_res[WG / 64][64];
  uint4* res = s_res[threadIdx.x >> 6];
  for (int64_t g0 = p0; g0 < p1; g0 += 64) {
    const int64_t g1 = g0 + 64 < p1 ? g0 + 64 : p1;
    int node = -1, ns = 0, stile = 0, sntile = 0;   // lane l: span l of the node (first tile, tiles)
    int64_t node_lo = 0, node_hi = 0, W = 0, spx = INT64_MAX;
    uint64_t seed = 0;
    uint32_t lnode = 0;
    for (int64_t p = g0; p < g1; ++p) {
      if (node < 0 || p >= node_hi) {
        node = __builtin_amdgcn_readfirstlane((int)find_node(off, N, p));
        node_lo = off[node];
        node_hi = off[node + 1];
        const NodeInfo& ni = d.ninfo[node];
        const SpanLane sl = d.span_lane[(size_t)node * 64 + lane];   // beside the NodeInfo loads
        seed = ni.seed;
        lnode = (uint32_t)ni.lnode;
        W = d.W_round[node];
        ns = ni.n_spans;
        spx = ns <= 64 ? sl.prefix : INT64_MAX;
        stile = sl.tile0;
        sntile = sl.ntiles;
      }
      const int64_t j = p - node_lo;
      const uint4 o = philox((uint32_t)j, lnode, (uint32_t)t, ST_PROBE | (1u << 8), seed);
      const uint64_t u = (uint64_t)o.x | ((uint64_t)o.y << 32);
      const int64_t T = (int64_t)mulhi64(u, (uint64_t)(W > 0 ? W : 0));
      int64_t slot = -1;
      uint32_t word = 0;
      if (W > 0) {
        if (ns <= 64) {
          const uint64_t ok = __ballot(spx <= T);      // monotone over the node's spans
          if (ok) {
            const int L = 63 - __clzll((unsigned long long)ok);
            slot = resolve_in_tiles<false, true>(d, __shfl(stile, L, 64), __shfl(sntile, L, 64), __shfl(spx, L, 64),
                                                 T, lane, &word);
          }
        } else {
          slot = resolve_probe<false, true>(d, node, T, d.span_prefix, lane, &word);
        }
      }
      if (lane == 0) {
        const uint64_t e = (slot < 0 ? SEL_SLOT : (uint64_t)slot) | ((uint64_t)node << 40);
        res[p - g0] = make_uint4((uint32_t)e, (uint32_t)(e >> 32), word, (uint32_t)j);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int64_t my_slot = -1, my_j = 0;
    uint32_t my_word = 0;
    int my_node = -1;
    if (g0 + lane < g1) {
      const uint4 r = res[lane];
      const uint64_t e = (uint64_t)r.x | ((uint64_t)r.y << 32);
      my_slot = (e & SEL_SLOT) == SEL_SLOT ? -1 : (int64_t)(e & SEL_SLOT);
      my_node = (int)(e >> 40);
      my_word = r.z;
      my_j = r.w;
    }
    // the group's claims, lane q for probe g0 + q
    const int64_t p = g0 + lane;
    int k = -1;
    bool won = false;
    uint4 rec = make_uint4(0u, 0u, 0u, 0u);
    float r0v[LP_MAX_STRAINS] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint4* at = nullptr;   // this probe's position in tomorrow's list, when reserved
    if (p < g1) {
      const NodeInfo& ni = d.ninfo[my_node];
      const int32_t rb = d.resv[my_node];
      if (rb >= 0) at = d.ei_list + (size_t)nxt * (size_t)d.capacity + ni.seg_begin + rb + my_j;
      int64_t w = 0;
      if (my_slot >= 0) {
        // everything the exposure reads, loaded beside the claim
        uint64_t thr[LP_MAX_STRAINS];
#pragma unroll
        for (int q = 0; q < LP_MAX_STRAINS; ++q) {
          thr[q] = d.thr_strain[(size_t)my_node * LP_MAX_STRAINS + q];
          r0v[q] = ni.r0k[q];
        }
        rec = d.rec[my_slot];   // the entry's record fields do not change when an agent is exposed
        w = weight_q24(d.risk[my_slot]);
        if (claim_from(d.state, my_slot, my_word)) {   // the first probe of an agent claims it
          won = true;
          k = strain_draw(d, thr, ni.seed, ni.slot_base, my_slot, t);
          *rec_byte(d, my_slot, RB_STRAIN) = (int8_t)k;
        }
      } else {
        atomicOr(d.err, ERR_SAMPLING);
      }
      d.sel_list[base + p] = (my_slot < 0 ? SEL_SLOT : (uint64_t)my_slot) | ((uint64_t)my_node << 40) | (won ? SEL_WON : 0);
      if (won) d.sel_word[base + p] = w < (int64_t)0xFFFFFFFF ? (uint32_t)w : 0xFFFFFFFFu;   // k_finish removes it
    }
    const int wn = won ? my_node : -1;
    const uint4 ent = won ? exposed_entry_rec(my_slot, rec, k, t, r0_sel(r0v, k)) : make_uint4(0u, 0u, 0u, 0u);
    if (at) *at = ent;
    // new exposures by strain: one atomic per strain when the group's winners share a node
    const int n0 = __shfl(wn, __ffsll((unsigned long long)__ballot(wn >= 0)) - 1, 64);
    const bool same = __ballot(wn >= 0 && wn != n0) == 0;
    if (same) {
      const uint64_t any = __ballot(wn >= 0);
      if (any) {
        for (int s = 0; s < d.K; ++s) {
          const int c = __popcll(__ballot(k == s));
          if (lane == 0 && c) atomicAdd(d.newexp + n0 * LP_MAX_STRAINS + s, c);
        }
        if (lane == 0) atomicAdd(d.n_uniq + n0, __popcll(any));
      }
    } else if (wn >= 0) {
      atomicAdd(d.newexp + wn * LP_MAX_STRAINS + k, 1);
      atomicAdd(d.n_uniq + wn, 1);
    }
    if (won && !at) append_entry(d, nxt, wn, ent);
  }
}

// Per node: take-all nodes (size >= #positive-weight susceptibles) expose every
// such agent; otherwise the node's later rounds (2, 3, ...) run here until
// `size` distinct agents are exposed or no weight is left.  A round's probes are
// kept in LDS up to FIN_LDS of them, else in the selected list.  The workgroup's
// threads pass values between the phases of a round only through LDS and through
// L2 (agent-scope atomics, atomic loads and stores), so __syncthreads (workgroup-
// scope release / acquire) orders them: a device-scope fence would also write back
// the XCD's whole L2, at a cost of microseconds.
constexpr int FIN_LDS = 2048;
__global__ void __launch_bounds__(WG) k_finish(Dev d) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const int n = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int64_t sh_base;
  __shared__ uint64_t s_sel[FIN_LDS];        // a round's selected entries when they fit
  __shared__ int32_t s_ne[LP_MAX_STRAINS];   // the node's new exposures by strain (one global add each)
  __shared__ int32_t s_claimed;
  // most nodes have nothing left after round 1: one round of loads decides
  const int32_t size = d.size[n], take_all = d.take_all[n], uniq0 = d.n_uniq[n], m1 = d.m_round[n];
  const bool later = size > 0 && (take_all || size - uniq0 > 0);
  const NodeInfo& ni = d.ninfo[n];
  // (a round 1 that overflowed the selected list ran no probes: probe_off[N] = 0)
  if (m1 > 0 && d.probe_off[d.n_nodes] > 0) {
    // round 1's winners (claimed by k_sample, state SEL_MARK): E now, and their weights
    // leave the chunk and span sums before any later round reads them
    // eight entries per thread with their loads in flight together; up to 64 spans the
    // span weights are summed in LDS and leave with one atomic per span
    __shared__ int64_t s_spw[64];
    const bool lds_spans = ni.n_spans <= 64;
    if (tid < 64) s_spw[tid] = 0;
    __syncthreads();
    const int64_t b1 = d.round_base[1] + d.probe_off[n];
    const int64_t tile0 = ni.seg_begin / LP_TILE;
    constexpr int U = 8;
    for (int j0 = 0; j0 < m1; j0 += U * WG) {
      uint64_t e[U];
      uint32_t wq[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * WG + tid;
        e[u] = j < m1 ? d.sel_list[b1 + j] : 0ull;
        wq[u] = j < m1 ? d.sel_word[b1 + j] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!(e[u] & SEL_WON)) continue;
        const int64_t slot = (int64_t)(e[u] & SEL_SLOT);
        const int64_t w = wq[u] != 0xFFFFFFFFu ? (int64_t)wq[u] : weight_q24(d.risk[slot]);
        d.state[slot] = 1;
        const unsigned long long dw = (unsigned long long)(-w);
        atomicAdd(reinterpret_cast<unsigned long long*>(d.chunk_sum + (slot >> 8)), dw);
        const int ls = (int)((slot / LP_TILE - tile0) >> d.tpw_sh);   // the node's span
        if (lds_spans) atomicAdd(reinterpret_cast<unsigned long long*>(&s_spw[ls]), dw);
        else atomicAdd(reinterpret_cast<unsigned long long*>(d.partials + (size_t)(ni.span_begin + ls) * NPART + P_SW), dw);
      }
    }
    __syncthreads();
    if (lds_spans && tid < ni.n_spans && s_spw[tid] != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(d.partials + (size_t)(ni.span_begin + tid) * NPART + P_SW),
                (unsigned long long)s_spw[tid]);
    if (!later) return;
    __syncthreads();   // the weights and states before the later rounds
  } else if (!later) {
    return;
  }
  if (tid < LP_MAX_STRAINS) s_ne[tid] = 0;
  if (tid == 0) s_claimed = 0;
  uint64_t thr[LP_MAX_STRAINS];   // the node's cumulative strain thresholds, loaded once
  float r0v[LP_MAX_STRAINS];
#pragma unroll
  for (int q = 0; q < LP_MAX_STRAINS; ++q) {
    thr[q] = d.thr_strain[(size_t)n * LP_MAX_STRAINS + q];
    r0v[q] = ni.r0k[q];
  }
  __syncthreads();
  const int nxt = (t + 1) & 1;
  if (take_all) {
    const int64_t lo = ni.seg_begin, hi = d.live_end[n];
    for (int64_t i0 = lo; i0 < hi; i0 += WG) {         // uniform trip count: whole waves append
      const int64_t i = i0 + tid;
      bool ex = false;
      int kx = 0;
      uint4 rec = make_uint4(0u, 0u, 0u, 0u);
      if (i < hi && d.state[i] == 0 && weight_q24(d.risk[i]) > 0) {
        rec = d.rec[i];
        unweight(d, n, i);
        kx = expose_agent_thr(d, thr, ni.seed, ni.slot_base, i, t);
        atomicAdd(&s_ne[kx], 1);
        ex = true;
      }
      wave_append_entry(d, nxt, n, ex ? exposed_entry_rec(i, rec, kx, t, r0_sel(r0v, kx)) : make_uint4(0u, 0u, 0u, 0u),
                        ex, lane);
    }
  } else {
    const int s0 = ni.span_begin, s1 = ni.span_begin + ni.n_spans;
    // up to 64 spans: lane l of every wave holds span l's tile range and, each round,
    // its prefix of the remaining weights (a wave scan of the span partials)
    const bool lane_spans = ni.n_spans <= 64;
    const bool my_span = lane_spans && lane < ni.n_spans;
    const int4 sp4 = my_span ? d.spans[s0 + lane] : make_int4(0, 0, 0, 0);
    for (int round = 2;; ++round) {
      if (round > MAX_ROUNDS) { if (tid == 0) atomicOr(d.err, ERR_SAMPLING); break; }
      const int32_t need = size - (uniq0 + s_claimed);
      if (need <= 0) break;
      int64_t W, spx = INT64_MAX;
      if (lane_spans) {
        const int64_t v = my_span ? __hip_atomic_load(d.partials + (size_t)(s0 + lane) * NPART + P_SW, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0;
        const int64_t incl = wave_incl_scan_i64(v, lane);
        W = __shfl(incl, 63, 64);
        if (my_span) spx = incl - v;
      } else {
        W = node_span_scan<true>(d.partials, d.fin_prefix, s0, s1);
      }
      if (W <= 0) break;
      // a round of up to FIN_LDS probes keeps them in LDS; a larger one reserves space in
      // the selected list (uniform over the workgroup)
      const bool in_lds = need <= FIN_LDS;
      if (tid == 0) {
        int64_t b = 0;
        if (!in_lds) {
          b = atomicAdd(reinterpret_cast<unsigned long long*>(d.sel_count), (unsigned long long)need);
          if (b + need > d.sel_cap) {
            atomicOr(d.err, ERR_OVERFLOW);
            b = -1;
          }
        }
        sh_base = b;
      }
      __syncthreads();
      const int64_t base = sh_base;
      if (base < 0) break;
      uint64_t* gsel = d.sel_list + base;
      for (int j = wave; j < need; j += WG / 64) {
        const uint4 o = philox((uint32_t)j, (uint32_t)ni.lnode, (uint32_t)t,
                               ST_PROBE | ((uint32_t)round << 8), ni.seed);
        const uint64_t u = (uint64_t)o.x | ((uint64_t)o.y << 32);
        const int64_t T = (int64_t)mulhi64(u, (uint64_t)W);
        int64_t slot = -1;
        if (lane_spans) {
          const uint64_t ok = __ballot(spx <= T);         // monotone over the node's spans
          if (ok) {
            const int L = 63 - __clzll((unsigned long long)ok);
            slot = resolve_in_tiles<true, false>(d, __shfl(sp4.y, L, 64), __shfl(sp4.z, L, 64), __shfl(spx, L, 64), T,
                                          lane);
          }
        } else {
          slot = resolve_probe<true, false>(d, n, T, d.fin_prefix, lane);
        }
        if (lane == 0) {
          const uint64_t e = slot < 0 ? SEL_INVALID : ((uint64_t)slot | ((uint64_t)n << 40));
          if (slot < 0) atomicOr(d.err, ERR_SAMPLING);
          if (in_lds) s_sel[j] = e;
          else __hip_atomic_store(gsel + j, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      __syncthreads();
      for (int j = tid; j < need; j += WG) {
        const uint64_t e = in_lds ? s_sel[j] : __hip_atomic_load(gsel + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e == SEL_INVALID) continue;
        const int64_t slot = (int64_t)(e & ((1ull << 40) - 1));
        if (claim(d.state, slot)) {
          unweight(d, n, slot);
          atomicAdd(&s_claimed, 1);
        } else if (in_lds) {
          s_sel[j] = SEL_INVALID;
        } else {
          __hip_atomic_store(gsel + j, SEL_INVALID, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      __syncthreads();   // the claims and weight updates before the exposures and the next round
      for (int j0 = 0; j0 < need; j0 += WG) {          // uniform trip count: whole waves append
        const int j = j0 + tid;
        const uint64_t e = j >= need ? SEL_INVALID
                           : in_lds ? s_sel[j]
                                    : __hip_atomic_load(gsel + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int64_t slot = (int64_t)(e & ((1ull << 40) - 1));
        const bool ex = e != SEL_INVALID;
        int kx = 0;
        uint4 rec = make_uint4(0u, 0u, 0u, 0u);
        if (ex) {
          rec = d.rec[slot];
          kx = expose_agent_thr(d, thr, ni.seed, ni.slot_base, slot, t);
          atomicAdd(&s_ne[kx], 1);
        }
        wave_append_entry(d, nxt, n, ex ? exposed_entry_rec(slot, rec, kx, t, r0_sel(r0v, kx)) : make_uint4(0u, 0u, 0u, 0u),
                          ex, lane);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (tid < LP_MAX_STRAINS && s_ne[tid]) atomicAdd(d.newexp + n * LP_MAX_STRAINS + tid, s_ne[tid]);
  if (tid == 0 && s_claimed) atomicAdd(d.n_uniq + n, s_claimed);
}

// Per node: transmission incidence and the log rows of day t, in two parts as
// day_begin: day_end_load (every input, the values the rows add to included, in flight
// together) and day_end_apply (stores only).
struct DayEndIn {
  int32_t sim, stop, flags, s_t, r_t, potp, pr;
  int32_t ne[LP_MAX_STRAINS], ek[LP_MAX_STRAINS], ik[LP_MAX_STRAINS];
  int32_t r_newexp, r_newexp_bs[LP_MAX_STRAINS], r_R;
  int64_t carry_s, carry_spos;
};
__device__ __forceinline__ DayEndIn day_end_load(const Dev& d, int n, int32_t t, int32_t day_flags_t, int mask,
                                                 int tx_adjust) {
  DayEndIn in;
  const int K = d.K;
  const size_t row = (size_t)t * d.n_nodes + n;
  const int64_t* tl = d.node_tally + (size_t)n * NPART;
  in.sim = d.node_sim[n];
  in.stop = d.stop_day[in.sim];
  in.flags = stage_day_flags(day_flags_t, mask);
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) {
    in.ne[k] = (tx_adjust && k < K) ? d.newexp[n * LP_MAX_STRAINS + k] : 0;
    in.ek[k] = k < K ? (int32_t)tl[P_E0 + k] : 0;
    in.ik[k] = k < K ? (int32_t)tl[P_I0 + k] : 0;
    in.r_newexp_bs[k] = k < K ? d.r_newexp_bs[row * K + k] : 0;
  }
  in.s_t = (int32_t)tl[P_S];
  in.r_t = (int32_t)tl[P_R];
  in.potp = d.potp_run[n];
  in.pr = d.p_run[n];
  in.r_newexp = d.r_newexp[row];
  in.r_R = d.r_R[row];
  in.carry_s = d.node_carry[(size_t)n * 3 + 0];
  in.carry_spos = d.node_carry[(size_t)n * 3 + 1];
  return in;
}
__device__ __forceinline__ void day_end_apply(const Dev& d, int n, int32_t t, int mask, int tx_adjust,
                                              const DayEndIn& in) {
  const size_t N = d.n_nodes, row = (size_t)t * N + n;
  const int K = d.K;
  if (in.stop < t || !(in.flags & LP_DAY_TX)) return;
  int32_t ne_tot = 0;
#pragma unroll
  for (int k = 0; k < LP_MAX_STRAINS; ++k) ne_tot += in.ne[k];
  if ((mask & LP_STAGE_TRANSMISSION) && tx_adjust) {
    d.r_newexp[row] = in.r_newexp + ne_tot;
#pragma unroll
    for (int k = 0; k < LP_MAX_STRAINS; ++k)
      if (k < K) d.r_newexp_bs[row * K + k] = in.r_newexp_bs[k] + in.ne[k];
    // every agent exposed by transmission left S and had a positive weight
    d.node_carry[(size_t)n * 3 + 0] = in.carry_s - ne_tot;
    d.node_carry[(size_t)n * 3 + 1] = in.carry_spos - ne_tot;
  }
  if (mask & LP_STAGE_LOG) {
    int32_t e = 0, i = 0;
#pragma unroll
    for (int k = 0; k < LP_MAX_STRAINS; ++k) {
      if (k >= K) break;
      d.r_Ebs[row * K + k] = in.ek[k] + in.ne[k];
      d.r_Ibs[row * K + k] = in.ik[k];
      e += in.ek[k] + in.ne[k]; i += in.ik[k];
    }
    d.r_S[row] = in.s_t - ne_tot;
    d.r_E[row] = e;
    d.r_I[row] = i;
    if (e + i > 0) d.sim_cases[in.sim] = 1;   // for the next day's early-stop rule
    d.r_R[row] = in.r_R + in.r_t;
    d.r_potp[row] = in.potp;
    d.r_p[row] = in.pr;
  }
}

// ei_piece_off = exclusive prefix over nodes of the pieces of list b (the last
// entry is the total); run by every thread of one workgroup.
__device__ void ei_plan(const Dev& d, int b) {
  const int N = d.n_nodes;
  const int32_t* cnt = d.ei_cnt + (size_t)b * N;
  const int64_t total = block_scan_nodes(
      N, [&](int n) { return (int64_t)((cnt[n] + EI_PIECE - 1) / EI_PIECE); },
      [&](int n, int64_t off) { d.ei_piece_off[n] = (int32_t)off; });
  if (threadIdx.x == 0) d.ei_piece_off[N] = (int32_t)total;
}

// Per node: transmission incidence and the log rows.  With `advance`: when the
// next tick is inside the current lp_run_days call (before day[1]), each node's
// thread also begins that day (day_begin_node) and the last workgroup runs the
// early-stop rule for it; the last workgroup plans the next day's E / I list
// pieces and moves the day counter.
__global__ void __launch_bounds__(WG) k_day_end(Dev d, int mask, int tx_adjust, int advance) {
  const int4 dy = today(d);
  const int32_t t = dy.x;
  const bool begin_next = advance && t + 1 < d.day[1];
  // the next day's record, loaded early (thread 0 of the last workgroup stores it)
  const int4 next = (advance && threadIdx.x == 0) ? day_record(d, t + 1) : make_int4(0, 0, 0, 0);
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < d.n_nodes) {
    // the next day's inputs are loaded before today's log rows are stored
    // every input of the day's end and of the next day's beginning loads before any store
    const DayEndIn ein = day_end_load(d, n, t, dy.y, mask, tx_adjust);
    const DayBeginIn in = begin_next ? day_begin_load(d, n, t + 1, mask) : DayBeginIn{};
    day_end_apply(d, n, t, mask, tx_adjust, ein);
    if (begin_next) day_begin_apply(d, n, t + 1, in);
  }
  if (advance) {
    // the last workgroup to get here moves the day counter: every workgroup has read it
    // (its load completed before the ticket, a workgroup-scope release), and nothing
    // else passes between the workgroups, so no device-scope fence is needed
    __shared__ bool am_last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t v = atomicAdd(d.ticket + 2, 1u);
      am_last = v == gridDim.x - 1;
      if (am_last) {
        atomicExch(d.ticket + 2, 0u);
        if (begin_next) d.sel_count[0] = 0;
        set_today(d, next);
      }
    }
  }
}

// The piece plan of today's E / I list for k_ei when the node count exceeds what its
// workgroups scan into LDS themselves (one workgroup).
__global__ void __launch_bounds__(WG) k_ei_plan(Dev d) { ei_plan(d, today(d).x & 1); }


}  // namespace lp
