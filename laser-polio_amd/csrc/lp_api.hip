// Host side of the C ABI declared in include/lp_step.h: binding table,
// validation, span tables, scratch, the per-day kernel sequence, hipGraph
// capture/replay and error reporting.
#include "lp_init.hip"
#include "lp_network.hip"
#include <string>
#include <vector>
#include <cstring>
#include <cstdio>
#include <algorithm>

using lp::Dev;

namespace {

enum Len { L_CAP, L_NODES, L_NODES1, L_SIMS, L_SIMPAR, L_KM, L_T, L_TS, L_E, L_EN2, L_TN, L_TNK, L_ANY };

struct BindSpec {
  const char* name;
  int dtype;
  Len len;
  bool required;
  size_t off;  // offset of the pointer field inside Dev
};

#define F(field) offsetof(Dev, field)
const BindSpec kBinds[] = {
  {"disease_state", LP_I8, L_CAP, true, F(state)},
  {"strain", LP_I8, L_CAP, true, F(strain)},
  {"node_id", LP_I16, L_CAP, true, F(node_id)},
  {"exposure_timer", LP_U16, L_CAP, true, F(exp_timer)},
  {"infection_timer", LP_U16, L_CAP, true, F(inf_timer)},
  {"paralysis_timer", LP_U16, L_CAP, true, F(par_timer)},
  {"potentially_paralyzed", LP_I8, L_CAP, true, F(pp)},
  {"paralyzed", LP_I8, L_CAP, true, F(paralyzed)},
  {"ipv_protected", LP_I8, L_CAP, true, F(ipv)},
  {"chronically_missed", LP_U8, L_CAP, true, F(missed)},
  {"acq_risk_multiplier", LP_F32, L_CAP, true, F(risk)},
  {"daily_infectivity", LP_F32, L_CAP, true, F(infect)},
  {"date_of_birth", LP_I32, L_CAP, true, F(dob)},
  {"date_of_death", LP_I32, L_CAP, true, F(dod)},
  {"ri_timer", LP_I16, L_CAP, true, F(ri_timer)},
  {"seg_begin", LP_I64, L_NODES, true, F(seg_begin)},
  {"live_end", LP_I64, L_NODES, true, F(live_end)},
  {"seg_end", LP_I64, L_NODES, true, F(seg_end)},
  {"node_sim", LP_I32, L_NODES, true, F(node_sim)},
  {"r0_scalar", LP_F32, L_NODES, true, F(r0s)},
  {"keep", LP_F32, L_NODES, true, F(keep)},
  {"birth_rate_step", LP_F64, L_NODES, true, F(birth_rate_step)},
  {"thr_ri", LP_U64, L_NODES, true, F(thr_ri)},
  {"thr_ipv", LP_U64, L_NODES, true, F(thr_ipv)},
  {"sim_params", LP_U8, L_SIMPAR, true, F(sims)},
  {"network_t", LP_F32, L_ANY, true, F(network_t)},
  {"km_cd", LP_I64, L_KM, true, F(km_cd)},
  {"day_flags", LP_I32, L_T, true, F(day_flags)},
  {"day_sia_begin", LP_I32, L_T, true, F(day_sia_begin)},
  {"day_sia_count", LP_I32, L_T, true, F(day_sia_count)},
  {"season", LP_F32, L_TS, true, F(season)},
  {"sia_strain", LP_I32, L_E, false, F(sia_strain)},
  {"sia_min_age", LP_I32, L_E, false, F(sia_min_age)},
  {"sia_max_age", LP_I32, L_E, false, F(sia_max_age)},
  {"sia_thr", LP_U64, L_EN2, false, F(sia_thr)},
  {"S", LP_I32, L_TN, true, F(r_S)},
  {"E", LP_I32, L_TN, true, F(r_E)},
  {"I", LP_I32, L_TN, true, F(r_I)},
  {"R", LP_I32, L_TN, true, F(r_R)},
  {"E_by_strain", LP_I32, L_TNK, true, F(r_Ebs)},
  {"I_by_strain", LP_I32, L_TNK, true, F(r_Ibs)},
  {"potentially_paralyzed_count", LP_I32, L_TN, true, F(r_potp)},
  {"paralyzed_count", LP_I32, L_TN, true, F(r_p)},
  {"new_potentially_paralyzed", LP_I32, L_TN, true, F(r_newpp)},
  {"new_paralyzed", LP_I32, L_TN, true, F(r_newp)},
  {"pop", LP_I32, L_TN, true, F(r_pop)},
  {"births", LP_I32, L_TN, true, F(r_births)},
  {"deaths", LP_I32, L_TN, true, F(r_deaths)},
  {"new_exposed", LP_I32, L_TN, true, F(r_newexp)},
  {"new_exposed_by_strain", LP_I32, L_TNK, true, F(r_newexp_bs)},
  {"ri_vaccinated", LP_I32, L_TN, true, F(r_riv)},
  {"ri_protected", LP_I32, L_TN, true, F(r_rip)},
  {"ri_new_exposed_by_strain", LP_I32, L_TNK, true, F(r_ri_newexp_bs)},
  {"ipv_vaccinated", LP_I32, L_TN, true, F(r_ipv)},
  {"sia_vaccinated", LP_I32, L_TN, true, F(r_siav)},
  {"sia_protected", LP_I32, L_TN, true, F(r_siap)},
  {"sia_new_exposed_by_strain", LP_I32, L_TNK, true, F(r_sia_newexp_bs)},
};
#undef F
constexpr int kNumBinds = sizeof(kBinds) / sizeof(kBinds[0]);
// Captured runs of 1, 2, 4 and 8 consecutive days of one kind.
constexpr int kRunLevels = 4;

std::string g_names;

}  // namespace

struct lp_ctx {
  lp_config cfg{};
  Dev d{};
  bool bound[kNumBinds] = {};
  int64_t bound_len[kNumBinds] = {};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool prepared = false;
  int32_t n_spans = 0;
  int32_t tpw = 4;
  int32_t agent_grid = 1;
  int32_t ei_grid = 2048, sample_grid = 2048;     // launch shapes of the list pass and the sampler
  bool mix_wide = false;                          // k_mix_foi<4> (more nodes than resident waves)
  int32_t mix_tiles = 0;                           // mix_mode 1: 16-row tiles of the largest simulation
  std::vector<void*> allocs;
  // per day kind (plain, FULL_* combinations, their list forms): graph[kind][j] runs 2^j
  // consecutive days of that kind, so a run of such days pays one graph launch per 8
  hipGraphExec_t graph[lp::N_KINDS][kRunLevels] = {};
  bool graph_ready[lp::N_KINDS][kRunLevels] = {};
  bool recount_pending = true;                     // next agent pass recounts carried statistics
  int32_t list_tick = -1;                          // the tick whose E / I list is complete (-1: none)
  bool rec_dirty = false;                          // agent records newer than the bound columns
  bool rec_actual = true;                          // records hold the actual timers of E / I agents
  bool ran = false;                                // a day or stage has been queued
  std::vector<int32_t> day_flags_h;
  std::vector<int32_t> ri_cnt_h;
  std::string err;
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  int64_t time_launches = 0;
  double time_ms = 0.0;
  lp_gravity_spec* net_specs = nullptr;            // lp_build_network inputs (n_sims), allocated on first use
  double* net_scratch = nullptr;                   // row scales and 1 / S^c per node
};

namespace {

int fail(lp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(c, LP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int64_t expected_len(const lp_ctx* c, Len l) {
  const lp_config& g = c->cfg;
  switch (l) {
    case L_CAP: return g.capacity;
    case L_NODES: return g.n_nodes;
    case L_NODES1: return g.n_nodes + 1;
    case L_SIMS: return g.n_sims;
    case L_SIMPAR: return (int64_t)g.n_sims * (int64_t)sizeof(lp_sim_params);
    case L_KM: return (int64_t)g.n_sims * LP_KM_LEN;
    case L_T: return g.nt;
    case L_TS: return (int64_t)g.nt * g.n_sims;
    case L_E: return g.n_sia_events;
    case L_EN2: return (int64_t)g.n_sia_events * g.n_nodes * 2;
    case L_TN: return (int64_t)g.nt * g.n_nodes;
    case L_TNK: return (int64_t)g.nt * g.n_nodes * g.n_strains;
    case L_ANY: return -1;
  }
  return -1;
}

template <typename T>
int alloc(lp_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  HIPCHK(c, hipMalloc(&q, n * sizeof(T) + 16));
  HIPCHK(c, hipMemsetAsync(q, 0, n * sizeof(T) + 16, c->stream));
  c->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return LP_OK;
}

int grid_of(int64_t n, int per) { return (int)((n + per - 1) / per); }

// Most workgroups of k_ei (4 wavefronts each; each wavefront loops over list pieces) and
// k_sample by default (fewer when fewer are resident at once); lp_config may set them
// (the batch autotuner does).
constexpr int kEiGrid = 2048;
constexpr int kSampleGrid = 2048;
// Workgroups per node of k_ri_list (each takes every kRiSplit-th batch of 1024 candidates).
constexpr int kRiSplit = 8;


// Host-side mirror of the device day classification (k_agents::is_plain_day).
int32_t host_stage_flags(int32_t flags, int mask) {
  int32_t f = 0;
  if (mask & LP_STAGE_VITAL) f |= flags & (LP_DAY_VD_STEP | LP_DAY_VD_COPY);
  if (mask & LP_STAGE_DISEASE) f |= flags & (LP_DAY_DS | LP_DAY_STOPCHK);
  if (mask & LP_STAGE_RI) f |= flags & LP_DAY_RI;
  if (mask & LP_STAGE_SIA) f |= flags & LP_DAY_SIA;
  if (mask & (LP_STAGE_TRANSMISSION | LP_STAGE_LOG)) f |= flags & LP_DAY_TX;
  return f;
}

// Day kind: KIND_RECOUNT, KIND_PLAIN, or the OR of FULL_VD / FULL_RI / FULL_SIA.
int day_kind(const lp_ctx* c, int32_t t, int mask, int recount) {
  if (recount) return lp::KIND_RECOUNT;
  const int32_t f = host_stage_flags(c->day_flags_h[t], mask);
  return ((f & LP_DAY_VD_STEP) ? lp::FULL_VD : 0) | ((f & LP_DAY_RI) ? lp::FULL_RI : 0) |
         ((f & LP_DAY_SIA) ? lp::FULL_SIA : 0);
}

// k_ei of one day; above EI_LDS_NODES nodes its piece plan comes from k_ei_plan.
void launch_ei(lp_ctx* c, hipStream_t s, int mask) {
  if (c->cfg.n_nodes > lp::EI_LDS_NODES) hipLaunchKernelGGL(lp::k_ei_plan, dim3(1), dim3(lp::WG), 0, s, c->d);
  hipLaunchKernelGGL(lp::k_ei, dim3(c->ei_grid), dim3(lp::WG), 0, s, c->d, mask);
}

// Agent pass of one day: k_plain over U = tiles per span on plain days,
// k_full<kind> on FULL days, k_agents in groups of 2 tiles when recounting.
template <int U>
void launch_agents_u(lp_ctx* c, hipStream_t s, int kind, int mask) {
  const dim3 g(c->n_spans), b(lp::WG);
  switch (kind) {
    case lp::KIND_PLAIN: hipLaunchKernelGGL((lp::k_plain<U>), g, b, 0, s, c->d, mask); break;
    case lp::KIND_PLAIN_LIST: launch_ei(c, s, mask); break;
    case 1: hipLaunchKernelGGL((lp::k_full<1, true>), g, b, 0, s, c->d, mask); break;
    case 2: hipLaunchKernelGGL((lp::k_full<2, true>), g, b, 0, s, c->d, mask); break;
    case 3: hipLaunchKernelGGL((lp::k_full<3, true>), g, b, 0, s, c->d, mask); break;
    case 4: hipLaunchKernelGGL((lp::k_full<4, true>), g, b, 0, s, c->d, mask); break;
    case 5: hipLaunchKernelGGL((lp::k_full<5, true>), g, b, 0, s, c->d, mask); break;
    case 6: hipLaunchKernelGGL((lp::k_full<6, true>), g, b, 0, s, c->d, mask); break;
    case 7: hipLaunchKernelGGL((lp::k_full<7, true>), g, b, 0, s, c->d, mask); break;
    case lp::KIND_FULL_LIST + 0: case lp::KIND_FULL_LIST + 1: case lp::KIND_FULL_LIST + 2:
    case lp::KIND_FULL_LIST + 3: case lp::KIND_FULL_LIST + 4: case lp::KIND_FULL_LIST + 5:
    case lp::KIND_FULL_LIST + 6: {
      // DiseaseState from the E / I lists, then births and deaths (only the chunks with a
      // death due or newborn slots), RI from the candidate lists, SIA over the targeted
      // nodes: each stage a pass of its own, in the reference's order, all tallying
      // into the per-node accumulators
      const int f = kind - lp::KIND_FULL_LIST + 1;
      launch_ei(c, s, mask);
      if (f & lp::FULL_VD)
        hipLaunchKernelGGL(lp::k_vd, dim3(c->cfg.n_nodes + grid_of(c->n_spans, lp::WG / 64)), b, 0, s, c->d, mask);
      if (f & lp::FULL_RI) hipLaunchKernelGGL(lp::k_ri_list, dim3(c->cfg.n_nodes, kRiSplit), b, 0, s, c->d, mask);
      if (f & lp::FULL_SIA) hipLaunchKernelGGL(lp::k_sia, g, b, 0, s, c->d, mask);
      break;
    }
    default: hipLaunchKernelGGL((lp::k_agents<2, lp::KIND_RECOUNT>), g, b, 0, s, c->d, mask); break;
  }
}

// Grid of the record pack / unpack kernels (grid-stride over quads of slots).
int rec_grid(const lp_ctx* c) {
  const int64_t nq = (c->cfg.capacity + 3) / 4;
  const int64_t g = (nq + lp::WG - 1) / lp::WG;
  return (int)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

// Day kinds that process the E / I list; afterwards the records of listed agents
// keep the timers of when they entered the list until enqueue_actualize.
bool list_kind(int kind) { return kind == lp::KIND_PLAIN_LIST || kind >= lp::KIND_FULL_LIST; }

// Records <- actual timers of the agents in the complete E / I list (that of
// list_tick), when days since the last call processed the list.
void enqueue_actualize(lp_ctx* c) {
  if (c->rec_actual) return;
  if (c->list_tick >= 0)
    hipLaunchKernelGGL(lp::k_ei_actual, dim3(c->cfg.n_nodes), dim3(lp::WG), 0, c->stream, c->d, c->list_tick & 1,
                       c->list_tick);
  c->rec_actual = true;
}

// Bound columns <- device-held state: agent records and actual ri_timer values.
void enqueue_sync_columns(lp_ctx* c) {
  enqueue_actualize(c);
  hipLaunchKernelGGL(lp::k_unpack, dim3(rec_grid(c)), dim3(lp::WG), 0, c->stream, c->d);
  hipLaunchKernelGGL(lp::k_ri_sync, dim3(c->n_spans), dim3(lp::WG), 0, c->stream, c->d);
  hipLaunchKernelGGL(lp::k_set_ri_ref, dim3(1), dim3(64), 0, c->stream, c->d, -2);
}

// Before the first day a context runs: bound ri_timer columns hold actual values as of
// the end of the tick before it (a run that starts at tick t > 0, e.g. from a checkpoint,
// has had every RI step through t - 1 applied to them).
void first_run(lp_ctx* c, int32_t t) {
  if (c->ran) return;
  c->ran = true;
  if (t > 0) hipLaunchKernelGGL(lp::k_set_ri_ref, dim3(1), dim3(64), 0, c->stream, c->d, t - 1);
}

// Whether every day in [a, b) runs as day kind `kind`, a list kind, inside lp_run_days
// (no recount pending; each day's list is completed by the day before it).
bool same_list_kind(const lp_ctx* c, int32_t a, int32_t b, int kind) {
  for (int32_t u = a; u < b; ++u) {
    if (!(host_stage_flags(c->day_flags_h[u], LP_STAGE_ALL) & LP_DAY_DS)) return false;
    int k = day_kind(c, u, LP_STAGE_ALL, 0);
    if (k == lp::KIND_PLAIN) k = lp::KIND_PLAIN_LIST;
    else if (k >= lp::KIND_FULL && k < lp::KIND_RECOUNT) k = lp::KIND_FULL_LIST + k - lp::KIND_FULL;
    if (k != kind) return false;
  }
  return true;
}

// The kernel sequence of one tick.  kind: agent-pass variant (KIND_RECOUNT
// recounts every carried statistic and the paralysis counts); mask: stages;
// tx_adjust: apply today's transmission exposures to the log; advance: move
// the day counter at the end.
int enqueue_day(lp_ctx* c, int32_t t, int kind, int mask, int tx_adjust, bool advance, bool begin) {
  const int recount = kind == lp::KIND_RECOUNT ? 1 : 0;
  const Dev& d = c->d;
  const int N = c->cfg.n_nodes;
  hipStream_t s = c->stream;
  // passes that read the records need their actual timers (never inside a captured
  // graph: run_days calls enqueue_actualize before capturing)
  if (!list_kind(kind)) enqueue_actualize(c);
  // a RECOUNT pass reads the bound columns: bring them up to date first; it rebuilds the
  // RI candidate lists
  if (recount && c->rec_dirty) enqueue_sync_columns(c);
  if (recount) HIPCHK(c, hipMemsetAsync(d.ri_lcnt, 0, 2 * (size_t)N * sizeof(int32_t), s));
  if (begin) hipLaunchKernelGGL(lp::k_day_begin, dim3(grid_of(N, 256)), dim3(256), 0, s, d, mask);
  const bool ev = c->timing;
  if (ev) {
    if (c->ev_used + 2 > c->ev.size()) {
      hipEvent_t a, b;
      HIPCHK(c, hipEventCreate(&a));
      HIPCHK(c, hipEventCreate(&b));
      c->ev.push_back(a); c->ev.push_back(b);
    }
    HIPCHK(c, hipEventRecord(c->ev[c->ev_used], s));
  }
  switch (c->tpw) {
    case 1: launch_agents_u<1>(c, s, kind, mask); break;
    case 2: launch_agents_u<2>(c, s, kind, mask); break;
    case 4: launch_agents_u<4>(c, s, kind, mask); break;
    case 8: launch_agents_u<8>(c, s, kind, mask); break;
    default: launch_agents_u<16>(c, s, kind, mask); break;
  }
  if (ev) { HIPCHK(c, hipEventRecord(c->ev[c->ev_used + 1], s)); c->ev_used += 2; }
  if (recount) {
    hipLaunchKernelGGL(lp::k_pack, dim3(rec_grid(c)), dim3(lp::WG), 0, s, d);
    hipLaunchKernelGGL(lp::k_node_info, dim3(grid_of(N, 256)), dim3(256), 0, s, d);
    // the pass left ri_timer actual as of today if it ran RI; otherwise it is actual as
    // of the latest applied RI step, as before the pass
    const bool ri = host_stage_flags(c->day_flags_h[t], mask) & LP_DAY_RI;
    hipLaunchKernelGGL(lp::k_set_ri_ref, dim3(1), dim3(64), 0, s, d, ri ? t : -2);
  }
  // list days tally into node_acc (span partials hold only the carried span weights);
  // the dense passes of the other kinds write every span's partial
  hipLaunchKernelGGL(lp::k_node_reduce, dim3(N), dim3(lp::WG), 0, s, d, mask, recount,
                     kind < lp::KIND_PLAIN_LIST ? 0 : 1);
  if (mask & LP_STAGE_TRANSMISSION) {
    if (c->cfg.mix_mode == 1)
      hipLaunchKernelGGL(lp::k_mix_mfma, dim3(c->cfg.n_sims, c->mix_tiles), dim3(64), 0, s, d);
    if (c->mix_wide)
      hipLaunchKernelGGL(lp::k_mix_foi<4>, dim3(grid_of(N, lp::WG / 64)), dim3(lp::WG), 0, s, d, mask);
    else
      hipLaunchKernelGGL(lp::k_mix_foi<1>, dim3(grid_of(N, lp::WG / 64)), dim3(lp::WG), 0, s, d, mask);
    // with the probe offsets in LDS k_sample plans round 1 itself
    const size_t off_lds = (N + 1 <= lp::LDS_OFF_MAX) ? (size_t)(N + 1) * sizeof(int64_t) : 0;
    if (!off_lds) hipLaunchKernelGGL(lp::k_plan, dim3(1), dim3(lp::WG), 0, s, d, 1);
    hipLaunchKernelGGL(lp::k_sample, dim3(c->sample_grid), dim3(lp::WG), off_lds, s, d);
    hipLaunchKernelGGL(lp::k_finish, dim3(N), dim3(lp::WG), 0, s, d);
  }
  hipLaunchKernelGGL(lp::k_day_end, dim3(grid_of(N, 256)), dim3(256), 0, s, d, mask, tx_adjust, advance ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  c->rec_dirty = true;
  return LP_OK;
}

// Captures `run` consecutive days of kind `kind` from day t into *g.  Nothing runs:
// the graph is launched later.
int capture_days(lp_ctx* c, hipGraphExec_t* g, int32_t t, int kind, int run) {
  hipGraph_t graph;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = LP_OK;
  for (int r = 0; r < run && rc == LP_OK; ++r) rc = enqueue_day(c, t + r, kind, LP_STAGE_ALL, 1, true, false);
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, LP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  HIPCHK(c, hipGraphInstantiate(g, graph, nullptr, nullptr, 0));
  HIPCHK(c, hipGraphDestroy(graph));
  return LP_OK;
}

}  // namespace

extern "C" {

int32_t lp_abi_version(void) { return LP_ABI_VERSION; }
int64_t lp_sizeof_sim_params(void) { return (int64_t)sizeof(lp_sim_params); }

const char* lp_bind_names(void) {
  if (g_names.empty()) {
    for (int i = 0; i < kNumBinds; ++i) { if (i) g_names += ","; g_names += kBinds[i].name; }
  }
  return g_names.c_str();
}

int lp_create(const lp_config* cfg, lp_ctx** out) {
  if (!cfg || !out) return LP_ERR_ARG;
  *out = nullptr;
  if (cfg->capacity <= 0 || cfg->capacity % LP_TILE != 0 || cfg->n_nodes <= 0 || cfg->n_sims <= 0 ||
      cfg->n_strains < 1 || cfg->n_strains > LP_MAX_STRAINS || cfg->nt < 1 || cfg->nt > (1 << 23) || cfg->n_sia_events < 0 ||
      cfg->n_nodes > 32767 || cfg->capacity >= (int64_t)1 << 32 || cfg->mix_mode < 0 || cfg->mix_mode > 1)
    return LP_ERR_ARG;
  for (int32_t g : {cfg->ei_grid, cfg->sample_grid, cfg->commit_grid})   // commit_grid: validated, unused
    if (g != 0 && (g < 64 || g > 65535)) return LP_ERR_ARG;
  lp_ctx* c = new lp_ctx();
  c->cfg = *cfg;
  c->ei_grid = cfg->ei_grid ? cfg->ei_grid : kEiGrid;
  c->sample_grid = cfg->sample_grid ? cfg->sample_grid : kSampleGrid;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return LP_ERR_HIP; }
  c->own_stream = true;
  *out = c;
  return LP_OK;
}

int lp_bind(lp_ctx* c, const char* name, void* ptr, int32_t dtype, int64_t n) {
  if (!c || !name) return LP_ERR_ARG;
  if (c->prepared) return fail(c, LP_ERR_STATE, "lp_bind after lp_prepare");
  for (int i = 0; i < kNumBinds; ++i) {
    if (std::strcmp(kBinds[i].name, name) != 0) continue;
    if (dtype != kBinds[i].dtype) return fail(c, LP_ERR_ARG, std::string("dtype mismatch for ") + name);
    const int64_t want = expected_len(c, kBinds[i].len);
    if (want >= 0 && n != want)
      return fail(c, LP_ERR_ARG, std::string("length mismatch for ") + name + ": got " + std::to_string(n) +
                                     ", want " + std::to_string(want));
    if (!ptr && n > 0) return fail(c, LP_ERR_ARG, std::string("null pointer for ") + name);
    *reinterpret_cast<void**>(reinterpret_cast<char*>(&c->d) + kBinds[i].off) = ptr;
    c->bound[i] = true;
    c->bound_len[i] = n;
    return LP_OK;
  }
  return fail(c, LP_ERR_ARG, std::string("unknown binding ") + name);
}

int lp_set_stream(lp_ctx* c, void* s) {
  if (!c) return LP_ERR_ARG;
  for (int k = 0; k < lp::N_KINDS; ++k)
    for (int j = 0; j < kRunLevels; ++j)
      if (c->graph_ready[k][j]) return fail(c, LP_ERR_STATE, "stream change after graph capture");
  const hipStream_t ns = reinterpret_cast<hipStream_t>(s);
  if (ns == c->stream) return LP_OK;
  // work queued on the current stream completes before anything is queued on the new one
  if (c->stream) HIPCHK(c, hipStreamSynchronize(c->stream));
  // the library's own stream is released; the caller's stream stays the caller's
  if (c->own_stream && c->stream) HIPCHK(c, hipStreamDestroy(c->stream));
  c->own_stream = false;
  c->stream = ns;
  return LP_OK;
}

int lp_owns_stream(lp_ctx* c) { return c ? (c->own_stream ? 1 : 0) : LP_ERR_ARG; }

int lp_prepare(lp_ctx* c) {
  if (!c) return LP_ERR_ARG;
  if (c->prepared) return fail(c, LP_ERR_STATE, "lp_prepare called twice");
  const lp_config& g = c->cfg;
  for (int i = 0; i < kNumBinds; ++i) {
    const bool need = kBinds[i].required || (g.n_sia_events > 0 && kBinds[i].len != L_ANY);
    if (need && !c->bound[i]) return fail(c, LP_ERR_UNBOUND, std::string("unbound buffer ") + kBinds[i].name);
  }
  const int N = g.n_nodes, S = g.n_sims;
  std::vector<int64_t> sb(N), se(N), le(N);
  std::vector<int32_t> nsim(N);
  std::vector<lp_sim_params> sims(S);
  HIPCHK(c, hipMemcpy(sb.data(), c->d.seg_begin, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(se.data(), c->d.seg_end, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(le.data(), c->d.live_end, N * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(nsim.data(), c->d.node_sim, N * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(sims.data(), c->d.sims, S * sizeof(lp_sim_params), hipMemcpyDeviceToHost));
  int64_t net_len = 0;
  for (int i = 0; i < kNumBinds; ++i) if (std::strcmp(kBinds[i].name, "network_t") == 0) net_len = c->bound_len[i];
  for (int n = 0; n < N; ++n) {
    if (sb[n] % LP_TILE || se[n] % LP_TILE || sb[n] > se[n] || se[n] > g.capacity || le[n] < sb[n] || le[n] > se[n] ||
        (n > 0 && sb[n] < se[n - 1]))
      return fail(c, LP_ERR_ARG, "bad node segment " + std::to_string(n));
    if (nsim[n] < 0 || nsim[n] >= S) return fail(c, LP_ERR_ARG, "bad node_sim " + std::to_string(n));
  }
  for (int s = 0; s < S; ++s) {
    const lp_sim_params& p = sims[s];
    if (p.n_nodes <= 0 || p.node_begin < 0 || p.node_begin + p.n_nodes > N)
      return fail(c, LP_ERR_ARG, "bad sim node range " + std::to_string(s));
    for (int n = p.node_begin; n < p.node_begin + p.n_nodes; ++n)
      if (nsim[n] != s) return fail(c, LP_ERR_ARG, "node_sim disagrees with sim_params " + std::to_string(s));
    if (p.net_off < 0 || p.net_off + (int64_t)p.n_nodes * p.n_nodes > net_len)
      return fail(c, LP_ERR_ARG, "network_t too short for sim " + std::to_string(s));
    if (p.slot_base < 0 || p.slot_base > sb[p.node_begin])
      return fail(c, LP_ERR_ARG, "bad slot_base for sim " + std::to_string(s));
    if (p.ri_strain < 0 || p.ri_strain >= g.n_strains || p.disp_r < 1 || p.ri_step < 1)
      return fail(c, LP_ERR_ARG, "bad scalar parameters for sim " + std::to_string(s));
  }
  if (g.n_sia_events > 0) {
    std::vector<int32_t> ss(g.n_sia_events);
    HIPCHK(c, hipMemcpy(ss.data(), c->d.sia_strain, g.n_sia_events * 4, hipMemcpyDeviceToHost));
    for (int e = 0; e < g.n_sia_events; ++e)
      if (ss[e] < 0 || ss[e] >= g.n_strains) return fail(c, LP_ERR_ARG, "SIA vaccine strain out of range");
  }
  int tpw = g.tiles_per_block > 0 ? g.tiles_per_block : 16;
  if (tpw != 1 && tpw != 2 && tpw != 4 && tpw != 8 && tpw != 16)
    return fail(c, LP_ERR_ARG, "tiles_per_block must be 1, 2, 4, 8 or 16");
  c->tpw = tpw;
  std::vector<int4> spans;
  std::vector<int32_t> nspan(N + 1, 0);
  for (int n = 0; n < N; ++n) {
    nspan[n] = (int32_t)spans.size();
    const int64_t t0 = sb[n] / LP_TILE, t1 = se[n] / LP_TILE;
    for (int64_t t = t0; t < t1; t += tpw) {
      const int64_t cnt = std::min<int64_t>(tpw, t1 - t);
      spans.push_back(make_int4(n, (int)t, (int)cnt, nsim[n]));
    }
  }
  nspan[N] = (int32_t)spans.size();
  c->n_spans = (int32_t)spans.size();
  if (c->n_spans == 0) return fail(c, LP_ERR_ARG, "no agent slots");
  Dev& d = c->d;
  d.capacity = g.capacity; d.n_nodes = N; d.n_sims = S; d.K = g.n_strains; d.nt = g.nt;
  d.n_spans = c->n_spans; d.tpw = tpw;
  d.tpw_sh = __builtin_ctz((unsigned)tpw);
  {
    int dev = 0, cus = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    c->agent_grid = std::max(1, std::min(c->n_spans, cus * 8));
    // default grids of k_ei and k_sample: as many workgroups as are resident at once
    // (occupancy x CUs, at most 2048), so none waits for another to finish before it starts
    auto resident = [&](const void* kernel, size_t lds, int cap) -> int {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, lp::WG, lds) != hipSuccess || nb < 1) return cap;
      return std::max(64, std::min(cap, nb * cus));
    };
    const size_t off_lds = (N + 1 <= lp::LDS_OFF_MAX) ? (size_t)(N + 1) * sizeof(int64_t) : 0;
    if (!c->cfg.ei_grid) c->ei_grid = resident(reinterpret_cast<const void*>(lp::k_ei), 0, kEiGrid);
    if (!c->cfg.sample_grid) c->sample_grid = resident(reinterpret_cast<const void*>(lp::k_sample), off_lds, kSampleGrid);
    // k_mix_foi: the variant held to 4 waves per SIMD once its workgroups outnumber those
    // of the unrestricted variant resident at once
    const int mix_blocks = (int)grid_of(N, lp::WG / 64);
    c->mix_wide = mix_blocks > resident(reinterpret_cast<const void*>(lp::k_mix_foi<1>), 0, 1 << 30);
  }
  int4* spans_d; int32_t* nspan_d;
  int rc;
  if ((rc = alloc(c, &spans_d, spans.size()))) return rc;
  if ((rc = alloc(c, &nspan_d, N + 1))) return rc;
  HIPCHK(c, hipMemcpyAsync(spans_d, spans.data(), spans.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(nspan_d, nspan.data(), (N + 1) * 4, hipMemcpyHostToDevice, c->stream));
  d.spans = spans_d; d.node_span_begin = nspan_d;
  const int64_t nch = g.capacity / LP_CHUNK + 1;
  if ((rc = alloc(c, &d.day, 8))) return rc;   // t, until, -, -, then the day record (16-byte aligned)
  if ((rc = alloc(c, &d.stop_day, S))) return rc;
  if ((rc = alloc(c, &d.err, 4))) return rc;
  if ((rc = alloc(c, &d.partials, (size_t)c->n_spans * lp::NPART))) return rc;
  if ((rc = alloc(c, &d.chunk_sum, nch))) return rc;
  if ((rc = alloc(c, &d.fin_prefix, c->n_spans))) return rc;
  if ((rc = alloc(c, &d.span_prefix, c->n_spans))) return rc;
  if ((rc = alloc(c, &d.span_lane, (size_t)N * 64))) return rc;
  HIPCHK(c, hipMemsetAsync(d.span_lane, 0, (size_t)N * 64 * sizeof(lp::SpanLane), c->stream));
  if ((rc = alloc(c, &d.nb_begin, N))) return rc;
  if ((rc = alloc(c, &d.births, N))) return rc;
  if ((rc = alloc(c, &d.node_tally, (size_t)N * lp::NPART))) return rc;
  if ((rc = alloc(c, &d.potp_run, N))) return rc;
  if ((rc = alloc(c, &d.p_run, N))) return rc;
  if ((rc = alloc(c, &d.node_carry, (size_t)N * 3))) return rc;
  if ((rc = alloc(c, &d.chunk_min_dod, nch))) return rc;
  if ((rc = alloc(c, &d.rec, (size_t)g.capacity))) return rc;
  if ((rc = alloc(c, &d.ei_list, 2 * (size_t)g.capacity))) return rc;
  if ((rc = alloc(c, &d.ei_cnt, 2 * (size_t)N))) return rc;
  if ((rc = alloc(c, &d.ei_piece_off, (size_t)N + 1))) return rc;
  if ((rc = alloc(c, &d.node_acc, (size_t)N * lp::NPART))) return rc;
  if ((rc = alloc(c, &d.beta_f, (size_t)N * LP_MAX_STRAINS))) return rc;
  d.mix_buf = nullptr;
  if (g.mix_mode == 1) {
    if ((rc = alloc(c, &d.mix_buf, (size_t)N * LP_MAX_STRAINS))) return rc;
    int max_n = 0;
    for (const auto& p : sims) max_n = std::max(max_n, (int)p.n_nodes);
    c->mix_tiles = (max_n + 15) / 16;
  }
  if ((rc = alloc(c, &d.sim_cases, (size_t)S))) return rc;
  if ((rc = alloc(c, &d.ninfo, (size_t)N))) return rc;
  if ((rc = alloc(c, &d.ri_list, 2 * (size_t)g.capacity))) return rc;
  if ((rc = alloc(c, &d.ri_lcnt, 2 * (size_t)N))) return rc;
  HIPCHK(c, hipMemsetAsync(d.beta_f, 0, (size_t)N * LP_MAX_STRAINS * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(d.ei_cnt, 0, 2 * (size_t)N * sizeof(int32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(d.ei_piece_off, 0, ((size_t)N + 1) * sizeof(int32_t), c->stream));
  HIPCHK(c, hipMemsetAsync(d.node_acc, 0, (size_t)N * lp::NPART * sizeof(int64_t), c->stream));
  if ((rc = alloc(c, &d.p_dbg, (size_t)N * LP_MAX_STRAINS))) return rc;
  if ((rc = alloc(c, &d.beta_mixed, (size_t)N * LP_MAX_STRAINS))) return rc;
  if ((rc = alloc(c, &d.expected_dbg, N))) return rc;
  if ((rc = alloc(c, &d.total_dbg, N))) return rc;
  if ((rc = alloc(c, &d.birth_dbg, 2 * (size_t)N))) return rc;
  if ((rc = alloc(c, &d.size, N))) return rc;
  if ((rc = alloc(c, &d.take_all, N))) return rc;
  if ((rc = alloc(c, &d.n_uniq, N))) return rc;
  if ((rc = alloc(c, &d.m_round, N))) return rc;
  if ((rc = alloc(c, &d.resv, N))) return rc;
  if ((rc = alloc(c, &d.W_round, N))) return rc;
  if ((rc = alloc(c, &d.thr_strain, (size_t)N * LP_MAX_STRAINS))) return rc;
  if ((rc = alloc(c, &d.newexp, (size_t)N * LP_MAX_STRAINS))) return rc;
  if ((rc = alloc(c, &d.probe_off, N + 1))) return rc;
  if ((rc = alloc(c, &d.round_base, 16))) return rc;
  if ((rc = alloc(c, &d.sel_count, 2))) return rc;
  if ((rc = alloc(c, &d.ticket, 4))) return rc;
  d.sel_cap = g.capacity / 4 + 65536;
  if ((rc = alloc(c, &d.sel_list, d.sel_cap))) return rc;
  if ((rc = alloc(c, &d.sel_word, d.sel_cap))) return rc;
  c->day_flags_h.resize(g.nt);
  HIPCHK(c, hipMemcpy(c->day_flags_h.data(), c->d.day_flags, g.nt * 4, hipMemcpyDeviceToHost));
  {
    // RI days among ticks 0..t, for the arithmetic on the reference-tick ri_timer values
    c->ri_cnt_h.assign((size_t)g.nt + 1, 0);
    for (int32_t t = 0; t < g.nt; ++t) c->ri_cnt_h[t + 1] = c->ri_cnt_h[t] + ((c->day_flags_h[t] & LP_DAY_RI) ? 1 : 0);
    int32_t* rc_d = nullptr;
    if ((rc = alloc(c, &rc_d, (size_t)g.nt + 1))) return rc;
    HIPCHK(c, hipMemcpyAsync(rc_d, c->ri_cnt_h.data(), ((size_t)g.nt + 1) * 4, hipMemcpyHostToDevice, c->stream));
    d.ri_cnt = rc_d;
    if ((rc = alloc(c, &d.ri_ref, 2))) return rc;
  }
  std::vector<int32_t> stop(S, INT_MAX);
  HIPCHK(c, hipMemcpyAsync(d.stop_day, stop.data(), S * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->prepared = true;
  return LP_OK;
}

int lp_run_days(lp_ctx* c, int32_t t_begin, int32_t t_end, int32_t use_graph) {
  if (!c) return LP_ERR_ARG;
  if (!c->prepared) return fail(c, LP_ERR_STATE, "lp_run_days before lp_prepare");
  if (t_begin < 0 || t_end > c->cfg.nt || t_begin > t_end) return fail(c, LP_ERR_ARG, "bad tick range");
  if (t_begin == t_end) return LP_OK;
  first_run(c, t_begin);
  // inside the call each day's k_day_end begins the next day; only the first is begun here
  hipLaunchKernelGGL(lp::k_set_day, dim3(1), dim3(64), 0, c->stream, c->d, t_begin, t_end);
  bool first = true;
  int rc;
  int32_t t = t_begin;
  if (t == 0) {
    if ((rc = enqueue_day(c, t, lp::KIND_RECOUNT, LP_STAGE_LOG, 0, true, first))) return rc;
    first = false;
    c->recount_pending = false;
    c->list_tick = 1;
    t = 1;
  }
  for (; t < t_end; ++t) {
    if (c->recount_pending) {
      if ((rc = enqueue_day(c, t, lp::KIND_RECOUNT, LP_STAGE_ALL, 1, true, first))) return rc;
      first = false;
      c->recount_pending = false;
      c->list_tick = t + 1;
      continue;
    }
    int kind = day_kind(c, t, LP_STAGE_ALL, 0);
    // the list kinds keep E / I timers as event days, which needs DiseaseState every day
    if (c->list_tick == t && (host_stage_flags(c->day_flags_h[t], LP_STAGE_ALL) & LP_DAY_DS)) {
      if (kind == lp::KIND_PLAIN) kind = lp::KIND_PLAIN_LIST;
      else if (kind >= lp::KIND_FULL && kind < lp::KIND_RECOUNT) kind = lp::KIND_FULL_LIST + kind - lp::KIND_FULL;
    }
    if (list_kind(kind)) c->rec_actual = false;
    else enqueue_actualize(c);
    c->list_tick = t + 1;   // every day kind run here completes the next tick's list
    if (use_graph && !c->timing && !first && c->stream != nullptr) {
      // the following days of the same list kind (each completes the next one's list,
      // so they need nothing from the host in between) join this day's graph: 2^j days
      int j = 0;
      if (list_kind(kind))
        while (j + 1 < kRunLevels && t + (2 << j) <= t_end && same_list_kind(c, t + (1 << j), t + (2 << j), kind)) ++j;
      const int run = 1 << j;
      if (!c->graph_ready[kind][j]) {
        if ((rc = capture_days(c, &c->graph[kind][j], t, kind, run))) return rc;
        c->graph_ready[kind][j] = true;
      }
      HIPCHK(c, hipGraphLaunch(c->graph[kind][j], c->stream));
      c->rec_dirty = true;   // as enqueue_day marks it: the columns are behind the records
      t += run - 1;
      c->list_tick = t + 1;
    } else {
      if ((rc = enqueue_day(c, t, kind, LP_STAGE_ALL, 1, true, first))) return rc;
    }
    first = false;
  }
  return LP_OK;
}

int lp_step_stage(lp_ctx* c, int32_t t, int32_t mask) {
  if (!c) return LP_ERR_ARG;
  if (!c->prepared) return fail(c, LP_ERR_STATE, "step before lp_prepare");
  if (t < 0 || t >= c->cfg.nt) return fail(c, LP_ERR_ARG, "bad tick");
  if (mask & ~LP_STAGE_ALL) return fail(c, LP_ERR_ARG, "bad stage mask");
  first_run(c, t);
  hipLaunchKernelGGL(lp::k_set_day, dim3(1), dim3(64), 0, c->stream, c->d, t, 0);
  const bool fused_tx_log = (mask & LP_STAGE_TRANSMISSION) && (mask & LP_STAGE_LOG);
  const bool log_only = (mask & LP_STAGE_LOG) && !(mask & LP_STAGE_TRANSMISSION);
  const int recount = (log_only || c->recount_pending) ? 1 : 0;
  c->recount_pending = false;
  enqueue_actualize(c);
  c->list_tick = -1;   // stage calls do not keep the E / I list
  return enqueue_day(c, t, day_kind(c, t, mask, recount), mask, fused_tx_log || !(mask & LP_STAGE_LOG) ? 1 : 0, false,
                     true);
}

int64_t lp_sizeof_init_sim(void) { return (int64_t)sizeof(lp_init_sim); }

int lp_init_population(lp_ctx* c, const lp_init_sim* sims, int32_t n_sims, const int64_t* infected, int64_t n_inf) {
  if (!c || !sims || n_inf < 0 || (n_inf > 0 && !infected)) return LP_ERR_ARG;
  if (!c->prepared) return fail(c, LP_ERR_STATE, "lp_init_population before lp_prepare");
  if (n_sims != c->cfg.n_sims) return fail(c, LP_ERR_ARG, "lp_init_population: one lp_init_sim per simulation");
  for (int s = 0; s < n_sims; ++s) {
    const lp_init_sim& p = sims[s];
    if (p.n_pyr < 0 || p.n_pyr > LP_INIT_PYR || (p.age_model == 1 && p.n_pyr < 1) || p.n_bands < 0 ||
        p.n_bands > LP_INIT_BANDS || p.het_model < 0 || p.het_model > 2 || p.age_model < 0 || p.age_model > 1)
      return fail(c, LP_ERR_ARG, "lp_init_population: bad spec for sim " + std::to_string(s));
  }
  lp_init_sim* specs_d = nullptr;
  int64_t* inf_d = nullptr;
  HIPCHK(c, hipMalloc(&specs_d, sizeof(lp_init_sim) * n_sims));
  if (n_inf > 0) HIPCHK(c, hipMalloc(&inf_d, sizeof(int64_t) * n_inf));
  HIPCHK(c, hipMemcpyAsync(specs_d, sims, sizeof(lp_init_sim) * n_sims, hipMemcpyHostToDevice, c->stream));
  if (n_inf > 0) HIPCHK(c, hipMemcpyAsync(inf_d, infected, sizeof(int64_t) * n_inf, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(lp::k_init, dim3(c->n_spans), dim3(lp::WG), 0, c->stream, c->d, specs_d);
  if (n_inf > 0)
    hipLaunchKernelGGL(lp::k_init_infected, dim3(grid_of(n_inf, 256) < 4096 ? grid_of(n_inf, 256) : 4096), dim3(256), 0,
                       c->stream, c->d, inf_d, n_inf);
  const hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(specs_d);
  if (inf_d) (void)hipFree(inf_d);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(c, LP_ERR_HIP, std::string("lp_init_population: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  // every carried statistic is recounted from the new columns by the first pass
  c->recount_pending = true;
  c->list_tick = -1;
  c->rec_dirty = false;
  return LP_OK;
}

int lp_build_network(lp_ctx* c, const lp_gravity_spec* specs, int32_t n_sims, const double* pop, const double* dist) {
  if (!c || !specs || !pop || !dist) return LP_ERR_ARG;
  if (!c->prepared) return fail(c, LP_ERR_STATE, "lp_build_network before lp_prepare");
  if (n_sims != c->cfg.n_sims) return fail(c, LP_ERR_ARG, "lp_build_network: one lp_gravity_spec per simulation");
  int max_n = 0;
  for (int s = 0; s < n_sims; ++s) {
    const lp_gravity_spec& g = specs[s];
    if (!(g.max_rowsum >= 0.0) || !(g.k >= 0.0)) return fail(c, LP_ERR_ARG, "lp_build_network: bad spec for sim " + std::to_string(s));
  }
  std::vector<lp_sim_params> sp((size_t)n_sims);
  HIPCHK(c, hipMemcpyAsync(sp.data(), c->d.sims, sizeof(lp_sim_params) * n_sims, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (const auto& p : sp) max_n = std::max(max_n, (int)p.n_nodes);
  int rc;
  if (!c->net_specs && (rc = alloc(c, &c->net_specs, (size_t)n_sims))) return rc;
  if (!c->net_scratch && (rc = alloc(c, &c->net_scratch, 2 * (size_t)c->cfg.n_nodes))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->net_specs, specs, sizeof(lp_gravity_spec) * n_sims, hipMemcpyHostToDevice, c->stream));
  double* scale = c->net_scratch;
  double* inv_sc = c->net_scratch + c->cfg.n_nodes;
  hipLaunchKernelGGL(lp::k_net_rows, dim3(c->cfg.n_nodes), dim3(lp::WG), 0, c->stream, c->d, c->net_specs, pop, dist, scale,
                     inv_sc);
  const int tiles = (max_n + lp::NET_TILE - 1) / lp::NET_TILE;
  hipLaunchKernelGGL(lp::k_net_tiles, dim3(tiles * tiles, n_sims), dim3(lp::WG), 0, c->stream, c->d, c->net_specs, pop,
                     dist, scale, inv_sc);
  HIPCHK(c, hipGetLastError());
  // the specs are copied before the call returns (pageable host memory)
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LP_OK;
}

int lp_invalidate(lp_ctx* c) {
  if (!c) return LP_ERR_ARG;
  if (c->rec_dirty)
    return fail(c, LP_ERR_STATE, "lp_invalidate with unsynced device state: call lp_sync before writing bound buffers");
  c->recount_pending = true;
  c->list_tick = -1;
  return LP_OK;
}

int lp_step_vital(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_VITAL); }
int lp_step_disease(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_DISEASE); }
int lp_step_ri(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_RI); }
int lp_step_sia(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_SIA); }
int lp_step_transmission(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_TRANSMISSION); }
int lp_log(lp_ctx* c, int32_t t) { return lp_step_stage(c, t, LP_STAGE_LOG); }

int lp_sync(lp_ctx* c) {
  if (!c) return LP_ERR_ARG;
  if (c->prepared && c->rec_dirty) {
    enqueue_sync_columns(c);
    HIPCHK(c, hipGetLastError());
    c->rec_dirty = false;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!c->prepared) return LP_OK;
  int32_t e = 0;
  HIPCHK(c, hipMemcpy(&e, c->d.err, 4, hipMemcpyDeviceToHost));
  if (e & lp::ERR_CAPACITY) return fail(c, LP_ERR_CAPACITY, "a node ran out of agent slots for births");
  if (e & lp::ERR_OVERFLOW) return fail(c, LP_ERR_OVERFLOW, "selected-agent list overflow");
  if (e & lp::ERR_SAMPLING) return fail(c, LP_ERR_STATE, "infection sampling found an inconsistent chunk prefix");
  if (e & lp::ERR_KIND) return fail(c, LP_ERR_STATE, "agent pass launched with the wrong day variant");
  if (e & lp::ERR_LIST) return fail(c, LP_ERR_STATE, "E/I list holds a slot outside the agent capacity");
  return LP_OK;
}

int lp_stop_days(lp_ctx* c, int32_t* out, int32_t n) {
  if (!c || !out || n != c->cfg.n_sims) return LP_ERR_ARG;
  if (!c->prepared) return fail(c, LP_ERR_STATE, "not prepared");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(out, c->d.stop_day, n * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) if (out[i] == INT_MAX) out[i] = -1;
  return LP_OK;
}

const void* lp_debug_buffer(lp_ctx* c, const char* name) {
  if (!c || !name || !c->prepared) return nullptr;
  if (!std::strcmp(name, "p")) return c->d.p_dbg;
  if (!std::strcmp(name, "expected")) return c->d.expected_dbg;
  if (!std::strcmp(name, "total")) return c->d.total_dbg;
  if (!std::strcmp(name, "beta_mixed")) return c->d.beta_mixed;
  if (!std::strcmp(name, "beta")) return c->d.beta_f;
  if (!std::strcmp(name, "mix")) return c->d.mix_buf;
  if (!std::strcmp(name, "size")) return c->d.size;
  if (!std::strcmp(name, "birth_expected")) return c->d.birth_dbg;
  return nullptr;
}

int lp_agent_pass_timing(lp_ctx* c, int64_t* launches, double* total_ms) {
  if (!c || !launches || !total_ms) return LP_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double ms = 0.0;
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float x = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&x, c->ev[i], c->ev[i + 1]));
    ms += x;
  }
  *launches = (int64_t)(c->ev_used / 2);
  *total_ms = ms;
  c->ev_used = 0;
  return LP_OK;
}

int64_t lp_agent_pass_launch_ms(lp_ctx* c, float* ms, int64_t cap) {
  if (!c || (!ms && cap > 0) || cap < 0) return LP_ERR_ARG;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return LP_ERR_HIP;
  const int64_t n = (int64_t)(c->ev_used / 2);
  for (int64_t i = 0; i < n && i < cap; ++i)
    if (hipEventElapsedTime(&ms[i], c->ev[2 * i], c->ev[2 * i + 1]) != hipSuccess) return LP_ERR_HIP;
  return n;
}

int lp_set_timing(lp_ctx* c, int32_t on) {
  if (!c) return LP_ERR_ARG;
  c->timing = on != 0;
  c->ev_used = 0;
  return LP_OK;
}

const char* lp_last_error(lp_ctx* c) { return c ? c->err.c_str() : "null context"; }

void lp_destroy(lp_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < lp::N_KINDS; ++k)
    for (int j = 0; j < kRunLevels; ++j)
      if (c->graph[k][j]) (void)hipGraphExecDestroy(c->graph[k][j]);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
#include "lp_meanfield.hip"
