/*
 * lp_step.h -- C ABI of the MI355X (gfx950) daily agent step of laser-polio.
 *
 * The library (liblpstep.so, built from laser-polio_amd/csrc) runs one or many
 * simulated days of the laser-polio agent-based model on one GPU.  Every agent,
 * node, schedule and results buffer is owned by the caller (device memory,
 * e.g. torch tensors) and handed over as a plain pointer + element count; the
 * library owns only scratch buffers, its HIP stream and its captured day graph.
 *
 * Reference interface replaced (InstituteforDiseaseModeling/laser-polio 0.2.31):
 *   lp_step_vital         VitalDynamics_ABM.step      src/laser_polio/model.py:1689-1757
 *                         (+ get_deaths               model.py:1767-1781)
 *   lp_step_disease       DiseaseState_ABM.step       model.py:715-795
 *                         (+ disease_state_step(_kernel) model.py:344-454)
 *   lp_step_ri            RI_ABM.step + fast_ri       model.py:1920-1985, 1784-1855
 *   lp_step_sia           SIA_ABM.step + fast_sia     model.py:2099-2156, 1994-2060
 *   lp_step_transmission  Transmission_ABM.step       model.py:1281-1459
 *                         (+ tx_step_prep model.py:932-1007, tx_infect_nb model.py:1010-1149)
 *   lp_log                Transmission_ABM.log + count_SEIRP  model.py:1461-1483, 869-929
 *   lp_run_days           SEIR_ABM.run day loop       model.py:246-282 (all of the above, fused)
 * The Python host package (laser-polio_amd) keeps the reference's component
 * classes, their names, run order, step()/log() methods and results arrays, and
 * calls these entry points through ctypes.
 *
 * Conventions: every entry point returns 0 (LP_OK) or a negative LP_ERR_* code;
 * lp_last_error() gives the message.  Calls on one context are not re-entrant
 * and come from one host thread.  Work is queued on the context's stream;
 * lp_sync() waits for it and reports device-side errors (capacity exhausted,
 * scratch overflow).
 */
#ifndef LP_STEP_H
#define LP_STEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LP_ABI_VERSION 1
#define LP_TILE 1024      /* agents per tile; node segments are multiples of this */
#define LP_CHUNK 256      /* agents per sampling chunk (one wavefront x 4 agents) */
#define LP_MAX_STRAINS 4
#define LP_KM_LEN 101     /* cumulative-deaths table length (ages 0..100 y) */

enum {
  LP_OK = 0,
  LP_ERR_ARG = -1,       /* bad argument / unknown binding name / wrong dtype or size */
  LP_ERR_HIP = -2,       /* HIP runtime error */
  LP_ERR_CAPACITY = -3,  /* a node ran out of agent slots for births */
  LP_ERR_STATE = -4,     /* call order violated (e.g. run before lp_prepare) */
  LP_ERR_UNBOUND = -5,   /* a required buffer was not bound */
  LP_ERR_OVERFLOW = -6   /* a scratch list overflowed (too many exposures in one day) */
};

enum lp_dtype { LP_I8 = 1, LP_U8 = 2, LP_I16 = 3, LP_U16 = 4, LP_I32 = 5, LP_U32 = 6,
                LP_I64 = 7, LP_U64 = 8, LP_F32 = 9,
                LP_F64 = 10 /* held as bit patterns: no kernel does float64 arithmetic */ };

/* Stages (bit mask) -- one per reference component, plus the log tally. */
enum {
  LP_STAGE_VITAL = 1 << 0,
  LP_STAGE_DISEASE = 1 << 1,
  LP_STAGE_RI = 1 << 2,
  LP_STAGE_SIA = 1 << 3,
  LP_STAGE_TRANSMISSION = 1 << 4,
  LP_STAGE_LOG = 1 << 5,
  LP_STAGE_ALL = 0x3f
};

/* Per-day schedule flags (int32 array "day_flags", one entry per tick). */
enum {
  LP_DAY_VD_STEP = 1 << 0,   /* VitalDynamics step day: deaths + births + pop update */
  LP_DAY_VD_COPY = 1 << 1,   /* VitalDynamics present, non-step day: pop[t] = pop[t-1] */
  LP_DAY_DS = 1 << 2,        /* DiseaseState present */
  LP_DAY_RI = 1 << 3,        /* RI step day (t % step == 0 and vx_prob_ri set) */
  LP_DAY_SIA = 1 << 4,       /* one or more SIA events today */
  LP_DAY_TX = 1 << 5,        /* Transmission present (step + log) */
  LP_DAY_STOPCHK = 1 << 6    /* DiseaseState early-stop rule enabled */
};

typedef struct {
  int64_t capacity;        /* agent slots over all sims; multiple of LP_TILE, below 2^32 */
  int32_t n_nodes;         /* nodes over all sims (node ids are global) */
  int32_t n_sims;          /* independent simulations in the batch */
  int32_t n_strains;       /* 1..LP_MAX_STRAINS */
  int32_t nt;              /* ticks = dur + 1, at most 2^23 */
  int32_t n_sia_events;    /* rows of the SIA event tables (may be 0) */
  int32_t tiles_per_block; /* tiles of 1024 slots per span: 1, 2, 4, 8 or 16 (0 = default 16) */
  int32_t mix_mode;        /* network mix: 0 = per-node fixed-order tree reduction (the oracle's order,
                              bit-identical replays); 1 = every simulation's prevalence x network
                              contraction on MFMA (within 1e-6 of float64, not replay-identical) */
  int32_t ei_grid;         /* workgroups of the E / I list pass, 64..65535 (0 = those resident at once, <= 2048) */
  int32_t sample_grid;     /* workgroups of the round-1 sampler, 64..65535 (0 = those resident at once, <= 2048) */
  int32_t commit_grid;     /* unused (the round-1 claims run inside the sampler); 0 or 64..65535 */
  int32_t reserved[4];
} lp_config;

/* Per-simulation parameters; array bound as "sim_params" (LP_U8, n_sims * sizeof). */
typedef struct {
  uint64_t seed;           /* Philox key */
  int64_t slot_base;       /* first agent slot of the sim (Philox agent index origin) */
  int64_t net_off;         /* offset of the sim's n x n transposed network in "network_t" */
  int32_t node_begin;      /* first global node id of the sim */
  int32_t n_nodes;
  float strain_r0[LP_MAX_STRAINS];
  uint64_t thr_paralysis;  /* ceil(p_paralysis * 2^32) */
  uint64_t thr_zi;         /* ceil(node_seeding_zero_inflation * 2^32) */
  float zi;                /* node_seeding_zero_inflation */
  int32_t disp_r;          /* max(1, round(node_seeding_dispersion)) */
  int32_t ri_strain;       /* strain given by RI OPV (1 Sabin, 2 nOPV2) */
  int32_t ri_step;         /* step_size_RI_ABM */
  int32_t newborn_ri_timer;/* ri_timer given to newborns; -1 when RI_ABM is absent */
  int32_t last_seed_day;   /* last tick with a seed_schedule entry, -1 if none */
} lp_sim_params;

typedef struct lp_ctx lp_ctx;

/* Library / struct layout checks for bindings. */
int32_t lp_abi_version(void);
int64_t lp_sizeof_sim_params(void);

/* Create a context for the given sizes.  Allocates nothing on the device yet. */
int lp_create(const lp_config* cfg, lp_ctx** out);

/* Bind a caller-owned device buffer by name (agent columns, node tables, sim
 * parameters, schedule, results).  The name table is listed in lp_bind_names(). */
int lp_bind(lp_ctx* ctx, const char* name, void* device_ptr, int32_t dtype, int64_t n_elems);
const char* lp_bind_names(void);

/* Validate bindings, build the tile/span tables from the node segments and
 * allocate scratch.  Must follow the last lp_bind and precede any step. */
int lp_prepare(lp_ctx* ctx);

/* Use this HIP stream (hipStream_t, owned by the caller; NULL = the default
 * stream) instead of the context's own.  Waits for the work queued on the
 * current stream, destroys the context's own stream if it still has it, and
 * never destroys the caller's.  Must precede the first graph capture.  On the
 * default stream lp_run_days launches the day kernels directly (a hipGraph is
 * not captured there). */
int lp_set_stream(lp_ctx* ctx, void* hip_stream);
/* 1 if the context still uses (and will destroy) its own stream, else 0. */
int lp_owns_stream(lp_ctx* ctx);

/* Fused days: runs ticks [t_begin, t_end) with every stage the schedule enables,
 * in reference order (VitalDynamics, DiseaseState, RI, SIA, Transmission, log).
 * Tick 0 only logs the initial state.  Days after a sim's early stop are no-ops.
 * use_graph != 0 replays captured hipGraphs: one per day kind, and one per run of 2, 4 or
 * 8 consecutive days of one kind. */
int lp_run_days(lp_ctx* ctx, int32_t t_begin, int32_t t_end, int32_t use_graph);

/* Single stages on tick t (any OR of LP_STAGE_*), for callers that step
 * components one at a time.  Same kernels as the fused path, masked. */
int lp_step_stage(lp_ctx* ctx, int32_t t, int32_t stage_mask);
int lp_step_vital(lp_ctx* ctx, int32_t t);
int lp_step_disease(lp_ctx* ctx, int32_t t);
int lp_step_ri(lp_ctx* ctx, int32_t t);
int lp_step_sia(lp_ctx* ctx, int32_t t);
int lp_step_transmission(lp_ctx* ctx, int32_t t);
int lp_log(lp_ctx* ctx, int32_t t);

/* The caller changed bound buffers between steps (the reference lets host code
 * edit sim.people / sim.results between component steps, e.g. seed_schedule
 * injections, model.py:772-786).  The next agent pass recounts the per-node and
 * per-chunk susceptible statistics the library otherwise carries from day to day. */
int lp_invalidate(lp_ctx* ctx);

/* Population initialisation on the device, in place of the host draws of the
 * components' constructors: SEIR_ABM.__init__ / DiseaseState_ABM init (ages,
 * immunity, initial infections, disease timers; model.py:561-713, 568-587),
 * populate_heterogeneous_values (model.py:816-866), VitalDynamics_ABM death dates
 * (model.py:1548-1669) and RI_ABM timers (model.py:1887-1894).  Every slot of every
 * node segment is written: initial agents [seg_begin, live_end), unborn slots after
 * them (timers and multipliers drawn, the rest at their defaults).  Draws come from
 * Philox streams keyed by (seed, sim-local slot): the distributions of the host
 * generator, not its numbers.  Exact-count selections (initial infections) are made
 * by the caller and passed as slots.  Call after lp_prepare, before the first step. */
enum { LP_DIST_CONSTANT = 0, LP_DIST_NORMAL = 1, LP_DIST_GAMMA = 2, LP_DIST_LOGNORMAL = 3, LP_DIST_POISSON = 4,
       LP_DIST_EXPONENTIAL = 5, LP_DIST_UNIFORM_INT = 6 };
/* constant(a), normal(mean a, std b), gamma(shape a, scale b), lognormal(mean a, std b),
 * poisson(lam a), exponential(scale a), uniform integers [a, b) */
typedef struct { int32_t kind; float a, b; } lp_dist;
#define LP_INIT_BANDS 8
#define LP_INIT_PYR 32
typedef struct {
  uint64_t seed;              /* Philox key of the initial draws */
  int32_t age_model;          /* 0: exponential (age_mean_days, capped at age_cap_days); 1: pyramid bins */
  float age_mean_days, age_cap_days;
  int32_t n_pyr;              /* pyramid: bin b drawn with cumulative weight pyr_cdf[b], age uniform in */
  float pyr_cdf[LP_INIT_PYR]; /* [pyr_lo[b], pyr_hi[b]) days */
  int32_t pyr_lo[LP_INIT_PYR], pyr_hi[LP_INIT_PYR];
  float init_immun;           /* probability that an initial agent starts recovered */
  int32_t n_bands;            /* recovered by age: age in [band_lo, band_hi) days with probability band_frac */
  int32_t band_lo[LP_INIT_BANDS], band_hi[LP_INIT_BANDS];
  float band_frac[LP_INIT_BANDS];
  lp_dist dur_exp, dur_inf, t_to_paralysis;
  int32_t het_model;          /* 0: multipliers 1 and inf_mean; 1: both lognormal; 2: lognormal risk, gamma(1) infectivity */
  float risk_mu, risk_sigma;  /* log-mean and log-sd of the risk (and, model 1, of infectivity / inf_mean) */
  float inf_mean;             /* mean daily infectivity */
  float rho;                  /* normal correlation of the copula, 2 sin(pi corr / 6) */
  int32_t ri;                 /* 1: ri_timer = date_of_birth + U[ri_lo, ri_hi) days */
  float ri_lo, ri_hi;
  int32_t vd;                 /* 1: date_of_death from the sim's mortality table (km_cd) */
  float missed_frac;          /* probability of being chronically missed */
  int32_t reserved[3];
} lp_init_sim;
int64_t lp_sizeof_init_sim(void);
int lp_init_population(lp_ctx* ctx, const lp_init_sim* sims, int32_t n_sims, const int64_t* infected_slots,
                       int64_t n_infected);

/* Gravity network of one simulation (Transmission_ABM._initialize_common, model.py:1213-1258:
 * gravity(init_pops, dist, k * 10^k_exponent, a, b, c) / sum(init_pops)^c, then
 * row_normalizer(net, max_migr_frac)). */
typedef struct {
  double k;            /* gravity_k * 10^gravity_k_exponent */
  double a, b, c;
  double max_rowsum;   /* max_migr_frac */
  int32_t reserved[2];
} lp_gravity_spec;

/* Build every simulation's network on the device into the bound "network_t" (each
 * simulation's block transposed, float32) and "keep" (1 - row sum), from one spec per
 * simulation (host array), the node populations (device, float64, one per node of the
 * context) and the distances (device, float64, each simulation's n x n block row-major at
 * the same offset as its "network_t" block).  Replaces the host construction of
 * model.py:1213-1258 for calibration loops that rebuild networks per trial.  Takes effect
 * from the next simulated day; queued on the context stream. */
int lp_build_network(lp_ctx* ctx, const lp_gravity_spec* specs, int32_t n_sims, const double* pop,
                     const double* dist);

/* Wait for queued work; returns LP_ERR_CAPACITY / LP_ERR_OVERFLOW if the device flagged one. */
int lp_sync(lp_ctx* ctx);

/* Per-sim status after lp_sync: out[s] = tick after which sim s stopped early, or -1. */
int lp_stop_days(lp_ctx* ctx, int32_t* out, int32_t n_sims);

/* Diagnostics of the last transmission step (device pointers of library scratch):
 * "p" float[n_nodes*n_strains], "expected" float[n_nodes], "total" int64[n_nodes],
 * "beta_mixed" float[n_nodes*n_strains]; of the last VitalDynamics step:
 * "birth_expected" int64[n_nodes][2] = floor and Bernoulli threshold ceil(frac * 2^32)
 * of the expected births; "beta" float[n_nodes][4], the infectious weight per node and
 * strain slot the network mix reads; with mix_mode 1, "mix" float[n_nodes][4], the
 * contraction sum_j N_ji beta_j done on MFMA.  Returns NULL for unknown names. */
const void* lp_debug_buffer(lp_ctx* ctx, const char* name);

/* Timing of the agent-pass kernel: with timing on, lp_run_days launches the day
 * kernels directly (no graph) and brackets every agent-pass launch with HIP
 * events on the context stream; lp_agent_pass_timing returns the number of
 * bracketed launches and their summed milliseconds since the last call. */
int lp_set_timing(lp_ctx* ctx, int32_t on);
int lp_agent_pass_timing(lp_ctx* ctx, int64_t* launches, double* total_ms);
/* The milliseconds of each bracketed launch since the last lp_agent_pass_timing call, in
 * launch order (up to cap of them written to ms); returns their number, or a negative
 * LP_ERR_* code.  Does not reset the record.  Diagnostics only: no reference
 * counterpart. */
int64_t lp_agent_pass_launch_ms(lp_ctx* ctx, float* ms, int64_t cap);

const char* lp_last_error(lp_ctx* ctx);
void lp_destroy(lp_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* LP_STEP_H */
