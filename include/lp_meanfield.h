/*
 * lp_meanfield.h -- C ABI of the mean-field rollout of laser-polio's node-level
 * transmission map and of its hand-written adjoint (liblpstep.so, gfx950).
 *
 * The map is the per-node force of infection of Transmission_ABM.step Part 2b
 * (src/laser_polio/model.py:1327-1363): infectious weight beta = iota * I,
 * network mix  m_i = beta_i * (1 - sum_j N_ij) + sum_j N_ji * beta_j,
 * g_i = m_i * season_t * r0_scalar_i / max(pop_i, 1),  p_i = 1 - exp(-g_i),
 * expected new exposures lambda_i = rho_i * S_i * p_i (rho = mean acquisition risk
 * of the node's susceptibles).  The rollout advances node compartments with
 * daily exits E -> I at rate 1/tau_E and I -> R at rate 1/tau_I:
 *   S' = S - lambda,  E' = E + lambda - E/tau_E,  I' = I + E/tau_E - I/tau_I.
 * The reference has no differentiable counterpart (it runs the stochastic ABM
 * only); these entry points exist for gradient-based calibration of r0_scalars,
 * the gravity parameters and seasonality against ABM output.
 *
 * Arrays are device pointers (float32, node-major rows of n).  All work is
 * queued on `hip_stream` (NULL = default stream).  Returns 0 or a negative
 * LP_ERR_* code of lp_step.h.
 */
#ifndef LP_MEANFIELD_H
#define LP_MEANFIELD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n;                /* nodes */
  int32_t reserved;
  const float* net;         /* [n*n] N[i*n+j] = migration fraction from node i to node j */
  const float* net_t;       /* [n*n] its transpose */
  const float* keep;        /* [n] 1 - sum_j N[i*n+j] */
  const float* r0_scalar;   /* [n] */
  const float* rho;         /* [n] mean acquisition risk of the node's susceptibles */
  const float* inv_pop;     /* [n] 1 / max(pop, 1) */
  const float* season;      /* [days] seasonality factor of each rolled day */
  float iota;               /* daily infectious weight per infectious agent */
  float inv_tau_e;          /* 1 / mean exposure duration (days) */
  float inv_tau_i;          /* 1 / mean infectious duration (days) */
  float pad;
} lp_mf_model;

/* Days [t0, t1) forward.  S, E, I: [(t1-t0+1)*n] trajectories whose row 0 holds the
 * state at t0 (rows 1.. are written); inc: [(t1-t0)*n] expected new exposures of
 * each day.  m_save, g_save (optional, may be NULL): [(t1-t0)*n] the mixed weight m
 * and the rate g of each day, for lp_mf_adjoint_days. */
int lp_mf_days(const lp_mf_model* model, int32_t t0, int32_t t1, float* S, float* E, float* I, float* inc,
               float* m_save, float* g_save, void* hip_stream);

/* Reverse-mode adjoint of days [t0, t1), given the forward trajectories of those days
 * (as written by lp_mf_days with m_save / g_save) and the loss adjoints a_inc
 * [(t1-t0)*n] of the daily exposures.  aS, aE, aI [n]: on entry the adjoints of the
 * state at t1, on return those of the state at t0.  Accumulates (+=) into r0_bar [n]
 * and net_bar [n*n] (the adjoint of N, diagonal untouched) and writes season_part
 * [(t1-t0)*n], whose row sums are the adjoints of season_t.  work: [2*n] scratch. */
int lp_mf_adjoint_days(const lp_mf_model* model, int32_t t0, int32_t t1, const float* S, const float* I,
                       const float* m_save, const float* g_save, const float* a_inc, float* aS, float* aE,
                       float* aI, float* r0_bar, float* net_bar, float* season_part, float* work,
                       void* hip_stream);

/* lp_mf_adjoint_days without the `work` argument: the same kernels, with their scratch
 * in a fixed device array of the library (2 * LP_MF_NOALLOC_MAX_NODES floats, 256 KiB
 * per GPU, held until the process ends).  It allocates nothing (no hipMalloc, no
 * library allocation), for callers that cannot: code under hipGraph capture, tight
 * calibration loops, harnesses that forbid allocation.  No counterpart in the
 * reference.  Returns LP_ERR_ARG if model->n > LP_MF_NOALLOC_MAX_NODES.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialised by the caller. */
#define LP_MF_NOALLOC_MAX_NODES 32768
int lp_mf_adjoint_days_noalloc(const lp_mf_model* model, int32_t t0, int32_t t1, const float* S, const float* I,
                               const float* m_save, const float* g_save, const float* a_inc, float* aS, float* aE,
                               float* aI, float* r0_bar, float* net_bar, float* season_part, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* LP_MEANFIELD_H */
