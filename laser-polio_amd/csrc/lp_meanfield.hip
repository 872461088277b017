// Mean-field rollout of the node-level transmission map and its adjoint
// (C ABI in include/lp_meanfield.h).  One workgroup per node for the network
// matvecs (the transposed network row is read coalesced), fixed-order LDS
// reductions, elementwise kernels for the compartment updates.
#include "../../include/lp_meanfield.h"
#include <mutex>

namespace lp_mf {

constexpr int WG = 256;

// Scratch of lp_mf_adjoint_days_noalloc: the adjoint's two per-node work rows
// (mixed-weight adjoint, carried I adjoint), 2 * LP_MF_NOALLOC_MAX_NODES floats
// = 256 KiB held by every GPU the library runs on until the process ends.
static __device__ __align__(256) float scratch[2 * LP_MF_NOALLOC_MAX_NODES];

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per device
// (each device holds its own copy of the array at its own address).
int scratch_of_current_device(float** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return LP_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return LP_ERR_HIP;
    g_scratch[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch[dev];
  return LP_OK;
}

// Fixed-order sum of the 256 per-thread partials of a workgroup.
__device__ __forceinline__ float wg_sum(float v, float* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = WG / 2; o > 0; o >>= 1) {
    if (tid < o) sh[tid] = sh[tid] + sh[tid + o];
    __syncthreads();
  }
  return sh[0];
}

// One day forward: row t of S/E/I -> row t+1; inc, m, g of the day.
__global__ void __launch_bounds__(WG) k_mf_day(lp_mf_model m, int t, const float* S0, const float* E0,
                                               const float* I0, float* S1, float* E1, float* I1, float* inc,
                                               float* msave, float* gsave) {
  __shared__ float sh[WG];
  const int i = blockIdx.x, n = m.n;
  const float* row = m.net_t + (size_t)i * n;
  float acc = 0.0f;
  for (int j = threadIdx.x; j < n; j += WG) acc = fmaf(row[j], I0[j], acc);
  const float mix = wg_sum(acc, sh);
  if (threadIdx.x != 0) return;
  const float season_t = m.season[t];
  const float mm = m.iota * fmaf(I0[i], m.keep[i], mix);
  const float g = ((mm * season_t) * m.r0_scalar[i]) * m.inv_pop[i];
  const float p = -lp::expm1f_lp(-g);
  const float lam = (m.rho[i] * S0[i]) * p;
  S1[i] = S0[i] - lam;
  E1[i] = (E0[i] + lam) - E0[i] * m.inv_tau_e;
  I1[i] = (I0[i] + E0[i] * m.inv_tau_e) - I0[i] * m.inv_tau_i;
  inc[i] = lam;
  if (msave) msave[i] = mm;
  if (gsave) gsave[i] = g;
}

// Adjoint, elementwise part: exposures, rate, r0 and season contributions; the
// adjoint of the mixed weight goes to mbar, the I adjoint's carried part to aIp.
__global__ void __launch_bounds__(WG) k_mf_adj_a(lp_mf_model m, int t, const float* S0, const float* msave,
                                                 const float* gsave, const float* ainc, float* aS, float* aE,
                                                 const float* aI, float* r0_bar, float* season_part, float* mbar,
                                                 float* aIp) {
  const int i = blockIdx.x * WG + threadIdx.x;
  if (i >= m.n) return;
  const float season_t = m.season[t];
  const float g = gsave[i], mm = msave[i];
  const float p = -lp::expm1f_lp(-g);
  const float q = lp::expf_lp(-g);
  const float lam_bar = (ainc[i] - aS[i]) + aE[i];
  const float s_bar = aS[i] + (lam_bar * m.rho[i]) * p;
  const float g_bar = ((lam_bar * m.rho[i]) * S0[i]) * q;
  r0_bar[i] += ((g_bar * mm) * season_t) * m.inv_pop[i];
  season_part[i] = ((g_bar * mm) * m.r0_scalar[i]) * m.inv_pop[i];
  mbar[i] = ((g_bar * season_t) * m.r0_scalar[i]) * m.inv_pop[i];
  const float e_bar = aE[i] * (1.0f - m.inv_tau_e) + aI[i] * m.inv_tau_e;
  aIp[i] = aI[i] * (1.0f - m.inv_tau_i);
  aS[i] = s_bar;
  aE[i] = e_bar;
}

// Adjoint, network part (one workgroup per node j): beta_bar_j = keep_j mbar_j +
// sum_i N_ji mbar_i; aI_j = aIp_j + iota beta_bar_j; N_bar[j][i] += beta_j (mbar_i - mbar_j).
__global__ void __launch_bounds__(WG) k_mf_adj_b(lp_mf_model m, const float* I0, const float* mbar, const float* aIp,
                                                 float* aI, float* net_bar) {
  __shared__ float sh[WG];
  const int j = blockIdx.x, n = m.n;
  const float* row = m.net + (size_t)j * n;
  float acc = 0.0f;
  for (int i = threadIdx.x; i < n; i += WG) acc = fmaf(row[i], mbar[i], acc);
  const float s = wg_sum(acc, sh);
  const float beta = m.iota * I0[j];
  const float mj = mbar[j];
  float* nb = net_bar + (size_t)j * n;
  for (int i = threadIdx.x; i < n; i += WG)
    if (i != j) nb[i] += beta * (mbar[i] - mj);
  if (threadIdx.x == 0) aI[j] = aIp[j] + m.iota * fmaf(m.keep[j], mj, s);
}

}  // namespace lp_mf

extern "C" {

int lp_mf_days(const lp_mf_model* model, int32_t t0, int32_t t1, float* S, float* E, float* I, float* inc,
               float* m_save, float* g_save, void* hip_stream) {
  if (!model || model->n <= 0 || t1 < t0 || !S || !E || !I || !inc) return LP_ERR_ARG;
  const lp_mf_model m = *model;
  const int n = m.n;
  hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
  for (int d = 0; d < t1 - t0; ++d) {
    const size_t r0 = (size_t)d * n, r1 = (size_t)(d + 1) * n;
    hipLaunchKernelGGL(lp_mf::k_mf_day, dim3(n), dim3(lp_mf::WG), 0, s, m, t0 + d, S + r0, E + r0, I + r0,
                       S + r1, E + r1, I + r1, inc + r0, m_save ? m_save + r0 : nullptr, g_save ? g_save + r0 : nullptr);
  }
  return hipGetLastError() == hipSuccess ? LP_OK : LP_ERR_HIP;
}

int lp_mf_adjoint_days_noalloc(const lp_mf_model* model, int32_t t0, int32_t t1, const float* S, const float* I,
                               const float* m_save, const float* g_save, const float* a_inc, float* aS, float* aE,
                               float* aI, float* r0_bar, float* net_bar, float* season_part, void* hip_stream) {
  if (!model || model->n <= 0 || model->n > LP_MF_NOALLOC_MAX_NODES) return LP_ERR_ARG;
  float* work = nullptr;
  const int rc = lp_mf::scratch_of_current_device(&work);
  if (rc) return rc;
  // every work row is written (k_mf_adj_a) before it is read (k_mf_adj_b), so the
  // previous call's contents need no clearing
  return lp_mf_adjoint_days(model, t0, t1, S, I, m_save, g_save, a_inc, aS, aE, aI, r0_bar, net_bar,
                            season_part, work, hip_stream);
}

int lp_mf_adjoint_days(const lp_mf_model* model, int32_t t0, int32_t t1, const float* S, const float* I,
                       const float* m_save, const float* g_save, const float* a_inc, float* aS, float* aE,
                       float* aI, float* r0_bar, float* net_bar, float* season_part, float* work,
                       void* hip_stream) {
  if (!model || model->n <= 0 || t1 < t0 || !S || !I || !m_save || !g_save || !a_inc || !aS || !aE || !aI ||
      !r0_bar || !net_bar || !season_part || !work)
    return LP_ERR_ARG;
  const lp_mf_model m = *model;
  const int n = m.n;
  hipStream_t s = reinterpret_cast<hipStream_t>(hip_stream);
  float* mbar = work;
  float* aIp = work + n;
  for (int d = t1 - t0 - 1; d >= 0; --d) {
    const size_t r = (size_t)d * n;
    hipLaunchKernelGGL(lp_mf::k_mf_adj_a, dim3((n + lp_mf::WG - 1) / lp_mf::WG), dim3(lp_mf::WG), 0, s, m, t0 + d,
                       S + r, m_save + r, g_save + r, a_inc + r, aS, aE, aI, r0_bar, season_part + r, mbar, aIp);
    hipLaunchKernelGGL(lp_mf::k_mf_adj_b, dim3(n), dim3(lp_mf::WG), 0, s, m, I + r, mbar, aIp, aI, net_bar);
  }
  return hipGetLastError() == hipSuccess ? LP_OK : LP_ERR_HIP;
}

}  // extern "C"
