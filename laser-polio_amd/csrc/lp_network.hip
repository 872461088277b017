// The gravity migration network built on the device (lp_build_network,
// include/lp_step.h), per simulation of the context, into the bound "network_t"
// (float32, each simulation's n x n block transposed) and "keep" (1 - row sum):
//   w_ij = k p_i^a p_j^b / d_ij^c for j != i (non-finite -> 0)     laser_core gravity as
//   w_ij /= (sum_j p_j)^c                                         model.py:1243-1251 calls it
//   rows whose sum exceeds max_rowsum scaled down to it           row_normalizer, model.py:1258
// in float64, cast to float32 at the end, as the host builds it (network.py).  Two
// passes: one workgroup per row for the row sums, scales and keep; then 64 x 64 tiles
// staged in LDS so both the distance reads (row-major) and the network_t writes
// (transposed) are contiguous.
#include "lp_kernels.hip"

namespace lp {

constexpr int NET_TILE = 64;

// w_ij before the row scale; pa_i = p_i^a, the caller's row term.
__device__ __forceinline__ double gravity_w(const lp_gravity_spec& g, double kpa_i, double p_j, double d_ij,
                                            double inv_sc, bool diag) {
  if (diag) return 0.0;
  double w = kpa_i * pow(p_j, g.b) / pow(d_ij, g.c);
  if (!isfinite(w)) w = 0.0;
  return w * inv_sc;
}

// Sum of v over the workgroup, in a fixed order (lane partials, then a tree in LDS).
__device__ double block_sum_f64(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = WG / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double out = red[0];
  __syncthreads();
  return out;
}

// One workgroup per node i (global index): its simulation's population total, its row
// sum, the row scale (row_scale[i]) and keep[i].
__global__ void __launch_bounds__(WG) k_net_rows(Dev d, const lp_gravity_spec* specs, const double* pop,
                                                 const double* dist, double* row_scale, double* inv_sc_out) {
  __shared__ double red[WG];
  const int n = blockIdx.x;
  const int sim = d.node_sim[n];
  const lp_sim_params& sp = d.sims[sim];
  const lp_gravity_spec g = specs[sim];
  const int nb = sp.node_begin, m = sp.n_nodes, i = n - nb;
  double s = 0.0;
  for (int j = threadIdx.x; j < m; j += WG) s += pop[nb + j];
  const double total = block_sum_f64(s, red);
  const double inv_sc = 1.0 / pow(total, g.c);
  const double kpa = g.k * pow(pop[n], g.a);
  const double* drow = dist + sp.net_off + (int64_t)i * m;
  double acc = 0.0;
  for (int j = threadIdx.x; j < m; j += WG) acc += gravity_w(g, kpa, pop[nb + j], drow[j], inv_sc, j == i);
  const double rs = block_sum_f64(acc, red);
  if (threadIdx.x == 0) {
    const double scale = rs > g.max_rowsum ? g.max_rowsum / rs : 1.0;
    row_scale[n] = scale;
    inv_sc_out[n] = inv_sc;
    const_cast<float*>(d.keep)[n] = (float)(1.0 - rs * scale);
  }
}

// Tile (ti, tj) of simulation blockIdx.y: rows i0..i0+63, columns j0..j0+63 of its network,
// written transposed into network_t[net_off + j * m + i].
__global__ void __launch_bounds__(WG) k_net_tiles(Dev d, const lp_gravity_spec* specs, const double* pop,
                                                  const double* dist, const double* row_scale, const double* inv_sc) {
  __shared__ float tile[NET_TILE][NET_TILE + 1];
  const int sim = blockIdx.y;
  const lp_sim_params& sp = d.sims[sim];
  const int m = sp.n_nodes, nb = sp.node_begin;
  const int tiles = (m + NET_TILE - 1) / NET_TILE;
  if ((int)blockIdx.x >= tiles * tiles) return;
  const lp_gravity_spec g = specs[sim];
  const int i0 = ((int)blockIdx.x / tiles) * NET_TILE, j0 = ((int)blockIdx.x % tiles) * NET_TILE;
  const int c = threadIdx.x & (NET_TILE - 1), r0 = threadIdx.x / NET_TILE;   // 4 rows per pass
  // rows of the tile: contiguous distance reads along j
  for (int r = r0; r < NET_TILE; r += WG / NET_TILE) {
    const int i = i0 + r, j = j0 + c;
    float v = 0.0f;
    if (i < m && j < m) {
      const double kpa = g.k * pow(pop[nb + i], g.a);
      const double w = gravity_w(g, kpa, pop[nb + j], dist[sp.net_off + (int64_t)i * m + j], inv_sc[nb + i], i == j);
      v = (float)(w * row_scale[nb + i]);
    }
    tile[r][c] = v;
  }
  __syncthreads();
  // columns of the tile: contiguous network_t writes along i
  float* out = const_cast<float*>(d.network_t) + sp.net_off;
  for (int r = r0; r < NET_TILE; r += WG / NET_TILE) {
    const int j = j0 + r, i = i0 + c;
    if (i < m && j < m) out[(int64_t)j * m + i] = tile[c][r];
  }
}

}  // namespace lp
