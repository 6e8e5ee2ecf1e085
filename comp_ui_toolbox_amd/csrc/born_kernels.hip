// HIP kernels of the Born operator pair for gfx950 (CDNA4).
//
// Operator (restated from operator_born.m:353-459 / cu:1099-1124):
//   u[m,k,f,r] = sum_a pref[m,k][f,a] * sum_p h_ref[fg, shift(p,e_a)] * p_inc[m][fm, p] * gamma[p,r]
// Factored as V[pair][e][r] = sum_p h_e[fg, p] * p_inc[pair, p] * gamma[p, r]
// (pair = (measurement, unique frequency of the measurement)), followed by a
// small mixing step.  On the structured 2-D path, for a fixed (frequency,
// depth row) the element transfer functions are one lateral row of h_ref
// read at offsets i - e*Np (mirrored at the left edge), so the per-row work is
// a Toeplitz-in-x complex GEMM: elements x points x (pairs*rhs) columns.
#include "born_internal.h"

#include <cmath>
#include <cstdio>

namespace born {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr double DBL_EPSILON_D = 2.220446049250313080847e-16;
constexpr int NCS = NC + 1;  // padded LDS row stride of column tiles

__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- layout
// MATLAB column-major complex double (rows fastest) -> row-major complex64
__global__ void k_convert_transpose(const double2* __restrict__ src, int rows, int cols,
                                    float2* __restrict__ dst) {
  __shared__ float2 tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;
    if (c < cols && r < rows) {
      const double2 v = src[(size_t)c * rows + r];
      tile[j][tx] = make_float2((float)v.x, (float)v.y);
    }
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    if (c < cols && r < rows) dst[(size_t)r * cols + c] = tile[tx][j];
  }
}

void launch_convert_transpose(const double2* src, int rows, int cols, float2* dst, hipStream_t s) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32), block(32, 8);
  hipLaunchKernelGGL(k_convert_transpose, grid, block, 0, s, src, rows, cols, dst);
}

__global__ void k_f64_to_f32(const double2* __restrict__ src, size_t n, float2* __restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

void launch_f64_to_f32(const double2* src, size_t n, float2* dst, hipStream_t s) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks), dim3(256), 0, s, src, n, dst);
}

// Occupancy (cu:466-488): a point takes part iff some column has |Re| or
// |Im| above DBL_EPSILON; other points are zeroed.
__global__ void k_mask_gamma(const float2* __restrict__ g, int N_pts, int N_rhs, float2* __restrict__ gm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N_pts) return;
  bool occ = false;
  for (int r = 0; r < N_rhs; ++r) {
    const float2 v = g[(size_t)r * N_pts + p];
    occ |= (fabs((double)v.x) > DBL_EPSILON_D) || (fabs((double)v.y) > DBL_EPSILON_D);
  }
  for (int r = 0; r < N_rhs; ++r) {
    const float2 v = g[(size_t)r * N_pts + p];
    gm[(size_t)r * N_pts + p] = occ ? v : make_float2(0.f, 0.f);
  }
}

// Zero fill by a kernel (no memset node, so the clear is safe to capture).
__global__ void k_zero_f2(float2* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = make_float2(0.f, 0.f);
}

void launch_zero_f2(float2* x, size_t n, hipStream_t s) {
  if (n == 0) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_f2, dim3((unsigned)blocks), dim3(256), 0, s, x, n);
}

void launch_mask_gamma(const float2* g, int N_pts, int N_rhs, float2* gm, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_gamma, dim3((N_pts + 255) / 256), dim3(256), 0, s, g, N_pts, N_rhs, gm);
}

// Row-block partials -> V[pair][e][r] (FP64 sum): V at element e = ex + Nex * ey
// of the pair and rhs in column pos of chunk ch is the sum over the chunk's
// blocks of slice ey.  Threads run over (chunk, slice, element, column), so
// neighbouring threads read neighbouring partials.
__global__ void k_fwd_reduce(const float2* __restrict__ vpart, const Chunk* __restrict__ chunks,
                             const int* __restrict__ col_pair, int n_chunks, int NZB, int Ne_pad, int Nex, int Ney,
                             int N_e, int N_rhs, double2* __restrict__ V, int expand_rhs) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)n_chunks * Ney * Ne_pad * NC) return;
  const int pos = (int)(idx % NC);
  size_t t = idx / NC;
  const int ex = (int)(t % Ne_pad);
  t /= Ne_pad;
  const int ey = (int)(t % Ney);
  const int ch = (int)(t / Ney);
  const Chunk c = chunks[ch];
  if (pos >= c.ncols || ex >= Nex) return;
  double sx = 0.0, sy = 0.0;
  const float2* src = vpart + (((size_t)ch * Ney + ey) * NZB * Ne_pad + ex) * NC + pos;
  for (int zb = 0; zb < NZB; ++zb) {
    const float2 v = src[(size_t)zb * Ne_pad * NC];
    sx += v.x;
    sy += v.y;
  }
  // expanded columns (expand_rhs > 0) carry pair * expand_rhs + rhs
  const int col = col_pair[c.col_start + pos];
  const int pair = expand_rhs > 0 ? col / expand_rhs : col;
  const int r = expand_rhs > 0 ? col - pair * expand_rhs : c.r;
  V[((size_t)pair * N_e + ey * Nex + ex) * N_rhs + r] = make_double2(sx, sy);
}

void launch_fwd_reduce_expanded(const Plan& P, int nblk, hipStream_t s) {
  const int Ney = P.nd == 3 ? P.Ney : 1;
  const int n_chunks = (int)P.fchunks.size();
  const size_t total = (size_t)n_chunks * Ney * P.Ne_pad * NC;
  hipLaunchKernelGGL(k_fwd_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P.d_fvpart, P.d_fchunks,
                     P.d_fcol, n_chunks, nblk, P.Ne_pad, P.Nex, Ney, P.N_e, P.ws_rhs, P.d_V,
                     P.fwd_table_expanded ? P.ws_rhs : 0);
}

// pexp[pair * N_rhs + r][p] = p_inc[pair][p] * gamma_masked[r][p]
__global__ void k_expand_pg(const float2* __restrict__ Pt, const float2* __restrict__ gm, int N_pairs, int N_rhs,
                            int N_pts, float2* __restrict__ pexp) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)N_pairs * N_pts) return;
  const int p = (int)(idx % N_pts);
  const size_t pair = idx / N_pts;
  const float2 a = Pt[idx];
  for (int r = 0; r < N_rhs; ++r) {
    const float2 g = gm[(size_t)r * N_pts + p];
    pexp[(pair * N_rhs + r) * N_pts + p] = make_float2(a.x * g.x - a.y * g.y, a.x * g.y + a.y * g.x);
  }
}

void launch_expand_pg(const Plan& P, hipStream_t s) {
  const size_t total = (size_t)P.N_pairs * P.N_pts;
  hipLaunchKernelGGL(k_expand_pg, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P.d_p, P.d_gm, P.N_pairs,
                     P.ws_rhs, P.N_pts, P.d_pexp);
}

void launch_fwd_reduce(const Plan& P, int N_rhs, int nblk, hipStream_t s) {
  (void)N_rhs;   // the rhs of every column comes from its chunk
  const int Ney = P.nd == 3 ? P.Ney : 1;
  const int n_chunks = (int)P.chunks.size();
  const size_t total = (size_t)n_chunks * Ney * P.Ne_pad * NC;
  hipLaunchKernelGGL(k_fwd_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P.d_vpart, P.d_chunks,
                     P.d_col_pair, n_chunks, nblk, P.Ne_pad, P.Nex, Ney, P.N_e, P.ws_rhs, P.d_V, 0);
}

// ------------------------------------------------------ general forward
// Any shift map: one block per (pair, element, rhs) gathers h through the
// map (the reference's Phi entries, cu:1114-1120, never materialized).
__global__ __launch_bounds__(256) void k_fwd_general(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ gm,
    const int* __restrict__ shift, const int* __restrict__ pair_fg, int N_pts, int H_plane, int N_e, int N_rhs,
    double2* __restrict__ V) {
  const int pair = blockIdx.x, e = blockIdx.y, r = blockIdx.z;
  const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
  const int* sh = shift + (size_t)e * N_pts;
  const float2* pr = P + (size_t)pair * N_pts;
  const float2* g = gm + (size_t)r * N_pts;
  float sx = 0.f, sy = 0.f;
  for (int p = threadIdx.x; p < N_pts; p += 256) {
    const float2 gv = g[p];
    if (gv.x == 0.f && gv.y == 0.f) continue;
    const float2 a = hr[sh[p]], b = pr[p];
    const float abx = a.x * b.x - a.y * b.y, aby = a.x * b.y + a.y * b.x;
    sx += abx * gv.x - aby * gv.y;
    sy += abx * gv.y + aby * gv.x;
  }
  __shared__ double rx[256], ry[256];
  rx[threadIdx.x] = sx;
  ry[threadIdx.x] = sy;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rx[threadIdx.x] += rx[threadIdx.x + o];
      ry[threadIdx.x] += ry[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) V[((size_t)pair * N_e + e) * N_rhs + r] = make_double2(rx[0], ry[0]);
}

void launch_fwd_general(const Plan& P, const float2* gm, int N_rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_fwd_general, dim3(P.N_pairs, P.N_e, N_rhs), dim3(256), 0, s, P.d_h, P.d_p, gm,
                     P.d_shift, P.d_pair_fg, P.N_pts, P.H_plane, P.N_e, N_rhs, P.d_V);
}

// ------------------------------------------------------------- mixing
// u[r][obs] = sum_a pref[mk][f][a] * V[pair(m, fmeas)][act[a]][r]
template <typename OutT>
__global__ void k_fwd_mix(const double2* __restrict__ V, const int* __restrict__ obs_mk,
                          const int* __restrict__ obs_f, const int* __restrict__ mk_meta,
                          const int64_t* __restrict__ mk_off, const int* __restrict__ fmeas,
                          const int* __restrict__ act, const double2* __restrict__ pref,
                          const int* __restrict__ meas_pair_base, int64_t N_obs, int N_e, int N_rhs,
                          OutT* __restrict__ u) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= N_obs * N_rhs) return;
  const int r = (int)(idx / N_obs);
  const int64_t obs = idx - (int64_t)r * N_obs;
  const int mk = obs_mk[obs], f = obs_f[obs];
  const int m = mk_meta[mk * 4 + 0], Nf = mk_meta[mk * 4 + 1], Nact = mk_meta[mk * 4 + 2];
  const int64_t f_off = mk_off[mk * 4 + 1], a_off = mk_off[mk * 4 + 2], p_off = mk_off[mk * 4 + 3];
  const int pair = meas_pair_base[m] + fmeas[f_off + f];
  double sx = 0.0, sy = 0.0;
  for (int a = 0; a < Nact; ++a) {
    const double2 c = pref[p_off + (int64_t)a * Nf + f];
    const double2 v = V[((size_t)pair * N_e + act[a_off + a]) * N_rhs + r];
    sx += c.x * v.x - c.y * v.y;
    sy += c.x * v.y + c.y * v.x;
  }
  OutT o;
  o.x = sx;
  o.y = sy;
  u[idx] = o;
}

void launch_fwd_mix(const Plan& P, int N_rhs, void* u, bool out_f64, hipStream_t s) {
  const int64_t total = P.N_obs * N_rhs;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  if (out_f64)
    hipLaunchKernelGGL(k_fwd_mix<double2>, grid, block, 0, s, P.d_V, P.d_obs_mk, P.d_obs_f, P.d_mk_meta,
                       P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, P.d_meas_pair_base, P.N_obs, P.N_e, N_rhs,
                       (double2*)u);
  else
    hipLaunchKernelGGL(k_fwd_mix<float2>, grid, block, 0, s, P.d_V, P.d_obs_mk, P.d_obs_f, P.d_mk_meta,
                       P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, P.d_meas_pair_base, P.N_obs, P.N_e, N_rhs,
                       (float2*)u);
}

// ------------------------------------------------------------- columns
// Columns of A: u[j][obs] = p_inc[pair][p_j] * sum_a pref[mk][f][a] *
// h_ref[fg(pair)][shift(p_j, e_a)], the forward of the unit vector at p_j
// (forward_quick_cpu with one occupied point, operator_born.m:428-452).
// One thread per (column, observation); FP64 sums.
template <typename OutT>
__global__ void k_columns(const float2* __restrict__ h, const float2* __restrict__ Pt, const int* __restrict__ shift,
                          const int* __restrict__ pair_fg, const int* __restrict__ obs_mk,
                          const int* __restrict__ obs_f, const int* __restrict__ mk_meta,
                          const int64_t* __restrict__ mk_off, const int* __restrict__ fmeas,
                          const int* __restrict__ act, const double2* __restrict__ pref,
                          const int* __restrict__ meas_pair_base, int64_t N_obs, int N_pts, int H_plane,
                          const int* __restrict__ pts, int n, OutT* __restrict__ u) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= N_obs * n) return;
  const int j = (int)(idx / N_obs);
  const int64_t obs = idx - (int64_t)j * N_obs;
  const int p = pts[j];
  const int mk = obs_mk[obs], f = obs_f[obs];
  const int m = mk_meta[mk * 4 + 0], Nf = mk_meta[mk * 4 + 1], Nact = mk_meta[mk * 4 + 2];
  const int64_t f_off = mk_off[mk * 4 + 1], a_off = mk_off[mk * 4 + 2], p_off = mk_off[mk * 4 + 3];
  const int pair = meas_pair_base[m] + fmeas[f_off + f];
  const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
  double qx = 0.0, qy = 0.0;
  for (int a = 0; a < Nact; ++a) {
    const double2 c = pref[p_off + (int64_t)a * Nf + f];
    const float2 hv = hr[shift[(size_t)act[a_off + a] * N_pts + p]];
    qx += c.x * hv.x - c.y * hv.y;
    qy += c.x * hv.y + c.y * hv.x;
  }
  const float2 pv = Pt[(size_t)pair * N_pts + p];
  OutT o;
  o.x = qx * pv.x - qy * pv.y;
  o.y = qx * pv.y + qy * pv.x;
  u[idx] = o;
}

void launch_columns(const Plan& P, const int* pts, int n, void* u, bool out_f64, hipStream_t s) {
  const int64_t total = P.N_obs * n;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  if (out_f64)
    hipLaunchKernelGGL(k_columns<double2>, grid, block, 0, s, P.d_h, P.d_p, P.d_shift, P.d_pair_fg, P.d_obs_mk,
                       P.d_obs_f, P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, P.d_meas_pair_base, P.N_obs,
                       P.N_pts, P.H_plane, pts, n, (double2*)u);
  else
    hipLaunchKernelGGL(k_columns<float2>, grid, block, 0, s, P.d_h, P.d_p, P.d_shift, P.d_pair_fg, P.d_obs_mk,
                       P.d_obs_f, P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, P.d_meas_pair_base, P.N_obs,
                       P.N_pts, P.H_plane, pts, n, (float2*)u);
}

// ------------------------------------------------------------- sparse forward
// u[r][obs] = sum_j A[obs, p_j] g[r][j] over the n occupied points p_j: the
// forward with the MEX's occupancy compaction (cu:454-499) as a gather, for
// inputs with few occupied points.  One thread per (rhs, observation).
__global__ void k_fwd_sparse(const float2* __restrict__ h, const float2* __restrict__ Pt,
                             const int* __restrict__ shift, const int* __restrict__ pair_fg,
                             const int* __restrict__ obs_mk, const int* __restrict__ obs_f,
                             const int* __restrict__ mk_meta, const int64_t* __restrict__ mk_off,
                             const int* __restrict__ fmeas, const int* __restrict__ act,
                             const double2* __restrict__ pref, const int* __restrict__ meas_pair_base,
                             int64_t N_obs, int N_pts, int H_plane, const int* __restrict__ pts, int n,
                             const double2* __restrict__ g, int N_rhs, double2* __restrict__ u) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= N_obs * N_rhs) return;
  const int r = (int)(idx / N_obs);
  const int64_t obs = idx - (int64_t)r * N_obs;
  const int mk = obs_mk[obs], f = obs_f[obs];
  const int m = mk_meta[mk * 4 + 0], Nf = mk_meta[mk * 4 + 1], Nact = mk_meta[mk * 4 + 2];
  const int64_t f_off = mk_off[mk * 4 + 1], a_off = mk_off[mk * 4 + 2], p_off = mk_off[mk * 4 + 3];
  const int pair = meas_pair_base[m] + fmeas[f_off + f];
  const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
  const float2* pr = Pt + (size_t)pair * N_pts;
  const double2* gr = g + (size_t)r * n;
  double sx = 0.0, sy = 0.0;
  for (int j = 0; j < n; ++j) {
    const int p = pts[j];
    double qx = 0.0, qy = 0.0;
    for (int a = 0; a < Nact; ++a) {
      const double2 c = pref[p_off + (int64_t)a * Nf + f];
      const float2 hv = hr[shift[(size_t)act[a_off + a] * N_pts + p]];
      qx += c.x * hv.x - c.y * hv.y;
      qy += c.x * hv.y + c.y * hv.x;
    }
    const float2 pv = pr[p];
    const double ax = qx * pv.x - qy * pv.y, ay = qx * pv.y + qy * pv.x;
    const double2 gv = gr[j];
    sx += ax * gv.x - ay * gv.y;
    sy += ax * gv.y + ay * gv.x;
  }
  u[idx] = make_double2(sx, sy);
}

void launch_fwd_sparse(const Plan& P, const int* pts, int n, const double2* g, int N_rhs, double2* u, hipStream_t s) {
  const int64_t total = P.N_obs * N_rhs;
  hipLaunchKernelGGL(k_fwd_sparse, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P.d_h, P.d_p, P.d_shift,
                     P.d_pair_fg, P.d_obs_mk, P.d_obs_f, P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref,
                     P.d_meas_pair_base, P.N_obs, P.N_pts, P.H_plane, pts, n, g, N_rhs, u);
}

// ------------------------------------------------------------- un-mixing
// W[pair][e][r] = sum over the (mk, f) reading this pair and the active
// slot a of element e: conj(pref[mk][f][a]) * u[r][obs(mk, f)], walked from
// the per-(pair, element) entry lists.  A gather, so the result is
// deterministic.  With Wc set, the result goes to the chunk layout
// Wc[chunk(pair, r)][e][column(pair, r)] of the structured kernels instead.
__global__ void k_adj_unmix(const float2* __restrict__ u, const int64_t* __restrict__ pe_off,
                            const int2* __restrict__ pe_ent, const double2* __restrict__ pref, int N_pairs,
                            int N_e, int N_rhs, int64_t N_obs, float2* __restrict__ W, float2* __restrict__ Wc,
                            const int* __restrict__ pr_chunk, const int* __restrict__ pr_pos, int Ne_pad, int Nex,
                            int Ney) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)N_pairs * N_e * N_rhs) return;
  const int r = (int)(idx % N_rhs);
  const size_t pe = idx / N_rhs;
  const int e = (int)(pe % N_e);
  const int pair = (int)(pe / N_e);
  double sx = 0.0, sy = 0.0;
  const float2* ur = u + (int64_t)r * N_obs;
  for (int64_t j = pe_off[pe]; j < pe_off[pe + 1]; ++j) {
    const int2 oe = pe_ent[j];
    const double2 c = pref[oe.y];
    const float2 v = ur[oe.x];
    sx += c.x * v.x + c.y * v.y;
    sy += c.x * v.y - c.y * v.x;
  }
  const float2 res = make_float2((float)sx, (float)sy);
  if (Wc) {
    const int ch = pr_chunk[(size_t)pair * N_rhs + r], pos = pr_pos[(size_t)pair * N_rhs + r];
    const int ex = e % Nex, ey = e / Nex;
    Wc[(((size_t)ch * Ney + ey) * Ne_pad + ex) * NC + pos] = res;
  } else {
    W[idx] = res;
  }
}

void launch_adj_unmix(const Plan& P, const float2* u, int N_rhs, bool chunked, hipStream_t s) {
  const size_t total = (size_t)P.N_pairs * P.N_e * N_rhs;
  hipLaunchKernelGGL(k_adj_unmix, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, u, P.d_pe_off,
                     P.d_pe_ent, P.d_pref, P.N_pairs, P.N_e, N_rhs, P.N_obs, P.d_W, chunked ? P.d_Wc : nullptr,
                     P.d_pr_chunk, P.d_pr_pos, P.Ne_pad, P.Nex, P.nd == 3 ? P.Ney : 1);
}

// ------------------------------------------------------ received energy
// E[p] = sum over the mixes k of measurement m and their frequencies f of
// |p_inc[f, p]|^2 * |sum_a pref[k][f, a] * h_ref[f, shift(p, e_a)]|^2, the
// squared norm of column p of the measurement's rows of A
// (energy_rx_canonical, operator_born.m:718-859, identity TGC).  One thread
// per point, FP64 sums.
__global__ __launch_bounds__(256) void k_energy_rx(
    const float2* __restrict__ h, const float2* __restrict__ Pt, const int* __restrict__ shift,
    const int* __restrict__ pair_fg, const int* __restrict__ mk_meta, const int64_t* __restrict__ mk_off,
    const int* __restrict__ fmeas, const int* __restrict__ act, const double2* __restrict__ pref, int mk0, int mk1,
    int pair_base, int N_pts, int H_plane, double* __restrict__ E) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= N_pts) return;
  double acc = 0.0;
  for (int mk = mk0; mk < mk1; ++mk) {
    const int Nf = mk_meta[mk * 4 + 1], Na = mk_meta[mk * 4 + 2];
    const int64_t f_off = mk_off[mk * 4 + 1], a_off = mk_off[mk * 4 + 2], pref_off = mk_off[mk * 4 + 3];
    for (int f = 0; f < Nf; ++f) {
      const int pair = pair_base + fmeas[f_off + f];
      const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
      double qx = 0.0, qy = 0.0;
      for (int a = 0; a < Na; ++a) {
        const float2 hv = hr[shift[(size_t)act[a_off + a] * N_pts + p]];
        const double2 c = pref[pref_off + (int64_t)a * Nf + f];
        qx += c.x * hv.x - c.y * hv.y;
        qy += c.x * hv.y + c.y * hv.x;
      }
      const float2 pv = Pt[(size_t)pair * N_pts + p];
      acc += ((double)pv.x * pv.x + (double)pv.y * pv.y) * (qx * qx + qy * qy);
    }
  }
  E[p] = acc;
}

void launch_energy_rx(const Plan& P, int m, double* E, hipStream_t s) {
  int mk0 = 0;
  while (mk0 < P.N_mk && P.mk_meas[mk0] != m) ++mk0;
  int mk1 = mk0;
  while (mk1 < P.N_mk && P.mk_meas[mk1] == m) ++mk1;
  hipLaunchKernelGGL(k_energy_rx, dim3((P.N_pts + 255) / 256), dim3(256), 0, s, P.d_h, P.d_p, P.d_shift, P.d_pair_fg,
                     P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, mk0, mk1, P.meas_pair_base[m], P.N_pts,
                     P.H_plane, E);
}

// ------------------------------------------------------ time-domain signals
// rf[r][mk][n] = sum over the frequencies f of mix mk of
//   w_q * Re(u[r][obs(mk, f)] * exp(+j 2 pi q n / N_t)),
// q = f / delta_f the integer frequency index, w_q = 1 for q = 0 and for the
// Nyquist index N_t / 2, else 2: N_t * ifft(X, N_t, 'symmetric') of the
// zero-padded spectrum (signal_matrix.m:379-391, lbs_q = 0).  The phase is
// reduced exactly in integers, (q n) mod N_t, before the FP32 sincos.
__global__ __launch_bounds__(256) void k_time_signals(const float2* __restrict__ u, const int* __restrict__ q_obs,
                                                      const int64_t* __restrict__ mk_off,
                                                      const int* __restrict__ mk_meta, int N_mk, int64_t N_obs,
                                                      int N_rhs, int N_t, float* __restrict__ rf) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)N_rhs * N_mk * N_t) return;
  const int n = (int)(idx % N_t);
  const size_t rs = idx / N_t;
  const int mk = (int)(rs % N_mk);
  const int r = (int)(rs / N_mk);
  const int64_t o0 = mk_off[mk * 4 + 0];
  const int Nf = mk_meta[mk * 4 + 1];
  const float2* ur = u + (size_t)r * N_obs;
  const float w0 = 6.283185307179586f / (float)N_t;
  float acc = 0.f;
  for (int f = 0; f < Nf; ++f) {
    const int q = q_obs[o0 + f];
    const float2 x = ur[o0 + f];
    const int ph = (int)(((int64_t)q * n) % N_t);
    float sn, cs;
    __sincosf(w0 * (float)ph, &sn, &cs);
    const float wq = (q == 0 || 2 * q == N_t) ? 1.f : 2.f;
    acc += wq * (x.x * cs - x.y * sn);
  }
  rf[idx] = acc;
}

void launch_time_signals(const Plan& P, const float2* u, const int* q_obs, int N_rhs, int N_t, float* rf,
                         hipStream_t s) {
  const size_t total = (size_t)N_rhs * P.N_mk * N_t;
  hipLaunchKernelGGL(k_time_signals, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, u, q_obs, P.d_mk_off,
                     P.d_mk_meta, P.N_mk, P.N_obs, N_rhs, N_t, rf);
}

// ------------------------------------------------------ time gain compensation
// One discrete convolution along the frequency coefficients of every mix,
// cut off to the mix's own length N (linear_transforms.convolutions.fft,
// fft.m:85-126 with cut_off; convolution.m:62-98).  Taps k[d], |d| <= M:
//   forward  y[i] = sum_d k[d] x[i - d]         (0 <= i - d < N)
//   adjoint  y[j] = sum_d conj(k[d]) x[j + d]   (0 <= j + d < N)
// A block computes 256 outputs of one (mix, right-hand side); the taps pass
// through LDS in chunks of 256 together with the 511 inputs a chunk reaches.
__device__ __forceinline__ void cfma(float2& a, float2 k, float2 x) {
  a.x = fmaf(k.x, x.x, fmaf(-k.y, x.y, a.x));
  a.y = fmaf(k.x, x.y, fmaf(k.y, x.x, a.y));
}
__device__ __forceinline__ void cfma(double2& a, double2 k, double2 x) {
  a.x = fma(k.x, x.x, fma(-k.y, x.y, a.x));
  a.y = fma(k.x, x.y, fma(k.y, x.x, a.y));
}

template <typename T, bool ADJ>
__global__ __launch_bounds__(256) void k_tgc(const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ taps,
                                             const int2* __restrict__ tiles, const int* __restrict__ half,
                                             const int64_t* __restrict__ tap_off, const int64_t* __restrict__ mk_off,
                                             const int* __restrict__ mk_meta, int64_t N_obs) {
  __shared__ T xs[511];
  __shared__ T ks[256];
  const int2 tl = tiles[blockIdx.x];
  const int mk = tl.x, i0 = tl.y;
  const int M = half[mk];
  const int N = mk_meta[mk * 4 + 1];
  const int64_t o0 = mk_off[mk * 4 + 0];
  const T* kt = taps + tap_off[mk] + M;   // kt[d], -M <= d <= M
  const T* xr = x + (size_t)blockIdx.y * N_obs + o0;
  const int li = threadIdx.x;
  T acc;
  acc.x = 0;
  acc.y = 0;
  for (int d0 = -M; d0 <= M; d0 += 256) {
    T k;
    k.x = 0;
    k.y = 0;
    if (d0 + li <= M) {
      k = kt[d0 + li];
      if (ADJ) k.y = -k.y;
    }
    ks[li] = k;
    for (int t = li; t < 511; t += 256) {
      const int j = ADJ ? i0 + d0 + t : i0 - d0 - 255 + t;
      T v;
      v.x = 0;
      v.y = 0;
      if (j >= 0 && j < N) v = xr[j];
      xs[t] = v;
    }
    __syncthreads();
    const int nd = min(256, M - d0 + 1);
    for (int dd = 0; dd < nd; ++dd) cfma(acc, ks[dd], ADJ ? xs[li + dd] : xs[li - dd + 255]);
    __syncthreads();
  }
  if (i0 + li < N) y[(size_t)blockIdx.y * N_obs + o0 + i0 + li] = acc;
}

void launch_tgc(const Plan& P, bool adjoint, bool f64, const void* in, void* out, int N_rhs, hipStream_t s) {
  const dim3 grid((unsigned)P.n_tgc_tiles, (unsigned)N_rhs);
  if (f64) {
    if (adjoint)
      hipLaunchKernelGGL((k_tgc<double2, true>), grid, dim3(256), 0, s, (const double2*)in, (double2*)out,
                         P.d_tgc_k64, P.d_tgc_tiles, P.d_tgc_half, P.d_tgc_off, P.d_mk_off, P.d_mk_meta, P.N_obs);
    else
      hipLaunchKernelGGL((k_tgc<double2, false>), grid, dim3(256), 0, s, (const double2*)in, (double2*)out,
                         P.d_tgc_k64, P.d_tgc_tiles, P.d_tgc_half, P.d_tgc_off, P.d_mk_off, P.d_mk_meta, P.N_obs);
  } else {
    if (adjoint)
      hipLaunchKernelGGL((k_tgc<float2, true>), grid, dim3(256), 0, s, (const float2*)in, (float2*)out,
                         P.d_tgc_k32, P.d_tgc_tiles, P.d_tgc_half, P.d_tgc_off, P.d_mk_off, P.d_mk_meta, P.N_obs);
    else
      hipLaunchKernelGGL((k_tgc<float2, false>), grid, dim3(256), 0, s, (const float2*)in, (float2*)out,
                         P.d_tgc_k32, P.d_tgc_tiles, P.d_tgc_half, P.d_tgc_off, P.d_mk_off, P.d_mk_meta, P.N_obs);
  }
}

// ------------------------------------------------------ received energy, TGC
// E[p] = sum over the mixes k of measurement m of sum_i |sum_j C_k[i, j] Phi_k[j, p]|^2
// with Phi_k[j, p] = p_inc[f_j, p] sum_a pref[k][j, a] h_ref[f_j, shift(p, e_a)]
// and C_k the mix's TGC convolution (C[i, j] = k[i - j], |i - j| <= M):
// energy_rx_canonical with a diagonal LT_tgc (operator_born.m:830-848).
// A block holds 64 points (the lanes) and loops over the mixes; its 4 waves
// first fill Phi[j][lane] for all rows j of the mix (FP32, in LDS, or in
// the block's slice of `spill` when the mix is too long for LDS), then each
// wave takes every 4th row i of the band product.  FP64 sums of |y|^2.
template <bool SPILL>
__global__ __launch_bounds__(256) void k_energy_rx_tgc(
    const float2* __restrict__ h, const float2* __restrict__ Pt, const int* __restrict__ shift,
    const int* __restrict__ pair_fg, const int* __restrict__ mk_meta, const int64_t* __restrict__ mk_off,
    const int* __restrict__ fmeas, const int* __restrict__ act, const double2* __restrict__ pref,
    const float2* __restrict__ taps, const int* __restrict__ half, const int64_t* __restrict__ tap_off, int mk0,
    int mk1, int pair_base, int N_pts, int H_plane, int n_tiles, float2* __restrict__ spill, int spill_rows,
    double* __restrict__ E) {
  extern __shared__ float2 phi_lds[];
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float2* phi = SPILL ? spill + (size_t)blockIdx.x * spill_rows * 64 : phi_lds;
  for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const int p = tile * 64 + lane;
    const bool valid = p < N_pts;
    double acc = 0.0;
    for (int mk = mk0; mk < mk1; ++mk) {
      const int Nf = mk_meta[mk * 4 + 1], Na = mk_meta[mk * 4 + 2];
      const int64_t f_off = mk_off[mk * 4 + 1], a_off = mk_off[mk * 4 + 2], pref_off = mk_off[mk * 4 + 3];
      for (int j = w; j < Nf; j += 4) {
        float2 v = make_float2(0.f, 0.f);
        if (valid) {
          const int pair = pair_base + fmeas[f_off + j];
          const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
          double qx = 0.0, qy = 0.0;
          for (int a = 0; a < Na; ++a) {
            const float2 hv = hr[shift[(size_t)act[a_off + a] * N_pts + p]];
            const double2 c = pref[pref_off + (int64_t)a * Nf + j];
            qx += c.x * hv.x - c.y * hv.y;
            qy += c.x * hv.y + c.y * hv.x;
          }
          const float2 pv = Pt[(size_t)pair * N_pts + p];
          v = make_float2((float)(qx * pv.x - qy * pv.y), (float)(qx * pv.y + qy * pv.x));
        }
        phi[j * 64 + lane] = v;
      }
      __syncthreads();
      const int M = half[mk];
      const float2* kt = taps + tap_off[mk] + M;   // kt[d], -M <= d <= M
      for (int i = w; i < Nf; i += 4) {
        const int j0 = max(0, i - M), j1 = min(Nf - 1, i + M);
        float2 y = make_float2(0.f, 0.f);
        for (int j = j0; j <= j1; ++j) cfma(y, kt[i - j], phi[j * 64 + lane]);
        acc += (double)y.x * y.x + (double)y.y * y.y;
      }
      __syncthreads();
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && valid) E[p] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    __syncthreads();
  }
}

// rows: the longest mix of the measurement
void launch_energy_rx_tgc(const Plan& P, int m, int rows, float2* spill, int spill_blocks, double* E,
                          hipStream_t s) {
  int mk0 = 0;
  while (mk0 < P.N_mk && P.mk_meas[mk0] != m) ++mk0;
  int mk1 = mk0;
  while (mk1 < P.N_mk && P.mk_meas[mk1] == m) ++mk1;
  const int n_tiles = (P.N_pts + 63) / 64;
  if (!spill) {
    const size_t lds = (size_t)std::max(rows, 1) * 64 * sizeof(float2);
    hipFuncSetAttribute((const void*)k_energy_rx_tgc<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        ETGC_LDS_ROWS * 64 * (int)sizeof(float2));
    hipLaunchKernelGGL(k_energy_rx_tgc<false>, dim3(n_tiles), dim3(256), lds, s, P.d_h, P.d_p, P.d_shift,
                       P.d_pair_fg, P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref, P.d_tgc_k32,
                       P.d_tgc_half, P.d_tgc_off, mk0, mk1, P.meas_pair_base[m], P.N_pts, P.H_plane, n_tiles,
                       (float2*)nullptr, 0, E);
  } else {
    hipLaunchKernelGGL(k_energy_rx_tgc<true>, dim3(std::min(n_tiles, spill_blocks)), dim3(256), 0, s, P.d_h,
                       P.d_p, P.d_shift, P.d_pair_fg, P.d_mk_meta, P.d_mk_off, P.d_fmeas, P.d_act, P.d_pref,
                       P.d_tgc_k32, P.d_tgc_half, P.d_tgc_off, mk0, mk1, P.meas_pair_base[m], P.N_pts, P.H_plane,
                       n_tiles, spill, rows, E);
  }
}

// ------------------------------------------------------ FP16 split scales
// FP16 has a 5-bit exponent, so the split form scales every operand by a
// power of 2 (exact) that brings its largest entry just below 2^14:
// s = 2^(14 - e) with max < 2^e.  The contraction results are scaled back by
// the inverse product.  Largest |re| or |im| of n floats, per block:
__global__ __launch_bounds__(256) void k_absmax_part(const float* __restrict__ x, size_t n, float* __restrict__ part) {
  __shared__ float red[256];
  float m = 0.f;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) m = fmaxf(m, fabsf(x[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__device__ __forceinline__ int max_exponent(float m) {   // m < 2^e (e = 0 for m = 0)
  int e = 0;
  if (m > 0.f) frexpf(m, &e);
  return e;
}

// which = 0: maxima of h_ref and p_inc -> maxima[0..1]; which = 1: forward
// scales from the maxima and max |gamma|; which = 2: adjoint scales from
// max |h_ref| and max |W|.  One block.
__global__ __launch_bounds__(256) void k_fp16_scales(const float* __restrict__ part, int n_part, int which,
                                                     float* __restrict__ maxima, float* __restrict__ fscale) {
  __shared__ float red[256];
  float m = 0.f;
  for (int i = threadIdx.x; i < n_part; i += 256) m = fmaxf(m, part[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const float x = red[0];
  if (which < 0) {
    maxima[-which - 1] = x;
    return;
  }
  const int eh = max_exponent(maxima[0]), ex = max_exponent(x);
  fscale[0] = ldexpf(1.f, 14 - eh);
  if (which == 1) {
    const int ep = max_exponent(maxima[1]);
    fscale[1] = ldexpf(1.f, 14 - ep - ex);
    fscale[2] = ldexpf(1.f, eh + ep + ex - 28);
  } else {
    fscale[3] = ldexpf(1.f, 14 - ex);
    fscale[4] = ldexpf(1.f, eh + ex - 28);
  }
}

// which: -1 / -2 store max |x| as the h_ref / p_inc maximum; 1 forward
// scales (x = gamma); 2 adjoint scales (x = W)
void launch_fp16_scales(const Plan& P, const float* x, size_t n_floats, int which, hipStream_t s) {
  const int nb = (int)std::min<size_t>(256, std::max<size_t>(1, (n_floats + 255) / 256));
  hipLaunchKernelGGL(k_absmax_part, dim3(nb), dim3(256), 0, s, x, n_floats, P.d_f16part);
  hipLaunchKernelGGL(k_fp16_scales, dim3(1), dim3(256), 0, s, P.d_f16part, nb, which, P.d_f16max, P.d_fscale);
}

// ------------------------------------------------------ general adjoint
__global__ __launch_bounds__(256) void k_adj_general(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ W,
    const int* __restrict__ shift, const int* __restrict__ pair_fg, int N_pts, int H_plane, int N_e, int N_pairs,
    int N_rhs, float2* __restrict__ g_out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (p >= N_pts) return;
  float gxs = 0.f, gys = 0.f;
  for (int pair = 0; pair < N_pairs; ++pair) {
    const float2* hr = h + (size_t)pair_fg[pair] * H_plane;
    float yx = 0.f, yy = 0.f;
    for (int e = 0; e < N_e; ++e) {
      const float2 a = hr[shift[(size_t)e * N_pts + p]];
      const float2 wv = W[((size_t)pair * N_e + e) * N_rhs + r];
      yx += a.x * wv.x + a.y * wv.y;
      yy += a.x * wv.y - a.y * wv.x;
    }
    const float2 pv = P[(size_t)pair * N_pts + p];
    gxs += pv.x * yx + pv.y * yy;
    gys += pv.x * yy - pv.y * yx;
  }
  g_out[(size_t)r * N_pts + p] = make_float2(gxs, gys);
}

void launch_adj_general(const Plan& P, int N_rhs, float2* g_out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_general, dim3((P.N_pts + 255) / 256, N_rhs), dim3(256), 0, s, P.d_h, P.d_p, P.d_W,
                     P.d_shift, P.d_pair_fg, P.N_pts, P.H_plane, P.N_e, P.N_pairs, N_rhs, g_out);
}

// --------------------------------------------------- table build (FP64)
struct cd {
  double re, im;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cd cinv(cd a) {
  const double d = a.re * a.re + a.im * a.im;
  return {a.re / d, -a.im / d};
}
__device__ __forceinline__ cd csqrt_(cd a) {
  const double r = cabs_(a);
  double x = sqrt(0.5 * (r + fabs(a.re)));
  if (x == 0.0) return {0.0, 0.0};
  const double y = a.im / (2.0 * x);
  if (a.re >= 0.0) return {x, y};
  return {fabs(y), a.im >= 0.0 ? x : -x};
}
__device__ __forceinline__ cd clog_(cd a) { return {log(cabs_(a)), atan2(a.im, a.re)}; }

// H0^(2)(z): Hankel asymptotic expansion (DLMF 10.17.4, 10.17.6) truncated at
// its smallest term for |z| >= 12; ascending series of J0 and Y0 (DLMF
// 10.2.2, 10.8.2) below.  The phase exp(-i(z - pi/4)) is reduced in FP64.
__device__ cd hankel2_0(cd z) {
  const double az = cabs_(z);
  if (az >= 12.0) {
    const cd iz = cinv(z);
    cd s = {1.0, 0.0}, term = {1.0, 0.0};
    double prev = 1e300;
    for (int k = 1; k < 60; ++k) {
      const double c = -(double)((2 * k - 1) * (2 * k - 1)) / (8.0 * k);
      const cd t = cmul(term, iz);
      term = {c * t.im, -c * t.re};  // c * (-i) * t
      const double at = cabs_(term);
      if (at >= prev) break;
      s = cadd(s, term);
      prev = at;
      if (at < 1e-17 * cabs_(s)) break;
    }
    const cd pre = csqrt_(cscale(iz, 2.0 / M_PI));
    const double ang = M_PI / 4.0 - z.re;
    double sn, cs;
    sincos(ang, &sn, &cs);
    const double ex = exp(z.im);
    return cmul(cmul(pre, cd{ex * cs, ex * sn}), s);
  }
  const cd q = cscale(cmul(z, z), 0.25);
  cd term = {1.0, 0.0}, j0 = {1.0, 0.0}, ys = {0.0, 0.0};
  double hk = 0.0;
  for (int k = 1; k < 80; ++k) {
    term = cscale(cmul(term, q), -1.0 / ((double)k * k));
    hk += 1.0 / k;
    j0 = cadd(j0, term);
    ys = cadd(ys, cscale(term, -hk));
    if (cabs_(term) < 1e-18 * fmax(1.0, cabs_(j0))) break;
  }
  const cd lg = clog_(cscale(z, 0.5));
  const cd y0 = cscale(cadd(cmul(cd{lg.re + 0.57721566490153286061, lg.im}, j0), ys), 2.0 / M_PI);
  return {j0.re + y0.im, j0.im - y0.re};  // J0 - i Y0
}

__device__ __forceinline__ double taper(int kind, double param, double flag) {
  if (kind == 1) {
    const double lb = M_PI * (1.0 - param), ub = M_PI;
    if (flag <= lb) return 1.0;
    if (flag >= ub) return 0.0;
    return 0.5 * (1.0 + cos(M_PI * (flag - lb) / (ub - lb)));
  }
  if (kind == 2) return 1.0 / (1.0 + exp(param * (flag - M_PI)));
  if (kind == 3) return flag < M_PI ? 1.0 : 0.0;
  return 1.0;
}

// h[fg][p] = -2 dA sum_face w G(k_fg |r_p - r_face|), times the taper of the
// reference element (setup.m:1216-1281, :916-1032; greens_function.m:88, :103)
__global__ __launch_bounds__(256) void k_build_h(GeomDev g, int N_fu, const double2* __restrict__ kk,
                                                 const double2* __restrict__ face_w, double2* __restrict__ h64) {
  const int N_tab = g.Nx_tab * g.Ny_tab * g.Nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int fg = blockIdx.y;
  if (p >= N_tab) return;
  const int ix = p % g.Nx_tab - g.ox;   // FOV subscript (negative before the FOV)
  const int rest = p / g.Nx_tab;
  const int iy = (g.nd == 3) ? rest % g.Ny_tab - g.oy : 0;
  const int iz = (g.nd == 3) ? rest / g.Ny_tab : rest;
  const double x = (ix - 0.5 * (g.Nx - 1)) * g.dx;
  const double y = (g.nd == 3) ? (iy - 0.5 * (g.Ny - 1)) * g.dy : 0.0;
  const double z = g.z0 + iz * g.dz;
  const double xc = -0.5 * (g.Nex - 1) * g.pitch_x;
  const double yc = (g.nd == 3) ? -0.5 * (g.Ney - 1) * g.pitch_y : 0.0;
  const cd k = {kk[fg].x, kk[fg].y};
  cd acc = {0.0, 0.0};
  const int Nfy = (g.nd == 3) ? g.Nfy : 1;
  for (int qy = 0; qy < Nfy; ++qy) {
    const double yf = (g.nd == 3) ? yc + (qy - 0.5 * (g.Nfy - 1)) * g.dfy : 0.0;
    for (int qx = 0; qx < g.Nfx; ++qx) {
      const double xf = xc + (qx - 0.5 * (g.Nfx - 1)) * g.dfx;
      const double dxx = x - xf, dyy = y - yf;
      const double D = sqrt(dxx * dxx + dyy * dyy + z * z);
      cd G;
      if (g.nd == 2) {
        const cd H = hankel2_0(cscale(k, D));
        G = {-0.25 * H.im, 0.25 * H.re};
      } else {
        const double ph = -k.re * D;
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double mag = -exp(k.im * D) / (4.0 * M_PI * D);
        G = {mag * cs, mag * sn};
      }
      if (face_w) {
        const double2 wv = face_w[(size_t)fg * g.Nfx * Nfy + (size_t)(qy * g.Nfx + qx)];  // N_face x N_f, col-major
        G = cmul(G, cd{wv.x, wv.y});
      }
      acc = cadd(acc, G);
    }
  }
  const double dA = (g.nd == 3) ? g.dfx * g.dfy : g.dfx;
  acc = cscale(acc, -2.0 * dA);
  if (g.filt_kind != 0) {
    const double rx = x - xc, ry = y - yc;
    const double R = sqrt(rx * rx + ry * ry + z * z);
    double t = 1.0;
    if (g.Nex > 1) t *= taper(g.filt_kind, g.filt_param, k.re * g.pitch_x * fabs(rx) / R);
    if (g.nd == 3 && g.Ney > 1) t *= taper(g.filt_kind, g.filt_param, k.re * g.pitch_y * fabs(ry) / R);
    if (g.filt_kind == 3 && g.nd == 3 && g.Nex > 1 && g.Ney > 1) {
      // boxcar: all dimensions must pass (boxcar.m:84) -- the product above already is 0 or 1
    }
    acc = cscale(acc, t);
  }
  h64[(size_t)fg * N_tab + p] = make_double2(acc.re, acc.im);
}

void launch_build_h(const GeomDev& g, int N_fu, const double2* k, const double2* face_w, double2* h64,
                    hipStream_t s) {
  const int N_tab = g.Nx_tab * g.Ny_tab * g.Nz;
  hipLaunchKernelGGL(k_build_h, dim3((N_tab + 255) / 256, N_fu), dim3(256), 0, s, g, N_fu, k, face_w, h64);
}

// p_in[pair][p] = sum_a h_tx[fg][shift(p, e_a)] v_d[a] in FP64, stored FP32
// (sequence.m:825-840).  v_d of measurement m: complex N_fm x N_tx, col-major.
__global__ __launch_bounds__(256) void k_synth_pinc(const double2* __restrict__ h64, const int* __restrict__ shift,
                                                    const int* __restrict__ pair_fg,
                                                    const int* __restrict__ pair_meas,
                                                    const int* __restrict__ meas_pair_base,
                                                    const int* __restrict__ meas_Nfm, const int* __restrict__ tx_off,
                                                    const int* __restrict__ tx_el, const double2* __restrict__ vd,
                                                    const int* __restrict__ vd_off, int N_pts, int H_plane,
                                                    float2* __restrict__ P) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int pair = blockIdx.y;
  if (p >= N_pts) return;
  const int m = pair_meas[pair];
  const int fm = pair - meas_pair_base[m];
  const int Nfm = meas_Nfm[m];
  const double2* hr = h64 + (size_t)pair_fg[pair] * H_plane;
  double sx = 0.0, sy = 0.0;
  for (int a = tx_off[m]; a < tx_off[m + 1]; ++a) {
    const double2 hv = hr[shift[(size_t)tx_el[a] * N_pts + p]];
    const double2 v = vd[vd_off[m] + (size_t)(a - tx_off[m]) * Nfm + fm];
    sx += hv.x * v.x - hv.y * v.y;
    sy += hv.x * v.y + hv.y * v.x;
  }
  P[(size_t)pair * N_pts + p] = make_float2((float)sx, (float)sy);
}

}  // namespace born

namespace born {
void launch_synth_pinc(const Plan& P, const double2* h64, const int* d_meas_Nfm, const int* d_tx_off,
                       const int* d_tx_el, const double2* d_vd, const int* d_vd_off, hipStream_t s) {
  hipLaunchKernelGGL(k_synth_pinc, dim3((P.N_pts + 255) / 256, P.N_pairs), dim3(256), 0, s, h64, P.d_shift,
                     P.d_pair_fg, P.d_pair_meas, P.d_meas_pair_base, d_meas_Nfm, d_tx_off, d_tx_el, d_vd, d_vd_off,
                     P.N_pts, P.H_plane, P.d_p);
}


// p_in as above, with the shift map computed from the separable lateral maps
// (structured geometries): element e = ex + Nex*ey sees point (ix, iy, iz) at
// extx[ix - ex*Np + offx] + HR*(exty[iy - ey*Npy + offy] + HY*iz).
__global__ __launch_bounds__(256) void k_synth_pinc_s(const double2* __restrict__ h64, const int* __restrict__ extx,
                                                      const int* __restrict__ exty, int Nx, int Ny, int Nex,
                                                      int Np, int Npy, int offx, int offy, int HR, int HY,
                                                      const int* __restrict__ pair_fg,
                                                      const int* __restrict__ pair_meas,
                                                      const int* __restrict__ meas_pair_base,
                                                      const int* __restrict__ meas_Nfm, const int* __restrict__ tx_off,
                                                      const int* __restrict__ tx_el, const double2* __restrict__ vd,
                                                      const int* __restrict__ vd_off, int N_pts, int H_plane,
                                                      float2* __restrict__ P) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int pair = blockIdx.y;
  if (p >= N_pts) return;
  const int ix = p % Nx, rest = p / Nx;
  const int iy = rest % Ny, iz = rest / Ny;
  const int m = pair_meas[pair];
  const int fm = pair - meas_pair_base[m];
  const int Nfm = meas_Nfm[m];
  const double2* hr = h64 + (size_t)pair_fg[pair] * H_plane;
  double sx = 0.0, sy = 0.0;
  for (int a = tx_off[m]; a < tx_off[m + 1]; ++a) {
    const int e = tx_el[a];
    const int ex = e % Nex, ey = e / Nex;
    const int ty = exty ? exty[iy - ey * Npy + offy] : 0;
    const double2 hv = hr[extx[ix - ex * Np + offx] + (size_t)HR * (ty + (size_t)HY * iz)];
    const double2 v = vd[vd_off[m] + (size_t)(a - tx_off[m]) * Nfm + fm];
    sx += hv.x * v.x - hv.y * v.y;
    sy += hv.x * v.y + hv.y * v.x;
  }
  P[(size_t)pair * N_pts + p] = make_float2((float)sx, (float)sy);
}

void launch_synth_pinc_structured(const Plan& P, const double2* h64, const int* d_meas_Nfm, const int* d_tx_off,
                                  const int* d_tx_el, const double2* d_vd, const int* d_vd_off, hipStream_t s) {
  const int Ny = P.nd == 3 ? P.Ny : 1, Ney = P.nd == 3 ? P.Ney : 1;
  hipLaunchKernelGGL(k_synth_pinc_s, dim3((P.N_pts + 255) / 256, P.N_pairs), dim3(256), 0, s, h64, P.d_ext,
                     P.nd == 3 ? P.d_exty : nullptr, P.Nx, Ny, P.Nex, P.Np, P.Npy, (P.Nex - 1) * P.Np,
                     (Ney - 1) * P.Npy, P.HR, P.HY, P.d_pair_fg, P.d_pair_meas, P.d_meas_pair_base, d_meas_Nfm,
                     d_tx_off, d_tx_el, d_vd, d_vd_off, P.N_pts, P.H_plane, P.d_p);
}

}  // namespace born

