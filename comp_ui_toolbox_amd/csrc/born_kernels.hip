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

void launch_mask_gamma(const float2* g, int N_pts, int N_rhs, float2* gm, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_gamma, dim3((N_pts + 255) / 256), dim3(256), 0, s, g, N_pts, N_rhs, gm);
}

// ------------------------------------------------- structured forward (2-D)
// One workgroup = (chunk, z-block).  Four waves split each depth row's
// points; each wave keeps all 16*MT elements x 16 columns in MFMA
// accumulators (A = elements x 4 points read from a polyphase LDS copy of the
// extended transfer-function row, B = 4 points x 16 columns of
// p_inc*gamma).  Row sums are folded into a second accumulator level, the
// four waves are summed through LDS, and the partial for the z-block is
// written out; k_fwd_reduce sums z-blocks in FP64.
template <int MT>
__global__ __launch_bounds__(256) void k_fwd_s2d(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ gm,
    const int* __restrict__ ext, const Chunk* __restrict__ chunks, const int* __restrict__ col_pair,
    int Nx, int Nz, int N_pts, int N_e, int Np, int N_ext, int NZB, int ZR, int S, int Ne_pad,
    float2* __restrict__ vpart) {
  extern __shared__ float2 lds[];
  const int XP = 16 * S;                       // padded row length (4 waves x S steps x 4 points)
  const int QL = (XP - 1) / Np + 16 * MT;      // polyphase row length
  float2* Hp = lds;                            // [Np][QL]
  float2* Bs = lds + Np * QL;                  // [XP][NCS]

  const int ci = blockIdx.x / NZB, zb = blockIdx.x - ci * NZB;
  const Chunk ch = chunks[ci];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int Qoff = 16 * MT - N_e;
  const int z_begin = zb * ZR, z_end = min(Nz, z_begin + ZR);
  const float2* hf = h + (size_t)ch.fg * N_pts;
  const float2* gr = gm + (size_t)ch.r * N_pts;

  v4f accr[MT], acci[MT], sr[MT], si[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    accr[mt] = v4f{0.f, 0.f, 0.f, 0.f};
    acci[mt] = accr[mt];
    sr[mt] = accr[mt];
    si[mt] = accr[mt];
  }

  for (int z = z_begin; z < z_end; ++z) {
    __syncthreads();
    const float2* hrow = hf + (size_t)z * Nx;
    for (int idx = tid; idx < Np * QL; idx += 256) {
      const int ph = idx / QL, q = idx - ph * QL;
      const int t = (q - Qoff) * Np + ph;
      Hp[idx] = (t >= 0 && t < N_ext) ? hrow[ext[t]] : make_float2(0.f, 0.f);
    }
    for (int idx = tid; idx < XP * NC; idx += 256) {
      const int c = idx / XP, i = idx - c * XP;
      float2 v = make_float2(0.f, 0.f);
      if (c < ch.ncols && i < Nx) {
        const int pair = col_pair[ch.col_start + c];
        const size_t p = (size_t)z * Nx + i;
        const float2 a = P[(size_t)pair * N_pts + p];
        const float2 b = gr[p];
        v = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
      }
      Bs[i * NCS + c] = v;
    }
    __syncthreads();

    const int kk = lane >> 4, col = lane & 15;
    for (int s = 0; s < S; ++s) {
      const int i = (w * S + s) * 4 + kk;
      const int ph = i % Np;
      const int qb = ph * QL + i / Np + 16 * MT - 1 - col;
      const float2 b = Bs[i * NCS + col];
      const float nby = -b.y;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const float2 a = Hp[qb - 16 * mt];
        accr[mt] = mfma16(a.x, b.x, accr[mt]);
        accr[mt] = mfma16(a.y, nby, accr[mt]);
        acci[mt] = mfma16(a.x, b.y, acci[mt]);
        acci[mt] = mfma16(a.y, b.x, acci[mt]);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      sr[mt] += accr[mt];
      si[mt] += acci[mt];
      accr[mt] = v4f{0.f, 0.f, 0.f, 0.f};
      acci[mt] = accr[mt];
    }
  }

  // cross-wave reduction through LDS: [w][mt][j][lane]
  __syncthreads();
  float* red = reinterpret_cast<float*>(lds);
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      red[((w * MT + mt) * 8 + v) * 64 + lane] = sr[mt][v];
      red[((w * MT + mt) * 8 + 4 + v) * 64 + lane] = si[mt][v];
    }
  }
  __syncthreads();
  const int per_wave = MT * 8 * 64;
  float* out = reinterpret_cast<float*>(vpart + ((size_t)ci * NZB + zb) * Ne_pad * NC);
  for (int idx = tid; idx < per_wave; idx += 256) {
    const float sum = red[idx] + red[idx + per_wave] + red[idx + 2 * per_wave] + red[idx + 3 * per_wave];
    const int ln = idx & 63, j = (idx >> 6) & 7, mt = idx >> 9;
    const int e = 16 * mt + 4 * (ln >> 4) + (j & 3);
    const int c = ln & 15;
    out[((size_t)e * NC + c) * 2 + (j >> 2)] = sum;
  }
}

static size_t fwd_s2d_lds_bytes(int MT, int S, int Np) {
  const int XP = 16 * S;
  const int QL = (XP - 1) / Np + 16 * MT;
  const size_t body = ((size_t)Np * QL + (size_t)XP * NCS) * sizeof(float2);
  const size_t red = (size_t)4 * MT * 8 * 64 * sizeof(float);
  return std::max(body, red);
}

void launch_fwd_structured(const Plan& P, const float2* gm, int N_rhs, hipStream_t s) {
  const int MT = P.Ne_pad / 16;
  const int S = (P.Nx + 15) / 16;
  const size_t lds = fwd_s2d_lds_bytes(MT, S, P.Np);
  const dim3 grid((unsigned)(P.chunks.size() * P.NZB)), block(256);
#define BORN_FWD_CASE(M)                                                                          \
  case M:                                                                                         \
    hipLaunchKernelGGL(k_fwd_s2d<M>, grid, block, lds, s, P.d_h, P.d_p, gm, P.d_ext, P.d_chunks, \
                       P.d_col_pair, P.Nx, P.Nz, P.N_pts, P.N_e, P.Np, P.N_ext, P.NZB, P.ZR, S,   \
                       P.Ne_pad, P.d_vpart);                                                      \
    break;
  switch (MT) {
    BORN_FWD_CASE(1)
    BORN_FWD_CASE(2)
    BORN_FWD_CASE(3)
    BORN_FWD_CASE(4)
    BORN_FWD_CASE(5)
    BORN_FWD_CASE(6)
    BORN_FWD_CASE(7)
    BORN_FWD_CASE(8)
    default: break;
  }
#undef BORN_FWD_CASE
}

// z-block partials -> V[pair][e][r] (FP64 sum)
__global__ void k_fwd_reduce(const float2* __restrict__ vpart, const int* __restrict__ pr_chunk,
                             const int* __restrict__ pr_pos, int N_pairs, int N_e, int N_rhs, int NZB,
                             int Ne_pad, double2* __restrict__ V) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t total = (size_t)N_pairs * N_e * N_rhs;
  if (idx >= total) return;
  const int r = (int)(idx % N_rhs);
  const size_t pe = idx / N_rhs;
  const int e = (int)(pe % N_e);
  const int pair = (int)(pe / N_e);
  const int ch = pr_chunk[pair * N_rhs + r], pos = pr_pos[pair * N_rhs + r];
  double sx = 0.0, sy = 0.0;
  for (int zb = 0; zb < NZB; ++zb) {
    const float2 v = vpart[(((size_t)ch * NZB + zb) * Ne_pad + e) * NC + pos];
    sx += v.x;
    sy += v.y;
  }
  V[idx] = make_double2(sx, sy);
}

void launch_fwd_reduce(const Plan& P, int N_rhs, hipStream_t s) {
  const size_t total = (size_t)P.N_pairs * P.N_e * N_rhs;
  hipLaunchKernelGGL(k_fwd_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P.d_vpart,
                     P.d_pr_chunk, P.d_pr_pos, P.N_pairs, P.N_e, N_rhs, P.NZB, P.Ne_pad, P.d_V);
}

// ------------------------------------------------------ general forward
// Any shift map: one block per (pair, element, rhs) gathers h through the
// map (the reference's Phi entries, cu:1114-1120, never materialized).
__global__ __launch_bounds__(256) void k_fwd_general(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ gm,
    const int* __restrict__ shift, const int* __restrict__ pair_fg, int N_pts, int N_e, int N_rhs,
    double2* __restrict__ V) {
  const int pair = blockIdx.x, e = blockIdx.y, r = blockIdx.z;
  const float2* hr = h + (size_t)pair_fg[pair] * N_pts;
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
                     P.d_shift, P.d_pair_fg, P.N_pts, P.N_e, N_rhs, P.d_V);
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

// ------------------------------------------------------------- un-mixing
// W[pair][e][r] = sum over the (mk, f) reading this pair and the active
// slot a of element e: conj(pref[mk][f][a]) * u[r][obs(mk, f)].  A gather,
// so the result is deterministic.
__global__ void k_adj_unmix(const float2* __restrict__ u, const int* __restrict__ pair_list_off,
                            const int2* __restrict__ pair_list, const int* __restrict__ slot,
                            const int* __restrict__ mk_meta, const int64_t* __restrict__ mk_off,
                            const double2* __restrict__ pref, int N_pairs, int N_e, int N_rhs, int64_t N_obs,
                            float2* __restrict__ W) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)N_pairs * N_e * N_rhs) return;
  const int r = (int)(idx % N_rhs);
  const size_t pe = idx / N_rhs;
  const int e = (int)(pe % N_e);
  const int pair = (int)(pe / N_e);
  double sx = 0.0, sy = 0.0;
  for (int j = pair_list_off[pair]; j < pair_list_off[pair + 1]; ++j) {
    const int2 mf = pair_list[j];
    const int a = slot[(size_t)mf.x * N_e + e];
    if (a < 0) continue;
    const int Nf = mk_meta[mf.x * 4 + 1];
    const double2 c = pref[mk_off[mf.x * 4 + 3] + (int64_t)a * Nf + mf.y];
    const float2 v = u[(int64_t)r * N_obs + mk_off[mf.x * 4 + 0] + mf.y];
    sx += c.x * v.x + c.y * v.y;
    sy += c.x * v.y - c.y * v.x;
  }
  W[idx] = make_float2((float)sx, (float)sy);
}

void launch_adj_unmix(const Plan& P, const float2* u, int N_rhs, hipStream_t s) {
  const size_t total = (size_t)P.N_pairs * P.N_e * N_rhs;
  hipLaunchKernelGGL(k_adj_unmix, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, u, P.d_pair_list_off,
                     P.d_pair_list, P.d_slot, P.d_mk_meta, P.d_mk_off, P.d_pref, P.N_pairs, P.N_e, N_rhs,
                     P.N_obs, P.d_W);
}

// ------------------------------------------------- structured adjoint (2-D)
// One workgroup = (depth row z, lateral segment of 64*MW points); it loops
// over every chunk (rhs-major, then frequency) in a fixed order and keeps its
// output in registers, so no atomics are used and results are bitwise
// repeatable.  Per chunk: Y[i][c] = sum_e conj(h_e[i]) W[e][c] on MFMA
// (A = 16 points x 4 elements from the extended row in LDS, B = 4 elements x
// 16 columns), then gamma_hat[i] += sum_c conj(p_inc[c][i]) Y[i][c].
// Sum the per-lane partials over the 16 column lanes of each point group and
// write gamma_hat for rhs r.
template <int MW>
__device__ __forceinline__ void adj_flush(float (&gx)[MW][4], float (&gy)[MW][4], int r, int ib, int kk, int col,
                                          int Nx, int N_pts, int z, float2* __restrict__ g_out) {
#pragma unroll
  for (int mw = 0; mw < MW; ++mw) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float x = gx[mw][v], y = gy[mw][v];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        x += __shfl_xor(x, o, 64);
        y += __shfl_xor(y, o, 64);
      }
      gx[mw][v] = x;
      gy[mw][v] = y;
    }
  }
  if (col < 4) {
#pragma unroll
    for (int mw = 0; mw < MW; ++mw) {
      const int i = ib + 16 * mw + 4 * kk + col;
      float x = gx[mw][0], y = gy[mw][0];
#pragma unroll
      for (int v = 1; v < 4; ++v) {
        if (col == v) {
          x = gx[mw][v];
          y = gy[mw][v];
        }
      }
      if (i < Nx) g_out[(size_t)r * N_pts + (size_t)z * Nx + i] = make_float2(x, y);
    }
  }
}

template <int MW>
__global__ __launch_bounds__(256) void k_adj_s2d(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ W,
    const int* __restrict__ ext, const Chunk* __restrict__ chunks, int n_chunks,
    const int* __restrict__ col_pair, int Nx, int N_pts, int N_e, int Ne_pad, int Np, int N_ext, int N_rhs,
    float2* __restrict__ g_out) {
  extern __shared__ float2 lds[];
  constexpr int SEG = 64 * MW;
  const int L = SEG + (Ne_pad - 1) * Np;
  float2* He = lds;        // [L]
  float2* Ws = lds + L;    // [Ne_pad][NCS]

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int z = blockIdx.y;
  const int i_s = blockIdx.x * SEG;
  const int ib = i_s + w * 16 * MW;
  const int kk = lane >> 4, col = lane & 15;
  const int t_shift = i_s - (Ne_pad - N_e) * Np;

  float gx[MW][4], gy[MW][4];
  int cur_r = -1;

  for (int ci = 0; ci < n_chunks; ++ci) {
    const Chunk ch = chunks[ci];
    if (ch.r != cur_r) {
      if (cur_r >= 0) adj_flush<MW>(gx, gy, cur_r, ib, kk, col, Nx, N_pts, z, g_out);
      cur_r = ch.r;
#pragma unroll
      for (int mw = 0; mw < MW; ++mw)
#pragma unroll
        for (int v = 0; v < 4; ++v) gx[mw][v] = gy[mw][v] = 0.f;
    }
    __syncthreads();
    const float2* hrow = h + (size_t)ch.fg * N_pts + (size_t)z * Nx;
    for (int u = tid; u < L; u += 256) {
      const int t = u + t_shift;
      He[u] = (t >= 0 && t < N_ext) ? hrow[ext[t]] : make_float2(0.f, 0.f);
    }
    for (int idx = tid; idx < Ne_pad * NC; idx += 256) {
      const int k = idx / NC, c = idx - k * NC;
      float2 v = make_float2(0.f, 0.f);
      if (c < ch.ncols && k < N_e) v = W[((size_t)col_pair[ch.col_start + c] * N_e + k) * N_rhs + ch.r];
      Ws[k * NCS + c] = v;
    }
    __syncthreads();

    v4f ar[MW], ai[MW];
#pragma unroll
    for (int mw = 0; mw < MW; ++mw) {
      ar[mw] = v4f{0.f, 0.f, 0.f, 0.f};
      ai[mw] = ar[mw];
    }
    const int ubase = ib - i_s + col + (Ne_pad - 1) * Np;
    for (int k0 = 0; k0 < Ne_pad; k0 += 4) {
      const int k = k0 + kk;
      const float2 b = Ws[k * NCS + col];
      const float nbx = -b.x;
      const int u0 = ubase - k * Np;
#pragma unroll
      for (int mw = 0; mw < MW; ++mw) {
        const float2 a = He[u0 + 16 * mw];
        ar[mw] = mfma16(a.x, b.x, ar[mw]);
        ar[mw] = mfma16(a.y, b.y, ar[mw]);
        ai[mw] = mfma16(a.x, b.y, ai[mw]);
        ai[mw] = mfma16(a.y, nbx, ai[mw]);
      }
    }
    if (col < ch.ncols) {
      const float2* pr = P + (size_t)col_pair[ch.col_start + col] * N_pts + (size_t)z * Nx;
#pragma unroll
      for (int mw = 0; mw < MW; ++mw) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = ib + 16 * mw + 4 * kk + v;
          const float2 p = (i < Nx) ? pr[i] : make_float2(0.f, 0.f);
          gx[mw][v] += p.x * ar[mw][v] + p.y * ai[mw][v];
          gy[mw][v] += p.x * ai[mw][v] - p.y * ar[mw][v];
        }
      }
    }
  }
  if (cur_r >= 0) adj_flush<MW>(gx, gy, cur_r, ib, kk, col, Nx, N_pts, z, g_out);
}

void launch_adj_structured(const Plan& P, int N_rhs, float2* g_out, hipStream_t s) {
  int MW = 8;
  if (P.Nx <= 128) MW = 2;
  else if (P.Nx <= 256) MW = 4;
  const int SEG = 64 * MW;
  const int L = SEG + (P.Ne_pad - 1) * P.Np;
  const size_t lds = ((size_t)L + (size_t)P.Ne_pad * NCS) * sizeof(float2);
  const dim3 grid((P.Nx + SEG - 1) / SEG, P.Nz), block(256);
  const int nch = (int)P.chunks.size();
#define BORN_ADJ_CASE(M)                                                                               \
  case M:                                                                                              \
    hipLaunchKernelGGL(k_adj_s2d<M>, grid, block, lds, s, P.d_h, P.d_p, P.d_W, P.d_ext, P.d_chunks, nch, \
                       P.d_col_pair, P.Nx, P.N_pts, P.N_e, P.Ne_pad, P.Np, P.N_ext, N_rhs, g_out);     \
    break;
  switch (MW) {
    BORN_ADJ_CASE(2)
    BORN_ADJ_CASE(4)
    BORN_ADJ_CASE(8)
    default: break;
  }
#undef BORN_ADJ_CASE
}

// ------------------------------------------------------ general adjoint
__global__ __launch_bounds__(256) void k_adj_general(
    const float2* __restrict__ h, const float2* __restrict__ P, const float2* __restrict__ W,
    const int* __restrict__ shift, const int* __restrict__ pair_fg, int N_pts, int N_e, int N_pairs, int N_rhs,
    float2* __restrict__ g_out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (p >= N_pts) return;
  float gxs = 0.f, gys = 0.f;
  for (int pair = 0; pair < N_pairs; ++pair) {
    const float2* hr = h + (size_t)pair_fg[pair] * N_pts;
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
                     P.d_shift, P.d_pair_fg, P.N_pts, P.N_e, P.N_pairs, N_rhs, g_out);
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
  const int N_pts = g.Nx * g.Ny * g.Nz;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int fg = blockIdx.y;
  if (p >= N_pts) return;
  const int ix = p % g.Nx;
  const int rest = p / g.Nx;
  const int iy = (g.nd == 3) ? rest % g.Ny : 0;
  const int iz = (g.nd == 3) ? rest / g.Ny : rest;
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
  h64[(size_t)fg * N_pts + p] = make_double2(acc.re, acc.im);
}

void launch_build_h(const GeomDev& g, int N_fu, const double2* k, const double2* face_w, double2* h64,
                    hipStream_t s) {
  const int N_pts = g.Nx * g.Ny * g.Nz;
  hipLaunchKernelGGL(k_build_h, dim3((N_pts + 255) / 256, N_fu), dim3(256), 0, s, g, N_fu, k, face_w, h64);
}

// p_in[pair][p] = sum_a h_tx[fg][shift(p, e_a)] v_d[a] in FP64, stored FP32
// (sequence.m:825-840).  v_d of measurement m: complex N_fm x N_tx, col-major.
__global__ __launch_bounds__(256) void k_synth_pinc(const double2* __restrict__ h64, const int* __restrict__ shift,
                                                    const int* __restrict__ pair_fg,
                                                    const int* __restrict__ pair_meas,
                                                    const int* __restrict__ meas_pair_base,
                                                    const int* __restrict__ meas_Nfm, const int* __restrict__ tx_off,
                                                    const int* __restrict__ tx_el, const double2* __restrict__ vd,
                                                    const int* __restrict__ vd_off, int N_pts,
                                                    float2* __restrict__ P) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int pair = blockIdx.y;
  if (p >= N_pts) return;
  const int m = pair_meas[pair];
  const int fm = pair - meas_pair_base[m];
  const int Nfm = meas_Nfm[m];
  const double2* hr = h64 + (size_t)pair_fg[pair] * N_pts;
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
                     P.N_pts, P.d_p);
}

// Allow dynamic LDS beyond the 64 KiB default for the structured kernels.
void set_max_lds(size_t bytes) {
  const int b = (int)bytes;
  hipFuncSetAttribute((const void*)k_fwd_s2d<1>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<2>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<3>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<4>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<5>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<6>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<7>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_fwd_s2d<8>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_adj_s2d<2>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_adj_s2d<4>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
  hipFuncSetAttribute((const void*)k_adj_s2d<8>, hipFuncAttributeMaxDynamicSharedMemorySize, b);
}
}  // namespace born
