// Structured contractions for chunks of at most 4 columns, on the vector ALUs
// (gfx950).
//
// A chunk holds the (wave, frequency) pairs of one frequency, so a plan with
// w plane waves fills w of the 16 MFMA columns of the contraction kernels
// (born_structured.hip, born_split.hip): with one wave, 15/16 of every MFMA
// is padding.  For w <= 4 the Toeplitz products run here as FP32 FMAs
// instead, W = 1, 2, 3 or 4 columns per chunk, with the same inputs and outputs
// as k_fwd_s2d / k_adj_s2d: the forward writes the same row-block partial
// sums (summed in FP64 by k_fwd_reduce), the adjoint writes gamma_hat once
// per right-hand side.
//
// Work split: a task is one phase ph of a row (points i = Np q + ph) and S
// consecutive q, against a block of E consecutive elements.  Element e sees
// point i at extended-row entry t = i - e Np + (N_e - 1) Np (the lateral map
// of setup_grid_symmetric.shift_lateral, setup_grid_symmetric.m:171-233).
// Rows are staged in LDS as in k_fwd_s2d: per phase, Hp[ph][q'] = Hext at
// t = (q' - (Ne_pad - N_e)) Np + ph, so element e at point q of phase ph
// reads Hp[ph][q + Ne_pad - 1 - e], and over a task's (s, j) = (q - S qb,
// e - E eb) these are S + E - 1 consecutive entries, held in registers for
// S x E x W complex products.  Complex products use three real
// multiplications (P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi)): three
// FMAs per complex multiply-add, all operands from registers.
#include "born_internal.h"

namespace born {

namespace {

// Geometry of the staged rows of one instance for a task shape (E, S).
struct FcGeom {
  int QN;    // points per phase (ceil(Nx / Np))
  int NQB;   // q-blocks of S per phase
  int QLF;   // staged entries per phase: S NQB + Ne_pad - 1
  int XPF;   // staged B points per row: Np S NQB
  int NEB;   // element blocks of E
  int tpr;   // tasks per row: Np NQB NEB
};
__host__ __device__ inline FcGeom fc_geom(int Nx, int Np, int Ne_pad, int E, int S) {
  FcGeom g;
  g.QN = (Nx + Np - 1) / Np;
  g.NQB = (g.QN + S - 1) / S;
  g.QLF = S * g.NQB + Ne_pad - 1;
  g.XPF = Np * S * g.NQB;
  g.NEB = Ne_pad / E;
  g.tpr = Np * g.NQB * g.NEB;
  return g;
}

// Forward: rows per step R (the smallest power of 2, at most 8, whose tasks
// fill the 256 threads); row stride in LDS Np QLF + XPF W entries.
__host__ __device__ inline int fc_fwd_rows(const FcGeom& g) {
  int R = 1;
  while (R < 8 && R * g.tpr < 256) R <<= 1;
  return R;
}
// Task shapes: forward W = 1, 2: E 8 x S 16, 3, 4: 8 x 8 (sums of E x W);
// adjoint W = 1: E 16 x S 8, 2: 16 x 8, 3, 4: 16 x 4 with at most fc_km
// tasks per thread (sums of S x W per task).  (W = 1 keeps four workgroups
// per compute unit at the cost of a 16-byte register spill: measured faster
// than E 8, and than three workgroups without the spill.)
constexpr int FC_LDS_MAX = 78 * 1024;   // two workgroups per compute unit
constexpr int FC_ROWS_MAX = 16;         // rows an adjoint workgroup stages (its row table in LDS)
void fc_fwd_shape(int fc_w, int& E, int& S) {
  E = 8;
  S = fc_w >= 3 ? 8 : 16;
}
void fc_adj_shape(int fc_w, int& E, int& S) {
  E = 16;
  S = fc_w >= 3 ? 4 : 8;
}
// tasks per thread: W = 1 one (4 waves per SIMD), W = 2, 4 two; one with
// right-hand sides packed per column (registers)
inline int fc_km(int fc_w, int kr) { return kr > 1 || fc_w == 1 ? 1 : 2; }
// cb staged units (windows, weights, p_inc of R rows, each padded against
// bank conflicts as the adjoint kernel lays them out), window offsets, row table
size_t fc_adj_lds_rows(const FcGeom& g, int Np, int Ne_pad, int fc_w, int R, int cb = 1) {
  int E, S;
  fc_adj_shape(fc_w, E, S);
  const size_t HS = (size_t)Np * ((g.QLF + S - 1) / S * S);
  const size_t row = HS + HS / S + (size_t)g.XPF * fc_w + g.NQB;
  const size_t unit = (size_t)R * row + (size_t)Ne_pad * fc_w + Ne_pad / E;
  return (size_t)cb * unit * sizeof(float2) + (HS + FC_ROWS_MAX) * sizeof(int);
}
// Adjoint work per workgroup, in tasks (BT) of the instance's list (row,
// phase, q-block, element block): R whole rows when their tasks fill at
// least 3/4 of at most fc_km rounds of 256 threads (the best filled, then
// the fewest rows), else 256 fc_km consecutive tasks (row segments);
// NR = the most rows a block touches (at most FC_ROWS_MAX), within the LDS
// budget.  BT = 0: none.
struct FcAdjBlock {
  int BT, NR;
};
FcAdjBlock fc_adj_block(int Nx, int Np, int Ne_pad, int fc_w, int kr, int n_rows) {
  int E, S;
  fc_adj_shape(fc_w, E, S);
  const FcGeom g = fc_geom(Nx, Np, Ne_pad, E, S);
  const int km = fc_km(fc_w, kr);
  FcAdjBlock b{0, 0};
  double best_u = 0.0;
  for (int R = 1; R <= 8; ++R) {
    const int K = (R * g.tpr + 255) / 256;
    if (K > km || fc_adj_lds_rows(g, Np, Ne_pad, fc_w, R) > FC_LDS_MAX) continue;
    const double u = (double)R * g.tpr / (256.0 * K);
    if (u > best_u + 1e-9) {
      best_u = u;
      b = FcAdjBlock{R * g.tpr, R};
    }
  }
  if (best_u >= 0.75) return b;
  const long T = (long)n_rows * g.tpr;
  const int BT = 256 * km;
  int NR = 1;
  for (long t0 = 0; t0 < T; t0 += BT) NR = std::max(NR, (int)((std::min(T, t0 + BT) - 1) / g.tpr - t0 / g.tpr + 1));
  if (NR > FC_ROWS_MAX || fc_adj_lds_rows(g, Np, Ne_pad, fc_w, NR) > FC_LDS_MAX) return b;
  return FcAdjBlock{BT, NR};
}

size_t fc_fwd_lds(int Nx, int Np, int Ne_pad, int fc_w) {
  int E, S;
  fc_fwd_shape(fc_w, E, S);
  const FcGeom g = fc_geom(Nx, Np, Ne_pad, E, S);
  // padded staged rows as the kernel lays them out (then the final
  // reduction), window slot offsets, row table
  const int PW = std::min(S, E);
  const size_t HS = (size_t)Np * ((g.QLF + PW - 1) / PW * PW);
  const size_t rows = (size_t)fc_fwd_rows(g) * (HS + HS / PW + (size_t)g.XPF * fc_w + g.NQB);
  return std::max(rows, (size_t)4 * 128 * fc_w) * sizeof(float2) + (HS + 8) * sizeof(int);
}
size_t fc_adj_lds(int Nx, int Np, int Ne_pad, int fc_w, int NR, int cb) {
  int E, S;
  fc_adj_shape(fc_w, E, S);
  return fc_adj_lds_rows(fc_geom(Nx, Np, Ne_pad, E, S), Np, Ne_pad, fc_w, NR, cb);
}

}  // namespace

int fewcol_fwd_rows(int Nx, int Np, int Ne_pad, int fc_w) {
  int E, S;
  fc_fwd_shape(fc_w, E, S);
  return fc_fwd_rows(fc_geom(Nx, Np, Ne_pad, E, S));
}

int fewcol_width(int max_cols) { return max_cols <= 4 ? std::max(max_cols, 1) : 0; }

bool fewcol_fits(int Nx, int Np, int Ne_pad, int fc_w, int n_rows) {
  if (fc_w <= 0 || !(Ne_pad == 16 || Ne_pad == 32 || Ne_pad == 64 || Ne_pad == 128)) return false;
  if (fc_adj_block(Nx, Np, Ne_pad, fc_w, 0, n_rows).BT == 0) return false;
  return fc_fwd_lds(Nx, Np, Ne_pad, fc_w) <= FC_LDS_MAX;
}

// ------------------------------------------------------------------ forward
// One workgroup per work item (instance, chunk, y slice, row block) -- the
// items of k_fwd_s2d (item_begin, NZB, ZR).  A step stages R rows; thread
// tid takes the tasks tid, tid + 256, ... of the step (element block
// eb = tid mod NEB throughout), and its sums over all steps are reduced over
// the threads of its element block at the end.
template <int W, int E, int S>
__global__ __launch_bounds__(256, 2) void k_fwd_fc(const Inst* __restrict__ inst, int n_inst, int Ne_pad,
                                                   const Chunk* __restrict__ chunks,
                                                   const int* __restrict__ chunk_pairs) {
  extern __shared__ float2 lds[];
  const int item = blockIdx.x;
  int k = 0;
  while (k + 1 < n_inst && item >= inst[k + 1].item_begin) ++k;
  const Inst Iv = inst[k];
  const Inst* I = &Iv;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

  const int li = item - I->item_begin;   // (chunk * Ney + ey) * NZB + zb
  const int ci = li / (I->Ney * I->NZB);
  const int rem = li - ci * (I->Ney * I->NZB);
  const int ey = rem / I->NZB;
  const int z0 = (rem - ey * I->NZB) * I->ZR;
  const int z1 = min(I->n_rows, z0 + I->ZR);
  const Chunk ch = chunks[I->chunk_base + ci];
  const int* cpr = chunk_pairs + (I->chunk_base + ci) * NC;

  const int Np = I->Np, Nx = I->Nx, N_pts = I->N_pts;
  const FcGeom g = fc_geom(Nx, Np, Ne_pad, E, S);
  const int R = fc_fwd_rows(g);
  // Staged row: the phase windows Hp [Np][QLS], then B[i][c], padded against
  // LDS bank conflicts as in the adjoint: the tasks of a wave read windows at
  // offsets S qb - E eb (multiples of PW = min(S, E) entries) and B at
  // (Np S W) qb, so one pad entry follows every PW window entries (phase rows
  // QLS = QLF rounded up to PW) and every Np S W entries of B.
  constexpr int PW = S < E ? S : E;
  const int QLS = (g.QLF + PW - 1) / PW * PW;
  const int HS = Np * QLS, BS = g.XPF * W;            // logical lengths
  const int HSP = HS + HS / PW, RS = HSP + BS + g.NQB; // padded row
  const int Qoff = Ne_pad - I->N_e;
  const float2* hf = I->h + (size_t)ch.fg * I->Hp;
  const float2* gmr = I->gm + (size_t)ch.r * N_pts;
  const int eb = tid % g.NEB;
  int pr[W];
#pragma unroll
  for (int c = 0; c < W; ++c) pr[c] = cpr[c];
  // table x index of each staged window slot (-1: outside the table), and
  // the table row of each staged row of the step (-1: past the block)
  const size_t rows_bytes = std::max((size_t)R * RS, (size_t)4 * 128 * W) * sizeof(float2);
  int* dh = reinterpret_cast<int*>(reinterpret_cast<char*>(lds) + rows_bytes);
  int* hr = dh + HS;
  for (int o = tid; o < HS; o += 256) {
    const int ph = o / QLS, q = o - ph * QLS;
    const int t = (q - Qoff) * Np + ph;
    dh[o] = q < g.QLF && t >= 0 && t < I->N_ext ? I->ext[t] : -1;
  }

  float a1[E][W], a2[E][W], a3[E][W];
#pragma unroll
  for (int j = 0; j < E; ++j)
#pragma unroll
    for (int c = 0; c < W; ++c) a1[j][c] = a2[j][c] = a3[j][c] = 0.f;

  for (int z = z0; z < z1; z += R) {
    if (tid < R) hr[tid] = z + tid < z1 ? I->hrow[ey * I->n_rows + z + tid] : -1;
    __syncthreads();
    for (int r = 0; r < R; ++r) {
      const int row = hr[r], zr = z + r;
      float2* dst = lds + r * RS;
      for (int o = tid; o < HS; o += 256) {
        const int e = dh[o];
        dst[o + o / PW] = row >= 0 && e >= 0 ? hf[(size_t)row * I->HR + e] : make_float2(0.f, 0.f);
      }
      for (int o = tid; o < BS; o += 256) {
        const int i = o / W, c = o - i * W;
        float2 v = make_float2(0.f, 0.f);
        if (row >= 0 && i < Nx) {
          int q = pr[0];
#pragma unroll
          for (int cc = 1; cc < W; ++cc)
            if (c == cc) q = pr[cc];
          const size_t p = (size_t)zr * Nx + i;
          const float2 a = I->P[(size_t)q * N_pts + p], b = gmr[p];
          v = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
        }
        dst[HSP + o + o / (Np * S * W)] = v;
      }
    }
    __syncthreads();
    for (int tk = tid; tk < R * g.tpr; tk += 256) {
      const int r = tk / g.tpr, t2 = (tk - r * g.tpr) / g.NEB;
      const int ph = t2 / g.NQB, qb = t2 - ph * g.NQB;
      const int hl = ph * QLS + S * qb + Ne_pad - E * eb - E;   // a multiple of PW
      const float2* Hr = lds + r * RS + hl + hl / PW;
      const float2* Br = lds + r * RS + HSP + qb * (Np * S * W + 1) + ph * W;
      float wr[S + E - 1], wi[S + E - 1], ws[S + E - 1];
#pragma unroll
      for (int u = 0; u < S + E - 1; ++u) {
        const float2 a = Hr[u + u / PW];
        wr[u] = a.x;
        wi[u] = a.y;
        ws[u] = a.x + a.y;
      }
#pragma unroll
      for (int s = 0; s < S; ++s) {
        float br[W], bi[W], bs[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
          const float2 b = Br[Np * s * W + c];
          br[c] = b.x;
          bi[c] = b.y;
          bs[c] = b.x + b.y;
        }
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const int u = s - j + E - 1;
#pragma unroll
          for (int c = 0; c < W; ++c) {
            a1[j][c] = fmaf(wr[u], br[c], a1[j][c]);
            a2[j][c] = fmaf(wi[u], bi[c], a2[j][c]);
            a3[j][c] = fmaf(ws[u], bs[c], a3[j][c]);
          }
        }
      }
    }
  }

  // Re = P1 - P2, Im = P3 - P1 - P2; sum over the lanes of one element block
  // (lane bits from NEB up), then over the 4 waves through LDS
  __syncthreads();
  float2* red = lds;   // [wave][element][column]
#pragma unroll
  for (int j = 0; j < E; ++j)
#pragma unroll
    for (int c = 0; c < W; ++c) {
      float x = a1[j][c] - a2[j][c], y = a3[j][c] - a1[j][c] - a2[j][c];
      for (int o = g.NEB; o < 64; o <<= 1) {
        x += __shfl_xor(x, o, 64);
        y += __shfl_xor(y, o, 64);
      }
      if (lane < g.NEB) red[(w * 128 + E * lane + j) * W + c] = make_float2(x, y);
    }
  __syncthreads();
  float2* out = I->vpart + (size_t)li * Ne_pad * NC;   // [element][column]
  for (int idx = tid; idx < Ne_pad * W; idx += 256) {
    const int e = idx / W, c = idx - e * W;
    float2 v = red[e * W + c];
#pragma unroll
    for (int ww = 1; ww < 4; ++ww) {
      const float2 u = red[(ww * 128 + e) * W + c];
      v.x += u.x;
      v.y += u.y;
    }
    out[(size_t)e * NC + c] = v;
  }
}

// ------------------------------------------------------------------ adjoint
// Workgroups of instance k: [wg_begin_f, wg_begin_f + blocks * G); workgroup
// wl owns the tasks [BT b, BT b + BT) of the instance's list (row, phase,
// q-block, element block; b = wl / G, BT = fc_bt: whole rows or a row
// segment) and the chunks g, g + G, g + 2G, ... (g = wl mod G; chunks are
// rhs-major, then frequency).  Its tasks go to the threads in rounds of 256,
// at most KM per thread.  Per chunk and y slice the workgroup stages the
// windows of the rows its tasks touch,
// the slice's weights and, with the first slice, the p_inc rows; then every
// task computes
//   Y[s][c]   = sum_e conj(Hext[t(s, e)]) W[e][c]      (its E elements)
//   Gh[r][s] += sum_c conj(p_inc[pair_c][i_s]) Y[s][c]  (r = rhs of column c)
// (p_inc does not depend on the slice, so each slice's Y is applied as it
// is done).  The NEB element-block lanes of a task (adjacent lanes) are
// summed when a right-hand side is complete within the group, and written
// to the group's partial gpart[g][r][point]; k_adj_fc_reduce sums the G
// partials in order (no atomics: bitwise repeatable).  KR = 1: every column
// belongs to the chunk's rhs; KR = col_rhs (2 or 4): column c belongs to rhs
// c mod KR, and all are written after the group's last chunk.
template <int W, int E, int S, int KR, int KM>
__global__ __launch_bounds__(256, (KM == 1 && W == 1 ? 4 : 2)) void k_adj_fc(const Inst* __restrict__ inst, int n_inst, int Ne_pad,
                                                   const Chunk* __restrict__ chunks,
                                                   const int* __restrict__ chunk_pairs) {
  extern __shared__ float2 lds[];
  int k = 0;
  while (k + 1 < n_inst && (int)blockIdx.x >= inst[k + 1].wg_begin_f) ++k;
  const Inst Iv = inst[k];
  const Inst* I = &Iv;
  const int tid = threadIdx.x;

  const int Np = I->Np, Nx = I->Nx, N_pts = I->N_pts, G = I->fc_g, BT = I->fc_bt;
  const FcGeom g = fc_geom(Nx, Np, Ne_pad, E, S);
  const int wl = (int)blockIdx.x - I->wg_begin_f;
  const int grp = wl % G;
  const int T = I->n_rows * g.tpr;
  const int t_first = (wl / G) * BT, t_end = min(T, t_first + BT);
  const int r_first = t_first / g.tpr, NR = (t_end - 1) / g.tpr - r_first + 1;
  const int t_off = t_first - r_first * g.tpr;   // first task's position in its row
  // Staged layouts, padded against LDS bank conflicts: the tasks of a wave
  // read their windows at offsets S qb - E eb (multiples of S entries), their
  // weights at E W eb and their p_inc at (Np S W) qb, strides that fall on
  // one or two bank positions.  One pad entry after every S window entries
  // (phase rows QLS = QLF rounded up to S, so every task's window starts at a
  // multiple of S), every E W weight entries and every Np S W p_inc entries
  // spreads them over the banks with fixed read offsets.
  const int QLS = (g.QLF + S - 1) / S * S;
  const int HS = Np * QLS, PS = g.XPF * W;             // logical row lengths
  const int HSP = HS + HS / S, PSP = PS + g.NQB;       // padded
  const int nH = NR * HSP, nW = Ne_pad * W + Ne_pad / E, nP = NR * PSP;
  // staged units, one per chunk of a step: windows [NR][Np][QLS], weights
  // [Ne_pad][W], p_inc [NR][XPF][W] (padded); then the window offsets and
  // row table
  const int unit = nH + nW + nP;
  int* dh = reinterpret_cast<int*>(lds + I->fc_cb * unit);   // window slot -> table x index (-1: outside)
  int* hr = dh + HS;                           // staged row -> table row (NR <= FC_ROWS_MAX)
  const int Qoff = Ne_pad - I->N_e;
  const int K = (BT + 255) / 256;              // task rounds (<= KM)
  const int ntask = t_end - t_first;
  const int eb = tid % g.NEB;

  // the thread's tasks: window, p_inc and output offsets
  int hoff[KM], poff[KM], oofs[KM], xk[KM];
  bool live[KM];
#pragma unroll
  for (int kk = 0; kk < KM; ++kk) {
    const int tk = tid + 256 * kk;
    live[kk] = kk < K && tk < ntask;
    const int tkc = (live[kk] ? tk : 0) + t_off;
    const int r = tkc / g.tpr, t2 = (tkc - r * g.tpr) / g.NEB;
    const int ph = t2 / g.NQB, qb = t2 - ph * g.NQB;
    const int hl = r * HS + ph * QLS + S * qb + Ne_pad - E * eb - E;   // a multiple of S
    hoff[kk] = hl + hl / S;
    poff[kk] = r * PSP + qb * (Np * S * W + 1) + ph * W;
    xk[kk] = Np * S * qb + ph;                          // x of s = 0
    oofs[kk] = (r_first + r) * Nx + xk[kk];
  }
  for (int o = tid; o < HS; o += 256) {
    const int p2 = o / QLS, q = o - p2 * QLS;
    const int t = (q - Qoff) * Np + p2;
    dh[o] = q < g.QLF && t >= 0 && t < I->N_ext ? I->ext[t] : -1;
  }

  float gr[KM][KR][S], gi[KM][KR][S];
#pragma unroll
  for (int kk = 0; kk < KM; ++kk)
#pragma unroll
    for (int q = 0; q < KR; ++q)
#pragma unroll
      for (int s = 0; s < S; ++s) gr[kk][q][s] = gi[kk][q][s] = 0.f;

  const int Ney = I->Ney, n_chunks = I->n_chunks, cb = I->fc_cb;
  // a step stages cb of the group's chunks (2-D; 3-D: one chunk and slice,
  // cb = 1)
  for (int m = 0; grp + G * m < n_chunks; m += cb) {
    const int nck = min(cb, (n_chunks - grp - G * m + G - 1) / G);
    for (int ey = 0; ey < Ney; ++ey) {
      if (tid < NR) hr[tid] = I->hrow[ey * I->n_rows + r_first + tid];
      __syncthreads();
      for (int j = 0; j < nck; ++j) {
        const int cix = grp + G * (m + j);
        const Chunk ch = chunks[I->chunk_base + cix];
        const float2* hf = I->h + (size_t)ch.fg * I->Hp;
        float2* Hs = lds + j * unit;
        float2* Ws = Hs + nH;
        float2* Ps = Ws + nW;
        for (int rr = 0; rr < NR; ++rr) {
          const int row = hr[rr];
          for (int o = tid; o < HS; o += 256) {
            const int e = dh[o];
            Hs[rr * HSP + o + o / S] = e >= 0 ? hf[(size_t)row * I->HR + e] : make_float2(0.f, 0.f);
          }
        }
        const float2* wc = I->Wc + ((size_t)cix * Ney + ey) * Ne_pad * NC;
        for (int o = tid; o < Ne_pad * W; o += 256) {
          const int e = o / W, c = o - e * W;
          Ws[o + o / (E * W)] = wc[e * NC + c];
        }
        if (ey == 0) {
          const int* cpr = chunk_pairs + (I->chunk_base + cix) * NC;
          int pr[W];
#pragma unroll
          for (int c = 0; c < W; ++c) pr[c] = cpr[c];
          for (int rr = 0; rr < NR; ++rr) {
            const size_t rowo = (size_t)(r_first + rr) * Nx;
            for (int o = tid; o < PS; o += 256) {
              const int i = o / W, c = o - i * W;
              int q = pr[0];
#pragma unroll
              for (int cc = 1; cc < W; ++cc)
                if (c == cc) q = pr[cc];
              float2 v = make_float2(0.f, 0.f);
              if (i < Nx) v = I->P[(size_t)q * N_pts + rowo + i];
              Ps[rr * PSP + o + o / (Np * S * W)] = v;
            }
          }
        }
      }
      __syncthreads();

      for (int j = 0; j < nck; ++j) {
        const float2* Hs = lds + j * unit;
        const float2* Ws = Hs + nH;
        const float2* Ps = Ws + nW;
#pragma unroll
        for (int kk = 0; kk < KM; ++kk) {
          if (kk >= K) continue;
          // conj(A) W: P1 = Ar Wr, P2 = Ai Wi, P3 = (Ar - Ai)(Wr + Wi)
          float y1[S][W], y2[S][W], y3[S][W];
#pragma unroll
          for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < W; ++c) y1[s][c] = y2[s][c] = y3[s][c] = 0.f;
          {
            const float2* Hr = Hs + hoff[kk];
            const float2* Wr = Ws + (E * W + 1) * eb;
            float wr[S + E - 1], wi[S + E - 1], wd[S + E - 1];
#pragma unroll
            for (int u = 0; u < S + E - 1; ++u) {
              const float2 a = Hr[u + u / S];
              wr[u] = a.x;
              wi[u] = a.y;
              wd[u] = a.x - a.y;
            }
#pragma unroll
            for (int jj = 0; jj < E; ++jj) {
              float vr[W], vi[W], vs[W];
#pragma unroll
              for (int c = 0; c < W; ++c) {
                const float2 v = Wr[jj * W + c];
                vr[c] = v.x;
                vi[c] = v.y;
                vs[c] = v.x + v.y;
              }
#pragma unroll
              for (int s = 0; s < S; ++s) {
                const int u = s - jj + E - 1;
#pragma unroll
                for (int c = 0; c < W; ++c) {
                  y1[s][c] = fmaf(wr[u], vr[c], y1[s][c]);
                  y2[s][c] = fmaf(wi[u], vi[c], y2[s][c]);
                  y3[s][c] = fmaf(wd[u], vs[c], y3[s][c]);
                }
              }
            }
          }
          // Re = P1 + P2, Im = P3 - P1 + P2; gamma_hat += conj(p) Y
          const float2* Pr = Ps + poff[kk];
#pragma unroll
          for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int c = 0; c < W; ++c) {
              const float2 pv = Pr[Np * s * W + c];
              const float yr = y1[s][c] + y2[s][c], yi = y3[s][c] - y1[s][c] + y2[s][c];
              gr[kk][c % KR][s] += pv.x * yr + pv.y * yi;
              gi[kk][c % KR][s] += pv.x * yi - pv.y * yr;
            }
          }
        }
        if (ey + 1 < Ney) continue;
        // a right-hand side is complete within the group after its last chunk
        // there (KR = 1), all of them after the group's last chunk (KR > 1)
        const int cix = grp + G * (m + j);
        const int rhs = chunks[I->chunk_base + cix].r;
        const bool last = cix + G >= n_chunks;
        const bool flush = last || (KR == 1 && chunks[I->chunk_base + cix + G].r != rhs);
        if (!flush) continue;
#pragma unroll
        for (int kk = 0; kk < KM; ++kk) {
          if (kk >= K) continue;
#pragma unroll
          for (int q = 0; q < KR; ++q) {
            if (KR > 1 && q >= I->n_rhs) continue;
            float2* out = I->gpart + ((size_t)grp * I->n_out + (KR > 1 ? q : rhs)) * N_pts + oofs[kk];
#pragma unroll
            for (int s = 0; s < S; ++s) {
              float x = gr[kk][q][s], y = gi[kk][q][s];
              for (int o = 1; o < g.NEB; o <<= 1) {
                x += __shfl_xor(x, o, 64);
                y += __shfl_xor(y, o, 64);
              }
              if (live[kk] && eb == 0 && xk[kk] + Np * s < Nx) out[Np * s] = make_float2(x, y);
              gr[kk][q][s] = gi[kk][q][s] = 0.f;
            }
          }
        }
      }
    }
  }
}

// gamma_hat[r][p] = sum over the G groups of gpart[g][r][p] (in group order).
__global__ void k_adj_fc_reduce(const float2* __restrict__ gpart, int G, size_t n, float2* __restrict__ out) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  float2 v = gpart[idx];
  for (int q = 1; q < G; ++q) {
    const float2 u = gpart[(size_t)q * n + idx];
    v.x += u.x;
    v.y += u.y;
  }
  out[idx] = v;
}

// ------------------------------------------------------------------ launch
void launch_fwd_fc(const LaunchSet& L, hipStream_t s) {
  const int Ne_pad = 16 * L.MT;
  size_t lds = 0;
  for (const Inst& I : L.inst) lds = std::max(lds, fc_fwd_lds(I.Nx, I.Np, Ne_pad, L.fc_w));
  const dim3 grid(L.n_items);
  const int n_inst = (int)L.inst.size();
#define BORN_FC_F(Wv, Ev, Sv)                                                                            \
  hipLaunchKernelGGL((k_fwd_fc<Wv, Ev, Sv>), grid, dim3(256), lds, s, L.d_inst, n_inst, Ne_pad, L.d_chunks, \
                     L.d_chunk_pairs)
  if (L.fc_w == 1) BORN_FC_F(1, 8, 16);
  else if (L.fc_w == 2) BORN_FC_F(2, 8, 16);
  else if (L.fc_w == 3) BORN_FC_F(3, 8, 8);
  else BORN_FC_F(4, 8, 8);
#undef BORN_FC_F
}

void launch_adj_fc(const LaunchSet& L, const OutPtrs& outs, hipStream_t s) {
  const int Ne_pad = 16 * L.MT;
  size_t lds = 0;
  for (const Inst& I : L.inst) lds = std::max(lds, fc_adj_lds(I.Nx, I.Np, Ne_pad, L.fc_w, I.fc_nr, I.fc_cb));
  const int n_inst = (int)L.inst.size();
  // the partials start at zero: a group writes only the right-hand sides it has chunks of
  for (const Inst& I : L.inst) launch_zero_f2(I.gpart, (size_t)I.fc_g * I.n_out * I.N_pts, s);
  const int kr = L.inst.empty() ? 0 : L.inst[0].col_rhs;
#define BORN_FC_A(Wv, Ev, Sv, KRv, KMv)                                                                   \
  hipLaunchKernelGGL((k_adj_fc<Wv, Ev, Sv, KRv, KMv>), dim3(L.n_wg_f), dim3(256), lds, s, L.d_inst, n_inst, \
                     Ne_pad, L.d_chunks, L.d_chunk_pairs)
  if (L.fc_w == 1) BORN_FC_A(1, 16, 8, 1, 1);
  else if (L.fc_w == 2 && kr == 2) BORN_FC_A(2, 16, 8, 2, 1);
  else if (L.fc_w == 2) BORN_FC_A(2, 16, 8, 1, 2);
  else if (L.fc_w == 3) BORN_FC_A(3, 16, 4, 1, 2);   // 3 columns: one rhs per chunk
  else if (kr == 4) BORN_FC_A(4, 16, 4, 4, 1);
  else if (kr == 2) BORN_FC_A(4, 16, 4, 2, 1);
  else BORN_FC_A(4, 16, 4, 1, 2);
#undef BORN_FC_A
  for (int k = 0; k < n_inst; ++k) {
    const Inst& I = L.inst[k];
    const size_t n = (size_t)I.n_out * I.N_pts;
    hipLaunchKernelGGL(k_adj_fc_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, I.gpart, I.fc_g, n,
                       outs.p[k]);
  }
}

// Adjoint tasks per workgroup and the most rows a workgroup stages (BT = 0: none fits).
void fewcol_adj_block(int Nx, int Np, int Ne_pad, int fc_w, int col_rhs, int n_rows, int& BT, int& NR) {
  const FcAdjBlock b = fc_adj_block(Nx, Np, Ne_pad, fc_w, col_rhs, n_rows);
  BT = b.BT;
  NR = b.NR;
}
// Chunks an adjoint workgroup stages per step: up to 4 (at most per_group),
// within the LDS that keeps the kernel's workgroups per compute unit (4 for
// W = 1 with one task per thread, else 2).  3-D plans (Ney > 1) take one: a
// step walks its chunks slice by slice, and the slices of one chunk must all
// be summed before the next chunk (possibly another right-hand side) adds to
// the same accumulators.
int fewcol_adj_cb(int Nx, int Np, int Ne_pad, int fc_w, int col_rhs, int NR, int per_group, int Ney) {
  if (Ney > 1) return 1;
  int E, S;
  fc_adj_shape(fc_w, E, S);
  const FcGeom g = fc_geom(Nx, Np, Ne_pad, E, S);
  const size_t budget = fc_w == 1 && fc_km(fc_w, col_rhs) == 1 ? FC_LDS_MAX / 2 : FC_LDS_MAX;
  int cb = 1;
  while (cb < 4 && cb < per_group && fc_adj_lds_rows(g, Np, Ne_pad, fc_w, NR, cb + 1) <= budget) ++cb;
  return cb;
}
// Tasks per row of the adjoint.
int fewcol_adj_tpr(int Nx, int Np, int Ne_pad, int fc_w) {
  int E, S;
  fc_adj_shape(fc_w, E, S);
  return fc_geom(Nx, Np, Ne_pad, E, S).tpr;
}

// Allow the few-column kernels their dynamic LDS.
void set_max_lds_fc(int bytes) {
#define BORN_SET(K) (void)hipFuncSetAttribute((const void*)(K), hipFuncAttributeMaxDynamicSharedMemorySize, bytes)
  BORN_SET((k_fwd_fc<1, 8, 16>));
  BORN_SET((k_fwd_fc<2, 8, 16>));
  BORN_SET((k_fwd_fc<3, 8, 8>));
  BORN_SET((k_fwd_fc<4, 8, 8>));
  BORN_SET((k_adj_fc<1, 16, 8, 1, 1>));
  BORN_SET((k_adj_fc<2, 16, 8, 1, 2>));
  BORN_SET((k_adj_fc<2, 16, 8, 2, 1>));
  BORN_SET((k_adj_fc<3, 16, 4, 1, 2>));
  BORN_SET((k_adj_fc<4, 16, 4, 1, 2>));
  BORN_SET((k_adj_fc<4, 16, 4, 2, 1>));
  BORN_SET((k_adj_fc<4, 16, 4, 4, 1>));
#undef BORN_SET
}

}  // namespace born
