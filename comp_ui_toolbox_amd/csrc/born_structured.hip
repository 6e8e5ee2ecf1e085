// Structured 2-D grid contractions of the Born operator pair (gfx950).
//
// setup_grid_symmetric.shift_lateral (setup_grid_symmetric.m:95-241) makes
// element e see grid point (i, z) at h_ref point (ext[i - e*Np + (N_e-1)*Np], z):
// for one (frequency, depth row) every element's transfer function is a
// window of one extended lateral row.  Per row the forward is therefore the
// Toeplitz-in-x complex GEMM
//     V[e][c] += sum_i Hext[i - e*Np + off] * B[i][c],  B[i][c] = p_inc[pair_c][i] gamma_r[i]
// and the adjoint its conjugate transpose.  Both run on 16x16x4 FP32 MFMA with
// the three-multiplication complex product (Re = P1 - P2 or P1 + P2, Im from
// the product of the sums), so each complex multiply-add costs three real
// MFMA products instead of four.  Launches whose chunks hold at most 8
// columns take the stacked form instead (STK): real and imaginary parts of
// the 8 columns side by side in the 16 MFMA columns, two MFMAs per tile.
//
// Data movement: the next row (forward) or next chunk (adjoint) is loaded
// from HBM into registers while the MFMAs of the current one run, then
// written to LDS between two barriers.  Two 4-wave workgroups per CU; the
// MFMA accumulators stay in architectural VGPRs.
#include "born_internal.h"

namespace born {

typedef float v4f __attribute__((ext_vector_type(4)));

namespace {
constexpr int NCS = NC + 1;   // padded LDS stride of a 16-column row
constexpr int HPT_MAX = 8;    // per-thread entries of the extended-row copy
constexpr int WPT_MAX = 8;    // per-thread entries of the W tile (Ne_pad <= 128)

__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f vzero() { return v4f{0.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
}  // namespace

// ------------------------------------------------------------------ forward
// Work item = (instance, chunk, z-block of ZR rows); workgroups are
// persistent and take items blockIdx.x, blockIdx.x + gridDim.x, ... (the
// items of instance k are [item_begin_k, item_begin_{k+1})).  Within a row
// the four waves split the work as EW element groups x PW point groups (2 x 2
// when MT >= 2): wave w holds element tiles [eg*MTW, (eg+1)*MTW) x 16 columns
// and walks the K-steps [pg*SW, (pg+1)*SW) of the row.  NP = points per pitch
// (4; 0 = runtime), PPT = row points per thread in the B-tile fill.
//
// The LDS layout follows the instance (row length), so the layout of the
// row held in LDS (used by the MFMA phase) and that of the prefetched row
// (used by the next fill) are tracked separately.
//
// DUAL (rows of at most 128 points): a step holds two consecutive rows, their
// B tiles side by side (points [0, XP) and [XP, 2 XP)) and their two
// extended h rows one after the other, so the point groups of the waves
// split at the row boundary and each barrier pair serves twice the MFMAs.
//
// STK (chunks of at most 8 columns): the B operand holds the real parts of
// the 8 columns in lanes 0-7 of each K row and their imaginary parts in lanes
// 8-15, so two MFMAs per tile and K-step, A_re [B_re | B_im] and
// A_im [B_re | B_im], give all four real products of 8 columns (three MFMAs
// over 16 columns otherwise); the item's sums are combined across lanes
// c and c + 8 when it is written.  Only the 8 columns are loaded and filled.
template <int MT, int NP, int PPT, bool DUAL, bool STK>
__global__ __launch_bounds__(256, 2) void k_fwd_s2d(const Inst* __restrict__ inst, int n_inst, int n_items,
                                                    int Ne_pad, const Chunk* __restrict__ chunks,
                                                    const int* __restrict__ chunk_pairs) {
  extern __shared__ float2 lds[];
  // entries of the polyphase copy per thread: XP <= 256*PPT, QL = (XP-1)/Np + 16*MT
  constexpr int HPT = NP ? PPT + (16 * MT * NP + 255) / 256 : HPT_MAX;
  constexpr int EW = MT >= 2 ? 2 : 1;        // element groups
  constexpr int PW = 4 / EW;                 // point groups
  constexpr int MTW = MT / EW;               // element tiles per wave
  constexpr int NCF = STK ? NC / 2 : NC;     // columns loaded and filled
  constexpr int NCB = NCF + 1;               // padded LDS stride of a B row
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int eg = w % EW, pg = w / EW;
  const int kk = lane >> 4, col = lane & 15;
  const int n_my = (n_items - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  if (n_my <= 0) return;

  // instance of the prefetched row ("n") and of the row in LDS ("c")
  int k = 0;
  Inst Iv = inst[0];            // record of the instance of the prefetched row
  const Inst* I = &Iv;
  int Np, XP, QL, HN;           // layout of the prefetched row
  int eidx[HPT];
  int fg, rr, pr[NC];           // chunk of the prefetched row
  int ey;                       // y slice of the prefetched row (3-D; 0 in 2-D)
  int z, z_end, item = blockIdx.x;
  constexpr int ZS = DUAL ? 2 : 1;   // rows per step

#define BORN_FWD_INST()                                                                   \
  {                                                                                       \
    while (k + 1 < n_inst && item >= inst[k + 1].item_begin) ++k;                         \
    Iv = inst[k];                                                                         \
    Np = NP ? NP : I->Np;                                                                 \
    XP = 16 * I->S;                                                                       \
    QL = (XP - 1) / Np + 16 * MT;                                                         \
    HN = Np * QL;                                                                         \
    const int Qoff_ = 16 * MT - I->N_e;                                                   \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) {                                    \
      const int idx_ = tid + 256 * j;                                                     \
      eidx[j] = -1;                                                                       \
      if (idx_ < HN) {                                                                    \
        const int ph_ = idx_ / QL, q_ = idx_ - ph_ * QL;                                  \
        const int t_ = (q_ - Qoff_) * Np + ph_;                                           \
        if (t_ >= 0 && t_ < I->N_ext) eidx[j] = I->ext[t_];                               \
      }                                                                                   \
    }                                                                                     \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) asm volatile("" ::"v"(eidx[j]));     \
  }
  // descriptor of the item's chunk (uniform: scalar registers); unused
  // columns read the zero row of the p_inc table
#define BORN_FWD_DESC()                                                                   \
  {                                                                                       \
    const int li_ = item - I->item_begin;      /* = (chunk * Ney + ey) * NZB + zb */       \
    const int ci_ = li_ / (I->Ney * I->NZB);                                              \
    const int rem_ = li_ - ci_ * (I->Ney * I->NZB);                                       \
    ey = rem_ / I->NZB;                                                                   \
    const Chunk c_ = chunks[I->chunk_base + ci_];                                         \
    fg = c_.fg;                                                                           \
    rr = c_.r;                                                                            \
    _Pragma("unroll") for (int q = 0; q < NC; ++q) pr[q] = chunk_pairs[(I->chunk_base + ci_) * NC + q]; \
    z = (rem_ - ey * I->NZB) * I->ZR;                                                     \
    z_end = min(I->n_rows, z + I->ZR);                                                    \
  }

  float2 hreg[HPT];
  float2 hreg2[DUAL ? HPT : 1];   // the step's second h row (DUAL)
  float2 praw[PPT][NCF];
  float2 greg[PPT];
  int gvalid = 0;                 // bit j: B point tid + 256 j inside its row (DUAL)

  // issue the loads of row zz of the current chunk into registers; every
  // address is valid (clamped), out-of-range entries are zeroed in the fill
#define BORN_FWD_PREFETCH(zz)                                                           \
  {                                                                                     \
    const float2* hrow_ = I->h + (size_t)fg * I->Hp + (size_t)I->hrow[ey * I->n_rows + (zz)] * I->HR; \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) hreg[j] = hrow_[max(eidx[j], 0)];  \
    if (DUAL) {                                                                         \
      /* second row of the step: zz + 1 inside the item, else row zz again (zeroed) */  \
      const int two_ = (zz) + 1 < z_end ? 1 : 0;                                        \
      const float2* hrow2_ = I->h + (size_t)fg * I->Hp + (size_t)I->hrow[ey * I->n_rows + (zz) + two_] * I->HR; \
      _Pragma("unroll") for (int j = 0; j < HPT; ++j) hreg2[j] = hrow2_[max(eidx[j], 0)]; \
      const int XP1_ = 16 * I->S;                                                       \
      gvalid = 0;                                                                       \
      _Pragma("unroll") for (int j = 0; j < PPT; ++j) {                                 \
        const int t_ = tid + 256 * j;                                                   \
        const int r_ = t_ >= XP1_ ? 1 : 0;                                              \
        const int x_ = t_ - r_ * XP1_;                                                  \
        if (t_ < 2 * XP1_ && x_ < I->Nx && r_ <= two_) gvalid |= 1 << j;                \
        const size_t o_ = (size_t)((zz) + (r_ & two_)) * I->Nx + min(x_, I->Nx - 1);    \
        greg[j] = I->gm[(size_t)rr * I->N_pts + o_];                                    \
        _Pragma("unroll") for (int c_ = 0; c_ < NCF; ++c_)                              \
          praw[j][c_] = I->P[(size_t)pr[c_] * I->N_pts + o_];                           \
      }                                                                                 \
    } else {                                                                            \
      const size_t rowo_ = (size_t)(zz) * I->Nx;                                        \
      const float2* grow_ = I->gm + (size_t)rr * I->N_pts + rowo_;                      \
      _Pragma("unroll") for (int j = 0; j < PPT; ++j) {                                 \
        const int i_ = min(tid + 256 * j, I->Nx - 1);                                   \
        greg[j] = grow_[i_];                                                            \
        _Pragma("unroll") for (int c_ = 0; c_ < NCF; ++c_)                              \
          praw[j][c_] = I->P[(size_t)pr[c_] * I->N_pts + rowo_ + i_];                   \
      }                                                                                 \
    }                                                                                   \
  }

  BORN_FWD_INST()
  BORN_FWD_DESC()
  BORN_FWD_PREFETCH(z)
  int n_Nx = I->Nx;              // row length of the prefetched row

  v4f SR[MTW], SI[MTW];
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt) {
    SR[mt] = vzero();
    SI[mt] = vzero();
  }
  const int b_base = kk * NCB + (STK ? (col & 7) : col);
  int it = 0;

  while (true) {
    // layout of the row about to be written to LDS
    const int cNp = Np, cQL = QL, cHN = HN, cXP1 = XP, cXP = ZS * XP;
    float2* Hp = lds;              // [ZS][Np][QL] polyphase extended rows
    float2* Bs = lds + ZS * cHN;   // [ZS * XP][NCB]
    __syncthreads();
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const int idx = tid + 256 * j;
      if (idx < cHN) {
        Hp[idx] = eidx[j] >= 0 ? hreg[j] : make_float2(0.f, 0.f);
        if (DUAL) Hp[cHN + idx] = eidx[j] >= 0 ? hreg2[j] : make_float2(0.f, 0.f);
      }
    }
    {
      const bool xfull = cXP == n_Nx;   // no padded points in the B tile
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int i = tid + 256 * j;
        if (i < cXP) {
          const bool ok = DUAL ? ((gvalid >> j) & 1) != 0 : (xfull || i < n_Nx);
          const float2 g = ok ? greg[j] : make_float2(0.f, 0.f);
#pragma unroll
          for (int c = 0; c < NCF; ++c) Bs[i * NCB + c] = cmulf(praw[j][c], g);
        }
      }
    }
    __syncthreads();

    // next step and its prefetch
    bool has_next = true;
    const int item_cur = item;
    float2* const vpart_cur = Iv.vpart;
    const int ib_cur = Iv.item_begin;
    const bool item_done = z + ZS >= z_end;
    if (!item_done) {
      z += ZS;
    } else if (it + 1 >= n_my) {
      has_next = false;
    } else {
      item = blockIdx.x + (it + 1) * gridDim.x;
      if (k + 1 < n_inst && item >= inst[k + 1].item_begin) BORN_FWD_INST()
      BORN_FWD_DESC()
    }
    if (has_next) {
      BORN_FWD_PREFETCH(z)
      n_Nx = I->Nx;
    }

    // MFMA over this wave's points of the row in LDS
    const int SW = cXP / (4 * PW);               // K-steps per wave and step
    // DUAL: a wave's K-steps lie in one row (cXP1 is a multiple of cXP / PW);
    // those of the second row read the second extended row
    const int rw = DUAL && 4 * pg * SW >= cXP1 ? 1 : 0;
    int a_base, a_step;
    if (NP == 4 || NP == 2 || NP == 1) {
      a_base = (kk % cNp) * cQL + kk / cNp + 16 * MT - 1 - col - 16 * MTW * eg;
      a_step = 4 / cNp;
      a_base += rw * (cHN - a_step * (cXP1 / 4));
    } else {
      a_base = 0;
      a_step = 0;
    }
    v4f P1[MTW], P2[MTW], P3[STK ? 1 : MTW];
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt) {
      P1[mt] = vzero();
      P2[mt] = vzero();
      if constexpr (!STK) P3[mt] = vzero();
    }
    for (int s = 0; s < SW; ++s) {
      const int ks = pg * SW + s;
      const float2 b = Bs[b_base + 4 * ks * NCB];
      const float bsum = b.x + b.y;
      const float bst = col < 8 ? b.x : b.y;   // STK: [B_re | B_im]
      int ai;
      if (NP == 4 || NP == 2 || NP == 1) {
        ai = a_base + a_step * ks;
      } else {
        const int i = 4 * ks + kk - rw * cXP1;
        ai = rw * cHN + (i % cNp) * cQL + i / cNp + 16 * MT - 1 - col - 16 * MTW * eg;
      }
#pragma unroll
      for (int mt = 0; mt < MTW; ++mt) {
        const float2 a = Hp[ai - 16 * mt];
        if constexpr (STK) {
          P1[mt] = mfma16(a.x, bst, P1[mt]);
          P2[mt] = mfma16(a.y, bst, P2[mt]);
        } else {
          const float asum = a.x + a.y;
          P1[mt] = mfma16(a.x, b.x, P1[mt]);
          P2[mt] = mfma16(a.y, b.y, P2[mt]);
          P3[mt] = mfma16(asum, bsum, P3[mt]);
        }
      }
    }
    // row done: Re = P1 - P2, Im = P3 - P1 - P2, folded into the item sums
    // (STK: the raw products, combined when the item is written)
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt) {
      if constexpr (STK) {
        SR[mt] += P1[mt];
        SI[mt] += P2[mt];
      } else {
        SR[mt] += P1[mt] - P2[mt];
        SI[mt] += P3[mt] - P1[mt] - P2[mt];
      }
    }

    if (item_done) {
      // item done: sum the point groups through LDS and write the partial
      __syncthreads();
      float* red = reinterpret_cast<float*>(lds);   // [w][mtw][j][lane]
#pragma unroll
      for (int mt = 0; mt < MTW; ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float re = SR[mt][v], im = SI[mt][v];
          if constexpr (STK) {
            // lane c < 8: A_re B_re, A_im B_re of column c; lane c + 8:
            // A_re B_im, A_im B_im.  Re = A_re B_re - A_im B_im,
            // Im = A_re B_im + A_im B_re; lanes 8-15 write zero columns.
            const float re_p = __shfl_xor(re, 8, 64), im_p = __shfl_xor(im, 8, 64);
            const bool lo = col < 8;
            const float r2 = lo ? re - im_p : 0.f, i2 = lo ? re_p + im : 0.f;
            re = r2;
            im = i2;
          }
          red[((w * MTW + mt) * 8 + v) * 64 + lane] = re;
          red[((w * MTW + mt) * 8 + 4 + v) * 64 + lane] = im;
        }
        SR[mt] = vzero();
        SI[mt] = vzero();
      }
      __syncthreads();
      const int li = item_cur - ib_cur;
      float* out = reinterpret_cast<float*>(vpart_cur + (size_t)li * Ne_pad * NC);   // [chunk][zb] = li
      for (int idx = tid; idx < MT * 512; idx += 256) {
        const int mt = idx >> 9, e_g = mt / MTW, mtw = mt - e_g * MTW, rest = idx & 511;
        float sum = 0.f;
#pragma unroll
        for (int p_g = 0; p_g < PW; ++p_g) sum += red[((p_g * EW + e_g) * MTW + mtw) * 512 + rest];
        const int ln = idx & 63, j = (idx >> 6) & 7;
        const int e = 16 * mt + 4 * (ln >> 4) + (j & 3);
        out[((size_t)e * NC + (ln & 15)) * 2 + (j >> 2)] = sum;
      }
      ++it;
    }
    if (!has_next) break;
  }
#undef BORN_FWD_PREFETCH
#undef BORN_FWD_DESC
#undef BORN_FWD_INST
}

// LDS of k_fwd_s2d: the step's polyphase rows and B tile (B rows of 8
// columns in the stacked form), or the item's reduction buffer
// [wave][element tile of the wave][8][64] floats.
static size_t fwd_lds_bytes(int MT, int S, int Np, bool dual = false, bool stk = false) {
  const int XP = 16 * S;
  const int QL = (XP - 1) / Np + 16 * MT;
  const int ncb = stk ? NC / 2 + 1 : NCS;
  const size_t body = (dual ? 2 : 1) * ((size_t)Np * QL + (size_t)XP * ncb) * sizeof(float2);
  const int mtw = MT >= 2 ? MT / 2 : 1;
  const size_t red = (size_t)4 * mtw * 8 * 64 * sizeof(float);
  return std::max(body, red);
}

bool fwd_structured_fits(int Nx, int Ne_pad, int Np) {
  const int S = (Nx + 15) / 16;
  const int XP = 16 * S;
  const int QL = (XP - 1) / Np + Ne_pad;
  return XP <= 1024 && Np * QL <= 256 * HPT_MAX && fwd_lds_bytes(Ne_pad / 16, S, Np) <= 160 * 1024 &&
         Ne_pad <= 128;
}

int fwd_ppt(int S) {
  const int XP = 16 * S;
  return XP <= 256 ? 1 : (XP <= 512 ? 2 : 4);
}

// Two rows per step when both fit the one-point-per-thread B fill.
static bool fwd_dual(int S) { return 16 * S <= 128; }

void launch_fwd_structured(const LaunchSet& L, int n_cu, hipStream_t s) {
  if (!L.inst.empty() && L.inst[0].sgrp_p > 1) {   // 3-D short rows: several y slices per item
    launch_fwd_s3d(L, n_cu, s);
    return;
  }
  const int PPT = fwd_ppt(L.S_max);
  const bool dual = fwd_dual(L.S_max);
  // the stacked kernels serve 4 and 2 points per pitch (as the launch below)
  const bool stk = L.stk && (L.Np == 4 || L.Np == 2);
  const size_t lds = fwd_lds_bytes(L.MT, L.S_max, L.Np, dual, stk);
  // the stacked kernels fit 3 workgroups per CU (registers and LDS), the others 2
  const int gm = L.grid_mult > 0 ? L.grid_mult : (stk ? 3 : 2);
  const int grid = std::max(1, std::min(L.n_items, gm * n_cu));
  const int Ne_pad = 16 * L.MT;
#define BORN_FWD_LAUNCH2(MTv, NPv, PPTv, Dv, Sv)                                                              \
  hipLaunchKernelGGL((k_fwd_s2d<MTv, NPv, PPTv, Dv, Sv>), dim3(grid), dim3(256), lds, s, L.d_inst, (int)L.inst.size(), \
                     L.n_items, Ne_pad, L.d_chunks, L.d_chunk_pairs)
#define BORN_FWD_LAUNCH(MTv, NPv, PPTv, Dv)                                \
  do {                                                                     \
    if (NPv != 0 && L.stk) BORN_FWD_LAUNCH2(MTv, NPv, PPTv, Dv, (NPv != 0)); \
    else BORN_FWD_LAUNCH2(MTv, NPv, PPTv, Dv, false);                      \
  } while (0)
#define BORN_FWD_PPT(MTv, NPv)                           \
  if (dual) BORN_FWD_LAUNCH(MTv, NPv, 1, true);          \
  else if (PPT <= 1) BORN_FWD_LAUNCH(MTv, NPv, 1, false); \
  else if (PPT <= 2) BORN_FWD_LAUNCH(MTv, NPv, 2, false); \
  else BORN_FWD_LAUNCH(MTv, NPv, 4, false);
#define BORN_FWD_NP(MTv)       \
  if (L.Np == 4) {             \
    BORN_FWD_PPT(MTv, 4)       \
  } else if (L.Np == 2) {      \
    BORN_FWD_PPT(MTv, 2)       \
  } else {                     \
    BORN_FWD_PPT(MTv, 0)       \
  }
  switch (L.MT) {
    case 1: BORN_FWD_NP(1) break;
    case 2: BORN_FWD_NP(2) break;
    case 4: BORN_FWD_NP(4) break;
    default: BORN_FWD_NP(8) break;
  }
#undef BORN_FWD_NP
#undef BORN_FWD_PPT
#undef BORN_FWD_LAUNCH
#undef BORN_FWD_LAUNCH2
}

// ------------------------------------------------------------------ adjoint
// One workgroup = (instance, lateral segment of 64*MW points, depth row z);
// the workgroups of instance k are [wg_begin_k, wg_begin_{k+1}).  It walks
// every chunk (rhs-major, then frequency) in a fixed order and keeps its
// output in registers: no atomics, bitwise repeatable.  Per chunk
//   Y[i][c] = sum_e conj(Hext[i - e*Np + off]) W[e][c]       (MFMA, A = points x elements)
//   gamma_hat[i] += sum_c conj(p_inc[pair_c][i]) Y[i][c]     (registers)
// STK (chunks of at most 8 columns): the B operand holds W_re of the 8
// columns in lanes 0-7 and W_im in lanes 8-15, two MFMAs per tile and
// K-step (A_re B, A_im B) instead of three over 16 columns; lane c < 8 then
// holds A_re W_re, A_im W_re of column c and lane c + 8 A_re W_im, A_im W_im,
// and each lane's share of conj(p) Y needs only its own products (the sum
// over the column lanes adds the two halves).  Only 8 p_inc rows are loaded.
template <int MT, int MW, int NP, bool STK>
__global__ __launch_bounds__(256, 2) void k_adj_s2d(const Inst* __restrict__ inst, int n_inst, OutPtrs outs,
                                                    const Chunk* __restrict__ chunks,
                                                    const int* __restrict__ chunk_pairs) {
  extern __shared__ float2 lds[];
  constexpr int SEG = 64 * MW;
  constexpr int PS = SEG + 4;                 // padded stride of the p_inc tile
  constexpr int NCP = STK ? NC / 2 : NC;      // p_inc rows loaded
  constexpr int PPT = NCP * SEG / 256;        // p_inc tile entries per thread
  constexpr int NEP = 16 * MT;                // padded element count
  constexpr int HPT = NP ? (SEG + (NEP - 1) * NP + 255) / 256 : HPT_MAX;
  constexpr int WPT = MT;                     // W tile entries per thread
  int k = 0;
  while (k + 1 < n_inst && (int)blockIdx.x >= inst[k + 1].wg_begin) ++k;
  const Inst Iv = inst[k];    // loop-invariant: held in scalar registers
  const Inst* I = &Iv;
  float2* __restrict__ g_out = outs.p[k];
  const int Np = NP ? NP : I->Np;
  const int L = SEG + (NEP - 1) * Np;
  float2* He = lds;                           // [L]
  float2* Ws = He + L;                        // [NEP][NCS]
  float2* Ps = Ws + NEP * NCS;                // [NCP][PS]

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int kk = lane >> 4, col = lane & 15;
  const int ccol = STK ? (col & 7) : col;     // column whose W and p_inc the lane reads
  const int lw = (int)blockIdx.x - I->wg_begin;
  const int z = lw / I->adj_segs;
  const int i_s = (lw - z * I->adj_segs) * SEG;
  const int ib = i_s + w * 16 * MW;
  const int t_shift = i_s - (NEP - I->N_e) * Np;

  // row-invariant gather indices of the extended-row window
  int eidx[HPT];
#pragma unroll
  for (int j = 0; j < HPT; ++j) {
    const int u = tid + 256 * j;
    const int t = u + t_shift;
    eidx[j] = (u < L && t >= 0 && t < I->N_ext) ? I->ext[t] : -1;
  }

  float2 hreg[HPT], wreg[WPT], preg[PPT];
  // A step is (chunk, y slice ey); 2-D has one slice.  Loads of a step: the
  // window of the table row of (ey, this row) at frequency FG and the weights
  // of the slice; the p_inc rows of the chunk's columns (PV) only with its
  // first slice.  All addresses are valid: window entries outside the table
  // are zeroed when written to LDS, unused columns meet zero weights, points
  // past the row end are never written out.
#define BORN_ADJ_PREFETCH_HW(cidx, EY, FG)                                                       \
  {                                                                                              \
    const float2* hrow_ = I->h + (size_t)(FG) * I->Hp + (size_t)I->hrow[(EY) * I->n_rows + z] * I->HR; \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) hreg[j] = hrow_[max(eidx[j], 0)];           \
    const float2* wblk_ = I->Wc + ((size_t)(cidx) * I->Ney + (EY)) * NEP * NC;                   \
    _Pragma("unroll") for (int j = 0; j < WPT; ++j) wreg[j] = wblk_[tid + 256 * j];             \
  }
#define BORN_ADJ_PREFETCH_P(PV)                                                                  \
  {                                                                                              \
    _Pragma("unroll") for (int j = 0; j < PPT; ++j) {                                           \
      const int idx_ = tid + 256 * j;                                                            \
      const int gi_ = min(i_s + idx_ % SEG, I->Nx - 1);                                          \
      preg[j] = I->P[(size_t)(PV)[j] * I->N_pts + (size_t)z * I->Nx + gi_];                      \
    }                                                                                            \
  }
#define BORN_ADJ_DESC(cidx, FG, R, PV)                                                           \
  {                                                                                              \
    const Chunk c_ = chunks[I->chunk_base + (cidx)];                                             \
    FG = c_.fg;                                                                                  \
    R = c_.r;                                                                                    \
    _Pragma("unroll") for (int j = 0; j < PPT; ++j)                                             \
      (PV)[j] = chunk_pairs[(I->chunk_base + (cidx)) * NC + (tid + 256 * j) / SEG];              \
  }

  const int Ney = I->Ney;
  int fg0, r0, pv0[PPT];
  BORN_ADJ_DESC(0, fg0, r0, pv0)
  BORN_ADJ_PREFETCH_HW(0, 0, fg0)
  BORN_ADJ_PREFETCH_P(pv0)
  int pf_fg = fg0;   // frequency of the chunk of the prefetched step
  int nfg = fg0, nr = r0, npv[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) npv[j] = pv0[j];
  if (I->n_chunks > 1) BORN_ADJ_DESC(1, nfg, nr, npv)
  int cur_r = r0;

  float gx[MW][4], gy[MW][4];
#pragma unroll
  for (int mw = 0; mw < MW; ++mw)
#pragma unroll
    for (int v = 0; v < 4; ++v) gx[mw][v] = gy[mw][v] = 0.f;

  const int u_base = ib - i_s + col + (NEP - 1) * Np;
  v4f P1[MW], P2[MW], P3[STK ? 1 : MW];
  const int n_steps = I->n_chunks * Ney;

  for (int st = 0; st < n_steps; ++st) {
    const int ci = st / Ney, ey = st - ci * Ney;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const int u = tid + 256 * j;
      if (u < L) He[u] = eidx[j] >= 0 ? hreg[j] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < WPT; ++j) Ws[((tid >> 4) + 16 * j) * NCS + (tid & 15)] = wreg[j];
    if (ey == 0) {
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int idx = tid + 256 * j;
        const int pc = idx / SEG, pi = idx - pc * SEG;
        Ps[pc * PS + pi] = preg[j];
      }
    }
    __syncthreads();
    const int r = cur_r;
    const bool has_next = st + 1 < n_steps;
    const bool last_slice = ey == Ney - 1;
    const int next_r = last_slice ? nr : cur_r;
    if (has_next) {
      if (!last_slice) {
        BORN_ADJ_PREFETCH_HW(ci, ey + 1, pf_fg)
      } else {
        BORN_ADJ_PREFETCH_HW(ci + 1, 0, nfg)
        BORN_ADJ_PREFETCH_P(npv)
        pf_fg = nfg;
        if (ci + 2 < I->n_chunks) BORN_ADJ_DESC(ci + 2, nfg, nr, npv)
      }
    }

    if (ey == 0) {
#pragma unroll
      for (int mw = 0; mw < MW; ++mw) {
        P1[mw] = vzero();
        P2[mw] = vzero();
        if constexpr (!STK) P3[mw] = vzero();
      }
    }
#pragma unroll 4
    for (int k0 = 0; k0 < NEP; k0 += 4) {
      const int k = k0 + kk;
      const float2 b = Ws[k * NCS + ccol];
      const float bsum = b.x + b.y;
      const float bst = col < 8 ? b.x : b.y;   // STK: [W_re | W_im]
      const int u0 = u_base - k * Np;
#pragma unroll
      for (int mw = 0; mw < MW; ++mw) {
        const float2 a = He[u0 + 16 * mw];
        if constexpr (STK) {
          P1[mw] = mfma16(a.x, bst, P1[mw]);
          P2[mw] = mfma16(a.y, bst, P2[mw]);
        } else {
          const float adiff = a.x - a.y;
          P1[mw] = mfma16(a.x, b.x, P1[mw]);
          P2[mw] = mfma16(a.y, b.y, P2[mw]);
          P3[mw] = mfma16(adiff, bsum, P3[mw]);
        }
      }
    }
    if (!last_slice) continue;
    // conj(A) W: Re = P1 + P2, Im = P3 - P1 + P2; then gamma_hat += conj(p) * Y.
    // STK: Re Y_c = P1[c] + P2[c + 8], Im Y_c = P1[c + 8] - P2[c], so lane c
    // contributes as Y = (P1, -P2) and lane c + 8 as Y = (P2, P1).
#pragma unroll
    for (int mw = 0; mw < MW; ++mw) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float yr, yi;
        if constexpr (STK) {
          const bool lo = col < 8;
          yr = lo ? P1[mw][v] : P2[mw][v];
          yi = lo ? -P2[mw][v] : P1[mw][v];
        } else {
          yr = P1[mw][v] + P2[mw][v];
          yi = P3[mw][v] - P1[mw][v] + P2[mw][v];
        }
        const float2 p = Ps[ccol * PS + (ib - i_s) + 16 * mw + 4 * kk + v];
        gx[mw][v] += p.x * yr + p.y * yi;
        gy[mw][v] += p.x * yi - p.y * yr;
      }
    }
    if (!has_next || next_r != r) {
      // sum over the column lanes of one rhs (all 16, or those with the same
      // column mod col_rhs) and write gamma_hat
      const int o0 = I->col_rhs > 0 ? I->col_rhs : 1;
#pragma unroll
      for (int mw = 0; mw < MW; ++mw) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float x = gx[mw][v], y = gy[mw][v];
          for (int o = o0; o < 16; o <<= 1) {
            x += __shfl_xor(x, o, 64);
            y += __shfl_xor(y, o, 64);
          }
          gx[mw][v] = x;
          gy[mw][v] = y;
        }
      }
      if (I->col_rhs > 0) {
        // lane col < n_rhs holds rhs col for every point of its rows
        if (col < I->n_rhs) {
#pragma unroll
          for (int mw = 0; mw < MW; ++mw)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int i = ib + 16 * mw + 4 * kk + v;
              if (i < I->Nx) g_out[(size_t)col * I->N_pts + (size_t)z * I->Nx + i] = make_float2(gx[mw][v], gy[mw][v]);
            }
        }
      } else if (col < 4) {
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
          if (i < I->Nx) g_out[(size_t)r * I->N_pts + (size_t)z * I->Nx + i] = make_float2(x, y);
        }
      }
#pragma unroll
      for (int mw = 0; mw < MW; ++mw)
#pragma unroll
        for (int v = 0; v < 4; ++v) gx[mw][v] = gy[mw][v] = 0.f;
    }
    cur_r = next_r;
  }
#undef BORN_ADJ_PREFETCH_HW
#undef BORN_ADJ_PREFETCH_P
#undef BORN_ADJ_DESC
}


// LDS of k_adj_s2d: window, W tile, p_inc rows (8 in the stacked form).
static size_t adj_lds_bytes(int MW, int Ne_pad, int Np, bool stk = false) {
  const int SEG = 64 * MW;
  const int L = SEG + (Ne_pad - 1) * Np;
  const int ncp = stk ? NC / 2 : NC;
  return ((size_t)L + (size_t)Ne_pad * NCS + (size_t)ncp * (SEG + 4)) * sizeof(float2);
}

bool adj_structured_fits(int Nx, int Ne_pad, int Np) {
  (void)Nx;
  const int L = 64 * 4 + (Ne_pad - 1) * Np;   // both segment widths (MW = 2, 4) must fit
  return L <= 256 * HPT_MAX && Ne_pad <= 16 * WPT_MAX && adj_lds_bytes(4, Ne_pad, Np) <= 160 * 1024;
}

// Adjoint segment of 128 points for rows of at most 128 points and for rows
// that 128-point segments pad less than 256-point ones (384 points: 3 x 128
// against 2 x 256), else 256.
int adj_mw_default(int Nx) {
  const int pad2 = (Nx + 127) / 128 * 128, pad4 = (Nx + 255) / 256 * 256;
  return Nx <= 128 || pad2 < pad4 ? 2 : 4;
}

void launch_adj_structured(const LaunchSet& L, const OutPtrs& outs, hipStream_t s) {
  const int MW = L.adj_mw;
  const bool stk = L.stk && (L.Np == 4 || L.Np == 2);   // as the launch below
  const size_t lds = adj_lds_bytes(MW, 16 * L.MT, L.Np, stk);
  const dim3 grid(L.n_wg);
#define BORN_ADJ_LAUNCH2(MTv, MWv, NPv, Sv)                                                                  \
  hipLaunchKernelGGL((k_adj_s2d<MTv, MWv, NPv, Sv>), grid, dim3(256), lds, s, L.d_inst, (int)L.inst.size(), outs, \
                     L.d_chunks, L.d_chunk_pairs)
#define BORN_ADJ_LAUNCH(MTv, MWv, NPv)                               \
  do {                                                               \
    if (NPv != 0 && L.stk) BORN_ADJ_LAUNCH2(MTv, MWv, NPv, (NPv != 0)); \
    else BORN_ADJ_LAUNCH2(MTv, MWv, NPv, false);                     \
  } while (0)
#define BORN_ADJ_MW(MTv, NPv)                     \
  if (MW == 2) BORN_ADJ_LAUNCH(MTv, 2, NPv);      \
  else BORN_ADJ_LAUNCH(MTv, 4, NPv);
#define BORN_ADJ_NP(MTv)        \
  if (L.Np == 4) {              \
    BORN_ADJ_MW(MTv, 4)         \
  } else if (L.Np == 2) {       \
    BORN_ADJ_MW(MTv, 2)         \
  } else {                      \
    BORN_ADJ_MW(MTv, 0)         \
  }
  switch (L.MT) {
    case 1: BORN_ADJ_NP(1) break;
    case 2: BORN_ADJ_NP(2) break;
    case 4: BORN_ADJ_NP(4) break;
    default: BORN_ADJ_NP(8) break;
  }
#undef BORN_ADJ_NP
#undef BORN_ADJ_MW
#undef BORN_ADJ_LAUNCH
#undef BORN_ADJ_LAUNCH2
}

// Allow dynamic LDS up to the CU's 160 KiB for the structured kernels.
template <int MT>
static void set_max_lds_mt(int b) {
#define BORN_SET(K) hipFuncSetAttribute((const void*)(K), hipFuncAttributeMaxDynamicSharedMemorySize, b)
  BORN_SET((k_fwd_s2d<MT, 4, 1, false, false>)); BORN_SET((k_fwd_s2d<MT, 4, 2, false, false>));
  BORN_SET((k_fwd_s2d<MT, 4, 4, false, false>)); BORN_SET((k_fwd_s2d<MT, 2, 1, false, false>));
  BORN_SET((k_fwd_s2d<MT, 2, 2, false, false>)); BORN_SET((k_fwd_s2d<MT, 2, 4, false, false>));
  BORN_SET((k_fwd_s2d<MT, 0, 1, false, false>)); BORN_SET((k_fwd_s2d<MT, 0, 2, false, false>));
  BORN_SET((k_fwd_s2d<MT, 0, 4, false, false>)); BORN_SET((k_fwd_s2d<MT, 4, 1, true, false>));
  BORN_SET((k_fwd_s2d<MT, 2, 1, true, false>)); BORN_SET((k_fwd_s2d<MT, 0, 1, true, false>));
  BORN_SET((k_fwd_s2d<MT, 4, 1, false, true>)); BORN_SET((k_fwd_s2d<MT, 4, 2, false, true>));
  BORN_SET((k_fwd_s2d<MT, 4, 4, false, true>)); BORN_SET((k_fwd_s2d<MT, 2, 1, false, true>));
  BORN_SET((k_fwd_s2d<MT, 2, 2, false, true>)); BORN_SET((k_fwd_s2d<MT, 2, 4, false, true>));
  BORN_SET((k_fwd_s2d<MT, 4, 1, true, true>)); BORN_SET((k_fwd_s2d<MT, 2, 1, true, true>));
  BORN_SET((k_adj_s2d<MT, 2, 4, false>)); BORN_SET((k_adj_s2d<MT, 4, 4, false>));
  BORN_SET((k_adj_s2d<MT, 2, 2, false>)); BORN_SET((k_adj_s2d<MT, 4, 2, false>));
  BORN_SET((k_adj_s2d<MT, 2, 0, false>)); BORN_SET((k_adj_s2d<MT, 4, 0, false>));
  BORN_SET((k_adj_s2d<MT, 2, 4, true>)); BORN_SET((k_adj_s2d<MT, 4, 4, true>));
  BORN_SET((k_adj_s2d<MT, 2, 2, true>)); BORN_SET((k_adj_s2d<MT, 4, 2, true>));
#undef BORN_SET
}

void set_max_lds(size_t bytes) {
  const int b = (int)bytes;
  set_max_lds_mt<1>(b);
  set_max_lds_mt<2>(b);
  set_max_lds_mt<4>(b);
  set_max_lds_mt<8>(b);
}

}  // namespace born
