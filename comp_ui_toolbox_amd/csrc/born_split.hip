// Split-precision structured contractions (gfx950).
//
// Same contractions as k_fwd_s2d / k_adj_s2d (born_structured.hip), with every
// FP32 operand x written as x_hi + x_lo (hi = x rounded to BF16, lo = the
// rounded remainder) and each real product taken as
//     x*y ~ x_hi*y_hi + x_hi*y_lo + x_lo*y_hi
// on 16x16x32 MFMA with FP32 accumulation.  The dropped x_lo*y_lo term and
// the roundings leave a relative error of about 2^-16 per product.
// The forward takes a complex multiply-add as three real products,
//     P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),
//     Re = P1 - P2, Im = P3 - P1 - P2,
// i.e. 9 MFMAs of K = 32 with 6 hi/lo vectors per operand (its time follows
// the MFMA count).  The adjoint keeps the four-product form (12 MFMAs, 4 hi/lo
// vectors per operand: Re = Ar Br - Ai Bi, Im = Ar Bi + Ai Br): its
// double-buffered operand planes and p_inc rows already fill the LDS, with
// no room for six planes.
#include "born_internal.h"

namespace born {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

namespace {

constexpr int SEG = 256;          // points per step (forward) / per workgroup (adjoint)

// 16x16x32 MFMA on BF16 or FP16 hi/lo operands, FP32 accumulation
template <bool F16>
__device__ __forceinline__ v4f mfma32(uint4 a, uint4 b, v4f c) {
  if constexpr (F16)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}
__device__ __forceinline__ uint4 join(uint2 a, uint2 b) { return make_uint4(a.x, a.y, b.x, b.y); }

__device__ __forceinline__ v4f vz() { return v4f{0.f, 0.f, 0.f, 0.f}; }

}  // namespace

// ======================================================================
// forward
// ======================================================================
// Round two floats to BF16 (nearest even) in one conversion; returns the
// packed pair (x0 in the low half) and the rounded values as floats.
__device__ __forceinline__ unsigned int bf_pack(float x0, float x1, float& r0, float& r1) {
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
  const unsigned int p = __builtin_bit_cast(unsigned int, __builtin_convertvector((v2f{x0, x1}), v2bf));
  r0 = __uint_as_float(p << 16);
  r1 = __uint_as_float(p & 0xffff0000u);
  return p;
}
// Round two floats to FP16 in one conversion; returns the packed pair and
// the rounded values as floats.
__device__ __forceinline__ unsigned int h_pack(float x0, float x1, float& r0, float& r1) {
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef _Float16 v2h __attribute__((ext_vector_type(2)));
  const v2h h = __builtin_convertvector((v2f{x0, x1}), v2h);
  r0 = (float)h[0];
  r1 = (float)h[1];
  return __builtin_bit_cast(unsigned int, h);
}
// hi/lo split of a pair: hi = t(x), lo = t(x - hi) for t = BF16 or FP16,
// each packed (x0 low)
template <bool F16>
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned int& hi, unsigned int& lo) {
  float h0, h1, d0, d1;
  if constexpr (F16) {
    hi = h_pack(x0, x1, h0, h1);
    lo = h_pack(x0 - h0, x1 - h1, d0, d1);
  } else {
    hi = bf_pack(x0, x1, h0, h1);
    lo = bf_pack(x0 - h0, x1 - h1, d0, d1);
  }
}
__device__ __forceinline__ const void* byte_offset(const void* base, unsigned int off) {
  return static_cast<const char*>(base) + off;
}

// Work item = (instance, chunk, y slice ey, row block, 256-point segment xs),
// taken persistently (the items of instance k are [item_begin_s_k,
// item_begin_s_{k+1})); a step is one row (y, z) of the item (rows are depth
// rows z in 2-D, Ney = 1).  The p_inc table of every instance has a zero row
// after its N_pairs rows.  Waves: 2 element groups x 2 point groups; wave
// (eg, pg) holds element tiles [eg*MTW, (eg+1)*MTW) and the K-blocks
// pg*KBW .. pg*KBW + KBW-1 of the segment.  NP = 2 or 4 points per pitch.
//
// Loads of a step depend on nothing loaded in the step before: per item the
// window gather offsets (registers) and the column rows (scalar registers,
// one p_inc row per column; unused columns read the zero row) are fixed, and
// per step only the row base moves.
//
// LDS (u16 units):
//   Aw[comp][copy][LAP]  window of the extended row (comp = re_hi, re_lo,
//                        im_hi, im_lo, sum_hi, sum_lo with sum = re + im);
//                        copy c is copy 0 shifted by NP*c entries (8 / NP
//                        copies) so every lane's 8-entry read is 16-byte aligned
//   Bt[comp][c][SEGP]    p_inc*gamma of the segment in the same 6 planes,
//                        points contiguous
// Copy length of the forward's window planes: at least LA + 12 entries,
// 16-byte multiple, and a stride between copies that puts the 16 lanes of
// one 16-byte read group on distinct banks of gfx950's 64-bank LDS (0 mod 64
// entries for 2 copies, 32 mod 64 for 4; measured: bank-conflict cycles of
// the forward down 3.5x against a stride tuned for 32 banks).
__host__ __device__ constexpr int fwd_split_lap(int LA, int NP) {
  return ((LA + 12 + 63) / 64) * 64 + (NP == 4 ? 0 : 32);
}

// HALF (rows of at most 128 points): the MFMAs cover the segment's first 128
// points only (its further points are padding), the waves splitting those.
template <int MT, bool EVEN, int NP, bool F16, bool HALF>
__global__ __launch_bounds__(256, 2) void k_fwd_s2d_split(const Inst* __restrict__ inst, int n_inst, int n_items,
                                                          int Ne_pad, const Chunk* __restrict__ chunks,
                                                          const int* __restrict__ chunk_pairs) {
  extern __shared__ unsigned short ldsu[];
  constexpr int NCOPY = 8 / NP;         // window copies, shifted by NP entries each
  constexpr int DSTEP = NP / 2;         // K-blocks of 32 points per element tile of 16
  constexpr int SEGP = SEG + 8;
  constexpr int EW = MT >= 2 ? 2 : 1;
  constexpr int PW = 4 / EW;
  constexpr int MTW = MT / EW;
  constexpr int KBW = (SEG / (HALF ? 64 : 32)) / PW;
  constexpr int LA = SEG + NP * (16 * MT - 1);
  constexpr int LAP = fwd_split_lap(LA, NP);
  constexpr int APT = 2 * ((LA + 511) / 512);   // window entries per thread: pairs (u, u + 1), u = 2 tid + 512 j
  unsigned short* Aw = ldsu;                     // [6][NCOPY][LAP]
  unsigned short* Bt = ldsu + 6 * NCOPY * LAP;   // [6][NC][SEGP]
  unsigned int* Bt32 = reinterpret_cast<unsigned int*>(Bt);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int eg = w % EW, pg = w / EW;
  const int kk = lane >> 4, col = lane & 15;
  const int n_my = (n_items - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  if (n_my <= 0) return;
  int ki = 0;
  Inst Iv = inst[0];            // record of the current instance, in scalar registers
  const Inst* I = &Iv;
  const int bc0 = (tid >> 7) * 8;    // columns of this thread's B loads: uniform per wave
  const int bi = 2 * (tid & 127);

  int it = 0, item = blockIdx.x;
  int ci, xs, z, z_end, fg, rr, ey, lid;
  bool hfull, gfull;                 // window inside the table / segment inside the row
  unsigned int hoff[APT];            // window gather byte offsets within a row
  unsigned int hval = 0;             // bit j: window entry j inside the table
  unsigned int goff0, goff1;         // byte offsets of points bi, bi+1 (clamped)
  int prow[8];                       // p_inc rows of the columns bc0 .. bc0+7
#define BORN_SPLIT_ITEM(itm)                                                              \
  {                                                                                       \
    if (ki + 1 < n_inst && (itm) >= inst[ki + 1].item_begin_s) {                          \
      while (ki + 1 < n_inst && (itm) >= inst[ki + 1].item_begin_s) ++ki;                 \
      Iv = inst[ki];                                                                      \
    }                                                                                     \
    const int tpad = (16 * MT - I->N_e) * NP;                                             \
    const int NBLK = I->NZB_s * I->nseg_s;   /* partial-sum blocks per chunk and slice */ \
    /* items run in reverse: the forward reads p_inc from the last frequency down, */      \
    /* where the adjoint before it stopped, so the tail it left in the cache is reused */  \
    lid = I->n_chunks * I->Ney * NBLK - 1 - ((itm) - I->item_begin_s);                    \
    const int li_ = lid; /* ((ci * Ney + ey) * NZB_s + zb) * nseg_s + xs */              \
    ci = li_ / (I->Ney * NBLK);                                                           \
    const int rem1_ = li_ - ci * (I->Ney * NBLK);                                         \
    ey = rem1_ / NBLK;                                                                    \
    const int rem_ = rem1_ - ey * NBLK;                                                   \
    const int zb_ = rem_ / I->nseg_s;                                                     \
    xs = rem_ - zb_ * I->nseg_s;                                                          \
    z = zb_ * I->ZR_s;                                                                    \
    z_end = min(I->n_rows, z + I->ZR_s);                                                  \
    const Chunk c_ = chunks[I->chunk_base + ci];                                          \
    fg = c_.fg;                                                                           \
    rr = c_.r;                                                                            \
    const int t0_ = xs * SEG - tpad;                                                      \
    hfull = t0_ >= 0 && t0_ + LA <= I->N_ext;                                              \
    gfull = xs * SEG + SEG <= I->Nx;                                                       \
    hval = 0;                                                                             \
    _Pragma("unroll") for (int j = 0; j < APT; ++j) {                                    \
      const int u_ = 2 * tid + 512 * (j >> 1) + (j & 1);                                  \
      const int t_ = u_ + t0_;                                                            \
      const bool ok_ = u_ < LA && t_ >= 0 && t_ < I->N_ext;                                \
      hoff[j] = ok_ ? 8u * (unsigned int)I->ext[t_] : 0u;                                  \
      hval |= ok_ ? (1u << j) : 0u;                                                       \
    }                                                                                     \
    const int gi_ = xs * SEG + bi;                                                        \
    goff0 = 8u * (unsigned int)(EVEN ? min(gi_, I->Nx - 2) : min(gi_, I->Nx - 1));          \
    goff1 = 8u * (unsigned int)min(gi_ + 1, I->Nx - 1);                                    \
    _Pragma("unroll") for (int q = 0; q < 8; ++q)                                        \
      prow[q] = __builtin_amdgcn_readfirstlane(chunk_pairs[(I->chunk_base + ci) * NC + bc0 + q]); \
    /* settle the per-item loads here, so the per-step prefetch waits on none */          \
    _Pragma("unroll") for (int j = 0; j < APT; ++j) asm volatile("" ::"v"(hoff[j]));     \
  }

  float2 hreg[APT];
  float2 preg[8][2];
  float2 greg[2];

#define BORN_SPLIT_PREFETCH(zz)                                                                  \
  {                                                                                              \
    const size_t rowo_ = (size_t)(zz) * I->Nx;                                                    \
    const float2* hrow_ = I->h + (size_t)fg * I->Hp + (size_t)I->hrow[ey * I->n_rows + (zz)] * I->HR; \
    _Pragma("unroll") for (int j = 0; j < APT; ++j)                                             \
      hreg[j] = *static_cast<const float2*>(byte_offset(hrow_, hoff[j]));                        \
    const float2* grow_ = I->gm + (size_t)rr * I->N_pts + rowo_;                                   \
    if (EVEN) {                                                                                  \
      const float4 v_ = *static_cast<const float4*>(byte_offset(grow_, goff0));                  \
      greg[0] = make_float2(v_.x, v_.y);                                                         \
      greg[1] = make_float2(v_.z, v_.w);                                                         \
    } else {                                                                                     \
      greg[0] = *static_cast<const float2*>(byte_offset(grow_, goff0));                          \
      greg[1] = *static_cast<const float2*>(byte_offset(grow_, goff1));                          \
    }                                                                                            \
    _Pragma("unroll") for (int q = 0; q < 8; ++q) {                                             \
      const float2* prow_ = I->P + (size_t)prow[q] * I->N_pts + rowo_;                             \
      if (EVEN) {                                                                                \
        const float4 v_ = *static_cast<const float4*>(byte_offset(prow_, goff0));                \
        preg[q][0] = make_float2(v_.x, v_.y);                                                    \
        preg[q][1] = make_float2(v_.z, v_.w);                                                    \
      } else {                                                                                   \
        preg[q][0] = *static_cast<const float2*>(byte_offset(prow_, goff0));                     \
        preg[q][1] = *static_cast<const float2*>(byte_offset(prow_, goff1));                     \
      }                                                                                          \
    }                                                                                            \
  }

  BORN_SPLIT_ITEM(item)
  BORN_SPLIT_PREFETCH(z)

  v4f C1[MTW], C2[MTW], C3[MTW];   // Ar Br, Ai Bi, (Ar + Ai)(Br + Bi)
#pragma unroll
  for (int m = 0; m < MTW; ++m) {
    C1[m] = vz();
    C2[m] = vz();
    C3[m] = vz();
  }

  // lane A position: e = 16*(eg*MTW + m) + col, u = 32*k + 8*kk + NP*(16*MT - 1 - e);
  // u mod 8 is a multiple of NP, fixed per lane: copy (u mod 8) / NP is aligned
  const int e0 = 16 * (eg * MTW) + col;
  const int u00 = 8 * kk + NP * (16 * MT - 1 - e0);
  const int acopy = (u00 & 7) / NP;
  const int abase = acopy * LAP + (u00 - NP * acopy);
  const int bbase = col * SEGP + 8 * kk;

  while (true) {
    __syncthreads();
    // FP16 operands are scaled by powers of 2 into range (see k_fp16_scales)
    const float s_h = F16 ? I->fscale[0] : 1.f, s_b = F16 ? I->fscale[1] : 1.f;
    // window -> hi/lo planes, NCOPY copies; a thread writes the entry pair
    // (u, u + 1) as one 32-bit word per plane (u and the copy shifts are even)
    {
      unsigned int* Aw32 = reinterpret_cast<unsigned int*>(Aw);
#pragma unroll
      for (int j = 0; j < APT; j += 2) {
        const int u = 2 * tid + 512 * (j >> 1);
        if (u < LA) {
          float2 h0 = hreg[j], h1 = hreg[j + 1];
          if (!hfull) {
            if (!((hval >> j) & 1u)) h0 = make_float2(0.f, 0.f);
            if (!((hval >> (j + 1)) & 1u)) h1 = make_float2(0.f, 0.f);
          }
          if (F16) {
            h0 = make_float2(h0.x * s_h, h0.y * s_h);
            h1 = make_float2(h1.x * s_h, h1.y * s_h);
          }
          unsigned int v[6];   // re, im, re + im as hi, lo; entry u in the low half
          split_pair<F16>(h0.x, h1.x, v[0], v[1]);
          split_pair<F16>(h0.y, h1.y, v[2], v[3]);
          split_pair<F16>(h0.x + h0.y, h1.x + h1.y, v[4], v[5]);
#pragma unroll
          for (int cp = 0; cp < 6; ++cp) {
#pragma unroll
            for (int c = 0; c < NCOPY; ++c)
              if (u >= NP * c) Aw32[((cp * NCOPY + c) * LAP + u - NP * c) >> 1] = v[cp];
          }
        }
      }
    }
    // B = p_inc * gamma -> hi/lo planes; a thread writes points bi, bi+1 of 8 columns
    {
      float2 g0 = greg[0], g1 = greg[1];
      if (!gfull) {
        if (xs * SEG + bi >= I->Nx) g0 = make_float2(0.f, 0.f);
        if (xs * SEG + bi + 1 >= I->Nx) g1 = make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float2 p0 = preg[q][0], p1 = preg[q][1];
        float re0 = p0.x * g0.x - p0.y * g0.y, im0 = p0.x * g0.y + p0.y * g0.x;
        float re1 = p1.x * g1.x - p1.y * g1.y, im1 = p1.x * g1.y + p1.y * g1.x;
        if (F16) {
          re0 *= s_b;
          im0 *= s_b;
          re1 *= s_b;
          im1 *= s_b;
        }
        unsigned int rh, rl, ih, il, sh, sl;
        split_pair<F16>(re0, re1, rh, rl);
        split_pair<F16>(im0, im1, ih, il);
        // FP16 range: with every component of p_inc below 2^ep and of gamma
        // below 2^ex, the scale 2^(14 - ep - ex) keeps |p_inc gamma| below
        // 2^15, so |re + im| <= sqrt(2) |p_inc gamma| < 46341 < 65504
        split_pair<F16>(re0 + im0, re1 + im1, sh, sl);
        const int o = ((bc0 + q) * SEGP + bi) >> 1;
        Bt32[0 * (NC * SEGP / 2) + o] = rh;
        Bt32[1 * (NC * SEGP / 2) + o] = rl;
        Bt32[2 * (NC * SEGP / 2) + o] = ih;
        Bt32[3 * (NC * SEGP / 2) + o] = il;
        Bt32[4 * (NC * SEGP / 2) + o] = sh;
        Bt32[5 * (NC * SEGP / 2) + o] = sl;
      }
    }
    __syncthreads();

    // next step: the next row of this item, else the first row of the next item
    const int lid_done = lid;
    float2* const vpart_done = Iv.vpart;
    const float unscale_done = F16 ? Iv.fscale[2] : 1.f;
    const bool item_done = z + 1 >= z_end;
    bool has_next = true;
    if (!item_done) {
      ++z;
    } else if (it + 1 >= n_my) {
      has_next = false;
    } else {
      item = blockIdx.x + (it + 1) * gridDim.x;
      BORN_SPLIT_ITEM(item)
    }
    if (has_next) BORN_SPLIT_PREFETCH(z)

    // Tile m at K-block kb reads the window at 32*(pg*KBW + kb) - 16*NP*m:
    // the same data as tile 0 at K-block kb - DSTEP*m.  Walk the diagonals
    // d = kb - DSTEP*m, holding the wave's B blocks in registers and
    // streaming one A set per diagonal (double-buffered).
    uint4 B6[KBW][6];
#pragma unroll
    for (int kb = 0; kb < KBW; ++kb) {
      const int k = pg * KBW + kb;
#pragma unroll
      for (int cp = 0; cp < 6; ++cp)
        B6[kb][cp] = *reinterpret_cast<const uint4*>(&Bt[cp * NC * SEGP + bbase + 32 * k]);
    }
    constexpr int D0 = -DSTEP * (MTW - 1);
    uint4 Acur[6], Anext[6];
    {
      const int ao = abase + 32 * (pg * KBW + D0);
#pragma unroll
      for (int cp = 0; cp < 6; ++cp) Acur[cp] = *reinterpret_cast<const uint4*>(&Aw[cp * NCOPY * LAP + ao]);
    }
#pragma unroll
    for (int d = D0; d < KBW; ++d) {
      if (d + 1 < KBW) {
        const int ao = abase + 32 * (pg * KBW + d + 1);
#pragma unroll
        for (int cp = 0; cp < 6; ++cp) Anext[cp] = *reinterpret_cast<const uint4*>(&Aw[cp * NCOPY * LAP + ao]);
      }
#pragma unroll
      for (int m = 0; m < MTW; ++m) {
        const int kb = d + DSTEP * m;
        if (kb >= 0 && kb < KBW) {
          C1[m] = mfma32<F16>(Acur[0], B6[kb][0], C1[m]);
          C1[m] = mfma32<F16>(Acur[0], B6[kb][1], C1[m]);
          C1[m] = mfma32<F16>(Acur[1], B6[kb][0], C1[m]);
          C2[m] = mfma32<F16>(Acur[2], B6[kb][2], C2[m]);
          C2[m] = mfma32<F16>(Acur[2], B6[kb][3], C2[m]);
          C2[m] = mfma32<F16>(Acur[3], B6[kb][2], C2[m]);
          C3[m] = mfma32<F16>(Acur[4], B6[kb][4], C3[m]);
          C3[m] = mfma32<F16>(Acur[4], B6[kb][5], C3[m]);
          C3[m] = mfma32<F16>(Acur[5], B6[kb][4], C3[m]);
        }
      }
      if (d + 1 < KBW) {
#pragma unroll
        for (int cp = 0; cp < 6; ++cp) Acur[cp] = Anext[cp];
      }
    }

    if (item_done) {
      __syncthreads();
      float* red = reinterpret_cast<float*>(ldsu);
#pragma unroll
      for (int m = 0; m < MTW; ++m) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          red[((w * MTW + m) * 8 + v) * 64 + lane] = C1[m][v] - C2[m][v];
          red[((w * MTW + m) * 8 + 4 + v) * 64 + lane] = C3[m][v] - C1[m][v] - C2[m][v];
        }
        C1[m] = vz();
        C2[m] = vz();
        C3[m] = vz();
      }
      __syncthreads();
      // vpart block of the item: [chunk][zb * nseg + xs] = local item index
      float* out = reinterpret_cast<float*>(vpart_done + (size_t)lid_done * Ne_pad * NC);
      for (int idx = tid; idx < MT * 512; idx += 256) {
        const int mt = idx >> 9, e_g = mt / MTW, mtw = mt - e_g * MTW, rest = idx & 511;
        float sum = 0.f;
#pragma unroll
        for (int p_g = 0; p_g < PW; ++p_g) sum += red[((p_g * EW + e_g) * MTW + mtw) * 512 + rest];
        const int ln = idx & 63, j = (idx >> 6) & 7;
        const int e = 16 * mt + 4 * (ln >> 4) + (j & 3);
        out[((size_t)e * NC + (ln & 15)) * 2 + (j >> 2)] = F16 ? sum * unscale_done : sum;
      }
      ++it;
    }
    if (!has_next) break;
  }
#undef BORN_SPLIT_PREFETCH
#undef BORN_SPLIT_ITEM
}

template <int MT, int NP>
static size_t fwd_split_lds_bytes() {
  constexpr int SEGP = SEG + 8;
  constexpr int LA = SEG + NP * (16 * MT - 1);
  constexpr int LAP = fwd_split_lap(LA, NP);
  const size_t body = ((size_t)6 * (8 / NP) * LAP + (size_t)6 * NC * SEGP) * sizeof(unsigned short);
  const size_t red = (size_t)4 * (MT >= 2 ? MT / 2 : MT) * 8 * 64 * sizeof(float);
  return std::max(body, red);
}

// ======================================================================
// adjoint
// ======================================================================
// One workgroup = (instance, 256-point segment, row (y, z)); the workgroups
// of instance k are [wg_begin_s_k, wg_begin_s_{k+1}).  Steps (chunk, y slice
// ey) are walked in a fixed order with gamma_hat kept in registers (no
// atomics); in 2-D a step is a chunk.
// The segment's points split into NP phases ph = i mod NP of SPH = 256 / NP
// points s; for one phase and slice the adjoint is the Toeplitz product
//     Y[s][c] = sum_e conj(H(NP s + ph - NP e + off)) W[e][c],
// summed over the slices of a chunk before gamma_hat += conj(p_inc) Y.
// With the phase row stored reversed, Hr[m] with m = SPH - s + e, tile t
// (points s = 16 t + r, r = 0..15) at K-block kb (elements 32 kb + 8 kk + j)
// reads Hr[SPH + 16 (2 kb - t) - r + 8 kk + j]: tile t + 2 at K-block kb + 1
// reads the same data as tile t at kb.  A compute wave owns 4 tiles of one
// phase and walks the diagonals delta = 2 kb - t, loading one A set per
// diagonal (double-buffered) with every K-block of W in registers.
// The reversed row is stored in 4 copies, copy d holding Hr[m] at position
// m - d + PADL, so lane r reads its 8 entries 8-byte aligned from copy
// d = (-r) mod 4; the copies start at offsets chosen so that the 16 lanes of
// a read cover all 32 banks once.
//
// The workgroup has 8 waves: waves 0-3 compute on one LDS buffer while
// waves 4-7 load the next step and write it, split into BF16 hi/lo planes,
// to the other buffer; one barrier per step swaps the buffers.  p_inc of a
// chunk has its own pair of buffers (by chunk parity), written with the
// chunk's first step.
//
// LDS: AW[2] = {Ar[comp][ph][CPB] (u16, 4 copies per block), Wt[comp][c][WS] (u16)}, Ps[2][c][PSS] (float2)
template <int MT, int NP>
struct AdjSplitLayout {
  static constexpr int NEP = 16 * MT;
  static constexpr int SPH = SEG / NP;                          // points per phase
  static constexpr int LQ = SPH + NEP;                          // reversed phase-row entries
  static constexpr int PADL = 4;                                // lead pad of every copy
  static constexpr int LR = ((LQ + PADL - 4 + 63) / 64) * 64 + 4;   // copy stride, 4 mod 64
  static constexpr int CPB = ((4 * LR + 40 - 16 + 63) / 64) * 64 + 16;  // (comp, phase) block, 16 mod 64
  static constexpr int WS = NEP + 8;                            // padded W row
  static constexpr int PSS = SEG + 1;                           // padded p_inc row (float2)
  static constexpr int AR_U16 = 4 * NP * CPB;
  static constexpr int AW_U16 = ((AR_U16 + 4 * NC * WS + 7) / 8) * 8;   // one A/W buffer
  static constexpr int PS_U16 = ((NC * PSS * 4 + 7) / 8) * 8;           // one p_inc buffer
  static constexpr int PS_OFF = 2 * AW_U16;
  static constexpr size_t bytes = (size_t)(2 * AW_U16 + 2 * PS_U16) * 2;
};
__device__ __forceinline__ int copy_offset(int d) { return d == 1 ? 16 : (d == 2 ? 28 : (d == 3 ? 40 : 0)); }

template <int MT, int NP, bool Y3, bool F16, bool HALF>
__global__ __launch_bounds__(512, 1) void k_adj_s2d_split(const Inst* __restrict__ inst, int n_inst, OutPtrs outs,
                                                          const Chunk* __restrict__ chunks,
                                                          const int* __restrict__ chunk_pairs) {
  extern __shared__ unsigned short ldsu[];
  typedef AdjSplitLayout<MT, NP> Lay;
  constexpr int NEP = Lay::NEP, SPH = Lay::SPH, LQ = Lay::LQ, PADL = Lay::PADL;
  constexpr int LR = Lay::LR, CPB = Lay::CPB, WS = Lay::WS, PSS = Lay::PSS;
  constexpr int KB = MT / 2;                    // K-blocks of 32 elements
  // extended-row window entries the compute waves read: with HALF only the
  // segment's first 128 points have tiles, which read reversed positions
  // m >= SPH - 64 + 1, i.e. window entries u < 128 + NP (NEP - 1)
  constexpr int L = (HALF ? SEG / 2 : SEG) + NP * (NEP - 1);
  // window fill: thread (p, r) = (ft mod NP, ft / NP) holds the entries
  // u = NP q + p, q = 4 r + i (i = 0..3): four consecutive reversed positions
  // of one phase row
  constexpr int HPT = 4;
  constexpr int WG4 = ((L + NP - 1) / NP + 3) / 4;   // 4-entry groups per phase
  static_assert(NP * WG4 <= 256, "window groups exceed the fill threads");
  static_assert(LQ % 4 == 0 && LR % 4 == 0 && CPB % 4 == 0, "copy alignment classes assume multiples of 4");
  constexpr int WPT = NEP * NC / 256;           // W entries per fill thread
  constexpr int PPT = NC * SEG / 256;           // p_inc entries per fill thread
  constexpr int WPP = 4 / NP;                   // compute waves per phase

  const int tid = threadIdx.x, lane = tid & 63, w = (tid >> 6) & 3;
  const bool filler = tid >= 256;
  const int ft = tid & 255;                     // thread index within its role
  const int kk = lane >> 4, col = lane & 15;
  int ki = 0;
  while (ki + 1 < n_inst && (int)blockIdx.x >= inst[ki + 1].wg_begin_s) ++ki;
  const Inst Iv = inst[ki];   // loop-invariant: held in scalar registers
  const Inst* I = &Iv;
  const int lw = (int)blockIdx.x - I->wg_begin_s;
  const int z = lw / I->adj_segs_s;             // row (y, z)
  const int i_s = (lw - z * I->adj_segs_s) * SEG;
  const int Ney = Y3 ? I->Ney : 1;
  const int n_steps = I->n_chunks * Ney;

  if (filler) {
    // ---------------------------------------------------------------- fill waves
    const int t_shift = i_s - (NEP - I->N_e) * NP;  // window entry u holds Hext[u + t_shift]
    // FP16 operands are scaled by powers of 2 into range (see k_fp16_scales)
    const float s_h = F16 ? I->fscale[0] : 1.f, s_w = F16 ? I->fscale[3] : 1.f;
    int eidx[HPT];
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const int u = NP * (4 * (ft / NP) + j) + ft % NP;
      const int t = u + t_shift;
      eidx[j] = (ft < NP * WG4 && u < L && t >= 0 && t < I->N_ext) ? I->ext[t] : -1;
    }
    // two register sets: the loads of step s + 2 are issued when step s has
    // been written to LDS, so each load has two step periods to arrive
    float2 hA[HPT], wA[WPT], pA[PPT], hB[HPT], wB[WPT], pB[PPT];
    const int gi = min(i_s + ft, I->Nx - 1);
    // Loads of step `st` = (chunk, slice), all from valid addresses: window
    // entries outside the table are zeroed when written to LDS; p_inc rows of
    // unused columns read the zero row; points past the row end are never
    // written out.  p_inc is loaded with a chunk's first slice.
#define BORN_ADJS_PREFETCH(ST, H, WV, PV)                                                        \
    {                                                                                            \
      const int st_ = (ST);                                                                      \
      const int c_ = Y3 ? st_ / Ney : st_, ey_ = Y3 ? st_ - c_ * Ney : 0;                        \
      const Chunk ch_ = chunks[I->chunk_base + c_];                                              \
      const int row_ = Y3 ? I->hrow[ey_ * I->n_rows + z] : z;                                    \
      const float2* hrow_ = I->h + (size_t)ch_.fg * I->Hp + (size_t)row_ * I->HR;                \
      _Pragma("unroll") for (int j = 0; j < HPT; ++j) H[j] = hrow_[max(eidx[j], 0)];             \
      const float2* wblk_ = I->Wc + (size_t)st_ * NEP * NC;                                      \
      _Pragma("unroll") for (int j = 0; j < WPT; ++j)                                           \
        WV[j] = wblk_[(2 * ((ft >> 4) + 16 * (j >> 1)) + (j & 1)) * NC + (ft & 15)];             \
      if (!Y3 || ey_ == 0) {                                                                     \
        const int* prs_ = chunk_pairs + (size_t)(I->chunk_base + c_) * NC;                       \
        _Pragma("unroll") for (int j = 0; j < PPT; ++j)                                          \
          PV[j] = I->P[(size_t)prs_[j] * I->N_pts + (size_t)z * I->Nx + gi];                     \
      }                                                                                          \
    }
#define BORN_ADJS_FILL(ST, H, WV, PV)                                                            \
    {                                                                                            \
      const int st_ = (ST);                                                                      \
      const int c_ = Y3 ? st_ / Ney : st_, ey_ = Y3 ? st_ - c_ * Ney : 0;                        \
      unsigned short* Ar = ldsu + (st_ & 1) * Lay::AW_U16;                                       \
      unsigned short* Wt = Ar + Lay::AR_U16;                                                     \
      /* window entry u -> phase p = u mod NP, reversed index m = LQ - 1 - u / NP;  */            \
      /* copy d holds m at m - d + PADL (>= 1).  The thread's entries q = 4 r + 3 .. 4 r */       \
      /* sit at positions P .. P + 3 of every copy; P mod 4 = -d mod 4, so copy 0 takes */       \
      /* one 8-byte store, copy 2 two 4-byte stores, copies 1 and 3 2 + 4 + 2 bytes.    */       \
      /* Entries outside the window (eidx < 0) are written as zero.                     */       \
      if (ft < NP * WG4) {                                                                       \
        const int p = ft % NP, r = ft / NP;                                                      \
        float xr[4], xi[4];   /* position order: index i holds q = 4 r + 3 - i */                \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                          \
          float2 hv = H[3 - i];                                                                  \
          if (eidx[3 - i] < 0) hv = make_float2(0.f, 0.f);                                       \
          if (F16) hv = make_float2(hv.x * s_h, hv.y * s_h);                                     \
          xr[i] = hv.x;                                                                          \
          xi[i] = -hv.y; /* conj */                                                              \
        }                                                                                        \
        unsigned int pr[4][2]; /* [comp][half]: half 0 = positions P, P+1; 1 = P+2, P+3 */       \
        split_pair<F16>(xr[0], xr[1], pr[0][0], pr[1][0]);                                       \
        split_pair<F16>(xr[2], xr[3], pr[0][1], pr[1][1]);                                       \
        split_pair<F16>(xi[0], xi[1], pr[2][0], pr[3][0]);                                       \
        split_pair<F16>(xi[2], xi[3], pr[2][1], pr[3][1]);                                       \
        _Pragma("unroll") for (int cp = 0; cp < 4; ++cp) {                                       \
          unsigned short* P0 = Ar + (cp * NP + p) * CPB + PADL + (LQ - 4 - 4 * r);               \
          const unsigned int A_ = pr[cp][0], B_ = pr[cp][1];                                     \
          const unsigned int mid_ = (A_ >> 16) | (B_ << 16);                                     \
          *reinterpret_cast<uint2*>(P0) = make_uint2(A_, B_);                                    \
          unsigned short* P1 = P0 + LR + 16 - 1;                                                 \
          P1[0] = (unsigned short)A_;                                                            \
          *reinterpret_cast<unsigned int*>(P1 + 1) = mid_;                                       \
          P1[3] = (unsigned short)(B_ >> 16);                                                    \
          unsigned short* P2 = P0 + 2 * LR + 28 - 2;                                             \
          *reinterpret_cast<unsigned int*>(P2) = A_;                                             \
          *reinterpret_cast<unsigned int*>(P2 + 2) = B_;                                         \
          unsigned short* P3 = P0 + 3 * LR + 40 - 3;                                             \
          P3[0] = (unsigned short)A_;                                                            \
          *reinterpret_cast<unsigned int*>(P3 + 1) = mid_;                                       \
          P3[3] = (unsigned short)(B_ >> 16);                                                    \
        }                                                                                        \
      }                                                                                          \
      /* weights: entries j and j + 1 of a thread are elements k and k + 1 of column */          \
      /* ft mod 16, packed into one 32-bit word per plane */                                     \
      _Pragma("unroll") for (int j = 0; j < WPT; j += 2) {                                       \
        const int k = 2 * ((ft >> 4) + 16 * (j >> 1)), wc = ft & 15;                             \
        unsigned int rh, rl, ih, il;                                                             \
        const float wsc_ = F16 ? s_w : 1.f;                                                      \
        split_pair<F16>(WV[j].x * wsc_, WV[j + 1].x * wsc_, rh, rl);                             \
        split_pair<F16>(WV[j].y * wsc_, WV[j + 1].y * wsc_, ih, il);                             \
        unsigned int* wd = reinterpret_cast<unsigned int*>(&Wt[wc * WS + k]);                    \
        wd[0 * NC * WS / 2] = rh;                                                                \
        wd[1 * NC * WS / 2] = rl;                                                                \
        wd[2 * NC * WS / 2] = ih;                                                                \
        wd[3 * NC * WS / 2] = il;                                                                \
      }                                                                                          \
      if (!Y3 || ey_ == 0) {                                                                     \
        float2* Ps = reinterpret_cast<float2*>(ldsu + Lay::PS_OFF + (c_ & 1) * Lay::PS_U16);     \
        _Pragma("unroll") for (int j = 0; j < PPT; ++j) Ps[j * PSS + ft] = PV[j]; /* column j, point ft */ \
      }                                                                                          \
    }
    // every step issues its loads (clamped to the last step), so in 2-D the
    // loads in flight are the same on every path and each fill waits only
    // for its own
    const int last = n_steps - 1;
    BORN_ADJS_PREFETCH(0, hA, wA, pA)
    BORN_ADJS_FILL(0, hA, wA, pA)
    BORN_ADJS_PREFETCH(min(1, last), hB, wB, pB)
    BORN_ADJS_PREFETCH(min(2, last), hA, wA, pA)
    __syncthreads();
    for (int st = 0; st < n_steps; st += 2) {
      // step st: step st + 1 (set B) -> its buffers; set B <- step st + 3
      if (st + 1 < n_steps) BORN_ADJS_FILL(st + 1, hB, wB, pB)
      BORN_ADJS_PREFETCH(min(st + 3, last), hB, wB, pB)
      __syncthreads();
      if (st + 1 >= n_steps) break;
      // step st + 1: step st + 2 (set A) -> its buffers; set A <- step st + 4
      if (st + 2 < n_steps) BORN_ADJS_FILL(st + 2, hA, wA, pA)
      BORN_ADJS_PREFETCH(min(st + 4, last), hA, wA, pA)
      __syncthreads();
    }
#undef BORN_ADJS_PREFETCH
#undef BORN_ADJS_FILL
    return;
  }

  // ---------------------------------------------------------------- compute waves
  float2* __restrict__ g_out = outs.p[ki];
  // tiles of this wave: 4, or 2 with HALF (rows of at most 128 points fill
  // only the first half of the segment, so the waves split that half)
  constexpr int TPW = HALF ? 2 : 4;
  const int ph = w / WPP, t0 = TPW * (w % WPP);   // phase and first tile of this wave
  // gamma_hat partials: tile t0 + t, v (0..3) -> point s = 16 (t0 + t) + 4 kk + v, i = NP s + ph
  float gx[4][4], gy[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) gx[a][v] = gy[a][v] = 0.f;
  // lane's A read position (relative diagonal 0, comp 0): copy d = (-r) mod 4 of phase ph
  const int dcp = (4 - (col & 3)) & 3;
  const int abase = ph * CPB + dcp * LR + copy_offset(dcp) + (SPH - col + 8 * kk - dcp) + PADL - 16 * t0;
  v4f YR[4], YI[4], YQ[4];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    YR[t] = vz();
    YI[t] = vz();
    YQ[t] = vz();
  }
  __syncthreads();
  for (int st = 0; st < n_steps; ++st) {
    const int c = Y3 ? st / Ney : st, ey = Y3 ? st - c * Ney : 0;
    const unsigned short* Ar = ldsu + (st & 1) * Lay::AW_U16;
    const unsigned short* Wt = Ar + Lay::AR_U16;

    // every K-block of W (B operand)
    uint4 W4[KB][4];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int cp = 0; cp < 4; ++cp)
        W4[kb][cp] = *reinterpret_cast<const uint4*>(&Wt[(cp * NC + col) * WS + 32 * kb + 8 * kk]);
    // Y of the 4 tiles, summed over the chunk's slices: conj(H) W with conj
    // folded into the table (imag part stored negated):
    // Re = Ar Wr - Ai' Wi (the Ai' Wi products in YQ, subtracted once per
    // chunk), Im = Ar Wi + Ai' Wr.  In 2-D each step starts a tile's sums
    // from zero with its first product.
    constexpr int DMIN = -(TPW - 1), DMAX = 2 * (KB - 1);
    uint4 Ac[4], An[4];
#define BORN_ADJS_LOAD_A(DST, DL)                                                                \
  _Pragma("unroll") for (int cp = 0; cp < 4; ++cp) {                                            \
    const unsigned short* a_ = Ar + cp * NP * CPB + abase + 16 * (DL);                           \
    DST[cp] = join(*reinterpret_cast<const uint2*>(a_), *reinterpret_cast<const uint2*>(a_ + 4)); \
  }
    BORN_ADJS_LOAD_A(Ac, DMIN)
#pragma unroll
    for (int dl = DMIN; dl <= DMAX; ++dl) {
      if (dl < DMAX) BORN_ADJS_LOAD_A(An, dl + 1)
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const int k2 = dl + t;
        if (k2 >= 0 && (k2 & 1) == 0 && k2 / 2 < KB) {
          const int kb = k2 / 2;
          // 2-D: the tile's first product of the step starts from zero
          const bool z0 = !Y3 && dl == -t;
          YR[t] = mfma32<F16>(Ac[0], W4[kb][0], z0 ? vz() : YR[t]);
          YR[t] = mfma32<F16>(Ac[0], W4[kb][1], YR[t]);
          YR[t] = mfma32<F16>(Ac[1], W4[kb][0], YR[t]);
          YQ[t] = mfma32<F16>(Ac[2], W4[kb][2], z0 ? vz() : YQ[t]);
          YQ[t] = mfma32<F16>(Ac[2], W4[kb][3], YQ[t]);
          YQ[t] = mfma32<F16>(Ac[3], W4[kb][2], YQ[t]);
          YI[t] = mfma32<F16>(Ac[0], W4[kb][2], z0 ? vz() : YI[t]);
          YI[t] = mfma32<F16>(Ac[0], W4[kb][3], YI[t]);
          YI[t] = mfma32<F16>(Ac[1], W4[kb][2], YI[t]);
          YI[t] = mfma32<F16>(Ac[2], W4[kb][0], YI[t]);
          YI[t] = mfma32<F16>(Ac[2], W4[kb][1], YI[t]);
          YI[t] = mfma32<F16>(Ac[3], W4[kb][0], YI[t]);
        }
      }
      if (dl < DMAX) {
#pragma unroll
        for (int cp = 0; cp < 4; ++cp) Ac[cp] = An[cp];
      }
    }
#undef BORN_ADJS_LOAD_A
    if (ey == Ney - 1) {
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        YR[t] -= YQ[t];
        if (Y3) YQ[t] = vz();
      }
      if (F16) {
        const float us = I->fscale[4];
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          YR[t] *= us;
          YI[t] *= us;
        }
      }
      // gamma_hat += conj(p) Y; lane holds rows 4*kk + v of tile t0 + t, column col
      const float2* Ps = reinterpret_cast<const float2*>(ldsu + Lay::PS_OFF + (c & 1) * Lay::PS_U16);
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int s = 16 * (t0 + t) + 4 * kk + v;
          const float2 p = Ps[col * PSS + NP * s + ph];
          gx[t][v] += p.x * YR[t][v] + p.y * YI[t][v];
          gy[t][v] += p.x * YI[t][v] - p.y * YR[t][v];
        }
        if (Y3) {
          YR[t] = vz();
          YI[t] = vz();
        }
      }
      const int r = chunks[I->chunk_base + c].r;
      const bool has_next = c + 1 < I->n_chunks;
      if (!has_next || chunks[I->chunk_base + c + 1].r != r) {
        // sum over the column lanes of one rhs (all 16, or those with the
        // same column mod col_rhs)
        const int o0 = I->col_rhs > 0 ? I->col_rhs : 1;
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            float x = gx[t][v], y = gy[t][v];
            for (int o = o0; o < 16; o <<= 1) {
              x += __shfl_xor(x, o, 64);
              y += __shfl_xor(y, o, 64);
            }
            gx[t][v] = x;
            gy[t][v] = y;
          }
        }
        if (I->col_rhs > 0) {
          // lane col < n_rhs holds rhs col for every point of its rows
          if (col < I->n_rhs) {
#pragma unroll
            for (int t = 0; t < TPW; ++t)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int i = i_s + NP * (16 * (t0 + t) + 4 * kk + v) + ph;
                if (i < I->Nx) g_out[(size_t)col * I->N_pts + (size_t)z * I->Nx + i] = make_float2(gx[t][v], gy[t][v]);
              }
          }
        } else if (col < 4) {
          // lane col < 4 writes v = col for every tile
#pragma unroll
          for (int t = 0; t < TPW; ++t) {
            float x = gx[t][0], y = gy[t][0];
#pragma unroll
            for (int v = 1; v < 4; ++v) {
              if (col == v) {
                x = gx[t][v];
                y = gy[t][v];
              }
            }
            const int s = 16 * (t0 + t) + 4 * kk + col;
            const int i = i_s + NP * s + ph;
            if (i < I->Nx) g_out[(size_t)r * I->N_pts + (size_t)z * I->Nx + i] = make_float2(x, y);
          }
        }
#pragma unroll
        for (int t = 0; t < TPW; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) gx[t][v] = gy[t][v] = 0.f;
      }
    }
    __syncthreads();
  }
}

template <int MT, int NP>
static size_t adj_split_lds_bytes() {
  return AdjSplitLayout<MT, NP>::bytes;
}

// ======================================================================
// launchers
// ======================================================================
bool fwd_split_supported(const Plan& P) {
  return P.structured && (P.Np == 4 || P.Np == 2) && P.Ne_pad >= 32 && P.Ne_pad <= 128;
}

void launch_fwd_split(const LaunchSet& L, int n_cu, hipStream_t s) {
  if (!L.inst.empty() && L.inst[0].sgrp > 1) {   // 3-D short rows: several y slices per item
    launch_fwd_split_slices(L, n_cu, s);
    return;
  }
  const int grid = std::max(1, std::min(L.n_items_s, (L.grid_mult > 0 ? L.grid_mult : 2) * n_cu));
  bool even = true;
  for (const Inst& I : L.inst) even = even && (I.Nx % 2) == 0;
  const int n_inst = (int)L.inst.size();
  const int Ne_pad = 16 * L.MT;
  const bool half = 16 * L.S_max <= 128;   // every instance's rows fit the first half of a segment
#define BORN_FWDS_K(MTv, EV, NPv, F16v, Hv)                                                            \
  hipLaunchKernelGGL((k_fwd_s2d_split<MTv, EV, NPv, F16v, Hv>), dim3(grid), dim3(256), (fwd_split_lds_bytes<MTv, NPv>()), s, \
                     L.d_inst, n_inst, L.n_items_s, Ne_pad, L.d_chunks, L.d_chunk_pairs)
#define BORN_FWDS_L(MTv, EV, NPv)                                        \
  do {                                                                   \
    if (L.f16) {                                                         \
      if (half) BORN_FWDS_K(MTv, EV, NPv, true, true);                   \
      else BORN_FWDS_K(MTv, EV, NPv, true, false);                       \
    } else {                                                             \
      if (half) BORN_FWDS_K(MTv, EV, NPv, false, true);                  \
      else BORN_FWDS_K(MTv, EV, NPv, false, false);                      \
    }                                                                    \
  } while (0)
#define BORN_FWDS_NP(MTv, EV)                 \
  if (L.Np == 2) BORN_FWDS_L(MTv, EV, 2);     \
  else BORN_FWDS_L(MTv, EV, 4);
#define BORN_FWDS(MTv)          \
  if (even) {                   \
    BORN_FWDS_NP(MTv, true)     \
  } else {                      \
    BORN_FWDS_NP(MTv, false)    \
  }
  switch (L.MT) {
    case 2: BORN_FWDS(2) break;
    case 4: BORN_FWDS(4) break;
    default: BORN_FWDS(8) break;
  }
#undef BORN_FWDS
#undef BORN_FWDS_NP
#undef BORN_FWDS_L
#undef BORN_FWDS_K
}

void launch_adj_split(const LaunchSet& L, const OutPtrs& outs, hipStream_t s) {
  if (!L.inst.empty() && L.inst[0].sgrp == 4 && L.MT == 2) {   // 3-D short rows, 32 elements per slice
    launch_adj_split_slices(L, outs, s);
    return;
  }
  const dim3 grid(L.n_wg_s);
  const int n_inst = (int)L.inst.size();
  bool y3 = false;
  for (const Inst& I : L.inst) y3 = y3 || I.Ney > 1;
  const bool half = 16 * L.S_max <= 128;   // every instance's rows fit the first half of a segment
#define BORN_ADJS_K(MTv, NPv, Y3v, F16v, Hv)                                                           \
  hipLaunchKernelGGL((k_adj_s2d_split<MTv, NPv, Y3v, F16v, Hv>), grid, dim3(512), (adj_split_lds_bytes<MTv, NPv>()), s, \
                     L.d_inst, n_inst, outs, L.d_chunks, L.d_chunk_pairs)
#define BORN_ADJS_L(MTv, NPv, Y3v)                                       \
  do {                                                                   \
    if (L.f16) {                                                         \
      if (half) BORN_ADJS_K(MTv, NPv, Y3v, true, true);                  \
      else BORN_ADJS_K(MTv, NPv, Y3v, true, false);                      \
    } else {                                                             \
      if (half) BORN_ADJS_K(MTv, NPv, Y3v, false, true);                 \
      else BORN_ADJS_K(MTv, NPv, Y3v, false, false);                     \
    }                                                                    \
  } while (0)
#define BORN_ADJS_NP(MTv)                                     \
  if (L.Np == 2) {                                            \
    if (y3) BORN_ADJS_L(MTv, 2, true); else BORN_ADJS_L(MTv, 2, false); \
  } else {                                                    \
    if (y3) BORN_ADJS_L(MTv, 4, true); else BORN_ADJS_L(MTv, 4, false); \
  }
  switch (L.MT) {
    case 2: BORN_ADJS_NP(2) break;
    case 4: BORN_ADJS_NP(4) break;
    default: BORN_ADJS_NP(8) break;
  }
#undef BORN_ADJS_NP
#undef BORN_ADJS_L
#undef BORN_ADJS_K
}

template <int MT, int NP>
static void set_max_lds_split_mt(int bytes) {
#define BORN_SET(K) hipFuncSetAttribute((const void*)(K), hipFuncAttributeMaxDynamicSharedMemorySize, bytes)
  BORN_SET((k_fwd_s2d_split<MT, true, NP, false, false>));
  BORN_SET((k_fwd_s2d_split<MT, false, NP, false, false>));
  BORN_SET((k_fwd_s2d_split<MT, true, NP, false, true>));
  BORN_SET((k_fwd_s2d_split<MT, false, NP, false, true>));
  BORN_SET((k_adj_s2d_split<MT, NP, false, false, false>));
  BORN_SET((k_adj_s2d_split<MT, NP, true, false, false>));
  BORN_SET((k_adj_s2d_split<MT, NP, false, false, true>));
  BORN_SET((k_adj_s2d_split<MT, NP, true, false, true>));
  BORN_SET((k_fwd_s2d_split<MT, true, NP, true, false>));
  BORN_SET((k_fwd_s2d_split<MT, false, NP, true, false>));
  BORN_SET((k_fwd_s2d_split<MT, true, NP, true, true>));
  BORN_SET((k_fwd_s2d_split<MT, false, NP, true, true>));
  BORN_SET((k_adj_s2d_split<MT, NP, false, true, false>));
  BORN_SET((k_adj_s2d_split<MT, NP, true, true, false>));
  BORN_SET((k_adj_s2d_split<MT, NP, false, true, true>));
  BORN_SET((k_adj_s2d_split<MT, NP, true, true, true>));
#undef BORN_SET
}

void set_max_lds_split(int bytes) {
  set_max_lds_split_mt<2, 2>(bytes);
  set_max_lds_split_mt<4, 2>(bytes);
  set_max_lds_split_mt<8, 2>(bytes);
  set_max_lds_split_mt<2, 4>(bytes);
  set_max_lds_split_mt<4, 4>(bytes);
  set_max_lds_split_mt<8, 4>(bytes);
}

}  // namespace born
