// Split-precision forward contraction of 3-D plans with short rows (gfx950).
//
// In 3-D the forward of one (frequency, row (y, z)) is, for every y slice
// ey of the array, the Toeplitz-in-x product of k_fwd_s2d_split
// (born_split.hip):
//     V[ey][ex][c] += sum_i Hext_ey[i - ex*Np + off] * B[i][c],
// where Hext_ey is the table row that slice ey sees at this row and
// B[i][c] = p_inc[pair_c][i] gamma_r[i] is the same for every slice.  With
// rows of at most 128 points and 32 elements per slice (a 32 x 32 matrix
// array), one slice per step leaves the MFMAs a small share of every step:
// B costs as many LDS writes as the slice's whole MFMA work reads.  Here a
// work item covers G slices: B is staged once per row and the G slices'
// windows side by side form the M dimension (G * 16 * MTS elements), so a
// barrier pair and one B fill serve G times the MFMAs.
//
// The numeric form is that of k_fwd_s2d_split: hi/lo BF16 (or scaled FP16)
// operands, 12 MFMA 16x16x32 per complex multiply-add and K-block, FP32
// accumulation, partial sums per (chunk, slice, row block) summed in FP64 by
// k_fwd_reduce.
#include "born_internal.h"

namespace born {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

namespace {

template <bool F16>
__device__ __forceinline__ v4f mfma32s(uint4 a, uint4 b, v4f c) {
  if constexpr (F16)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}
__device__ __forceinline__ uint4 neg8s(uint4 v) {
  return make_uint4(v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u);
}
__device__ __forceinline__ v4f vzs() { return v4f{0.f, 0.f, 0.f, 0.f}; }

// hi = t(x), lo = t(x - hi) for t = BF16 or FP16 (nearest even), a pair of
// values packed per 32-bit word (x0 in the low half)
template <bool F16>
__device__ __forceinline__ void split_pair_s(float x0, float x1, unsigned int& hi, unsigned int& lo) {
  typedef float v2f __attribute__((ext_vector_type(2)));
  if constexpr (F16) {
    typedef _Float16 v2h __attribute__((ext_vector_type(2)));
    const v2h h = __builtin_convertvector((v2f{x0, x1}), v2h);
    const v2h l = __builtin_convertvector((v2f{x0 - (float)h[0], x1 - (float)h[1]}), v2h);
    hi = __builtin_bit_cast(unsigned int, h);
    lo = __builtin_bit_cast(unsigned int, l);
  } else {
    typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
    const unsigned int p = __builtin_bit_cast(unsigned int, __builtin_convertvector((v2f{x0, x1}), v2bf));
    const float r0 = __uint_as_float(p << 16), r1 = __uint_as_float(p & 0xffff0000u);
    hi = p;
    lo = __builtin_bit_cast(unsigned int, __builtin_convertvector((v2f{x0 - r0, x1 - r1}), v2bf));
  }
}

constexpr int SEGH = 128;          // points per row (rows of at most 128 points)
constexpr int SEGHP = SEGH + 8;    // padded B row (u16): 16-byte reads of 16 lanes cover the 64 banks once

}  // namespace

// Copy stride of the window planes (u16 entries): at least LA + 12, a 16-byte
// multiple, 0 mod 64 for 2 copies and 32 mod 64 for 4, as for the row windows
// of k_fwd_s2d_split.
__host__ __device__ constexpr int fwd_split3_lap(int LA, int NP) {
  return ((LA + 12 + 63) / 64) * 64 + (NP == 4 ? 0 : 32);
}

// Work item = (instance, chunk, slice group sg, row block zb), taken
// persistently in reverse order (the forward then reads p_inc from the last
// frequency down, where the adjoint before it stopped).  A step is one row
// (y, z) of the block.  Waves: 2 element groups x 2 point groups; wave (eg,
// pg) holds tiles [eg*MTW, (eg+1)*MTW) -- tile m belongs to slice m / MTS of
// the group, element tile m % MTS of that slice -- and the K-blocks
// 2 pg, 2 pg + 1 of the 4 K-blocks of 32 points.
//
// LDS (u16 units):
//   Aw[comp][slice][copy][LAP]  the G windows (comp = re_hi, re_lo, im_hi,
//                               im_lo); copy c is copy 0 shifted by NP*c
//                               entries so every lane's 8-entry read is
//                               16-byte aligned
//   Bt[comp][c][SEGHP]          p_inc*gamma of the row, points contiguous
template <int MTS, int G, bool EVEN, int NP, bool F16>
__global__ __launch_bounds__(256, 2) void k_fwd_split_slices(const Inst* __restrict__ inst, int n_inst, int n_items,
                                                             const Chunk* __restrict__ chunks,
                                                             const int* __restrict__ chunk_pairs) {
  extern __shared__ unsigned short ldsu[];
  constexpr int MT = MTS * G;           // element tiles of a step
  constexpr int NEPS = 16 * MTS;        // padded elements of one slice
  constexpr int NCOPY = 8 / NP;
  constexpr int DSTEP = NP / 2;         // K-blocks of 32 points per element tile of 16
  constexpr int EW = 2, PW = 2;
  constexpr int MTW = MT / EW;          // tiles per wave
  constexpr int KBW = (SEGH / 32) / PW; // K-blocks per wave
  constexpr int LA = SEGH + NP * (NEPS - 1);      // window entries of one slice
  constexpr int LAP = fwd_split3_lap(LA, NP);
  constexpr int NPAIR = (LA + 1) / 2;             // entry pairs of one window
  constexpr int TPS = 256 / G;                    // fill threads per slice window
  constexpr int APT = (NPAIR + TPS - 1) / TPS;    // pairs per thread
  static_assert(MT % EW == 0 && MTW % MTS == 0, "a wave holds whole slices");
  static_assert(256 % G == 0, "fill threads split evenly over the slices");
  unsigned short* Aw = ldsu;                              // [4][G][NCOPY][LAP]
  unsigned short* Bt = ldsu + 4 * G * NCOPY * LAP;        // [4][NC][SEGHP]
  unsigned int* Aw32 = reinterpret_cast<unsigned int*>(Aw);
  unsigned int* Bt32 = reinterpret_cast<unsigned int*>(Bt);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int eg = w % EW, pg = w / EW;
  const int kk = lane >> 4, col = lane & 15;
  const int n_my = (n_items - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  if (n_my <= 0) return;
  int ki = 0;
  Inst Iv = inst[0];
  const Inst* I = &Iv;
  // fill roles: the window of slice fs = tid / TPS, entry pairs q = tid % TPS + TPS j;
  // B: columns bc0 .. bc0 + 3, points bi, bi + 1
  const int fs = tid / TPS, fq = tid % TPS;
  const int bc0 = (tid >> 6) * 4;
  const int bi = 2 * (tid & 63);

  int it = 0, item = blockIdx.x;
  int ci, sg, zb, z, z_end, fg, rr, lid;
  bool hfull;
  unsigned int hoff[2 * APT];        // window gather byte offsets within a table row
  unsigned int hval = 0;             // bit j: window entry j inside the table
  unsigned int goff0, goff1;
  int prow[4];
#define BORN_S3_ITEM(itm)                                                                   \
  {                                                                                         \
    if (ki + 1 < n_inst && (itm) >= inst[ki + 1].item_begin_s) {                            \
      while (ki + 1 < n_inst && (itm) >= inst[ki + 1].item_begin_s) ++ki;                   \
      Iv = inst[ki];                                                                        \
    }                                                                                       \
    const int tpad_ = (NEPS - I->N_e) * NP;                                                 \
    const int NSG_ = I->Ney / G;                                                            \
    lid = I->n_chunks * NSG_ * I->NZB_s - 1 - ((itm) - I->item_begin_s);                    \
    slice_item(lid, NSG_, I->NZB_s, ci, sg, zb);                                            \
    z = zb * I->ZR_s;                                                                       \
    z_end = min(I->n_rows, z + I->ZR_s);                                                    \
    const Chunk c_ = chunks[I->chunk_base + ci];                                            \
    fg = c_.fg;                                                                             \
    rr = c_.r;                                                                              \
    const int t0_ = -tpad_;                                                                 \
    hfull = t0_ >= 0 && t0_ + LA <= I->N_ext;                                                \
    hval = 0;                                                                               \
    _Pragma("unroll") for (int j = 0; j < 2 * APT; ++j) {                                  \
      const int u_ = 2 * (fq + TPS * (j >> 1)) + (j & 1);                                   \
      const int t_ = u_ + t0_;                                                              \
      const bool ok_ = u_ < LA && t_ >= 0 && t_ < I->N_ext;                                  \
      hoff[j] = ok_ ? 8u * (unsigned int)I->ext[t_] : 0u;                                    \
      hval |= ok_ ? (1u << j) : 0u;                                                         \
    }                                                                                       \
    goff0 = 8u * (unsigned int)(EVEN ? min(bi, I->Nx - 2) : min(bi, I->Nx - 1));            \
    goff1 = 8u * (unsigned int)min(bi + 1, I->Nx - 1);                                      \
    _Pragma("unroll") for (int q = 0; q < 4; ++q)                                          \
      prow[q] = __builtin_amdgcn_readfirstlane(chunk_pairs[(I->chunk_base + ci) * NC + bc0 + q]); \
    _Pragma("unroll") for (int j = 0; j < 2 * APT; ++j) asm volatile("" ::"v"(hoff[j]));  \
  }

  float2 hreg[2 * APT];
  float2 preg[4][2];
  float2 greg[2];

#define BORN_S3_PREFETCH(zz)                                                                       \
  {                                                                                                \
    const size_t rowo_ = (size_t)(zz) * I->Nx;                                                      \
    const int ey_ = sg * G + fs;                                                                   \
    const float2* hrow_ = I->h + (size_t)fg * I->Hp + (size_t)I->hrow[ey_ * I->n_rows + (zz)] * I->HR; \
    _Pragma("unroll") for (int j = 0; j < 2 * APT; ++j)                                           \
      hreg[j] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(hrow_) + hoff[j]);   \
    const float2* grow_ = I->gm + (size_t)rr * I->N_pts + rowo_;                                     \
    if (EVEN) {                                                                                    \
      const float4 v_ = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(grow_) + goff0); \
      greg[0] = make_float2(v_.x, v_.y);                                                           \
      greg[1] = make_float2(v_.z, v_.w);                                                           \
    } else {                                                                                       \
      greg[0] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(grow_) + goff0);     \
      greg[1] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(grow_) + goff1);     \
    }                                                                                              \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                                               \
      const float2* prow_ = I->P + (size_t)prow[q] * I->N_pts + rowo_;                               \
      if (EVEN) {                                                                                  \
        const float4 v_ = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(prow_) + goff0); \
        preg[q][0] = make_float2(v_.x, v_.y);                                                      \
        preg[q][1] = make_float2(v_.z, v_.w);                                                      \
      } else {                                                                                     \
        preg[q][0] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(prow_) + goff0); \
        preg[q][1] = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(prow_) + goff1); \
      }                                                                                            \
    }                                                                                              \
  }

  BORN_S3_ITEM(item)
  BORN_S3_PREFETCH(z)

  v4f CR[MTW], CI[MTW];
#pragma unroll
  for (int m = 0; m < MTW; ++m) {
    CR[m] = vzs();
    CI[m] = vzs();
  }

  // A read position of lane (kk, col) for local tile 0 of a slice, K-block 0:
  // u = 8 kk + NP (NEPS - 1 - col); the copy (u mod 8) / NP makes it aligned
  const int u00 = 8 * kk + NP * (NEPS - 1 - col);
  const int acopy = (u00 & 7) / NP;
  const int abase = acopy * LAP + (u00 - NP * acopy);
  const int bbase = col * SEGHP + 8 * kk;

  while (true) {
    __syncthreads();
    const float s_h = F16 ? I->fscale[0] : 1.f, s_b = F16 ? I->fscale[1] : 1.f;
    // windows -> hi/lo planes, NCOPY copies; entry pair (u, u + 1) per 32-bit word
#pragma unroll
    for (int j = 0; j < 2 * APT; j += 2) {
      const int u = 2 * (fq + TPS * (j >> 1));
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
        unsigned int v[4];
        split_pair_s<F16>(h0.x, h1.x, v[0], v[1]);
        split_pair_s<F16>(h0.y, h1.y, v[2], v[3]);
#pragma unroll
        for (int cp = 0; cp < 4; ++cp) {
#pragma unroll
          for (int c = 0; c < NCOPY; ++c)
            if (u >= NP * c) Aw32[(((cp * G + fs) * NCOPY + c) * LAP + u - NP * c) >> 1] = v[cp];
        }
      }
    }
    // B = p_inc * gamma -> hi/lo planes: points bi, bi + 1 of columns bc0 .. bc0 + 3
    {
      float2 g0 = greg[0], g1 = greg[1];
      if (bi >= I->Nx) g0 = make_float2(0.f, 0.f);
      if (bi + 1 >= I->Nx) g1 = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float2 p0 = preg[q][0], p1 = preg[q][1];
        float re0 = p0.x * g0.x - p0.y * g0.y, im0 = p0.x * g0.y + p0.y * g0.x;
        float re1 = p1.x * g1.x - p1.y * g1.y, im1 = p1.x * g1.y + p1.y * g1.x;
        if (F16) {
          re0 *= s_b;
          im0 *= s_b;
          re1 *= s_b;
          im1 *= s_b;
        }
        unsigned int rh, rl, ih, il;
        split_pair_s<F16>(re0, re1, rh, rl);
        split_pair_s<F16>(im0, im1, ih, il);
        const int o = ((bc0 + q) * SEGHP + bi) >> 1;
        Bt32[0 * (NC * SEGHP / 2) + o] = rh;
        Bt32[1 * (NC * SEGHP / 2) + o] = rl;
        Bt32[2 * (NC * SEGHP / 2) + o] = ih;
        Bt32[3 * (NC * SEGHP / 2) + o] = il;
      }
    }
    __syncthreads();

    const int ci_done = ci, sg_done = sg, zb_done = zb;
    float2* const vpart_done = Iv.vpart;
    const int Ney_done = I->Ney, NZB_done = I->NZB_s * I->nseg_s;
    const float unscale_done = F16 ? Iv.fscale[2] : 1.f;
    const bool item_done = z + 1 >= z_end;
    bool has_next = true;
    if (!item_done) {
      ++z;
    } else if (it + 1 >= n_my) {
      has_next = false;
    } else {
      item = blockIdx.x + (it + 1) * gridDim.x;
      BORN_S3_ITEM(item)
    }
    if (has_next) BORN_S3_PREFETCH(z)

    // B of the wave's K-blocks in registers, shared by every slice
    uint4 B4[KBW][4], NB[KBW][2];
#pragma unroll
    for (int kb = 0; kb < KBW; ++kb) {
      const int k = pg * KBW + kb;
#pragma unroll
      for (int cp = 0; cp < 4; ++cp)
        B4[kb][cp] = *reinterpret_cast<const uint4*>(&Bt[cp * NC * SEGHP + bbase + 32 * k]);
      NB[kb][0] = neg8s(B4[kb][2]);
      NB[kb][1] = neg8s(B4[kb][3]);
    }
    // per slice of the wave: local tile lt at K-block kb reads the slice's
    // window at 32 (pg KBW + kb) - 16 NP lt, the data of tile 0 at K-block
    // kb - DSTEP lt: walk the diagonals d = kb - DSTEP lt
    constexpr int D0 = -DSTEP * (MTS - 1);
#pragma unroll
    for (int sl = 0; sl < MTW / MTS; ++sl) {
      const int s = (eg * MTW) / MTS + sl;            // slice within the group
#pragma unroll
      for (int d = D0; d < KBW; ++d) {
        uint4 Ac[4];
        const int ao = abase + 32 * (pg * KBW + d);
#pragma unroll
        for (int cp = 0; cp < 4; ++cp)
          Ac[cp] = *reinterpret_cast<const uint4*>(&Aw[(cp * G + s) * NCOPY * LAP + ao]);
#pragma unroll
        for (int lt = 0; lt < MTS; ++lt) {
          const int kb = d + DSTEP * lt;
          const int m = sl * MTS + lt;
          if (kb >= 0 && kb < KBW) {
            CR[m] = mfma32s<F16>(Ac[0], B4[kb][0], CR[m]);
            CR[m] = mfma32s<F16>(Ac[0], B4[kb][1], CR[m]);
            CR[m] = mfma32s<F16>(Ac[1], B4[kb][0], CR[m]);
            CR[m] = mfma32s<F16>(Ac[2], NB[kb][0], CR[m]);
            CR[m] = mfma32s<F16>(Ac[2], NB[kb][1], CR[m]);
            CR[m] = mfma32s<F16>(Ac[3], NB[kb][0], CR[m]);
            CI[m] = mfma32s<F16>(Ac[0], B4[kb][2], CI[m]);
            CI[m] = mfma32s<F16>(Ac[0], B4[kb][3], CI[m]);
            CI[m] = mfma32s<F16>(Ac[1], B4[kb][2], CI[m]);
            CI[m] = mfma32s<F16>(Ac[2], B4[kb][0], CI[m]);
            CI[m] = mfma32s<F16>(Ac[2], B4[kb][1], CI[m]);
            CI[m] = mfma32s<F16>(Ac[3], B4[kb][0], CI[m]);
          }
        }
      }
    }

    if (item_done) {
      __syncthreads();
      float* red = reinterpret_cast<float*>(ldsu);
#pragma unroll
      for (int m = 0; m < MTW; ++m) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          red[((w * MTW + m) * 8 + v) * 64 + lane] = CR[m][v];
          red[((w * MTW + m) * 8 + 4 + v) * 64 + lane] = CI[m][v];
        }
        CR[m] = vzs();
        CI[m] = vzs();
      }
      __syncthreads();
      // each slice of the group to its own partial-sum block:
      // [(chunk * Ney + ey) * NZB + zb][element][column]
      for (int idx = tid; idx < MT * 512; idx += 256) {
        const int mt = idx >> 9, e_g = mt / MTW, mtw = mt - e_g * MTW, rest = idx & 511;
        float sum = 0.f;
#pragma unroll
        for (int p_g = 0; p_g < PW; ++p_g) sum += red[((p_g * EW + e_g) * MTW + mtw) * 512 + rest];
        const int s = mt / MTS, lt = mt - s * MTS;
        const int ey = sg_done * G + s;
        float* out = reinterpret_cast<float*>(
            vpart_done + ((size_t)(ci_done * Ney_done + ey) * NZB_done + zb_done) * NEPS * NC);
        const int ln = idx & 63, j = (idx >> 6) & 7;
        const int e = 16 * lt + 4 * (ln >> 4) + (j & 3);
        out[((size_t)e * NC + (ln & 15)) * 2 + (j >> 2)] = F16 ? sum * unscale_done : sum;
      }
      ++it;
    }
    if (!has_next) break;
  }
#undef BORN_S3_PREFETCH
#undef BORN_S3_ITEM
}

template <int MTS, int G, int NP>
static size_t fwd_split3_lds_bytes() {
  constexpr int LA = SEGH + NP * (16 * MTS - 1);
  constexpr int LAP = fwd_split3_lap(LA, NP);
  const size_t body = ((size_t)4 * G * (8 / NP) * LAP + (size_t)4 * NC * SEGHP) * sizeof(unsigned short);
  const size_t red = (size_t)4 * (MTS * G / 2) * 8 * 64 * sizeof(float);
  return std::max(body, red);
}

// ======================================================================
// adjoint over several y slices per step
// ======================================================================
// One workgroup = (instance, row (y, z)) of at most 128 points.  For one
// chunk the adjoint sums over the y slices ey of the Toeplitz products of
// k_adj_s2d_split (born_split.hip):
//     Y[s][c] = sum_ey sum_e conj(H_ey(NP s + ph - NP e + off)) W[ey][e][c],
// then gamma_hat += conj(p_inc) Y.  With 32 elements per slice one slice is a
// single K-block of 32, and a step of one slice is too short for the loads
// two steps ahead to arrive: the steps wait on memory latency.  Here a step
// is (chunk, group of G slices): G windows and G weight blocks per buffer,
// the compute waves walk the slices of the group, so every latency-bound step
// carries G times the MFMAs.
//
// Layout per buffer (u16): G slices of {Ar[comp][ph][CPB], Wt[comp][c][WS]}
// with the reversed-row copies of the one-slice kernel, here for a 128-point
// segment; then two p_inc buffers Ps[c][PSS] (float2) by chunk parity.
template <int NP>
struct AdjSplit3Layout {
  static constexpr int G = 4;                                   // slices per step
  static constexpr int NEP = 32;                                // elements per slice (two tiles)
  static constexpr int SPH = SEGH / NP;                         // points per phase
  static constexpr int LQ = SPH + NEP;                          // reversed phase-row entries
  static constexpr int PADL = 4;                                // lead pad of every copy
  static constexpr int LR = ((LQ + PADL - 4 + 63) / 64) * 64 + 4;   // copy stride, 4 mod 64
  // (comp, phase) block, 16 mod 64; with 2 phases 16 entries more, which
  // measured a quarter fewer LDS bank conflicts of the fill waves' window
  // stores (the 4-phase layout has no LDS left for it)
  static constexpr int CPB = ((4 * LR + 40 - 16 + 63) / 64) * 64 + 16 + (NP == 2 ? 16 : 0);
  static constexpr int WS = NEP + 8;                            // padded W row
  static constexpr int PSS = SEGH + 1;                          // padded p_inc row (float2)
  static constexpr int AR_U16 = 4 * NP * CPB;
  static constexpr int SL_U16 = AR_U16 + 4 * NC * WS;           // one slice: window copies + weights
  static constexpr int AW_U16 = ((G * SL_U16 + 7) / 8) * 8;     // one buffer
  static constexpr int PS_U16 = ((NC * PSS * 4 + 7) / 8) * 8;   // one p_inc buffer
  static constexpr int PS_OFF = 2 * AW_U16;
  static constexpr size_t bytes = (size_t)(2 * AW_U16 + 2 * PS_U16) * 2;
};
__device__ __forceinline__ int copy_offset3(int d) { return d == 1 ? 16 : (d == 2 ? 28 : (d == 3 ? 40 : 0)); }
__device__ __forceinline__ uint4 join3(uint2 a, uint2 b) { return make_uint4(a.x, a.y, b.x, b.y); }

template <int NP, bool F16>
__global__ __launch_bounds__(512, 1) void k_adj_split_slices(const Inst* __restrict__ inst, int n_inst, OutPtrs outs,
                                                             const Chunk* __restrict__ chunks,
                                                             const int* __restrict__ chunk_pairs) {
  extern __shared__ unsigned short ldsu[];
  typedef AdjSplit3Layout<NP> Lay;
  constexpr int G = Lay::G, NEP = Lay::NEP, SPH = Lay::SPH, LQ = Lay::LQ, PADL = Lay::PADL;
  constexpr int LR = Lay::LR, CPB = Lay::CPB, WS = Lay::WS, PSS = Lay::PSS;
  constexpr int L = SEGH + NP * (NEP - 1);      // window entries of one slice
  constexpr int HPT = 4;
  constexpr int WG4 = ((L + NP - 1) / NP + 3) / 4;   // 4-entry groups per phase
  constexpr int TPSW = 256 / G;                      // fill threads per slice window
  static_assert(NP * WG4 <= TPSW, "window groups exceed the fill threads of a slice");
  static_assert(LQ % 4 == 0 && LR % 4 == 0 && CPB % 4 == 0, "copy alignment classes assume multiples of 4");
  constexpr int WPT = NEP * NC / 256;           // W entries per fill thread and slice (2)
  constexpr int PPT = NC * SEGH / 256;          // p_inc entries per fill thread (8)
  constexpr int WPP = 4 / NP;                   // compute waves per phase
  constexpr int TPW = 2;                        // tiles per compute wave (64 points of 4 phase tiles)

  const int tid = threadIdx.x, lane = tid & 63, w = (tid >> 6) & 3;
  const bool filler = tid >= 256;
  const int ft = tid & 255;
  const int kk = lane >> 4, col = lane & 15;
  int ki = 0;
  while (ki + 1 < n_inst && (int)blockIdx.x >= inst[ki + 1].wg_begin_s) ++ki;
  const Inst Iv = inst[ki];
  const Inst* I = &Iv;
  const int z = (int)blockIdx.x - I->wg_begin_s;   // one 128-point segment per row
  const int NSG = I->Ney / G;
  const int n_steps = I->n_chunks * NSG;

  if (filler) {
    const int t_shift = -(NEP - I->N_e) * NP;   // window entry u holds Hext[u + t_shift]
    const float s_h = F16 ? I->fscale[0] : 1.f, s_w = F16 ? I->fscale[3] : 1.f;
    const int fsl = ft / TPSW, fw = ft % TPSW;  // window role: slice of the group, thread within it
    const bool wact = fw < NP * WG4;
    const int wp = fw % NP, wr = fw / NP;
    int eidx[HPT];
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const int u = NP * (4 * wr + j) + wp;
      const int t = u + t_shift;
      eidx[j] = (wact && u < L && t >= 0 && t < I->N_ext) ? I->ext[t] : -1;
    }
    const int pt = ft & (SEGH - 1), pcg = ft >> 7;   // p_inc role: point, column group of 8
    const int gi = min(pt, I->Nx - 1);
    float2 hA[HPT], wA[G][WPT], pA[PPT], hB[HPT], wB[G][WPT], pB[PPT];
#define BORN_AS3_PREFETCH(ST, H, WV, PV)                                                        \
    {                                                                                           \
      const int st_ = (ST);                                                                     \
      const int c_ = st_ / NSG, sg_ = st_ - c_ * NSG;                                           \
      const Chunk ch_ = chunks[I->chunk_base + c_];                                             \
      const int row_ = I->hrow[(sg_ * G + fsl) * I->n_rows + z];                                \
      const float2* hrow_ = I->h + (size_t)ch_.fg * I->Hp + (size_t)row_ * I->HR;               \
      _Pragma("unroll") for (int j = 0; j < HPT; ++j) H[j] = hrow_[max(eidx[j], 0)];            \
      _Pragma("unroll") for (int s_ = 0; s_ < G; ++s_) {                                       \
        const float2* wblk_ = I->Wc + ((size_t)c_ * I->Ney + sg_ * G + s_) * NEP * NC;           \
        _Pragma("unroll") for (int j = 0; j < WPT; ++j)                                         \
          WV[s_][j] = wblk_[(2 * (ft >> 4) + j) * NC + (ft & 15)];                              \
      }                                                                                         \
      if (sg_ == 0) {                                                                           \
        const int* prs_ = chunk_pairs + (size_t)(I->chunk_base + c_) * NC + 8 * pcg;             \
        _Pragma("unroll") for (int j = 0; j < PPT; ++j)                                         \
          PV[j] = I->P[(size_t)prs_[j] * I->N_pts + (size_t)z * I->Nx + gi];                    \
      }                                                                                         \
    }
#define BORN_AS3_FILL(ST, H, WV, PV)                                                            \
    {                                                                                           \
      const int st_ = (ST);                                                                     \
      const int c_ = st_ / NSG, sg_ = st_ - c_ * NSG;                                           \
      unsigned short* buf_ = ldsu + (st_ & 1) * Lay::AW_U16;                                    \
      if (wact) {                                                                               \
        unsigned short* Ar = buf_ + fsl * Lay::SL_U16;                                          \
        float xr[4], xi[4];   /* position order: index i holds q = 4 wr + 3 - i */              \
        _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                         \
          float2 hv = H[3 - i];                                                                 \
          if (eidx[3 - i] < 0) hv = make_float2(0.f, 0.f);                                      \
          if (F16) hv = make_float2(hv.x * s_h, hv.y * s_h);                                    \
          xr[i] = hv.x;                                                                         \
          xi[i] = -hv.y; /* conj */                                                             \
        }                                                                                       \
        unsigned int pr[4][2];                                                                  \
        split_pair_s<F16>(xr[0], xr[1], pr[0][0], pr[1][0]);                                    \
        split_pair_s<F16>(xr[2], xr[3], pr[0][1], pr[1][1]);                                    \
        split_pair_s<F16>(xi[0], xi[1], pr[2][0], pr[3][0]);                                    \
        split_pair_s<F16>(xi[2], xi[3], pr[2][1], pr[3][1]);                                    \
        _Pragma("unroll") for (int cp = 0; cp < 4; ++cp) {                                      \
          unsigned short* P0 = Ar + (cp * NP + wp) * CPB + PADL + (LQ - 4 - 4 * wr);            \
          const unsigned int A_ = pr[cp][0], B_ = pr[cp][1];                                    \
          const unsigned int mid_ = (A_ >> 16) | (B_ << 16);                                    \
          *reinterpret_cast<uint2*>(P0) = make_uint2(A_, B_);                                   \
          unsigned short* P1 = P0 + LR + 16 - 1;                                                \
          P1[0] = (unsigned short)A_;                                                           \
          *reinterpret_cast<unsigned int*>(P1 + 1) = mid_;                                      \
          P1[3] = (unsigned short)(B_ >> 16);                                                   \
          unsigned short* P2 = P0 + 2 * LR + 28 - 2;                                            \
          *reinterpret_cast<unsigned int*>(P2) = A_;                                            \
          *reinterpret_cast<unsigned int*>(P2 + 2) = B_;                                        \
          unsigned short* P3 = P0 + 3 * LR + 40 - 3;                                            \
          P3[0] = (unsigned short)A_;                                                           \
          *reinterpret_cast<unsigned int*>(P3 + 1) = mid_;                                      \
          P3[3] = (unsigned short)(B_ >> 16);                                                   \
        }                                                                                       \
      }                                                                                         \
      /* weights: elements k, k + 1 of column ft mod 16, one 32-bit word per plane */           \
      _Pragma("unroll") for (int s_ = 0; s_ < G; ++s_) {                                       \
        unsigned short* Wt = buf_ + s_ * Lay::SL_U16 + Lay::AR_U16;                             \
        const int k = 2 * (ft >> 4), wc = ft & 15;                                              \
        unsigned int rh, rl, ih, il;                                                            \
        const float wsc_ = F16 ? s_w : 1.f;                                                     \
        split_pair_s<F16>(WV[s_][0].x * wsc_, WV[s_][1].x * wsc_, rh, rl);                      \
        split_pair_s<F16>(WV[s_][0].y * wsc_, WV[s_][1].y * wsc_, ih, il);                      \
        unsigned int* wd = reinterpret_cast<unsigned int*>(&Wt[wc * WS + k]);                   \
        wd[0 * NC * WS / 2] = rh;                                                               \
        wd[1 * NC * WS / 2] = rl;                                                               \
        wd[2 * NC * WS / 2] = ih;                                                               \
        wd[3 * NC * WS / 2] = il;                                                               \
      }                                                                                         \
      if (sg_ == 0) {                                                                           \
        float2* Ps = reinterpret_cast<float2*>(ldsu + Lay::PS_OFF + (c_ & 1) * Lay::PS_U16);    \
        _Pragma("unroll") for (int j = 0; j < PPT; ++j) Ps[(8 * pcg + j) * PSS + pt] = PV[j];   \
      }                                                                                         \
    }
    const int last = n_steps - 1;
    BORN_AS3_PREFETCH(0, hA, wA, pA)
    BORN_AS3_FILL(0, hA, wA, pA)
    BORN_AS3_PREFETCH(min(1, last), hB, wB, pB)
    BORN_AS3_PREFETCH(min(2, last), hA, wA, pA)
    __syncthreads();
    for (int st = 0; st < n_steps; st += 2) {
      if (st + 1 < n_steps) BORN_AS3_FILL(st + 1, hB, wB, pB)
      BORN_AS3_PREFETCH(min(st + 3, last), hB, wB, pB)
      __syncthreads();
      if (st + 1 >= n_steps) break;
      if (st + 2 < n_steps) BORN_AS3_FILL(st + 2, hA, wA, pA)
      BORN_AS3_PREFETCH(min(st + 4, last), hA, wA, pA)
      __syncthreads();
    }
#undef BORN_AS3_PREFETCH
#undef BORN_AS3_FILL
    return;
  }

  // ---------------------------------------------------------------- compute waves
  float2* __restrict__ g_out = outs.p[ki];
  const int ph = w / WPP, t0 = TPW * (w % WPP);
  float gx[TPW][4], gy[TPW][4];
#pragma unroll
  for (int a = 0; a < TPW; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) gx[a][v] = gy[a][v] = 0.f;
  const int dcp = (4 - (col & 3)) & 3;
  const int abase = ph * CPB + dcp * LR + copy_offset3(dcp) + (SPH - col + 8 * kk - dcp) + PADL - 16 * t0;
  v4f YR[TPW], YI[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    YR[t] = vzs();
    YI[t] = vzs();
  }
  __syncthreads();
  for (int st = 0; st < n_steps; ++st) {
    const int c = st / NSG, sg = st - c * NSG;
    // the group's slices: tile t at diagonal dl reads the window at
    // 16 (2 kb - t) with kb = 0, so (t, dl) = (1, -1) and (0, 0)
#pragma unroll
    for (int s = 0; s < G; ++s) {
      const unsigned short* Ar = ldsu + (st & 1) * Lay::AW_U16 + s * Lay::SL_U16;
      const unsigned short* Wt = Ar + Lay::AR_U16;
      uint4 W4[4];
#pragma unroll
      for (int cp = 0; cp < 4; ++cp) W4[cp] = *reinterpret_cast<const uint4*>(&Wt[(cp * NC + col) * WS + 8 * kk]);
#pragma unroll
      for (int dl = -1; dl <= 0; ++dl) {
        uint4 Ac[4];
#pragma unroll
        for (int cp = 0; cp < 4; ++cp) {
          const unsigned short* a_ = Ar + cp * NP * CPB + abase + 16 * dl;
          Ac[cp] = join3(*reinterpret_cast<const uint2*>(a_), *reinterpret_cast<const uint2*>(a_ + 4));
        }
        const uint4 nih = neg8s(Ac[2]), nil = neg8s(Ac[3]);
        const int t = -dl;
        YR[t] = mfma32s<F16>(Ac[0], W4[0], YR[t]);
        YR[t] = mfma32s<F16>(Ac[0], W4[1], YR[t]);
        YR[t] = mfma32s<F16>(Ac[1], W4[0], YR[t]);
        YR[t] = mfma32s<F16>(nih, W4[2], YR[t]);
        YR[t] = mfma32s<F16>(nih, W4[3], YR[t]);
        YR[t] = mfma32s<F16>(nil, W4[2], YR[t]);
        YI[t] = mfma32s<F16>(Ac[0], W4[2], YI[t]);
        YI[t] = mfma32s<F16>(Ac[0], W4[3], YI[t]);
        YI[t] = mfma32s<F16>(Ac[1], W4[2], YI[t]);
        YI[t] = mfma32s<F16>(Ac[2], W4[0], YI[t]);
        YI[t] = mfma32s<F16>(Ac[2], W4[1], YI[t]);
        YI[t] = mfma32s<F16>(Ac[3], W4[0], YI[t]);
      }
    }
    if (sg == NSG - 1) {
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
          const int sp = 16 * (t0 + t) + 4 * kk + v;
          const float2 p = Ps[col * PSS + NP * sp + ph];
          gx[t][v] += p.x * YR[t][v] + p.y * YI[t][v];
          gy[t][v] += p.x * YI[t][v] - p.y * YR[t][v];
        }
        YR[t] = vzs();
        YI[t] = vzs();
      }
      const int r = chunks[I->chunk_base + c].r;
      const bool has_next = c + 1 < I->n_chunks;
      if (!has_next || chunks[I->chunk_base + c + 1].r != r) {
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
          if (col < I->n_rhs) {
#pragma unroll
            for (int t = 0; t < TPW; ++t)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int i = NP * (16 * (t0 + t) + 4 * kk + v) + ph;
                if (i < I->Nx) g_out[(size_t)col * I->N_pts + (size_t)z * I->Nx + i] = make_float2(gx[t][v], gy[t][v]);
              }
          }
        } else if (col < 4) {
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
            const int i = NP * (16 * (t0 + t) + 4 * kk + col) + ph;
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

void launch_adj_split_slices(const LaunchSet& L, const OutPtrs& outs, hipStream_t s) {
  const dim3 grid(L.n_wg_s);
  const int n_inst = (int)L.inst.size();
#define BORN_AS3_K(NPv, F16v)                                                                              \
  hipLaunchKernelGGL((k_adj_split_slices<NPv, F16v>), grid, dim3(512), AdjSplit3Layout<NPv>::bytes, s, L.d_inst, \
                     n_inst, outs, L.d_chunks, L.d_chunk_pairs)
  if (L.Np == 2) {
    if (L.f16) BORN_AS3_K(2, true); else BORN_AS3_K(2, false);
  } else {
    if (L.f16) BORN_AS3_K(4, true); else BORN_AS3_K(4, false);
  }
#undef BORN_AS3_K
}

// Slices per work item for a plan: 3-D, rows of at most 128 points, 16 or
// 32 padded elements per slice (G * MTS = 8 tiles), a slice count G
// divides; 1 = the one-slice kernel.
int fwd_split_slice_group(int nd, int Nx, int Ne_pad, int Ney, int Np) {
  if (nd != 3 || Nx > SEGH || (Np != 2 && Np != 4)) return 1;
  const int G = Ne_pad == 32 ? 4 : (Ne_pad == 16 ? 8 : 1);
  if (G == 1 || Ney % G != 0) return 1;
  return G;
}

void launch_fwd_split_slices(const LaunchSet& L, int n_cu, hipStream_t s) {
  const int grid = std::max(1, std::min(L.n_items_s, (L.grid_mult > 0 ? L.grid_mult : 2) * n_cu));
  bool even = true;
  for (const Inst& I : L.inst) even = even && (I.Nx % 2) == 0;
  const int n_inst = (int)L.inst.size();
  const int G = L.inst[0].sgrp;
#define BORN_S3_K(MTSv, Gv, EV, NPv, F16v)                                                                  \
  hipLaunchKernelGGL((k_fwd_split_slices<MTSv, Gv, EV, NPv, F16v>), dim3(grid), dim3(256),                   \
                     (fwd_split3_lds_bytes<MTSv, Gv, NPv>()), s, L.d_inst, n_inst, L.n_items_s, L.d_chunks,  \
                     L.d_chunk_pairs)
#define BORN_S3_F(MTSv, Gv, EV, NPv)                                         \
  do {                                                                       \
    if (L.f16) BORN_S3_K(MTSv, Gv, EV, NPv, true);                           \
    else BORN_S3_K(MTSv, Gv, EV, NPv, false);                                \
  } while (0)
#define BORN_S3_NP(MTSv, Gv, EV)                     \
  if (L.Np == 2) BORN_S3_F(MTSv, Gv, EV, 2);         \
  else BORN_S3_F(MTSv, Gv, EV, 4);
#define BORN_S3_E(MTSv, Gv)                          \
  if (even) {                                        \
    BORN_S3_NP(MTSv, Gv, true)                       \
  } else {                                           \
    BORN_S3_NP(MTSv, Gv, false)                      \
  }
  if (G == 4)
    BORN_S3_E(2, 4)
  else
    BORN_S3_E(1, 8)
#undef BORN_S3_E
#undef BORN_S3_NP
#undef BORN_S3_F
#undef BORN_S3_K
}

template <int MTS, int G, int NP>
static void set_max_lds_split3_mt(int bytes) {
#define BORN_SET(K) hipFuncSetAttribute((const void*)(K), hipFuncAttributeMaxDynamicSharedMemorySize, bytes)
  BORN_SET((k_fwd_split_slices<MTS, G, true, NP, false>));
  BORN_SET((k_fwd_split_slices<MTS, G, false, NP, false>));
  BORN_SET((k_fwd_split_slices<MTS, G, true, NP, true>));
  BORN_SET((k_fwd_split_slices<MTS, G, false, NP, true>));
#undef BORN_SET
}

void set_max_lds_split3(int bytes) {
  hipFuncSetAttribute((const void*)(k_adj_split_slices<2, false>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  hipFuncSetAttribute((const void*)(k_adj_split_slices<2, true>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  hipFuncSetAttribute((const void*)(k_adj_split_slices<4, false>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  hipFuncSetAttribute((const void*)(k_adj_split_slices<4, true>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  set_max_lds_split3_mt<2, 4, 2>(bytes);
  set_max_lds_split3_mt<2, 4, 4>(bytes);
  set_max_lds_split3_mt<1, 8, 2>(bytes);
  set_max_lds_split3_mt<1, 8, 4>(bytes);
}

}  // namespace born
