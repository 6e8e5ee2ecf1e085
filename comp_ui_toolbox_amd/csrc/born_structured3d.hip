// Plain FP32 forward contraction of 3-D plans with short rows (gfx950).
//
// The forward of one (frequency, row (y, z)) is, for every y slice ey, the
// Toeplitz-in-x product of k_fwd_s2d (born_structured.hip),
//     V[ey][ex][c] += sum_i Hext_ey[i - ex*Np + off] * B[i][c],
// with B[i][c] = p_inc[pair_c][i] gamma_r[i] the same for every slice.  With
// rows of at most 128 points and 32 elements per slice, one slice per step
// leaves each wave one A and one B LDS read per three MFMAs.  Here a work
// item covers G = 4 slices: the row's B tile is staged once and the 4
// slices' polyphase windows side by side make the M dimension (128
// elements), so per K-step a wave reads one B entry for its 4 tiles and every
// barrier pair serves 4 times the MFMAs.  Numeric form as k_fwd_s2d: 16x16x4
// FP32 MFMA with the three-multiplication complex product, partial sums per
// (chunk, slice, row block) summed in FP64 by k_fwd_reduce.
#include "born_internal.h"

namespace born {

typedef float v4f __attribute__((ext_vector_type(4)));

namespace {
constexpr int NCS3 = NC + 1;       // padded LDS stride of a 16-column row
constexpr int XPH = 128;           // rows of at most 128 points
constexpr int MTS3 = 2;            // element tiles of one slice (32 elements)
constexpr int G3 = 4;              // slices per work item

__device__ __forceinline__ v4f mfma16s(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f vz3() { return v4f{0.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ float2 cmul3(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
}  // namespace

// Work item = (instance, chunk, slice group sg, row block zb); persistent
// workgroups take items blockIdx.x, blockIdx.x + gridDim.x, ...  A step is
// one row (y, z).  Waves: 2 element groups x 2 point groups; wave (eg, pg)
// holds tiles [4 eg, 4 eg + 4) -- slices 2 eg, 2 eg + 1 of the group, two
// tiles each -- and the K-steps [pg SW, (pg + 1) SW) of the row.
//
// LDS (float2): Hp[G][Np][QL] the slices' polyphase extended rows, Bs[XP][NCS]
// the row's B tile.
template <int NP>
__global__ __launch_bounds__(256, 2) void k_fwd_s3d(const Inst* __restrict__ inst, int n_inst, int n_items,
                                                    const Chunk* __restrict__ chunks,
                                                    const int* __restrict__ chunk_pairs) {
  extern __shared__ float2 lds[];
  constexpr int NEPS = 16 * MTS3;
  constexpr int QL = (XPH - 1) / NP + NEPS;           // polyphase row length of a slice (for XP = 128)
  constexpr int HN = NP * QL;                         // window entries of a slice
  constexpr int HPT = (G3 * HN + 255) / 256;          // window entries per thread
  constexpr int EW = 2, PW = 2, MTW = MTS3 * G3 / EW; // 4 tiles per wave
  constexpr int A_STEP = 4 / NP;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int eg = w % EW, pg = w / EW;
  const int kk = lane >> 4, col = lane & 15;
  const int n_my = (n_items - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  if (n_my <= 0) return;

  int k = 0;
  Inst Iv = inst[0];
  const Inst* I = &Iv;
  int eidx[HPT];                  // table x index of window entry tid + 256 j (-1: outside)
  int eslc[HPT];                  // its slice within the group
  int fg, rr, pr[NC];
  int ci, sg, zb;
  int z, z_end, item = blockIdx.x;

#define BORN_F3_INST()                                                                    \
  {                                                                                       \
    while (k + 1 < n_inst && item >= inst[k + 1].item_begin) ++k;                         \
    Iv = inst[k];                                                                         \
    const int Qoff_ = NEPS - I->N_e;                                                      \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) {                                    \
      const int idx_ = tid + 256 * j;                                                     \
      eidx[j] = -1;                                                                       \
      eslc[j] = idx_ / HN;                                                                \
      if (idx_ < G3 * HN) {                                                               \
        const int e_ = idx_ - eslc[j] * HN;                                               \
        const int ph_ = e_ / QL, q_ = e_ - ph_ * QL;                                      \
        const int t_ = (q_ - Qoff_) * NP + ph_;                                           \
        if (t_ >= 0 && t_ < I->N_ext) eidx[j] = I->ext[t_];                               \
      } else {                                                                            \
        eslc[j] = 0;                                                                      \
      }                                                                                   \
    }                                                                                     \
  }
#define BORN_F3_DESC()                                                                    \
  {                                                                                       \
    slice_item(item - I->item_begin, I->Ney / G3, I->NZB, ci, sg, zb);                   \
    const Chunk c_ = chunks[I->chunk_base + ci];                                          \
    fg = c_.fg;                                                                           \
    rr = c_.r;                                                                            \
    _Pragma("unroll") for (int q = 0; q < NC; ++q) pr[q] = chunk_pairs[(I->chunk_base + ci) * NC + q]; \
    z = zb * I->ZR;                                                                       \
    z_end = min(I->n_rows, z + I->ZR);                                                    \
  }

  float2 hreg[HPT];
  float2 praw[NC / 2];
  float2 greg;

#define BORN_F3_PREFETCH(zz)                                                              \
  {                                                                                       \
    _Pragma("unroll") for (int j = 0; j < HPT; ++j) {                                    \
      const int row_ = I->hrow[(sg * G3 + eslc[j]) * I->n_rows + (zz)];                    \
      hreg[j] = I->h[(size_t)fg * I->Hp + (size_t)row_ * I->HR + max(eidx[j], 0)];        \
    }                                                                                     \
    /* B tile: thread = (column half tid / 128, point tid % 128) */                       \
    const size_t o_ = (size_t)(zz) * I->Nx + min(tid & (XPH - 1), I->Nx - 1);             \
    greg = I->gm[(size_t)rr * I->N_pts + o_];                                             \
    _Pragma("unroll") for (int c_ = 0; c_ < NC / 2; ++c_) {                               \
      const int q_ = tid >= XPH ? pr[NC / 2 + c_] : pr[c_];                               \
      praw[c_] = I->P[(size_t)q_ * I->N_pts + o_];                                        \
    }                                                                                     \
  }

  BORN_F3_INST()
  BORN_F3_DESC()
  BORN_F3_PREFETCH(z)

  v4f SR[MTW], SI[MTW];
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt) {
    SR[mt] = vz3();
    SI[mt] = vz3();
  }
  const int b_base = kk * NCS3 + col;
  // A position of lane (kk, col) in a slice's polyphase row for tile 0, K-step 0
  const int a_base = (kk % NP) * QL + kk / NP + NEPS - 1 - col;
  int it = 0;

  while (true) {
    float2* Hp = lds;                  // [G][HN]
    float2* Bs = lds + G3 * HN;        // [XPH][NCS]
    const int nx = I->Nx;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < HPT; ++j) {
      const int idx = tid + 256 * j;
      if (idx < G3 * HN) Hp[idx] = eidx[j] >= 0 ? hreg[j] : make_float2(0.f, 0.f);
    }
    {
      const int x = tid & (XPH - 1), c0 = (tid / XPH) * (NC / 2);
      const float2 g = x < nx ? greg : make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < NC / 2; ++c) Bs[x * NCS3 + c0 + c] = cmul3(praw[c], g);
    }
    __syncthreads();

    bool has_next = true;
    const int ci_cur = ci, sg_cur = sg, zb_cur = zb;
    float2* const vpart_cur = Iv.vpart;
    const int Ney_cur = I->Ney, NZB_cur = I->NZB;
    const bool item_done = z + 1 >= z_end;
    if (!item_done) {
      ++z;
    } else if (it + 1 >= n_my) {
      has_next = false;
    } else {
      item = blockIdx.x + (it + 1) * gridDim.x;
      if (k + 1 < n_inst && item >= inst[k + 1].item_begin) BORN_F3_INST()
      BORN_F3_DESC()
    }
    if (has_next) BORN_F3_PREFETCH(z)

    // MFMA over this wave's K-steps of the row: one B read, four A reads, twelve MFMAs
    constexpr int SW = XPH / (4 * PW);
    v4f P1[MTW], P2[MTW], P3[MTW];
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt) {
      P1[mt] = vz3();
      P2[mt] = vz3();
      P3[mt] = vz3();
    }
#pragma unroll 4
    for (int s = 0; s < SW; ++s) {
      const int ks = pg * SW + s;
      const float2 b = Bs[b_base + 4 * ks * NCS3];
      const float bsum = b.x + b.y;
#pragma unroll
      for (int mt = 0; mt < MTW; ++mt) {
        const int m = eg * MTW + mt, sl = m / MTS3, lt = m - sl * MTS3;
        const float2 a = Hp[sl * HN + a_base + A_STEP * ks - 16 * lt];
        const float asum = a.x + a.y;
        P1[mt] = mfma16s(a.x, b.x, P1[mt]);
        P2[mt] = mfma16s(a.y, b.y, P2[mt]);
        P3[mt] = mfma16s(asum, bsum, P3[mt]);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt) {
      SR[mt] += P1[mt] - P2[mt];
      SI[mt] += P3[mt] - P1[mt] - P2[mt];
    }

    if (item_done) {
      __syncthreads();
      float* red = reinterpret_cast<float*>(lds);   // [w][mtw][j][lane]
#pragma unroll
      for (int mt = 0; mt < MTW; ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          red[((w * MTW + mt) * 8 + v) * 64 + lane] = SR[mt][v];
          red[((w * MTW + mt) * 8 + 4 + v) * 64 + lane] = SI[mt][v];
        }
        SR[mt] = vz3();
        SI[mt] = vz3();
      }
      __syncthreads();
      // each slice of the group to its own block [(chunk * Ney + ey) * NZB + zb]
      for (int idx = tid; idx < MTS3 * G3 * 512; idx += 256) {
        const int mt = idx >> 9, e_g = mt / MTW, mtw = mt - e_g * MTW, rest = idx & 511;
        float sum = 0.f;
#pragma unroll
        for (int p_g = 0; p_g < PW; ++p_g) sum += red[((p_g * EW + e_g) * MTW + mtw) * 512 + rest];
        const int sl = mt / MTS3, lt = mt - sl * MTS3;
        const int ey = sg_cur * G3 + sl;
        float* out = reinterpret_cast<float*>(
            vpart_cur + ((size_t)(ci_cur * Ney_cur + ey) * NZB_cur + zb_cur) * NEPS * NC);
        const int ln = idx & 63, j = (idx >> 6) & 7;
        const int e = 16 * lt + 4 * (ln >> 4) + (j & 3);
        out[((size_t)e * NC + (ln & 15)) * 2 + (j >> 2)] = sum;
      }
      ++it;
    }
    if (!has_next) break;
  }
#undef BORN_F3_PREFETCH
#undef BORN_F3_DESC
#undef BORN_F3_INST
}

int fwd_plain_slice_group(int nd, int Nx, int Ne_pad, int Ney, int Np) {
  return nd == 3 && Nx <= XPH && (Np == 2 || Np == 4) && Ne_pad == 16 * MTS3 && Ney % G3 == 0 ? G3 : 1;
}

template <int NP>
static size_t fwd_s3d_lds_bytes() {
  constexpr int HN = NP * ((XPH - 1) / NP + 16 * MTS3);
  const size_t body = ((size_t)G3 * HN + (size_t)XPH * NCS3) * sizeof(float2);
  const size_t red = (size_t)4 * (MTS3 * G3 / 2) * 8 * 64 * sizeof(float);
  return std::max(body, red);
}

void launch_fwd_s3d(const LaunchSet& L, int n_cu, hipStream_t s) {
  const int grid = std::max(1, std::min(L.n_items, (L.grid_mult > 0 ? L.grid_mult : 2) * n_cu));
  if (L.Np == 2)
    hipLaunchKernelGGL((k_fwd_s3d<2>), dim3(grid), dim3(256), fwd_s3d_lds_bytes<2>(), s, L.d_inst,
                       (int)L.inst.size(), L.n_items, L.d_chunks, L.d_chunk_pairs);
  else
    hipLaunchKernelGGL((k_fwd_s3d<4>), dim3(grid), dim3(256), fwd_s3d_lds_bytes<4>(), s, L.d_inst,
                       (int)L.inst.size(), L.n_items, L.d_chunks, L.d_chunk_pairs);
}

}  // namespace born
