// Internal plan layout and kernel launchers of libborn (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

namespace born {

// Column tiles of the grid contraction: one chunk is (frequency, rhs, <=16
// measurement/frequency pairs).
constexpr int NC = 16;

struct Chunk {
  int fg;         // global unique frequency index
  int r;          // right-hand side
  int col_start;  // into col_pair
  int ncols;      // <= NC
};

// Pointers that kernels read out of device memory are marked as global
// memory in device code, so loads through them are global (not flat) loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define BORN_GLOBAL __attribute__((address_space(1)))
#else
#define BORN_GLOBAL
#endif

// One operator instance of a structured launch.  A launch covers one plan
// or a batch of plans that share the element padding and the points per
// pitch; every per-instance size and pointer comes from this record.
struct Inst {
  BORN_GLOBAL const float2* h;
  BORN_GLOBAL const float2* P;
  BORN_GLOBAL const float2* gm;
  BORN_GLOBAL const int* ext;
  BORN_GLOBAL const float2* Wc;
  BORN_GLOBAL float2* vpart;
  int Nx, Nz, N_pts, N_e, Np, N_ext, Hp, HR;
  int n_chunks;
  int max_cols;             // widest chunk of the instance (columns)
  int chunk_base;           // first chunk of the instance in the launch's chunk tables
  int Ney, n_rows;          // y-element slices, rows (Ny * Nz)
  BORN_GLOBAL const int* hrow;   // [Ney][n_rows] table row of (slice, row)
  int NZB, ZR, S;           // plain forward: z-blocks of ZR rows, S = ceil(Nx / 16)
  int NZB_s, ZR_s, nseg_s;  // split forward: z-blocks x 256-point segments
  int sgrp;                 // split forward: y slices per work item (k_fwd_split_slices when > 1)
  int sgrp_p;               // plain forward: y slices per work item (k_fwd_s3d when > 1)
  int item_begin, item_begin_s;  // first forward work item of the instance (plain, split)
  int wg_begin, wg_begin_s;      // first adjoint workgroup of the instance (plain, split)
  int adj_segs, adj_segs_s;      // adjoint segments per row (plain, split)
  int wg_begin_f;                // first adjoint workgroup of the instance (few-column kernels)
  int fc_g;                      // few-column adjoint: chunk groups (partial sums summed after)
  int fc_bt, fc_nr;              // few-column adjoint: tasks per workgroup, most rows it stages
  int fc_cb;                     // few-column adjoint: chunks staged per step
  int n_out;                     // right-hand sides the adjoint writes
  BORN_GLOBAL float2* gpart;     // few-column adjoint partials [fc_g][n_out][N_pts]
  // adjoint chunk columns: col_rhs = 0: every column of a chunk belongs to
  // the chunk's rhs; col_rhs = k (power of 2): column c belongs to rhs
  // c mod k (< n_rhs), and gamma_hat of all right-hand sides is written once
  // after the last chunk
  int col_rhs, n_rhs;
  // FP16 split operand scales (powers of 2): [0] h_ref, [1] p_inc * gamma,
  // [2] forward unscale 1 / ([0] [1]), [3] adjoint weights, [4] adjoint
  // unscale 1 / ([0] [3])
  BORN_GLOBAL const float* fscale;
};

// Work items of the 3-D forwards over several y slices per item: instance-
// local index lid -> (chunk ci, slice group sg of NSG, row block zb of NZB).
// The NSG items of one (chunk, row block) read the same p_inc * gamma rows;
// they lie XCDS apart in item order, so on a persistent grid that is a
// multiple of XCDS (workgroups are dispatched round-robin over the XCDs)
// they run at the same time on one XCD and share its L2.  Row blocks past
// the last multiple of XCDS follow with their slice groups adjacent.
constexpr int XCDS = 8;
__device__ __forceinline__ void slice_item(int lid, int NSG, int NZB, int& ci, int& sg, int& zb) {
  ci = lid / (NSG * NZB);
  const int r = lid - ci * (NSG * NZB);
  const int NZB8 = NZB - NZB % XCDS;
  if (r < NSG * NZB8) {
    const int t = r / XCDS;
    sg = t % NSG;
    zb = (t / NSG) * XCDS + (r - t * XCDS);
  } else {
    const int r2 = r - NSG * NZB8;
    sg = r2 % NSG;
    zb = NZB8 + r2 / NSG;
  }
}

constexpr int MAX_BATCH = 64;
struct OutPtrs {  // adjoint outputs, passed by value
  float2* p[MAX_BATCH];
};

// Launch parameters of one structured launch over instances.
struct LaunchSet {
  Inst* d_inst = nullptr;         // device copy of inst
  std::vector<Inst> inst;
  // chunk descriptors and column rows of all instances, concatenated
  // (kernel arguments, so uniform reads of them are scalar loads)
  const Chunk* d_chunks = nullptr;
  const int* d_chunk_pairs = nullptr;
  bool owns_chunks = false;
  int MT = 1, Np = 4;
  int S_max = 1;                  // largest S (plain forward LDS and registers)
  int adj_mw = 4;                 // plain adjoint segment = 64 * adj_mw points
  bool stk = false;               // plain kernels: chunks of at most 8 columns (stacked B operand)
  int n_items = 0, n_items_s = 0;
  int n_wg = 0, n_wg_s = 0;
  int grid_mult = 0;              // persistent forward workgroups per CU (0: the kernel's default)
  bool f16 = false;               // split kernels: FP16 (true) or BF16 hi/lo operands
  int fc_w = 0;                   // > 0: chunks of at most fc_w columns run on the few-column kernels
  bool fc_allowed = true;         // false: MFMA kernels only (the allocation-free layout)
  int n_wg_f = 0;                 // adjoint workgroups of the few-column kernel
};

// Device tables and index maps.  All device arrays are owned by the plan.
struct Plan {
  int device = 0;
  int n_cu = 256;           // compute units of the device (persistent grids)
  int precision = 0;
  bool force_general = false;
  bool structured = false;  // 2-D lateral Toeplitz path usable

  int nd = 2;
  int Nx = 0, Ny = 1, Nz = 0;
  int N_pts = 0;
  int N_e = 0, Nex = 0, Ney = 1;
  int N_fu = 0;
  // h_ref holds H_plane entries per frequency: the FOV grid (N_pts, the
  // reference's mirrored layout) or, for a FOV narrower than the aperture,
  // the grid extended laterally by (N_elements_axis - 1) pitches, indexed by
  // lateral offset.  HR = H_plane / Nz entries per depth row.
  int H_plane = 0, HR = 0;
  int HY = 1;       // x-rows of h_ref per depth plane (3-D: table grid Ny)
  bool extended = false;
  // structured path: element e = ex + Nex * ey sees point (ix, iy, iz) at
  // h_ref entry extx[ix - ex*Np + (Nex-1)*Np] + HR * (exty[iy - ey*Npy + (Ney-1)*Npy] + HY * iz).
  // A "row" is (iy, iz); hrow[ey][row] = exty[...] + HY * iz is its table row.
  int Np = 1;       // grid points per pitch along x
  int Npy = 1;      // along y (3-D)
  int N_ext = 0;    // extended lateral length Nx + (Nex-1)*Np
  int n_rows = 0;   // Ny * Nz
  std::vector<int> extx, exty;
  int* d_hrow = nullptr;          // [Ney][n_rows]

  // (measurement, unique-frequency-of-measurement) pairs: one p_inc row each
  int N_pairs = 0;
  std::vector<int> pair_fg, pair_meas;
  std::vector<int> meas_pair_base;  // first pair of each measurement
  std::vector<int> meas_Nfm;
  // pairs grouped by global frequency
  std::vector<int> fg_pair_off, fg_pairs;

  // mixes, flattened over (measurement, mix)
  int N_mk = 0;
  int64_t N_obs = 0;
  std::vector<int> mk_meas, mk_Nf, mk_Nact;
  std::vector<int64_t> mk_obs_off, mk_f_off, mk_a_off, mk_pref_off;
  std::vector<int> fmeas;   // 0-based measurement-unique freq per (mk, f)
  std::vector<int> act;     // 0-based element per (mk, a)
  std::vector<double> pref; // complex (re, im) per (mk, f, a), f fastest

  // device copies
  float2* d_h = nullptr;          // [N_fu][H_plane]
  float2* d_p = nullptr;          // [N_pairs + 1][N_pts], last row zero
  int* d_shift = nullptr;         // [N_e][N_pts], 0-based (general path)
  int* d_ext = nullptr;           // [N_ext] lateral x map (structured path)
  int* d_exty = nullptr;          // [Ny + (Ney-1)*Npy] lateral y map (3-D structured path)
  int* d_pair_fg = nullptr;
  int* d_pair_meas = nullptr;
  int* d_obs_mk = nullptr;        // [N_obs]
  int* d_obs_f = nullptr;         // [N_obs]
  int* d_mk_meta = nullptr;       // [N_mk][4]: meas, Nf, Nact, pad
  int64_t* d_mk_off = nullptr;    // [N_mk][4]: obs_off, f_off, a_off, pref_off
  int* d_fmeas = nullptr;
  int* d_act = nullptr;
  double2* d_pref = nullptr;
  int* d_meas_pair_base = nullptr;
  // adjoint un-mixing gather lists: for each pair, the (mk, f) that read it
  int64_t* d_pe_off = nullptr;    // [N_pairs * N_e + 1]: un-mixing entries of (pair, element)
  int2* d_pe_ent = nullptr;       // (observation, prefactor index)
  std::vector<int> h_shift;       // host copy of the 0-based shift map [N_e][N_pts]

  // per-N_rhs workspace
  int ws_rhs = -1;
  bool ws_split = false;          // built for a split precision (the forward's expansion rule)
  uint64_t ws_gen = 0;            // counts workspace builds: batches holding its pointers compare it
  std::vector<Chunk> chunks;
  std::vector<int> col_pair;
  Chunk* d_chunks = nullptr;
  int* d_col_pair = nullptr;
  int* d_chunk_pairs = nullptr;   // [n_chunks][NC]: pair of each column, N_pairs (zero row) past ncols
  int* d_pr_chunk = nullptr;      // [N_pairs*N_rhs]: chunk of (pair, r)
  int* d_pr_pos = nullptr;        // [N_pairs*N_rhs]: column position in chunk
  int NZB = 1, ZR = 1;            // forward z-blocking (plain FP32 kernel: whole rows)
  int NZB_s = 1, ZR_s = 1, nseg_s = 1;  // split forward: z-blocks x 256-point segments
  int vblk = 1;                   // partial-sum blocks per chunk allocated in vpart
  int Ne_pad = 16;
  float2* d_vpart = nullptr;      // [n_chunks][vblk][Ne_pad][NC]
  float2* d_gpart = nullptr;      // few-column adjoint partials [FC_GROUPS][N_rhs][N_pts]
  double2* d_V = nullptr;         // [N_pairs][N_e][N_rhs]
  float2* d_W = nullptr;          // [N_pairs][N_e][N_rhs] (general path)
  float2* d_Wc = nullptr;         // [n_chunks][Ne_pad][NC] (structured path; unused slots stay zero)
  float2* d_gm = nullptr;         // masked gamma [N_rhs][N_pts]
  int items_target = 8192;        // forward work items per launch (tunable)
  LaunchSet ls;                   // structured launch of this plan alone
  // FP16 split form: operand scales [5], table maxima [2] (h_ref, p_inc),
  // block partial maxima [256]
  float* d_fscale = nullptr;
  float* d_f16max = nullptr;
  float* d_f16part = nullptr;
  bool f16_maxima = false;        // table maxima computed
  // Forward with several right-hand sides and fewer than NC pairs per
  // frequency: the products p_inc * gamma of every (pair, rhs) column are
  // formed first (d_pexp, a zero row after them) and the forward runs on
  // chunks that pack the columns of all right-hand sides of a frequency.
  bool fwd_expand = false;         // the forward runs on ls_fwd (below)
  bool fwd_table_expanded = false; // ls_fwd reads d_pexp (else d_p with the rhs of each chunk)
  std::vector<Chunk> fchunks;     // chunks over expanded columns ep = pair * N_rhs + r
  std::vector<int> fcol;          // expanded column of each chunk column
  Chunk* d_fchunks = nullptr;
  int* d_fcol = nullptr;
  int* d_fchunk_pairs = nullptr;  // [n_fchunks][NC]: expanded column, N_pairs * N_rhs (zero row) past ncols
  float2* d_pexp = nullptr;       // [N_pairs * N_rhs + 1][N_pts]
  float2* d_ones = nullptr;       // [N_pts] of 1
  float2* d_fvpart = nullptr;     // [n_fchunks][fvblk][Ne_pad][NC]
  int fvblk = 1;
  LaunchSet ls_fwd;               // forward launch over the expanded columns
  // host forward: the gather path for at most sparse_max occupied points
  // (-1: the cost rule, 0: never)
  int sparse_max = -1;
  // time gain compensation: per mix the half length M (after the cut-off
  // truncation) and the offset of its 2M + 1 taps; tiles (mix, first row) of
  // 256 outputs
  bool tgc = false;
  int n_tgc_tiles = 0;
  int2* d_tgc_tiles = nullptr;
  int* d_tgc_half = nullptr;
  int64_t* d_tgc_off = nullptr;
  float2* d_tgc_k32 = nullptr;
  double2* d_tgc_k64 = nullptr;
  size_t io_bytes = 0;
  void* d_io_in = nullptr;        // host-path staging
  void* d_io_out = nullptr;
  hipStream_t stream = nullptr;   // host-path stream

  // Allocation-free apply (born_apply_dev_noalloc): a one-rhs chunk layout
  // built at plan creation.  Its scratch -- masked gamma, per-element sums
  // V, un-mixed weights Wc, forward partial sums -- lives in the library's
  // static device array; the members carry the names of the plan's own
  // workspace so that an apply can swap the two.
  struct NoAlloc {
    bool ok = false;
    std::string why;                // the reason when !ok
    size_t bytes = 0;               // scratch bytes the layout needs
    size_t Wc_n = 0;                // float2 entries of Wc (cleared on every adjoint)
    std::vector<Chunk> chunks;
    std::vector<int> col_pair;
    Chunk* d_chunks = nullptr;
    int* d_col_pair = nullptr;
    int* d_chunk_pairs = nullptr;
    int* d_pr_chunk = nullptr;
    int* d_pr_pos = nullptr;
    int NZB = 1, ZR = 1, NZB_s = 1, ZR_s = 1, nseg_s = 1, vblk = 1;
    float2* d_gm = nullptr;         // in the scratch array
    double2* d_V = nullptr;
    float2* d_Wc = nullptr;
    float2* d_vpart = nullptr;
    LaunchSet ls;
    int ws_rhs = 1;
    bool fwd_expand = false, fwd_table_expanded = false;
  } na;

  // optional event timing of the contraction kernels: (start, stop, kind)
  bool timing = false;
  struct Timed {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Timed> timed;

  ~Plan();
};

// ------------------------------------------------------------------ kernels
void launch_convert_transpose(const double2* src, int rows, int cols, float2* dst, hipStream_t s);
void launch_f64_to_f32(const double2* src, size_t n, float2* dst, hipStream_t s);
void launch_mask_gamma(const float2* g, int N_pts, int N_rhs, float2* gm, hipStream_t s);
void launch_zero_f2(float2* x, size_t n, hipStream_t s);
// 3-D forwards over several y slices per work item (born_split3d.hip,
// born_structured3d.hip)
int fwd_split_slice_group(int nd, int Nx, int Ne_pad, int Ney, int Np);
int fwd_plain_slice_group(int nd, int Nx, int Ne_pad, int Ney, int Np);
void launch_fwd_s3d(const LaunchSet& L, int n_cu, hipStream_t s);
void launch_fwd_split_slices(const LaunchSet& L, int n_cu, hipStream_t s);
void launch_adj_split_slices(const LaunchSet& L, const OutPtrs& outs, hipStream_t s);
void set_max_lds_split3(int bytes);

void launch_fwd_structured(const LaunchSet& L, int n_cu, hipStream_t s);
// chunks of at most 4 columns on the vector ALUs (born_fewcol.hip)
int fewcol_width(int max_cols);
bool fewcol_fits(int Nx, int Np, int Ne_pad, int fc_w, int n_rows);
void fewcol_adj_block(int Nx, int Np, int Ne_pad, int fc_w, int col_rhs, int n_rows, int& BT, int& NR);
int fewcol_adj_tpr(int Nx, int Np, int Ne_pad, int fc_w);
int fewcol_adj_cb(int Nx, int Np, int Ne_pad, int fc_w, int col_rhs, int NR, int per_group, int Ney);
int fewcol_fwd_rows(int Nx, int Np, int Ne_pad, int fc_w);
constexpr int FC_GROUPS = 8;    // most chunk groups of the few-column adjoint
void set_max_lds_fc(int bytes);
void launch_fwd_fc(const LaunchSet& L, hipStream_t s);
void launch_adj_fc(const LaunchSet& L, const OutPtrs& outs, hipStream_t s);
void launch_fwd_general(const Plan& P, const float2* gm, int N_rhs, hipStream_t s);
void launch_fwd_reduce(const Plan& P, int N_rhs, int nblk, hipStream_t s);
void launch_fwd_reduce_expanded(const Plan& P, int nblk, hipStream_t s);
void launch_expand_pg(const Plan& P, hipStream_t s);
void launch_energy_rx(const Plan& P, int m, double* E, hipStream_t s);
void launch_fp16_scales(const Plan& P, const float* x, size_t n_floats, int which, hipStream_t s);
void launch_time_signals(const Plan& P, const float2* u, const int* q_obs, int N_rhs, int N_t, float* rf,
                         hipStream_t s);
void launch_fwd_mix(const Plan& P, int N_rhs, void* u, bool out_f64, hipStream_t s);
void launch_tgc(const Plan& P, bool adjoint, bool f64, const void* in, void* out, int N_rhs, hipStream_t s);
void launch_columns(const Plan& P, const int* pts, int n, void* u, bool out_f64, hipStream_t s);
void launch_fwd_sparse(const Plan& P, const int* pts, int n, const double2* g, int N_rhs, double2* u, hipStream_t s);
// received energy under the plan's TGC; mixes of up to ETGC_LDS_ROWS rows keep
// their rows in LDS, longer ones in `spill` (spill_blocks x rows x 64 float2)
constexpr int ETGC_LDS_ROWS = 304;   // 304 rows x 64 points x 8 B = 152 KiB
void launch_energy_rx_tgc(const Plan& P, int m, int rows, float2* spill, int spill_blocks, double* E,
                          hipStream_t s);

void launch_adj_unmix(const Plan& P, const float2* u, int N_rhs, bool chunked, hipStream_t s);
void launch_adj_structured(const LaunchSet& L, const OutPtrs& outs, hipStream_t s);
void launch_adj_general(const Plan& P, int N_rhs, float2* g_out, hipStream_t s);

struct GeomDev {
  int nd;
  int Nx, Ny, Nz;          // FOV grid
  int Nx_tab, Ny_tab;      // lateral size of the table grid (FOV grid plus ox, oy points before it)
  int ox, oy;
  double dx, dy, dz, z0;
  int Nex, Ney;
  double pitch_x, pitch_y;
  int Nfx, Nfy;
  double dfx, dfy;
  int filt_kind;
  double filt_param;
};
void launch_build_h(const GeomDev& g, int N_fu, const double2* k, const double2* face_w,
                    double2* h64, hipStream_t s);
void launch_synth_pinc(const Plan& P, const double2* h64, const int* d_meas_Nfm, const int* d_tx_off,
                       const int* d_tx_el, const double2* d_vd, const int* d_vd_off, hipStream_t s);
void launch_synth_pinc_structured(const Plan& P, const double2* h64, const int* d_meas_Nfm, const int* d_tx_off,
                                  const int* d_tx_el, const double2* d_vd, const int* d_vd_off, hipStream_t s);
void set_max_lds(size_t bytes);
bool fwd_structured_fits(int Nx, int Ne_pad, int Np);
bool fwd_split_supported(const Plan& P);
void launch_fwd_split(const LaunchSet& L, int n_cu, hipStream_t s);
void launch_adj_split(const LaunchSet& L, const OutPtrs& outs, hipStream_t s);
int fwd_ppt(int S);              // B-tile points per thread of the plain forward
int adj_mw_default(int Nx);
void set_max_lds_split(int bytes);
bool adj_structured_fits(int Nx, int Ne_pad, int Np);

}  // namespace born
