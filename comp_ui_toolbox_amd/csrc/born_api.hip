// C ABI of libborn: plan management, argument validation and dispatch.
//
// Replaces the MEX gateway combined_quick_gpu (combined_quick_gpu.cu:66-1094):
// the validation and error identifiers follow cu:201-441 and h:31-44, the
// observation ordering cu:773/:978, the forward occupancy rule cu:454-499.
// Tables are converted and uploaded once per plan instead of once per call
// (cu:33 TODO, cu:548-593, :817-847).
#include "born.h"
#include "born_internal.h"

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace born {

static thread_local std::string g_last_error;

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define HIP_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw Error(BORN_ERR_DEVICE, std::string("HIP error '") + hipGetErrorString(e_) + "' in " + \
                                       #expr + " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

// live device allocations of the library (error-path tests check it returns
// to its value before a failed call)
static std::atomic<int64_t> g_live_alloc{0};

template <typename T>
static void dfree(T*& p) {
  if (p) {
    hipFree((void*)p);
    g_live_alloc.fetch_sub(1);
  }
  p = nullptr;
}

// Error-path testing: the g_fail_alloc-th device allocation from now fails
// (0: the next one); negative: off.
static std::atomic<int64_t> g_fail_alloc{-1};

template <typename T>
static T* dalloc(size_t n) {
  if (n == 0) n = 1;
  if (g_fail_alloc.load() >= 0 && g_fail_alloc.fetch_sub(1) == 0)
    throw Error(BORN_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(n * sizeof(T)) +
                                            " bytes failed (injected)");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) {
    hipGetLastError();
    throw Error(BORN_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  g_live_alloc.fetch_add(1);
  return (T*)p;
}

// Small index tables: synchronous copies, so the host vectors may die on return.
template <typename T>
static T* dupload(const std::vector<T>& v, hipStream_t) {
  T* p = dalloc<T>(v.size());
  try {
    if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  } catch (...) {
    dfree(p);
    throw;
  }
  return p;
}

// p_inc rows of all pairs plus one zero row (index N_pairs) that unused
// columns of the structured kernels read.
static float2* alloc_p_table(const Plan& P) {
  float2* d = dalloc<float2>((size_t)(P.N_pairs + 1) * P.N_pts);
  try {
    HIP_CHECK(hipMemset(d + (size_t)P.N_pairs * P.N_pts, 0, (size_t)P.N_pts * sizeof(float2)));
  } catch (...) {
    dfree(d);
    throw;
  }
  return d;
}

static void free_workspace(Plan& P) {
  dfree(P.d_chunks);
  dfree(P.d_col_pair);
  dfree(P.d_chunk_pairs);
  dfree(P.d_pr_chunk);
  dfree(P.d_pr_pos);
  dfree(P.d_vpart);
  dfree(P.d_gpart);
  dfree(P.d_V);
  dfree(P.d_W);
  dfree(P.d_Wc);
  dfree(P.d_gm);
  dfree(P.d_fchunks);
  dfree(P.d_fcol);
  dfree(P.d_fchunk_pairs);
  dfree(P.d_pexp);
  dfree(P.d_ones);
  dfree(P.d_fvpart);
  P.fwd_expand = false;
  P.fwd_table_expanded = false;
  P.fchunks.clear();
  P.fcol.clear();
  P.ws_rhs = -1;
  P.chunks.clear();
  P.col_pair.clear();
}

Plan::~Plan() {
  hipSetDevice(device);
  if (stream) hipStreamSynchronize(stream);
  free_workspace(*this);
  dfree(na.d_chunks);
  dfree(na.d_col_pair);
  dfree(na.d_chunk_pairs);
  dfree(na.d_pr_chunk);
  dfree(na.d_pr_pos);
  dfree(na.ls.d_inst);
  dfree(ls.d_inst);
  dfree(ls_fwd.d_inst);
  dfree(d_fscale);
  dfree(d_f16max);
  dfree(d_f16part);
  dfree(d_exty);
  dfree(d_hrow);
  dfree(d_h);
  dfree(d_p);
  dfree(d_shift);
  dfree(d_ext);
  dfree(d_pair_fg);
  dfree(d_pair_meas);
  dfree(d_obs_mk);
  dfree(d_obs_f);
  dfree(d_mk_meta);
  dfree(d_mk_off);
  dfree(d_fmeas);
  dfree(d_act);
  dfree(d_pref);
  dfree(d_meas_pair_base);
  dfree(d_pe_off);
  dfree(d_pe_ent);
  dfree(d_tgc_tiles);
  dfree(d_tgc_half);
  dfree(d_tgc_off);
  dfree(d_tgc_k32);
  dfree(d_tgc_k64);
  dfree(d_io_in);
  dfree(d_io_out);
  if (stream) hipStreamDestroy(stream);
}

// ------------------------------------------------------------ validation
static int idx1(double v, int n, const char* what) {
  // 1-based double index as MATLAB holds it (cu:317, :329-333)
  if (!(v >= 1.0) || !(v <= (double)n) || v != std::floor(v))
    throw Error(BORN_ERR_INVALID_ARGUMENT, std::string(what) + " index out of range or not an integer");
  return (int)v - 1;
}

// Measurement/mix index structures shared by both creation paths.
static void build_index_structures(Plan& P, int N_meas, const born_measurement_desc* meas) {
  if (N_meas <= 0 || !meas) throw Error(BORN_ERR_NO_OPERATOR_BORN, "no selected measurements");
  P.meas_pair_base.assign(N_meas, 0);
  P.meas_Nfm.assign(N_meas, 0);
  P.pair_fg.clear();
  P.pair_meas.clear();
  for (int m = 0; m < N_meas; ++m) {
    const born_measurement_desc& M = meas[m];
    if (M.N_f_unique <= 0 || !M.indices_f_to_unique)
      throw Error(BORN_ERR_INVALID_ARGUMENT, "measurement without unique frequencies");
    P.meas_pair_base[m] = (int)P.pair_fg.size();
    P.meas_Nfm[m] = M.N_f_unique;
    for (int fm = 0; fm < M.N_f_unique; ++fm) {
      P.pair_fg.push_back(idx1(M.indices_f_to_unique[fm], P.N_fu, "indices_f_to_unique{m}"));
      P.pair_meas.push_back(m);
    }
  }
  P.N_pairs = (int)P.pair_fg.size();
  // pairs grouped by global frequency
  P.fg_pair_off.assign(P.N_fu + 1, 0);
  for (int fg : P.pair_fg) P.fg_pair_off[fg + 1]++;
  for (int f = 0; f < P.N_fu; ++f) P.fg_pair_off[f + 1] += P.fg_pair_off[f];
  P.fg_pairs.assign(P.N_pairs, 0);
  {
    std::vector<int> fill(P.fg_pair_off.begin(), P.fg_pair_off.end() - 1);
    for (int pr = 0; pr < P.N_pairs; ++pr) P.fg_pairs[fill[P.pair_fg[pr]]++] = pr;
  }
  // mixes
  P.mk_meas.clear(); P.mk_Nf.clear(); P.mk_Nact.clear();
  P.mk_obs_off.clear(); P.mk_f_off.clear(); P.mk_a_off.clear(); P.mk_pref_off.clear();
  P.fmeas.clear(); P.act.clear(); P.pref.clear();
  int64_t obs = 0;
  for (int m = 0; m < N_meas; ++m) {
    const born_measurement_desc& M = meas[m];
    if (M.N_mix <= 0 || !M.mixes) throw Error(BORN_ERR_INVALID_ARGUMENT, "measurement without mixes");
    for (int k = 0; k < M.N_mix; ++k) {
      const born_mix_desc& X = M.mixes[k];
      if (X.N_f <= 0 || X.N_active <= 0 || !X.indices_f_to_unique || !X.indices_active || !X.prefactors)
        throw Error(BORN_ERR_INVALID_ARGUMENT, "incomplete mix description");
      P.mk_meas.push_back(m);
      P.mk_Nf.push_back(X.N_f);
      P.mk_Nact.push_back(X.N_active);
      P.mk_obs_off.push_back(obs);
      P.mk_f_off.push_back((int64_t)P.fmeas.size());
      P.mk_a_off.push_back((int64_t)P.act.size());
      P.mk_pref_off.push_back((int64_t)(P.pref.size() / 2));
      for (int f = 0; f < X.N_f; ++f) P.fmeas.push_back(idx1(X.indices_f_to_unique[f], M.N_f_unique, "indices_f_to_unique{k}"));
      std::vector<char> seen(P.N_e, 0);
      for (int a = 0; a < X.N_active; ++a) {
        const int e = idx1(X.indices_active[a], P.N_e, "indices_active");
        if (seen[e]) throw Error(BORN_ERR_UNSUPPORTED, "an element appears twice in one mix");
        seen[e] = 1;
        P.act.push_back(e);
      }
      P.pref.insert(P.pref.end(), X.prefactors, X.prefactors + 2 * (size_t)X.N_f * X.N_active);
      obs += X.N_f;
    }
  }
  P.N_mk = (int)P.mk_meas.size();
  P.N_obs = obs;
}

static void upload_index_structures(Plan& P, hipStream_t s) {
  P.d_pair_fg = dupload(P.pair_fg, s);
  P.d_pair_meas = dupload(P.pair_meas, s);
  P.d_meas_pair_base = dupload(P.meas_pair_base, s);
  std::vector<int> obs_mk(P.N_obs), obs_f(P.N_obs);
  std::vector<int> meta(4 * (size_t)P.N_mk);
  std::vector<int64_t> off(4 * (size_t)P.N_mk);
  for (int mk = 0; mk < P.N_mk; ++mk) {
    meta[4 * mk + 0] = P.mk_meas[mk];
    meta[4 * mk + 1] = P.mk_Nf[mk];
    meta[4 * mk + 2] = P.mk_Nact[mk];
    meta[4 * mk + 3] = 0;
    off[4 * mk + 0] = P.mk_obs_off[mk];
    off[4 * mk + 1] = P.mk_f_off[mk];
    off[4 * mk + 2] = P.mk_a_off[mk];
    off[4 * mk + 3] = P.mk_pref_off[mk];
    for (int f = 0; f < P.mk_Nf[mk]; ++f) {
      obs_mk[P.mk_obs_off[mk] + f] = mk;
      obs_f[P.mk_obs_off[mk] + f] = f;
    }
  }
  P.d_obs_mk = dupload(obs_mk, s);
  P.d_obs_f = dupload(obs_f, s);
  P.d_mk_meta = dupload(meta, s);
  P.d_mk_off = dupload(off, s);
  P.d_fmeas = dupload(P.fmeas, s);
  P.d_act = dupload(P.act, s);
  std::vector<double2> pref(P.pref.size() / 2);
  for (size_t i = 0; i < pref.size(); ++i) pref[i] = make_double2(P.pref[2 * i], P.pref[2 * i + 1]);
  P.d_pref = dupload(pref, s);
  // un-mixing gather lists: for every (pair, element) the (observation,
  // prefactor) entries that read it, in (mix, frequency) order
  std::vector<int64_t> peoff((size_t)P.N_pairs * P.N_e + 1, 0);
  for (int mk = 0; mk < P.N_mk; ++mk)
    for (int f = 0; f < P.mk_Nf[mk]; ++f) {
      const int pr = P.meas_pair_base[P.mk_meas[mk]] + P.fmeas[P.mk_f_off[mk] + f];
      for (int a = 0; a < P.mk_Nact[mk]; ++a) peoff[(size_t)pr * P.N_e + P.act[P.mk_a_off[mk] + a] + 1]++;
    }
  for (size_t i = 0; i + 1 < peoff.size(); ++i) peoff[i + 1] += peoff[i];
  if (peoff.back() > INT32_MAX || P.N_obs > INT32_MAX || (int64_t)P.pref.size() / 2 > INT32_MAX)
    throw Error(BORN_ERR_UNSUPPORTED, "too many observations for the un-mixing lists");
  std::vector<int2> ent((size_t)peoff.back());
  {
    std::vector<int64_t> fill(peoff.begin(), peoff.end() - 1);
    for (int mk = 0; mk < P.N_mk; ++mk)
      for (int f = 0; f < P.mk_Nf[mk]; ++f) {
        const int pr = P.meas_pair_base[P.mk_meas[mk]] + P.fmeas[P.mk_f_off[mk] + f];
        for (int a = 0; a < P.mk_Nact[mk]; ++a) {
          const size_t pe = (size_t)pr * P.N_e + P.act[P.mk_a_off[mk] + a];
          ent[fill[pe]++] = make_int2((int)(P.mk_obs_off[mk] + f),
                                      (int)(P.mk_pref_off[mk] + (int64_t)a * P.mk_Nf[mk] + f));
        }
      }
  }
  P.d_pe_off = dupload(peoff, s);
  P.d_pe_ent = dupload(ent, s);
}

// The lateral Toeplitz structure of setup_grid_symmetric.shift_lateral
// (setup_grid_symmetric.m:95-241): in 2-D, element e sees point (i, z) at
// h_ref point (ext[i - e*Np + (N_e-1)*Np], z).  Derived from the map itself
// and verified entry by entry; any mismatch selects the general path.
// Structured path: the element's transfer function is a window of one
// extended lateral row (2-D) or of one extended x-row chosen by the y offset
// (3-D).  finish_structure takes the lateral maps extx (and exty), sizes the
// tiles, builds the table row of every (y slice, row) and uploads the maps.
static void finish_structure(Plan& P) {
  P.structured = false;
  const int Nex = P.Nex, Ney = (P.nd == 3) ? P.Ney : 1;
  if (Nex > 128) return;
  P.N_ext = P.Nx + (Nex - 1) * P.Np;
  {
    int mt = (Nex + 15) / 16;  // element tiles of a slice, rounded up to 1, 2, 4 or 8
    mt = mt <= 1 ? 1 : (mt <= 2 ? 2 : (mt <= 4 ? 4 : 8));
    P.Ne_pad = 16 * mt;
  }
  if (!fwd_structured_fits(P.Nx, P.Ne_pad, P.Np) || !adj_structured_fits(P.Nx, P.Ne_pad, P.Np)) return;
  const int Ny = (P.nd == 3) ? P.Ny : 1;
  P.n_rows = Ny * P.Nz;
  const int off_y = (Ney - 1) * P.Npy;
  std::vector<int> hrow((size_t)Ney * P.n_rows);
  for (int ey = 0; ey < Ney; ++ey)
    for (int iz = 0; iz < P.Nz; ++iz)
      for (int iy = 0; iy < Ny; ++iy) {
        const int ty = (P.nd == 3) ? P.exty[iy - ey * P.Npy + off_y] : 0;
        hrow[(size_t)ey * P.n_rows + iy + (size_t)Ny * iz] = ty + P.HY * iz;
      }
  dfree(P.d_ext);
  dfree(P.d_exty);
  dfree(P.d_hrow);
  P.d_ext = dupload(P.extx, P.stream);
  if (P.nd == 3) P.d_exty = dupload(P.exty, P.stream);
  P.d_hrow = dupload(hrow, P.stream);
  P.structured = true;
}

// Lateral maps derived from a given shift map, verified entry by entry; any
// mismatch leaves the plan on the general gather path.
static void detect_structure(Plan& P) {
  P.structured = false;
  if (P.nd == 3 && P.Nex * P.Ney != P.N_e) return;
  if (P.nd == 2 && P.Nex != P.N_e) return;
  const int Nx = P.Nx, Ny = (P.nd == 3) ? P.Ny : 1, Nz = P.Nz;
  const int Nex = P.Nex, Ney = (P.nd == 3) ? P.Ney : 1;
  const int* sh = P.h_shift.data();
  const int HR = P.HR, HY = (P.nd == 3) ? P.HY : 1;
  // points per pitch: element 1 (x) and element Nex (y) see the last point of
  // the first x-row / the first point of the last y-row one pitch closer
  int Np = 1, Npy = 1;
  if (Nex >= 2) {
    const int p = Nx - 1;
    Np = sh[p] - sh[(size_t)1 * P.N_pts + p];
    if (Np < 1 || Np > HR - 1) return;
  }
  if (Ney >= 2) {
    const int p = Nx * (Ny - 1);
    const int d = sh[p] - sh[(size_t)Nex * P.N_pts + p];
    if (d % HR != 0) return;
    Npy = d / HR;
    if (Npy < 1 || Npy > HY - 1) return;
  }
  const int N_ext = Nx + (Nex - 1) * Np, N_exty = Ny + (Ney - 1) * Npy;
  std::vector<int> extx(N_ext, -1), exty(N_exty, -1);
  for (int ex = 0; ex < Nex; ++ex)
    for (int i = 0; i < Nx; ++i) {
      const int t = i - ex * Np + (Nex - 1) * Np;
      const int v = sh[(size_t)ex * P.N_pts + i];
      if (v < 0 || v >= HR) return;
      if (extx[t] < 0) extx[t] = v;
      else if (extx[t] != v) return;
    }
  if (P.nd == 3) {
    for (int ey = 0; ey < Ney; ++ey)
      for (int iy = 0; iy < Ny; ++iy) {
        const int t = iy - ey * Npy + (Ney - 1) * Npy;
        const int v = sh[(size_t)(ey * Nex) * P.N_pts + (size_t)iy * Nx] - extx[(Nex - 1) * Np];
        if (v % HR != 0) return;
        const int r = v / HR;
        if (r < 0 || r >= HY) return;
        if (exty[t] < 0) exty[t] = r;
        else if (exty[t] != r) return;
      }
    for (int t = 0; t < N_exty; ++t)
      if (exty[t] < 0) return;
  }
  for (int t = 0; t < N_ext; ++t)
    if (extx[t] < 0) return;
  for (int ey = 0; ey < Ney; ++ey)
    for (int ex = 0; ex < Nex; ++ex) {
      const int* se = sh + (size_t)(ex + Nex * ey) * P.N_pts;
      for (int z = 0; z < Nz; ++z)
        for (int iy = 0; iy < Ny; ++iy) {
          const int ty = (P.nd == 3) ? exty[iy - ey * Npy + (Ney - 1) * Npy] : 0;
          const int base = HR * (ty + HY * z);
          const int* row = se + (size_t)Nx * (iy + (size_t)Ny * z);
          for (int i = 0; i < Nx; ++i)
            if (row[i] != base + extx[i - ex * Np + (Nex - 1) * Np]) return;
        }
    }
  P.Np = Np;
  P.Npy = Npy;
  P.extx = std::move(extx);
  P.exty = std::move(exty);
  finish_structure(P);
}

static bool use_structured(const Plan& P) { return P.structured && !P.force_general; }

static void materialize_shift(Plan& P);

static void ensure_shift_on_device(Plan& P) {
  if (P.d_shift) return;
  if (P.h_shift.empty()) materialize_shift(P);
  P.d_shift = dupload(P.h_shift, P.stream);
}

// Forward blocking: about items_target work items per launch, (chunk,
// z-block) for the plain kernel and (chunk, z-block, 256-point segment) for
// the split kernel.
struct Blocking {
  int NZB, ZR, NZB_s, ZR_s, nseg_s, vblk;
};
static Blocking blocking_for(const Plan& P, int n_chunks, int target) {
  Blocking b;
  const int Ney = P.nd == 3 ? P.Ney : 1;
  const int rows = P.n_rows > 0 ? P.n_rows : P.Nz;
  // the plain forward takes sgrp_p y slices per work item
  const int sgrp_p = fwd_plain_slice_group(P.nd, P.Nx, P.Ne_pad, Ney, P.Np);
  int NZB = std::max(1, std::min(rows, target / std::max(1, n_chunks * (Ney / sgrp_p))));
  // several slices per item: whole groups of XCDS row blocks (from XCDS / 2
  // up), so that the slice groups of a row block run together on one XCD
  // (slice_item)
  if (sgrp_p > 1 && NZB >= XCDS / 2) NZB = std::min(rows, (NZB + XCDS - 1) / XCDS * XCDS);
  b.ZR = (rows + NZB - 1) / NZB;
  b.NZB = (rows + b.ZR - 1) / b.ZR;
  b.nseg_s = (P.Nx + 255) / 256;
  // the split forward takes sgrp y slices per work item
  const int sgrp = fwd_split_slice_group(P.nd, P.Nx, P.Ne_pad, Ney, P.Np);
  NZB = std::max(1, std::min(rows, target / std::max(1, n_chunks * (Ney / sgrp) * b.nseg_s)));
  if (sgrp > 1 && NZB >= XCDS / 2) NZB = std::min(rows, (NZB + XCDS - 1) / XCDS * XCDS);
  b.ZR_s = (rows + NZB - 1) / NZB;
  b.NZB_s = (rows + b.ZR_s - 1) / b.ZR_s;
  b.vblk = Ney * std::max(b.NZB, b.NZB_s * b.nseg_s);   // partial-sum blocks per chunk
  return b;
}
static Blocking blocking_for(const Plan& P, int n_chunks) { return blocking_for(P, n_chunks, P.items_target); }
static void set_blocking(Plan& P) {
  const Blocking b = blocking_for(P, (int)P.chunks.size());
  P.NZB = b.NZB;
  P.ZR = b.ZR;
  P.NZB_s = b.NZB_s;
  P.ZR_s = b.ZR_s;
  P.nseg_s = b.nseg_s;
  P.vblk = b.vblk;
}

static int chunk_max_cols(const std::vector<Chunk>& ch) {
  int m = 0;
  for (const Chunk& c : ch) m = std::max(m, c.ncols);
  return m;
}

// The record of a plan inside a structured launch (workspace must exist).
static Inst make_inst(const Plan& P) {
  Inst I;
  std::memset(&I, 0, sizeof(I));
  I.h = (decltype(I.h))P.d_h;
  I.P = (decltype(I.P))P.d_p;
  I.gm = (decltype(I.gm))P.d_gm;
  I.ext = (decltype(I.ext))P.d_ext;
  I.Wc = (decltype(I.Wc))P.d_Wc;
  I.vpart = (decltype(I.vpart))P.d_vpart;
  I.Nx = P.Nx;
  I.Nz = P.Nz;
  I.N_pts = P.N_pts;
  I.N_e = P.Nex;   // elements of one slice
  I.Np = P.Np;
  I.Ney = P.nd == 3 ? P.Ney : 1;
  I.n_rows = P.n_rows;
  I.hrow = (decltype(I.hrow))P.d_hrow;
  I.fscale = (decltype(I.fscale))P.d_fscale;
  I.N_ext = P.N_ext;
  I.Hp = P.H_plane;
  I.HR = P.HR;
  I.n_chunks = (int)P.chunks.size();
  I.NZB = P.NZB;
  I.ZR = P.ZR;
  I.S = (P.Nx + 15) / 16;
  I.NZB_s = P.NZB_s;
  I.ZR_s = P.ZR_s;
  I.nseg_s = P.nseg_s;
  I.sgrp = fwd_split_slice_group(P.nd, P.Nx, P.Ne_pad, I.Ney, P.Np);
  I.sgrp_p = fwd_plain_slice_group(P.nd, P.Nx, P.Ne_pad, I.Ney, P.Np);
  I.max_cols = chunk_max_cols(P.chunks);
  I.gpart = (decltype(I.gpart))P.d_gpart;
  I.n_out = 1;
  for (const Chunk& c : P.chunks) I.n_out = std::max(I.n_out, c.r + 1);
  return I;
}

// Work item and workgroup ranges of every instance, totals, device copy.
static void finalize_launchset(LaunchSet& L) {
  int it = 0, its = 0, wg = 0, wgs = 0;
  L.S_max = 1;
  // the several-slice forwards serve a launch only when every instance
  // takes the same slice group
  bool same_group = true, same_group_p = true;
  for (const Inst& I : L.inst) {
    same_group = same_group && I.sgrp == L.inst[0].sgrp;
    same_group_p = same_group_p && I.sgrp_p == L.inst[0].sgrp_p;
  }
  for (Inst& I : L.inst) {
    if (!same_group) I.sgrp = 1;
    if (!same_group_p) I.sgrp_p = 1;
  }
  // chunks of at most 4 columns: the few-column kernels (one work item per
  // (chunk, y slice, row block), like the one-slice plain forward)
  int max_cols = 0;
  for (const Inst& I : L.inst) max_cols = std::max(max_cols, I.max_cols);
  L.fc_w = L.inst.empty() || !L.fc_allowed ? 0 : fewcol_width(max_cols);
  // the plain MFMA kernels take the 8-column stacked form when every chunk fits
  L.stk = max_cols > 0 && max_cols <= NC / 2;
  for (const Inst& I : L.inst) {
    if (!I.gpart || !fewcol_fits(I.Nx, I.Np, 16 * L.MT, L.fc_w, I.n_rows)) L.fc_w = 0;
    // the 3-column kernels take one rhs per chunk (right-hand sides packed
    // per column give chunks of a multiple of 2 columns, never 3)
    if (L.fc_w == 3 && I.col_rhs > 1) L.fc_w = 0;
  }
  for (const Inst& I : L.inst) {
    int bt = 0, nr = 0;
    if (L.fc_w > 0) fewcol_adj_block(I.Nx, I.Np, 16 * L.MT, L.fc_w, I.col_rhs, I.n_rows, bt, nr);
    if (bt == 0) L.fc_w = 0;
  }
  // few-column forward items of at least 4 steps of R staged rows (fewer,
  // longer items than the MFMA kernels' blocking)
  if (L.fc_w > 0)
    for (Inst& I : L.inst) {
      I.sgrp_p = 1;
      const int R = fewcol_fwd_rows(I.Nx, I.Np, 16 * L.MT, L.fc_w);
      if (I.ZR < 4 * R) {
        I.ZR = std::min(I.n_rows, 4 * R);
        I.NZB = (I.n_rows + I.ZR - 1) / I.ZR;
      }
    }
  // adjoint: task blocks x chunk groups, about 2048 workgroups per instance
  // and at least 2 chunks per group
  int wgf = 0;
  for (Inst& I : L.inst) {
    I.wg_begin_f = wgf;
    I.fc_g = 1;
    I.fc_bt = 256;
    I.fc_nr = 1;
    I.fc_cb = 1;
    if (L.fc_w == 0) continue;
    fewcol_adj_block(I.Nx, I.Np, 16 * L.MT, L.fc_w, I.col_rhs, I.n_rows, I.fc_bt, I.fc_nr);
    const long T = (long)I.n_rows * fewcol_adj_tpr(I.Nx, I.Np, 16 * L.MT, L.fc_w);
    const int nb = (int)((T + I.fc_bt - 1) / I.fc_bt);
    I.fc_g = std::max(1, std::min({FC_GROUPS, (2048 + nb - 1) / nb, I.n_chunks / 2}));
    I.fc_cb = fewcol_adj_cb(I.Nx, I.Np, 16 * L.MT, L.fc_w, I.col_rhs, I.fc_nr, (I.n_chunks + I.fc_g - 1) / I.fc_g, I.Ney);
    wgf += nb * I.fc_g;
  }
  L.n_wg_f = wgf;
  for (Inst& I : L.inst) {
    I.item_begin = it;
    it += I.n_chunks * (I.Ney / std::max(1, I.sgrp_p)) * I.NZB;
    I.item_begin_s = its;
    its += I.n_chunks * (I.Ney / std::max(1, I.sgrp)) * I.NZB_s * I.nseg_s;
    I.adj_segs = (I.Nx + 64 * L.adj_mw - 1) / (64 * L.adj_mw);
    I.wg_begin = wg;
    wg += I.adj_segs * I.n_rows;
    I.adj_segs_s = (I.Nx + 255) / 256;
    I.wg_begin_s = wgs;
    wgs += I.adj_segs_s * I.n_rows;
    L.S_max = std::max(L.S_max, I.S);
  }
  L.n_items = it;
  L.n_items_s = its;
  L.n_wg = wg;
  L.n_wg_s = wgs;
  if (!L.d_inst) L.d_inst = dalloc<Inst>(MAX_BATCH);
  HIP_CHECK(hipMemcpy(L.d_inst, L.inst.data(), L.inst.size() * sizeof(Inst), hipMemcpyHostToDevice));
}

// With several right-hand sides, chunks that pack the (pair, rhs) columns
// of all right-hand sides of a frequency need fewer MFMA column tiles than
// one chunk set per rhs.  The forward then runs on the products
// p_inc * gamma per column (one extra pass over p_inc), when they fit in
// the free device memory with 1 GiB to spare.
// Whether packing the right-hand sides per column lane (setup_column_rhs_adjoint)
// needs fewer adjoint chunks than one chunk set per rhs.
static bool column_rhs_saves(const Plan& P, int N_rhs) {
  int k = 1;
  while (k < N_rhs) k <<= 1;
  if (N_rhs <= 1 || k > NC) return false;
  const int per = NC / k;
  size_t n2 = 0;
  for (int fg = 0; fg < P.N_fu; ++fg) n2 += (P.fg_pair_off[fg + 1] - P.fg_pair_off[fg] + per - 1) / per;
  return n2 < P.chunks.size();
}

// The forward's own launch set over the rhs-major chunks of the workspace
// (the p_inc table, gamma of each chunk's rhs, the plan's partial sums), so
// that the adjoint can replace the plan's chunks by its packed ones.
static void setup_rhs_major_forward(Plan& P) {
  const int n_f = (int)P.chunks.size();
  std::vector<int> fpairs((size_t)n_f * NC, P.N_pairs);   // unused columns: the zero row
  for (int c = 0; c < n_f; ++c)
    for (int q = 0; q < P.chunks[c].ncols; ++q) fpairs[(size_t)c * NC + q] = P.col_pair[P.chunks[c].col_start + q];
  P.fchunks = P.chunks;
  P.fcol = P.col_pair;
  P.d_fchunks = dupload(P.fchunks, P.stream);
  P.d_fcol = dupload(P.fcol, P.stream);
  P.d_fchunk_pairs = dupload(fpairs, P.stream);
  P.fvblk = P.vblk;
  Inst I = make_inst(P);
  P.d_fvpart = P.d_vpart;   // the plan's partials move to the forward's launch set
  P.d_vpart = nullptr;
  P.ls_fwd.inst.assign(1, I);
  P.ls_fwd.MT = P.Ne_pad / 16;
  P.ls_fwd.Np = P.Np;
  P.ls_fwd.adj_mw = P.ls.adj_mw;
  P.ls_fwd.d_chunks = P.d_fchunks;
  P.ls_fwd.d_chunk_pairs = P.d_fchunk_pairs;
  finalize_launchset(P.ls_fwd);
  P.fwd_expand = true;
  P.fwd_table_expanded = false;
}

static void setup_expanded_forward(Plan& P, int N_rhs) {
  if (N_rhs <= 1) return;
  std::vector<Chunk> fch;
  std::vector<int> fcol;
  for (int fg = 0; fg < P.N_fu; ++fg) {
    const int b = P.fg_pair_off[fg], e = P.fg_pair_off[fg + 1];
    const int ncol = (e - b) * N_rhs;
    for (int j = 0; j < ncol; j += NC) {
      Chunk c;
      c.fg = fg;
      c.r = 0;
      c.col_start = (int)fcol.size();
      c.ncols = std::min(NC, ncol - j);
      for (int q = 0; q < c.ncols; ++q) {
        const int k = j + q;   // column k of the frequency: pair b + k / N_rhs, rhs k % N_rhs
        fcol.push_back(P.fg_pairs[b + k / N_rhs] * N_rhs + k % N_rhs);
      }
      fch.push_back(c);
    }
  }
  // The expanded table costs a write and a read of N_rhs tables per forward
  // on top of the reads it replaces.  The FP32 contraction is bound by its
  // MFMA issue, so any saved column tile pays; the split contraction is bound
  // by its operand traffic (LDS and power), so the table pays only where it
  // at least halves the column tiles (few pairs per frequency, as with 5
  // waves and 4 rhs).  Otherwise, when the adjoint packs the right-hand
  // sides per column lane, the forward keeps the rhs-major chunks in a
  // launch set of its own.
  const size_t rows = (size_t)P.N_pairs * N_rhs + 1;
  const size_t bytes = rows * P.N_pts * sizeof(float2);
  size_t free_b = 0, total_b = 0;
  HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const bool pays = P.ws_split ? 2 * fch.size() <= P.chunks.size() : fch.size() < P.chunks.size();
  if (!pays || bytes + ((size_t)1 << 30) > free_b) {
    if (column_rhs_saves(P, N_rhs)) setup_rhs_major_forward(P);
    return;
  }
  const int n_f = (int)fch.size();
  std::vector<int> fpairs((size_t)n_f * NC, (int)(rows - 1));   // unused columns: the zero row
  for (int c = 0; c < n_f; ++c)
    for (int q = 0; q < fch[c].ncols; ++q) fpairs[(size_t)c * NC + q] = fcol[fch[c].col_start + q];
  const Blocking bl = blocking_for(P, n_f);
  P.fchunks = fch;
  P.fcol = fcol;
  P.d_fchunks = dupload(fch, P.stream);
  P.d_fcol = dupload(fcol, P.stream);
  P.d_fchunk_pairs = dupload(fpairs, P.stream);
  P.d_pexp = dalloc<float2>(rows * P.N_pts);
  HIP_CHECK(hipMemset(P.d_pexp + (rows - 1) * P.N_pts, 0, P.N_pts * sizeof(float2)));
  std::vector<float2> ones(P.N_pts, make_float2(1.f, 0.f));
  P.d_ones = dupload(ones, P.stream);
  P.fvblk = bl.vblk;
  P.d_fvpart = dalloc<float2>((size_t)n_f * bl.vblk * P.Ne_pad * NC);
  Inst I = make_inst(P);
  I.P = (decltype(I.P))P.d_pexp;
  I.gm = (decltype(I.gm))P.d_ones;
  I.vpart = (decltype(I.vpart))P.d_fvpart;
  I.n_chunks = n_f;
  I.max_cols = chunk_max_cols(fch);
  I.NZB = bl.NZB;
  I.ZR = bl.ZR;
  I.NZB_s = bl.NZB_s;
  I.ZR_s = bl.ZR_s;
  I.nseg_s = bl.nseg_s;
  P.ls_fwd.inst.assign(1, I);
  P.ls_fwd.MT = P.Ne_pad / 16;
  P.ls_fwd.Np = P.Np;
  P.ls_fwd.adj_mw = P.ls.adj_mw;
  P.ls_fwd.d_chunks = P.d_fchunks;
  P.ls_fwd.d_chunk_pairs = P.d_fchunk_pairs;
  finalize_launchset(P.ls_fwd);
  P.fwd_expand = true;
  P.fwd_table_expanded = true;
}

// With the forward on expanded columns, the adjoint can pack the right-hand
// sides the same way: chunks of NC / k pairs of one frequency with k (the
// rhs count rounded up to a power of 2) columns each, column = slot * k + r.
// The adjoint kernels keep one rhs per column lane and write gamma_hat once.
static void setup_column_rhs_adjoint(Plan& P, int N_rhs) {
  int k = 1;
  while (k < N_rhs) k <<= 1;
  if (N_rhs <= 1 || k > NC) return;
  const int per = NC / k;
  size_t n2 = 0;
  for (int fg = 0; fg < P.N_fu; ++fg) n2 += (P.fg_pair_off[fg + 1] - P.fg_pair_off[fg] + per - 1) / per;
  if (n2 >= P.chunks.size()) return;
  std::vector<Chunk> ch;
  std::vector<int> colp;
  std::vector<int> pr_chunk((size_t)P.N_pairs * N_rhs, -1), pr_pos((size_t)P.N_pairs * N_rhs, -1);
  std::vector<int> cpairs(n2 * NC, P.N_pairs);   // unused columns: the zero row
  for (int fg = 0; fg < P.N_fu; ++fg) {
    const int b = P.fg_pair_off[fg], e = P.fg_pair_off[fg + 1];
    for (int j = b; j < e; j += per) {
      Chunk c;
      c.fg = fg;
      c.r = 0;
      c.col_start = (int)colp.size();
      const int ns = std::min(per, e - j);
      c.ncols = ns * k;
      for (int q = 0; q < ns; ++q) {
        const int pr = P.fg_pairs[j + q];
        for (int r = 0; r < k; ++r) {
          colp.push_back(pr);
          if (r < N_rhs) {
            pr_chunk[(size_t)pr * N_rhs + r] = (int)ch.size();
            pr_pos[(size_t)pr * N_rhs + r] = q * k + r;
            cpairs[ch.size() * NC + q * k + r] = pr;
          }
        }
      }
      ch.push_back(c);
    }
  }
  dfree(P.d_chunks);
  dfree(P.d_col_pair);
  dfree(P.d_chunk_pairs);
  dfree(P.d_pr_chunk);
  dfree(P.d_pr_pos);
  dfree(P.d_Wc);
  dfree(P.d_vpart);   // the forward runs on the expanded chunks
  P.chunks = ch;
  P.col_pair = colp;
  P.d_chunks = dupload(ch, P.stream);
  P.d_col_pair = dupload(colp, P.stream);
  P.d_chunk_pairs = dupload(cpairs, P.stream);
  P.d_pr_chunk = dupload(pr_chunk, P.stream);
  P.d_pr_pos = dupload(pr_pos, P.stream);
  const size_t n = n2 * (P.nd == 3 ? P.Ney : 1) * P.Ne_pad * NC;
  P.d_Wc = dalloc<float2>(n);
  HIP_CHECK(hipMemset(P.d_Wc, 0, n * sizeof(float2)));  // unused columns and padding stay zero
  Inst I = make_inst(P);
  I.col_rhs = k;
  I.n_rhs = N_rhs;
  I.n_out = N_rhs;
  P.ls.inst.assign(1, I);
  P.ls.d_chunks = P.d_chunks;
  P.ls.d_chunk_pairs = P.d_chunk_pairs;
  finalize_launchset(P.ls);
}

// Chunks of <= NC (pair, rhs) columns per frequency, rhs-major.
static void build_workspace(Plan& P, int N_rhs);

// A failure while building frees the partial workspace, so a plan that
// could not get its workspace holds none.
static void ensure_workspace(Plan& P, int N_rhs) {
  const bool split = P.precision != BORN_PRECISION_FP32;
  if (P.ws_rhs == N_rhs && (N_rhs == 1 || P.ws_split == split)) return;
  free_workspace(P);
  P.ws_split = split;
  try {
    build_workspace(P, N_rhs);
  } catch (...) {
    free_workspace(P);
    throw;
  }
}

static void build_workspace(Plan& P, int N_rhs) {
  P.chunks.clear();
  P.col_pair.clear();
  std::vector<int> pr_chunk((size_t)P.N_pairs * N_rhs, -1), pr_pos((size_t)P.N_pairs * N_rhs, -1);
  for (int r = 0; r < N_rhs; ++r)
    for (int fg = 0; fg < P.N_fu; ++fg) {
      const int b = P.fg_pair_off[fg], e = P.fg_pair_off[fg + 1];
      for (int j = b; j < e; j += NC) {
        Chunk c;
        c.fg = fg;
        c.r = r;
        c.col_start = (int)P.col_pair.size();
        c.ncols = std::min(NC, e - j);
        for (int q = 0; q < c.ncols; ++q) {
          const int pr = P.fg_pairs[j + q];
          pr_chunk[(size_t)pr * N_rhs + r] = (int)P.chunks.size();
          pr_pos[(size_t)pr * N_rhs + r] = q;
          P.col_pair.push_back(pr);
        }
        P.chunks.push_back(c);
      }
    }
  const int n_chunks = (int)P.chunks.size();
  std::vector<int> chunk_pairs((size_t)n_chunks * NC, P.N_pairs);  // unused: the zero row
  for (int c = 0; c < n_chunks; ++c)
    for (int q = 0; q < P.chunks[c].ncols; ++q) chunk_pairs[(size_t)c * NC + q] = P.col_pair[P.chunks[c].col_start + q];
  set_blocking(P);
  P.d_chunks = dupload(P.chunks, P.stream);
  P.d_col_pair = dupload(P.col_pair, P.stream);
  P.d_chunk_pairs = dupload(chunk_pairs, P.stream);
  P.d_pr_chunk = dupload(pr_chunk, P.stream);
  P.d_pr_pos = dupload(pr_pos, P.stream);
  if (P.structured) P.d_vpart = dalloc<float2>((size_t)n_chunks * P.vblk * P.Ne_pad * NC);
  P.d_V = dalloc<double2>((size_t)P.N_pairs * P.N_e * N_rhs);
  P.d_W = dalloc<float2>((size_t)P.N_pairs * P.N_e * N_rhs);
  if (P.structured) {
    const size_t n = (size_t)n_chunks * (P.nd == 3 ? P.Ney : 1) * P.Ne_pad * NC;
    P.d_Wc = dalloc<float2>(n);
    HIP_CHECK(hipMemset(P.d_Wc, 0, n * sizeof(float2)));  // padding columns/elements stay zero
  }
  P.d_gm = dalloc<float2>((size_t)N_rhs * P.N_pts);
  // partial sums of the few-column adjoint (plans with at most 4 pairs per frequency)
  int ppf = 0;
  for (int fg = 0; fg < P.N_fu; ++fg) ppf = std::max(ppf, P.fg_pair_off[fg + 1] - P.fg_pair_off[fg]);
  if (P.structured && ppf <= 4) P.d_gpart = dalloc<float2>((size_t)FC_GROUPS * N_rhs * P.N_pts);
  if (P.structured) {
    P.ls.inst.assign(1, make_inst(P));
    P.ls.MT = P.Ne_pad / 16;
    P.ls.Np = P.Np;
    P.ls.adj_mw = adj_mw_default(P.Nx);
    P.ls.d_chunks = P.d_chunks;
    P.ls.d_chunk_pairs = P.d_chunk_pairs;
    finalize_launchset(P.ls);
    setup_expanded_forward(P, N_rhs);
    if (P.fwd_expand) setup_column_rhs_adjoint(P, N_rhs);
  }
  HIP_CHECK(hipDeviceSynchronize());  // workspace ready before work on any stream
  // valid only once complete: a failure above leaves ws_rhs = -1, so the next
  // call frees the partial workspace and builds it again
  P.ws_rhs = N_rhs;
  ++P.ws_gen;
}

// ------------------------------------------------ allocation-free apply
// The scratch of born_apply_dev_noalloc: one static array per device, held
// until the process ends.  Host code never uses its name as an address:
// scratch_base() looks the device address up once per device (with that
// device current) and every launch takes that pointer.
static __device__ __align__(256) unsigned char scratch[BORN_NOALLOC_SCRATCH_BYTES];

static void* scratch_base(int device) {
  static std::mutex mu;
  static std::vector<void*> base;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)base.size() <= device) base.resize(device + 1, nullptr);
  if (!base[device]) {
    int cur = -1;
    HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) throw Error(BORN_ERR_DEVICE, "scratch lookup on a device that is not current");
    void* ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(BORN_ERR_DEVICE, "no device address for the scratch array");
    base[device] = ptr;
  }
  return base[device];
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Scratch bytes of the one-rhs layout with `n_chunks` chunks and blocking b:
// masked gamma, V, Wc, forward partial sums, each 256-byte aligned.
static size_t noalloc_bytes(const Plan& P, int n_chunks, const Blocking& b) {
  const int Ney = P.nd == 3 ? P.Ney : 1;
  return align256((size_t)P.N_pts * sizeof(float2)) + align256((size_t)P.N_pairs * P.N_e * sizeof(double2)) +
         align256((size_t)n_chunks * Ney * P.Ne_pad * NC * sizeof(float2)) +
         align256((size_t)n_chunks * b.vblk * P.Ne_pad * NC * sizeof(float2));
}

// The one-rhs chunk layout of the allocation-free apply, built at plan
// creation (structured plans).  Its blocking takes the largest work-item
// count, up to the plan's, whose scratch fits the static array; a plan that
// does not fit even with one z-block per chunk keeps na.ok = false.
static void build_noalloc_layout(Plan& P) {
  Plan::NoAlloc& A = P.na;
  if (!P.structured) {
    A.why = "the allocation-free apply needs the structured path";
    return;
  }
  A.chunks.clear();
  A.col_pair.clear();
  std::vector<int> pr_chunk((size_t)P.N_pairs, -1), pr_pos((size_t)P.N_pairs, -1);
  for (int fg = 0; fg < P.N_fu; ++fg) {
    const int b = P.fg_pair_off[fg], e = P.fg_pair_off[fg + 1];
    for (int j = b; j < e; j += NC) {
      Chunk c;
      c.fg = fg;
      c.r = 0;
      c.col_start = (int)A.col_pair.size();
      c.ncols = std::min(NC, e - j);
      for (int q = 0; q < c.ncols; ++q) {
        const int pr = P.fg_pairs[j + q];
        pr_chunk[pr] = (int)A.chunks.size();
        pr_pos[pr] = q;
        A.col_pair.push_back(pr);
      }
      A.chunks.push_back(c);
    }
  }
  const int n_chunks = (int)A.chunks.size();
  int target = P.items_target;
  Blocking bl = blocking_for(P, n_chunks, target);
  while (noalloc_bytes(P, n_chunks, bl) > BORN_NOALLOC_SCRATCH_BYTES && target > 1) {
    target /= 2;
    bl = blocking_for(P, n_chunks, target);
  }
  A.bytes = noalloc_bytes(P, n_chunks, bl);
  if (A.bytes > BORN_NOALLOC_SCRATCH_BYTES) {
    A.why = "the plan needs " + std::to_string(A.bytes) + " bytes of scratch; the allocation-free apply holds " +
            std::to_string((size_t)BORN_NOALLOC_SCRATCH_BYTES);
    return;
  }
  A.NZB = bl.NZB;
  A.ZR = bl.ZR;
  A.NZB_s = bl.NZB_s;
  A.ZR_s = bl.ZR_s;
  A.nseg_s = bl.nseg_s;
  A.vblk = bl.vblk;
  unsigned char* base = static_cast<unsigned char*>(scratch_base(P.device));
  const int Ney = P.nd == 3 ? P.Ney : 1;
  size_t off = 0;
  A.d_gm = reinterpret_cast<float2*>(base + off);
  off += align256((size_t)P.N_pts * sizeof(float2));
  A.d_V = reinterpret_cast<double2*>(base + off);
  off += align256((size_t)P.N_pairs * P.N_e * sizeof(double2));
  A.d_Wc = reinterpret_cast<float2*>(base + off);
  A.Wc_n = (size_t)n_chunks * Ney * P.Ne_pad * NC;
  off += align256(A.Wc_n * sizeof(float2));
  A.d_vpart = reinterpret_cast<float2*>(base + off);
  std::vector<int> chunk_pairs((size_t)n_chunks * NC, P.N_pairs);  // unused: the zero row
  for (int c = 0; c < n_chunks; ++c)
    for (int q = 0; q < A.chunks[c].ncols; ++q) chunk_pairs[(size_t)c * NC + q] = A.col_pair[A.chunks[c].col_start + q];
  A.d_chunks = dupload(A.chunks, P.stream);
  A.d_col_pair = dupload(A.col_pair, P.stream);
  A.d_chunk_pairs = dupload(chunk_pairs, P.stream);
  A.d_pr_chunk = dupload(pr_chunk, P.stream);
  A.d_pr_pos = dupload(pr_pos, P.stream);
  Inst I = make_inst(P);
  I.gm = (decltype(I.gm))A.d_gm;
  I.Wc = (decltype(I.Wc))A.d_Wc;
  I.vpart = (decltype(I.vpart))A.d_vpart;
  I.n_chunks = n_chunks;
  I.max_cols = chunk_max_cols(A.chunks);
  I.NZB = A.NZB;
  I.ZR = A.ZR;
  I.NZB_s = A.NZB_s;
  I.ZR_s = A.ZR_s;
  I.nseg_s = A.nseg_s;
  A.ls.inst.assign(1, I);
  A.ls.fc_allowed = false;   // its scratch holds no few-column partials
  A.ls.MT = P.Ne_pad / 16;
  A.ls.Np = P.Np;
  A.ls.adj_mw = adj_mw_default(P.Nx);
  A.ls.d_chunks = A.d_chunks;
  A.ls.d_chunk_pairs = A.d_chunk_pairs;
  finalize_launchset(A.ls);
  A.ok = true;
}

// Swaps the plan's per-rhs workspace with its allocation-free layout for the
// duration of one apply (and back, also on an exception); event timing is
// off meanwhile, since it creates events.
struct NoAllocView {
  Plan& P;
  bool timing;
  explicit NoAllocView(Plan& P_) : P(P_), timing(P_.timing) {
    swap_all();
    P.timing = false;
  }
  ~NoAllocView() {
    swap_all();
    P.timing = timing;
  }
  void swap_all() {
    Plan::NoAlloc& A = P.na;
    std::swap(P.chunks, A.chunks);
    std::swap(P.col_pair, A.col_pair);
    std::swap(P.d_chunks, A.d_chunks);
    std::swap(P.d_col_pair, A.d_col_pair);
    std::swap(P.d_chunk_pairs, A.d_chunk_pairs);
    std::swap(P.d_pr_chunk, A.d_pr_chunk);
    std::swap(P.d_pr_pos, A.d_pr_pos);
    std::swap(P.NZB, A.NZB);
    std::swap(P.ZR, A.ZR);
    std::swap(P.NZB_s, A.NZB_s);
    std::swap(P.ZR_s, A.ZR_s);
    std::swap(P.nseg_s, A.nseg_s);
    std::swap(P.vblk, A.vblk);
    std::swap(P.d_gm, A.d_gm);
    std::swap(P.d_V, A.d_V);
    std::swap(P.d_Wc, A.d_Wc);
    std::swap(P.d_vpart, A.d_vpart);
    std::swap(P.ls, A.ls);
    std::swap(P.ws_rhs, A.ws_rhs);
    std::swap(P.fwd_expand, A.fwd_expand);
    std::swap(P.fwd_table_expanded, A.fwd_table_expanded);
  }
};

static void ensure_io(Plan& P, size_t in_bytes, size_t out_bytes) {
  const size_t need = std::max(in_bytes, out_bytes);
  if (P.io_bytes >= need) return;
  dfree(P.d_io_in);
  dfree(P.d_io_out);
  P.d_io_in = P.d_io_out = nullptr;
  P.io_bytes = 0;
  P.d_io_in = dalloc<char>(need);
  P.d_io_out = dalloc<char>(need);
  P.io_bytes = need;
}

static void common_dims(Plan& P, int nd, const int32_t* Npa, const int32_t* Nea, int N_points, int N_elements) {
  if (nd != 2 && nd != 3) throw Error(BORN_ERR_NO_SYMMETRIC_GRIDS_SETUP, "grid must have 2 or 3 dimensions");
  P.nd = nd;
  P.Nx = Npa[0];
  if (nd == 2) {
    P.Ny = 1;
    P.Nz = Npa[1];
  } else {
    P.Ny = Npa[1];
    P.Nz = Npa[2];
  }
  P.Nex = Nea[0];
  P.Ney = (nd == 3) ? Nea[1] : 1;
  if (P.Nx <= 0 || P.Ny <= 0 || P.Nz <= 0 || (int64_t)P.Nx * P.Ny * P.Nz != N_points)
    throw Error(BORN_ERR_INVALID_ARGUMENT, "N_points_axis does not match N_points");
  if (P.Nex <= 0 || P.Ney <= 0 || P.Nex * P.Ney != N_elements)
    throw Error(BORN_ERR_INVALID_ARGUMENT, "N_elements_axis does not match N_elements");
  P.N_pts = N_points;
  P.N_e = N_elements;
  P.H_plane = N_points;
  P.HR = P.Nx;
  P.HY = P.Ny;
  P.n_rows = P.Ny * P.Nz;
}

static void build_noalloc_layout(Plan& P);

static void finish_plan(Plan& P, bool structure_known = false) {
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, P.device) == hipSuccess && ncu > 0)
    P.n_cu = ncu;
  if (!structure_known) detect_structure(P);
  if (!P.structured) ensure_shift_on_device(P);
  P.d_fscale = dupload(std::vector<float>(5, 1.f), P.stream);
  P.d_f16max = dalloc<float>(2);
  P.d_f16part = dalloc<float>(256);
  set_max_lds(160 * 1024);
  set_max_lds_split(160 * 1024);
  set_max_lds_split3(160 * 1024);
  set_max_lds_fc(160 * 1024);
  build_noalloc_layout(P);
  HIP_CHECK(hipStreamSynchronize(P.stream));
}

// ------------------------------------------------------------ plan paths
static Plan* create_tables(const born_tables_desc* d) {
  if (!d) throw Error(BORN_ERR_NUMBER_ARGUMENTS, "descriptor is NULL");
  if (!d->h_ref) throw Error(BORN_ERR_NO_COMPLEX_DOUBLES, "operator_born.sequence.h_ref must be complex doubles!");
  if (!d->indices_grid_FOV_shift)
    throw Error(BORN_ERR_NO_SYMMETRIC_GRIDS_SETUP, "sequence.setup must be scattering.sequences.setups.setup_grid_symmetric!");
  if (d->precision != BORN_PRECISION_FP32 && d->precision != BORN_PRECISION_SPLIT_BF16 &&
      d->precision != BORN_PRECISION_SPLIT_FP16)
    throw Error(BORN_ERR_INVALID_ARGUMENT, "unknown precision");
  std::unique_ptr<Plan> P(new Plan());
  P->device = d->device;
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev) throw Error(BORN_ERR_DEVICE, "invalid device index");
  HIP_CHECK(hipSetDevice(d->device));
  HIP_CHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  P->precision = d->precision;
  if (d->N_f_unique <= 0) throw Error(BORN_ERR_INVALID_ARGUMENT, "N_f_unique must be positive");
  P->N_fu = d->N_f_unique;
  common_dims(*P, d->N_dimensions, d->N_points_axis, d->N_elements_axis, d->N_points, d->N_elements);
  build_index_structures(*P, d->N_measurements, d->measurements);
  for (int m = 0; m < d->N_measurements; ++m)
    if (!d->measurements[m].p_incident)
      throw Error(BORN_ERR_NO_COMPLEX_DOUBLES, "operator_born.incident_waves.p_incident must be complex doubles!");
  // shift map, 1-based double -> 0-based int [e][p]
  P->h_shift.resize((size_t)P->N_e * P->N_pts);
  for (int e = 0; e < P->N_e; ++e)
    for (int p = 0; p < P->N_pts; ++p)
      P->h_shift[(size_t)e * P->N_pts + p] = idx1(d->indices_grid_FOV_shift[(size_t)e * P->N_pts + p], P->N_pts,
                                                  "indices_grid_FOV_shift");
  upload_index_structures(*P, P->stream);
  // tables: FP64 column-major -> FP32 row-major on the device, through one
  // staging buffer sized for the largest matrix
  size_t max_rows = P->N_fu;
  for (int m = 0; m < d->N_measurements; ++m) max_rows = std::max(max_rows, (size_t)d->measurements[m].N_f_unique);
  double2* stage = dalloc<double2>(max_rows * P->N_pts);
  try {
    P->d_h = dalloc<float2>((size_t)P->N_fu * P->N_pts);
    HIP_CHECK(hipMemcpyAsync(stage, d->h_ref, sizeof(double2) * (size_t)P->N_fu * P->N_pts, hipMemcpyHostToDevice,
                             P->stream));
    launch_convert_transpose(stage, P->N_fu, P->N_pts, P->d_h, P->stream);
    P->d_p = alloc_p_table(*P);
    for (int m = 0; m < d->N_measurements; ++m) {
      const int nf = d->measurements[m].N_f_unique;
      HIP_CHECK(hipMemcpyAsync(stage, d->measurements[m].p_incident, sizeof(double2) * (size_t)nf * P->N_pts,
                               hipMemcpyHostToDevice, P->stream));
      launch_convert_transpose(stage, nf, P->N_pts, P->d_p + (size_t)P->meas_pair_base[m] * P->N_pts, P->stream);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(P->stream));
  } catch (...) {
    hipStreamSynchronize(P->stream);   // conversions already enqueued read the staging buffer
    dfree(stage);
    throw;
  }
  dfree(stage);
  finish_plan(*P);
  return P.release();
}

// shift_lateral restated (setup_grid_symmetric.m:171-233): subscripts shifted
// by (element subscript - 1) * points-per-pitch, mirrored when <= 0 about
// index_offset = N_points_axis - (N_elements_axis - 1) * Np + 1.
// Element e sees grid point p at h_ref entry shift[e][p].  Mirrored layout
// (setup_grid_symmetric.m:184-226): the lateral subscript moves by the
// element offset and reflects at the FOV edge, so h_ref covers the FOV grid.
// Extended layout (FOV narrower than the aperture, where the reflection is
// undefined): h_ref covers the grid extended by (N_elements_axis - 1)
// pitches before the FOV, and the lateral subscript is the offset itself.
// Both are separable: per lateral axis, the table subscript of offset
// t = i - e*Np + (Ne-1)*Np is ext[t].
static void compute_layout(Plan& P, const int* Np_axis) {
  const int nlat = P.nd - 1;
  const int Na[2] = {P.Nx, P.Ny};
  const int Ne[2] = {P.Nex, P.Ney};
  P.extended = false;
  for (int d = 0; d < nlat; ++d)
    if (Na[d] - (Ne[d] - 1) * Np_axis[d] < 1) P.extended = true;
  std::vector<int> ext[2];
  int Nt[2] = {P.Nx, P.Ny};
  for (int d = 0; d < nlat; ++d) {
    const int off = (Ne[d] - 1) * Np_axis[d];
    ext[d].resize(Na[d] + off);
    for (int t = 0; t < Na[d] + off; ++t) {
      int v = t - off + 1;  // 1-based subscript s - e*Np
      if (P.extended) v = t + 1;
      else if (v <= 0) v = (Na[d] - (Ne[d] - 1) * Np_axis[d] + 1) - v;
      ext[d][t] = v - 1;
    }
    if (P.extended) Nt[d] = Na[d] + off;
  }
  const int64_t plane = (int64_t)Nt[0] * (P.nd == 3 ? Nt[1] : 1) * P.Nz;
  if (plane > INT32_MAX) throw Error(BORN_ERR_UNSUPPORTED, "table too large");
  P.H_plane = (int)plane;
  P.HR = Nt[0];
  P.HY = (P.nd == 3) ? Nt[1] : 1;
  P.Np = Np_axis[0];
  P.Npy = (P.nd == 3) ? Np_axis[1] : 1;
  P.extx = ext[0];
  P.exty = (P.nd == 3) ? ext[1] : std::vector<int>{0};
  P.h_shift.clear();
}

// The full shift map [N_e][N_pts] from the lateral maps (general path only).
static void materialize_shift(Plan& P) {
  const int Ny = (P.nd == 3) ? P.Ny : 1, Ney = (P.nd == 3) ? P.Ney : 1;
  const int offx = (P.Nex - 1) * P.Np, offy = (Ney - 1) * P.Npy;
  P.h_shift.resize((size_t)P.N_e * P.N_pts);
  for (int e = 0; e < P.N_e; ++e) {
    const int ex = e % P.Nex, ey = e / P.Nex;
    int* out = P.h_shift.data() + (size_t)e * P.N_pts;
    for (int z = 0; z < P.Nz; ++z)
      for (int iy = 0; iy < Ny; ++iy) {
        const int ty = (P.nd == 3) ? P.exty[iy - ey * P.Npy + offy] : 0;
        const int base = P.HR * (ty + P.HY * z);
        int* row = out + (size_t)P.Nx * (iy + (size_t)Ny * z);
        for (int i = 0; i < P.Nx; ++i) row[i] = base + P.extx[i - ex * P.Np + offx];
      }
  }
}

static Plan* create_geometry(const born_geometry_desc* d) {
  if (!d) throw Error(BORN_ERR_NUMBER_ARGUMENTS, "descriptor is NULL");
  std::unique_ptr<Plan> P(new Plan());
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev) throw Error(BORN_ERR_DEVICE, "invalid device index");
  P->device = d->device;
  HIP_CHECK(hipSetDevice(d->device));
  HIP_CHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  P->precision = d->precision;
  if (d->N_f_unique <= 0 || !d->frequencies || !d->wavenumbers)
    throw Error(BORN_ERR_INVALID_ARGUMENT, "frequencies and wavenumbers are required");
  P->N_fu = d->N_f_unique;
  int Np_axis[2] = {1, 1};
  const int nlat = d->N_dimensions - 1;
  const int N_elem = d->N_elements_axis[0] * (d->N_dimensions == 3 ? d->N_elements_axis[1] : 1);
  int N_points = d->N_points_axis[0] * d->N_points_axis[1] * (d->N_dimensions == 3 ? d->N_points_axis[2] : 1);
  common_dims(*P, d->N_dimensions, d->N_points_axis, d->N_elements_axis, N_points, N_elem);
  for (int a = 0; a < nlat; ++a) {
    const double r = d->pitch_axis[a] / d->delta_axis[a];
    const int n = (int)std::lround(r);
    if (n < 1 || std::fabs(r - n) > 1e-9 * n)
      throw Error(BORN_ERR_NO_SYMMETRIC_GRIDS_SETUP, "lateral spacing is not an integer fraction of the pitch");
    Np_axis[a] = n;
  }
  build_index_structures(*P, d->N_measurements, d->measurements);
  compute_layout(*P, Np_axis);
  upload_index_structures(*P, P->stream);
  hipStream_t s = P->stream;

  GeomDev g;
  g.nd = d->N_dimensions;
  g.Nx = P->Nx; g.Ny = P->Ny; g.Nz = P->Nz;
  g.ox = P->extended ? (P->Nex - 1) * Np_axis[0] : 0;
  g.oy = (P->extended && g.nd == 3) ? (P->Ney - 1) * Np_axis[1] : 0;
  g.Nx_tab = P->Nx + g.ox;
  g.Ny_tab = P->Ny + g.oy;
  g.dx = d->delta_axis[0];
  g.dy = (g.nd == 3) ? d->delta_axis[1] : 0.0;
  g.dz = d->delta_axis[g.nd - 1];
  g.z0 = d->z0;
  g.Nex = P->Nex; g.Ney = P->Ney;
  g.pitch_x = d->pitch_axis[0];
  g.pitch_y = (g.nd == 3) ? d->pitch_axis[1] : 0.0;
  g.Nfx = d->N_face_points_axis[0];
  g.Nfy = (g.nd == 3) ? d->N_face_points_axis[1] : 1;
  g.dfx = d->face_delta_axis[0];
  g.dfy = (g.nd == 3) ? d->face_delta_axis[1] : 0.0;
  if (g.Nfx <= 0 || g.Nfy <= 0) throw Error(BORN_ERR_INVALID_ARGUMENT, "face discretization must be positive");

  std::vector<double2> kv(P->N_fu);
  for (int f = 0; f < P->N_fu; ++f) kv[f] = make_double2(d->wavenumbers[2 * f], d->wavenumbers[2 * f + 1]);
  std::vector<double2> fw;
  if (d->face_weights) {
    const size_t nfw = (size_t)g.Nfx * g.Nfy * P->N_fu;
    fw.resize(nfw);
    for (size_t i = 0; i < nfw; ++i) fw[i] = make_double2(d->face_weights[2 * i], d->face_weights[2 * i + 1]);
  }
  double2* dk = nullptr;
  double2* dfw = nullptr;
  double2* h64 = nullptr;
  double2* h64tx = nullptr;
  int* d_meas_Nfm = nullptr;
  int* d_tx_off = nullptr;
  int* d_tx_el = nullptr;
  double2* d_vd = nullptr;
  int* d_vd_off = nullptr;
  try {
    dk = dupload(kv, s);
    if (d->face_weights) dfw = dupload(fw, s);
    const size_t nh = (size_t)P->N_fu * P->H_plane;
    h64 = dalloc<double2>(nh);
    g.filt_kind = d->filter_rx_kind;
    g.filt_param = d->filter_rx_param;
    launch_build_h(g, P->N_fu, dk, dfw, h64, s);
    P->d_h = dalloc<float2>(nh);
    launch_f64_to_f32(h64, nh, P->d_h, s);
    const double2* htx = h64;
    if (d->filter_tx_kind != d->filter_rx_kind || d->filter_tx_param != d->filter_rx_param) {
      h64tx = dalloc<double2>(nh);
      g.filt_kind = d->filter_tx_kind;
      g.filt_param = d->filter_tx_param;
      launch_build_h(g, P->N_fu, dk, dfw, h64tx, s);
      htx = h64tx;
    }
    // transmit lists and normal velocities
    if (!d->N_tx_active || !d->tx_indices_active || !d->v_d)
      throw Error(BORN_ERR_INVALID_ARGUMENT, "transmit settings are required");
    std::vector<int> tx_off(d->N_measurements + 1, 0), tx_el, vd_off(d->N_measurements, 0);
    std::vector<double2> vd;
    for (int m = 0; m < d->N_measurements; ++m) {
      const int n = d->N_tx_active[m];
      if (n <= 0) throw Error(BORN_ERR_INVALID_ARGUMENT, "no active transmit element");
      for (int a = 0; a < n; ++a) tx_el.push_back(idx1(d->tx_indices_active[m][a], P->N_e, "tx indices_active"));
      tx_off[m + 1] = tx_off[m] + n;
      vd_off[m] = (int)vd.size();
      const size_t cnt = (size_t)n * P->meas_Nfm[m];
      for (size_t i = 0; i < cnt; ++i) vd.push_back(make_double2(d->v_d[m][2 * i], d->v_d[m][2 * i + 1]));
    }
    d_meas_Nfm = dupload(P->meas_Nfm, s);
    d_tx_off = dupload(tx_off, s);
    d_tx_el = dupload(tx_el, s);
    d_vd = dupload(vd, s);
    d_vd_off = dupload(vd_off, s);
    finish_structure(*P);
    P->d_p = alloc_p_table(*P);
    if (P->structured) {
      launch_synth_pinc_structured(*P, htx, d_meas_Nfm, d_tx_off, d_tx_el, d_vd, d_vd_off, s);
    } else {
      ensure_shift_on_device(*P);
      launch_synth_pinc(*P, htx, d_meas_Nfm, d_tx_off, d_tx_el, d_vd, d_vd_off, s);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    hipStreamSynchronize(s);
    dfree(dk); dfree(dfw); dfree(h64); dfree(h64tx);
    dfree(d_meas_Nfm); dfree(d_tx_off); dfree(d_tx_el); dfree(d_vd); dfree(d_vd_off);
    throw;
  }
  dfree(dk); dfree(dfw); dfree(h64); dfree(h64tx);
  dfree(d_meas_Nfm); dfree(d_tx_off); dfree(d_tx_el); dfree(d_vd); dfree(d_vd_off);
  finish_plan(*P, true);
  if (P->structured) dfree(P->d_shift);  // only the general path reads it
  return P.release();
}

// ----------------------------------------------------------- apply paths
struct TimedLaunch {
  Plan& P;
  hipStream_t s;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(Plan& P_, hipStream_t s_, int k) : P(P_), s(s_), kind(k) {
    if (!P.timing) return;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, s));
  }
  ~TimedLaunch() {
    if (!P.timing || !a) return;
    hipEventRecord(b, s);
    P.timed.push_back({a, b, kind});
  }
};

// FP16 split form: maxima of the tables once per plan, then the operand
// scales of one apply (forward: from the masked gamma; adjoint: from W)
static void fp16_scales(Plan& P, bool forward, int N_rhs, hipStream_t s) {
  if (P.precision != BORN_PRECISION_SPLIT_FP16) return;
  if (!P.f16_maxima) {
    launch_fp16_scales(P, (const float*)P.d_h, (size_t)2 * P.N_fu * P.H_plane, -1, s);
    launch_fp16_scales(P, (const float*)P.d_p, (size_t)2 * (P.N_pairs + 1) * P.N_pts, -2, s);
    P.f16_maxima = true;
  }
  if (forward) {
    launch_fp16_scales(P, (const float*)P.d_gm, (size_t)2 * N_rhs * P.N_pts, 1, s);
  } else {
    const size_t n = P.chunks.size() * (size_t)(P.nd == 3 ? P.Ney : 1) * P.Ne_pad * NC;
    launch_fp16_scales(P, (const float*)P.d_Wc, 2 * n, 2, s);
  }
}

// Contraction launches of a launch set in a numeric form.  Few-column sets
// run on the FP32 vector kernels in the FP32 form, and in the split forms
// when every chunk has at most two columns (where the split MFMA kernels
// fill 1-2 columns in 16 and the FP32 kernels are more accurate and no
// slower).
static bool fc_path(const LaunchSet& L, int precision) {
  return L.fc_w > 0 && (precision == BORN_PRECISION_FP32 || L.fc_w <= 2);
}
static void fwd_contract(const LaunchSet& L, int precision, int n_cu, hipStream_t s) {
  if (L.inst.empty()) return;
  if (fc_path(L, precision)) launch_fwd_fc(L, s);
  else if (precision == BORN_PRECISION_FP32) launch_fwd_structured(L, n_cu, s);
  else launch_fwd_split(L, n_cu, s);
}
static void adj_contract(const LaunchSet& L, int precision, const OutPtrs& outs, hipStream_t s) {
  if (L.inst.empty()) return;
  if (fc_path(L, precision)) launch_adj_fc(L, outs, s);
  else if (precision == BORN_PRECISION_FP32) launch_adj_structured(L, outs, s);
  else launch_adj_split(L, outs, s);
}
// Row-block partial sums per (chunk, y slice) that the forward of L wrote.
static int fwd_blocks(const LaunchSet& L, int precision, const Inst& I) {
  return fc_path(L, precision) || precision == BORN_PRECISION_FP32 ? I.NZB : I.NZB_s * I.nseg_s;
}

static void forward_device(Plan& P, const float2* gm, int N_rhs, void* u, bool out_f64, hipStream_t s) {
  if (use_structured(P)) {
    P.ls.f16 = P.ls_fwd.f16 = P.precision == BORN_PRECISION_SPLIT_FP16;
    fp16_scales(P, true, N_rhs, s);
  }
  if (use_structured(P) && P.fwd_expand) {
    {
      TimedLaunch t(P, s, 0);
      if (P.fwd_table_expanded) launch_expand_pg(P, s);
      fwd_contract(P.ls_fwd, P.precision, P.n_cu, s);
    }
    launch_fwd_reduce_expanded(P, fwd_blocks(P.ls_fwd, P.precision, P.ls_fwd.inst[0]), s);
  } else if (use_structured(P)) {
    (void)gm;   // the launch record points at the masked input P.d_gm
    {
      TimedLaunch t(P, s, 0);
      fwd_contract(P.ls, P.precision, P.n_cu, s);
    }
    launch_fwd_reduce(P, N_rhs, fwd_blocks(P.ls, P.precision, P.ls.inst[0]), s);  // per slice
  } else {
    ensure_shift_on_device(P);
    TimedLaunch t(P, s, 0);
    launch_fwd_general(P, gm, N_rhs, s);
  }
  launch_fwd_mix(P, N_rhs, u, out_f64, s);
  HIP_CHECK(hipGetLastError());
}

static void adjoint_device(Plan& P, const float2* u, int N_rhs, float2* g, hipStream_t s) {
  launch_adj_unmix(P, u, N_rhs, use_structured(P), s);
  if (use_structured(P)) {
    P.ls.f16 = P.precision == BORN_PRECISION_SPLIT_FP16;
    fp16_scales(P, false, N_rhs, s);
  }
  TimedLaunch t(P, s, 1);
  if (use_structured(P)) {
    OutPtrs outs;
    outs.p[0] = g;
    adj_contract(P.ls, P.precision, outs, s);
  } else {
    ensure_shift_on_device(P);
    launch_adj_general(P, N_rhs, g, s);
  }
  HIP_CHECK(hipGetLastError());
}

static void check_precision(const Plan& P) {
  if (P.precision != BORN_PRECISION_FP32 && !(fwd_split_supported(P) && !P.force_general))
    throw Error(BORN_ERR_UNSUPPORTED,
                "split precision needs the structured path with 2 or 4 grid points per pitch and 17 to 128 elements per x row");
}

// Gather path or contraction for a host forward with n_occ occupied points:
// the gather costs about n_occ * sum over mixes of (N_f * N_active) table
// reads, the contraction about N_pairs * N_e * N_pts multiply-adds; the
// gather is taken when clearly cheaper (factor 512), or by the plan's limit.
static bool use_sparse_forward(const Plan& P, size_t n_occ) {
  if (P.sparse_max == 0) return false;
  if (P.sparse_max > 0) return n_occ <= (size_t)P.sparse_max;
  double gather = 0.0;
  for (int mk = 0; mk < P.N_mk; ++mk) gather += (double)P.mk_Nf[mk] * P.mk_Nact[mk];
  return (double)n_occ * gather * 512.0 <= (double)P.N_pairs * P.N_e * P.N_pts;
}

// Host forward through k_fwd_sparse: the occupied points (the MEX rule on
// the doubles, cu:466-488) with their FP64 values, FP64 sums over FP32 tables.
static void host_forward_sparse(Plan& P, const double* gamma, int N_rhs, size_t n_occ, double* u) {
  std::vector<int> pts;
  pts.reserve(n_occ);
  for (int p = 0; p < P.N_pts; ++p) {
    bool occ = false;
    for (int r = 0; r < N_rhs && !occ; ++r) {
      const double* d = gamma + 2 * ((size_t)r * P.N_pts + p);
      occ = std::fabs(d[0]) > 2.220446049250313080847e-16 || std::fabs(d[1]) > 2.220446049250313080847e-16;
    }
    if (occ) pts.push_back(p);
  }
  const int n = (int)pts.size();
  std::vector<double2> g((size_t)n * N_rhs);
  for (int r = 0; r < N_rhs; ++r)
    for (int j = 0; j < n; ++j) {
      const double* d = gamma + 2 * ((size_t)r * P.N_pts + pts[j]);
      g[(size_t)r * n + j] = make_double2(d[0], d[1]);
    }
  const size_t nout = (size_t)P.N_obs * N_rhs;
  ensure_shift_on_device(P);
  ensure_io(P, 0, nout * sizeof(double2));
  int* d_pts = dupload(pts, P.stream);
  double2* d_g = nullptr;
  try {
    d_g = dupload(g, P.stream);
    launch_fwd_sparse(P, d_pts, n, d_g, N_rhs, (double2*)P.d_io_out, P.stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(u, P.d_io_out, nout * sizeof(double2), hipMemcpyDeviceToHost, P.stream));
    HIP_CHECK(hipStreamSynchronize(P.stream));
  } catch (...) {
    hipStreamSynchronize(P.stream);
    dfree(d_pts);
    dfree(d_g);
    throw;
  }
  dfree(d_pts);
  dfree(d_g);
}

static void host_forward(Plan& P, const double* gamma, int N_rhs, double* u) {
  if (!gamma || !u) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input must be a double matrix with N_points rows!");
  if (N_rhs <= 0) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input must be a nonempty double matrix!");
  check_precision(P);
  HIP_CHECK(hipSetDevice(P.device));
  const size_t npr = (size_t)P.N_pts * N_rhs;
  // occupancy on the doubles, as the MEX does before converting (cu:466-488)
  std::vector<float2> buf(npr);
  size_t n_occ = 0;
  for (int p = 0; p < P.N_pts; ++p) {
    bool occ = false;
    for (int r = 0; r < N_rhs && !occ; ++r) {
      const double* v = gamma + 2 * ((size_t)r * P.N_pts + p);
      occ = std::fabs(v[0]) > 2.220446049250313080847e-16 || std::fabs(v[1]) > 2.220446049250313080847e-16;
    }
    n_occ += occ;
    for (int r = 0; r < N_rhs; ++r) {
      const double* v = gamma + 2 * ((size_t)r * P.N_pts + p);
      buf[(size_t)r * P.N_pts + p] = occ ? make_float2((float)v[0], (float)v[1]) : make_float2(0.f, 0.f);
    }
  }
  const size_t nout = (size_t)P.N_obs * N_rhs;
  if (n_occ == 0) {  // quick exit (cu:496): zeros
    std::memset(u, 0, sizeof(double) * 2 * nout);
    return;
  }
  if (use_sparse_forward(P, n_occ)) {
    host_forward_sparse(P, gamma, N_rhs, n_occ, u);
    return;
  }
  ensure_workspace(P, N_rhs);
  ensure_io(P, npr * sizeof(float2), nout * sizeof(double2));
  // the occupancy mask is applied on the host: upload straight to the masked-input buffer
  HIP_CHECK(hipMemcpyAsync(P.d_gm, buf.data(), npr * sizeof(float2), hipMemcpyHostToDevice, P.stream));
  forward_device(P, P.d_gm, N_rhs, P.d_io_out, true, P.stream);
  HIP_CHECK(hipMemcpyAsync(u, P.d_io_out, nout * sizeof(double2), hipMemcpyDeviceToHost, P.stream));
  HIP_CHECK(hipStreamSynchronize(P.stream));
}

static void host_adjoint(Plan& P, const double* uin, int N_rhs, double* g) {
  if (!uin || !g) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input must be a double matrix with N_observations rows!");
  if (N_rhs <= 0) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input must be a nonempty double matrix!");
  check_precision(P);
  HIP_CHECK(hipSetDevice(P.device));
  const size_t nin = (size_t)P.N_obs * N_rhs, nout = (size_t)P.N_pts * N_rhs;
  std::vector<float2> buf(std::max(nin, nout));
  for (size_t i = 0; i < nin; ++i) buf[i] = make_float2((float)uin[2 * i], (float)uin[2 * i + 1]);
  ensure_workspace(P, N_rhs);
  ensure_io(P, nin * sizeof(float2), nout * sizeof(float2));
  HIP_CHECK(hipMemcpyAsync(P.d_io_in, buf.data(), nin * sizeof(float2), hipMemcpyHostToDevice, P.stream));
  adjoint_device(P, (const float2*)P.d_io_in, N_rhs, (float2*)P.d_io_out, P.stream);
  HIP_CHECK(hipMemcpyAsync(buf.data(), P.d_io_out, nout * sizeof(float2), hipMemcpyDeviceToHost, P.stream));
  HIP_CHECK(hipStreamSynchronize(P.stream));
  for (size_t i = 0; i < nout; ++i) {
    g[2 * i] = buf[i].x;
    g[2 * i + 1] = buf[i].y;
  }
}

// ------------------------------------------------------------- batches
// A batch runs several plans (operator instances of uneven size) with one
// persistent contraction launch per direction: the instance records of all
// plans go into one launch set, the chunk tables are concatenated.  The
// per-plan masking, reduction, mixing and un-mixing kernels are launched
// around it.  With graph replay on, the whole sequence is captured once per
// (stream, buffers) into a hipGraph and replayed.
struct Batch {
  int device = 0;
  int n_cu = 256;
  int precision = 0;
  std::vector<Plan*> plans;
  std::vector<uint64_t> gen;     // workspace generation of each plan when the records were made
  // instances whose chunks hold at most 1, 2, 4 or 3 columns run in a launch
  // of their own per width on the few-column kernels (sets 1-3, 7), the others in
  // MFMA launches: set 0, and in the FP32 form set 4 for the instances whose
  // rows fill 128-point adjoint segments better than 256-point ones (rows of
  // at most 128 points; 384 points: 3 segments of 128 instead of 2 of 256,
  // one half empty); FP32 instances of at most 8 columns per chunk take the
  // stacked kernels in sets 5 (like 0) and 6 (like 4)
  static constexpr int N_SETS = 8;
  LaunchSet sets[N_SETS];
  std::vector<int> set_of;         // per plan: its launch set
  std::vector<int> slot;           // per plan: its index in its launch set
  Chunk* d_chunks[N_SETS] = {};
  int* d_chunk_pairs[N_SETS] = {};
  LaunchSet& set(int k) { return sets[k]; }
  static bool mfma_set(int k) { return k == 0 || (k >= 4 && k <= 6); }
  bool graph = false;
  struct Captured {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    std::vector<const void*> sig;
  } cap[2];
  bool timing = false;
  std::vector<Plan::Timed> timed;
  // graphs cannot be captured on the legacy default stream: work submitted
  // there runs on this stream, ordered by events
  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  ~Batch();
};

static void drop_graph(Batch::Captured& c) {
  if (c.exec) hipGraphExecDestroy(c.exec);
  if (c.g) hipGraphDestroy(c.g);
  c.exec = nullptr;
  c.g = nullptr;
  c.sig.clear();
}

Batch::~Batch() {
  hipSetDevice(device);
  for (auto& c : cap) drop_graph(c);
  for (auto& t : timed) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  for (int k = 0; k < N_SETS; ++k) {
    dfree(set(k).d_inst);
    dfree(d_chunks[k]);
    dfree(d_chunk_pairs[k]);
  }
  if (ev_in) hipEventDestroy(ev_in);
  if (ev_out) hipEventDestroy(ev_out);
  if (own) {
    hipStreamSynchronize(own);
    hipStreamDestroy(own);
  }
}

// Instance records, concatenated chunk tables and ranges of the batch's
// launch sets.
static void build_batch_set(Batch& B, int set) {
  LaunchSet& L = B.set(set);
  Chunk*& d_chunks = B.d_chunks[set];
  int*& d_chunk_pairs = B.d_chunk_pairs[set];
  L.inst.clear();
  size_t n_chunks = 0;
  for (size_t i = 0; i < B.plans.size(); ++i)
    if (B.set_of[i] == set) n_chunks += B.plans[i]->chunks.size();
  dfree(d_chunks);
  dfree(d_chunk_pairs);
  d_chunks = nullptr;
  d_chunk_pairs = nullptr;
  if (n_chunks == 0) {
    L.fc_w = 0;
    return;
  }
  d_chunks = dalloc<Chunk>(n_chunks);
  d_chunk_pairs = dalloc<int>(n_chunks * NC);
  size_t base = 0;
  bool small = true;
  const Plan* first = nullptr;
  for (size_t i = 0; i < B.plans.size(); ++i) {
    if (B.set_of[i] != set) continue;
    Plan* P = B.plans[i];
    const size_t nc = P->chunks.size();
    HIP_CHECK(hipMemcpy(d_chunks + base, P->d_chunks, nc * sizeof(Chunk), hipMemcpyDeviceToDevice));
    HIP_CHECK(hipMemcpy(d_chunk_pairs + base * NC, P->d_chunk_pairs, nc * NC * sizeof(int),
                        hipMemcpyDeviceToDevice));
    Inst I = make_inst(*P);
    I.chunk_base = (int)base;
    B.slot[i] = (int)L.inst.size();
    L.inst.push_back(I);
    base += nc;
    small = small && P->Nx <= 128;
    if (!first) first = P;
  }
  L.d_chunks = d_chunks;
  L.d_chunk_pairs = d_chunk_pairs;
  L.MT = first->Ne_pad / 16;
  L.Np = first->Np;
  if (set == 4 || set == 6) L.adj_mw = 2;
  else if (L.adj_mw != 2 && L.adj_mw != 4) L.adj_mw = small ? 2 : 4;
  finalize_launchset(L);
}

static void build_batch_launch(Batch& B) {
  B.gen.clear();
  B.set_of.assign(B.plans.size(), 0);
  B.slot.assign(B.plans.size(), 0);
  for (size_t i = 0; i < B.plans.size(); ++i) {
    const Plan& P = *B.plans[i];
    const int w = fewcol_width(chunk_max_cols(P.chunks));
    // (the split forms run few-column sets on the vector kernels only for at
    // most two columns, so their 3-4-column plans join the MFMA set)
    const bool fc = w > 0 && (B.precision == BORN_PRECISION_FP32 || w <= 2);
    if (fc && P.d_gpart && fewcol_fits(P.Nx, P.Np, P.Ne_pad, w, P.n_rows)) B.set_of[i] = w == 1 ? 1 : (w == 2 ? 2 : (w == 3 ? 7 : 3));
    else if (B.precision == BORN_PRECISION_FP32) {
      const bool mw2 = adj_mw_default(P.Nx) == 2, stk = chunk_max_cols(P.chunks) <= NC / 2;
      B.set_of[i] = stk ? (mw2 ? 6 : 5) : (mw2 ? 4 : 0);
    }
    B.gen.push_back(P.ws_gen);
  }
  for (int k = 0; k < Batch::N_SETS; ++k) build_batch_set(B, k);
  for (auto& c : B.cap) drop_graph(c);
}

static void batch_forward_enqueue(Batch& B, const void* const* in, void* const* out, hipStream_t s) {
  for (size_t i = 0; i < B.plans.size(); ++i) {
    Plan& P = *B.plans[i];
    launch_mask_gamma((const float2*)in[i], P.N_pts, 1, P.d_gm, s);
    fp16_scales(P, true, 1, s);
  }
  for (int k = 0; k < Batch::N_SETS; ++k) B.set(k).f16 = B.precision == BORN_PRECISION_SPLIT_FP16;
  hipEvent_t a = nullptr, b = nullptr;
  if (B.timing) {
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, s));
  }
  for (int k = 0; k < Batch::N_SETS; ++k) fwd_contract(B.set(k), B.precision, B.n_cu, s);
  if (B.timing) {
    HIP_CHECK(hipEventRecord(b, s));
    B.timed.push_back({a, b, 0});
  }
  for (size_t i = 0; i < B.plans.size(); ++i) {
    Plan& P = *B.plans[i];
    const LaunchSet& L = B.set(B.set_of[i]);
    launch_fwd_reduce(P, 1, fwd_blocks(L, B.precision, L.inst[B.slot[i]]), s);
    launch_fwd_mix(P, 1, out[i], false, s);
  }
  HIP_CHECK(hipGetLastError());
}

static void batch_adjoint_enqueue(Batch& B, const void* const* in, void* const* out, hipStream_t s) {
  OutPtrs outs[Batch::N_SETS];
  for (size_t i = 0; i < B.plans.size(); ++i) {
    Plan& P = *B.plans[i];
    launch_adj_unmix(P, (const float2*)in[i], 1, true, s);
    fp16_scales(P, false, 1, s);
    outs[B.set_of[i]].p[B.slot[i]] = (float2*)out[i];
  }
  for (int k = 0; k < Batch::N_SETS; ++k) B.set(k).f16 = B.precision == BORN_PRECISION_SPLIT_FP16;
  hipEvent_t a = nullptr, b = nullptr;
  if (B.timing) {
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(a, s));
  }
  for (int k = 0; k < Batch::N_SETS; ++k) adj_contract(B.set(k), B.precision, outs[k], s);
  if (B.timing) {
    HIP_CHECK(hipEventRecord(b, s));
    B.timed.push_back({a, b, 1});
  }
  HIP_CHECK(hipGetLastError());
}

// Run one direction, replaying a captured graph when the stream and the
// buffers are those it was captured with.
// A plan of the batch applied on its own with another rhs count gets a new
// workspace, which frees the one the batch's records point at: rebuild the
// plan's one-rhs workspace and the records (and drop the captured graphs).
static void batch_refresh(Batch& B) {
  bool stale = false;
  for (size_t i = 0; i < B.plans.size(); ++i) {
    Plan& P = *B.plans[i];
    if (P.precision != B.precision) throw Error(BORN_ERR_UNSUPPORTED, "a plan of the batch changed its precision");
    if (P.ws_rhs != 1 || P.ws_gen != B.gen[i]) {
      ensure_workspace(P, 1);
      stale = true;
    }
  }
  if (stale) build_batch_launch(B);
}

static void batch_apply(Batch& B, int dir, const void* const* in, void* const* out, hipStream_t s) {
  batch_refresh(B);
  const size_t n = B.plans.size();
  if (!B.graph || B.timing) {
    if (dir == 0) batch_forward_enqueue(B, in, out, s);
    else batch_adjoint_enqueue(B, in, out, s);
    return;
  }
  hipStream_t user = s;
  if (!user) {
    HIP_CHECK(hipEventRecord(B.ev_in, nullptr));
    HIP_CHECK(hipStreamWaitEvent(B.own, B.ev_in, 0));
    s = B.own;
  }
  std::vector<const void*> sig(in, in + n);
  sig.insert(sig.end(), out, out + n);
  Batch::Captured& c = B.cap[dir];
  if (!c.exec || c.stream != s || c.sig != sig) {
    drop_graph(c);
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      if (dir == 0) batch_forward_enqueue(B, in, out, s);
      else batch_adjoint_enqueue(B, in, out, s);
    } catch (...) {
      hipGraph_t g = nullptr;
      hipStreamEndCapture(s, &g);
      if (g) hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(s, &c.g));
    HIP_CHECK(hipGraphInstantiate(&c.exec, c.g, nullptr, nullptr, 0));
    c.stream = s;
    c.sig = sig;
  }
  HIP_CHECK(hipGraphLaunch(c.exec, s));
  if (!user) {
    HIP_CHECK(hipEventRecord(B.ev_out, s));
    HIP_CHECK(hipStreamWaitEvent(nullptr, B.ev_out, 0));
  }
}

// Autotuner: times the contraction launches of the candidate work-item
// counts and persistent grid sizes (forward) and segment widths (plain
// adjoint), keeps the fastest.  The timed launches read the plans' current
// workspaces; their results are discarded.
static float time_launches(const std::function<void()>& f, hipStream_t s, int reps) {
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  f();  // warm-up
  HIP_CHECK(hipEventRecord(a, s));
  for (int r = 0; r < reps; ++r) f();
  HIP_CHECK(hipEventRecord(b, s));
  HIP_CHECK(hipEventSynchronize(b));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms / reps;
}

// New forward blocking for every plan; vpart grows when needed.
static void set_items_target(std::vector<Plan*>& plans, int target) {
  for (Plan* P : plans) {
    if (P->items_target == target) continue;
    P->items_target = target;
    const size_t have = P->chunks.size() * (size_t)P->vblk;
    set_blocking(*P);
    if (P->chunks.size() * (size_t)P->vblk > have) {
      HIP_CHECK(hipDeviceSynchronize());
      dfree(P->d_vpart);
      P->d_vpart = dalloc<float2>(P->chunks.size() * (size_t)P->vblk * P->Ne_pad * NC);
    } else {
      P->vblk = (int)(have / std::max<size_t>(1, P->chunks.size()));   // capacity of the allocation
    }
    P->ls.inst.assign(1, make_inst(*P));
    finalize_launchset(P->ls);
  }
}

struct TuneResult {
  int items, grid_mult, adj_mw;
  float fwd_ms, adj_ms;
};

static TuneResult autotune(std::vector<Plan*>& plans, LaunchSet& L, int precision, int n_cu,
                           const std::function<void()>& rebuild, hipStream_t s) {
  const int items_c[] = {1024, 2048, 4096, 8192};
  const int grid_c[] = {1, 2, 3};   // persistent workgroups per CU (3: where registers and LDS allow)
  TuneResult best{plans[0]->items_target, L.grid_mult, L.adj_mw, 1e30f, 1e30f};
  L.f16 = precision == BORN_PRECISION_SPLIT_FP16;
  OutPtrs outs;
  std::vector<float2*> scratch;
  try {
    // instance i of L need not be plan i (a batch splits its plans over two
    // launch sets): every buffer holds the largest plan
    int max_pts = 0;
    for (const Plan* P : plans) max_pts = std::max(max_pts, P->N_pts);
    for (size_t i = 0; i < plans.size(); ++i) {
      float2* g = dalloc<float2>(max_pts);
      scratch.push_back(g);
      outs.p[i] = g;
    }
    for (int items : items_c) {
      set_items_target(plans, items);
      rebuild();
      for (int gm : grid_c) {
        L.grid_mult = gm;
        const float t = time_launches([&] { fwd_contract(L, precision, n_cu, s); }, s, 3);
        if (t < best.fwd_ms) {
          best.fwd_ms = t;
          best.items = items;
          best.grid_mult = gm;
        }
      }
    }
    set_items_target(plans, best.items);
    L.grid_mult = best.grid_mult;
    const int mw_c[] = {2, 4};
    for (int mw : mw_c) {
      if (precision != BORN_PRECISION_FP32 && mw != 4) continue;   // the split adjoint has one width
      L.adj_mw = mw;
      rebuild();
      const float t = time_launches([&] { adj_contract(L, precision, outs, s); }, s, 3);
      if (t < best.adj_ms) {
        best.adj_ms = t;
        best.adj_mw = mw;
      }
    }
    L.adj_mw = best.adj_mw;
    rebuild();
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    hipStreamSynchronize(s);
    for (float2* g : scratch) dfree(g);
    throw;
  }
  for (float2* g : scratch) dfree(g);
  return best;
}

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return BORN_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return BORN_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BORN_ERR_DEVICE;
  }
}

}  // namespace born

using namespace born;

extern "C" {

int born_plan_create_tables(const born_tables_desc* desc, born_plan_t* plan) {
  if (!plan) return BORN_ERR_NUMBER_ARGUMENTS;
  *plan = nullptr;
  return guarded([&] { *plan = reinterpret_cast<born_plan_t>(create_tables(desc)); });
}

int born_plan_create_geometry(const born_geometry_desc* desc, born_plan_t* plan) {
  if (!plan) return BORN_ERR_NUMBER_ARGUMENTS;
  *plan = nullptr;
  return guarded([&] { *plan = reinterpret_cast<born_plan_t>(create_geometry(desc)); });
}

int born_plan_destroy(born_plan_t plan) {
  return guarded([&] { delete reinterpret_cast<Plan*>(plan); });
}

int born_plan_size(born_plan_t plan, int64_t* N_observations, int64_t* N_points) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
    Plan* P = reinterpret_cast<Plan*>(plan);
    if (N_observations) *N_observations = P->N_obs;
    if (N_points) *N_points = P->N_pts;
  });
}

int born_plan_set_precision(born_plan_t plan, int32_t precision) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
    if (precision < 0 || precision > 2) throw Error(BORN_ERR_INVALID_ARGUMENT, "unknown precision");
    reinterpret_cast<Plan*>(plan)->precision = precision;
  });
}

int born_forward(born_plan_t plan, const double* gamma, int32_t N_rhs, double* u) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    host_forward(*reinterpret_cast<Plan*>(plan), gamma, N_rhs, u);
  });
}

int born_adjoint(born_plan_t plan, const double* u, int32_t N_rhs, double* gamma_hat) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    host_adjoint(*reinterpret_cast<Plan*>(plan), u, N_rhs, gamma_hat);
  });
}

int born_forward_dev(born_plan_t plan, const void* gamma_c64, int32_t N_rhs, void* u_c64, void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
    if (!gamma_c64 || !u_c64 || N_rhs <= 0) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "invalid device buffers");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    check_precision(P);
    HIP_CHECK(hipSetDevice(P.device));
    hipStream_t s = (hipStream_t)stream;
    ensure_workspace(P, N_rhs);
    launch_mask_gamma((const float2*)gamma_c64, P.N_pts, N_rhs, P.d_gm, s);
    forward_device(P, P.d_gm, N_rhs, u_c64, false, s);
  });
}

int born_adjoint_dev(born_plan_t plan, const void* u_c64, int32_t N_rhs, void* gamma_hat_c64, void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
    if (!u_c64 || !gamma_hat_c64 || N_rhs <= 0) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "invalid device buffers");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    check_precision(P);
    HIP_CHECK(hipSetDevice(P.device));
    ensure_workspace(P, N_rhs);
    adjoint_device(P, (const float2*)u_c64, N_rhs, (float2*)gamma_hat_c64, (hipStream_t)stream);
  });
}

int born_apply_dev_noalloc(born_plan_t plan, int32_t mode, const void* in_c64, int32_t N_rhs, void* out_c64,
                           void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
    if (!in_c64 || !out_c64 || N_rhs <= 0) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "invalid device buffers");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    check_precision(P);
    if (!use_structured(P)) throw Error(BORN_ERR_UNSUPPORTED, "the allocation-free apply needs the structured path");
    if (!P.na.ok) throw Error(BORN_ERR_UNSUPPORTED, P.na.why);
    HIP_CHECK(hipSetDevice(P.device));
    hipStream_t s = (hipStream_t)stream;
    const bool forward = mode == 1;
    const size_t rows_in = forward ? (size_t)P.N_pts : (size_t)P.N_obs;
    const size_t rows_out = forward ? (size_t)P.N_obs : (size_t)P.N_pts;
    NoAllocView view(P);
    for (int r = 0; r < N_rhs; ++r) {   // one right-hand side at a time through the one-rhs layout
      const float2* in = static_cast<const float2*>(in_c64) + r * rows_in;
      float2* out = static_cast<float2*>(out_c64) + r * rows_out;
      if (forward) {
        launch_mask_gamma(in, P.N_pts, 1, P.d_gm, s);
        forward_device(P, P.d_gm, 1, out, false, s);
      } else {
        launch_zero_f2(P.d_Wc, P.na.Wc_n, s);   // unused chunk columns and element padding read as zero
        adjoint_device(P, in, 1, out, s);
      }
    }
  });
}

static void energy_rx(Plan& P, int m, double* E_dev, hipStream_t s) {
  if (m < 0 || m >= (int)P.meas_pair_base.size())
    throw Error(BORN_ERR_INVALID_ARGUMENT, "operator_born.indices_measurement_sel must be a single index!");
  ensure_shift_on_device(P);
  launch_energy_rx(P, m, E_dev, s);
  HIP_CHECK(hipGetLastError());
}

int born_energy_rx(born_plan_t plan, int32_t measurement, double* energy) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!energy) throw Error(BORN_ERR_INVALID_ARGUMENT, "energy output is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    double* d = dalloc<double>((size_t)P.N_pts);
    try {
      energy_rx(P, measurement, d, P.stream);
      HIP_CHECK(hipMemcpyAsync(energy, d, (size_t)P.N_pts * sizeof(double), hipMemcpyDeviceToHost, P.stream));
      HIP_CHECK(hipStreamSynchronize(P.stream));
    } catch (...) {
      dfree(d);
      throw;
    }
    dfree(d);
  });
}

// Energy under the plan's TGC; synchronous on s (a spill buffer for long
// mixes lives for the launch only).
static void energy_rx_tgc(Plan& P, int m, double* E_dev, hipStream_t s) {
  if (m < 0 || m >= (int)P.meas_pair_base.size())
    throw Error(BORN_ERR_INVALID_ARGUMENT, "operator_born.indices_measurement_sel must be a single index!");
  if (!P.tgc) throw Error(BORN_ERR_INVALID_ARGUMENT, "no TGC kernels are set on this operator");
  ensure_shift_on_device(P);
  int rows = 1;
  for (int mk = 0; mk < P.N_mk; ++mk)
    if (P.mk_meas[mk] == m) rows = std::max(rows, P.mk_Nf[mk]);
  if (rows <= ETGC_LDS_ROWS) {
    launch_energy_rx_tgc(P, m, rows, nullptr, 0, E_dev, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  const int blocks = std::min((P.N_pts + 63) / 64, 1024);
  float2* spill = dalloc<float2>((size_t)blocks * rows * 64);
  try {
    launch_energy_rx_tgc(P, m, rows, spill, blocks, E_dev, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    dfree(spill);
    throw;
  }
  dfree(spill);
}

int born_energy_rx_tgc(born_plan_t plan, int32_t measurement, double* energy) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!energy) throw Error(BORN_ERR_INVALID_ARGUMENT, "energy output is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    double* d = dalloc<double>((size_t)P.N_pts);
    try {
      energy_rx_tgc(P, measurement, d, P.stream);
      HIP_CHECK(hipMemcpy(energy, d, (size_t)P.N_pts * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
      dfree(d);
      throw;
    }
    dfree(d);
  });
}

int born_energy_rx_tgc_dev(born_plan_t plan, int32_t measurement, void* energy_f64, void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!energy_f64) throw Error(BORN_ERR_INVALID_ARGUMENT, "energy output is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    energy_rx_tgc(P, measurement, (double*)energy_f64, (hipStream_t)stream);
  });
}

int born_energy_rx_dev(born_plan_t plan, int32_t measurement, void* energy_f64, void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!energy_f64) throw Error(BORN_ERR_INVALID_ARGUMENT, "energy output is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    energy_rx(P, measurement, (double*)energy_f64, (hipStream_t)stream);
  });
}

int born_time_signals(born_plan_t plan, const double* frequencies, double delta_f, double delta_t,
                      const void* u_c64, int32_t N_rhs, void* rf_f32, int64_t capacity, int32_t* N_t_out,
                      void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!frequencies || !u_c64 || !rf_f32 || N_rhs <= 0)
      throw Error(BORN_ERR_INVALID_ARGUMENT, "frequencies, voltages and output are required");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    // regular quantized frequency axis: f = q * delta_f with integer q
    if (delta_f <= 0.0) {
      std::vector<double> f(frequencies, frequencies + P.N_fu);
      std::sort(f.begin(), f.end());
      delta_f = 0.0;
      for (size_t i = 0; i + 1 < f.size(); ++i) {
        const double d = f[i + 1] - f[i];
        if (d > 0.0 && (delta_f == 0.0 || d < delta_f)) delta_f = d;
      }
      if (delta_f == 0.0) delta_f = f[0];
    }
    std::vector<int> q_fg(P.N_fu);
    int q_max = 0;
    for (int i = 0; i < P.N_fu; ++i) {
      const double q = frequencies[i] / delta_f;
      if (q < -0.5 || std::fabs(q - std::round(q)) > 1e-6 * std::max(1.0, std::fabs(q)))
        throw Error(BORN_ERR_UNSUPPORTED, "signal: the frequencies are not integer multiples of delta_f");
      q_fg[i] = (int)std::lround(q);
      q_max = std::max(q_max, q_fg[i]);
    }
    int N_t;
    if (delta_t > 0.0) {
      const double n = 1.0 / (delta_f * delta_t);
      N_t = (int)std::lround(n);
      if (std::fabs(n - N_t) > 1e-6 * std::max(1.0, n))
        throw Error(BORN_ERR_INVALID_ARGUMENT, "Ts_rec must be integer multiples of delta!");
    } else {
      N_t = 2 * q_max;   // delta = 1 / (2 q_ub delta_f) (signal_matrix.m:333)
    }
    if (N_t <= 0 || q_max > N_t / 2)
      throw Error(BORN_ERR_INVALID_ARGUMENT, "Number of spectral samples exceeds the maximum!");
    const size_t need = (size_t)N_rhs * P.N_mk * N_t;
    if (capacity < (int64_t)need) throw Error(BORN_ERR_INVALID_ARGUMENT, "output capacity too small");
    std::vector<int> q_obs(P.N_obs);
    for (int mk = 0; mk < P.N_mk; ++mk) {
      std::vector<int> qs(P.mk_Nf[mk]);
      for (int f = 0; f < P.mk_Nf[mk]; ++f) {
        const int pr = P.meas_pair_base[P.mk_meas[mk]] + P.fmeas[P.mk_f_off[mk] + f];
        qs[f] = q_obs[P.mk_obs_off[mk] + f] = q_fg[P.pair_fg[pr]];
      }
      std::sort(qs.begin(), qs.end());
      if (std::adjacent_find(qs.begin(), qs.end()) != qs.end())
        throw Error(BORN_ERR_UNSUPPORTED, "signal: a mix repeats a frequency, so it has no frequency axis");
    }
    const hipStream_t s = (hipStream_t)stream;
    int* d_q = dupload(q_obs, s);
    try {
      launch_time_signals(P, (const float2*)u_c64, d_q, N_rhs, N_t, (float*)rf_f32, s);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      dfree(d_q);
      throw;
    }
    dfree(d_q);
    if (N_t_out) *N_t_out = N_t;
  });
}

// Columns of A at 1-based points; out complex FP64 column-major N_obs x n
// (host) or complex64 n x N_obs (device); complete on return.
static void columns(Plan& P, const double* points, int32_t n, void* out, bool host, hipStream_t s) {
  if (!points || !out) throw Error(BORN_ERR_INVALID_ARGUMENT, "points and output are required");
  if (n <= 0) throw Error(BORN_ERR_INVALID_ARGUMENT, "at least one point is required");
  if ((double)n * (double)P.N_obs > 2.0e11) throw Error(BORN_ERR_INVALID_ARGUMENT, "too many columns for one call");
  std::vector<int> pts(n);
  for (int j = 0; j < n; ++j) pts[j] = idx1(points[j], P.N_pts, "points");
  ensure_shift_on_device(P);
  int* d_pts = dupload(pts, s);
  void* d_out = nullptr;
  try {
    const size_t bytes = (size_t)n * P.N_obs * (host ? sizeof(double2) : sizeof(float2));
    d_out = host ? (void*)dalloc<char>(bytes) : out;
    launch_columns(P, d_pts, n, d_out, host, s);
    HIP_CHECK(hipGetLastError());
    if (host) HIP_CHECK(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    hipStreamSynchronize(s);
    dfree(d_pts);
    if (host) dfree(d_out);
    throw;
  }
  dfree(d_pts);
  if (host) dfree(d_out);
}

int born_columns(born_plan_t plan, const double* points, int32_t n, double* out) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    columns(P, points, n, out, true, P.stream);
  });
}

int born_columns_dev(born_plan_t plan, const double* points, int32_t n, void* out_c64, void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    columns(P, points, n, out_c64, false, (hipStream_t)stream);
  });
}

static void tgc_clear(Plan& P) {
  dfree(P.d_tgc_tiles);
  dfree(P.d_tgc_half);
  dfree(P.d_tgc_off);
  dfree(P.d_tgc_k32);
  dfree(P.d_tgc_k64);
  P.tgc = false;
  P.n_tgc_tiles = 0;
}

int born_tgc_set(born_plan_t plan, const int32_t* half_lengths, const double* kernels_c128) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    if (P.stream) HIP_CHECK(hipStreamSynchronize(P.stream));
    HIP_CHECK(hipDeviceSynchronize());   // no launch may still read the old taps
    tgc_clear(P);
    if (!half_lengths) return;           // identity (regularization.tgc.off)
    if (!kernels_c128) throw Error(BORN_ERR_INVALID_ARGUMENT, "TGC kernels are NULL");
    std::vector<int> half(P.N_mk);
    std::vector<int64_t> off(P.N_mk);
    std::vector<double2> k64;
    std::vector<int2> tiles;
    int64_t src = 0;
    for (int mk = 0; mk < P.N_mk; ++mk) {
      const int M = half_lengths[mk];
      const int N = P.mk_Nf[mk];
      if (M < 0) throw Error(BORN_ERR_INVALID_ARGUMENT, "TGC kernels must have odd lengths 2 M + 1");
      // a kernel reaching past the mix is cut to 2 N - 1 taps (convolution.m:76-90)
      const int Mc = std::min(M, N - 1);
      const int64_t cut = M - Mc;
      half[mk] = Mc;
      off[mk] = (int64_t)k64.size();
      for (int64_t t = cut; t <= 2 * (int64_t)M - cut; ++t) {
        const double* k = kernels_c128 + 2 * (src + t);
        if (!std::isfinite(k[0]) || !std::isfinite(k[1]))
          throw Error(BORN_ERR_INVALID_ARGUMENT, "TGC kernels must be finite");
        k64.push_back(make_double2(k[0], k[1]));
      }
      src += 2 * (int64_t)M + 1;
      for (int i0 = 0; i0 < N; i0 += 256) tiles.push_back(make_int2(mk, i0));
    }
    std::vector<float2> k32(k64.size());
    for (size_t i = 0; i < k64.size(); ++i) k32[i] = make_float2((float)k64[i].x, (float)k64[i].y);
    try {
      P.d_tgc_half = dupload(half, P.stream);
      P.d_tgc_off = dupload(off, P.stream);
      P.d_tgc_tiles = dupload(tiles, P.stream);
      P.d_tgc_k64 = dupload(k64, P.stream);
      P.d_tgc_k32 = dupload(k32, P.stream);
    } catch (...) {
      tgc_clear(P);
      throw;
    }
    P.n_tgc_tiles = (int)tiles.size();
    P.tgc = true;
  });
}

static void tgc_check(const Plan& P, int32_t N_rhs) {
  if (!P.tgc) throw Error(BORN_ERR_INVALID_ARGUMENT, "no TGC kernels are set on this operator");
  if (N_rhs <= 0 || N_rhs > 65535) throw Error(BORN_ERR_INVALID_ARGUMENT, "N_rhs must lie in [1, 65535]");
}

int born_tgc_apply(born_plan_t plan, int32_t mode, const double* in, int32_t N_rhs, double* out) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!in || !out) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input and output are required");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    tgc_check(P, N_rhs);
    const size_t n = (size_t)N_rhs * P.N_obs;
    double2* d = dalloc<double2>(2 * n);
    try {
      HIP_CHECK(hipMemcpyAsync(d, in, n * sizeof(double2), hipMemcpyHostToDevice, P.stream));
      launch_tgc(P, mode != 1, true, d, d + n, N_rhs, P.stream);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(out, d + n, n * sizeof(double2), hipMemcpyDeviceToHost, P.stream));
      HIP_CHECK(hipStreamSynchronize(P.stream));
    } catch (...) {
      dfree(d);
      throw;
    }
    dfree(d);
  });
}

int born_tgc_apply_dev(born_plan_t plan, int32_t mode, const void* in_c64, void* out_c64, int32_t N_rhs,
                       void* stream) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    if (!in_c64 || !out_c64) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "input and output are required");
    if (in_c64 == out_c64) throw Error(BORN_ERR_INVALID_ARGUMENT, "TGC input and output must not be the same buffer");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    tgc_check(P, N_rhs);
    launch_tgc(P, mode != 1, false, in_c64, out_c64, N_rhs, (hipStream_t)stream);
    HIP_CHECK(hipGetLastError());
  });
}

static void table_shape(const Plan& P, int which, int index, int64_t& rows, int64_t& cols) {
  if (which == 0) {
    rows = P.N_fu;
    cols = P.H_plane;
  } else if (which == 1) {
    if (index < 0 || index >= (int)P.meas_Nfm.size()) throw Error(BORN_ERR_INVALID_ARGUMENT, "measurement index");
    rows = P.meas_Nfm[index];
    cols = P.N_pts;
  } else {
    throw Error(BORN_ERR_INVALID_ARGUMENT, "unknown table");
  }
}

int born_plan_table_shape(born_plan_t plan, int32_t which, int32_t index, int64_t* rows, int64_t* cols) {
  return guarded([&] {
    if (!plan || !rows || !cols) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan and outputs are required");
    table_shape(*reinterpret_cast<Plan*>(plan), which, index, *rows, *cols);
  });
}

int born_plan_get_table(born_plan_t plan, int32_t which, int32_t index, void* out_c64, int64_t capacity) {
  return guarded([&] {
    if (!plan || !out_c64) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan and output are required");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    int64_t rows = 0, cols = 0;
    table_shape(P, which, index, rows, cols);
    const float2* src = which == 0 ? P.d_h : P.d_p + (size_t)P.meas_pair_base[index] * P.N_pts;
    const size_t n = (size_t)rows * cols;
    if ((int64_t)n > capacity) throw Error(BORN_ERR_INVALID_ARGUMENT, "output capacity too small");
    std::vector<float2> tmp(n);
    HIP_CHECK(hipMemcpy(tmp.data(), src, n * sizeof(float2), hipMemcpyDeviceToHost));
    float2* o = (float2*)out_c64;
    for (int64_t r = 0; r < rows; ++r)
      for (int64_t p = 0; p < cols; ++p) o[(size_t)p * rows + r] = tmp[(size_t)r * cols + p];
  });
}

int born_plan_path(born_plan_t plan) {
  if (!plan) return -1;
  return use_structured(*reinterpret_cast<Plan*>(plan)) ? 0 : 1;
}

int born_plan_set_sparse_forward(born_plan_t plan, int32_t max_points) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_NO_OPERATOR_BORN, "operator_born must be a single scattering.operator_born!");
    reinterpret_cast<Plan*>(plan)->sparse_max = max_points < 0 ? -1 : max_points;
  });
}

int born_plan_force_general(born_plan_t plan, int32_t on) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    P.force_general = on != 0;
    if (P.force_general) ensure_shift_on_device(P);
    HIP_CHECK(hipStreamSynchronize(P.stream));
  });
}

int born_plan_set_timing(born_plan_t plan, int32_t on) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan is NULL");
    reinterpret_cast<Plan*>(plan)->timing = on != 0;
  });
}

int born_plan_get_timing(born_plan_t plan, double* ms, int64_t* launches) {
  return guarded([&] {
    if (!plan || !ms || !launches) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan and outputs are required");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    HIP_CHECK(hipSetDevice(P.device));
    ms[0] = ms[1] = 0.0;
    launches[0] = launches[1] = 0;
    for (auto& t : P.timed) {
      HIP_CHECK(hipEventSynchronize(t.b));
      float e = 0.f;
      HIP_CHECK(hipEventElapsedTime(&e, t.a, t.b));
      ms[t.kind] += e;
      launches[t.kind] += 1;
      hipEventDestroy(t.a);
      hipEventDestroy(t.b);
    }
    P.timed.clear();
  });
}


int born_batch_create(const born_plan_t* plans, int32_t n_plans, born_batch_t* batch) {
  return guarded([&] {
    if (!plans || !batch || n_plans < 1) throw Error(BORN_ERR_INVALID_ARGUMENT, "plans and output are required");
    if (n_plans > MAX_BATCH) throw Error(BORN_ERR_UNSUPPORTED, "too many plans in one batch");
    *batch = nullptr;
    std::unique_ptr<Batch> B(new Batch());
    for (int i = 0; i < n_plans; ++i) {
      if (!plans[i]) throw Error(BORN_ERR_NO_OPERATOR_BORN, "plan is NULL");
      B->plans.push_back(reinterpret_cast<Plan*>(plans[i]));
    }
    Plan& P0 = *B->plans[0];
    B->device = P0.device;
    B->n_cu = P0.n_cu;
    B->precision = P0.precision;
    HIP_CHECK(hipSetDevice(B->device));
    for (Plan* P : B->plans) {
      if (P->device != B->device) throw Error(BORN_ERR_UNSUPPORTED, "plans of a batch must share the device");
      if (!use_structured(*P)) throw Error(BORN_ERR_UNSUPPORTED, "a batch needs plans on the structured path");
      if (P->Ne_pad != P0.Ne_pad || P->Np != P0.Np)
        throw Error(BORN_ERR_UNSUPPORTED, "plans of a batch must share element padding and points per pitch");
      if (P->precision != B->precision) throw Error(BORN_ERR_UNSUPPORTED, "plans of a batch must share the precision");
      check_precision(*P);
      ensure_workspace(*P, 1);
    }
    HIP_CHECK(hipStreamCreateWithFlags(&B->own, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
    build_batch_launch(*B);
    *batch = reinterpret_cast<born_batch_t>(B.release());
  });
}

int born_batch_destroy(born_batch_t batch) {
  return guarded([&] { delete reinterpret_cast<Batch*>(batch); });
}

int born_batch_forward_dev(born_batch_t batch, const void* const* gamma_c64, void* const* u_c64, void* stream) {
  return guarded([&] {
    if (!batch || !gamma_c64 || !u_c64) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch and buffers are required");
    Batch& B = *reinterpret_cast<Batch*>(batch);
    for (size_t i = 0; i < B.plans.size(); ++i)
      if (!gamma_c64[i] || !u_c64[i]) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "invalid device buffers");
    HIP_CHECK(hipSetDevice(B.device));
    batch_apply(B, 0, gamma_c64, u_c64, (hipStream_t)stream);
  });
}

int born_batch_adjoint_dev(born_batch_t batch, const void* const* u_c64, void* const* gamma_hat_c64, void* stream) {
  return guarded([&] {
    if (!batch || !u_c64 || !gamma_hat_c64) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch and buffers are required");
    Batch& B = *reinterpret_cast<Batch*>(batch);
    for (size_t i = 0; i < B.plans.size(); ++i)
      if (!u_c64[i] || !gamma_hat_c64[i]) throw Error(BORN_ERR_NO_VALID_DOUBLE_MATRIX, "invalid device buffers");
    HIP_CHECK(hipSetDevice(B.device));
    batch_apply(B, 1, u_c64, gamma_hat_c64, (hipStream_t)stream);
  });
}

int born_batch_set_graph(born_batch_t batch, int32_t on) {
  return guarded([&] {
    if (!batch) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch is NULL");
    Batch& B = *reinterpret_cast<Batch*>(batch);
    B.graph = on != 0;
    if (!B.graph) {
      HIP_CHECK(hipSetDevice(B.device));
      for (auto& c : B.cap) drop_graph(c);
    }
  });
}

int born_batch_autotune(born_batch_t batch, void* stream, double* fwd_ms, double* adj_ms) {
  return guarded([&] {
    if (!batch) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch is NULL");
    Batch& B = *reinterpret_cast<Batch*>(batch);
    HIP_CHECK(hipSetDevice(B.device));
    hipStream_t s = (hipStream_t)stream;
    batch_refresh(B);
    HIP_CHECK(hipDeviceSynchronize());
    // each MFMA launch set over its own plans: work items and grid (and the
    // adjoint width where the set does not fix it at 128 points)
    TuneResult r{0, 0, 0, 0.f, 0.f};
    for (int k = 0; k < Batch::N_SETS; ++k) {
      if (!Batch::mfma_set(k)) continue;
      std::vector<Plan*> pk;
      for (size_t i = 0; i < B.plans.size(); ++i)
        if (B.set_of[i] == k) pk.push_back(B.plans[i]);
      if (pk.empty()) continue;
      const TuneResult rk = autotune(pk, B.set(k), B.precision, B.n_cu, [&] { build_batch_launch(B); }, s);
      r.fwd_ms += rk.fwd_ms;
      r.adj_ms += rk.adj_ms;
    }
    if (fwd_ms) *fwd_ms = r.fwd_ms;
    if (adj_ms) *adj_ms = r.adj_ms;
  });
}

int born_plan_autotune(born_plan_t plan, void* stream, double* fwd_ms, double* adj_ms) {
  return guarded([&] {
    if (!plan) throw Error(BORN_ERR_INVALID_ARGUMENT, "plan is NULL");
    Plan& P = *reinterpret_cast<Plan*>(plan);
    if (!use_structured(P)) throw Error(BORN_ERR_UNSUPPORTED, "autotuning covers the structured path");
    check_precision(P);
    HIP_CHECK(hipSetDevice(P.device));
    ensure_workspace(P, 1);
    hipStream_t s = (hipStream_t)stream;
    HIP_CHECK(hipDeviceSynchronize());
    std::vector<Plan*> one{&P};
    const TuneResult r = autotune(one, P.ls, P.precision, P.n_cu, [&] {
      P.ls.inst.assign(1, make_inst(P));
      finalize_launchset(P.ls);
    }, s);
    if (fwd_ms) *fwd_ms = r.fwd_ms;
    if (adj_ms) *adj_ms = r.adj_ms;
  });
}

int born_batch_set_timing(born_batch_t batch, int32_t on) {
  return guarded([&] {
    if (!batch) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch is NULL");
    reinterpret_cast<Batch*>(batch)->timing = on != 0;
  });
}

int born_batch_get_timing(born_batch_t batch, double* ms, int64_t* launches) {
  return guarded([&] {
    if (!batch || !ms || !launches) throw Error(BORN_ERR_INVALID_ARGUMENT, "batch and outputs are required");
    Batch& B = *reinterpret_cast<Batch*>(batch);
    HIP_CHECK(hipSetDevice(B.device));
    ms[0] = ms[1] = 0.0;
    launches[0] = launches[1] = 0;
    for (auto& t : B.timed) {
      HIP_CHECK(hipEventSynchronize(t.b));
      float e = 0.f;
      HIP_CHECK(hipEventElapsedTime(&e, t.a, t.b));
      ms[t.kind] += e;
      launches[t.kind] += 1;
      hipEventDestroy(t.a);
      hipEventDestroy(t.b);
    }
    B.timed.clear();
  });
}

const char* born_last_error(void) { return g_last_error.c_str(); }

int born_debug_fail_allocation(int64_t n) {
  g_fail_alloc.store(n < 0 ? -1 : n);
  return BORN_OK;
}

int64_t born_debug_live_allocations(void) { return g_live_alloc.load(); }

const char* born_status_id(int status) {
  switch (status) {
    case BORN_OK: return "";
    case BORN_ERR_NUMBER_ARGUMENTS: return "combined_quick_gpu:NumberArguments";
    case BORN_ERR_NO_SYMMETRIC_GRIDS_SETUP: return "combined_quick_gpu:NoSymmetricGridsSetup";
    case BORN_ERR_NO_GPU_ACTIVE: return "combined_quick_gpu:NoScatteringOptionsGPUActive";
    case BORN_ERR_NO_COMPLEX_DOUBLES: return "combined_quick_gpu:NoComplexDoubles";
    case BORN_ERR_NO_OPERATOR_BORN: return "combined_quick_gpu:NoOperatorBorn";
    case BORN_ERR_NO_NUMERIC_SCALAR_MODE: return "combined_quick_gpu:NoNumericScalarMode";
    case BORN_ERR_NO_VALID_DOUBLE_MATRIX: return "combined_quick_gpu:NoValidDoubleMatrix";
    case BORN_ERR_CONVERSION_FAILED: return "combined_quick_gpu:ConversionFailed";
    case BORN_ERR_DEVICE: return "adjoint_quick_gpu:ErrorCUDA";
    case BORN_ERR_INVALID_ARGUMENT: return "born:InvalidArgument";
    case BORN_ERR_UNSUPPORTED: return "born:Unsupported";
    case BORN_ERR_OUT_OF_MEMORY: return "born:OutOfMemory";
    default: return "born:Unknown";
  }
}

}  // extern "C"
