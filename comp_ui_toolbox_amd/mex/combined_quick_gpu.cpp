// combined_quick_gpu.cpp -- MEX shim over libborn (MI355X).
//
// Drop-in for the CUDA MEX +scattering/combined_quick_gpu.cu: same call
// syntax (combined_quick_gpu.m:11-23), same property paths (cu:201-391,
// :817-823) and error identifiers.  The plan with its device tables persists
// between calls, keyed by a fingerprint of every input it depends on, and is
// freed by mexAtExit.  An optional fourth argument selects the numeric form
// of the contraction: 'fp32' (default), 'split_bf16' or 'split_fp16' (or the
// codes 0, 1, 2).  Build (MATLAB R2018a or newer, interleaved complex API):
//   mex -R2018a -I<repo>/include combined_quick_gpu.cpp -L<repo>/comp_ui_toolbox_amd -lborn
#include "mex.h"
#include "born.h"
#include <cstdint>
#include <cstring>
#include <vector>

static born_plan_t g_plan = nullptr;
static uint64_t g_key = 0;

static void release() { if (g_plan) { born_plan_destroy(g_plan); g_plan = nullptr; } }

[[noreturn]] static void fail(int status) {
  mexErrMsgIdAndTxt(born_status_id(status), "%s", born_last_error());
}
[[noreturn]] static void fail(const char* id, const char* msg) { mexErrMsgIdAndTxt(id, "%s", msg); }

static const mxArray* prop(const mxArray* obj, mwIndex i, const char* name) {
  const mxArray* p = mxGetProperty(obj, i, name);
  if (!p) fail("combined_quick_gpu:NoOperatorBorn", "missing property");
  return p;
}
static const mxArray* cell(const mxArray* c, mwIndex i) {
  const mxArray* p = mxGetCell(c, i);
  if (!p) fail("combined_quick_gpu:NoOperatorBorn", "missing cell entry");
  return p;
}
static const mxArray* values(const mxArray* signal_matrix) {   // .samples.values
  return prop(prop(signal_matrix, 0, "samples"), 0, "values");
}
static const double* complex_doubles(const mxArray* a) {
  if (!mxIsDouble(a) || !mxIsComplex(a))
    fail("combined_quick_gpu:NoComplexDoubles", "tables must be complex doubles!");
  return reinterpret_cast<const double*>(mxGetComplexDoubles(a));
}

// Plan fingerprint: FNV-1a over 64-bit words.  Small inputs (sizes,
// selection, frequency maps, active sets, prefactors) enter in full; the big
// tables (h_ref, shift map, p_incident) by their size and 65536 strided
// entries each, so that a call costs no full pass over gigabytes of tables.
struct Fingerprint {
  uint64_t h = 1469598103934665603ull;
  void word(uint64_t w) {
    for (int b = 0; b < 8; ++b) h = (h ^ ((w >> (8 * b)) & 0xffu)) * 1099511628211ull;
  }
  void full(const double* p, size_t n) {
    word(n);
    for (size_t i = 0; i < n; ++i) { uint64_t w; std::memcpy(&w, p + i, 8); word(w); }
  }
  void sampled(const double* p, size_t n) {
    word(n);
    const size_t step = n > 65536 ? n / 65536 : 1;
    for (size_t i = 0; i < n; i += step) { uint64_t w; std::memcpy(&w, p + i, 8); word(w); }
    if (n) { uint64_t w; std::memcpy(&w, p + n - 1, 8); word(w); }
  }
};

static int32_t precision_arg(const mxArray* a) {
  if (mxIsChar(a)) {
    char buf[32] = {0};
    if (mxGetString(a, buf, sizeof buf) == 0) {
      if (!std::strcmp(buf, "fp32")) return BORN_PRECISION_FP32;
      if (!std::strcmp(buf, "split_bf16")) return BORN_PRECISION_SPLIT_BF16;
      if (!std::strcmp(buf, "split_fp16")) return BORN_PRECISION_SPLIT_FP16;
    }
  } else if (mxIsNumeric(a) && !mxIsComplex(a) && mxGetNumberOfElements(a) == 1) {
    const int c = (int)mxGetScalar(a);
    if (c >= BORN_PRECISION_FP32 && c <= BORN_PRECISION_SPLIT_FP16) return c;
  }
  fail("combined_quick_gpu:NumberArguments",
       "the optional fourth input must be 'fp32', 'split_bf16' or 'split_fp16'!");
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  mexAtExit(release);
  if ((nrhs != 3 && nrhs != 4) || nlhs != 1)                                  // cu:201
    fail("combined_quick_gpu:NumberArguments", "Three inputs and one output are required!");
  const mxArray* op = prhs[0];
  if (!mxIsClass(op, "scattering.operator_born") || mxGetNumberOfElements(op) != 1)
    fail("combined_quick_gpu:NoOperatorBorn", "The first input must be a single scattering.operator_born!");
  const mxArray* seq = prop(op, 0, "sequence");
  const mxArray* setup = prop(seq, 0, "setup");
  if (!mxIsClass(setup, "scattering.sequences.setups.setup_grid_symmetric"))   // cu:228
    fail("combined_quick_gpu:NoSymmetricGridsSetup", "The setup must be setup_grid_symmetric!");
  const mxArray* gpu = prop(prop(prop(op, 0, "options"), 0, "momentary"), 0, "gpu");
  if (!mxIsClass(gpu, "scattering.options.gpu_active"))                        // cu:372
    fail("combined_quick_gpu:NoScatteringOptionsGPUActive", "GPU options must be gpu_active!");
  const mxArray* mode_a = prhs[1];
  if (!mxIsNumeric(mode_a) || mxIsComplex(mode_a) || mxGetNumberOfElements(mode_a) != 1)
    fail("combined_quick_gpu:NoNumericScalarMode", "mode must be a real numeric scalar!");
  const int mode = (int)mxGetScalar(mode_a) == 1 ? 1 : 2;                      // cu:404-409
  const int32_t precision = nrhs == 4 ? precision_arg(prhs[3]) : BORN_PRECISION_FP32;

  // tables, read by the property paths of cu:211-391 and :817-823
  const mxArray* shift = prop(setup, 0, "indices_grid_FOV_shift");
  const mxArray* h_ref = values(prop(seq, 0, "h_ref"));
  const mxArray* sel = prop(op, 0, "indices_measurement_sel");
  const mxArray* fov_grid = prop(prop(prop(setup, 0, "FOV"), 0, "shape"), 0, "grid");
  const mxArray* n_axis = prop(fov_grid, 0, "N_points_axis");
  const mxArray* xdc = prop(setup, 0, "xdc_array");
  const mxArray* ne_axis = prop(xdc, 0, "N_elements_axis");

  born_tables_desc d{};
  d.N_dimensions = (int32_t)mxGetNumberOfElements(n_axis);
  if (d.N_dimensions < 2 || d.N_dimensions > 3)
    fail("born:InvalidArgument", "the FOV grid must have 2 or 3 axes");
  for (int i = 0; i < d.N_dimensions; ++i) d.N_points_axis[i] = (int32_t)mxGetPr(n_axis)[i];
  d.N_elements_axis[0] = (int32_t)mxGetPr(ne_axis)[0];
  d.N_elements_axis[1] = mxGetNumberOfElements(ne_axis) > 1 ? (int32_t)mxGetPr(ne_axis)[1] : 1;
  d.N_points = (int32_t)mxGetN(h_ref);
  d.N_elements = (int32_t)mxGetScalar(prop(xdc, 0, "N_elements"));             // cu:234
  d.N_f_unique = (int32_t)mxGetM(h_ref);
  d.h_ref = complex_doubles(h_ref);
  d.indices_grid_FOV_shift = mxGetPr(shift);
  d.device = (int32_t)mxGetScalar(prop(gpu, 0, "index"));                      // cu:378
  d.precision = precision;

  const mxArray* f_to_unique = prop(seq, 0, "indices_f_to_unique");
  const mxArray* prefactors = prop(seq, 0, "prefactors");
  const mxArray* settings = prop(seq, 0, "settings");
  const mxArray* waves = prop(op, 0, "incident_waves");
  d.N_measurements = (int32_t)mxGetNumberOfElements(sel);
  std::vector<born_measurement_desc> meas(d.N_measurements);
  std::vector<std::vector<born_mix_desc>> mixes(d.N_measurements);

  Fingerprint fp;
  const int32_t sizes[] = {d.N_dimensions, d.N_points_axis[0], d.N_points_axis[1], d.N_points_axis[2],
                           d.N_elements_axis[0], d.N_elements_axis[1], d.N_points, d.N_elements,
                           d.N_f_unique, d.device, d.precision, d.N_measurements};
  for (int32_t v : sizes) fp.word((uint64_t)(uint32_t)v);
  fp.sampled(d.h_ref, 2 * mxGetNumberOfElements(h_ref));
  fp.sampled(d.indices_grid_FOV_shift, mxGetNumberOfElements(shift));
  fp.full(mxGetPr(sel), d.N_measurements);
  const size_t n_seq = mxGetNumberOfElements(waves);
  for (int i = 0; i < d.N_measurements; ++i) {
    const double s = mxGetPr(sel)[i];                                          // cu:238-241
    if (!(s >= 1.0) || s > (double)n_seq || s != (double)(mwIndex)s)
      fail("born:InvalidArgument", "indices_measurement_sel must index the sequence's measurements");
    const mwIndex m = (mwIndex)s - 1;
    const mxArray* fu = cell(f_to_unique, m);                                  // cu:278-281
    const mxArray* p_in = values(prop(waves, m, "p_incident"));                // cu:807-823: waves(index_measurement)
    const mxArray* rx = prop(settings, m, "rx");
    const mxArray* fk = prop(settings, m, "indices_f_to_unique");
    const mxArray* pref = cell(prefactors, m);
    born_measurement_desc& md = meas[i];
    md.N_f_unique = (int32_t)mxGetNumberOfElements(fu);
    md.indices_f_to_unique = mxGetPr(fu);
    md.p_incident = complex_doubles(p_in);
    md.N_mix = (int32_t)mxGetNumberOfElements(rx);
    fp.full(md.indices_f_to_unique, md.N_f_unique);
    fp.sampled(md.p_incident, 2 * mxGetNumberOfElements(p_in));
    fp.word((uint64_t)md.N_mix);
    mixes[i].resize(md.N_mix);
    for (int k = 0; k < md.N_mix; ++k) {                                       // cu:298-352
      const mxArray* act = prop(rx, k, "indices_active");
      const mxArray* fkk = cell(fk, k);
      const mxArray* pv = prop(prop(pref, k, "samples"), 0, "values");
      born_mix_desc& x = mixes[i][k];
      x.N_f = (int32_t)mxGetNumberOfElements(fkk);
      x.indices_f_to_unique = mxGetPr(fkk);
      x.N_active = (int32_t)mxGetNumberOfElements(act);
      x.indices_active = mxGetPr(act);
      x.prefactors = complex_doubles(pv);
      fp.full(x.indices_f_to_unique, x.N_f);
      fp.full(x.indices_active, x.N_active);
      fp.full(x.prefactors, 2 * mxGetNumberOfElements(pv));
    }
    md.mixes = mixes[i].data();
  }
  d.measurements = meas.data();

  if (!g_plan || fp.h != g_key) {
    release();
    int st = born_plan_create_tables(&d, &g_plan);
    if (st != BORN_OK) { g_plan = nullptr; fail(st); }
    g_key = fp.h;
  }

  int64_t N_obs = 0, N_pts = 0;
  born_plan_size(g_plan, &N_obs, &N_pts);
  const mxArray* in = prhs[2];
  const int64_t rows = mode == 1 ? N_pts : N_obs;
  if (!mxIsDouble(in) || mxGetNumberOfDimensions(in) != 2 || (int64_t)mxGetM(in) != rows)
    fail("combined_quick_gpu:NoValidDoubleMatrix", "input must be a double matrix of matching rows!");
  const mwSize N_rhs = mxGetN(in);
  // real input is made complex, as at cu:421-437
  std::vector<double> tmp;
  const double* x;
  if (mxIsComplex(in)) {
    x = reinterpret_cast<const double*>(mxGetComplexDoubles(in));
  } else {
    tmp.assign(2 * rows * N_rhs, 0.0);
    const double* r = mxGetPr(in);
    for (int64_t i = 0; i < rows * (int64_t)N_rhs; ++i) tmp[2 * i] = r[i];
    x = tmp.data();
  }
  plhs[0] = mxCreateDoubleMatrix(mode == 1 ? N_obs : N_pts, N_rhs, mxCOMPLEX);  // cu:447-451
  double* y = reinterpret_cast<double*>(mxGetComplexDoubles(plhs[0]));
  const int st = mode == 1 ? born_forward(g_plan, x, (int32_t)N_rhs, y)
                           : born_adjoint(g_plan, x, (int32_t)N_rhs, y);
  if (st != BORN_OK) fail(st);
}
