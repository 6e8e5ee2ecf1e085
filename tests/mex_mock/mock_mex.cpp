// In-memory MATLAB object model behind tests/mex_mock/mex.h -- TEST
// INFRASTRUCTURE ONLY.  Arrays are double matrices (real or interleaved
// complex), cell arrays, or object arrays of a named class whose elements
// hold named properties.  A C API lets Python build operator objects and
// call the shim's mexFunction; mexErrMsgIdAndTxt becomes a C++ exception
// that mock_call turns into the error identifier.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct mxArray_tag {
  enum Kind { DOUBLE, CELL, OBJECT, CHAR } kind = DOUBLE;
  std::string text;                                   // CHAR arrays
  std::string cls = "double";
  size_t M = 0, N = 0;
  bool cplx = false;
  std::vector<double> data;                           // real, or (re, im) interleaved
  std::vector<mxArray*> cells;
  std::vector<std::map<std::string, mxArray*>> props;  // one map per object element
};

namespace {
std::vector<std::unique_ptr<mxArray>> g_arena;
std::vector<void (*)(void)> g_atexit;
struct MexError {
  std::string id, msg;
};
mxArray* make() {
  g_arena.emplace_back(new mxArray());
  return g_arena.back().get();
}
}  // namespace

extern "C" {
int mexAtExit(void (*fn)(void)) {
  for (auto f : g_atexit)
    if (f == fn) return 0;
  g_atexit.push_back(fn);
  return 0;
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw MexError{id ? id : "", buf};
}

mxArray* mxGetProperty(const mxArray* obj, mwIndex index, const char* name) {
  if (!obj || obj->kind != mxArray::OBJECT || index >= obj->props.size()) return nullptr;
  auto it = obj->props[index].find(name);
  return it == obj->props[index].end() ? nullptr : it->second;
}
mxArray* mxGetCell(const mxArray* cell, mwIndex index) {
  if (!cell || cell->kind != mxArray::CELL || index >= cell->cells.size()) return nullptr;
  return cell->cells[index];
}
bool mxIsDouble(const mxArray* a) { return a && a->kind == mxArray::DOUBLE; }
bool mxIsComplex(const mxArray* a) { return a && a->kind == mxArray::DOUBLE && a->cplx; }
bool mxIsNumeric(const mxArray* a) { return a && a->kind == mxArray::DOUBLE; }
bool mxIsClass(const mxArray* a, const char* name) { return a && a->cls == name; }
bool mxIsChar(const mxArray* a) { return a && a->kind == mxArray::CHAR; }
int mxGetString(const mxArray* a, char* buf, mwSize buflen) {
  if (!a || a->kind != mxArray::CHAR || buflen == 0) return 1;
  std::snprintf(buf, buflen, "%s", a->text.c_str());
  return a->text.size() + 1 > buflen ? 1 : 0;
}
size_t mxGetNumberOfElements(const mxArray* a) {
  if (!a) return 0;
  if (a->kind == mxArray::OBJECT) return a->props.size();
  if (a->kind == mxArray::CELL) return a->cells.size();
  return a->M * a->N;
}
mwSize mxGetNumberOfDimensions(const mxArray*) { return 2; }
size_t mxGetM(const mxArray* a) { return a->M; }
size_t mxGetN(const mxArray* a) { return a->N; }
double mxGetScalar(const mxArray* a) { return a && !a->data.empty() ? a->data[0] : 0.0; }
double* mxGetPr(const mxArray* a) {
  if (!a || a->kind != mxArray::DOUBLE || a->cplx) throw MexError{"mock:mxGetPr", "not a real double array"};
  return const_cast<double*>(a->data.data());
}
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) {
  if (!a || a->kind != mxArray::DOUBLE || !a->cplx)
    throw MexError{"mock:mxGetComplexDoubles", "not a complex double array"};
  return reinterpret_cast<mxComplexDouble*>(const_cast<double*>(a->data.data()));
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  mxArray* a = make();
  a->M = m;
  a->N = n;
  a->cplx = c == mxCOMPLEX;
  a->data.assign(m * n * (a->cplx ? 2 : 1), 0.0);
  return a;
}

// ---------------------------------------------------------------- builders
mxArray* mock_double(size_t m, size_t n, int cplx, const double* data) {
  mxArray* a = mxCreateDoubleMatrix(m, n, cplx ? mxCOMPLEX : mxREAL);
  if (data) std::memcpy(a->data.data(), data, a->data.size() * sizeof(double));
  return a;
}
mxArray* mock_char(const char* text) {
  mxArray* a = make();
  a->kind = mxArray::CHAR;
  a->cls = "char";
  a->text = text;
  a->M = 1;
  a->N = a->text.size();
  return a;
}
mxArray* mock_cell(size_t n) {
  mxArray* a = make();
  a->kind = mxArray::CELL;
  a->cls = "cell";
  a->M = 1;
  a->N = n;
  a->cells.assign(n, nullptr);
  return a;
}
void mock_cell_set(mxArray* cell, size_t i, mxArray* v) { cell->cells.at(i) = v; }
mxArray* mock_object(const char* cls, size_t n) {
  mxArray* a = make();
  a->kind = mxArray::OBJECT;
  a->cls = cls;
  a->M = 1;
  a->N = n;
  a->props.resize(n);
  return a;
}
void mock_prop_set(mxArray* obj, size_t i, const char* name, mxArray* v) { obj->props.at(i)[name] = v; }
void mock_shape(const mxArray* a, size_t* m, size_t* n, int* cplx) {
  *m = a->M;
  *n = a->N;
  *cplx = a->cplx ? 1 : 0;
}
const double* mock_data(const mxArray* a) { return a->data.data(); }

// Calls mexFunction; returns 0, or 1 with the error identifier and message.
int mock_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs, char* err_id, size_t id_cap, char* err_msg,
              size_t msg_cap) {
  try {
    mexFunction(nlhs, plhs, nrhs, prhs);
    return 0;
  } catch (const MexError& e) {
    std::snprintf(err_id, id_cap, "%s", e.id.c_str());
    std::snprintf(err_msg, msg_cap, "%s", e.msg.c_str());
    return 1;
  }
}

// `clear mex`: run the registered exit handlers, then free every array.
void mock_clear() {
  for (auto f : g_atexit) f();
  g_atexit.clear();
  g_arena.clear();
}
}
