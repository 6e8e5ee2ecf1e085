/*
 * born_ref_loop.c -- TEST INFRASTRUCTURE ONLY (oracle / CPU baseline).
 *
 * C restatement, in FP64, of the reference CPU operator pair for one
 * measurement: forward_quick_cpu (operator_born.m:353-459) and
 * adjoint_quick_cpu (operator_born.m:511-593).  It keeps the reference's
 * loop nest -- mix -> active element -- and, with `fused` = 0, its cost
 * structure: per active element the matrix Phi = h(fmap, shift) .* p(fmap, :)
 * .* pref(:, a) is materialized (N_f x N_occ, f fastest) and then multiplied
 * (u += Phi * gamma, or gamma_hat += Phi' * u).  `fused` = 1 forms each Phi
 * entry on the fly instead.  Occupancy follows the MEX rule
 * (combined_quick_gpu.cu:466-488).  OpenMP parallelizes over grid points.
 *
 * Complex numbers are interleaved (re, im) doubles; matrices column-major.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define DBL_EPS 2.220446049250313080847e-16

/* mode 1: out[obs] = forward of this measurement, obs ordered mix-major then f.
 * mode 2: out[p] += adjoint contribution of this measurement's u (in). */
int ref_apply(int mode, int N_pts, int N_fu, const double* h_, const int* shift, int N_e,
              int N_fm, const int* fg_of_fm, const double* p_,
              int N_mix, const int* mix_Nf, const int* mix_fmeas, const int* mix_Nact, const int* mix_act,
              const double* mix_pref_, const double* in_, double* out_, int fused, int nthreads) {
  const cplx* h = (const cplx*)h_;
  const cplx* p = (const cplx*)p_;
  const cplx* pref_all = (const cplx*)mix_pref_;
  const cplx* in = (const cplx*)in_;
  cplx* out = (cplx*)out_;
  (void)N_fu;
  (void)N_e;
  (void)N_fm;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  /* occupied points (forward) or all points (adjoint) */
  int* occ = (int*)malloc(sizeof(int) * (size_t)N_pts);
  if (!occ) return 1;
  int N_occ = 0;
  for (int q = 0; q < N_pts; ++q) {
    if (mode != 1) {
      occ[N_occ++] = q;
      continue;
    }
    const double re = creal(in[q]), im = cimag(in[q]);
    if (re > DBL_EPS || re < -DBL_EPS || im > DBL_EPS || im < -DBL_EPS) occ[N_occ++] = q;
  }
  int Nf_max = 0;
  for (int k = 0; k < N_mix; ++k) Nf_max = mix_Nf[k] > Nf_max ? mix_Nf[k] : Nf_max;
  cplx* Phi = NULL;
  if (!fused && N_occ > 0) {
    Phi = (cplx*)malloc(sizeof(cplx) * (size_t)Nf_max * N_occ);
    if (!Phi) {
      free(occ);
      return 1;
    }
  }
  size_t f_off = 0, a_off = 0, pref_off = 0, obs_off = 0;
  for (int k = 0; k < N_mix; ++k) {
    const int Nf = mix_Nf[k], Na = mix_Nact[k];
    const int* fm = mix_fmeas + f_off;
    const int* act = mix_act + a_off;
    const cplx* pref = pref_all + pref_off;
    if (mode == 1)
      for (int f = 0; f < Nf; ++f) out[obs_off + f] = 0;
    for (int a = 0; a < Na; ++a) {
      const int e = act[a];
      const int* sh = shift + (size_t)e * N_pts;
      if (!fused) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N_occ; ++j) {
          const int q = occ[j];
          const cplx* hcol = h + (size_t)sh[q] * N_fu;
          const cplx* pcol = p + (size_t)q * N_fm;
          cplx* phi = Phi + (size_t)j * Nf;
          for (int f = 0; f < Nf; ++f) phi[f] = hcol[fg_of_fm[fm[f]]] * pcol[fm[f]] * pref[(size_t)a * Nf + f];
        }
        if (mode == 1) {
          /* u(:) += Phi * gamma_occ : reduce over points per thread */
#pragma omp parallel
          {
            cplx* acc = (cplx*)calloc((size_t)Nf, sizeof(cplx));
#pragma omp for schedule(static)
            for (int j = 0; j < N_occ; ++j) {
              const cplx g = in[occ[j]];
              const cplx* phi = Phi + (size_t)j * Nf;
              for (int f = 0; f < Nf; ++f) acc[f] += phi[f] * g;
            }
#pragma omp critical
            for (int f = 0; f < Nf; ++f) out[obs_off + f] += acc[f];
            free(acc);
          }
        } else {
          const cplx* uu = in + obs_off;
#pragma omp parallel for schedule(static)
          for (int j = 0; j < N_occ; ++j) {
            const cplx* phi = Phi + (size_t)j * Nf;
            cplx s = 0;
            for (int f = 0; f < Nf; ++f) s += conj(phi[f]) * uu[f];
            out[occ[j]] += s;
          }
        }
      } else {
        if (mode == 1) {
#pragma omp parallel
          {
            cplx* acc = (cplx*)calloc((size_t)Nf, sizeof(cplx));
#pragma omp for schedule(static)
            for (int j = 0; j < N_occ; ++j) {
              const int q = occ[j];
              const cplx* hcol = h + (size_t)sh[q] * N_fu;
              const cplx* pcol = p + (size_t)q * N_fm;
              const cplx g = in[q];
              for (int f = 0; f < Nf; ++f)
                acc[f] += hcol[fg_of_fm[fm[f]]] * pcol[fm[f]] * pref[(size_t)a * Nf + f] * g;
            }
#pragma omp critical
            for (int f = 0; f < Nf; ++f) out[obs_off + f] += acc[f];
            free(acc);
          }
        } else {
          const cplx* uu = in + obs_off;
#pragma omp parallel for schedule(static)
          for (int j = 0; j < N_occ; ++j) {
            const int q = occ[j];
            const cplx* hcol = h + (size_t)sh[q] * N_fu;
            const cplx* pcol = p + (size_t)q * N_fm;
            cplx s = 0;
            for (int f = 0; f < Nf; ++f)
              s += conj(hcol[fg_of_fm[fm[f]]] * pcol[fm[f]] * pref[(size_t)a * Nf + f]) * uu[f];
            out[q] += s;
          }
        }
      }
    }
    f_off += Nf;
    a_off += Na;
    pref_off += (size_t)Nf * Na;
    obs_off += Nf;
  }
  free(Phi);
  free(occ);
  return 0;
}

/* One frequency of the 3-D spatial transfer function of every element at
 * arbitrary points, before the anti-aliasing taper (setup.m:1216-1281):
 *     out[p][e] = -2 dA sum_q G(k, |r_p - f_eq|),  G(k, D) = -exp(-j k D) / (4 pi D)
 * (greens_function.m:103), face points f_eq = (ctr_e + off_q, 0).
 * pos: n_pts x 3, ctr: n_e x 2, off: n_q x 2 (row-major); out interleaved
 * complex, n_pts x n_e.  k = k_re + j k_im (k_im <= 0: attenuation). */
void ref_face_sum3(long n_pts, const double* pos, int n_e, const double* ctr, int n_q, const double* off,
                   double dA, double k_re, double k_im, double* out) {
  const double c0 = 2.0 * dA / (4.0 * 3.14159265358979323846);   /* -2 dA * (-1 / (4 pi)) */
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n_pts; ++p) {
    const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
    for (int e = 0; e < n_e; ++e) {
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < n_q; ++q) {
        const double dx = x - (ctr[2 * e] + off[2 * q]), dy = y - (ctr[2 * e + 1] + off[2 * q + 1]);
        const double D = sqrt(dx * dx + dy * dy + z * z);
        /* exp(-j k D) = exp(k_im D) (cos(k_re D) - j sin(k_re D)) */
        const double a = exp(k_im * D) / D;
        sr += a * cos(k_re * D);
        si -= a * sin(k_re * D);
      }
      out[2 * (p * (long)n_e + e)] = c0 * sr;
      out[2 * (p * (long)n_e + e) + 1] = c0 * si;
    }
  }
}
