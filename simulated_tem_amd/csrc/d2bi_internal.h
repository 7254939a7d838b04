// d2bi_internal.h -- declarations shared by the host setup and the HIP code.
#pragma once
#include <cstdint>
#include "../../include/d2bi_gpu.h"

namespace d2bi {

// Record an error message; returns code (so callers can `return seterr(...)`).
int seterr(int code, const char* fmt, ...);

double det3(const double* M);
void inverse3(const double* M, double* N);
void matmul3(const double* A, const double* B, double* C);
double norm2_3(const double* v);
double wavelength(double V);
double velocity(double V);
double deviation_parameter(double E_keV, const double* g);
int estimate_max_count(double nCells, int nTotal);
int plan_image(const d2bi_record* r, int64_t nAtoms, double precAngle, int nPrec, d2bi_plan* p);
int aperture_points(int n, double* xy);
int aperture_setup(const d2bi_record* r, double precAngle, int nPrec, double apertureAngle, int nApp,
                   double* gshift, double* aa);
void rotate_to_standard_frame(const double* z0, const double* A, const double* hkl, double* U);
int setup_tilt_sg(d2bi_record* r, const double* k, double sg, int mxy, int mzz, const double* d);
int setup_tilt_ur(d2bi_record* r, const double* U, const double* R, int mxy, int mzz, const double* d);
int setup_tilt_ng(d2bi_record* r, const double* k, double ng, int latt, const char* xifile, int mxy, int mzz,
                  double* d);
int synth_bcc(const d2bi_synth_params* p, double* x, int64_t cap, int64_t* nAtoms, double* box);

}  // namespace d2bi
