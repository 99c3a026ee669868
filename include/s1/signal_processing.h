/*
 * signal_processing.h -- drop-in building blocks of the reference's
 * signal_processing.h used on the L0 -> SLC path, computed on the GPU
 * through the C ABI (FP64).  Signatures follow signal_processing.h:15-180.
 * get_fftw_plans / destroy_fftw_plans return opaque plan handles
 * (s1g_plan) in place of FFTW's; code that includes these headers must not
 * also include fftw3.h.
 */
#pragma once

#include "misc_types.h"

/* linspace, endpoint included (signal_processing.cpp:18-29). */
D_VEC_1D linspace(const double& start, const double& end, const int& size);

/* Hann-windowed sinc interpolation at fractional index t over 2 L + 1
 * samples (signal_processing.cpp:238-257). */
std::complex<double> sinc_interpolate(const CF_VEC_1D& signal, double t, int L = 8);

CF_VEC_2D transpose(const CF_VEC_2D& arr);

/* DFTs with FFTW's conventions: forward unnormalised, inverse x 1/fft_size
 * (signal_processing.cpp:516-681).  fft_size = 0 (or <= 0 for the 1-D
 * forms) means the full length; a shorter fft_size transforms the leading
 * fft_size samples of each row. */
CF_VEC_1D compute_1d_dft(const CF_VEC_1D& signal, int fft_size = 0, const bool& inverse = false);
void compute_1d_dft_in_place(CF_VEC_1D& signal, int fft_size = 0, const bool& inverse = false);
void _compute_axis_dft(CF_VEC_2D& signals, int& fft_size, const bool& inverse);
CF_VEC_2D compute_axis_dft(CF_VEC_2D& signals, int fft_size = 0, const int& axis = 0, const bool& inverse = false);
void compute_axis_dft_in_place(CF_VEC_2D& signals, int fft_size = 0, const int& axis = 0,
                               const bool& inverse = false);

/* Resampling and fitting on the host, as the reference writes them
 * (signal_processing.cpp:32-233,501-513): linear_resample walks a running
 * abscissa over the input segments and ends on the last input;
 * quadratic_resample is three-point Lagrange (its first output interpolates
 * through a repeated point, 0 / 0, when the lengths differ); polyfit is the
 * degree-2 least-squares fit by normal equations, coefficients of 1, x, x^2;
 * fftfreq is numpy's. */
D_VEC_1D linear_resample(const D_VEC_1D& arr, const int& num_output_samples);
CF_VEC_1D linear_resample(const CF_VEC_1D& arr, const int& num_output_samples);
CF_VEC_1D quadratic_resample(const CF_VEC_1D& arr, const int& num_output_samples);
D_VEC_1D polyfit(const D_VEC_1D& x, const D_VEC_1D& y);
double polyval(const D_VEC_1D& p, const double& x);
D_VEC_1D fftfreq(int n, double d = 1.0);

/* std::rotate by size/2 (signal_processing.cpp:713-727). */
void fftshift_in_place(CF_VEC_1D& signal);
void fftshift_in_place(CF_VEC_2D& signals);

/* ECCM RFI detection and notch (signal_processing.cpp:913-1004); rx_pol
 * is unused, as in the reference. */
void eccm(CF_VEC_2D& signals, const int& fft_size, const int& stride, const double& detection_threshold,
          const double& mitigation_threshold, const char& rx_pol = 'V');

/* Per-row FFT plans (signal_processing.cpp:684-710): plan k binds row k of
 * `signals` in place, unnormalised in both directions, like the FFTW plans
 * it replaces; fftw_execute(plan) transforms the bound row on the GPU. */
namespace s1 {
struct fft_plan_rec;
}
typedef s1::fft_plan_rec* fftw_plan;
#ifndef FFTW_FORWARD
#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)
#endif
std::vector<fftw_plan> get_fftw_plans(CF_VEC_2D& signals, const int& direction = FFTW_FORWARD);
void destroy_fftw_plans(std::vector<fftw_plan>& plans);
void fftw_execute(const fftw_plan plan);

/* Display scaling (signal_processing.cpp:367-460,1007-1071).  The 2-D
 * forms run on the GPU (s1g_scale): norm_2d gives (|x| + 1e-12) / max|x|,
 * or 20 log10 of it, as float; magnitude_2d |x|^2 as float; scale(image,
 * mode) flattens row-major for mode "norm_log", "norm", "mag", "real" or
 * "imag" (std::invalid_argument otherwise).  The 1-D forms scale one line
 * on the host in double, as the reference's do (no 1e-12 offset). */
std::vector<std::vector<float>> norm_2d(const CF_VEC_2D& complex_values, const bool& log_scale = false);
F_VEC_2D magnitude_2d(const CF_VEC_2D& complex_values);
std::vector<float> scale(const CF_VEC_2D& signal, const std::string& scaling_mode);
D_VEC_1D norm_1d(const CF_VEC_1D& complex_values, const bool& log_scale = false);
D_VEC_1D magnitude_1d(const CF_VEC_1D& complex_values);
D_VEC_1D scale(const CF_VEC_1D& signal, const std::string& scaling_mode);
