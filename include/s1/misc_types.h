/*
 * misc_types.h -- value types of the drop-in C++ surface.
 *
 * The same aliases as the reference's misc_types.h (misc_types.h:25-50), so
 * code written against the reference's signatures compiles unchanged: rows
 * of std::complex<double> per radar line, rows of lines per burst.
 */
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../s1gpu.h"

template <typename T>
using VEC_1D = std::vector<T>;
template <typename T>
using VEC_2D = std::vector<std::vector<T>>;

using F_VEC_1D = VEC_1D<float>;
using F_VEC_2D = VEC_2D<float>;
using D_VEC_1D = VEC_1D<double>;
using D_VEC_2D = VEC_2D<double>;
using INT_VEC_1D = VEC_1D<int>;
using UINT8_VEC_1D = VEC_1D<std::uint8_t>;
using STRING_VEC_1D = VEC_1D<std::string>;

using cdouble = std::complex<double>;
using CF_VEC_1D = VEC_1D<cdouble>;
using CF_VEC_2D = VEC_2D<cdouble>;

namespace s1 {

/* The GPU the drop-in functions run on, per host thread (default 0).  Each
 * host thread gets its own s1g context (stream, plans, caches) on first
 * use; every call returns after its results are back on the host. */
void set_device(int device);

/* Precision of the burst-level functions (s1_decoder.h): S1G_F64 (the
 * reference's arithmetic, default) or S1G_F32 (the fast chain).  The
 * building-block functions always compute in FP64, like the reference. */
void set_burst_precision(int prec);

/* The calling thread's context, for mixing with the C ABI (include/s1gpu.h). */
s1g_ctx* context();

}  // namespace s1
