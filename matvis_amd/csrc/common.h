// Shared definitions for the libmatvis_hip kernels (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/matvis_hip.h"

namespace mv {

// Speed of light, m/s (astropy.constants.c, the value the reference divides by).
constexpr double kC = 299792458.0;
constexpr double kPi = 3.14159265358979323846;

// Records an error message for mv_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...);

// Checks the launch that just happened on this thread.
int check_launch(const char *what);

inline hipStream_t as_stream(mv_stream_t s) {
  return reinterpret_cast<hipStream_t>(s);
}

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) {
  return (a + b - 1) / b;
}

// Vector types handed to the MFMA builtins.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename T> struct cplx {
  T re, im;
};

}  // namespace mv

#define MV_REQUIRE(cond, ...)                                                \
  do {                                                                       \
    if (!(cond)) return ::mv::set_error(MV_ERR_ARG, __VA_ARGS__);            \
  } while (0)
