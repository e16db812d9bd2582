// common.h -- shared definitions of the MI355X EZmock engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>

#include "EZmock.h"

// Error codes: numbering of the reference's errmsg.h:33-88 (OpenMP build).
enum {
  EZMOCK_SUCCESS = 0, EZMOCK_ERR_MEMORY, EZMOCK_ERR_ARG_ECODE,
  EZMOCK_ERR_ARG_EZ, EZMOCK_ERR_ARG_COSMO, EZMOCK_ERR_ARG_PK,
  EZMOCK_ERR_ARG_MESH, EZMOCK_ERR_ARG_LBOX, EZMOCK_ERR_ARG_NGRID,
  EZMOCK_ERR_ARG_RNG, EZMOCK_ERR_ARG_SEED, EZMOCK_ERR_ARG_THREAD,
  EZMOCK_ERR_ARG_NULL, EZMOCK_ERR_ARG_NBIN, EZMOCK_ERR_ARG_PARAMS,
  EZMOCK_ERR_ARG_LOWEXP, EZMOCK_ERR_ARG_HIEXP, EZMOCK_ERR_ARG_NDATA,
  EZMOCK_ERR_ARG_FNAME, EZMOCK_ERR_ARG_RSD, EZMOCK_ERR_COS_GROWTH,
  EZMOCK_ERR_PK_FINITE, EZMOCK_ERR_PK_NONPOS, EZMOCK_ERR_PK_NONASC,
  EZMOCK_ERR_PK_RANGE, EZMOCK_ERR_PK_NBIN, EZMOCK_ERR_PK_MODBAO,
  EZMOCK_ERR_RNG_SET, EZMOCK_ERR_RNG_RESET, EZMOCK_ERR_RNG_JUMP,
  EZMOCK_ERR_PAR_COSMO, EZMOCK_ERR_PAR_RCUT, EZMOCK_ERR_PAR_REXP,
  EZMOCK_ERR_PAR_PBASE, EZMOCK_ERR_PAR_SIGV, EZMOCK_ERR_RHO_USED,
  EZMOCK_ERR_RHO_LESS, EZMOCK_ERR_RHO_PBASE, EZMOCK_ERR_RHO_NPART,
  EZMOCK_ERR_FFT_INIT, EZMOCK_ERR_FILE, EZMOCK_ERR_UNKNOWN
};

// Model and configuration constants (reference config.h:39-92).
#define EZ_PI 3.14159265358979323846
#define EZ_DOUBLE_TOL 1e-8
#define EZ_MAX_GRID 65536
#define EZ_MAX_PK_NBIN 1073741824
#define EZ_MIN_NUM_TRACER 128
#define EZ_MAX_NUM_TRACER_PER_CELL 1024
#define EZ_MAX_K_BAO 0.5
#define EZ_NUM_LOGB_BIN 64
#define EZ_RHO_SAT 100
#define EZ_GAUSS_WIDTH 10

#ifdef SINGLE_PREC
#define EZ_REAL_MIN 1.17549435082228750797e-38F
#define EZ_REAL_TOL 1e-5
typedef float2 cplx;
typedef float4 real4;
typedef uint32_t rkey;          // order-preserving key of a positive real
#else
#define EZ_REAL_MIN 2.2250738585072013830902e-308
#define EZ_REAL_TOL 1e-10
typedef double2 cplx;
typedef double4 real4;
typedef uint64_t rkey;
#endif

// Random-stream positions of the reference (draw units; N = Ng^3):
//   noise: cell idx >= 1 of the half-complex grid -> (idx-1)*2 (amp, phase)
//          or (idx-1) with fixed amplitudes   perturb.c:112-118,153-164
//   bias pair p                  -> 2N + 2p   pop_tracer.c:233-237,256
//   attach-to-particle draws     -> 4N        pop_tracer.c:781
//   tracer t position            -> 5N + 3t   pop_tracer.c:919
//   velocity scatter, pair q     -> 8N + 6q   pop_tracer.c:1071

// Every entry point of the C ABI that calls HIP holds this guard (shared,
// once per thread however deeply the entry points nest); the capture of a
// mock's HIP graphs holds the same lock exclusively, since HIP refuses work
// on the legacy stream, from any thread, while a stream is being captured.
struct ApiGuard {
  bool held;
  ApiGuard();
  ~ApiGuard();
};

#define HIP_OK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "EZmock: HIP error %s at %s:%d\n",                    \
              hipGetErrorString(e_), __FILE__, __LINE__);                   \
      return EZMOCK_ERR_UNKNOWN;                                            \
    }                                                                       \
  } while (0)

static inline unsigned ez_blocks(size_t n, unsigned per) {
  return (unsigned) ((n + per - 1) / per);
}

// Division of 32-bit unsigned integers by a run-time invariant divisor d
// (1 <= d < 2^32) with one high multiply, exact for every numerator
// (T. Granlund and P. L. Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, fig. 4.1): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, t = mulhi(m, n),
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
struct FastDiv {
  uint32_t d, m;
  int s1, s2;
};

static inline FastDiv fastdiv_make(uint32_t d) {
  FastDiv f;
  int l = 0;
  while (l < 32 && (1ull << l) < d) l++;
  f.d = d;
  f.m = (uint32_t) ((((unsigned __int128) 1 << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

__host__ __device__ __forceinline__ uint32_t fastdiv_q(uint32_t n, const FastDiv &f) {
  const uint32_t t = (uint32_t) (((uint64_t) f.m * n) >> 32);   // mul_hi
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// n = q d + r
__host__ __device__ __forceinline__ uint32_t fastdiv_qr(uint32_t n, const FastDiv &f,
    uint32_t &r) {
  const uint32_t q = fastdiv_q(n, f);
  r = n - q * f.d;
  return q;
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the
// MI355X (workgroup b runs on XCD b % 8), each with its own L2.  For kernels
// whose neighbouring work items share cache lines (gathers), map workgroup b
// to an item so that every XCD gets one contiguous range of items:
// XCD x takes items [x*per + min(x, rem), ... + per + (x < rem)).
#define EZ_NUM_XCD 8
__device__ __forceinline__ unsigned xcd_item(unsigned b, unsigned n) {
  const unsigned per = n / EZ_NUM_XCD, rem = n % EZ_NUM_XCD;
  const unsigned x = b % EZ_NUM_XCD, q = b / EZ_NUM_XCD;
  return x * per + (x < rem ? x : rem) + q;
}
