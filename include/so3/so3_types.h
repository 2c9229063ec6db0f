/*
 * so3_types.h -- parameter struct, enums and constants of the SO(3) Wigner
 * transform C API.
 *
 * ABI-identical replacement for the reference header
 * include/so3/so3_types.h:50-201: same enum orders, same field order and
 * widths of so3_parameters_t (11 x 4-byte fields), same SO3_COMPLEX macro.
 *
 * The reference takes ssht_dl_method_t from the SSHT library header.  That
 * library is not needed here: when <ssht/ssht.h> is not on the include path
 * this header declares a layout-compatible enum (RISBO = 0, TRAPANI = 1, the
 * order the reference's Python binding declares at src/so3/bindings.pyx:93-94).
 */
#ifndef SO3_TYPES
#define SO3_TYPES

#if defined(__has_include)
#if __has_include(<ssht/ssht.h>)
#include <ssht/ssht.h>
#define SO3_HAVE_SSHT_HEADER 1
#endif
#endif

#ifndef SO3_HAVE_SSHT_HEADER
#ifndef SSHT_DL_METHOD_T_DEFINED
#define SSHT_DL_METHOD_T_DEFINED
/* Wigner small-d recursion selector (Risbo 1996; Trapani & Navaza 2006). */
typedef enum { SSHT_DL_RISBO = 0, SSHT_DL_TRAPANI = 1 } ssht_dl_method_t;
#endif
#endif

#ifdef __cplusplus
#include <complex>
#define SO3_COMPLEX(TYPE) std::complex<TYPE>
#else
#include <complex.h>
#define SO3_COMPLEX(TYPE) TYPE complex
#endif

#define SO3_PI 3.141592653589793238462643383279502884197
#define SO3_PION2 1.570796326794896619231321691639751442099
#define SO3_SQRT2 1.41421356237309504880168872420969807856967
#define SO3_PROMPT "[so3] "

/* Storage order of the flm blocks in n: 0, -1, 1, -2, 2, ... or -N+1 .. N-1. */
typedef enum {
  SO3_N_ORDER_ZERO_FIRST,
  SO3_N_ORDER_NEGATIVE_FIRST,
  SO3_N_ORDER_SIZE
} so3_n_order_t;

/* Padded: every n block holds L*L entries.  Compact: l < |n| is not stored. */
typedef enum { SO3_STORAGE_PADDED, SO3_STORAGE_COMPACT, SO3_STORAGE_SIZE } so3_storage_t;

/* Which n blocks may be non-zero. */
typedef enum {
  SO3_N_MODE_ALL,
  SO3_N_MODE_EVEN,
  SO3_N_MODE_ODD,
  SO3_N_MODE_MAXIMUM,
  SO3_N_MODE_L,
  SO3_N_MODE_SIZE
} so3_n_mode_t;

/* MW: (2L-1) x L x (2N-1) samples.  MW_SS: 2L x (L+1) x (2N-1) samples. */
typedef enum { SO3_SAMPLING_MW, SO3_SAMPLING_MW_SS, SO3_SAMPLING_SIZE } so3_sampling_t;

/* Zero-initialise ("so3_parameters_t p = {0};") so that unset fields take
 * their defaults: MW, zero-first, padded, all n, Risbo, not steerable. */
typedef struct {
  int verbosity;                   /* diagnostic output level 0..5 */
  int reality;                     /* non-zero: the signal f is real */
  int L0;                          /* lower harmonic band-limit */
  int L;                           /* upper harmonic band-limit */
  int N;                           /* orientational band-limit */
  so3_sampling_t sampling_scheme;  /* MW or MW_SS */
  so3_n_order_t n_order;           /* n order of the flm blocks */
  so3_storage_t storage;           /* padded or compact */
  so3_n_mode_t n_mode;             /* which n are non-zero */
  ssht_dl_method_t dl_method;      /* Risbo or Trapani Delta recursion */
  int steerable;                   /* non-zero: steerable signal */
} so3_parameters_t;

#endif
