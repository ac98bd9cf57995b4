/*
 * cvl_hip_ext.h -- entry points of the HIP CVL backend that have no
 * counterpart in the reference CVL API (nesl/include/cvl.h).  They serve
 * hosts that embed the library (stream selection, error reporting) and the
 * segmented-operation scheduler (strategies, autotuner, learned dispatcher).
 * None of them is needed by the VCODE interpreter.
 */
#ifndef cvl_hip_ext_h
#define cvl_hip_ext_h

#include "cvl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------- */
/* Queue all CVL work on HIP stream `s` (a hipStream_t); NULL = null stream. */
void cvlhip_set_stream(void *s);
void *cvlhip_get_stream(void);
/* 1 (default): an error prints a message and exits, as the reference CUDA
 * backend does.  0: errors are recorded and returned by cvlhip_last_error. */
void cvlhip_set_error_mode(int fatal_on_error);
/* Returns 0 if no error since the last call, else a non-zero code and the
 * message in buf. */
int cvlhip_last_error(char *buf, int len);
/* Non-zero when a kernel met an out-of-range permutation index since the
 * last call (the access was skipped); clears the flag. */
unsigned cvlhip_check_device_errors(void);
void cvlhip_synchronize(void);
/* Grid of the tile kernels (scan, pack): cap > 0 keeps at most `cap`
 * resident blocks per compute unit, each looping over tiles; 0 = as many
 * resident blocks as the occupancy allows; -1 = one block per tile. */
void cvlhip_set_blocks_per_cu(int cap);

/* ---- random numbers -------------------------------------------------- */
/* The 48-bit lrand48 state behind RAND (rnd_wuz).  It starts at 0, as an
 * unseeded glibc lrand48 does; rnd_foz leaves it alone, as the serial CVL
 * on Linux does (its srandom call does not touch lrand48). */
void cvlhip_rand_set_state(unsigned long long state48);
unsigned long long cvlhip_rand_get_state(void);

/* ---- synthetic inputs ------------------------------------------------ */
/* splitmix64 counter streams, element k = mix(seed + (k+1)*golden):
 * u32 = high 32 bits; f64 = lo + (hi-lo) * (top 53 bits * 2^-53). */
void cvlhip_fill_splitmix_u32(vec_p d, long long n, unsigned long long seed);
void cvlhip_fill_splitmix_f64(vec_p d, long long n, unsigned long long seed, double lo, double hi);

#ifdef __cplusplus
}
#endif

#endif /* cvl_hip_ext_h */
