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
/* Device address of the error word kernels raise (generated fused
 * operations report into it too); read it with cvlhip_check_device_errors. */
void *cvlhip_device_error_word(void);

/* ---- segment descriptors ------------------------------------------------
 * Byte offset of the m int32 segment lengths inside a descriptor built by
 * mke_fov (after a 256-byte header that records n and m); a host may read
 * them there instead of copying them out with len_fos. */
int cvlhip_segd_lengths_offset(void);

/* ---- HIP graphs ------------------------------------------------------
 * A host can record a sequence of CVL calls that return no host data
 * (vector operations, not reductions to a scalar, ext_vu*, v2c, ...) into
 * a HIP graph and replay it: the VCODE interpreter does this for basic
 * blocks (option -G).  Capture needs a stream of its own. */
/* A new non-blocking stream (default priority), for cvlhip_set_stream. */
void *cvlhip_stream_create(void);
/* Start capturing the work queued on the current stream; 0 on success. */
int cvlhip_capture_begin(void);
/* End the capture and instantiate it; *exec receives the executable graph.
 * Returns 0 on success; on failure nothing was executed. */
int cvlhip_capture_end(void **exec);
/* Replay an executable graph on the current stream; 0 on success. */
int cvlhip_graph_launch(void *exec);
void cvlhip_graph_destroy(void *exec);

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

/* ---- segmented-op schedules ---------------------------------------- */
/* A segmented scan / reduction can run under several schedules ("arms"):
 * 0 flat look-back tiles of 8 rows per wave, 1 of 2 rows, 2 of 1 row,
 * 3 thread per segment, 4 wave per segment, 5 workgroup per segment.  The
 * CVL entries (add_sez, max_red, ...) run the arm mke_fov picked for their
 * descriptor (cvlhip_segd_pick).  The calls below run the arm given: op 0 +,
 * 1 *, 2 min, 3 max; ty 0 int, 1 double; same arguments and results as the
 * CVL functions; scratch: cvlhip_seg_scratch units. */
int cvlhip_seg_arms(void);
void cvlhip_seg_scan(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
void cvlhip_seg_reduce(int op, int ty, int arm, vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int cvlhip_seg_scratch(int n, int m);
/* add_sed and max_red of the same values in one pass, reading s once:
 * scan_d gets add_sed(s, sd), red_d gets max_red(s, sd) (n elements, m
 * segments; scratch: cvlhip_add_sed_max_red_scratch units).  No
 * counterpart in the reference, whose programs call the two CVL functions
 * (nesl/cvl/serial/vprims.c:62-109 and 155-201) one after the other; the
 * scan differs from add_sed only in the association of its float sums. */
void cvlhip_add_sed_max_red(vec_p scan_d, vec_p red_d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int cvlhip_add_sed_max_red_scratch(int n, int m);
/* Descriptor statistics for the dispatcher: the features its network reads
 * (host floats, cvlhip_seg_features_count of them; the last two, 8-byte
 * values and reduction, are left 0 for the caller to set). */
int cvlhip_seg_features_count(void);
void cvlhip_segd_features(vec_p sd, int n, int m, float *out);
/* The arm a descriptor's pick names for a scan (reduce 0) or reduction of
 * elem_bytes-byte values, or -1 when it holds none (fewer than
 * cvlhip_sched_min_n elements, or built with the dispatcher off). */
int cvlhip_segd_pick(vec_p sd, int reduce, int elem_bytes);
int cvlhip_sched_min_n(void);
/* The network mke_fov runs: the library is built with the trained one
 * (neslgpu_amd/dispatch_model.npz); set_params installs another (device
 * pointer, the P floats of cvlhip_mlp_dims) for the descriptors built
 * after it, or none (NULL); get_params copies the current one out (device
 * pointer). */
void cvlhip_sched_set_params(const float *params);
void cvlhip_sched_get_params(float *params);
int cvlhip_sched_has_model(void);
/* 0: mke_fov picks nothing (every segmented op runs arm 0); 1: default.
 * CVLHIP_SCHED=0 in the environment sets 0 when the library loads. */
void cvlhip_sched_enable(int on);

/* ---- callers that cannot allocate -------------------------------------
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc).  add_sez (same d, s, sd, n, m and result) with its scratch in
 * a static 4 MiB device array per GPU instead of a scratch argument; it
 * calls no allocator.  n may be at most 1073545216: a larger n is rejected
 * with an error before any work is queued.  Calls that overlap on one GPU
 * race, as they share that GPU's array. */
void cvlhip_add_sez_noalloc(vec_p d, vec_p s, vec_p sd, int n, int m);

/* ---- learned dispatcher network (device pointers, fp32) -------------- */
/* Sizes: F input features, H hidden units (two ReLU layers), A outputs (one
 * predicted log-time per schedule), P parameters in the flat array. */
void cvlhip_mlp_dims(int *features, int *hidden, int *arms, int *params);
/* out: B*A predictions; acts: 2*B*H floats kept for the backward pass. */
void cvlhip_mlp_forward(const float *params, const float *X, int B, float *out, float *acts);
/* Masked mean-squared-error gradients of one batch: T targets, M mask
 * (B*A each), mask_sum = sum of M; grads: P floats; loss: 1 float;
 * acts: 2*B*H floats; work: B*(2H+2A) floats. */
void cvlhip_mlp_gradients(const float *params, const float *X, const float *T, const float *M, int B,
                          float mask_sum, float *grads, float *loss, float *acts, float *work);
/* One Adam step, t >= 1. */
void cvlhip_mlp_adam(float *params, const float *grads, float *m, float *v, int t, float lr);

#ifdef __cplusplus
}
#endif

#endif /* cvl_hip_ext_h */
