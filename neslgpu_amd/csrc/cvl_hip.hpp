// cvl_hip.hpp -- shared definitions of the MI355X (gfx950) CVL backend.
//
// Operators and identities follow the serial CVL
// (ref nesl/cvl/serial/defins.h:61-113, 218-228; vprims.c:30-49).  Every
// operator here is applied as op(earlier, later).  Scans and reductions are
// computed with an *internal neutral* element that is a true identity of the
// operator (0, 1, ~0, INT_MAX/INT_MIN, +-inf for double min/max) and the CVL
// identity is folded in once, on the left, when a result is written:
//     result = op(cvl_identity, fold(x_0 .. x_k))
// which equals the serial left fold because each operator is associative.
// For double max/min this matters: the CVL identity (float-rounded 1e36) is
// not an identity for |x| > 1e36, so it may only ever appear leftmost.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

namespace cvlhip {

// ------------------------------------------------------------------ runtime
extern hipStream_t g_stream;
inline hipStream_t stream() { return g_stream; }

[[noreturn]] void fatal(const char *fmt, ...);
void hip_fail(hipError_t e, const char *expr, const char *file, int line);
// Records a usage error (bad size, bad argument).  Fatal unless the caller
// switched to recorded errors (cvlhip_set_error_mode).
void usage_error(const char *fmt, ...);
// Device word raised by kernels that met an out-of-range index (the write or
// read is skipped instead of faulting).
unsigned *device_error_word();
// The error word of the no-allocation entry's static array on the current
// device, or null before that entry's first call there.
unsigned *noalloc_error_word();

#define HIPCHECK(x)                                                        \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) ::cvlhip::hip_fail(e_, #x, __FILE__, __LINE__); \
    } while (0)
#define LAUNCHCHECK() HIPCHECK(hipGetLastError())

// Size of one CVL memory unit: the union of int, float and double
// (ref nesl/cvl/serial/facilt.c:92-93).
constexpr int kUnit = 8;

// Units needed to hold `bytes` bytes.
inline int units_for_bytes(int64_t bytes) { return (int)((bytes + kUnit - 1) / kUnit); }

// ------------------------------------------------------------------ launch
constexpr int kBlock = 256;     // 4 waves of 64

inline int grid_for(int64_t n, int per_block) {
    int64_t g = (n + per_block - 1) / per_block;
    return (int)(g < 1 ? 1 : g);
}

// ------------------------------------------------------------------ operators
// Each operator: T neutral() (true identity), T cvl_id() (the serial
// identity), T operator()(T a, T b) with a the earlier operand.

struct AddZ {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return (int)((unsigned)a + (unsigned)b); }
};
struct MulZ {
    using T = int;
    __host__ __device__ static T neutral() { return 1; }
    __host__ __device__ static T cvl_id() { return 1; }
    __host__ __device__ T operator()(T a, T b) const { return (int)((unsigned)a * (unsigned)b); }
};
struct MinZ {
    using T = int;
    __host__ __device__ static T neutral() { return INT_MAX; }
    __host__ __device__ static T cvl_id() { return INT_MAX; }
    __host__ __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct MaxZ {
    using T = int;
    __host__ __device__ static T neutral() { return INT_MIN; }
    __host__ __device__ static T cvl_id() { return INT_MIN; }
    __host__ __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};
struct AndZ {
    using T = int;
    __host__ __device__ static T neutral() { return ~0; }
    __host__ __device__ static T cvl_id() { return ~0; }
    __host__ __device__ T operator()(T a, T b) const { return a & b; }
};
struct IorZ {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return a | b; }
};
struct XorZ {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return a ^ b; }
};
// Logical operators on cvl_bool.  Inputs are normalised to 0/1 when loaded
// (Op::load), after which &, | and ^ are the logical &&, || and lxor.
struct AndB {
    using T = int;
    __host__ __device__ static T neutral() { return 1; }
    __host__ __device__ static T cvl_id() { return 1; }
    __host__ __device__ T operator()(T a, T b) const { return a & b; }
};
struct IorB {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return a | b; }
};
struct XorB {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return a ^ b; }
};
struct AddD {
    using T = double;
    __host__ __device__ static T neutral() { return 0.0; }
    __host__ __device__ static T cvl_id() { return 0.0; }
    __host__ __device__ T operator()(T a, T b) const { return a + b; }
};
struct MulD {
    using T = double;
    __host__ __device__ static T neutral() { return 1.0; }
    __host__ __device__ static T cvl_id() { return 1.0; }
    __host__ __device__ T operator()(T a, T b) const { return a * b; }
};
struct MinD {
    using T = double;
    __host__ __device__ static T neutral() { return __builtin_inf(); }
    __host__ __device__ static T cvl_id() { return (double)(float)1.0e36; }
    __host__ __device__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct MaxD {
    using T = double;
    __host__ __device__ static T neutral() { return -__builtin_inf(); }
    __host__ __device__ static T cvl_id() { return (double)(float)-1.0e36; }
    __host__ __device__ T operator()(T a, T b) const { return a > b ? a : b; }
};

struct AddU {
    using T = unsigned;
    __host__ __device__ static T neutral() { return 0u; }
    __host__ __device__ static T cvl_id() { return 0u; }
    __host__ __device__ T operator()(T a, T b) const { return a + b; }
};

// Number of true flags: cvl_bool values normalised to 0/1, then added
// (ref nesl/cvl/serial/library.c:66-75, 114-133 count `if (flag)`).
struct CountB {
    using T = int;
    __host__ __device__ static T neutral() { return 0; }
    __host__ __device__ static T cvl_id() { return 0; }
    __host__ __device__ T operator()(T a, T b) const { return (int)((unsigned)a + (unsigned)b); }
};

template <class Op> struct is_logical { static constexpr bool value = false; };
template <> struct is_logical<CountB> { static constexpr bool value = true; };
template <> struct is_logical<AndB> { static constexpr bool value = true; };
template <> struct is_logical<IorB> { static constexpr bool value = true; };
template <> struct is_logical<XorB> { static constexpr bool value = true; };

// Value as seen by the operator (cvl_bool normalised to 0/1).
template <class Op>
__device__ __forceinline__ typename Op::T load_val(typename Op::T x) {
    if constexpr (is_logical<Op>::value) return x != 0;
    else return x;
}

// ------------------------------------------------------------------ wave64
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// 32-bit DPP move with `old` kept where the source lane is out of range.
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ int dpp_i32(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROW_MASK, BANK_MASK, false);
}

// A value seen as its 32-bit words (for moving wider types across lanes).
template <class T>
struct Words {
    static_assert(sizeof(T) % 4 == 0, "cross-lane moves of whole 32-bit words");
    int w[sizeof(T) / 4];
};

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf, class T>
__device__ __forceinline__ T dpp_move(T old, T src) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_i32<CTRL, ROW_MASK, BANK_MASK>(
                                           __builtin_bit_cast(int, old), __builtin_bit_cast(int, src)));
    } else if constexpr (sizeof(T) == 8) {
        uint64_t o = __builtin_bit_cast(uint64_t, old), s = __builtin_bit_cast(uint64_t, src);
        int lo = dpp_i32<CTRL, ROW_MASK, BANK_MASK>((int)(uint32_t)o, (int)(uint32_t)s);
        int hi = dpp_i32<CTRL, ROW_MASK, BANK_MASK>((int)(uint32_t)(o >> 32), (int)(uint32_t)(s >> 32));
        return __builtin_bit_cast(T, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    } else {
        const Words<T> o = __builtin_bit_cast(Words<T>, old), s = __builtin_bit_cast(Words<T>, src);
        Words<T> r;
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); i++) r.w[i] = dpp_i32<CTRL, ROW_MASK, BANK_MASK>(o.w[i], s.w[i]);
        return __builtin_bit_cast(T, r);
    }
}

// DPP control codes (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

// Inclusive scan across the 64 lanes of a wave (lane 0 first).  Lanes that
// have no source in a step combine with the neutral element.
template <class Op, class T>
__device__ __forceinline__ T wave_inclusive_scan(T x, Op op) {
    const T e = Op::neutral();
    x = op(dpp_move<kRowShr1>(e, x), x);
    x = op(dpp_move<kRowShr2>(e, x), x);
    x = op(dpp_move<kRowShr4>(e, x), x);
    x = op(dpp_move<kRowShr8>(e, x), x);
    x = op(dpp_move<kRowBcast15, 0xa>(e, x), x);
    x = op(dpp_move<kRowBcast31, 0xc>(e, x), x);
    return x;
}

template <class T>
__device__ __forceinline__ T shfl(T v, int src_lane) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __shfl(__builtin_bit_cast(int, v), src_lane, kWave));
    } else if constexpr (sizeof(T) == 8) {
        uint64_t u = __builtin_bit_cast(uint64_t, v);
        int lo = __shfl((int)(uint32_t)u, src_lane, kWave);
        int hi = __shfl((int)(uint32_t)(u >> 32), src_lane, kWave);
        return __builtin_bit_cast(T, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    } else {
        Words<T> u = __builtin_bit_cast(Words<T>, v);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(T) / 4); i++) u.w[i] = __shfl(u.w[i], src_lane, kWave);
        return __builtin_bit_cast(T, u);
    }
}

template <class T>
__device__ __forceinline__ T shfl_up1(T v) {
    int l = lane_id();
    return shfl(v, l > 0 ? l - 1 : 0);
}

template <class T>
__device__ __forceinline__ T bcast_last(T v) { return shfl(v, kWave - 1); }

// Reduction over the lanes of a wave; every lane gets the result.  Lanes
// are combined in index order (lower lane = earlier operand).
template <class Op, class T>
__device__ __forceinline__ T wave_reduce(T x, Op op) {
    return bcast_last(wave_inclusive_scan(x, op));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// ------------------------------------------------------------------ XCD-aware tiles
// MI355X spreads consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  xcd_tile maps block b of a grid of g blocks to a tile
// index such that the blocks of one XCD (b % 8 == x) take a contiguous range
// of tiles; neighbouring tiles' partial-line writes (histogram columns, the
// ends of bucket runs) then meet in the same L2 and merge before they reach
// HBM.  It is a bijection on [0, g).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t g) {
    const int64_t q = g / kXcds;
    if (b >= kXcds * q) return b;
    return (b % kXcds) * q + b / kXcds;
}

// ------------------------------------------------------------------ coherent status words
// Tile status words of single-pass scans live in global memory and are
// written and polled by different workgroups, possibly on different XCDs:
// agent-scope relaxed atomics keep them coherent without fences, because a
// value and its flag always travel in the same 64-bit word.
__device__ __forceinline__ uint64_t ld_status(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// ------------------------------------------------------------------ fills
void fill_u32(void *p, uint32_t v, int64_t count);   // stream-ordered fill kernel
void zero2_u32(void *p1, int64_t n1, void *p2, int64_t n2);   // two regions, one launch
void zero3_u32(void *p1, int64_t n1, void *p2, int64_t n2, void *p3, int64_t n3);

// ------------------------------------------------------------------ scalar results to the host
// A CVL call that returns a scalar (a reduction, ext_vu*) has its kernel
// post the value into a word pair of pinned, device-mapped host memory;
// the host waits on the sequence number instead of copying the value out
// and synchronizing the stream.  Posts are ordered: the value is written
// before the sequence number, both at system scope.
struct Mailbox {
    unsigned long long seq;
    unsigned long long value;
};
Mailbox *mailbox_device();                              // device address (allocated on first use)
unsigned long long mailbox_next();                      // sequence number of the next post
unsigned long long mailbox_wait(unsigned long long seq);   // host: the value posted with seq

__device__ __forceinline__ void mailbox_post(Mailbox *mb, unsigned long long seq, unsigned long long bits) {
    __hip_atomic_store(&mb->value, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__host__ __device__ __forceinline__ unsigned long long scalar_bits(T v) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(unsigned long long, v);
    else return (unsigned long long)__builtin_bit_cast(unsigned, v);
}
template <class T>
__host__ __device__ __forceinline__ T scalar_from_bits(unsigned long long b) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, b);
    else return __builtin_bit_cast(T, (unsigned)b);
}

}  // namespace cvlhip
