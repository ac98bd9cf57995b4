/*
 * vcode_fused.h -- ABI of a fused-operation library, the shared object a
 * VCODE interpreter loads (option -l) to execute `FUSED n` instructions.
 *
 * The layout is the reference's, so libraries and interpreters of either
 * side interoperate: the descriptor vopdes_t and the function triple
 * cvl_triple_t of nesl/vcode/vcode.h:73-103, the TYPE enumeration of
 * vcode.h:12-14, the call-shape enumeration cvl_desc_t of vcode.h:58-71 and
 * the argument-constraint bits of vcode.h:36-52.  A library exports
 *
 *     vopdes_t     vops[];        descriptor of fused op n at index n
 *     cvl_triple_t cvl_funs[];    its function, scratch and in-place hooks
 *     int          vops_count;    number of entries (optional)
 *     void         init(void *);  optional set-up hook (a no-op here: a
 *                                 vec_p is already a device address)
 *
 * Fused op n with k vector arguments is called, like every CVL elementwise
 * function, as   fusedN(vec_p d, vec_p s0, ..., vec_p s(k-1), int len,
 * vec_p scratch);  where len is the element count of argument 0 (of the
 * vector it describes, when argument 0 is a segment descriptor).  The
 * generator of such libraries is neslgpu_amd/fusion (HIP for gfx950).
 */
#ifndef vcode_fused_h
#define vcode_fused_h

#include "cvl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum type {
    Illegal = 0, Pair, None, Bool, Int, Float, Segdes, String, Agree1, Agree2, Agree3, Agree4, Given
} TYPE;

#define MAX_ARGS 40
#define MAX_OUT 3

/* argument constraints (arg_len) and result lengths (result_len) */
#define NONE     0x0000
#define AGREE1   0x0001 /* length == length of argument 1 */
#define AGREE2   0x0002
#define AGREE3   0x0004
#define AGREE4   0x0008
#define AGREE5   0x0010
#define COMPAT1  0x0020 /* a segment descriptor describing argument 1 */
#define COMPAT2  0x0040
#define COMPAT3  0x0080
#define COMPAT4  0x0100
#define COMPAT5  0x0200
#define IS_SCALAR 0x0400

typedef enum CVL_DESC_T {
    Elwise0, Elwise1, Elwise2, Elwise3, Elwise4, Elwise5, Elwise6, Elwise7, Elwise8, Elwise9,
    Elwise10, Elwise11, Elwise12, Elwise13, Elwise14, Elwise15, Elwise16, Elwise17, Elwise18, Elwise19,
    Elwise20, Elwise21, Elwise22, Elwise23, Elwise24, Elwise25, Elwise26, Elwise27, Elwise28, Elwise29,
    Elwise30, Elwise31, Elwise32, Elwise33, Elwise34, Elwise35, Elwise36, Elwise37, Elwise38, Elwise39,
    Elwise40,
    Scan, Reduce, Special, SegOp,
    Permute, Dpermute, Fpermute, Bpermute, Bfpermute, Dfpermute,
    Extract, Replace, Dist, Index, Pack, NotSupported
} cvl_desc_t;

typedef struct {
    int vop;                          /* opcode token (informational) */
    const char *vopname;
    int arg_num;
    int result_num;
    TYPE arg_type[MAX_ARGS];
    unsigned arg_len[MAX_ARGS];
    TYPE result_type[MAX_OUT];
    unsigned result_len[MAX_OUT];
    int arg_for_scratch[MAX_ARGS];    /* 1-based arguments sizing the scratch */
    cvl_desc_t cvl_desc;
} vopdes_t;

typedef struct {
    void (*function)(void);
    int (*scratch_function)(void);
    unsigned (*inplace_function)(void);
} cvl_funct_t;

typedef struct {
    cvl_funct_t cvl_int;
    cvl_funct_t cvl_float;
    cvl_funct_t cvl_boolean;
} cvl_triple_t;

/* value of the FUSED token in the reference grammar's numbering
 * (nesl/vcode/grammar.yy:42-96, tokens numbered from 258 in order) */
#define VCODE_FUSED_TOKEN 370

#ifdef __cplusplus
}
#endif

#endif /* vcode_fused_h */
