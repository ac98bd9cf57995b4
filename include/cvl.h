/*
 * cvl.h -- C API of the MI355X (gfx950, HIP) backend of NESL's CVL vector
 * library.
 *
 * This header is a drop-in for the reference CVL header
 * (nesl/include/cvl.h:121-741): the same function names, argument order and
 * return types, so the VCODE interpreter (nesl/vcode) and fused-operation
 * libraries compile against it unchanged.  What differs is the memory model:
 *
 *   - vec_p is a DEVICE pointer (HIP global memory).  The interpreter obtains
 *     one arena from alo_foz() and sub-allocates it, exactly as with the
 *     serial backend (nesl/vcode/vstack.c:156-160); it must not dereference
 *     vec_p on the host.  Host data moves through v2c_fu* / c2v_fu* and the
 *     scalar entry points (reductions, ext_vu*, rep_vu*, dis_vu*).
 *   - One CVL unit is 8 bytes (the int/float/double union of the serial
 *     backend, nesl/cvl/serial/facilt.c:92-93), so doubles stay 8-byte
 *     aligned and keep full double precision.
 *   - cvl_bool is a 4-byte int (nesl/include/cvl.h:60-68).
 *   - A segment descriptor holds a small header (n, m), the m segment
 *     lengths (the whole of the serial backend's descriptor) and derived
 *     data (offsets, a map from element blocks to segments, and a
 *     head-flag bitmap) built once by mke_fov; siz_fos(n, m) gives its
 *     size.
 *   - Every operation's X_scratch() gives the device scratch, in CVL units,
 *     that the caller passes as `scratch` (the interpreter guarantees it,
 *     nesl/vcode/check_args.c:264-311).
 *   - Work is queued on one HIP stream (the null stream unless
 *     cvlhip_set_stream() names another); functions that return host data
 *     wait for it.  Errors print a message and exit, as the reference CUDA
 *     backend does (nesl/include/cutil_inline.h:23-53).
 *
 * Each prototype replaces the reference prototype of the same name in
 * nesl/include/cvl.h; the families below follow the order of that file.
 */
#ifndef cvl_h
#define cvl_h

#include <time.h>

typedef void *vec_p;
typedef int cvl_bool;
typedef struct timespec cvl_timer_t;

#define CVL_SCRATCH_NULL ((vec_p) 0)
#define INPLACE_NONE 0
#define INPLACE_1 0x001
#define INPLACE_2 0x002
#define INPLACE_3 0x004
#define INPLACE_4 0x008
#define INPLACE_5 0x010
#define INPLACE_6 0x020
#define INPLACE_7 0x040
#define INPLACE_8 0x080
#define INPLACE_9 0x100
#define INPLACE_10 0x200
#define INPLACE_11 0x400
#define INPLACE_12 0x800

#ifdef __cplusplus
extern "C" {
#endif

/* Library start-up.  The argument names the fused-operation library the
 * interpreter will load (nesl/vcode/main.c:182-184); this backend keeps no
 * base pointer, so fused libraries need no init call from here. */
void CVL_init(char *fused_library);

/* Elementwise operations (d[i] = op(s1[i], s2[i]...)); comparisons write cvl_bool */
void max_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int max_wuz_scratch(int len);
unsigned int max_wuz_inplace(void);
void max_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int max_wud_scratch(int len);
unsigned int max_wud_inplace(void);
void min_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int min_wuz_scratch(int len);
unsigned int min_wuz_inplace(void);
void min_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int min_wud_scratch(int len);
unsigned int min_wud_inplace(void);
void add_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int add_wuz_scratch(int len);
unsigned int add_wuz_inplace(void);
void add_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int add_wud_scratch(int len);
unsigned int add_wud_inplace(void);
void sub_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int sub_wuz_scratch(int len);
unsigned int sub_wuz_inplace(void);
void sub_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int sub_wud_scratch(int len);
unsigned int sub_wud_inplace(void);
void mul_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int mul_wuz_scratch(int len);
unsigned int mul_wuz_inplace(void);
void mul_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int mul_wud_scratch(int len);
unsigned int mul_wud_inplace(void);
void div_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int div_wuz_scratch(int len);
unsigned int div_wuz_inplace(void);
void div_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int div_wud_scratch(int len);
unsigned int div_wud_inplace(void);
void grt_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int grt_wuz_scratch(int len);
unsigned int grt_wuz_inplace(void);
void grt_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int grt_wud_scratch(int len);
unsigned int grt_wud_inplace(void);
void les_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int les_wuz_scratch(int len);
unsigned int les_wuz_inplace(void);
void les_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int les_wud_scratch(int len);
unsigned int les_wud_inplace(void);
void geq_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int geq_wuz_scratch(int len);
unsigned int geq_wuz_inplace(void);
void geq_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int geq_wud_scratch(int len);
unsigned int geq_wud_inplace(void);
void leq_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int leq_wuz_scratch(int len);
unsigned int leq_wuz_inplace(void);
void leq_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int leq_wud_scratch(int len);
unsigned int leq_wud_inplace(void);
void lsh_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int lsh_wuz_scratch(int len);
unsigned int lsh_wuz_inplace(void);
void rsh_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int rsh_wuz_scratch(int len);
unsigned int rsh_wuz_inplace(void);
void mod_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int mod_wuz_scratch(int len);
unsigned int mod_wuz_inplace(void);
void rnd_wuz(vec_p d, vec_p s, int len, vec_p scratch);
int rnd_wuz_scratch(int len);
unsigned int rnd_wuz_inplace(void);
void eql_wub(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int eql_wub_scratch(int len);
unsigned int eql_wub_inplace(void);
void eql_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int eql_wuz_scratch(int len);
unsigned int eql_wuz_inplace(void);
void eql_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int eql_wud_scratch(int len);
unsigned int eql_wud_inplace(void);
void neq_wub(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int neq_wub_scratch(int len);
unsigned int neq_wub_inplace(void);
void neq_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int neq_wuz_scratch(int len);
unsigned int neq_wuz_inplace(void);
void neq_wud(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int neq_wud_scratch(int len);
unsigned int neq_wud_inplace(void);
void sel_wuz(vec_p d, vec_p s1, vec_p s2, vec_p s3, int len, vec_p scratch);
int sel_wuz_scratch(int len);
unsigned int sel_wuz_inplace(void);
void sel_wub(vec_p d, vec_p s1, vec_p s2, vec_p s3, int len, vec_p scratch);
int sel_wub_scratch(int len);
unsigned int sel_wub_inplace(void);
void sel_wud(vec_p d, vec_p s1, vec_p s2, vec_p s3, int len, vec_p scratch);
int sel_wud_scratch(int len);
unsigned int sel_wud_inplace(void);
void not_wub(vec_p d, vec_p s, int len, vec_p scratch);
int not_wub_scratch(int len);
unsigned int not_wub_inplace(void);
void xor_wub(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int xor_wub_scratch(int len);
unsigned int xor_wub_inplace(void);
void ior_wub(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int ior_wub_scratch(int len);
unsigned int ior_wub_inplace(void);
void and_wub(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int and_wub_scratch(int len);
unsigned int and_wub_inplace(void);
void ior_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int ior_wuz_scratch(int len);
unsigned int ior_wuz_inplace(void);
void and_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int and_wuz_scratch(int len);
unsigned int and_wuz_inplace(void);
void not_wuz(vec_p d, vec_p s, int len, vec_p scratch);
int not_wuz_scratch(int len);
unsigned int not_wuz_inplace(void);
void xor_wuz(vec_p d, vec_p s1, vec_p s2, int len, vec_p scratch);
int xor_wuz_scratch(int len);
unsigned int xor_wuz_inplace(void);
void int_wud(vec_p d, vec_p s, int len, vec_p scratch);
int int_wud_scratch(int len);
unsigned int int_wud_inplace(void);
void int_wub(vec_p d, vec_p s, int len, vec_p scratch);
int int_wub_scratch(int len);
unsigned int int_wub_inplace(void);
void dbl_wuz(vec_p d, vec_p s, int len, vec_p scratch);
int dbl_wuz_scratch(int len);
unsigned int dbl_wuz_inplace(void);
void boo_wuz(vec_p d, vec_p s, int len, vec_p scratch);
int boo_wuz_scratch(int len);
unsigned int boo_wuz_inplace(void);
void flr_wud(vec_p d, vec_p s, int len, vec_p scratch);
int flr_wud_scratch(int len);
unsigned int flr_wud_inplace(void);
void cei_wud(vec_p d, vec_p s, int len, vec_p scratch);
int cei_wud_scratch(int len);
unsigned int cei_wud_inplace(void);
void trn_wud(vec_p d, vec_p s, int len, vec_p scratch);
int trn_wud_scratch(int len);
unsigned int trn_wud_inplace(void);
void rou_wud(vec_p d, vec_p s, int len, vec_p scratch);
int rou_wud_scratch(int len);
unsigned int rou_wud_inplace(void);
void exp_wud(vec_p d, vec_p s, int len, vec_p scratch);
int exp_wud_scratch(int len);
unsigned int exp_wud_inplace(void);
void log_wud(vec_p d, vec_p s, int len, vec_p scratch);
int log_wud_scratch(int len);
unsigned int log_wud_inplace(void);
void sqt_wud(vec_p d, vec_p s, int len, vec_p scratch);
int sqt_wud_scratch(int len);
unsigned int sqt_wud_inplace(void);
void sin_wud(vec_p d, vec_p s, int len, vec_p scratch);
int sin_wud_scratch(int len);
unsigned int sin_wud_inplace(void);
void cos_wud(vec_p d, vec_p s, int len, vec_p scratch);
int cos_wud_scratch(int len);
unsigned int cos_wud_inplace(void);
void tan_wud(vec_p d, vec_p s, int len, vec_p scratch);
int tan_wud_scratch(int len);
unsigned int tan_wud_inplace(void);
void asn_wud(vec_p d, vec_p s, int len, vec_p scratch);
int asn_wud_scratch(int len);
unsigned int asn_wud_inplace(void);
void acs_wud(vec_p d, vec_p s, int len, vec_p scratch);
int acs_wud_scratch(int len);
unsigned int acs_wud_inplace(void);
void atn_wud(vec_p d, vec_p s, int len, vec_p scratch);
int atn_wud_scratch(int len);
unsigned int atn_wud_inplace(void);
void snh_wud(vec_p d, vec_p s, int len, vec_p scratch);
int snh_wud_scratch(int len);
unsigned int snh_wud_inplace(void);
void csh_wud(vec_p d, vec_p s, int len, vec_p scratch);
int csh_wud_scratch(int len);
unsigned int csh_wud_inplace(void);
void tnh_wud(vec_p d, vec_p s, int len, vec_p scratch);
int tnh_wud_scratch(int len);
unsigned int tnh_wud_inplace(void);
void cpy_wuz(vec_p d, vec_p s, int len, vec_p scratch);
int cpy_wuz_scratch(int len);
unsigned int cpy_wuz_inplace(void);
void cpy_wub(vec_p d, vec_p s, int len, vec_p scratch);
int cpy_wub_scratch(int len);
unsigned int cpy_wub_inplace(void);
void cpy_wud(vec_p d, vec_p s, int len, vec_p scratch);
int cpy_wud_scratch(int len);
unsigned int cpy_wud_inplace(void);
void cpy_wus(vec_p d, vec_p s, int n, int m, vec_p scratch);
int cpy_wus_scratch(int n, int m);
unsigned int cpy_wus_inplace(void);

/* Unsegmented exclusive scans */
void add_suz(vec_p d, vec_p s, int len, vec_p scratch);
int add_suz_scratch(int len);
unsigned int add_suz_inplace(void);
void add_sud(vec_p d, vec_p s, int len, vec_p scratch);
int add_sud_scratch(int len);
unsigned int add_sud_inplace(void);
void mul_suz(vec_p d, vec_p s, int len, vec_p scratch);
int mul_suz_scratch(int len);
unsigned int mul_suz_inplace(void);
void mul_sud(vec_p d, vec_p s, int len, vec_p scratch);
int mul_sud_scratch(int len);
unsigned int mul_sud_inplace(void);
void min_suz(vec_p d, vec_p s, int len, vec_p scratch);
int min_suz_scratch(int len);
unsigned int min_suz_inplace(void);
void min_sud(vec_p d, vec_p s, int len, vec_p scratch);
int min_sud_scratch(int len);
unsigned int min_sud_inplace(void);
void max_suz(vec_p d, vec_p s, int len, vec_p scratch);
int max_suz_scratch(int len);
unsigned int max_suz_inplace(void);
void max_sud(vec_p d, vec_p s, int len, vec_p scratch);
int max_sud_scratch(int len);
unsigned int max_sud_inplace(void);
void and_sub(vec_p d, vec_p s, int len, vec_p scratch);
int and_sub_scratch(int len);
unsigned int and_sub_inplace(void);
void and_suz(vec_p d, vec_p s, int len, vec_p scratch);
int and_suz_scratch(int len);
unsigned int and_suz_inplace(void);
void ior_sub(vec_p d, vec_p s, int len, vec_p scratch);
int ior_sub_scratch(int len);
unsigned int ior_sub_inplace(void);
void ior_suz(vec_p d, vec_p s, int len, vec_p scratch);
int ior_suz_scratch(int len);
unsigned int ior_suz_inplace(void);
void xor_sub(vec_p d, vec_p s, int len, vec_p scratch);
int xor_sub_scratch(int len);
unsigned int xor_sub_inplace(void);
void xor_suz(vec_p d, vec_p s, int len, vec_p scratch);
int xor_suz_scratch(int len);
unsigned int xor_suz_inplace(void);

/* Segmented exclusive scans (sd = segment descriptor with n elements in m segments) */
void add_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int add_sez_scratch(int vec_len, int seg_len);
unsigned int add_sez_inplace(void);
void add_sed(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int add_sed_scratch(int vec_len, int seg_len);
unsigned int add_sed_inplace(void);
void mul_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int mul_sez_scratch(int vec_len, int seg_len);
unsigned int mul_sez_inplace(void);
void mul_sed(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int mul_sed_scratch(int vec_len, int seg_len);
unsigned int mul_sed_inplace(void);
void min_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int min_sez_scratch(int vec_len, int seg_len);
unsigned int min_sez_inplace(void);
void min_sed(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int min_sed_scratch(int vec_len, int seg_len);
unsigned int min_sed_inplace(void);
void max_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int max_sez_scratch(int vec_len, int seg_len);
unsigned int max_sez_inplace(void);
void max_sed(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int max_sed_scratch(int vec_len, int seg_len);
unsigned int max_sed_inplace(void);
void and_seb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int and_seb_scratch(int vec_len, int seg_len);
unsigned int and_seb_inplace(void);
void and_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int and_sez_scratch(int vec_len, int seg_len);
unsigned int and_sez_inplace(void);
void ior_seb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int ior_seb_scratch(int vec_len, int seg_len);
unsigned int ior_seb_inplace(void);
void ior_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int ior_sez_scratch(int vec_len, int seg_len);
unsigned int ior_sez_inplace(void);
void xor_seb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int xor_seb_scratch(int vec_len, int seg_len);
unsigned int xor_seb_inplace(void);
void xor_sez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int xor_sez_scratch(int vec_len, int seg_len);
unsigned int xor_sez_inplace(void);

/* Unsegmented reductions (return a host scalar) */
int add_ruz(vec_p s, int len, vec_p scratch);
int add_ruz_scratch(int len);
unsigned int add_ruz_inplace(void);
double add_rud(vec_p s, int len, vec_p scratch);
int add_rud_scratch(int len);
unsigned int add_rud_inplace(void);
int mul_ruz(vec_p s, int len, vec_p scratch);
int mul_ruz_scratch(int len);
unsigned int mul_ruz_inplace(void);
double mul_rud(vec_p s, int len, vec_p scratch);
int mul_rud_scratch(int len);
unsigned int mul_rud_inplace(void);
int min_ruz(vec_p s, int len, vec_p scratch);
int min_ruz_scratch(int len);
unsigned int min_ruz_inplace(void);
double min_rud(vec_p s, int len, vec_p scratch);
int min_rud_scratch(int len);
unsigned int min_rud_inplace(void);
int max_ruz(vec_p s, int len, vec_p scratch);
int max_ruz_scratch(int len);
unsigned int max_ruz_inplace(void);
double max_rud(vec_p s, int len, vec_p scratch);
int max_rud_scratch(int len);
unsigned int max_rud_inplace(void);
cvl_bool and_rub(vec_p s, int len, vec_p scratch);
int and_rub_scratch(int len);
unsigned int and_rub_inplace(void);
int and_ruz(vec_p s, int len, vec_p scratch);
int and_ruz_scratch(int len);
unsigned int and_ruz_inplace(void);
cvl_bool ior_rub(vec_p s, int len, vec_p scratch);
int ior_rub_scratch(int len);
unsigned int ior_rub_inplace(void);
int ior_ruz(vec_p s, int len, vec_p scratch);
int ior_ruz_scratch(int len);
unsigned int ior_ruz_inplace(void);
cvl_bool xor_rub(vec_p s, int len, vec_p scratch);
int xor_rub_scratch(int len);
unsigned int xor_rub_inplace(void);
int xor_ruz(vec_p s, int len, vec_p scratch);
int xor_ruz_scratch(int len);
unsigned int xor_ruz_inplace(void);

/* Segmented reductions (one value per segment) */
void add_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int add_rez_scratch(int vec_len, int seg_len);
unsigned int add_rez_inplace(void);
void add_red(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int add_red_scratch(int vec_len, int seg_len);
unsigned int add_red_inplace(void);
void mul_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int mul_rez_scratch(int vec_len, int seg_len);
unsigned int mul_rez_inplace(void);
void mul_red(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int mul_red_scratch(int vec_len, int seg_len);
unsigned int mul_red_inplace(void);
void min_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int min_rez_scratch(int vec_len, int seg_len);
unsigned int min_rez_inplace(void);
void min_red(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int min_red_scratch(int vec_len, int seg_len);
unsigned int min_red_inplace(void);
void max_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int max_rez_scratch(int vec_len, int seg_len);
unsigned int max_rez_inplace(void);
void max_red(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int max_red_scratch(int vec_len, int seg_len);
unsigned int max_red_inplace(void);
void and_reb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int and_reb_scratch(int vec_len, int seg_len);
unsigned int and_reb_inplace(void);
void and_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int and_rez_scratch(int vec_len, int seg_len);
unsigned int and_rez_inplace(void);
void ior_reb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int ior_reb_scratch(int vec_len, int seg_len);
unsigned int ior_reb_inplace(void);
void ior_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int ior_rez_scratch(int vec_len, int seg_len);
unsigned int ior_rez_inplace(void);
void xor_reb(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int xor_reb_scratch(int vec_len, int seg_len);
unsigned int xor_reb_inplace(void);
void xor_rez(vec_p d, vec_p s, vec_p sd, int n, int m, vec_p scratch);
int xor_rez_scratch(int vec_len, int seg_len);
unsigned int xor_rez_inplace(void);

/* Extract, replace, distribute */
int ext_vuz(vec_p V, int i, int len, vec_p scratch);
int ext_vuz_scratch(int len);
unsigned int ext_vuz_inplace(void);
cvl_bool ext_vub(vec_p V, int i, int len, vec_p scratch);
int ext_vub_scratch(int len);
unsigned int ext_vub_inplace(void);
double ext_vud(vec_p V, int i, int len, vec_p scratch);
int ext_vud_scratch(int len);
unsigned int ext_vud_inplace(void);
void ext_vez(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int ext_vez_scratch(int vec_len, int seg_len);
unsigned int ext_vez_inplace(void);
void ext_veb(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int ext_veb_scratch(int vec_len, int seg_len);
unsigned int ext_veb_inplace(void);
void ext_ved(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int ext_ved_scratch(int vec_len, int seg_len);
unsigned int ext_ved_inplace(void);
void rep_vuz(vec_p V, int i, int val, int len, vec_p scratch);
int rep_vuz_scratch(int len);
unsigned int rep_vuz_inplace(void);
void rep_vub(vec_p V, int i, cvl_bool val, int len, vec_p scratch);
int rep_vub_scratch(int len);
unsigned int rep_vub_inplace(void);
void rep_vud(vec_p V, int i, double val, int len, vec_p scratch);
int rep_vud_scratch(int len);
unsigned int rep_vud_inplace(void);
void rep_vez(vec_p d, vec_p s, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int rep_vez_scratch(int vec_len, int seg_len);
unsigned int rep_vez_inplace(void);
void rep_veb(vec_p d, vec_p s, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int rep_veb_scratch(int vec_len, int seg_len);
unsigned int rep_veb_inplace(void);
void rep_ved(vec_p d, vec_p s, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int rep_ved_scratch(int vec_len, int seg_len);
unsigned int rep_ved_inplace(void);
void dis_vuz(vec_p d, int v, int len, vec_p scratch);
int dis_vuz_scratch(int len);
unsigned int dis_vuz_inplace(void);
void dis_vub(vec_p d, cvl_bool v, int len, vec_p scratch);
int dis_vub_scratch(int len);
unsigned int dis_vub_inplace(void);
void dis_vud(vec_p d, double v, int len, vec_p scratch);
int dis_vud_scratch(int len);
unsigned int dis_vud_inplace(void);
void dis_vez(vec_p d, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int dis_vez_scratch(int vec_len, int seg_len);
unsigned int dis_vez_inplace(void);
void dis_veb(vec_p d, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int dis_veb_scratch(int vec_len, int seg_len);
unsigned int dis_veb_inplace(void);
void dis_ved(vec_p d, vec_p v, vec_p sd, int n, int m, vec_p scratch);
int dis_ved_scratch(int vec_len, int seg_len);
unsigned int dis_ved_inplace(void);

/* Permutes: permute (d[i[k]] = s[k]), back-permute (d[k] = s[i[k]]), flagged and default variants */
void smp_puz(vec_p d, vec_p s, vec_p i, int len, vec_p scratch);
int smp_puz_scratch(int len);
unsigned int smp_puz_inplace(void);
void smp_pub(vec_p d, vec_p s, vec_p i, int len, vec_p scratch);
int smp_pub_scratch(int len);
unsigned int smp_pub_inplace(void);
void smp_pud(vec_p d, vec_p s, vec_p i, int len, vec_p scratch);
int smp_pud_scratch(int len);
unsigned int smp_pud_inplace(void);
void smp_pez(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int smp_pez_scratch(int vec_len, int seg_len);
unsigned int smp_pez_inplace(void);
void smp_peb(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int smp_peb_scratch(int vec_len, int seg_len);
unsigned int smp_peb_inplace(void);
void smp_ped(vec_p d, vec_p s, vec_p i, vec_p sd, int n, int m, vec_p scratch);
int smp_ped_scratch(int vec_len, int seg_len);
unsigned int smp_ped_inplace(void);
void bck_puz(vec_p d, vec_p s, vec_p i, int s_len, int d_len, vec_p scratch);
int bck_puz_scratch(int s_len, int d_len);
unsigned int bck_puz_inplace(void);
void bck_pub(vec_p d, vec_p s, vec_p i, int s_len, int d_len, vec_p scratch);
int bck_pub_scratch(int s_len, int d_len);
unsigned int bck_pub_inplace(void);
void bck_pud(vec_p d, vec_p s, vec_p i, int s_len, int d_len, vec_p scratch);
int bck_pud_scratch(int s_len, int d_len);
unsigned int bck_pud_inplace(void);
void bck_pez(vec_p d, vec_p s, vec_p i, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bck_pez_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bck_pez_inplace(void);
void bck_peb(vec_p d, vec_p s, vec_p i, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bck_peb_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bck_peb_inplace(void);
void bck_ped(vec_p d, vec_p s, vec_p i, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bck_ped_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bck_ped_inplace(void);
void fpm_puz(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int fpm_puz_scratch(int s_len, int d_len);
unsigned int fpm_puz_inplace(void);
void fpm_pub(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int fpm_pub_scratch(int s_len, int d_len);
unsigned int fpm_pub_inplace(void);
void fpm_pud(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int fpm_pud_scratch(int s_len, int d_len);
unsigned int fpm_pud_inplace(void);
void fpm_pez(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int fpm_pez_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int fpm_pez_inplace(void);
void fpm_peb(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int fpm_peb_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int fpm_peb_inplace(void);
void fpm_ped(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int fpm_ped_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int fpm_ped_inplace(void);
void bfp_puz(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int bfp_puz_scratch(int s_len, int d_len);
unsigned int bfp_puz_inplace(void);
void bfp_pub(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int bfp_pub_scratch(int s_len, int d_len);
unsigned int bfp_pub_inplace(void);
void bfp_pud(vec_p d, vec_p s, vec_p i, vec_p f, int len_src, int len_dest, vec_p scratch);
int bfp_pud_scratch(int s_len, int d_len);
unsigned int bfp_pud_inplace(void);
void bfp_pez(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bfp_pez_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bfp_pez_inplace(void);
void bfp_peb(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bfp_peb_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bfp_peb_inplace(void);
void bfp_ped(vec_p d, vec_p s, vec_p i, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int bfp_ped_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int bfp_ped_inplace(void);
void dpe_puz(vec_p d, vec_p s, vec_p i, vec_p v, int len_src, int len_dest, vec_p scratch);
int dpe_puz_scratch(int s_len, int d_len);
unsigned int dpe_puz_inplace(void);
void dpe_pub(vec_p d, vec_p s, vec_p i, vec_p v, int len_src, int len_dest, vec_p scratch);
int dpe_pub_scratch(int s_len, int d_len);
unsigned int dpe_pub_inplace(void);
void dpe_pud(vec_p d, vec_p s, vec_p i, vec_p v, int len_src, int len_dest, vec_p scratch);
int dpe_pud_scratch(int s_len, int d_len);
unsigned int dpe_pud_inplace(void);
void dpe_pez(vec_p d, vec_p s, vec_p i, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dpe_pez_scratch(int n_in, int m_in, int n_out, int m_out);
unsigned int dpe_pez_inplace(void);
void dpe_peb(vec_p d, vec_p s, vec_p i, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dpe_peb_scratch(int n_in, int m_in, int n_out, int m_out);
unsigned int dpe_peb_inplace(void);
void dpe_ped(vec_p d, vec_p s, vec_p i, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dpe_ped_scratch(int n_in, int m_in, int n_out, int m_out);
unsigned int dpe_ped_inplace(void);
void dfp_puz(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, int len_src, int len_dest, vec_p scratch);
int dfp_puz_scratch(int s_len, int d_len);
unsigned int dfp_puz_inplace(void);
void dfp_pub(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, int len_src, int len_dest, vec_p scratch);
int dfp_pub_scratch(int s_len, int d_len);
unsigned int dfp_pub_inplace(void);
void dfp_pud(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, int len_src, int len_dest, vec_p scratch);
int dfp_pud_scratch(int s_len, int d_len);
unsigned int dfp_pud_inplace(void);
void dfp_pez(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dfp_pez_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int dfp_pez_inplace(void);
void dfp_peb(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dfp_peb_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int dfp_peb_inplace(void);
void dfp_ped(vec_p d, vec_p s, vec_p i, vec_p f, vec_p v, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int dfp_ped_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int dfp_ped_inplace(void);

/* Facilities: timers, sizes (in CVL units), segment descriptors, memory, host copies */
void tgt_fos(cvl_timer_t *cvl_timer_p);
double tdf_fos(cvl_timer_t *t1, cvl_timer_t *t2);
int siz_fob(int length);
int siz_foz(int length);
int siz_fod(int length);
int siz_fos(int n, int m);
void mke_fov(vec_p sd, vec_p l, int n, int m, vec_p scratch);
int mke_fov_scratch(int n, int m);
unsigned int mke_fov_inplace(void);
void len_fos(vec_p l, vec_p sd, int n, int m, vec_p scratch);
int len_fos_scratch(int n, int m);
unsigned int len_fos_inplace(void);
vec_p alo_foz(int size);
void fre_fov(vec_p pointer);
void mov_fov(vec_p d, vec_p s, int size, vec_p scratch);
int mov_fov_scratch(int size);
vec_p add_fov(vec_p v, int a);
int sub_fov(vec_p v1, vec_p v2);
cvl_bool eql_fov(vec_p v1, vec_p v2);
int cmp_fov(vec_p v1, vec_p v2);
void v2c_fuz(int *d, vec_p s, int len, vec_p scratch);
int v2c_fuz_scratch(int len);
unsigned int v2c_fuz_inplace(void);
void v2c_fub(cvl_bool *d, vec_p s, int len, vec_p scratch);
int v2c_fub_scratch(int len);
unsigned int v2c_fub_inplace(void);
void v2c_fud(double *d, vec_p s, int len, vec_p scratch);
int v2c_fud_scratch(int len);
unsigned int v2c_fud_inplace(void);
void c2v_fuz(vec_p d, int *s, int len, vec_p scratch);
int c2v_fuz_scratch(int len);
unsigned int c2v_fuz_inplace(void);
void c2v_fub(vec_p d, cvl_bool *s, int len, vec_p scratch);
int c2v_fub_scratch(int len);
unsigned int c2v_fub_inplace(void);
void c2v_fud(vec_p d, double *s, int len, vec_p scratch);
int c2v_fud_scratch(int len);
unsigned int c2v_fud_inplace(void);

/* Library: index, pack */
void ind_luz(vec_p d, int init, int stride, int count, vec_p scratch);
int ind_luz_scratch(int len);
unsigned int ind_luz_inplace(void);
void ind_lez(vec_p d, vec_p init, vec_p stride, vec_p dest_segd, int vec_len, int seg_count, vec_p scratch);
int ind_lez_scratch(int vec_len, int seg_len);
unsigned int ind_lez_inplace(void);
int pk1_luv(vec_p f, int len, vec_p scratch);
int pk1_luv_scratch(int len);
unsigned int pk1_luv_inplace(void);
void pk2_luz(vec_p d, vec_p s, vec_p f, int src_len, int dest_len, vec_p scratch);
int pk2_luz_scratch(int len);
unsigned int pk2_luz_inplace(void);
void pk2_lub(vec_p d, vec_p s, vec_p f, int src_len, int dest_len, vec_p scratch);
int pk2_lub_scratch(int len);
unsigned int pk2_lub_inplace(void);
void pk2_lud(vec_p d, vec_p s, vec_p f, int src_len, int dest_len, vec_p scratch);
int pk2_lud_scratch(int len);
unsigned int pk2_lud_inplace(void);
void pk1_lev(vec_p ds, vec_p f, vec_p sd, int n, int m, vec_p scratch);
int pk1_lev_scratch(int vec_len, int seg_len);
unsigned int pk1_lev_inplace(void);
void pk2_lez(vec_p d, vec_p s, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int pk2_lez_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int pk2_lez_inplace(void);
void pk2_leb(vec_p d, vec_p s, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int pk2_leb_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int pk2_leb_inplace(void);
void pk2_led(vec_p d, vec_p s, vec_p f, vec_p sd_s, int n_s, int m_s, vec_p sd_d, int n_d, int m_d, vec_p scratch);
int pk2_led_scratch(int s_vec_len, int s_seg_len, int d_vec_len, int d_seg_len);
unsigned int pk2_led_inplace(void);

/* Rank: stable radix rank, ascending (rku) or descending (rkd) */
void rkd_led(vec_p d, vec_p s, vec_p segd, int vec_len, int seg_count, vec_p scratch);
int rkd_led_scratch(int vec_len, int seg_count);
unsigned int rkd_led_inplace(void);
void rku_led(vec_p d, vec_p s, vec_p segd, int vec_len, int seg_count, vec_p scratch);
int rku_led_scratch(int vec_len, int seg_count);
unsigned int rku_led_inplace(void);
void rkd_lez(vec_p d, vec_p s, vec_p segd, int vec_len, int seg_count, vec_p scratch);
int rkd_lez_scratch(int vec_len, int seg_count);
unsigned int rkd_lez_inplace(void);
void rku_lez(vec_p d, vec_p s, vec_p segd, int vec_len, int seg_count, vec_p scratch);
int rku_lez_scratch(int vec_len, int seg_count);
unsigned int rku_lez_inplace(void);
void rkd_lud(vec_p d, vec_p s, int len, vec_p scratch);
int rkd_lud_scratch(int len);
unsigned int rkd_lud_inplace(void);
void rku_lud(vec_p d, vec_p s, int len, vec_p scratch);
int rku_lud_scratch(int len);
unsigned int rku_lud_inplace(void);
void rkd_luz(vec_p d, vec_p s, int len, vec_p scratch);
int rkd_luz_scratch(int len);
unsigned int rkd_luz_inplace(void);
void rku_luz(vec_p d, vec_p s, int len, vec_p scratch);
int rku_luz_scratch(int len);
unsigned int rku_luz_inplace(void);

/* Random numbers */
void rnd_foz(int seed);
#ifdef __cplusplus
}
#endif

#endif /* cvl_h */
