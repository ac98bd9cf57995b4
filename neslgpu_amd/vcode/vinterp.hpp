// vinterp.hpp -- a VCODE interpreter over the CVL C API.
//
// VCODE is the stack-machine language the NESL compiler emits (the
// reference interpreter is nesl/vcode; its grammar is grammar.yy:42-472 and
// its lexer tokens.flex:27-107).  This interpreter reads VCODE (and the
// fused-op form .fcode), links it, and executes it by calling a CVL backend
// loaded at run time with dlopen -- by default libcvl_hip.so of this
// package, but any library exporting the CVL API of nesl/include/cvl.h
// works, which is how the serial reference CVL produces golden outputs for
// the same program.  The interpreter never touches vector memory itself:
// every vector lives in one arena obtained from alo_foz and is read and
// written only through CVL calls (v2c/c2v, ext/rep and the vector ops).
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace vcode {

// Element types; the order of the first entries follows the TYPE enum of
// nesl/vcode/vcode.h:12-14 because fused libraries store these values in
// their descriptor tables.
enum Type : int {
    T_Illegal = 0, T_Pair, T_None, T_Bool, T_Int, T_Float, T_Segdes, T_String,
    T_Agree1, T_Agree2, T_Agree3, T_Agree4, T_Given
};

const char *type_name(int t);

// Kinds of CVL call shapes, in the order of cvl_desc_t
// (nesl/vcode/vcode.h:58-71): Elwise0..Elwise40 then the vector ops.
enum Desc : int {
    D_Elwise0 = 0, D_Elwise40 = 40,
    D_Scan, D_Reduce, D_Special, D_SegOp,
    D_Permute, D_Dpermute, D_Fpermute, D_Bpermute, D_Bfpermute, D_Dfpermute,
    D_Extract, D_Replace, D_Dist, D_Index, D_Pack, D_NotSupported
};

// Length rules of results and argument constraints (nesl/vcode/vcode.h:36-52).
enum : unsigned {
    L_NONE = 0, L_AGREE1 = 0x1, L_AGREE2 = 0x2, L_AGREE3 = 0x4, L_AGREE4 = 0x8, L_AGREE5 = 0x10,
    L_COMPAT1 = 0x20, L_COMPAT2 = 0x40, L_COMPAT3 = 0x80, L_COMPAT4 = 0x100, L_COMPAT5 = 0x200,
    L_SCALAR = 0x400
};

enum Op : int {
    // elementwise
    OP_PLUS, OP_MINUS, OP_TIMES, OP_DIV, OP_MOD, OP_LT, OP_LEQ, OP_GT, OP_GEQ, OP_EQ, OP_NEQ,
    OP_LSHIFT, OP_RSHIFT, OP_NOT, OP_AND, OP_OR, OP_XOR, OP_SELECT, OP_RAND,
    OP_FLOOR, OP_CEIL, OP_TRUNC, OP_ROUND, OP_LOG, OP_SQRT, OP_EXP,
    OP_SIN, OP_COS, OP_TAN, OP_ASIN, OP_ACOS, OP_ATAN, OP_SINH, OP_COSH, OP_TANH,
    OP_I_TO_F, OP_I_TO_B, OP_B_TO_I,
    // vector ops
    OP_PLUS_SCAN, OP_MULT_SCAN, OP_MAX_SCAN, OP_MIN_SCAN, OP_AND_SCAN, OP_OR_SCAN, OP_XOR_SCAN,
    OP_PLUS_REDUCE, OP_MULT_REDUCE, OP_MAX_REDUCE, OP_MIN_REDUCE, OP_AND_REDUCE, OP_OR_REDUCE,
    OP_XOR_REDUCE,
    OP_PERMUTE, OP_DPERMUTE, OP_FPERMUTE, OP_BPERMUTE, OP_BFPERMUTE, OP_DFPERMUTE,
    OP_EXTRACT, OP_REPLACE, OP_DIST, OP_INDEX, OP_RANK_UP, OP_RANK_DOWN, OP_PACK,
    OP_LENGTH, OP_MAKE_SEGDES, OP_LENGTHS,
    // control
    OP_COPY, OP_POP, OP_CPOP, OP_PAIR, OP_UNPAIR, OP_CALL, OP_RET, OP_IF, OP_ELSE, OP_CONST,
    OP_EXIT, OP_START_TIMER, OP_STOP_TIMER,
    OP_READ, OP_WRITE, OP_FOPEN, OP_FCLOSE, OP_FREAD, OP_FREAD_CHAR, OP_FWRITE, OP_SPAWN, OP_SRAND,
    OP_FUSED,
    OP_COUNT
};

// Which type token, if any, follows an opcode in the source.
enum TypeSyntax { TS_NONE, TS_1 /*INT|FLOAT*/, TS_2 /*INT|BOOL*/, TS_3 /*INT|BOOL|FLOAT*/,
                  TS_5 /*INT|BOOL|FLOAT|SEGDES|CHAR*/, TS_INT /*INT only*/ };

// Static description of one opcode: arguments, result, CVL call shape and
// the CVL function family (the backend's function is base + type suffix,
// z/d/b for Int/Float/Bool).
struct OpInfo {
    Op op;
    const char *name;        // VCODE spelling
    TypeSyntax syntax;
    int fixed_type;          // instruction type when the syntax has none (T_Illegal: none)
    int nargs;
    int nres;
    int arg_type[6];         // T_Given = the instruction type
    int res_type[2];
    unsigned res_len[2];
    int desc;
    const char *cvl;         // CVL family prefix, e.g. "add_se"
    int scratch_args[2];     // 1-based args whose lengths size the scratch (0 = none)
};

const OpInfo &op_info(Op op);
bool lookup_keyword(const std::string &s, int *kind, int *value);

// Fused-op descriptor, binary compatible with vopdes_t of
// nesl/vcode/vcode.h:73-84 (MAX_ARGS = 40, MAX_OUT = 3), which fused
// libraries export as the array `vops`.
constexpr int kMaxArgs = 40;
constexpr int kMaxOut = 3;
struct VopDes {
    int vop;
    const char *vopname;
    int arg_num;
    int result_num;
    int arg_type[kMaxArgs];
    unsigned arg_len[kMaxArgs];
    int result_type[kMaxOut];
    unsigned result_len[kMaxOut];
    int arg_for_scratch[kMaxArgs];
    int cvl_desc;
};
struct CvlFunct {
    void (*function)();
    int (*scratch_function)();
    unsigned (*inplace_function)();
};
struct CvlTriple {
    CvlFunct cvl_int, cvl_float, cvl_boolean;
};

// One parsed instruction.
struct Instr {
    Op op = OP_COUNT;
    int type = T_Illegal;
    int line = 0;
    int a = 0, b = 0;            // COPY/POP/CPOP counts, FUSED number
    int branch = -1;             // CALL / IF / ELSE target
    std::string label;           // CALL target name
    // constants
    bool pooled = false;         // vector constant in the pool
    int pool_index = -1;
    long long ival = 0;
    double dval = 0.0;
};

struct ConstVec {
    int type;                    // T_Int, T_Float, T_Bool, T_Segdes (lengths), T_String
    std::vector<long long> ints;
    std::vector<double> dbls;
};

struct Program {
    std::vector<Instr> code;
    std::map<std::string, int> labels;
    std::vector<ConstVec> pool;
    std::string entry;           // first function, or MAIN
    bool uses_fused = false;
};

// Parse and link a VCODE source; returns false (message on stderr) on error.
// lex_trace prints every token, link_trace every resolved call
// (-d lex, -d link; nesl/vcode/main.c:194-214).
bool parse_program(const std::string &text, Program &prog, bool lex_trace = false, bool link_trace = false);

// The linked program, one instruction per line with its index (-d program).
void show_program(const Program &prog, FILE *out);

struct Options {
    std::string cvl_library;     // CVL backend shared library
    std::string fused_library;   // -l
    long long arena_units = 0;   // -m (0: default)
    bool time_run = false;       // -t
    bool trace = false;          // -d command: every instruction executed
    bool runtime_trace = false;  // -d runtime: the call stack
    bool stack_trace = false;    // -d stack: the vector stack after each instruction
    bool value_trace = false;    // -d values: and the first values of each entry
    bool heap_trace = false;     // -d heap: arena use after each instruction
    bool abort_on_error = false; // -d abort: abort() (core dump) instead of exit on an error
    bool garbage_notify = false; // -g: accepted; this arena never compacts
    bool graphs = false;         // -G: replay repeated basic blocks as HIP graphs
    std::string profile_path;    // -p: per-function times written here (.vprof)
};

int run(const Program &prog, const Options &opt);

}  // namespace vcode
