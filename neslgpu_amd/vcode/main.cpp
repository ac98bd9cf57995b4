// main.cpp -- command line of the VCODE interpreter.
//
//   vinterp [-cgptGP] [-m units] [-l fused.so] [-B cvl.so] [-d switch] file.vcode
//
// Options as in the reference interpreter (nesl/vcode/main.c:130-160):
// -t times the run, -m sizes the vector arena in CVL units (8 bytes),
// -l names the fused-operation library that FUSED n instructions call,
// -d lex|link|program|runtime|command|stack|values|heap|abort trace the
// lexer, the linker, the parsed program, the call stack, each instruction,
// the vector stack (with its first values), the arena's use, or make an
// error abort (core dump) instead of exit; -c is accepted (arguments are
// always checked) and so is -g (this arena never compacts).  -p writes the
// per-function profile the reference builds with VCODE_PROFILE
// (vcode-profile.c: calls, exclusive time, min/avg/max, cumulative) to
// <program>.vprof, as the reference names it.
// -B names the CVL backend (default: libcvl_hip.so next to this binary,
// or $VINTERP_CVL).  -P only parses and links the program and prints its
// instruction count.  -G records each basic block of device-only
// instructions and replays it as a HIP graph when the same block runs again
// on the same vectors (a loop body, an iterated kernel).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "vinterp.hpp"

static void usage() {
    std::fprintf(stderr, "vinterp [-cgptGP] [-m <units>] [-l <fused library>] [-B <CVL library>] [-d <debug-switch>] "
                         "file-name\n");
    std::fprintf(stderr, "\t -d abort\t dump core on internal error\n");
    std::fprintf(stderr, "\t -d lex\t\t lex trace\n");
    std::fprintf(stderr, "\t -d link\t link trace\n");
    std::fprintf(stderr, "\t -d program\t dump program parse\n");
    std::fprintf(stderr, "\t -d runtime\t trace call stack\n");
    std::fprintf(stderr, "\t -d command\t trace commands executed\n");
    std::fprintf(stderr, "\t -d stack\t trace vector stack\n");
    std::fprintf(stderr, "\t -d values\t trace vector stack values\n");
    std::fprintf(stderr, "\t -d heap\t trace vector arena use\n");
    std::fprintf(stderr, "\t -p\t\t per-function profile to <file-name>.vprof\n");
    std::exit(1);
}

// <program>.vprof, the reference's name for the profile (vcode-profile.c:93-101)
static std::string vprof_name(const std::string &src) {
    std::string out = src;
    const size_t dot = out.rfind('.');
    if (dot != std::string::npos) {
        const std::string ext = out.substr(dot);
        if (ext == ".vcode" || ext == ".ucode" || ext == ".fcode") out.erase(dot);
    }
    return out + ".vprof";
}

static std::string default_backend(const char *argv0) {
    if (const char *e = std::getenv("VINTERP_CVL")) return e;
    char buf[4096];
    ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    std::string dir = ".";
    if (n > 0) {
        buf[n] = 0;
        dir = buf;
        dir = dir.substr(0, dir.rfind('/'));
    } else if (std::strchr(argv0, '/')) {
        dir = std::string(argv0).substr(0, std::string(argv0).rfind('/'));
    }
    return dir + "/../libcvl_hip.so";
}

int main(int argc, char **argv) {
    vcode::Options opt;
    int c;
    bool parse_only = false, lex_trace = false, link_trace = false, program_dump = false;
    bool profile = false;
    while ((c = getopt(argc, argv, "tm:l:B:d:cgPGp")) != -1) {
        switch (c) {
            case 't': opt.time_run = true; break;
            case 'm': opt.arena_units = std::atoll(optarg); break;
            case 'l': opt.fused_library = optarg; break;
            case 'B': opt.cvl_library = optarg; break;
            case 'd': {
                const std::string d = optarg;
                if (d == "command") opt.trace = true;
                else if (d == "lex") lex_trace = true;
                else if (d == "link") link_trace = true;
                else if (d == "program") program_dump = true;
                else if (d == "runtime") opt.runtime_trace = true;
                else if (d == "stack") opt.stack_trace = true;
                else if (d == "values") opt.value_trace = true;
                else if (d == "heap") opt.heap_trace = true;
                else if (d == "abort") opt.abort_on_error = true;
                else usage();
                break;
            }
            case 'c': break;                       // arguments are always checked
            case 'g': opt.garbage_notify = true; break;
            case 'P': parse_only = true; break;
            case 'G': opt.graphs = true; break;
            case 'p': profile = true; break;
            default: usage();
        }
    }
    if (optind + 1 != argc) usage();
    if (profile) opt.profile_path = vprof_name(argv[optind]);
    if (opt.cvl_library.empty()) opt.cvl_library = default_backend(argv[0]);
    std::ifstream in(argv[optind], std::ios::binary);
    if (!in) {
        std::fprintf(stderr, "vinterp: Error opening input file %s\n", argv[optind]);
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    vcode::Program prog;
    if (!vcode::parse_program(ss.str(), prog, lex_trace, link_trace)) {
        std::fprintf(stderr, "Parse failure\n");
        return 1;
    }
    if (program_dump) vcode::show_program(prog, stderr);
    if (parse_only) {
        std::printf("%zu instructions, %zu functions, %zu constants\n", prog.code.size(), prog.labels.size(),
                    prog.pool.size());
        return 0;
    }
    return vcode::run(prog, opt);
}
