// main.cpp -- command line of the VCODE interpreter.
//
//   vinterp [-t] [-G] [-m units] [-l fused.so] [-B cvl.so] [-d command] [-P] file.vcode
//
// Options as in the reference interpreter (nesl/vcode/main.c:138-160):
// -t times the run, -m sizes the vector arena in CVL units (8 bytes),
// -l names the fused-operation library that FUSED n instructions call.
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
    std::fprintf(stderr, "vinterp [-t] [-G] [-P] [-m <units>] [-l <fused library>] [-B <CVL library>] [-d command] file-name\n");
    std::exit(1);
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
    bool parse_only = false;
    while ((c = getopt(argc, argv, "tm:l:B:d:cPG")) != -1) {
        switch (c) {
            case 't': opt.time_run = true; break;
            case 'm': opt.arena_units = std::atoll(optarg); break;
            case 'l': opt.fused_library = optarg; break;
            case 'B': opt.cvl_library = optarg; break;
            case 'd': if (std::strcmp(optarg, "command") == 0) opt.trace = true; break;
            case 'c': break;                       // arguments are always checked
            case 'P': parse_only = true; break;
            case 'G': opt.graphs = true; break;
            default: usage();
        }
    }
    if (optind + 1 != argc) usage();
    if (opt.cvl_library.empty()) opt.cvl_library = default_backend(argv[0]);
    std::ifstream in(argv[optind], std::ios::binary);
    if (!in) {
        std::fprintf(stderr, "vinterp: Error opening input file %s\n", argv[optind]);
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    vcode::Program prog;
    if (!vcode::parse_program(ss.str(), prog)) {
        std::fprintf(stderr, "Parse failure\n");
        return 1;
    }
    if (parse_only) {
        std::printf("%zu instructions, %zu functions, %zu constants\n", prog.code.size(), prog.labels.size(),
                    prog.pool.size());
        return 0;
    }
    return vcode::run(prog, opt);
}
