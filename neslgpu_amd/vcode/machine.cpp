// machine.cpp -- the VCODE stack machine.
//
// Semantics follow the reference interpreter: stack manipulation
// (COPY / POP / CPOP / PAIR / UNPAIR, nesl/vcode/stack.c:188-388), the
// dispatch of vector instructions to CVL calls with the argument order of
// each call shape (nesl/vcode/main.c:414-1895), result allocation with
// in-place reuse of a dying argument when the CVL function allows it
// (nesl/vcode/check_args.c:138-249), scratch sizing from the operation's
// *_scratch function (check_args.c:264-311), MAKE_SEGDES / LENGTHS /
// LENGTH / PACK as host-driven sequences (main.c:473-657), constants
// (nesl/vcode/constant.c:183-247), and the text I/O of nesl/vcode/io.c
// (write_value: "( 1 2 )", " %.14e" for floats, "[ lengths ]" for segment
// descriptors; WRITE goes to stderr, the final stack to stdout).
//
// Differences from the reference, on purpose:
//   - vectors are reference counted objects over a best-fit arena instead of
//     the reference's compacting vector stack; the arena comes from one
//     alo_foz call, and scratch is carved from it per instruction;
//   - argument types and lengths are always checked before a CVL call (the
//     reference checks only under -c), because a mismatched length handed
//     to a device kernel would read or write out of bounds;
//   - BPERMUTE passes the destination segment count as m_d (the reference
//     passes the source count, main.c:1750; the two are equal in programs
//     the NESL compiler emits);
//   - DPERMUTE / DFPERMUTE always write a fresh result (the reference's
//     storage shortcut for DFPERMUTE reuses the wrong argument,
//     main.c:1800-1805).
#include "vinterp.hpp"

#include <dlfcn.h>
#include <cstddef>

#include "../../include/vcode_fused.h"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <functional>
#include <limits>
#include <set>
#include <utility>

namespace vcode {

// the interpreter's view of a fused library's tables is the public ABI
static_assert(sizeof(VopDes) == sizeof(vopdes_t), "vopdes_t layout");
static_assert(offsetof(VopDes, arg_type) == offsetof(vopdes_t, arg_type), "vopdes_t layout");
static_assert(offsetof(VopDes, result_len) == offsetof(vopdes_t, result_len), "vopdes_t layout");
static_assert(offsetof(VopDes, cvl_desc) == offsetof(vopdes_t, cvl_desc), "vopdes_t layout");
static_assert(sizeof(CvlTriple) == sizeof(cvl_triple_t), "cvl_triple_t layout");
static_assert(T_Segdes == (int)Segdes && T_Given == (int)Given, "TYPE values");
static_assert(D_Scan == (int)Scan && D_NotSupported == (int)NotSupported, "cvl_desc_t values");

static bool g_abort_on_error = false;      // -d abort

[[noreturn]] static void die(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "vinterp: ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
    std::fflush(stdout);
    if (g_abort_on_error) std::abort();
    std::exit(1);
}

// ------------------------------------------------------------------ backend
class Backend {
  public:
    void open(const std::string &path) {
        h_ = dlopen(path.c_str(), RTLD_NOW | RTLD_GLOBAL);
        if (!h_) die("cannot load CVL library %s: %s", path.c_str(), dlerror());
#define NEED(name) name = reinterpret_cast<decltype(name)>(need(#name))
        NEED(alo_foz); NEED(siz_foz); NEED(siz_fob); NEED(siz_fod); NEED(siz_fos);
        NEED(add_ruz); NEED(add_ruz_scratch);
        NEED(mke_fov); NEED(mke_fov_scratch); NEED(len_fos); NEED(len_fos_scratch);
        NEED(pk1_lev); NEED(pk1_lev_scratch);
        NEED(rep_vuz); NEED(rep_vub); NEED(rep_vud);
        NEED(rep_vuz_scratch); NEED(rep_vub_scratch); NEED(rep_vud_scratch);
        NEED(ext_vuz); NEED(ext_vub); NEED(ext_vud);
        NEED(ext_vuz_scratch); NEED(ext_vub_scratch); NEED(ext_vud_scratch);
        NEED(v2c_fuz); NEED(v2c_fub); NEED(v2c_fud);
        NEED(c2v_fuz); NEED(c2v_fub); NEED(c2v_fud);
        NEED(v2c_fuz_scratch); NEED(v2c_fub_scratch); NEED(v2c_fud_scratch);
        NEED(c2v_fuz_scratch); NEED(c2v_fub_scratch); NEED(c2v_fud_scratch);
        NEED(cpy_wuz); NEED(cpy_wub); NEED(cpy_wud);
        NEED(cpy_wuz_scratch); NEED(cpy_wub_scratch); NEED(cpy_wud_scratch);
        NEED(tgt_fos); NEED(tdf_fos); NEED(rnd_foz);
#undef NEED
        CVL_init = reinterpret_cast<decltype(CVL_init)>(dlsym(h_, "CVL_init"));
    }

    void *sym(const std::string &name) { return h_ ? dlsym(h_, name.c_str()) : nullptr; }
    void *need(const char *name) {
        void *p = sym(name);
        if (!p) die("CVL library lacks %s", name);
        return p;
    }

    vec_p (*alo_foz)(int);
    int (*siz_foz)(int);
    int (*siz_fob)(int);
    int (*siz_fod)(int);
    int (*siz_fos)(int, int);
    int (*add_ruz)(vec_p, int, vec_p);
    int (*add_ruz_scratch)(int);
    void (*mke_fov)(vec_p, vec_p, int, int, vec_p);
    int (*mke_fov_scratch)(int, int);
    void (*len_fos)(vec_p, vec_p, int, int, vec_p);
    int (*len_fos_scratch)(int, int);
    void (*pk1_lev)(vec_p, vec_p, vec_p, int, int, vec_p);
    int (*pk1_lev_scratch)(int, int);
    void (*rep_vuz)(vec_p, int, int, int, vec_p);
    void (*rep_vub)(vec_p, int, cvl_bool, int, vec_p);
    void (*rep_vud)(vec_p, int, double, int, vec_p);
    int (*rep_vuz_scratch)(int);
    int (*rep_vub_scratch)(int);
    int (*rep_vud_scratch)(int);
    int (*ext_vuz)(vec_p, int, int, vec_p);
    cvl_bool (*ext_vub)(vec_p, int, int, vec_p);
    double (*ext_vud)(vec_p, int, int, vec_p);
    int (*ext_vuz_scratch)(int);
    int (*ext_vub_scratch)(int);
    int (*ext_vud_scratch)(int);
    void (*v2c_fuz)(int *, vec_p, int, vec_p);
    void (*v2c_fub)(cvl_bool *, vec_p, int, vec_p);
    void (*v2c_fud)(double *, vec_p, int, vec_p);
    void (*c2v_fuz)(vec_p, int *, int, vec_p);
    void (*c2v_fub)(vec_p, cvl_bool *, int, vec_p);
    void (*c2v_fud)(vec_p, double *, int, vec_p);
    int (*v2c_fuz_scratch)(int);
    int (*v2c_fub_scratch)(int);
    int (*v2c_fud_scratch)(int);
    int (*c2v_fuz_scratch)(int);
    int (*c2v_fub_scratch)(int);
    int (*c2v_fud_scratch)(int);
    void (*cpy_wuz)(vec_p, vec_p, int, vec_p);
    void (*cpy_wub)(vec_p, vec_p, int, vec_p);
    void (*cpy_wud)(vec_p, vec_p, int, vec_p);
    int (*cpy_wuz_scratch)(int);
    int (*cpy_wub_scratch)(int);
    int (*cpy_wud_scratch)(int);
    void (*tgt_fos)(void *);
    double (*tdf_fos)(void *, void *);
    void (*rnd_foz)(int);
    void (*CVL_init)(char *) = nullptr;

  private:
    void *h_ = nullptr;
};

// ------------------------------------------------------------------ arena
// Best-fit allocator over one alo_foz region, in CVL units (8 bytes), with
// every block a multiple of 32 units so vectors start 256-byte aligned.
class Arena {
  public:
    static constexpr size_t kGrain = 32;

    void init(char *base, size_t units) {
        base_ = base;
        total_ = units / kGrain * kGrain;
        if (total_) insert(0, total_);
    }

    // returns the unit offset, or SIZE_MAX when no block is large enough
    size_t alloc(size_t units) {
        units = (units + kGrain - 1) / kGrain * kGrain;
        if (units == 0) units = kGrain;
        auto it = by_size_.lower_bound({units, 0});
        if (it == by_size_.end()) return SIZE_MAX;
        size_t len = it->first, off = it->second;
        erase(off, len);
        if (len > units) insert(off + units, len - units);
        used_ += units;
        if (used_ > peak_) peak_ = used_;
        return off;
    }

    void release(size_t off, size_t units) {
        units = (units + kGrain - 1) / kGrain * kGrain;
        if (units == 0) units = kGrain;
        used_ -= units;
        auto next = by_off_.lower_bound(off);
        if (next != by_off_.end() && next->first == off + units) {
            size_t nl = next->second;
            erase(next->first, nl);
            units += nl;
        }
        auto prev = by_off_.lower_bound(off);
        if (prev != by_off_.begin()) {
            --prev;
            if (prev->first + prev->second == off) {
                size_t po = prev->first, pl = prev->second;
                erase(po, pl);
                off = po;
                units += pl;
            }
        }
        insert(off, units);
    }

    vec_p addr(size_t off) const { return base_ + off * 8; }
    size_t peak() const { return peak_; }
    size_t total() const { return total_; }
    size_t used() const { return used_; }
    size_t free_blocks() const { return by_off_.size(); }
    size_t largest_free() const { return by_size_.empty() ? 0 : by_size_.rbegin()->first; }

  private:
    void insert(size_t off, size_t len) {
        by_off_[off] = len;
        by_size_.insert({len, off});
    }
    void erase(size_t off, size_t len) {
        by_off_.erase(off);
        by_size_.erase({len, off});
    }
    char *base_ = nullptr;
    size_t total_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> by_off_;
    std::set<std::pair<size_t, size_t>> by_size_;
};

// ------------------------------------------------------------------ values
struct Vec {
    int type = T_Int;
    int len = 0;          // elements; for a segment descriptor, the segment count
    int seg_len = -1;     // for a segment descriptor, the element count it describes
    size_t off = 0, units = 0;
    vec_p p = nullptr;
    int refs = 1;
    Vec *first = nullptr, *second = nullptr;   // pairs
    Vec *owner = nullptr;   // a view into another vector's memory (LENGTHS of a descriptor)
};

// The call signature of one instruction: from the opcode table and the
// backend, or from a fused library's descriptor.
struct Sig {
    int nargs = 0;
    int arg_type[kMaxArgs] = {0};
    unsigned arg_len[kMaxArgs] = {0};
    int res_type = T_Illegal;
    unsigned res_len = 0;
    int desc = D_NotSupported;
    int scratch_args[2] = {0, 0};
    void *fn = nullptr;
    void *scratch_fn = nullptr;
    unsigned (*inplace_fn)() = nullptr;
    std::string what;
    bool resolved = false;
};

// ------------------------------------------------------------------ calls
template <size_t... I>
static void call_elwise_k(void *f, vec_p d, vec_p *s, int len, vec_p scr, std::index_sequence<I...>) {
    using Fn = void (*)(vec_p, decltype((void)I, vec_p())..., int, vec_p);
    reinterpret_cast<Fn>(f)(d, s[I]..., len, scr);
}
template <size_t K>
static void call_elwise_fixed(void *f, vec_p d, vec_p *s, int len, vec_p scr) {
    call_elwise_k(f, d, s, len, scr, std::make_index_sequence<K>{});
}
typedef void (*ElwiseCaller)(void *, vec_p, vec_p *, int, vec_p);
template <size_t... K>
static constexpr std::array<ElwiseCaller, sizeof...(K)> elwise_table(std::index_sequence<K...>) {
    return {{&call_elwise_fixed<K>...}};
}
static const auto kElwise = elwise_table(std::make_index_sequence<kMaxArgs + 1>{});

class Machine {
  public:
    Machine(const Program &prog, const Options &opt) : prog_(prog), opt_(opt), sigs_(prog.code.size()) {}

    int run() {
        g_abort_on_error = opt_.abort_on_error;
        be_.open(opt_.cvl_library);
        if (be_.CVL_init) {
            static char empty[1] = {0};
            be_.CVL_init(empty);
        }
        long long units = opt_.arena_units > 0 ? opt_.arena_units : (1LL << 28);
        if (units > std::numeric_limits<int>::max()) units = std::numeric_limits<int>::max();
        vec_p base = be_.alo_foz((int)units);
        if (!base) die("cannot allocate a vector arena of %lld units", units);
        arena_.init(static_cast<char *>(base), (size_t)units);
        if (prog_.uses_fused) load_fused();
        if (opt_.graphs) setup_graphs();
        // backends whose descriptors hold the m lengths contiguously say where
        auto lens_off = reinterpret_cast<int (*)()>(be_.sym("cvlhip_segd_lengths_offset"));
        lengths_offset_ = lens_off ? lens_off() : -1;
        install_constants();

        auto it = prog_.labels.find(prog_.entry);
        if (it == prog_.labels.end()) die("start function %s not found", prog_.entry.c_str());
        alignas(64) unsigned char t0[64], t1[64];
        if (opt_.time_run) be_.tgt_fos(t0);
        if (!opt_.profile_path.empty()) profile_init();
        execute(it->second);
        if (!opt_.profile_path.empty()) profile_report();
        if (opt_.time_run) {
            be_.tgt_fos(t1);
            std::fprintf(stderr, "elapsed time = %f seconds\n", be_.tdf_fos(t1, t0));
            std::fprintf(stderr, "instructions executed = %llu, peak vector memory = %zu units\n",
                         executed_, arena_.peak());
            if (graphs_)
                std::fprintf(stderr, "graphs: %llu basic blocks captured, %llu replayed\n", graph_captures_,
                             graph_replays_);
        }
        for (Vec *v : st_) {
            if (v->type == T_Pair)
                std::fprintf(stdout, "pair(%s,%s) \n", type_name(v->first->type), type_name(v->second->type));
            else
                write_value(v, stdout);
        }
        std::fflush(stdout);
        return 0;
    }

  private:
    // ---------------------------------------------------------- device calls
    // Every CVL call that queues device work without returning host data goes
    // through dev(): executed at once, or, while a basic block is being
    // recorded for a HIP graph, kept (with its arguments as the graph key).
    struct Recorder {
        bool on = false;
        std::vector<std::function<void()>> calls;
        std::vector<uint64_t> key;
    };
    struct GraphEntry {
        std::vector<uint64_t> key;
        void *exec = nullptr;
    };

    void key_push(const void *p) { rec_.key.push_back((uint64_t)(uintptr_t)p); }
    void key_push(int v) { rec_.key.push_back((uint64_t)(uint32_t)v); }
    void key_push(double v) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        rec_.key.push_back(u);
    }

    template <class F, class... A>
    void dev(F f, A... a) {
        if (!rec_.on) {
            f(a...);
            return;
        }
        key_push(reinterpret_cast<const void *>(f));
        (key_push(a), ...);
        rec_.calls.push_back([=]() { f(a...); });
    }

    void setup_graphs() {
        using StreamCreate = void *(*)();
        using SetStream = void (*)(void *);
        auto create = reinterpret_cast<StreamCreate>(be_.sym("cvlhip_stream_create"));
        auto set = reinterpret_cast<SetStream>(be_.sym("cvlhip_set_stream"));
        cap_begin_ = reinterpret_cast<int (*)()>(be_.sym("cvlhip_capture_begin"));
        cap_end_ = reinterpret_cast<int (*)(void **)>(be_.sym("cvlhip_capture_end"));
        graph_launch_ = reinterpret_cast<int (*)(void *)>(be_.sym("cvlhip_graph_launch"));
        graph_destroy_ = reinterpret_cast<void (*)(void *)>(be_.sym("cvlhip_graph_destroy"));
        if (!create || !set || !cap_begin_ || !cap_end_ || !graph_launch_ || !graph_destroy_) {
            std::fprintf(stderr, "vinterp: the CVL library cannot capture graphs; -G ignored\n");
            return;
        }
        set(create());
        graphs_ = true;
    }

    // Which instructions a graph may hold: stack manipulation, constants and
    // vector operations that only queue device work.  Control flow, anything
    // that returns data to the host (MAKE_SEGDES and PACK need a length, IF a
    // flag, I/O), RAND (it advances a host-side generator state) and rank
    // (it inspects key histograms on the host) end a block.
    bool capturable(int pc, const Instr &in) {
        switch (in.op) {
            case OP_COPY: case OP_POP: case OP_CPOP: case OP_PAIR: case OP_UNPAIR: case OP_CONST:
            case OP_LENGTH: case OP_LENGTHS:
                return true;
            case OP_RAND: case OP_RANK_UP: case OP_RANK_DOWN:
                return false;
            case OP_FUSED: {
                const Sig &s = signature(pc, in);
                return s.desc >= D_Elwise0 && s.desc <= D_Elwise40;
            }
            default: break;
        }
        const int d = op_info(in.op).desc;
        return (d >= D_Elwise0 && d <= D_Elwise40) || d == D_Scan || d == D_Reduce || d == D_Permute ||
               d == D_Dpermute || d == D_Fpermute || d == D_Bpermute || d == D_Bfpermute || d == D_Dfpermute ||
               d == D_Extract || d == D_Replace || d == D_Dist || d == D_Index;
    }

    void run_recorded() {
        for (auto &c : rec_.calls) c();
    }

    // end of a recorded block: replay its graph, or capture one
    void flush() {
        rec_.on = false;
        if (rec_.calls.size() < 2) {
            run_recorded();
            return;
        }
        uint64_t h = 1469598103934665603ull;
        for (uint64_t w : rec_.key) h = (h ^ w) * 1099511628211ull;
        auto it = graphs_cache_.find(h);
        if (it != graphs_cache_.end() && it->second.key == rec_.key) {
            if (graph_launch_(it->second.exec) != 0) die("graph launch failed");
            ++graph_replays_;
            return;
        }
        if (cap_begin_() != 0) {
            run_recorded();
            return;
        }
        run_recorded();
        void *exec = nullptr;
        if (cap_end_(&exec) != 0) die("capturing a basic block failed");
        if (graphs_cache_.size() >= 512) {
            graph_destroy_(graphs_cache_.begin()->second.exec);
            graphs_cache_.erase(graphs_cache_.begin());
        }
        if (it != graphs_cache_.end()) {          // hash collision: keep the newer block
            graph_destroy_(it->second.exec);
            graphs_cache_.erase(it);
        }
        graphs_cache_[h] = GraphEntry{rec_.key, exec};
        if (graph_launch_(exec) != 0) die("graph launch failed");
        ++graph_captures_;
    }

    // ---------------------------------------------------------- memory
    int units_of(int type, int len, int seg_len) {
        switch (type) {
            case T_Int: case T_String: return be_.siz_foz(len);
            case T_Bool: return be_.siz_fob(len);
            case T_Float: return be_.siz_fod(len);
            case T_Segdes: return be_.siz_fos(seg_len, len);
            default: die("cannot allocate a vector of type %s", type_name(type));
        }
    }

    Vec *alloc(int type, int len, int seg_len = -1) {
        if (len < 0 || (type == T_Segdes && seg_len < 0)) die("negative vector length (%d, %d)", len, seg_len);
        Vec *v = new Vec;
        v->type = type == T_String ? T_Int : type;
        v->len = len;
        v->seg_len = type == T_Segdes ? seg_len : -1;
        v->units = (size_t)std::max(units_of(v->type, len, seg_len), 1);
        v->off = arena_.alloc(v->units);
        if (v->off == SIZE_MAX)
            die("out of vector memory (%zu units wanted, arena %zu units; raise -m)", v->units, arena_.total());
        v->p = arena_.addr(v->off);
        return v;
    }

    void unref(Vec *v) {
        if (--v->refs > 0) return;
        if (v->type == T_Pair) {
            unref(v->first);
            unref(v->second);
        } else if (v->owner) {
            unref(v->owner);
        } else {
            arena_.release(v->off, v->units);
        }
        delete v;
    }

    // scratch of `units` CVL units, released when the guard dies
    struct Scratch {
        Scratch(Machine &m, int units) : m_(m) {
            units_ = (size_t)std::max(units, 1);
            off_ = m.arena_.alloc(units_);
            if (off_ == SIZE_MAX) die("out of vector memory for %d units of scratch (raise -m)", units);
            p = m.arena_.addr(off_);
        }
        ~Scratch() { m_.arena_.release(off_, units_); }
        Machine &m_;
        size_t off_, units_;
        vec_p p;
    };

    // ---------------------------------------------------------- stack
    void push(Vec *v) { st_.push_back(v); }
    Vec *top() {
        if (st_.empty()) die("line %d: stack empty", line_);
        return st_.back();
    }
    void pop_top() {
        Vec *v = top();
        st_.pop_back();
        unref(v);
    }
    void need_stack(int n, const char *what) {
        if ((int)st_.size() < n)
            die("line %d: only %zu values on stack; %s requires %d", line_, st_.size(), what, n);
    }

    void push_int_scalar(int v) {
        Vec *r = alloc(T_Int, 1);
        Scratch s(*this, be_.rep_vuz_scratch(1));
        dev(be_.rep_vuz, r->p, 0, v, 1, s.p);
        push(r);
    }
    void push_bool_scalar(int v) {
        Vec *r = alloc(T_Bool, 1);
        Scratch s(*this, be_.rep_vub_scratch(1));
        dev(be_.rep_vub, r->p, 0, v ? 1 : 0, 1, s.p);
        push(r);
    }
    void push_float_scalar(double v) {
        Vec *r = alloc(T_Float, 1);
        Scratch s(*this, be_.rep_vud_scratch(1));
        dev(be_.rep_vud, r->p, 0, v, 1, s.p);
        push(r);
    }
    int int_scalar(Vec *v, const char *what) {
        if (v->type != T_Int || v->len != 1) die("line %d: %s: expected an integer scalar", line_, what);
        Scratch s(*this, be_.ext_vuz_scratch(1));
        return be_.ext_vuz(v->p, 0, 1, s.p);
    }

    // ---------------------------------------------------------- constants
    void install_constants() {
        for (const ConstVec &c : prog_.pool) {
            Vec *v = nullptr;
            const int n = (int)(c.type == T_Float ? c.dbls.size() : c.ints.size());
            if (c.type == T_Float) {
                v = alloc(T_Float, n);
                std::vector<double> h(c.dbls);
                Scratch s(*this, be_.c2v_fud_scratch(n));
                be_.c2v_fud(v->p, h.data(), n, s.p);
            } else if (c.type == T_Bool) {
                v = alloc(T_Bool, n);
                std::vector<cvl_bool> h(c.ints.begin(), c.ints.end());
                Scratch s(*this, be_.c2v_fub_scratch(n));
                be_.c2v_fub(v->p, h.data(), n, s.p);
            } else {
                v = alloc(T_Int, n);
                std::vector<int> h(c.ints.begin(), c.ints.end());
                {
                    Scratch s(*this, be_.c2v_fuz_scratch(n));
                    be_.c2v_fuz(v->p, h.data(), n, s.p);
                }
                if (c.type == T_Segdes) {
                    Vec *sd = make_segdes(v);
                    unref(v);
                    v = sd;
                }
            }
            pool_.push_back(v);               // the pool keeps one reference for good
        }
    }

    // segment descriptor from an integer lengths vector (MAKE_SEGDES)
    Vec *make_segdes(Vec *lens) {
        int n;
        {
            Scratch s(*this, be_.add_ruz_scratch(lens->len));
            n = be_.add_ruz(lens->p, lens->len, s.p);
        }
        if (n < 0) die("line %d: MAKE_SEGDES: negative total length %d", line_, n);
        Vec *sd = alloc(T_Segdes, lens->len, n);
        Scratch s(*this, be_.mke_fov_scratch(n, lens->len));
        be_.mke_fov(sd->p, lens->p, n, lens->len, s.p);
        return sd;
    }

    // ---------------------------------------------------------- fused library
    void load_fused() {
        if (opt_.fused_library.empty())
            die("FUSED calls found in VCODE, but no fused library was provided with -l");
        void *h = dlopen(opt_.fused_library.c_str(), RTLD_NOW);
        if (!h) die("unable to load fused library %s: %s", opt_.fused_library.c_str(), dlerror());
        vops_ = static_cast<const VopDes *>(dlsym(h, "vops"));
        funs_ = static_cast<const CvlTriple *>(dlsym(h, "cvl_funs"));
        const int *count = static_cast<const int *>(dlsym(h, "vops_count"));
        if (!vops_ || !funs_) die("fused library %s lacks vops or cvl_funs", opt_.fused_library.c_str());
        nfused_ = count ? *count : -1;
        auto init = reinterpret_cast<void (*)(void *)>(dlsym(h, "init"));
        if (init) init(nullptr);
    }

    // ---------------------------------------------------------- signatures
    static char suffix(int type) { return type == T_Float ? 'd' : type == T_Bool ? 'b' : 'z'; }

    Sig &signature(int pc, const Instr &in) {
        Sig &s = sigs_[pc];
        if (s.resolved) return s;
        if (in.op == OP_FUSED) {
            if (nfused_ >= 0 && in.a >= nfused_) die("line %d: FUSED %d not in the fused library", in.line, in.a);
            const VopDes &d = vops_[in.a];
            if (d.arg_num < 0 || d.arg_num > kMaxArgs || d.result_num != 1)
                die("line %d: fused op %d has an unsupported descriptor", in.line, in.a);
            s.nargs = d.arg_num;
            for (int k = 0; k < d.arg_num; ++k) {
                s.arg_type[k] = d.arg_type[k];
                s.arg_len[k] = d.arg_len[k];
            }
            s.res_type = d.result_type[0];
            s.res_len = d.result_len[0];
            s.desc = d.cvl_desc;
            s.scratch_args[0] = d.arg_for_scratch[0] > 0 ? d.arg_for_scratch[0] : 0;
            s.scratch_args[1] = s.scratch_args[0] && d.arg_for_scratch[1] > 0 ? d.arg_for_scratch[1] : 0;
            s.fn = reinterpret_cast<void *>(funs_[in.a].cvl_int.function);
            s.scratch_fn = reinterpret_cast<void *>(funs_[in.a].cvl_int.scratch_function);
            s.inplace_fn = funs_[in.a].cvl_int.inplace_function;
            s.what = d.vopname ? d.vopname : "fused";
        } else {
            const OpInfo &oi = op_info(in.op);
            s.nargs = oi.nargs;
            for (int k = 0; k < oi.nargs; ++k) s.arg_type[k] = oi.arg_type[k];
            length_rules(in.op, s.arg_len);
            s.res_type = oi.res_type[0];
            s.res_len = oi.res_len[0];
            s.desc = oi.desc;
            s.scratch_args[0] = oi.scratch_args[0];
            s.scratch_args[1] = oi.scratch_args[1];
            std::string name = std::string(oi.cvl) + suffix(in.type);
            s.what = std::string(oi.name) + " " + type_name(in.type);
            s.fn = be_.sym(name);
            s.scratch_fn = be_.sym(name + "_scratch");
            s.inplace_fn = reinterpret_cast<unsigned (*)()>(be_.sym(name + "_inplace"));
            if (!s.fn || !s.scratch_fn || !s.inplace_fn)
                die("line %d: operation %s not supported by the CVL library (%s)", in.line, s.what.c_str(),
                    name.c_str());
        }
        s.resolved = true;
        return s;
    }

    // Argument length constraints: AGREEk = same length as argument k
    // (a segment descriptor's length is its segment count), COMPATk = a
    // segment descriptor describing argument k.
    static void length_rules(Op op, unsigned *r) {
        const OpInfo &oi = op_info(op);
        if (oi.desc >= D_Elwise0 && oi.desc <= D_Elwise40) {
            for (int k = 1; k < oi.nargs; ++k) r[k] = L_AGREE1;
            return;
        }
        switch (oi.desc) {
            case D_Scan: case D_Reduce: r[1] = L_COMPAT1; break;
            case D_Permute: r[1] = L_AGREE1; r[2] = L_COMPAT1; break;
            case D_Dpermute: r[1] = L_AGREE1; r[3] = L_COMPAT1; r[4] = L_COMPAT3 | L_AGREE4; break;
            case D_Fpermute: r[1] = L_AGREE1; r[2] = L_AGREE1; r[3] = L_COMPAT1; r[4] = L_AGREE4; break;
            case D_Bpermute: r[1] = 0; r[2] = L_COMPAT1; r[3] = L_COMPAT2 | L_AGREE3; break;
            case D_Bfpermute: r[2] = L_AGREE2; r[3] = L_COMPAT1; r[4] = L_COMPAT2 | L_AGREE4; break;
            case D_Dfpermute: r[1] = L_AGREE1; r[2] = L_AGREE1; r[4] = L_COMPAT1; r[5] = L_COMPAT4 | L_AGREE5; break;
            case D_Extract: r[2] = L_COMPAT1 | L_AGREE2; break;
            case D_Replace: r[2] = L_AGREE2; r[3] = L_COMPAT1 | L_AGREE2; break;
            case D_Dist: r[1] = L_AGREE1; break;
            case D_Index: r[1] = L_AGREE1; r[2] = L_AGREE1; break;
            case D_Pack: r[1] = L_AGREE1; r[2] = L_COMPAT1; break;
            default: break;
        }
    }

    void check_args(const Sig &s, int itype, Vec **args) {
        for (int k = 0; k < s.nargs; ++k) {
            Vec *a = args[k];
            int want = s.arg_type[k] == T_Given ? itype : s.arg_type[k];
            if (a->type == T_Pair) die("line %d: %s: argument %d is a pair", line_, s.what.c_str(), k + 1);
            if (want != T_Illegal && want != T_None && want != a->type && !(want == T_String && a->type == T_Int))
                die("line %d: %s: argument %d is %s, expected %s", line_, s.what.c_str(), k + 1,
                    type_name(a->type), type_name(want));
            unsigned r = s.arg_len[k];
            for (int j = 0; j < 5; ++j) {
                if ((r & (L_AGREE1 << j)) && j < s.nargs && a->len != args[j]->len)
                    die("line %d: %s: argument %d has length %d, argument %d has %d", line_, s.what.c_str(), k + 1,
                        a->len, j + 1, args[j]->len);
                if ((r & (L_COMPAT1 << j)) && j < s.nargs) {
                    if (a->type != T_Segdes)
                        die("line %d: %s: argument %d is not a segment descriptor", line_, s.what.c_str(), k + 1);
                    if (a->seg_len != args[j]->len)
                        die("line %d: %s: segment descriptor (argument %d) describes %d elements, argument %d has %d",
                            line_, s.what.c_str(), k + 1, a->seg_len, j + 1, args[j]->len);
                }
            }
        }
    }

    int result_type(const Sig &s, Vec **args) {
        switch (s.res_type) {
            case T_Bool: case T_Int: case T_Float: return s.res_type;
            case T_Agree1: case T_Agree2: case T_Agree3: case T_Agree4: {
                int k = s.res_type - T_Agree1;
                if (k >= s.nargs) die("line %d: %s: bad result type", line_, s.what.c_str());
                return args[k]->type;
            }
            default: die("line %d: %s: result type %s cannot be allocated", line_, s.what.c_str(),
                         type_name(s.res_type));
        }
    }

    int result_len(const Sig &s, Vec **args) {
        unsigned r = s.res_len;
        for (int j = 0; j < 5; ++j)
            if (r & (L_AGREE1 << j)) return args[j]->len;
        for (int j = 0; j < 5; ++j)
            if (r & (L_COMPAT1 << j)) {
                if (args[j]->type != T_Segdes) die("line %d: %s: bad result length", line_, s.what.c_str());
                return args[j]->seg_len;
            }
        if (r == L_SCALAR) return 1;
        die("line %d: %s: bad result length rule", line_, s.what.c_str());
    }

    int scratch_units(const Sig &s, Vec **args) {
        if (!s.scratch_args[0]) return 1;
        Vec *a = args[s.scratch_args[0] - 1];
        if (!s.scratch_args[1]) {
            if (a->type != T_Segdes) return reinterpret_cast<int (*)(int)>(s.scratch_fn)(a->len);
            return reinterpret_cast<int (*)(int, int)>(s.scratch_fn)(a->seg_len, a->len);
        }
        Vec *b = args[s.scratch_args[1] - 1];
        if (a->type != T_Segdes) return reinterpret_cast<int (*)(int, int)>(s.scratch_fn)(a->len, b->len);
        return reinterpret_cast<int (*)(int, int, int, int)>(s.scratch_fn)(a->seg_len, a->len, b->seg_len, b->len);
    }

    // ---------------------------------------------------------- vector ops
    void vector_op(int pc, const Instr &in) {
        Sig &s = signature(pc, in);
        need_stack(s.nargs, s.what.c_str());
        Vec *args[kMaxArgs];
        const size_t base = st_.size() - s.nargs;
        for (int k = 0; k < s.nargs; ++k) args[k] = st_[base + k];
        check_args(s, in.type, args);

        if (s.desc == D_SegOp) {                     // fused MAKE_SEGDES
            seg_make_segdes();
            return;
        }
        const int rtype = result_type(s, args);
        const int rlen = result_len(s, args);
        Vec *d = nullptr;
        if (s.desc == D_Replace) {
            if (args[0]->refs == 1 && !args[0]->owner) {   // REPLACE into the dying source
                d = args[0];
                d->refs++;
            } else {
                d = alloc(rtype, rlen);
                copy_vector(d, args[0]);
            }
        } else if (s.desc != D_Dpermute && s.desc != D_Dfpermute) {
            unsigned mask = s.inplace_fn();
            for (int j = 0; mask && j < s.nargs; ++j, mask >>= 1) {
                Vec *a = args[j];
                if ((mask & 1) && a->refs == 1 && !a->owner && a->type == rtype && a->len == rlen) {
                    d = a;
                    d->refs++;
                    break;
                }
            }
        }
        if (!d) d = alloc(rtype, rlen);
        {
            Scratch scr(*this, scratch_units(s, args));
            call(s, d, args, scr.p);
        }
        for (int k = 0; k < s.nargs; ++k) unref(args[k]);
        st_.resize(base);
        push(d);
    }

    void copy_vector(Vec *d, Vec *s) {
        switch (s->type) {
            case T_Int: { Scratch x(*this, be_.cpy_wuz_scratch(s->len)); dev(be_.cpy_wuz, d->p, s->p, s->len, x.p); break; }
            case T_Bool: { Scratch x(*this, be_.cpy_wub_scratch(s->len)); dev(be_.cpy_wub, d->p, s->p, s->len, x.p); break; }
            case T_Float: { Scratch x(*this, be_.cpy_wud_scratch(s->len)); dev(be_.cpy_wud, d->p, s->p, s->len, x.p); break; }
            default: die("line %d: illegal operand for REPLACE", line_);
        }
    }

    void call(const Sig &s, Vec *d, Vec **a, vec_p scr) {
        typedef void (*F6)(vec_p, vec_p, vec_p, int, int, vec_p);
        typedef void (*F7)(vec_p, vec_p, vec_p, vec_p, int, int, vec_p);
        typedef void (*F10)(vec_p, vec_p, vec_p, vec_p, int, int, vec_p, int, int, vec_p);
        typedef void (*F11)(vec_p, vec_p, vec_p, vec_p, vec_p, int, int, vec_p, int, int, vec_p);
        typedef void (*F12)(vec_p, vec_p, vec_p, vec_p, vec_p, vec_p, int, int, vec_p, int, int, vec_p);
        if (s.desc >= D_Elwise0 && s.desc <= D_Elwise40) {
            const int k = s.desc - D_Elwise0;
            if (k > s.nargs) die("line %d: %s: call shape needs %d arguments", line_, s.what.c_str(), k);
            int len;
            if (k == 0) len = d->type == T_Segdes ? d->seg_len : d->len;
            else len = a[0]->type == T_Segdes ? a[0]->seg_len : a[0]->len;
            std::array<vec_p, kMaxArgs> src{};
            for (int j = 0; j < k; ++j) src[j] = a[j]->p;
            if (!rec_.on) {
                kElwise[k](s.fn, d->p, src.data(), len, scr);
                return;
            }
            key_push(s.fn);
            key_push(d->p);
            for (int j = 0; j < k; ++j) key_push(src[j]);
            key_push(len);
            key_push(scr);
            void *fn = s.fn;
            vec_p dp = d->p;
            rec_.calls.push_back([=]() mutable { kElwise[k](fn, dp, src.data(), len, scr); });
            return;
        }
        switch (s.desc) {
            case D_Scan: case D_Reduce:
                dev(reinterpret_cast<F6>(s.fn), d->p, a[0]->p, a[1]->p, a[1]->seg_len, a[1]->len, scr);
                return;
            case D_Permute:
                dev(reinterpret_cast<F7>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[2]->seg_len, a[2]->len, scr);
                return;
            case D_Bpermute:
                dev(reinterpret_cast<F10>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[2]->seg_len, a[2]->len, a[3]->p,
                                            a[3]->seg_len, a[3]->len, scr);
                return;
            case D_Fpermute: case D_Bfpermute: case D_Dpermute:
                dev(reinterpret_cast<F11>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[3]->p, a[3]->seg_len, a[3]->len,
                                            a[4]->p, a[4]->seg_len, a[4]->len, scr);
                return;
            case D_Dfpermute:
                dev(reinterpret_cast<F12>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[3]->p, a[4]->p, a[4]->seg_len,
                                            a[4]->len, a[5]->p, a[5]->seg_len, a[5]->len, scr);
                return;
            case D_Extract:
                dev(reinterpret_cast<F7>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[2]->seg_len, a[2]->len, scr);
                return;
            case D_Replace:
                dev(reinterpret_cast<F7>(s.fn), d->p, a[1]->p, a[2]->p, a[3]->p, a[3]->seg_len, a[3]->len, scr);
                return;
            case D_Dist:
                dev(reinterpret_cast<F6>(s.fn), d->p, a[0]->p, a[1]->p, a[1]->seg_len, a[1]->len, scr);
                return;
            case D_Index:
                dev(reinterpret_cast<F7>(s.fn), d->p, a[0]->p, a[1]->p, a[2]->p, a[2]->seg_len, a[2]->len, scr);
                return;
            default:
                die("line %d: operation %s not supported", line_, s.what.c_str());
        }
    }

    // ---------------------------------------------------------- segment ops
    void seg_make_segdes() {
        Vec *lens = top();
        if (lens->type != T_Int) die("line %d: not applying MAKE_SEGDES to an integer vector", line_);
        Vec *sd = make_segdes(lens);
        pop_top();
        push(sd);
    }

    void seg_op(const Instr &in) {
        switch (in.op) {
            case OP_MAKE_SEGDES: need_stack(1, "MAKE_SEGDES"); seg_make_segdes(); return;
            case OP_LENGTHS: {
                need_stack(1, "LENGTHS");
                Vec *sd = top();
                if (sd->type != T_Segdes) die("line %d: applying LENGTHS to a non-segment descriptor", line_);
                if (lengths_offset_ >= 0) {
                    // the descriptor holds its lengths: a read-only view, no copy
                    Vec *r = new Vec;
                    r->type = T_Int;
                    r->len = sd->len;
                    r->p = static_cast<char *>(sd->p) + lengths_offset_;
                    r->owner = sd;
                    sd->refs++;
                    pop_top();
                    push(r);
                    return;
                }
                Vec *r = alloc(T_Int, sd->len);
                {
                    Scratch s(*this, be_.len_fos_scratch(sd->seg_len, sd->len));
                    dev(be_.len_fos, r->p, sd->p, sd->seg_len, sd->len, s.p);
                }
                pop_top();
                push(r);
                return;
            }
            case OP_LENGTH: {
                need_stack(1, "LENGTH");
                Vec *v = top();
                if (v->type == T_Segdes || v->type == T_Pair)
                    die("line %d: applying LENGTH to a %s", line_, type_name(v->type));
                int n = v->len;
                pop_top();
                push_int_scalar(n);
                return;
            }
            default: die("line %d: internal error in segment op", line_);
        }
    }

    void pack(int pc, const Instr &in) {
        Sig &s = signature(pc, in);
        need_stack(3, "PACK");
        const size_t base = st_.size() - 3;
        Vec *args[3] = {st_[base], st_[base + 1], st_[base + 2]};
        check_args(s, in.type, args);
        Vec *src = args[0], *flags = args[1], *sd = args[2];
        Vec *lens = alloc(T_Int, sd->len);
        {
            Scratch x(*this, be_.pk1_lev_scratch(sd->seg_len, sd->len));
            be_.pk1_lev(lens->p, flags->p, sd->p, sd->seg_len, sd->len, x.p);
        }
        Vec *dsd = make_segdes(lens);
        unref(lens);
        Vec *dst = alloc(src->type, dsd->seg_len);
        {
            typedef int (*S4)(int, int, int, int);
            typedef void (*F10)(vec_p, vec_p, vec_p, vec_p, int, int, vec_p, int, int, vec_p);
            Scratch x(*this, reinterpret_cast<S4>(s.scratch_fn)(sd->seg_len, sd->len, dsd->seg_len, dsd->len));
            dev(reinterpret_cast<F10>(s.fn), dst->p, src->p, flags->p, sd->p, sd->seg_len, sd->len, dsd->p,
                                        dsd->seg_len, dsd->len, x.p);
        }
        for (int k = 0; k < 3; ++k) unref(args[k]);
        st_.resize(base);
        push(dst);
        push(dsd);
    }

    // ---------------------------------------------------------- stack ops
    void do_copy(int count, int start) {
        if ((size_t)(count + start) > st_.size())
            die("line %d: can't do a COPY %d %d; only %zu items on the stack", line_, count, start, st_.size());
        const size_t from = st_.size() - start - count;
        for (int k = 0; k < count; ++k) {
            Vec *v = st_[from + k];
            v->refs++;
            st_.push_back(v);
        }
    }
    void do_pop(int count, int start) {
        if ((size_t)(count + start) > st_.size())
            die("line %d: can't do a POP %d %d; only %zu items on the stack", line_, count, start, st_.size());
        const size_t from = st_.size() - start - count;
        for (int k = 0; k < count; ++k) unref(st_[from + k]);
        st_.erase(st_.begin() + from, st_.begin() + from + count);
    }
    void do_cpop(int count, int start) {
        if ((size_t)(count + start) > st_.size())
            die("line %d: can't do a CPOP %d %d; only %zu items on the stack", line_, count, start, st_.size());
        const size_t from = st_.size() - start - count;
        std::vector<Vec *> moved(st_.begin() + from, st_.begin() + from + count);
        st_.erase(st_.begin() + from, st_.begin() + from + count);
        st_.insert(st_.end(), moved.begin(), moved.end());
    }

    // ---------------------------------------------------------- I/O
    void write_value(Vec *v, FILE *f) {
        Vec *pv = v;
        if (v->type == T_Segdes) {
            pv = alloc(T_Int, v->len);
            Scratch s(*this, be_.len_fos_scratch(v->seg_len, v->len));
            be_.len_fos(pv->p, v->p, v->seg_len, v->len, s.p);
            std::fputc('[', f);
        } else {
            std::fputc('(', f);
        }
        const int n = pv->len;
        switch (pv->type) {
            case T_Int: {
                std::vector<int> h(n);
                Scratch s(*this, be_.v2c_fuz_scratch(n));
                be_.v2c_fuz(h.data(), pv->p, n, s.p);
                for (int i = 0; i < n; ++i) std::fprintf(f, " %d", h[i]);
                break;
            }
            case T_Bool: {
                std::vector<cvl_bool> h(n);
                Scratch s(*this, be_.v2c_fub_scratch(n));
                be_.v2c_fub(h.data(), pv->p, n, s.p);
                for (int i = 0; i < n; ++i) {
                    std::fputc(' ', f);
                    std::fputc(h[i] ? 'T' : 'F', f);
                }
                break;
            }
            case T_Float: {
                std::vector<double> h(n);
                Scratch s(*this, be_.v2c_fud_scratch(n));
                be_.v2c_fud(h.data(), pv->p, n, s.p);
                for (int i = 0; i < n; ++i) std::fprintf(f, " %.14e", h[i]);
                break;
            }
            default: die("internal error: illegally typed vector");
        }
        std::fputc(' ', f);
        if (v->type == T_Segdes) {
            std::fputc(']', f);
            unref(pv);
        } else {
            std::fputc(')', f);
        }
        std::fputc('\n', f);
        if (std::ferror(f)) set_io_error(std::strerror(errno));
        else std::fflush(f);
    }

    void write_string(Vec *v, FILE *f) {
        std::vector<int> h(v->len);
        Scratch s(*this, be_.v2c_fuz_scratch(v->len));
        be_.v2c_fuz(h.data(), v->p, v->len, s.p);
        for (int c : h) std::fputc((char)c, f);
        if (std::ferror(f)) set_io_error(std::strerror(errno));
        else std::fflush(f);
    }

    void set_io_error(const std::string &m) {
        io_error_ = true;
        io_msg_ = m;
    }

    // push the error string and the success flag of an I/O instruction
    void push_io_status() {
        if (io_error_) {
            std::vector<int> h(io_msg_.begin(), io_msg_.end());
            push_int_vector(h);
            push_bool_scalar(0);
        } else {
            push(alloc(T_Int, 0));
            push_bool_scalar(1);
        }
    }

    void push_int_vector(const std::vector<int> &h) {
        Vec *r = alloc(T_Int, (int)h.size());
        std::vector<int> tmp(h);
        Scratch s(*this, be_.c2v_fuz_scratch((int)h.size()));
        be_.c2v_fuz(r->p, tmp.data(), (int)h.size(), s.p);
        push(r);
    }

    void do_write(const Instr &in, FILE *f) {
        io_error_ = false;
        need_stack(1, "WRITE");
        Vec *v = top();
        if (v->type == T_Pair) die("line %d: cannot WRITE a pair", line_);
        if (in.type == T_String) write_string(v, f);
        else write_value(v, f);
        pop_top();
        push_io_status();
    }

    void do_read(const Instr &in, FILE *f) {
        io_error_ = false;
        const int type = in.type;
        const int stack_type = (type == T_String || type == T_Segdes) ? T_Int : type;
        const char open = type == T_String ? '"' : type == T_Segdes ? '[' : '(';
        const char close = type == T_String ? '"' : type == T_Segdes ? ']' : ')';
        std::vector<int> ints;
        std::vector<double> dbls;
        int c;
        while ((c = std::getc(f)) != EOF && std::isspace(c)) {}
        if (c != open) {
            set_io_error(std::string("vinterp: input error: vector must start with ") + open);
        } else if (type == T_Float) {
            for (;;) {
                double x;
                if (std::fscanf(f, " %lf", &x) != 1) { c = std::getc(f); break; }
                dbls.push_back(x);
            }
        } else if (type == T_Bool) {
            for (;;) {
                c = std::getc(f);
                if (c == ' ' || c == '\t' || c == '\n') continue;
                if (c == 'T') ints.push_back(1);
                else if (c == 'F') ints.push_back(0);
                else break;
            }
        } else if (type == T_String) {
            while ((c = std::getc(f)) != EOF && c != '"') ints.push_back(c);
        } else {
            for (;;) {
                while ((c = std::getc(f)) != EOF && std::isspace(c)) {}
                bool neg = false;
                if (c == '-') { neg = true; c = std::getc(f); }
                else if (c == '+') c = std::getc(f);
                if (!std::isdigit(c)) break;
                unsigned v = 0;
                while (std::isdigit(c)) { v = 10 * v + (unsigned)(c - '0'); c = std::getc(f); }
                ints.push_back(neg ? -(int)v : (int)v);
                std::ungetc(c, f);
            }
        }
        if (!io_error_ && c != close)
            set_io_error(std::string("input error: vector must end with ") + close);
        if (io_error_) {
            push(alloc(stack_type, 0));
        } else if (type == T_Float) {
            Vec *r = alloc(T_Float, (int)dbls.size());
            Scratch s(*this, be_.c2v_fud_scratch(r->len));
            be_.c2v_fud(r->p, dbls.data(), r->len, s.p);
            push(r);
        } else if (type == T_Bool) {
            Vec *r = alloc(T_Bool, (int)ints.size());
            std::vector<cvl_bool> h(ints.begin(), ints.end());
            Scratch s(*this, be_.c2v_fub_scratch(r->len));
            be_.c2v_fub(r->p, h.data(), r->len, s.p);
            push(r);
        } else {
            push_int_vector(ints);
            if (type == T_Segdes) seg_make_segdes();
        }
        push_io_status();
    }

    FILE *file_of(int fd) {
        if (fd < 0 || fd >= (int)files_.size()) return nullptr;
        return files_[fd];
    }

    void io_op(const Instr &in) {
        switch (in.op) {
            case OP_READ: do_read(in, stdin); return;
            case OP_WRITE: do_write(in, stderr); return;
            case OP_FOPEN: {
                io_error_ = false;
                need_stack(2, "FOPEN");
                int mode = int_scalar(top(), "FOPEN");
                pop_top();
                const char *m = mode == 1 ? "r" : mode == 2 ? "w" : mode == 3 ? "a" : nullptr;
                if (!m) die("line %d: FOPEN: top of stack (%d) not legal mode", line_, mode);
                Vec *nv = top();
                std::vector<int> h(nv->len);
                {
                    Scratch s(*this, be_.v2c_fuz_scratch(nv->len));
                    be_.v2c_fuz(h.data(), nv->p, nv->len, s.p);
                }
                pop_top();
                std::string name(h.begin(), h.end());
                int fd = -1;
                FILE *f = std::fopen(name.c_str(), m);
                if (!f) {
                    set_io_error(std::strerror(errno));
                } else {
                    for (size_t k = 3; k < files_.size(); ++k)
                        if (!files_[k]) { files_[k] = f; fd = (int)k; break; }
                    if (fd < 0) { files_.push_back(f); fd = (int)files_.size() - 1; }
                }
                push_int_scalar(fd);
                push_io_status();
                return;
            }
            case OP_FCLOSE: {
                io_error_ = false;
                need_stack(1, "FCLOSE");
                int fd = int_scalar(top(), "FCLOSE");
                pop_top();
                FILE *f = file_of(fd);
                if (!f) set_io_error("illegal VCODE file descriptor in FCLOSE");
                else if (std::fclose(f) == EOF) set_io_error("vinterp: FCLOSE: could not close file");
                else files_[fd] = nullptr;
                push_io_status();
                return;
            }
            case OP_FREAD: {
                io_error_ = false;
                need_stack(1, "FREAD");
                int fd = int_scalar(top(), "FREAD");
                pop_top();
                FILE *f = file_of(fd);
                if (!f) {
                    set_io_error("illegal VCODE file descriptor for FREAD");
                    push(alloc(in.type == T_String || in.type == T_Segdes ? T_Int : in.type, 0));
                    push_io_status();
                } else {
                    do_read(in, f);
                }
                return;
            }
            case OP_FWRITE: {
                io_error_ = false;
                need_stack(2, "FWRITE");
                int fd = int_scalar(top(), "FWRITE");
                pop_top();
                FILE *f = file_of(fd);
                if (!f) {
                    set_io_error("illegal VCODE file descriptor in FWRITE");
                    pop_top();
                    push_io_status();
                } else {
                    do_write(in, f);
                }
                return;
            }
            case OP_FREAD_CHAR: {
                io_error_ = false;
                need_stack(3, "FREAD_CHAR");
                int fd = int_scalar(top(), "FREAD_CHAR");
                pop_top();
                FILE *f = file_of(fd);
                if (!f) {
                    pop_top();
                    pop_top();
                    push(alloc(T_Int, 0));
                    push(alloc(T_Int, 0));
                    set_io_error("illegal VCODE file descriptor in FREAD");
                    push_io_status();
                    return;
                }
                int max_char = int_scalar(top(), "FREAD_CHAR");
                if (max_char < 0) max_char = 1 << 30;
                pop_top();
                Vec *dv = top();
                std::vector<int> delims(dv->len);
                {
                    Scratch s(*this, be_.v2c_fuz_scratch(dv->len));
                    be_.v2c_fuz(delims.data(), dv->p, dv->len, s.p);
                }
                pop_top();
                bool mask[256] = {false};
                for (int d : delims) {
                    if (d < 0 || d >= 256) die("line %d: FREAD_CHAR delimiter %d is not a character", line_, d);
                    mask[d] = true;
                }
                std::vector<int> got;
                int c = 0;
                while ((int)got.size() < max_char && (c = std::getc(f)) != EOF && !mask[c]) got.push_back(c);
                int status = (int)got.size() == max_char ? -2 : c == EOF ? -1 : c;
                push_int_vector(got);
                push_int_scalar(status);
                push(alloc(T_Int, 0));
                push_bool_scalar(1);
                return;
            }
            case OP_SPAWN: {                       // process spawning is not offered
                need_stack(4, "SPAWN");
                for (int k = 0; k < 4; ++k) pop_top();
                push_int_scalar(-1);
                push_int_scalar(-1);
                push_int_scalar(-1);
                io_error_ = true;
                io_msg_ = "SPAWN not permitted";
                push_io_status();
                return;
            }
            default: die("line %d: internal error in I/O op", line_);
        }
    }

    // ---------------------------------------------------------- main loop
    // ---------------------------------------------------------- diagnostics
    // -d runtime, -d stack, -d values, -d heap (nesl/vcode/main.c:130-139)
    // and the per-function profile the reference builds with VCODE_PROFILE
    // (vcode-profile.c:79-136): one line per function, sorted by decreasing
    // exclusive time.  A function's time runs from its CALL to its RET; the
    // library's stream is synchronized at both, so device work is charged
    // to the function that queued it.
    struct FuncTime {
        std::string name;
        long long calls = 0;
        double time = 0, min = 1e9, max = 0, cum = 0;
    };
    struct TimeFrame {
        int func;
        double start, child;
    };

    void profile_init() {
        sync_ = reinterpret_cast<void (*)()>(be_.sym("cvlhip_synchronize"));
        for (const auto &l : prog_.labels) {
            func_of_pc_[l.second] = (int)funcs_.size();
            funcs_.push_back(FuncTime{l.first});
        }
    }
    double now() {
        if (sync_) sync_();
        timespec t;
        clock_gettime(CLOCK_REALTIME, &t);
        return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
    }
    void profile_call(int target_pc) {
        auto it = func_of_pc_.find(target_pc);
        times_.push_back(TimeFrame{it == func_of_pc_.end() ? -1 : it->second, now(), 0.0});
    }
    void profile_return() {
        if (times_.empty()) return;
        const TimeFrame top = times_.back();
        times_.pop_back();
        const double dt = now() - top.start;
        if (!times_.empty()) times_.back().child += dt;
        if (top.func < 0) return;
        FuncTime &f = funcs_[top.func];
        f.cum += dt;
        const double self = dt - top.child;
        f.time += self;
        if (self < f.min) f.min = self;
        if (self > f.max) f.max = self;
        f.calls++;
    }
    void profile_report() {
        std::vector<int> order(funcs_.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        std::sort(order.begin(), order.end(), [&](int a, int b) {
            const FuncTime &p = funcs_[a], &q = funcs_[b];
            if ((p.calls > 0) != (q.calls > 0)) return p.calls > 0;
            if (p.calls > 0 && p.time != q.time) return p.time > q.time;
            return p.name < q.name;
        });
        FILE *out = std::fopen(opt_.profile_path.c_str(), "w");
        if (!out) {
            std::fprintf(stderr, "unable to open \"%s\"\n", opt_.profile_path.c_str());
            return;
        }
        std::fprintf(stderr, "dumping profile information to %s\n", opt_.profile_path.c_str());
        std::fprintf(out, "# Calls     Time         Min.        Avg.        Max.    Cumulative   Name\n");
        std::fprintf(out, "-------  ----------  ----------  ----------  ----------  ----------   ---------------------\n");
        for (int i : order) {
            const FuncTime &p = funcs_[i];
            if (p.calls > 0)
                std::fprintf(out, " %6lld  %10.6f  %10.6f  %10.6f  %10.6f  %10.6f   %s\n", p.calls, p.time, p.min,
                             p.time / (double)p.calls, p.max, p.cum, p.name.c_str());
            else
                std::fprintf(out, "      0     ----        ----        ----        ----        ----      %s\n",
                             p.name.c_str());
        }
        std::fclose(out);
    }

    // One stack entry: type and length, and with -d values its first values.
    void trace_entry(const Vec *v) {
        if (v->type == T_Pair) {
            std::fprintf(stderr, " <pair>");
            return;
        }
        if (v->type == T_Segdes) {
            std::fprintf(stderr, " SEGDES[%d segments, %d elements]", v->len, v->seg_len);
            return;
        }
        std::fprintf(stderr, " %s[%d]", type_name(v->type), v->len);
        if (!opt_.value_trace || v->len <= 0) return;
        const int n = v->len < 8 ? v->len : 8;
        std::fprintf(stderr, "(");
        if (v->type == T_Float) {
            double h[8];
            Scratch s(*this, be_.v2c_fud_scratch(n));
            be_.v2c_fud(h, v->p, n, s.p);
            for (int i = 0; i < n; ++i) std::fprintf(stderr, " %.6g", h[i]);
        } else if (v->type == T_Bool) {
            cvl_bool h[8];
            Scratch s(*this, be_.v2c_fub_scratch(n));
            be_.v2c_fub(h, v->p, n, s.p);
            for (int i = 0; i < n; ++i) std::fprintf(stderr, " %c", h[i] ? 'T' : 'F');
        } else {
            int h[8];
            Scratch s(*this, be_.v2c_fuz_scratch(n));
            be_.v2c_fuz(h, v->p, n, s.p);
            for (int i = 0; i < n; ++i) std::fprintf(stderr, " %d", h[i]);
        }
        std::fprintf(stderr, "%s )", v->len > n ? " ..." : "");
    }
    void trace_after(const Instr &in) {
        if (opt_.stack_trace || opt_.value_trace) {
            std::fprintf(stderr, "stack after line %d (%zu):", in.line, st_.size());
            for (const Vec *v : st_) trace_entry(v);
            std::fprintf(stderr, "\n");
        }
        if (opt_.heap_trace)
            std::fprintf(stderr, "heap after line %d: %zu of %zu units in use, %zu free blocks, largest %zu\n",
                         in.line, arena_.used(), arena_.total(), arena_.free_blocks(), arena_.largest_free());
    }

    void execute(int pc) {
        std::vector<int> ret;
        ret.push_back(-1);
        std::vector<const std::string *> frames;             // -d runtime
        if (opt_.runtime_trace) {
            std::fprintf(stderr, "call %s\n", prog_.entry.c_str());
            frames.push_back(&prog_.entry);
        }
        const bool profiling = !opt_.profile_path.empty();
        if (profiling) profile_call(pc);
        alignas(64) unsigned char t_start[64] = {0}, t_end[64];
        while (pc >= 0) {
            const Instr &in = prog_.code[pc];
            line_ = in.line;
            ++executed_;
            if (graphs_) {
                const bool cap = capturable(pc, in);
                if (cap && !rec_.on) {
                    rec_.on = true;
                    rec_.calls.clear();
                    rec_.key.clear();
                } else if (!cap && rec_.on) {
                    flush();
                }
            }
            if (opt_.trace) std::fprintf(stderr, "%d: %s %s (stack %zu)\n", in.line, op_info(in.op).name,
                                         type_name(in.type), st_.size());
            int next = pc + 1;
            switch (in.op) {
                case OP_COPY: do_copy(in.a, in.b); break;
                case OP_POP: do_pop(in.a, in.b); break;
                case OP_CPOP: do_cpop(in.a, in.b); break;
                case OP_PAIR: {
                    need_stack(2, "PAIR");
                    Vec *p = new Vec;
                    p->type = T_Pair;
                    p->second = st_.back();
                    st_.pop_back();
                    p->first = st_.back();
                    st_.back() = p;
                    break;
                }
                case OP_UNPAIR: {
                    Vec *p = top();
                    if (p->type != T_Pair) die("line %d: attempting to UNPAIR a non-pair (%s)", line_,
                                               type_name(p->type));
                    st_.pop_back();
                    p->first->refs++;
                    p->second->refs++;
                    st_.push_back(p->first);
                    st_.push_back(p->second);
                    unref(p);
                    break;
                }
                case OP_CALL:
                    if (opt_.runtime_trace) {
                        std::fprintf(stderr, "%*scall %s (line %d)\n", 2 * (int)frames.size(), "",
                                     in.label.c_str(), in.line);
                        frames.push_back(&in.label);
                    }
                    if (profiling) profile_call(in.branch);
                    ret.push_back(pc + 1);
                    next = in.branch;
                    break;
                case OP_RET:
                    if (opt_.runtime_trace && !frames.empty()) {
                        const std::string *f = frames.back();
                        frames.pop_back();
                        std::fprintf(stderr, "%*sreturn from %s\n", 2 * (int)frames.size(), "", f->c_str());
                    }
                    if (profiling) profile_return();
                    next = ret.back();
                    ret.pop_back();
                    break;
                case OP_IF: {
                    Vec *f = top();
                    if (f->type != T_Bool || f->len != 1) die("line %d: IF: top of stack not bool scalar", line_);
                    int v;
                    {
                        Scratch s(*this, be_.ext_vub_scratch(1));
                        v = be_.ext_vub(f->p, 0, 1, s.p);
                    }
                    pop_top();
                    if (!v) next = in.branch;
                    break;
                }
                case OP_ELSE: next = in.branch; break;
                case OP_CONST:
                    if (in.pooled) {
                        Vec *v = pool_[in.pool_index];
                        v->refs++;
                        push(v);
                    } else if (in.type == T_Int) {
                        push_int_scalar((int)in.ival);
                    } else if (in.type == T_Float) {
                        push_float_scalar(in.dval);
                    } else if (in.type == T_Bool) {
                        push_bool_scalar((int)in.ival);
                    } else {
                        die("line %d: illegal type for CONST", line_);
                    }
                    break;
                case OP_EXIT:
                    if (rec_.on) flush();
                    std::fflush(stdout);
                    std::exit(0);
                case OP_START_TIMER: be_.tgt_fos(t_start); break;
                case OP_STOP_TIMER: {
                    be_.tgt_fos(t_end);
                    push_float_scalar(be_.tdf_fos(t_end, t_start));
                    break;
                }
                case OP_SRAND: {
                    need_stack(1, "SRAND");
                    int seed = int_scalar(top(), "SRAND");
                    be_.rnd_foz(seed);
                    pop_top();
                    push_bool_scalar(1);
                    break;
                }
                case OP_READ: case OP_WRITE: case OP_FOPEN: case OP_FCLOSE: case OP_FREAD:
                case OP_FREAD_CHAR: case OP_FWRITE: case OP_SPAWN:
                    io_op(in);
                    break;
                case OP_MAKE_SEGDES: case OP_LENGTHS: case OP_LENGTH: seg_op(in); break;
                case OP_PACK: pack(pc, in); break;
                default: vector_op(pc, in); break;
            }
            if (opt_.stack_trace || opt_.value_trace || opt_.heap_trace) trace_after(in);
            pc = next;
        }
        if (rec_.on) flush();
    }

    const Program &prog_;
    const Options &opt_;
    Backend be_;
    Arena arena_;
    std::vector<Sig> sigs_;
    std::vector<Vec *> st_;
    std::vector<Vec *> pool_;
    std::vector<FILE *> files_{stdin, stdout, stderr};
    const VopDes *vops_ = nullptr;
    const CvlTriple *funs_ = nullptr;
    int nfused_ = -1;
    int line_ = 0;
    unsigned long long executed_ = 0;
    int lengths_offset_ = -1;
    Recorder rec_;
    bool graphs_ = false;
    std::map<uint64_t, GraphEntry> graphs_cache_;
    unsigned long long graph_captures_ = 0, graph_replays_ = 0;
    int (*cap_begin_)() = nullptr;
    int (*cap_end_)(void **) = nullptr;
    int (*graph_launch_)(void *) = nullptr;
    void (*graph_destroy_)(void *) = nullptr;
    bool io_error_ = false;
    std::string io_msg_;
    // -p
    void (*sync_)() = nullptr;
    std::map<int, int> func_of_pc_;
    std::vector<FuncTime> funcs_;
    std::vector<TimeFrame> times_;
};

int run(const Program &prog, const Options &opt) {
    Machine m(prog, opt);
    return m.run();
}

}  // namespace vcode
