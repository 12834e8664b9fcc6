// Host side of the C ABI: context, input binding, the `fix metad` keyword
// parser, DIM expression compiler, HILLS I/O and the per-step launch
// sequence (plain launches for the drop-in path, one captured HIP graph per
// step for the standalone/batched driver).
//
// Mirrors of the reference interface (paths relative to the Metad-toy tree):
//   FixMetadynamics ctor / init / post_force   src/core/fix_crystallize.cu:32-472
//   Steinhardt::create (STEINH keyword list)   src/cvs/CVs_Steinhardt.cu:33-161
//   MetaDimensionManager (SYMBOL, DIM)         src/grids/DimSet.cu:28-97
//   GH_t0_uniformGrid (grid, HILLS, restart)   src/grids/GH_t0_uniformGrid.cu:73-220
#include <algorithm>
#include <cctype>
#include <cmath>
#include <functional>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "engine.h"
#include "mdt.h"

using namespace mdt;

namespace {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
struct Err : std::exception {
    int code;
    std::string msg;
    Err(int c, std::string m) : code(c), msg(std::move(m)) {}
};

std::string fmt(const char* f, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw Err(MDT_EHIP, fmt("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                    __FILE__, __LINE__));                                 \
    } while (0)

template <class T>
struct DBuf {  // device buffer
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Err(MDT_ENOMEM, fmt("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e)));
        }
        n = count;
    }
    // blocking: the host staging vectors are usually freed right after
    void upload(const T* h, size_t count, hipStream_t st, size_t off = 0) {
        if (!count) return;
        HIPCHK(hipMemcpyAsync(p + off, h, count * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
};

double parse_num(const char* s, const char* what) {
    char* end = nullptr;
    double v = strtod(s, &end);
    if (!s[0] || (end && *end)) throw Err(MDT_EINVAL, fmt("expected a number for %s, got '%s'", what, s));
    return v;
}
long long parse_int(const char* s, const char* what) {
    char* end = nullptr;
    long long v = strtoll(s, &end, 10);
    if (!s[0] || (end && *end)) throw Err(MDT_EINVAL, fmt("expected an integer for %s, got '%s'", what, s));
    return v;
}

// ---------------------------------------------------------------------------
// Expression compiler: infix -> reverse Polish over symbols (DIM expressions;
// the reference compiles them with exprtk, src/grids/DimSet.cu:54-66).
// Supported: numbers, symbols, pi, + - * / ^, unary -, parentheses and
// sin cos tan asin acos atan sinh cosh tanh exp log log10 sqrt abs pow min max.
// ---------------------------------------------------------------------------
struct ExprCompiler {
    const std::string& src;
    const std::vector<std::string>& syms;
    size_t pos = 0;
    ExprProg prog{};
    ExprCompiler(const std::string& s, const std::vector<std::string>& y) : src(s), syms(y) { prog.n = 0; }

    void emit(int op, double c = 0) {
        if (prog.n >= kMaxProg) throw Err(MDT_EINVAL, "DIM expression too long");
        prog.op[prog.n] = op;
        prog.c[prog.n] = c;
        prog.n++;
    }
    void ws() { while (pos < src.size() && isspace((unsigned char)src[pos])) ++pos; }
    bool eat(char ch) { ws(); if (pos < src.size() && src[pos] == ch) { ++pos; return true; } return false; }
    [[noreturn]] void fail(const char* why) {
        throw Err(MDT_EINVAL, fmt("DIM expression '%s': %s at column %zu", src.c_str(), why, pos + 1));
    }
    void expr() {  // sum
        term();
        for (;;) {
            if (eat('+')) { term(); emit(OP_ADD); }
            else if (eat('-')) { term(); emit(OP_SUB); }
            else break;
        }
    }
    void term() {
        unary();
        for (;;) {
            if (eat('*')) { unary(); emit(OP_MUL); }
            else if (eat('/')) { unary(); emit(OP_DIV); }
            else break;
        }
    }
    void unary() {
        if (eat('-')) { unary(); emit(OP_NEG); return; }
        if (eat('+')) { unary(); return; }
        power();
    }
    void power() {
        primary();
        if (eat('^')) { unary(); emit(OP_POW); }
    }
    void primary() {
        ws();
        if (pos >= src.size()) fail("unexpected end");
        char ch = src[pos];
        if (ch == '(') { ++pos; expr(); if (!eat(')')) fail("missing ')'"); return; }
        if (isdigit((unsigned char)ch) || ch == '.') {
            const char* b = src.c_str() + pos;
            char* e = nullptr;
            double v = strtod(b, &e);
            if (e == b) fail("bad number");
            pos += (size_t)(e - b);
            emit(OP_CONST, v);
            return;
        }
        if (isalpha((unsigned char)ch) || ch == '_') {
            size_t b = pos;
            while (pos < src.size() && (isalnum((unsigned char)src[pos]) || src[pos] == '_')) ++pos;
            std::string name = src.substr(b, pos - b);
            static const std::map<std::string, int> f1 = {
                {"sin", OP_SIN}, {"cos", OP_COS}, {"tan", OP_TAN}, {"exp", OP_EXP}, {"log", OP_LOG},
                {"sqrt", OP_SQRT}, {"abs", OP_ABS}, {"tanh", OP_TANH}, {"asin", OP_ASIN},
                {"acos", OP_ACOS}, {"atan", OP_ATAN}, {"sinh", OP_SINH}, {"cosh", OP_COSH},
                {"log10", OP_LOG10}};
            static const std::map<std::string, int> f2 = {{"pow", OP_POW}, {"min", OP_MIN}, {"max", OP_MAX}};
            ws();
            if (pos < src.size() && src[pos] == '(') {
                ++pos;
                auto a = f1.find(name);
                if (a != f1.end()) { expr(); if (!eat(')')) fail("missing ')'"); emit(a->second); return; }
                auto b2 = f2.find(name);
                if (b2 != f2.end()) {
                    expr(); if (!eat(',')) fail("expected ','");
                    expr(); if (!eat(')')) fail("missing ')'");
                    emit(b2->second);
                    return;
                }
                fail("unknown function");
            }
            if (name == "pi") { emit(OP_CONST, 3.14159265358979323846); return; }
            for (size_t k = 0; k < syms.size(); ++k)
                if (syms[k] == name) { emit(OP_SYM, (double)k); return; }
            throw Err(MDT_EINVAL, fmt("DIM expression '%s': unknown symbol '%s'", src.c_str(), name.c_str()));
        }
        fail("unexpected character");
    }
    ExprProg compile() {
        expr();
        ws();
        if (pos != src.size()) fail("trailing characters");
        int depth = 0, maxd = 0;
        for (int k = 0; k < prog.n; ++k) {
            int op = prog.op[k];
            if (op == OP_CONST || op == OP_SYM) depth++;
            else if ((op >= OP_ADD && op <= OP_POW) || op == OP_MIN || op == OP_MAX) depth--;
            maxd = std::max(maxd, depth);
        }
        if (maxd > kStack) throw Err(MDT_EINVAL, "DIM expression nests too deeply");
        return prog;
    }
};

// ---------------------------------------------------------------------------
// walkers and context
// ---------------------------------------------------------------------------
struct WalkerIn {
    bool has_box = false, has_x = false, has_nl = false;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int64_t n_local = 0, n_all = 0, nnz = 0;
    std::vector<double> x;
    std::vector<int32_t> off, idx, owner;
    // group memberships of local atoms by group name ("" = the default, bound by
    // mdt_bind_group, used by any group name without a binding of its own)
    std::map<std::string, std::vector<uint8_t>> groups;
    std::vector<uint8_t> share;       // this context's share of the CV sums (empty: every atom)
    std::vector<int32_t> types;       // atom types of local atoms (LAMMPS atom->type), empty: unbound
    const double* x_src = nullptr;   // attached engine buffer read at the start of every step
    // after commit
    int64_t atom_off = 0, ghost_off = 0, nnz_off = 0;
    int32_t tile_off = 0, ntiles = 0;
};

}  // namespace

struct mdt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<WalkerIn> w;
    std::string err;
    bool committed = false;
    uint64_t generation = 0;   // bumps on every commit (invalidates graphs)
    uint64_t generation_attach = 0;   // bumps on every attach (invalidates graphs)
    int64_t total_local = 0, total_ghost = 0, total_nnz = 0;
    bool any_group = false;
    DBuf<double> x;
    DBuf<double4> x4;
    DBuf<float4> x4f;
    DBuf<int32_t> off, idx, owner;
    // owners and reverse positions of the lists the Steinhardt kernels walk: the bound
    // lists when none is longer than kMaxList, else the near lists
    DBuf<int32_t> eown;
    DBuf<uint8_t> revpos;
    int64_t max_list = 0;       // longest bound list
    // the near lists (see NearDev): in place at the bound offsets, the build positions,
    // the epoch flags; built when a bound list is longer than kMaxList
    DBuf<int32_t> nidx, ncnt, nflag, noverflow;
    DBuf<float4> xpr;
    DBuf<int64_t> nepoch;
    double near_rp2 = 0.0;      // 0: no near list
    double near_rc = 0.0;       // the slack-widened cutoff the near lists serve
    uint64_t near_gen = 0;      // generation it was built for
    // device memberships by group name, concatenated over walkers; per-walker counts
    std::map<std::string, DBuf<uint8_t>> dgroups;
    std::map<std::string, std::vector<double>> gcount;
    std::map<std::string, std::string> gmissing;   // name -> the first walker without a membership
    DBuf<uint8_t> dshare;                          // concatenated shares (when any walker has one)
    bool any_share = false;
    DBuf<Tile> tiles;
    DBuf<double> f;
    std::vector<Tile> h_tiles;
    std::vector<int32_t> tile_begin;
    std::vector<mdt_fix*> fixes;   // fixes created on this context (detached on destroy)
    // the last device neighbour build (mdt_neighbor_build)
    DBuf<double> nl_xall;
    DBuf<int32_t> nl_off, nl_idx, nl_owner;
    int64_t nl_n = 0, nl_n_all = 0, nl_nnz = 0;
    double nl_lo[3] = {0, 0, 0}, nl_hi[3] = {0, 0, 0};
    bool nl_ok = false;

    // the bound lists (no topology): the kernels that walk whole lists
    Geo geo() const {
        Geo g{};
        g.x4 = x4.p; g.x4f = x4f.p; g.off = off.p; g.idx = idx.p; g.owner = owner.p;
        g.x64 = engine_x64();
        g.eown = nullptr; g.revpos = nullptr; g.cnt = nullptr;
        g.group = nullptr;
        g.share = any_share ? dshare.p : nullptr;
        g.tiles = tiles.p; g.ntiles = (int32_t)h_tiles.size(); g.n_local_total = (int32_t)total_local;
        return g;
    }
    bool near_on() const { return near_rp2 > 0.0 && near_gen == generation; }
    // the lists of the N-nearest selection and its gradient, with their topology
    Geo geo_near() const {
        Geo g = geo();
        g.eown = eown.p; g.revpos = revpos.p;
        if (near_on()) { g.idx = nidx.p; g.cnt = ncnt.p; }
        return g;
    }
    NearDev near_dev() {
        NearDev d{};
        d.off = off.p; d.idx = idx.p; d.nidx = nidx.p; d.ncnt = ncnt.p;
        d.xpr = xpr.p; d.flag = nflag.p; d.epoch = nepoch.p; d.overflow = noverflow.p;
        d.rp2 = near_rp2;
        // a pair can cross from beyond Rp into rc only when its separation changed by more
        // than Rp - rc, i.e. when an atom moved more than half of that
        const double half = 0.5 * (std::sqrt(near_rp2) - near_rc);
        d.move2 = (float)(0.95 * half * half);
        d.n_local = (int32_t)total_local; d.n_all = (int32_t)(total_local + total_ghost);
        return d;
    }
    // (re)build the near lists for radius^2 rp2 when they are missing, stale or too short
    void ensure_near(double rc2);
    // bring the near lists up to date inside the stream (a rebuild runs only when some
    // atom moved past the threshold since the last one)
    void refresh_near() {
        if (!near_on()) return;
        const NearDev nd = near_dev();
        HIPCHK(launch_prune(geo(), nd, 0, stream));
        HIPCHK(launch_topology_near(geo_near(), nd, eown.p, revpos.p, 0, stream));
    }
    // device membership of a group name (nullptr for "all"); a name without a binding
    // of its own falls back to the default group
    const std::string& group_key(const std::string& name) const {
        static const std::string dflt;
        if (dgroups.count(name)) return name;
        if (dgroups.count(dflt)) return dflt;
        throw Err(MDT_EINVAL, fmt("group '%s' used but no group membership was bound (mdt_bind_group)", name.c_str()));
    }
    const uint8_t* group_ptr(const std::string& name) const {
        if (name == "all") return nullptr;
        const std::string& k = group_key(name);
        auto m = gmissing.find(k);
        if (m != gmissing.end())
            throw Err(MDT_EINVAL, fmt("group '%s': %s", name.c_str(), m->second.c_str()));
        return dgroups.at(k).p;
    }
    double group_count(const std::string& name, int walker) const {
        if (name == "all") return (double)w[walker].n_local;
        group_ptr(name);
        return gcount.at(group_key(name))[walker];
    }
    // atoms of a group in this context's share of a walker (the count its CV sums cover)
    double share_count(const std::string& name, int walker) const {
        const WalkerIn& a = w[walker];
        if (a.share.empty()) return group_count(name, walker);
        const std::vector<uint8_t>* m = nullptr;
        if (name != "all") {
            group_ptr(name);
            auto it = a.groups.find(name);
            if (it == a.groups.end()) it = a.groups.find(std::string());
            if (it != a.groups.end()) m = &it->second;
        }
        double c = 0;
        for (int64_t i = 0; i < a.n_local; ++i) c += (a.share[i] && (!m || (*m)[i])) ? 1 : 0;
        return c;
    }
    WalkerIn& walker(int k) {
        if (k < 0 || k >= (int)w.size()) throw Err(MDT_EINVAL, fmt("walker %d out of range [0,%zu)", k, w.size()));
        return w[k];
    }
    void commit();
    // the attached engine buffer of a one-walker context: its FP64 positions in the x4
    // order (locals then ghosts), read in place by the selection's exact tests
    const double* engine_x64() const { return w.size() == 1 ? w[0].x_src : nullptr; }
    bool x4_stale = false;      // the last pack wrote only x4f (x4 holds older positions)
    // pack walker k's positions x (n_all*3 device doubles, locals then ghosts); without
    // with_x4 (one attached walker only) the FP64 copy x4 is not written
    void pack(int k, const double* xd, bool with_x4 = true) {
        WalkerIn& a = w[k];
        if (!with_x4 && !near_on() && engine_x64() == xd) {
            HIPCHK(launch_pack_walker(xd, a.n_local, a.n_all, nullptr, x4f.p + a.atom_off,
                                      nullptr, x4f.p + total_local + a.ghost_off, stream));
            x4_stale = true;
            return;
        }
        x4_stale = false;
        if (near_on())
            HIPCHK(launch_pack_walker_near(xd, a.n_local, a.n_all, x4.p + a.atom_off, x4f.p + a.atom_off,
                                           x4.p + total_local + a.ghost_off, x4f.p + total_local + a.ghost_off,
                                           xpr.p + a.atom_off, xpr.p + total_local + a.ghost_off, near_dev(), stream));
        else
            HIPCHK(launch_pack_walker(xd, a.n_local, a.n_all, x4.p + a.atom_off, x4f.p + a.atom_off,
                                      x4.p + total_local + a.ghost_off, x4f.p + total_local + a.ghost_off, stream));
    }
};

void mdt_ctx::commit() {
    int64_t tl = 0, tg = 0, tn = 0, ml = 0;
    for (size_t k = 0; k < w.size(); ++k) {
        WalkerIn& a = w[k];
        if (!a.has_box || !a.has_x || !a.has_nl)
            throw Err(MDT_ESTATE, fmt("walker %zu needs box, positions and neighbour lists before the first step", k));
        if ((int64_t)a.off.size() != a.n_local + 1)
            throw Err(MDT_EINVAL, fmt("walker %zu: neighbour list has %zu centers, expected n_local=%lld",
                                      k, a.off.size() - 1, (long long)a.n_local));
        for (int64_t i = 0; i < a.n_local; ++i) {
            if (a.off[i + 1] < a.off[i]) throw Err(MDT_EINVAL, fmt("walker %zu: offsets decrease at atom %lld", k, (long long)i));
            ml = std::max<int64_t>(ml, (int64_t)a.off[i + 1] - a.off[i]);
        }
        a.atom_off = tl; a.ghost_off = tg; a.nnz_off = tn;
        tl += a.n_local; tg += a.n_all - a.n_local; tn += a.nnz;
    }
    if (tl + tg > INT32_MAX || tn > INT32_MAX) throw Err(MDT_ENOTSUP, "more than 2^31 atoms or list entries");
    total_local = tl; total_ghost = tg; total_nnz = tn;
    std::vector<double> hx((size_t)(tl + tg) * 3);
    std::vector<int32_t> hoff((size_t)tl + 1), hidx((size_t)tn), hown((size_t)std::max<int64_t>(tg, 1));
    std::map<std::string, std::vector<uint8_t>> hgroups;
    for (auto& a : w)
        for (auto& kv : a.groups) hgroups[kv.first].assign((size_t)tl, 1);
    any_group = !hgroups.empty();
    gcount.clear();
    gmissing.clear();
    h_tiles.clear();
    tile_begin.assign(w.size() + 1, 0);
    for (size_t k = 0; k < w.size(); ++k) {
        WalkerIn& a = w[k];
        std::memcpy(&hx[3 * a.atom_off], a.x.data(), sizeof(double) * 3 * a.n_local);
        std::memcpy(&hx[3 * (tl + a.ghost_off)], a.x.data() + 3 * a.n_local, sizeof(double) * 3 * (a.n_all - a.n_local));
        for (int64_t i = 0; i < a.n_local; ++i) hoff[a.atom_off + i] = (int32_t)(a.nnz_off + a.off[i]);
        for (int64_t e = 0; e < a.nnz; ++e) {
            int32_t j = a.idx[e];
            if (j < 0 || j >= a.n_all)
                throw Err(MDT_EINVAL, fmt("walker %zu: list entry %lld = %d outside [0,%lld)", k, (long long)e, j, (long long)a.n_all));
            hidx[a.nnz_off + e] = j < a.n_local ? (int32_t)(a.atom_off + j) : (int32_t)(tl + a.ghost_off + (j - a.n_local));
        }
        for (int64_t g = 0; g < a.n_all - a.n_local; ++g) {
            int32_t o = a.owner[g];
            if (o < 0 || o >= a.n_local) throw Err(MDT_EINVAL, fmt("walker %zu: ghost %lld has owner %d outside [0,n_local)", k, (long long)g, o));
            hown[a.ghost_off + g] = (int32_t)(a.atom_off + o);
        }
        for (auto& kv : hgroups) {
            auto it = a.groups.find(kv.first);
            if (it == a.groups.end()) it = a.groups.find(std::string());
            std::vector<double>& cnt = gcount[kv.first];
            cnt.resize(w.size(), 0.0);
            if (it != a.groups.end()) {
                std::memcpy(&kv.second[a.atom_off], it->second.data(), a.n_local);
                double c = 0;
                for (auto v : it->second) c += v ? 1 : 0;
                cnt[k] = c;
            } else {
                cnt[k] = (double)a.n_local;
                if (!gmissing.count(kv.first))
                    gmissing[kv.first] = fmt("walker %zu has no membership bound for it (nor a default group)", k);
            }
        }
        a.tile_off = (int32_t)h_tiles.size();
        for (int64_t s = 0; s < a.n_local; s += kTile) {
            Tile t{};
            t.start = (int32_t)(a.atom_off + s);
            t.count = (int32_t)std::min<int64_t>(kTile, a.n_local - s);
            t.walker = (int32_t)k;
            h_tiles.push_back(t);
        }
        a.ntiles = (int32_t)h_tiles.size() - a.tile_off;
        tile_begin[k + 1] = (int32_t)h_tiles.size();
    }
    hoff[tl] = (int32_t)tn;
    HIPCHK(hipSetDevice(device));
    x.alloc(hx.size()); x.upload(hx.data(), hx.size(), stream);
    off.alloc(hoff.size()); off.upload(hoff.data(), hoff.size(), stream);
    idx.alloc(std::max<size_t>(hidx.size(), 1)); idx.upload(hidx.data(), hidx.size(), stream);
    owner.alloc(hown.size()); owner.upload(hown.data(), tg, stream);
    x4.alloc((size_t)(tl + tg));
    x4f.alloc((size_t)(tl + tg));
    HIPCHK(launch_pack_x4(x.p, x4.p, x4f.p, tl + tg, stream));
    x4_stale = false;
    any_share = false;
    for (auto& a : w) any_share |= !a.share.empty();
    if (any_share) {
        std::vector<uint8_t> hs((size_t)tl, 1);
        for (auto& a : w)
            if (!a.share.empty()) std::memcpy(&hs[a.atom_off], a.share.data(), a.n_local);
        dshare.alloc(hs.size());
        dshare.upload(hs.data(), hs.size(), stream);
    } else {
        dshare.release();
    }
    dgroups.clear();
    for (auto& kv : hgroups) {
        DBuf<uint8_t>& d = dgroups[kv.first];
        d.alloc(kv.second.size());
        d.upload(kv.second.data(), kv.second.size(), stream);
    }
    tiles.alloc(std::max<size_t>(h_tiles.size(), 1)); tiles.upload(h_tiles.data(), h_tiles.size(), stream);
    max_list = ml;
    eown.alloc(std::max<size_t>(hidx.size(), 1));
    revpos.alloc(std::max<size_t>(hidx.size(), 1));
    if (ml <= kMaxList) HIPCHK(launch_topology(geo(), tl, eown.p, revpos.p, stream));
    f.alloc((size_t)std::max<int64_t>(tl, 1) * 3);
    HIPCHK(launch_fill(f.p, tl * 3, 0.0, stream));
    HIPCHK(hipStreamSynchronize(stream));
    committed = true;
    generation++;
}

void mdt_ctx::ensure_near(double rc2) {
    const double rc = std::sqrt(rc2);
    if (near_on() && near_rc >= rc) return;
    const size_t nl = (size_t)std::max<int64_t>(total_local, 1);
    const size_t na = (size_t)std::max<int64_t>(total_local + total_ghost, 1);
    nidx.alloc((size_t)std::max<int64_t>(total_nnz, 1));
    ncnt.alloc(nl); xpr.alloc(na);
    nflag.alloc(2); noverflow.alloc(1); nepoch.alloc(1);
    const int32_t z2[2] = {0, 0};
    const int64_t z64 = 0;
    nflag.upload(z2, 2, stream);
    nepoch.upload(&z64, 1, stream);
    near_rc = rc;
    near_gen = generation;
    // the widest margin above rc (rc/8, rc/16, ...) whose lists all fit kMaxList entries:
    // a narrower margin only means earlier rebuilds
    for (double f = 0.125; ; f *= 0.5) {
        const bool last = f < 1.0 / 1024.0;
        const double rp = last ? rc : rc * (1.0 + f);
        near_rp2 = rp * rp;
        noverflow.upload(z2, 1, stream);
        HIPCHK(launch_prune(geo(), near_dev(), 1, stream));
        int32_t of = 0;
        HIPCHK(hipMemcpyAsync(&of, noverflow.p, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (!of) break;
        if (last) {
            near_rp2 = 0.0;
            throw Err(MDT_ENOTSUP, fmt("an atom has more than %d list entries within the Steinhardt cutoff %.6g; "
                                       "at most %d are supported", kMaxList, rc, kMaxList));
        }
    }
    HIPCHK(launch_topology_near(geo_near(), near_dev(), eown.p, revpos.p, 1, stream));
    HIPCHK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// fix
// ---------------------------------------------------------------------------
namespace {

enum CvKind { CV_Q = 0, CV_LD = 1, CV_SK = 2, CV_DIST = 3, CV_SF = 4, CV_LQ = 5 };

struct EnvH {
    double rc = 0;
    int N = 0;
    std::string group;
    int lmask = 0;
    bool biased = false;
    EnvDev dev{};
    DBuf<float> rec, ql, extra;
    DBuf<int32_t> nsel, cnt;
    DBuf<uint64_t> selmask;
};

struct CvH {
    std::string name;
    CvKind kind = CV_Q;
    int l = 0;
    int env = -1;
    int slot = -1;          // value slot
    int psum = -1;          // partial slot (Q, LD)
    bool fused = false;     // LD: unbiased, accumulated by the force pass of its environment
    bool biased = false;
    // LD
    DBuf<float> lq, bq;
    // SK
    std::vector<double> kvec;   // nk*3
    DBuf<double> kd, krev, apart, amp, sterm, nnorm;
    DBuf<SkK> kpack;
    DBuf<float> ampf;
    DBuf<int32_t> chunk, chunk_begin, natoms, abegin, perm, tbeg;
    bool sk_wide = false;   // tile-relative phases may exceed 128 revolutions
    DBuf<double4> tref;
    DBuf<float4> dr;
    int nchunk = 0, max_atoms = 0;
    // DISTANCE: local indices of the two atoms, per-walker offsets, boxes and (d, r)
    int a1 = -1, a2 = -1;
    DBuf<int64_t> doff;
    DBuf<double> dbox, dvec;
    // STRU_FACTOR: cutoff, q, group; per-atom S_i, per-tile sums, per-walker 1/N_group
    double sf_rc = 8.0, sf_q = 0.0;
    std::string sf_group;
    bool sf_has_sw = false;
    SwFn sf_sw{};
    // chemical weights w_type of STRU_FACTOR ... Chemical and of WEIGHT_CHEM: Chem_map values,
    // else exp(-(type - Chem_ctarget)^2 / (2 Chem_sigma^2)) (CVs_Strufactor_chem.cu:44-66)
    int sf_pair = 0;                // 0: sin(q r)/(q r) s(r^2); 1: s(r^2) (WEIGHT_CHEM)
    bool sf_chem = false;
    double chem_c = 0.0, chem_sigma = 1.0;
    std::map<int, double> chem_map;
    DBuf<double> sf_aw;
    std::string func;       // the SYMBOL function bound to this CAL (AVE unless a SYMBOL says otherwise)
    DBuf<double> sf_S, sf_part, sf_inv;
    DBuf<int32_t> sf_tb;
    // STEINH L: sigma(r), f(LQ_l), eps, group; per-atom state, per-entry pair gradients
    SwFn lq_swr{}, lq_swcv{};
    double lq_eps = 1e-6;
    std::string lq_group;
    DBuf<double> lq_S, lq_q, lq_LQ, lq_ql, lq_K, lq_g, lq_part, lq_sums, lq_ng;
    DBuf<double2> lq_es, lq_hk;
    DBuf<double> lq_H, lq_u;
    DBuf<int32_t> lq_tb, lq_eidx, lq_eown;
    DBuf<int64_t> lq_apos, lq_erev;
};

// STEINH L functions (CVs_Steinhardt_Local.cu:631-690): MEAN_SOLID (alias AVE), NSOLID
// (alias SW_FUNC), FRAC; returns the mode (0, 1, 2) or -1
int lq_mode(const std::string& f) {
    if (f == "MEAN_SOLID" || f == "AVE") return 0;
    if (f == "NSOLID" || f == "SW_FUNC") return 1;
    if (f == "FRAC") return 2;
    return -1;
}

struct SymH {
    std::string name;
    int cv = -1;
    std::string func;
};

}  // namespace

struct mdt_fix {
    mdt_ctx* ctx = nullptr;
    int device = 0;
    // parsed keywords
    double sigma = 0.05, height0 = 0.1, biasf = 10.0, temp = 300.0, boltz = 8.617343e-5;
    int pace = 100, cv_dim = 1, hill_type = 0;
    bool restart = false, wt = false;
    std::string hills_file = "HILLS";
    std::vector<double> sigmas;                 // per-dimension (extension)
    // learned bias (Gaussian_Hill_type 3)
    int64_t mlp_refit = 500, mlp_points = 1000000, mlp_holdout = 65536, mlp_batch = 4096;
    int mlp_epochs = 8, mlp_epochs_first = 40;
    double mlp_lr = 1e-2, mlp_tol = 0.1, mlp_spread = 2.0;
    double mlp_lr_refit = -1.0;   // learning rate of warm-started refits (< 0: MLP_LR / 10)
    int mlp_graph = 1;            // 1: replay a captured training epoch; 0: plain launches
    uint64_t mlp_seed = 20241203;
    int64_t mlp_round = 0;
    std::vector<double> lo, hi;
    std::vector<int> nb;
    std::vector<std::string> dim_expr;
    std::vector<ExprProg> h_progs;
    std::vector<EnvH*> envs;
    std::vector<CvH*> cvs;
    std::vector<SymH> syms;
    std::map<std::string, SwFn> swreg;          // CAL NAME <name> SW_FUNC ... (sw_registry)
    // device state
    bool inited = false;
    uint64_t built_gen = 0;
    uint64_t builds = 0;           // fix_build calls (device buffers may move)
    int nslot_sum = 0;
    DBuf<double> partial, direct, grid, hills, state, wcv, cv_inv, bwork;
    DBuf<int64_t> nhills, step;
    DBuf<int32_t> tile_begin, mlp_ok;
    DBuf<ExprProg> progs;
    DBuf<double> aff;              // [kMaxDim][kMaxSlotCV] affine DIM coefficients per value slot
    DBuf<MlpNet> mlp_live, mlp_stage;
    DBuf<float> mlp_m, mlp_v, mlp_gpart;
    DBuf<int64_t> mlp_fits, mlp_fallbacks, mlp_t;
    std::vector<int64_t> mlp_bound_fit;   // per walker: hills bound at the last fit
    std::vector<int64_t> mlp_roff;        // per walker: next training block to replace
    DBuf<double> mlp_err, mlp_box_lo, mlp_box_hi, mlp_vdata, mlp_blkerr;
    DBuf<float4> mlp_sdata;
    std::vector<int> mlp_has_fit;
    // split reductions (mdt_fix_partials / mdt_fix_resume): the CV sums of this context's
    // share of the atoms leave the step, sums over every share come back
    bool split = false, built_split = false;
    DBuf<double> psums;             // [walker][kMaxSlotCV] raw sums
    DBuf<int32_t> psum_of;          // [kMaxSlotCV] partial slot of each CV (-1: none)
    MlpDev mdev{};
    ReduceDev red{};
    int64_t hills_cap = 0;
    std::vector<int64_t> hills_written;   // per walker, hills already in the HILLS file
    std::vector<int64_t> hills_bound;     // per walker upper bound of deposited hills
    std::vector<FILE*> hfiles;
    // captured runs of k consecutive steps, by k (the per-step state, the step
    // counter included, lives on the device, so steps replay back to back)
    std::map<int, hipGraphExec_t> run_graphs;
    uint64_t graph_gen = 0;
    uint64_t graph_attach = 0;
    // captured training loops of the learned bias, per (walker, first fit)
    std::map<std::pair<int, int>, hipGraphExec_t> mlp_graphs;
    uint64_t mlp_graph_gen = 0;
    // host staging
    double* pin = nullptr;        // pinned scratch for small transfers
    size_t pin_n = 0;
    std::vector<double> h_state;
    bool state_valid = false;

    ~mdt_fix() {
        for (auto& kv : run_graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto& kv : mlp_graphs) (void)hipGraphExecDestroy(kv.second);
        for (FILE* fp : hfiles) if (fp) fclose(fp);
        for (auto* e : envs) delete e;
        for (auto* c : cvs) delete c;
        if (pin) (void)hipHostFree(pin);
    }
    int D() const { return cv_dim; }
    // the bias is a grid: the uniform grid (type 0) or the sparse-hash grid held dense (type 1)
    bool grid_bias() const { return hill_type == 0 || hill_type == 1; }
    int nw() const { return (int)ctx->w.size(); }
};

namespace {

void affine_form(ExprProg& p);

void parse_fix(mdt_fix* fx, int argc, const char* const* argv) {
    int i = 0;
    auto need = [&](int k, const char* what) {
        if (i + k >= argc) throw Err(MDT_EINVAL, fmt("Error: %s", what));
    };
    bool have_dim_kw = false;
    std::map<int, std::pair<std::string, std::vector<double>>> dims;
    while (i < argc) {
        std::string kw = argv[i];
        if (kw == "GAUSSIAN") {
            need(3, "GAUSSIAN command requires 3 arguments: sigma, height, biasf.");
            fx->sigma = parse_num(argv[i + 1], "GAUSSIAN sigma");
            fx->height0 = parse_num(argv[i + 2], "GAUSSIAN height");
            fx->biasf = parse_num(argv[i + 3], "GAUSSIAN biasf");
            i += 4;
        } else if (kw == "PACE") {
            need(1, "PACE command requires 1 argument: integer timesteps.");
            fx->pace = (int)parse_int(argv[i + 1], "PACE");
            i += 2;
        } else if (kw == "RECORD") {
            // accepted for compatibility; the reference opens the file and never writes it
            int a = i + 1;
            while (a < argc) {
                if (!strcmp(argv[a], "FILE_NAME") || !strcmp(argv[a], "REC_PACE")) {
                    if (a + 1 >= argc) throw Err(MDT_EINVAL, fmt("Error: '%s' keyword requires a value", argv[a]));
                    a += 2;
                } else break;
            }
            i = a;
        } else if (kw == "CV_dim") {
            need(1, "CV_dim command requires 1 argument.");
            fx->cv_dim = (int)parse_int(argv[i + 1], "CV_dim");
            if (fx->cv_dim < 1 || fx->cv_dim > kMaxDim)
                throw Err(MDT_ENOTSUP, fmt("CV_dim %d: only 1-%d dimensions are supported", fx->cv_dim, kMaxDim));
            have_dim_kw = true;
            i += 2;
        } else if (kw == "TEMP") {
            need(1, "TEMP command requires 1 argument: target temperature.");
            fx->temp = parse_num(argv[i + 1], "TEMP");
            i += 2;
        } else if (kw == "BOLTZ") {
            need(1, "BOLTZ requires 1 argument (Boltzmann constant in energy units per K).");
            fx->boltz = parse_num(argv[i + 1], "BOLTZ");
            i += 2;
        } else if (kw == "HILLS_FILE") {
            need(1, "HILLS_FILE requires a path (NONE disables the file).");
            fx->hills_file = argv[i + 1];
            i += 2;
        } else if (kw.rfind("MLP_", 0) == 0) {
            // learned-bias options (Gaussian_Hill_type 3)
            need(1, "MLP_* options take one value");
            const char* v = argv[i + 1];
            if (kw == "MLP_REFIT") fx->mlp_refit = parse_int(v, kw.c_str());
            else if (kw == "MLP_POINTS") fx->mlp_points = parse_int(v, kw.c_str());
            else if (kw == "MLP_HOLDOUT") fx->mlp_holdout = parse_int(v, kw.c_str());
            else if (kw == "MLP_BATCH") fx->mlp_batch = parse_int(v, kw.c_str());
            else if (kw == "MLP_EPOCHS") fx->mlp_epochs = (int)parse_int(v, kw.c_str());
            else if (kw == "MLP_EPOCHS_FIRST") fx->mlp_epochs_first = (int)parse_int(v, kw.c_str());
            else if (kw == "MLP_LR") fx->mlp_lr = parse_num(v, kw.c_str());
            else if (kw == "MLP_LR_REFIT") {
                fx->mlp_lr_refit = parse_num(v, kw.c_str());
                if (!(fx->mlp_lr_refit >= 0)) throw Err(MDT_EINVAL, fmt("%s: value out of range", kw.c_str()));
            }
            else if (kw == "MLP_TOL") fx->mlp_tol = parse_num(v, kw.c_str());
            else if (kw == "MLP_SPREAD") fx->mlp_spread = parse_num(v, kw.c_str());
            else if (kw == "MLP_SEED") fx->mlp_seed = (uint64_t)parse_int(v, kw.c_str());
            else if (kw == "MLP_GRAPH") fx->mlp_graph = parse_int(v, kw.c_str()) != 0 ? 1 : 0;
            else throw Err(MDT_EINVAL, fmt("Error: Unknown keyword in fix metadynamics command: %s", kw.c_str()));
            if (fx->mlp_refit < 1 || fx->mlp_points < 64 || fx->mlp_holdout < 64 || fx->mlp_batch < 64 ||
                fx->mlp_epochs < 0 || fx->mlp_epochs_first < 0 || !(fx->mlp_lr >= 0) || !(fx->mlp_tol >= 0))
                throw Err(MDT_EINVAL, fmt("%s: value out of range", kw.c_str()));
            i += 2;
        } else if (kw == "SIGMAS") {
            // per-dimension hill widths for the hills-sum and learned biases: the numbers
            // that follow (one per dimension, checked against CV_dim after parsing)
            fx->sigmas.clear();
            int k = i + 1;
            while (k < argc && fx->sigmas.size() < (size_t)kMaxDim) {
                char* end = nullptr;
                double v = strtod(argv[k], &end);
                if (!argv[k][0] || (end && *end)) break;
                fx->sigmas.push_back(v);
                ++k;
            }
            if (fx->sigmas.empty()) throw Err(MDT_EINVAL, "SIGMAS requires one value per dimension.");
            i = k;
        } else if (kw == "CAL") {
            need(3, "CAL requires NAME <name> <type> ...");
            if (strcmp(argv[i + 1], "NAME")) throw Err(MDT_EINVAL, "Error: CAL requires NAME keyword.");
            std::string name = argv[i + 2], type = argv[i + 3];
            if (type == "SW_FUNC") {
                if (fx->swreg.count(name)) throw Err(MDT_EINVAL, fmt("Error: CAL name '%s' is duplicated.", name.c_str()));
            } else {
                for (auto* c : fx->cvs)
                    if (c->name == name) throw Err(MDT_EINVAL, fmt("Error: CAL name '%s' is duplicated.", name.c_str()));
            }
            i += 4;
            // <type> [r_0 v] [d_0 v] [alpha v] [n i] [m i] with the reference's defaults per
            // type (include/utils/zqc_switch_function.h:85-151); i at the type on entry
            auto parse_swspec = [&]() {
                if (i >= argc) throw Err(MDT_EINVAL, "Error: SW_FUNC command requires a type (LINE, STEP, FERMI, TANH, RATIONAL).");
                std::string t = argv[i];
                SwFn w{};
                if (t == "STEP") { w.type = 1; w.r0 = 1.0; }
                else if (t == "FERMI") { w.type = 2; w.r0 = 1.0; w.alpha = 20.0; }
                else if (t == "TANH") { w.type = 3; w.r0 = 1.0; w.alpha = 20.0; }
                else if (t == "RATIONAL") { w.type = 4; w.r0 = 1.25; w.d0 = 0.0; w.n = 6; w.m = 12; }
                else throw Err(MDT_EINVAL, "Error: Unknown SW_func type. Choose from FERMI, TANH, RATIONAL.");
                i += 1;
                while (i < argc) {
                    std::string p2 = argv[i];
                    if (p2 == "r_0") { need(1, "Error: 'r_0' requires a numeric value"); w.r0 = parse_num(argv[i + 1], "r_0"); i += 2; }
                    else if (p2 == "d_0") { need(1, "Error: 'd_0' requires a numeric value"); w.d0 = parse_num(argv[i + 1], "d_0"); i += 2; }
                    else if (p2 == "alpha") { need(1, "Error: 'alpha' requires a numeric value"); w.alpha = parse_num(argv[i + 1], "alpha"); i += 2; }
                    else if (p2 == "n") { need(1, "Error: 'n' requires an integer value"); w.n = (int)parse_int(argv[i + 1], "n"); i += 2; }
                    else if (p2 == "m") { need(1, "Error: 'm' requires an integer value"); w.m = (int)parse_int(argv[i + 1], "m"); i += 2; }
                    else break;
                }
                if (w.type == 4 && !(w.r0 != 0.0)) throw Err(MDT_EINVAL, "RATIONAL r_0 must be nonzero");
                return w;
            };
            // the chemical options at argv[i]; false when argv[i] is none of them
            auto parse_chem = [&](CvH* c) {
                const std::string k2 = argv[i];
                if (k2 == "Chemical") { c->sf_chem = true; i += 1; return true; }
                if (k2 == "Chem_ctarget") {
                    need(1, "Error: 'Chem_ctarget' keyword requires a number");
                    c->chem_c = parse_num(argv[i + 1], "Chem_ctarget"); i += 2; return true;
                }
                if (k2 == "Chem_sigma") {
                    need(1, "Error: 'sigma' keyword requires a number");
                    c->chem_sigma = parse_num(argv[i + 1], "Chem_sigma"); i += 2; return true;
                }
                if (k2 != "Chem_map") return false;
                i += 1;
                // (type weight) pairs: "(t,w)" or "(t w)" as one word, or "(t" "w)" as two
                while (i < argc && argv[i][0] == '(') {
                    std::string a = argv[i];
                    std::string body;
                    if (a.size() > 1 && a.back() == ')') {
                        body = a.substr(1, a.size() - 2);
                        i += 1;
                    } else {
                        if (i + 1 >= argc) throw Err(MDT_EINVAL, fmt("Error: Invalid Chem_map format near %s", argv[i]));
                        const std::string b = argv[i + 1];
                        if (b.empty() || b.back() != ')')
                            throw Err(MDT_EINVAL, "Error: Chem_map pairs must be formatted as (type weight)");
                        body = a.substr(1) + " " + b.substr(0, b.size() - 1);
                        i += 2;
                    }
                    const size_t sep = body.find_first_of(" ,");
                    if (sep == std::string::npos)
                        throw Err(MDT_EINVAL, "Error: Chem_map pairs must be formatted as (type weight)");
                    const std::string ts = body.substr(0, sep), wsv = body.substr(sep + 1);
                    c->chem_map[(int)parse_int(ts.c_str(), "Chem_map type")] = parse_num(wsv.c_str(), "Chem_map weight");
                }
                return true;
            };
            auto find_sw = [&](const char* nm) {
                auto it = fx->swreg.find(nm);
                if (it == fx->swreg.end()) throw Err(MDT_EINVAL, fmt("Error: SwitchFunction named %s not found in registry!", nm));
                return it->second;
            };
            if (type == "STEINH") {
                if (i + 2 >= argc)
                    throw Err(MDT_EINVAL, "Error: STEINH command requires \"STEINH <Q/L/LD> <3/4/6> <group> \".");
                std::string qt = argv[i];
                int l = (int)parse_int(argv[i + 1], "STEINH order");
                std::string group = argv[i + 2];
                i += 3;
                if (l != 3 && l != 4 && l != 6 && l != 8 && l != 12)
                    throw Err(MDT_EINVAL, "Error: Steinhardt order L must be 3, 4, 6, 8, or 12.");
                if (l == 8 || l == 12) throw Err(MDT_ENOTSUP, fmt("STEINH order %d is accepted by the reference parser but has no implementation there either", l));
                double rc = 4.0, eps = 1e-6;
                int N = 12;
                SwFn swr{}, swcv{};          // LINE unless named (get_default_step is LINE)
                while (i < argc) {
                    std::string k2 = argv[i];
                    if (k2 == "cutoff_r") { need(1, "'cutoff_r' keyword requires a value"); rc = parse_num(argv[i + 1], "cutoff_r"); i += 2; }
                    else if (k2 == "cutoff_Natoms") { need(1, "'cutoff_Natoms' keyword requires an integer"); N = (int)parse_int(argv[i + 1], "cutoff_Natoms"); i += 2; }
                    else if (k2 == "cutoff_eps") { need(1, "'cutoff_eps' keyword requires a value"); eps = parse_num(argv[i + 1], "cutoff_eps"); i += 2; }
                    else if (k2 == "d_block_size") { need(1, "'d_block_size' keyword requires an integer");
                        if (parse_int(argv[i + 1], "d_block_size") <= 0) throw Err(MDT_EINVAL, "Error: 'd_block_size' must be > 0");
                        i += 2; }
                    else if (k2 == "SW_FUNC_r") { need(1, "Error: 'SW_FUNC_r' keyword requires a value"); swr = find_sw(argv[i + 1]); i += 2; }
                    else if (k2 == "SW_FUNC_cv") { need(1, "Error: 'SW_FUNC_cv' keyword requires a value"); swcv = find_sw(argv[i + 1]); i += 2; }
                    else break;
                }
                if (qt == "L") {
                    // sigma applies to every list entry; cutoff_r and cutoff_Natoms only size the
                    // reference's own neighbour request (CVs_Steinhardt.cu:125-148)
                    if (!(eps >= 0)) throw Err(MDT_EINVAL, "cutoff_eps must be >= 0");
                    CvH* c = new CvH();
                    c->name = name; c->kind = CV_LQ; c->l = l; c->lq_group = group;
                    c->lq_swr = swr; c->lq_swcv = swcv; c->lq_eps = eps;
                    fx->cvs.push_back(c);
                    continue;
                }
                if (qt != "Q" && qt != "LD") throw Err(MDT_EINVAL, "Error: Steinhardt type must be 'Q', 'L' or 'LD'.");
                if (N < 1 || N > kMaxSlots) throw Err(MDT_ENOTSUP, fmt("cutoff_Natoms %d outside [1,%d]", N, kMaxSlots));
                if (!(rc > 0)) throw Err(MDT_EINVAL, "cutoff_r must be positive");
                int ei = -1;
                for (size_t k = 0; k < fx->envs.size(); ++k)
                    if (fx->envs[k]->rc == rc && fx->envs[k]->N == N && fx->envs[k]->group == group) ei = (int)k;
                if (ei < 0) {
                    EnvH* e = new EnvH();
                    e->rc = rc; e->N = N; e->group = group;
                    fx->envs.push_back(e);
                    ei = (int)fx->envs.size() - 1;
                }
                fx->envs[ei]->lmask |= l == 3 ? kL3 : l == 4 ? kL4 : kL6;
                CvH* c = new CvH();
                c->name = name; c->kind = qt == "Q" ? CV_Q : CV_LD; c->l = l; c->env = ei;
                fx->cvs.push_back(c);
            } else if (type == "SK") {
                // SK <group> [FCC111 a] [KVEC kx ky kz]... [NRAND n KMIN a KMAX b SEED s]
                if (i >= argc) throw Err(MDT_EINVAL, "SK requires a group");
                std::string group = argv[i++];
                if (group != "all") throw Err(MDT_ENOTSUP, "SK supports group all only");
                CvH* c = new CvH();
                c->name = name; c->kind = CV_SK;
                long long nrand = 0, seed = 12345;
                double kmin = 1.5, kmax = 3.5;
                while (i < argc) {
                    std::string k2 = argv[i];
                    if (k2 == "FCC111") {
                        need(1, "FCC111 requires the lattice constant");
                        double a = parse_num(argv[i + 1], "FCC111");
                        double k0 = 2.0 * 3.14159265358979323846 / a;
                        for (int sx = -1; sx <= 1; sx += 2)
                            for (int sy = -1; sy <= 1; sy += 2)
                                for (int sz = -1; sz <= 1; sz += 2) {
                                    c->kvec.push_back(sx * k0); c->kvec.push_back(sy * k0); c->kvec.push_back(sz * k0);
                                }
                        i += 2;
                    } else if (k2 == "KVEC") {
                        need(3, "KVEC requires 3 components");
                        for (int d = 0; d < 3; ++d) c->kvec.push_back(parse_num(argv[i + 1 + d], "KVEC"));
                        i += 4;
                    } else if (k2 == "NRAND") { need(1, "NRAND requires a count"); nrand = parse_int(argv[i + 1], "NRAND"); i += 2; }
                    else if (k2 == "KMIN") { need(1, "KMIN requires a value"); kmin = parse_num(argv[i + 1], "KMIN"); i += 2; }
                    else if (k2 == "KMAX") { need(1, "KMAX requires a value"); kmax = parse_num(argv[i + 1], "KMAX"); i += 2; }
                    else if (k2 == "SEED") { need(1, "SEED requires an integer"); seed = parse_int(argv[i + 1], "SEED"); i += 2; }
                    else break;
                }
                // seeded isotropic random k: |k| ~ U[kmin,kmax], direction uniform on the sphere
                uint64_t st = (uint64_t)seed;
                auto next = [&]() {
                    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
                    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                    z ^= z >> 31;
                    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
                };
                for (long long q = 0; q < nrand; ++q) {
                    double km = kmin + (kmax - kmin) * next();
                    double cz = 2.0 * next() - 1.0, ph = 2.0 * 3.14159265358979323846 * next();
                    double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
                    c->kvec.push_back(km * sz * std::cos(ph));
                    c->kvec.push_back(km * sz * std::sin(ph));
                    c->kvec.push_back(km * cz);
                }
                if (c->kvec.empty()) { delete c; throw Err(MDT_EINVAL, "SK needs at least one k vector"); }
                fx->cvs.push_back(c);
            } else if (type == "SW_FUNC") {
                // CAL NAME <name> SW_FUNC <type> ... (fix_crystallize.cu:144-151)
                fx->swreg[name] = parse_swspec();
            } else if (type == "DISTANCE") {
                // DISTANCE id1 id2 (CVs_DISTANCE.cu:17-30): 1-based atom IDs, the same in every walker
                if (i + 1 >= argc) throw Err(MDT_EINVAL, "Error: DISTANCE command requires 2 atom IDs.");
                const long long id1 = parse_int(argv[i], "DISTANCE"), id2 = parse_int(argv[i + 1], "DISTANCE");
                if (id1 < 1 || id2 < 1) throw Err(MDT_EINVAL, "DISTANCE atom IDs start at 1");
                if (id1 == id2) throw Err(MDT_EINVAL, "DISTANCE needs two different atoms");
                i += 2;
                CvH* c = new CvH();
                c->name = name; c->kind = CV_DIST; c->a1 = (int)(id1 - 1); c->a2 = (int)(id2 - 1);
                fx->cvs.push_back(c);
            } else if (type == "STRU_FACTOR") {
                // STRU_FACTOR <group> [cutoff_r r] [q_factor q] [d_block_size b] [SW_func ...] [Chemical]
                // [Chem_ctarget c] [Chem_sigma s] [Chem_map (type weight)...] (CVs_Strufactor.cu:34-190)
                if (i >= argc) throw Err(MDT_EINVAL, "Error: STRU_FACTOR command requires at least \"STRU_FACTOR <group> \".");
                CvH* c = new CvH();
                c->name = name; c->kind = CV_SF; c->sf_group = argv[i++];
                bool have_q = false;
                try {
                    while (i < argc) {
                        std::string k2 = argv[i];
                        if (k2 == "cutoff_r") {
                            need(1, "Error: 'cutoff_r' keyword requires a value");
                            c->sf_rc = parse_num(argv[i + 1], "cutoff_r"); i += 2;
                        } else if (k2 == "q_factor") {
                            need(1, "Error: 'q_factor' keyword requires an integer");
                            c->sf_q = parse_num(argv[i + 1], "q_factor"); have_q = true; i += 2;
                            if (!(c->sf_q > 0)) throw Err(MDT_EINVAL, "Error: 'q_factor' must be > 0");
                        } else if (k2 == "d_block_size") {
                            need(1, "Error: 'd_block_size' keyword requires an integer");
                            if (parse_int(argv[i + 1], "d_block_size") <= 0) throw Err(MDT_EINVAL, "Error: 'd_block_size' must be > 0");
                            i += 2;
                        } else if (k2 == "SW_func") {
                            // SW_func <type> [r_0 v] [d_0 v] [alpha v] [n i] [m i] (CVs_Strufactor.cu:73-124)
                            need(1, "Error: 'SW_func' keyword requires a type (FERMI, TANH, RATIONAL)");
                            i += 1;
                            c->sf_sw = parse_swspec();
                            c->sf_has_sw = true;
                        } else if (parse_chem(c)) {
                            // Chemical, Chem_ctarget, Chem_sigma, Chem_map
                        } else {
                            break;
                        }
                    }
                    if (!have_q) throw Err(MDT_EINVAL, "STRU_FACTOR needs q_factor (the reference leaves it unset)");
                    if (!(c->sf_rc > 0)) throw Err(MDT_EINVAL, "STRU_FACTOR cutoff_r must be positive");
                } catch (...) { delete c; throw; }
                fx->cvs.push_back(c);
            } else if (type == "WEIGHT_CHEM") {
                // WEIGHT_CHEM <group> [cutoff_r r] [d_block_size b] [SW_func ...] [Chem_ctarget c]
                // [Chem_sigma s] [Chem_map (type weight)...] (CV_Weighted_Chemical_Pair_Order.cu:32-172):
                // per atom 1 + w_i sum_j w_j s(r^2), always chemically weighted
                if (i >= argc) throw Err(MDT_EINVAL, "Error: CHEM_PAIR command requires at least \"CHEM_PAIR <group> \".");
                CvH* c = new CvH();
                c->name = name; c->kind = CV_SF; c->sf_group = argv[i++];
                c->sf_pair = 1; c->sf_chem = true;
                try {
                    while (i < argc) {
                        std::string k2 = argv[i];
                        if (k2 == "cutoff_r") {
                            need(1, "Error: 'cutoff_r' keyword requires a value");
                            c->sf_rc = parse_num(argv[i + 1], "cutoff_r"); i += 2;
                        } else if (k2 == "d_block_size") {
                            need(1, "Error: 'd_block_size' keyword requires an integer");
                            if (parse_int(argv[i + 1], "d_block_size") <= 0) throw Err(MDT_EINVAL, "Error: 'd_block_size' must be > 0");
                            i += 2;
                        } else if (k2 == "SW_func") {
                            need(1, "Error: 'SW_func' keyword requires a type (FERMI, TANH, RATIONAL)");
                            i += 1;
                            c->sf_sw = parse_swspec();
                            c->sf_has_sw = true;
                        } else if (k2 != "Chemical" && parse_chem(c)) {
                            // Chem_ctarget, Chem_sigma, Chem_map
                        } else {
                            break;
                        }
                    }
                    if (!(c->sf_rc > 0)) throw Err(MDT_EINVAL, "WEIGHT_CHEM cutoff_r must be positive");
                } catch (...) { delete c; throw; }
                fx->cvs.push_back(c);
            } else {
                throw Err(MDT_EINVAL, fmt("Metad-toy Error: Unknown or unregistered object type '%s' specified in CAL '%s'.",
                                          type.c_str(), name.c_str()));
            }
        } else if (kw == "SYMBOL" || kw == "SIMBOL") {
            need(2, "SYMBOL requires 2 arguments: SYMBOL <symbol_name> <cv_method>.");
            std::string target = argv[i + 2];
            size_t dot = target.find('.');
            SymH s;
            s.name = argv[i + 1];
            std::string cname = dot == std::string::npos ? target : target.substr(0, dot);
            s.func = dot == std::string::npos ? "value" : target.substr(dot + 1);
            for (size_t k = 0; k < fx->cvs.size(); ++k) if (fx->cvs[k]->name == cname) s.cv = (int)k;
            if (s.cv < 0) throw Err(MDT_EINVAL, fmt("SYMBOL %s: CAL '%s' is not defined", s.name.c_str(), cname.c_str()));
            // a DISTANCE CV has one value whatever the function is called (CVs_DISTANCE.cu:125-128);
            // STRU_FACTOR has AVE and COUNT (CVs_Strufactor.cu:616-636)
            {
                CvH* cc = fx->cvs[s.cv];
                if (cc->kind == CV_LQ) {
                    if (lq_mode(s.func) < 0)
                        throw Err(MDT_EINVAL, fmt("Unknown STEIN_LocalQL func %s. Use MEAN_SOLID (alias AVE), NSOLID (alias SW_FUNC), or FRAC.", s.func.c_str()));
                    if (!cc->func.empty() && lq_mode(cc->func) != lq_mode(s.func))
                        throw Err(MDT_ENOTSUP, fmt("CAL %s is bound as %s and %s; define one CAL per function", cc->name.c_str(),
                                                   cc->func.c_str(), s.func.c_str()));
                    cc->func = s.func;
                }
                const bool ok = s.func == "AVE" || cc->kind == CV_DIST || (cc->kind == CV_SF && s.func == "COUNT") || cc->kind == CV_LQ;
                if (!ok) throw Err(MDT_EINVAL, fmt("We can't find the func %s.", s.func.c_str()));
                if (cc->kind == CV_SF) {
                    if (!cc->func.empty() && cc->func != s.func)
                        throw Err(MDT_ENOTSUP, fmt("CAL %s is bound as %s and %s; define one CAL per function", cc->name.c_str(),
                                                   cc->func.c_str(), s.func.c_str()));
                    if (s.func == "COUNT" && !cc->sf_has_sw)
                        throw Err(MDT_EINVAL, fmt("%s.COUNT needs SW_func on the CAL", cc->name.c_str()));
                    cc->func = s.func;
                }
            }
            for (auto& o : fx->syms) if (o.name == s.name) throw Err(MDT_EINVAL, fmt("SYMBOL %s defined twice", s.name.c_str()));
            if ((int)fx->syms.size() >= kMaxSym) throw Err(MDT_ENOTSUP, fmt("at most %d symbols", kMaxSym));
            fx->syms.push_back(s);
            i += 3;
        } else if (kw == "DIM") {
            need(5, "DIM command requires 5 arguments: index, lower, upper, bins, cv_expr.");
            int d = (int)parse_int(argv[i + 1], "DIM index") - 1;
            if (d < 0 || d >= fx->cv_dim) throw Err(MDT_EINVAL, "Error: DIM index out of range or unsupported dimension.");
            double lo = parse_num(argv[i + 2], "DIM lower"), hi = parse_num(argv[i + 3], "DIM upper");
            int nbins = (int)parse_int(argv[i + 4], "DIM bins");
            dims[d] = {argv[i + 5], {lo, hi, (double)nbins}};
            i += 6;
        } else if (kw == "Gaussian_Hill_type") {
            need(1, "Gaussian_Hill_type requires 1 argument.");
            fx->hill_type = (int)parse_int(argv[i + 1], "Gaussian_Hill_type");
            i += 2;
        } else if (kw == "METAD_RESTART") {
            need(1, "METAD_RESTART command requires 1 argument: 0 or 1.");
            fx->restart = parse_int(argv[i + 1], "METAD_RESTART") != 0;
            i += 2;
        } else if (kw == "WT") {
            need(1, "WT requires 1 argument: 0 or 1.");
            fx->wt = parse_int(argv[i + 1], "WT") != 0;
            i += 2;
        } else {
            throw Err(MDT_EINVAL, fmt("Error: Unknown keyword in fix metadynamics command: %s", argv[i]));
        }
    }
    (void)have_dim_kw;
    for (int d = 0; d < fx->cv_dim; ++d) {
        auto it = dims.find(d);
        if (it == dims.end()) throw Err(MDT_EINVAL, fmt("CV_dim is %d but DIM %d is missing", fx->cv_dim, d + 1));
        fx->dim_expr.push_back(it->second.first);
        fx->lo.push_back(it->second.second[0]);
        fx->hi.push_back(it->second.second[1]);
        fx->nb.push_back((int)it->second.second[2]);
        if (!(fx->hi[d] > fx->lo[d])) throw Err(MDT_EINVAL, fmt("DIM %d: upper bound must exceed lower", d + 1));
    }
    if (fx->hill_type < 0 || fx->hill_type > 3)
        throw Err(MDT_EINVAL, fmt("Gaussian_Hill_type %d unknown", fx->hill_type));
    if (fx->grid_bias()) {
        // type 0: the uniform grid (GH_t0_uniformGrid, 1-D and 2-D); type 1: the
        // sparse-hash grid (GH_t1_sparseHash, up to 3-D) held as a dense device grid
        if (fx->hill_type == 0 && fx->cv_dim > 2)
            throw Err(MDT_ENOTSUP, "the uniform grid supports 1 or 2 dimensions; use Gaussian_Hill_type 1, 2 or 3 for 3-D");
        double bins = 1.0;
        for (int d = 0; d < fx->cv_dim; ++d) {
            if (fx->nb[d] < 4) throw Err(MDT_EINVAL, "DIM bins must be >= 4 for Catmull-Rom interpolation");
            bins *= fx->nb[d];
        }
        if (bins > (double)kMaxGridBins) throw Err(MDT_ENOTSUP, fmt("a bias grid of %.0f bins exceeds %lld per walker", bins, (long long)kMaxGridBins));
    }
    if (fx->hill_type == 3 && fx->cv_dim != 3) throw Err(MDT_ENOTSUP, "the learned bias is built for CV_dim 3");
    if (fx->sigmas.empty()) fx->sigmas.assign(fx->cv_dim, fx->sigma);
    if ((int)fx->sigmas.size() != fx->cv_dim)
        throw Err(MDT_EINVAL, fmt("SIGMAS has %zu values but CV_dim is %d", fx->sigmas.size(), fx->cv_dim));
    for (double v : fx->sigmas) if (!(v > 0)) throw Err(MDT_EINVAL, "SIGMAS must be positive");
    if (fx->grid_bias())
        for (double v : fx->sigmas)
            if (v != fx->sigma)
                throw Err(MDT_ENOTSUP, "SIGMAS (per-dimension widths) applies to Gaussian_Hill_type 2 and 3; "
                                       "the grids deposit isotropic hills of the GAUSSIAN sigma");
    if (fx->syms.empty()) throw Err(MDT_EINVAL, "fix metad needs at least one SYMBOL");
    if ((int)fx->cvs.size() > kMaxSlotCV) throw Err(MDT_ENOTSUP, fmt("at most %d CALs", kMaxSlotCV));
    // compile the DIM expressions against the symbol table (DimSet.cu:54-66)
    std::vector<std::string> names;
    for (auto& s : fx->syms) names.push_back(s.name);
    fx->h_progs.clear();
    for (int d = 0; d < fx->cv_dim; ++d) {
        ExprProg pg = ExprCompiler(fx->dim_expr[d], names).compile();
        affine_form(pg);
        fx->h_progs.push_back(pg);
    }
}

// Affine reading of a program: when every operation keeps the value of the form
// c0 + sum_q c_q v_q the program is evaluated on the device as that sum.
void affine_form(ExprProg& p) {
    struct Aff { double c0; double c[kMaxSym]; bool cst; };
    std::vector<Aff> st;
    p.linear = 0;
    auto cst = [](const Aff& a) { return a.cst; };
    for (int k = 0; k < p.n; ++k) {
        const int op = p.op[k];
        if (op == OP_CONST || op == OP_SYM) {
            Aff a{};
            if (op == OP_CONST) { a.c0 = p.c[k]; a.cst = true; }
            else { a.c[(int)p.c[k]] = 1.0; a.cst = false; }
            st.push_back(a);
            continue;
        }
        if (op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV) {
            Aff b = st.back(); st.pop_back();
            Aff& a = st.back();
            if (op == OP_ADD || op == OP_SUB) {
                const double sg = op == OP_ADD ? 1.0 : -1.0;
                a.c0 += sg * b.c0;
                for (int q = 0; q < kMaxSym; ++q) a.c[q] += sg * b.c[q];
                a.cst = a.cst && b.cst;
            } else if (op == OP_MUL && (cst(a) || cst(b))) {
                const Aff& v = cst(a) ? b : a;
                const double f = cst(a) ? a.c0 : b.c0;
                Aff r{};
                r.c0 = v.c0 * f;
                for (int q = 0; q < kMaxSym; ++q) r.c[q] = v.c[q] * f;
                r.cst = v.cst;
                a = r;
            } else if (op == OP_DIV && cst(b) && b.c0 != 0.0) {
                a.c0 /= b.c0;
                for (int q = 0; q < kMaxSym; ++q) a.c[q] /= b.c0;
            } else {
                return;   // not affine
            }
            continue;
        }
        if (op == OP_NEG) {
            Aff& a = st.back();
            a.c0 = -a.c0;
            for (int q = 0; q < kMaxSym; ++q) a.c[q] = -a.c[q];
            continue;
        }
        return;           // functions, powers, min/max: use the interpreter
    }
    if (st.size() != 1) return;
    p.linear = 1;
    p.lin0 = st[0].c0;
    for (int q = 0; q < kMaxSym; ++q) p.lin[q] = st[0].c[q];
}

bool prog_uses(const ExprProg& p, int sym) {
    for (int k = 0; k < p.n; ++k) if (p.op[k] == OP_SYM && (int)p.c[k] == sym) return true;
    return false;
}

void open_hills(mdt_fix* fx);
void ensure_hills(mdt_fix* fx, int64_t extra);

// Position of the cell (X[0], X[1], X[2]) (each below 2^b) along the 3-D Hilbert curve,
// by Skilling's transform to the transposed index ("Programming the Hilbert curve",
// AIP Conf. Proc. 707, 381 (2004)) and the interleaving of its bits.
static uint64_t hilbert_key(uint32_t X[3], int b) {
    const uint32_t M = 1u << (b - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int j = b - 1; j >= 0; --j)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> j) & 1u);
    return key;
}

// A walker's local atoms (0 .. n_local-1) ordered along a Hilbert curve over its box
// cut into 2^b cells per dimension, the cells about kSkCell length units long (ties in
// atom order), so that every run of 256 consecutive atoms of the order lies within a
// few neighbouring cells: the structure-factor passes take their phases relative to
// the first atom of each run.  The grid covers the box exactly, so the curve never
// leaves it between two occupied cells.
constexpr double kSkCell = 8.0;
static std::vector<int32_t> hilbert_order(const WalkerIn& a) {
    const int64_t n = a.n_local;
    int b = 1;
    for (int d = 0; d < 3; ++d) {
        const double L = a.hi[d] - a.lo[d];
        while (b < 20 && (double)(1 << b) * kSkCell < L) ++b;
    }
    const int nc = 1 << b;
    std::vector<std::pair<uint64_t, int32_t>> key((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t X[3];
        for (int d = 0; d < 3; ++d) {
            const double L = a.hi[d] - a.lo[d];
            const double u = L > 0.0 ? (a.x[3 * i + d] - a.lo[d]) / L : 0.0;
            const int c = (int)std::floor(u * nc);
            X[d] = (uint32_t)std::min(std::max(c, 0), nc - 1);
        }
        key[(size_t)i] = {hilbert_key(X, b), (int32_t)i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> ord((size_t)n);
    for (int64_t i = 0; i < n; ++i) ord[(size_t)i] = key[(size_t)i].second;
    return ord;
}

void fix_build(mdt_fix* fx) {
    mdt_ctx* ctx = fx->ctx;
    if (!ctx->committed) ctx->commit();
    const int nw = fx->nw();
    const int64_t NT = ctx->total_local;
    hipStream_t st = ctx->stream;
    const std::vector<ExprProg>& progs = fx->h_progs;
    // slots
    int nsum = 0;
    std::vector<int> walias(kMaxSlotCV, -1);
    for (size_t c = 0; c < fx->cvs.size(); ++c) fx->cvs[c]->slot = (int)c;
    for (auto* c : fx->cvs) c->biased = false;
    for (size_t q = 0; q < fx->syms.size(); ++q) {
        bool used = false;
        for (auto& p : progs) used |= prog_uses(p, (int)q);
        if (used) fx->cvs[fx->syms[q].cv]->biased = true;
    }
    // group counts per walker
    auto group_counts = [&](const std::string& name) {
        std::vector<double> v(nw);
        for (int k = 0; k < nw; ++k) v[k] = ctx->group_count(name, k);
        return v;
    };
    // lists longer than kMaxList: the Steinhardt kernels walk near lists cut just above the
    // largest slack-widened Steinhardt cutoff
    if (!fx->envs.empty() && ctx->max_list > kMaxList) {
        double rc2 = 0.0;
        for (auto* e : fx->envs) rc2 = std::max(rc2, e->rc * e->rc * (1.0 + kRcSlack));
        ctx->ensure_near(rc2);
    }
    // environments
    for (auto* e : fx->envs) {
        const int lm = e->lmask;
        const int RS = rec_stride(lm), NL = rec_nl(lm);
        // the force pass addresses records and list words with 32-bit byte offsets
        if ((uint64_t)NT * (uint64_t)RS * 4u >= (1ull << 31))
            throw Err(MDT_ENOTSUP, fmt("%lld local atoms in one context: at most %lld are supported",
                                       (long long)NT, (long long)((1ull << 31) / (4u * (uint64_t)RS)) - 1));
        if ((uint64_t)ctx->total_nnz * 4u >= (1ull << 31))
            throw Err(MDT_ENOTSUP, fmt("%lld list entries in one context: at most %lld are supported",
                                       (long long)ctx->total_nnz, (long long)((1ull << 29) - 1)));
        e->rec.alloc((size_t)NT * RS);
        e->ql.alloc((size_t)NT * NL);
        e->nsel.alloc(NT); e->cnt.alloc(NT); e->selmask.alloc((size_t)NT * 4);
        EnvDev& d = e->dev;
        d = EnvDev{};
        d.rc2 = e->rc * e->rc; d.N = e->N; d.lmask = lm;
        d.rcs = (float)std::sqrt(d.rc2 * (1.0 + kRcSlack));
        d.rec = e->rec.p; d.ql = e->ql.p; d.nsel = e->nsel.p; d.cnt_rc = e->cnt.p; d.selmask = e->selmask.p;
        int li = 0;
        for (int l : {3, 4, 6}) {
            int bit = l == 3 ? kL3 : l == 4 ? kL4 : kL6;
            if (!(lm & bit)) continue;
            d.slot_ql[li] = nsum++;
            d.wslot[li] = -1;
            // CALs with identical definitions share this degree: its force pass takes the
            // sum of their weights, kept at the first one's combined slot
            int rep_slot = -1;
            for (auto* c : fx->cvs)
                if (c->kind == CV_Q && fx->envs[c->env] == e && c->l == l) {
                    c->psum = d.slot_ql[li];
                    if (c->biased) {
                        if (rep_slot < 0) rep_slot = c->slot;
                        walias[c->slot] = rep_slot;
                        d.wslot[li] = kMaxSlotCV + rep_slot;
                    }
                }
            li++;
        }
        e->biased = false;
        for (int k = 0; k < NL; ++k) e->biased |= d.wslot[k] >= 0;
        bool ld_biased = false;
        for (auto* c : fx->cvs) if (c->kind == CV_LD && fx->envs[c->env] == e && c->biased) ld_biased = true;
        if (ld_biased) {
            e->extra.alloc((size_t)NT * RS);
            HIPCHK(launch_fill_f(e->extra.p, (int64_t)NT * RS, 0.0f, st));
            d.extra = e->extra.p;
            e->biased = true;
        } else {
            e->extra.release();
            d.extra = nullptr;
        }
    }
    // an unbiased local-Q CV on a biased environment without a biased local-Q rides
    // on that environment's force pass (one per environment)
    for (auto* e : fx->envs) { e->dev.fld_l = 0; e->dev.fld_slot = -1; e->dev.fld_lq = nullptr; }
    for (auto* c : fx->cvs) {
        c->fused = false;
        if (c->kind != CV_LD || c->biased || fx->split) continue;
        EnvH* e = fx->envs[c->env];
        if (e->biased && e->dev.extra == nullptr && e->dev.fld_l == 0) { c->fused = true; e->dev.fld_l = c->l; }
    }
    for (auto* c : fx->cvs) {
        if (c->kind == CV_LD) {
            c->psum = nsum++;
            const int W = 2 * c->l + 1, BS = (W + 3) / 4 * 4;
            c->lq.alloc(NT);
            c->bq.alloc((size_t)NT * BS);
            if (c->fused) {
                EnvH* e = fx->envs[c->env];
                e->dev.fld_slot = c->psum;
                e->dev.fld_lq = c->lq.p;
            }
        } else if (c->kind == CV_SK) {
            const int nk = (int)(c->kvec.size() / 3);
            std::vector<double> krev(c->kvec.size());
            for (size_t q = 0; q < krev.size(); ++q) krev[q] = c->kvec[q] / (2.0 * 3.14159265358979323846);
            c->kd.alloc(c->kvec.size()); c->kd.upload(c->kvec.data(), c->kvec.size(), st);
            c->krev.alloc(krev.size()); c->krev.upload(krev.data(), krev.size(), st);
            std::vector<int32_t> ch, cb(nw + 1, 0), na(nw), ab(nw);
            // atoms per chunk of the amplitude pass: about 2,048 workgroups (k tiles x chunks)
            // whatever the k count, between 2,048 and 16,384 atoms (at least 8 atoms per
            // thread, so that few k vectors do not leave every workgroup a block sum for
            // one atom per thread, and the fold few chunk partials)
            int64_t tot_atoms = 0;
            for (int k = 0; k < nw; ++k) tot_atoms += ctx->w[k].n_local;
            const int64_t ktiles = (nk + 7) / 8;
            int CH = 2048;
            while (CH < 16384 && (tot_atoms + CH - 1) / CH * ktiles > 2048) CH *= 2;
            c->max_atoms = 0;
            for (int k = 0; k < nw; ++k) {
                WalkerIn& a = ctx->w[k];
                na[k] = (int32_t)a.n_local; ab[k] = (int32_t)a.atom_off;
                c->max_atoms = std::max<int>(c->max_atoms, (int)a.n_local);
                for (int64_t s = 0; s < a.n_local; s += CH) {
                    ch.push_back((int32_t)(a.atom_off + s));
                    ch.push_back((int32_t)std::min<int64_t>(CH, a.n_local - s));
                    ch.push_back(k);
                }
                cb[k + 1] = (int32_t)(ch.size() / 3);
            }
            c->nchunk = (int)(ch.size() / 3);
            c->chunk.alloc(ch.size()); c->chunk.upload(ch.data(), ch.size(), st);
            c->chunk_begin.alloc(cb.size()); c->chunk_begin.upload(cb.data(), cb.size(), st);
            c->natoms.alloc(nw); c->natoms.upload(na.data(), nw, st);
            c->abegin.alloc(nw); c->abegin.upload(ab.data(), nw, st);
            c->apart.alloc((size_t)c->nchunk * nk * 2);
            c->amp.alloc((size_t)nw * nk * 2);
            c->sterm.alloc((size_t)nw * nk);
            c->kpack.alloc((size_t)nw * nk);
            if (fx->split) c->nnorm.alloc(nw); else c->nnorm.release();
            if (nk > kSkTiledK) {
                // tile-relative phases: slot -> atom, each walker's atoms in Hilbert order
                // within its own range, and the per-step tile references and offsets
                std::vector<int32_t> perm((size_t)std::max<int64_t>(ctx->total_local, 1), 0);
                double ext = 0.0, kmax = 0.0;   // largest |r_j - r_t| of a tile, largest |k| / 2pi
                for (int k = 0; k < nw; ++k) {
                    const WalkerIn& a = ctx->w[k];
                    const std::vector<int32_t> ord = hilbert_order(a);
                    for (int64_t s2 = 0; s2 < a.n_local; ++s2) {
                        perm[(size_t)(a.atom_off + s2)] = (int32_t)(a.atom_off + ord[(size_t)s2]);
                        const int64_t t = ord[(size_t)(s2 / 256 * 256)], j = ord[(size_t)s2];
                        double d2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            const double dd = a.x[3 * j + d] - a.x[3 * t + d];
                            d2 += dd * dd;
                        }
                        ext = std::max(ext, std::sqrt(d2));
                    }
                }
                for (int q = 0; q < nk; ++q)
                    kmax = std::max(kmax, std::sqrt(krev[3 * q] * krev[3 * q] + krev[3 * q + 1] * krev[3 * q + 1] +
                                                    krev[3 * q + 2] * krev[3 * q + 2]));
                // atoms move a few length units between list bindings: 16 of margin
                c->sk_wide = (ext + 16.0) * kmax + 1.0 > 128.0;
                c->perm.alloc(perm.size()); c->perm.upload(perm.data(), perm.size(), st);
                std::vector<int32_t> tb(nw + 1, 0);
                for (int k = 0; k < nw; ++k) tb[k + 1] = tb[k] + (int32_t)((ctx->w[k].n_local + 255) / 256);
                c->tbeg.alloc(tb.size()); c->tbeg.upload(tb.data(), tb.size(), st);
                c->tref.alloc((size_t)std::max(tb[nw], 1));
                c->dr.alloc(perm.size());
            } else {
                c->perm.release(); c->tbeg.release(); c->tref.release(); c->dr.release();
            }
        } else if (c->kind == CV_SF) {
            const std::vector<double> ngs = group_counts(c->sf_group);
            std::vector<double> inv(nw);
            for (int k = 0; k < nw; ++k) inv[k] = ngs[k] > 0 ? 1.0 / ngs[k] : 0.0;
            c->sf_inv.alloc(nw); c->sf_inv.upload(inv.data(), nw, st);
            if (c->sf_chem) {
                // w of every atom, ghost images taking their owner's type
                std::vector<double> aw((size_t)std::max<int64_t>(ctx->total_local + ctx->total_ghost, 1), 1.0);
                auto wt = [&](int t) {
                    auto it = c->chem_map.find(t);
                    if (it != c->chem_map.end()) return it->second;
                    const double dd = (double)t - c->chem_c;
                    return std::exp(-(dd * dd) / (2.0 * c->chem_sigma * c->chem_sigma));
                };
                for (int k = 0; k < nw; ++k) {
                    const WalkerIn& a = ctx->w[k];
                    if ((int64_t)a.types.size() != a.n_local)
                        throw Err(MDT_EINVAL, fmt("CAL %s is chemically weighted: bind the atom types (mdt_bind_types)",
                                                  c->name.c_str()));
                    for (int64_t q = 0; q < a.n_local; ++q) aw[(size_t)(a.atom_off + q)] = wt(a.types[q]);
                    for (int64_t q = 0; q < a.n_all - a.n_local; ++q)
                        aw[(size_t)(ctx->total_local + a.ghost_off + q)] = wt(a.types[a.owner[q]]);
                }
                c->sf_aw.alloc(aw.size()); c->sf_aw.upload(aw.data(), aw.size(), st);
            } else {
                c->sf_aw.release();
            }
            c->sf_S.alloc((size_t)std::max<int64_t>(NT, 1));
            c->sf_part.alloc((size_t)std::max<size_t>(ctx->h_tiles.size(), 1));
            c->sf_tb.alloc(nw + 1); c->sf_tb.upload(ctx->tile_begin.data(), nw + 1, st);
        } else if (c->kind == CV_LQ) {
            const std::vector<double> ng = group_counts(c->lq_group);
            const size_t LS = 2 * (size_t)(c->l + 1), n1 = (size_t)std::max<int64_t>(NT, 1);
            c->lq_ng.alloc(nw); c->lq_ng.upload(ng.data(), nw, st);
            c->lq_S.alloc(n1); c->lq_ql.alloc(n1);
            c->lq_q.alloc(n1 * LS); c->lq_LQ.alloc(n1 * LS); c->lq_K.alloc(n1 * LS); c->lq_H.alloc(n1 * LS);
            c->lq_hk.alloc(n1);
            // per-tile ELL layout of the per-entry words: tile t holds 64 x (its longest list)
            std::vector<int32_t> hoff((size_t)NT + 1, 0);
            HIPCHK(hipStreamSynchronize(st));
            if (NT > 0) HIPCHK(hipMemcpy(hoff.data(), ctx->off.p, sizeof(int32_t) * (NT + 1), hipMemcpyDeviceToHost));
            std::vector<int64_t> apos((size_t)std::max<int64_t>(NT, 1), 0);
            int64_t ell = 0;
            for (const Tile& t : ctx->h_tiles) {
                int mx = 0;
                for (int a = 0; a < t.count; ++a) {
                    mx = std::max(mx, hoff[t.start + a + 1] - hoff[t.start + a]);
                    apos[t.start + a] = ell + a;
                }
                ell += 64 * (int64_t)mx;
            }
            const size_t ell1 = (size_t)std::max<int64_t>(ell, 1);
            c->lq_apos.alloc(apos.size()); c->lq_apos.upload(apos.data(), apos.size(), st);
            c->lq_g.alloc(ell1 * 3);
            c->lq_es.alloc(ell1);
            c->lq_u.alloc(ell1);
            c->lq_erev.alloc(ell1);
            c->lq_eidx.alloc(ell1);
            c->lq_eown.alloc(ell1);
            {
                // owners and reverse positions on the bound lists, for the per-entry layout
                DBuf<int32_t> reown, rrev;
                reown.alloc((size_t)std::max<int64_t>(ctx->total_nnz, 1));
                rrev.alloc((size_t)std::max<int64_t>(ctx->total_nnz, 1));
                HIPCHK(launch_topology_raw(ctx->geo(), NT, reown.p, rrev.p, st));
                HIPCHK(launch_lq_erev(ctx->geo(), NT, reown.p, rrev.p, c->lq_apos.p, c->lq_erev.p, c->lq_eidx.p,
                                      c->lq_eown.p, st));
                HIPCHK(hipStreamSynchronize(st));
            }
            c->lq_part.alloc(2 * (size_t)std::max<size_t>(ctx->h_tiles.size(), 1));
            c->lq_sums.alloc(2 * (size_t)nw);
            c->lq_tb.alloc(nw + 1); c->lq_tb.upload(ctx->tile_begin.data(), nw + 1, st);
        } else if (c->kind == CV_DIST) {
            std::vector<int64_t> off(nw);
            std::vector<double> box((size_t)nw * 3);
            for (int k = 0; k < nw; ++k) {
                WalkerIn& a = ctx->w[k];
                if (c->a1 >= a.n_local || c->a2 >= a.n_local)
                    throw Err(MDT_EINVAL, fmt("DISTANCE %s: atom ID beyond walker %d's %lld atoms", c->name.c_str(), k,
                                              (long long)a.n_local));
                off[k] = a.atom_off;
                for (int d = 0; d < 3; ++d) box[3 * k + d] = a.hi[d] - a.lo[d];
            }
            c->doff.alloc(nw); c->doff.upload(off.data(), nw, st);
            c->dbox.alloc((size_t)nw * 3); c->dbox.upload(box.data(), (size_t)nw * 3, st);
            c->dvec.alloc((size_t)nw * 4);
        }
    }
    fx->nslot_sum = nsum;
    const int ntiles = (int)ctx->h_tiles.size();
    fx->partial.alloc((size_t)std::max(1, nsum) * std::max(1, ntiles));
    // zero, so the first step's reduce kernel, which sums the fused local-Q partials of
    // the step before, reads numbers
    HIPCHK(launch_fill(fx->partial.p, (int64_t)std::max(1, nsum) * std::max(1, ntiles), 0.0, st));
    fx->direct.alloc((size_t)nw * kMaxSlotCV);
    HIPCHK(launch_fill(fx->direct.p, (int64_t)nw * kMaxSlotCV, 0.0, st));
    fx->state.alloc((size_t)nw * kStStride);
    HIPCHK(launch_fill(fx->state.p, (int64_t)nw * kStStride, 0.0, st));
    fx->wcv.alloc((size_t)nw * kWcvStride);
    HIPCHK(launch_fill(fx->wcv.p, (int64_t)nw * kWcvStride, 0.0, st));
    fx->bwork.alloc((size_t)nw * kBwStride);
    HIPCHK(launch_fill(fx->bwork.p, (int64_t)nw * kBwStride, 0.0, st));
    // 1 / group count per walker of every CV whose value is a sum over its group
    {
        std::vector<double> h((size_t)nw * kMaxSlotCV, 0.0);
        for (auto* c : fx->cvs) {
            if (c->kind != CV_Q && c->kind != CV_LD) continue;
            for (int k = 0; k < nw; ++k) {
                const double ng = ctx->share_count(fx->envs[c->env]->group, k);
                h[(size_t)k * kMaxSlotCV + c->slot] = ng > 0 ? 1.0 / ng : 0.0;
            }
        }
        fx->cv_inv.alloc(h.size());
        fx->cv_inv.upload(h.data(), h.size(), st);
    }
    fx->tile_begin.alloc(nw + 1); fx->tile_begin.upload(ctx->tile_begin.data(), nw + 1, st);
    fx->progs.alloc(progs.size()); fx->progs.upload(progs.data(), progs.size(), st);
    // affine DIM programs as coefficients per value slot: s_d = lin0_d + sum_c A[d][c] v_c,
    // A[d][c] = sum of lin_d[q] over the symbols q bound to slot c
    std::vector<double> h_aff((size_t)kMaxDim * kMaxSlotCV, 0.0);
    bool all_affine = !progs.empty();
    for (size_t d = 0; d < progs.size(); ++d) {
        if (!progs[d].linear) { all_affine = false; continue; }
        for (size_t q = 0; q < fx->syms.size(); ++q)
            h_aff[d * kMaxSlotCV + fx->cvs[fx->syms[q].cv]->slot] += progs[d].lin[q];
    }
    fx->aff.alloc(h_aff.size()); fx->aff.upload(h_aff.data(), h_aff.size(), st);
    if (!fx->inited) {
        // bias storage (first run only, like init_set_mode)
        const int D = fx->D();
        int64_t gsz = 1;
        if (fx->grid_bias()) for (int d = 0; d < D; ++d) gsz *= fx->nb[d];
        else gsz = 1;
        fx->grid.alloc((size_t)nw * gsz);
        HIPCHK(launch_fill(fx->grid.p, (int64_t)nw * gsz, 0.0, st));
        fx->hills_cap = 4096;
        fx->hills.alloc((size_t)nw * fx->hills_cap * (D + 3));
        fx->nhills.alloc(nw);
        std::vector<int64_t> z(nw, 0);
        fx->nhills.upload(z.data(), nw, st);
        fx->step.alloc(nw);
        fx->step.upload(z.data(), nw, st);
        fx->mlp_ok.alloc(nw);
        std::vector<int32_t> zi(nw, 0);
        fx->mlp_ok.upload(zi.data(), nw, st);
        if (fx->hill_type == 3) {
            fx->mlp_live.alloc(nw); fx->mlp_stage.alloc(nw);
            fx->mlp_m.alloc((size_t)nw * kMlpP); fx->mlp_v.alloc((size_t)nw * kMlpP);
            fx->mlp_fits.alloc(nw); fx->mlp_fallbacks.alloc(nw); fx->mlp_t.alloc(nw);
            fx->mlp_bound_fit.assign(nw, 0); fx->mlp_roff.assign(nw, 0);
            fx->mlp_err.alloc(nw); fx->mlp_box_lo.alloc((size_t)3 * nw); fx->mlp_box_hi.alloc((size_t)3 * nw);
            fx->mlp_fits.upload(z.data(), nw, st);
            fx->mlp_fallbacks.upload(z.data(), nw, st);
            HIPCHK(launch_fill(fx->mlp_err.p, nw, 0.0, st));
            const int64_t ntot = fx->mlp_points + fx->mlp_holdout;
            fx->mlp_sdata.alloc((size_t)ntot);
            fx->mlp_vdata.alloc((size_t)ntot);
            fx->mlp_gpart.alloc((size_t)mlp_grad_rows(fx->mlp_batch) * kMlpP);
            fx->mlp_blkerr.alloc((size_t)((fx->mlp_holdout + 63) / 64));
            for (int w = 0; w < nw; ++w) HIPCHK(mlp_init(fx->mlp_live.p + w, fx->mlp_seed + 977 * w, st));
            fx->mlp_has_fit.assign(nw, 0);
        }
        fx->hills_written.assign(nw, 0);
        fx->hills_bound.assign(nw, 0);
        HIPCHK(hipStreamSynchronize(st));
    }
    // reduce descriptor
    ReduceDev& r = fx->red;
    r = ReduceDev{};
    r.n_walkers = nw; r.nslot_sum = nsum; r.ntiles = ntiles; r.tiles = ctx->tiles.p;
    r.tile_begin = fx->tile_begin.p; r.partial = fx->partial.p;
    r.direct = fx->direct.p;
    r.ncv = (int)fx->cvs.size();
    // every CV divides by its own group count: Q/LD partial sums use the env's inv_ng
    for (size_t c = 0; c < fx->cvs.size(); ++c)
        r.cv_src[c] = (fx->split || fx->cvs[c]->kind == CV_SK || fx->cvs[c]->kind == CV_DIST || fx->cvs[c]->kind == CV_SF ||
                       fx->cvs[c]->kind == CV_LQ) ? -1 : fx->cvs[c]->fused ? -2 - fx->cvs[c]->psum : fx->cvs[c]->psum;
    if (fx->split) {
        // the sums over every share arrive from the host: each CV's value is a direct one
        for (auto* c : fx->cvs)
            if (c->kind == CV_DIST || c->kind == CV_SF || c->kind == CV_LQ)
                throw Err(MDT_ENOTSUP, fmt("CAL %s: split reductions cover STEINH Q/LD and SK", c->name.c_str()));
        std::vector<int32_t> po(kMaxSlotCV, -1);
        for (auto* c : fx->cvs) if (c->kind == CV_Q || c->kind == CV_LD) po[c->slot] = c->psum;
        fx->psum_of.alloc(kMaxSlotCV); fx->psum_of.upload(po.data(), kMaxSlotCV, st);
        fx->psums.alloc((size_t)nw * kMaxSlotCV);
    }
    fx->built_split = fx->split;
    r.cv_inv = fx->cv_inv.p;
    for (int c = 0; c < kMaxSlotCV; ++c) r.walias[c] = c < (int)fx->cvs.size() ? (walias[c] >= 0 ? walias[c] : c) : -1;
    r.nsym = (int)fx->syms.size();
    for (size_t q = 0; q < fx->syms.size(); ++q) r.sym_slot[q] = fx->cvs[fx->syms[q].cv]->slot;
    r.D = fx->D();
    r.prog = fx->progs.p;
    r.affine = all_affine ? 1 : 0;
    for (int d = 0; d < fx->D(); ++d) r.aff0[d] = all_affine ? progs[d].lin0 : 0.0;
    r.aff = fx->aff.p;
    r.tile_args = nw <= kRedArgWalkers ? 1 : 0;
    if (r.tile_args) for (int k = 0; k <= nw; ++k) r.tile_range[k] = ctx->tile_begin[k];
    r.kind = fx->grid_bias() ? BIAS_GRID : fx->hill_type == 2 ? BIAS_HILLS : BIAS_MLP;
    r.pace = fx->pace;
    for (int d = 0; d < fx->D(); ++d) {
        r.lo[d] = fx->lo[d]; r.hi[d] = fx->hi[d]; r.nb[d] = fx->nb[d];
        double dcv = (fx->hi[d] - fx->lo[d]) / fx->nb[d];
        r.radius[d] = (int)std::ceil(4.0 * fx->sigma / dcv) + 1;
        r.sigma[d] = fx->sigmas[d];
    }
    if (r.kind == BIAS_GRID) for (int d = 0; d < fx->D(); ++d) r.sigma[d] = fx->sigma;
    r.grid_size = 1;
    if (r.kind == BIAS_GRID) for (int d = 0; d < fx->D(); ++d) r.grid_size *= fx->nb[d];
    r.grid = fx->grid.p;
    r.h0 = fx->height0; r.biasf = fx->biasf; r.kT = fx->temp * fx->boltz; r.wt = fx->wt ? 1 : 0;
    r.hills = fx->hills.p; r.nhills = fx->nhills.p; r.hills_cap = fx->hills_cap;
    r.step = fx->step.p; r.state = fx->state.p; r.wcv = fx->wcv.p;
    r.epoch = ctx->near_on() ? ctx->nepoch.p : nullptr;
    r.mlp = fx->mlp_live.p; r.mlp_ok = fx->mlp_ok.p; r.mlp_err = fx->mlp_err.p;
    r.bwork = fx->bwork.p;
    if (fx->hill_type == 3) {
        MlpDev& m = fx->mdev;
        m = MlpDev{};
        m.D = fx->D(); m.hills_cap = fx->hills_cap; m.hills = fx->hills.p; m.nhills = fx->nhills.p;
        for (int d = 0; d < 3; ++d) m.sigma[d] = fx->sigmas[d];
        m.spread = fx->mlp_spread;
        m.live = fx->mlp_live.p; m.stage = fx->mlp_stage.p;
        m.adam_m = fx->mlp_m.p; m.adam_v = fx->mlp_v.p;
        m.ok = fx->mlp_ok.p; m.err = fx->mlp_err.p; m.fits = fx->mlp_fits.p; m.fallbacks = fx->mlp_fallbacks.p;
        m.box_lo_out = fx->mlp_box_lo.p; m.box_hi_out = fx->mlp_box_hi.p;
        m.adam_t = fx->mlp_t.p;
    }
    for (auto* e : fx->envs) e->dev.partial = fx->partial.p, e->dev.wcv = fx->wcv.p;
    if (!fx->pin) {
        fx->pin_n = 1 << 16;
        HIPCHK(hipHostMalloc((void**)&fx->pin, fx->pin_n * sizeof(double), hipHostMallocDefault));
    }
    fx->built_gen = ctx->generation;
    fx->builds++;
    if (!fx->inited) {
        fx->inited = true;
        open_hills(fx);
    }
}

// Restart from HILLS: "step cv... h sigma" per line, '#' comments
// (GH_t0_uniformGrid.cu:100-192); replayed hills use the file's h and sigma.
void replay_hills(mdt_fix* fx, int w, const std::string& path) {
    FILE* fp = fopen(path.c_str(), "r");
    if (!fp) return;
    const int D = fx->D();
    std::vector<double> s, h;
    std::vector<double> sig;
    char line[4096];
    while (fgets(line, sizeof(line), fp)) {
        if (line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
        char* p = line;
        char* q = nullptr;
        strtoll(p, &q, 10);
        p = q;
        double cv[kMaxDim];
        for (int d = 0; d < D; ++d) { cv[d] = strtod(p, &q); p = q; }
        double hh = strtod(p, &q); p = q;
        double sg = strtod(p, &q);
        for (int d = 0; d < D; ++d) s.push_back(cv[d]);
        h.push_back(hh);
        sig.push_back(sg);
    }
    fclose(fp);
    if (h.empty()) return;
    hipStream_t st = fx->ctx->stream;
    // a hills-list bias keeps every replayed hill: grow its device log first
    if (!fx->grid_bias()) ensure_hills(fx, (int64_t)h.size());
    DBuf<double> ds, dh;
    // hills with different sigma are deposited in runs of equal sigma
    size_t a = 0;
    while (a < h.size()) {
        size_t b = a;
        while (b < h.size() && sig[b] == sig[a]) ++b;
        ds.alloc((b - a) * D); ds.upload(&s[a * D], (b - a) * D, st);
        dh.alloc(b - a); dh.upload(&h[a], b - a, st);
        HIPCHK(launch_bias_add(fx->red, w, (int64_t)(b - a), ds.p, dh.p, sig[a], st));
        HIPCHK(hipStreamSynchronize(st));
        a = b;
    }
    fx->hills_bound[w] += (int64_t)h.size();
}

std::string hills_path(mdt_fix* fx, int w) {
    if (fx->nw() == 1) return fx->hills_file;
    return fx->hills_file + "." + std::to_string(w);
}

void open_hills(mdt_fix* fx) {
    const int nw = fx->nw();
    fx->hfiles.assign(nw, nullptr);
    if (fx->hills_file == "NONE") return;
    for (int w = 0; w < nw; ++w) {
        std::string p = hills_path(fx, w);
        bool exists = false;
        if (fx->restart) {
            FILE* t = fopen(p.c_str(), "r");
            if (t) { fseek(t, 0, SEEK_END); exists = ftell(t) > 0; fclose(t); }
        }
        if (exists) {
            replay_hills(fx, w, p);
            // replayed hills are not re-written
            int64_t nh = 0;
            HIPCHK(hipMemcpy(&nh, fx->nhills.p + w, sizeof(int64_t), hipMemcpyDeviceToHost));
            fx->hills_written[w] = nh;
            fx->hfiles[w] = fopen(p.c_str(), "a");
        } else {
            fx->hfiles[w] = fopen(p.c_str(), "w");
            if (fx->hfiles[w]) {
                if (fx->D() == 2) fprintf(fx->hfiles[w], "# step cv1 cv2 height sigma\n");
                if (fx->D() == 1) fprintf(fx->hfiles[w], "# step cv_values height sigma\n");
            }
        }
        if (!fx->hfiles[w]) throw Err(MDT_EINVAL, fmt("Cannot open HILLS file %s for writing/appending.", p.c_str()));
    }
}

// grow the device hill log so that `extra` more hills fit for every walker
void ensure_hills(mdt_fix* fx, int64_t extra) {
    int64_t need = 0;
    for (auto b : fx->hills_bound) need = std::max(need, b + extra);
    if (need <= fx->hills_cap) return;
    int64_t cap = fx->hills_cap;
    while (cap < need) cap *= 2;
    const int D = fx->D(), nw = fx->nw();
    DBuf<double> nh;
    nh.alloc((size_t)nw * cap * (D + 3));
    for (int w = 0; w < nw; ++w)
        HIPCHK(hipMemcpyAsync(nh.p + (size_t)w * cap * (D + 3), fx->hills.p + (size_t)w * fx->hills_cap * (D + 3),
                              sizeof(double) * fx->hills_cap * (D + 3), hipMemcpyDeviceToDevice, fx->ctx->stream));
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    std::swap(fx->hills.p, nh.p);
    std::swap(fx->hills.n, nh.n);
    fx->hills_cap = cap;
    fx->red.hills = fx->hills.p;
    fx->red.hills_cap = cap;
    fx->mdev.hills = fx->hills.p;
    fx->mdev.hills_cap = cap;
    fx->graph_gen = 0;   // invalidate the captured graph
}

// write hills deposited since the last flush to the HILLS files
void flush_hills(mdt_fix* fx) {
    const int nw = fx->nw(), D = fx->D();
    std::vector<int64_t> nh(nw);
    HIPCHK(hipMemcpyAsync(nh.data(), fx->nhills.p, sizeof(int64_t) * nw, hipMemcpyDeviceToHost, fx->ctx->stream));
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    for (int w = 0; w < nw; ++w) {
        fx->hills_bound[w] = nh[w];
        int64_t a = fx->hills_written[w];
        if (nh[w] <= a) continue;
        std::vector<double> buf((size_t)(nh[w] - a) * (D + 3));
        HIPCHK(hipMemcpy(buf.data(), fx->hills.p + ((size_t)w * fx->hills_cap + a) * (D + 3),
                         sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        FILE* fp = fx->hfiles.empty() ? nullptr : fx->hfiles[w];
        if (fp) {
            for (int64_t k = 0; k < nh[w] - a; ++k) {
                const double* rec = &buf[(size_t)k * (D + 3)];
                fprintf(fp, "%ld", (long)rec[0]);
                for (int d = 0; d < D; ++d) fprintf(fp, " %.16g", rec[1 + d]);
                if (fx->grid_bias() || D == 1)
                    fprintf(fp, " %.16g %.16g\n", rec[1 + D], rec[2 + D]);
                else {
                    fprintf(fp, " %.16g", rec[1 + D]);
                    for (int d = 0; d < D; ++d) fprintf(fp, " %.16g", fx->sigmas[d]);
                    fprintf(fp, "\n");
                }
            }
            fflush(fp);
        }
        fx->hills_written[w] = nh[w];
    }
}

// Fit the learned bias of walker w to the hills deposited so far (asynchronous,
// on the context stream): training points with exact FP64 labels, Adam epochs on the
// MFMA kernels, held-out max |dV| gate against MLP_TOL kT.  The first fit draws every
// point; a refit keeps the points, adds the hills deposited since the last fit to their
// labels and redraws a block around those hills (its share of all hills, 1/64 to 1/4 of
// the set).  One epoch of training steps is captured as a graph and replayed once per
// epoch: the Adam step and the learning-rate schedule come from a device counter.
void mlp_refit(mdt_fix* fx, int w) {
    hipStream_t st = fx->ctx->stream;
    const MlpDev& m = fx->mdev;
    MlpNet* stage = fx->mlp_stage.p + w;
    const bool first = !fx->mlp_has_fit[w];
    // a refit continues from the last fitted network, kept in the staging slot
    // whether or not the held-out gate adopted it
    if (first) HIPCHK(mlp_init(stage, fx->mlp_seed + 977 * w, st));
    HIPCHK(launch_fill_f(fx->mlp_m.p + (size_t)w * kMlpP, kMlpP, 0.0f, st));
    HIPCHK(launch_fill_f(fx->mlp_v.p + (size_t)w * kMlpP, kMlpP, 0.0f, st));
    HIPCHK(hipMemsetAsync(fx->mlp_t.p + w, 0, sizeof(int64_t), st));
    const uint64_t seed = fx->mlp_seed * 0x9E3779B97F4A7C15ull + (uint64_t)(fx->mlp_round++) * 7919 + w;
    const int64_t M = fx->mlp_points, Hn = fx->mlp_holdout;
    if (first) {
        HIPCHK(mlp_box(m, w, st));
        HIPCHK(mlp_data(m, w, M, seed, fx->mlp_sdata.p, fx->mlp_vdata.p, 0, st));
    } else {
        HIPCHK(mlp_data(m, w, M, seed, fx->mlp_sdata.p, fx->mlp_vdata.p, 1, st));
        const int64_t now = std::max<int64_t>(fx->hills_bound[w], 1), prev = fx->mlp_bound_fit[w];
        int64_t R = (int64_t)((double)M * (double)std::max<int64_t>(now - prev, 0) / (double)now);
        R = std::min(std::max(R, M / 64), M / 4);
        int64_t off = fx->mlp_roff[w] % M;
        const int64_t r1 = std::min(R, M - off);
        HIPCHK(mlp_data(m, w, r1, seed ^ 0x5bd1e995ull, fx->mlp_sdata.p + off, fx->mlp_vdata.p + off, 2, st));
        if (R > r1) HIPCHK(mlp_data(m, w, R - r1, seed ^ 0x27d4eb2full, fx->mlp_sdata.p, fx->mlp_vdata.p, 2, st));
        fx->mlp_roff[w] = (off + R) % M;
        HIPCHK(mlp_box(m, w, st));
    }
    fx->mlp_bound_fit[w] = fx->hills_bound[w];
    HIPCHK(mlp_data(m, w, Hn, seed ^ 0xA5A5A5A5A5A5A5A5ull, fx->mlp_sdata.p + M, fx->mlp_vdata.p + M, 0, st));
    HIPCHK(mlp_stats(m, w, M, fx->mlp_vdata.p, st));
    const int epochs = first ? fx->mlp_epochs_first : fx->mlp_epochs;
    const int64_t B = fx->mlp_batch;
    const int64_t nb = (M + B - 1) / B;
    if (fx->mlp_graph_gen != fx->builds) {
        for (auto& kv : fx->mlp_graphs) (void)hipGraphExecDestroy(kv.second);
        fx->mlp_graphs.clear();
        fx->mlp_graph_gen = fx->builds;
    }
    // one epoch: nb training steps over the batches in order; a refit starts from the
    // live network and fine-tunes at a lower rate
    const double lr0 = first ? fx->mlp_lr : (fx->mlp_lr_refit >= 0.0 ? fx->mlp_lr_refit : 0.1 * fx->mlp_lr);
    const int64_t nit = std::max<int64_t>((int64_t)epochs * nb, 1);
    if (!fx->mlp_graph) {
        for (int e = 0; e < epochs; ++e)
            for (int64_t b0 = 0; b0 < M; b0 += B)
                HIPCHK(mlp_train_step(m, w, b0, std::min(B, M - b0), M, fx->mlp_sdata.p, fx->mlp_vdata.p,
                                      fx->mlp_gpart.p, (float)lr0, nullptr, 0, nit, st));
    }
    const std::pair<int, int> key(w, first ? 1 : 0);
    auto git = fx->mlp_graphs.find(key);
    if (fx->mlp_graph && git == fx->mlp_graphs.end()) {
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (int64_t b0 = 0; b0 < M && ce == hipSuccess; b0 += B)
            ce = mlp_train_step(m, w, b0, std::min(B, M - b0), M, fx->mlp_sdata.p, fx->mlp_vdata.p,
                                fx->mlp_gpart.p, (float)lr0, nullptr, 0, nit, st);
        hipError_t ee = hipStreamEndCapture(st, &gr);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            throw Err(MDT_EHIP, fmt("capturing the training epoch: %s", hipGetErrorString(ce != hipSuccess ? ce : ee)));
        }
        hipGraphExec_t ex = nullptr;
        hipError_t ie = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) throw Err(MDT_EHIP, fmt("hipGraphInstantiate: %s", hipGetErrorString(ie)));
        git = fx->mlp_graphs.emplace(key, ex).first;
    }
    if (fx->mlp_graph)
        for (int e = 0; e < epochs; ++e) HIPCHK(hipGraphLaunch(git->second, st));
    const int nblk = (int)((Hn + 63) / 64);
    HIPCHK(mlp_forward(stage, Hn, fx->mlp_sdata.p + M, fx->mlp_vdata.p + M, nullptr, fx->mlp_blkerr.p, st));
    HIPCHK(mlp_commit(m, w, nblk, fx->mlp_blkerr.p, fx->mlp_tol * fx->temp * fx->boltz, st));
    fx->mlp_has_fit[w] = 1;
}

void mlp_refit_all(mdt_fix* fx) {
    for (int w = 0; w < fx->nw(); ++w) mlp_refit(fx, w);
}

// Optional per-launch timing (HIP events on the context stream).
struct Timer {
    bool on = false;
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<int64_t> count;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<int> pending_id;
    int id_of(const std::string& n) {
        for (size_t k = 0; k < names.size(); ++k) if (names[k] == n) return (int)k;
        names.push_back(n); ms.push_back(0.0); count.push_back(0);
        return (int)names.size() - 1;
    }
    template <class F>
    void run(const char* name, hipStream_t st, F&& fn) {
        if (!on) { fn(); return; }
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, st));
        fn();
        HIPCHK(hipEventRecord(b, st));
        pending.push_back({a, b});
        pending_id.push_back(id_of(name));
    }
    void collect() {
        for (size_t k = 0; k < pending.size(); ++k) {
            HIPCHK(hipEventSynchronize(pending[k].second));
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, pending[k].first, pending[k].second));
            ms[pending_id[k]] += t;
            count[pending_id[k]] += 1;
            (void)hipEventDestroy(pending[k].first);
            (void)hipEventDestroy(pending[k].second);
        }
        pending.clear();
        pending_id.clear();
    }
};

DistDev dist_dev(mdt_fix* fx, CvH* c) {
    DistDev d{};
    d.n_walkers = fx->nw(); d.a1 = c->a1; d.a2 = c->a2;
    d.atom_off = c->doff.p; d.box = c->dbox.p; d.dvec = c->dvec.p;
    d.direct = fx->direct.p; d.wcv = fx->wcv.p; d.slot = c->slot;
    return d;
}

SfDev sf_dev(mdt_fix* fx, CvH* c) {
    SfDev d{};
    d.count = c->func == "COUNT" ? 1 : 0;
    d.pair = c->sf_pair;
    d.aw = c->sf_chem ? c->sf_aw.p : nullptr;
    d.sw = c->sf_sw;
    d.rc2 = c->sf_rc * c->sf_rc; d.q = c->sf_q;
    d.S = c->sf_S.p; d.part = c->sf_part.p; d.tile_begin = c->sf_tb.p; d.inv_ng = c->sf_inv.p;
    d.direct = fx->direct.p; d.wcv = fx->wcv.p; d.slot = c->slot; d.n_walkers = fx->nw();
    return d;
}

LqDev lq_dev(mdt_fix* fx, CvH* c) {
    LqDev d{};
    d.l = c->l; d.mode = lq_mode(c->func.empty() ? "MEAN_SOLID" : c->func);
    d.swr = c->lq_swr; d.swcv = c->lq_swcv; d.eps = c->lq_eps;
    d.S = c->lq_S.p; d.q = c->lq_q.p; d.LQ = c->lq_LQ.p; d.ql = c->lq_ql.p; d.K = c->lq_K.p; d.g = c->lq_g.p;
    d.es = c->lq_es.p; d.erev = c->lq_erev.p; d.apos = c->lq_apos.p;
    d.H = c->lq_H.p; d.hk = c->lq_hk.p; d.u = c->lq_u.p;
    d.eidx = c->lq_eidx.p; d.eown = c->lq_eown.p;
    d.part = c->lq_part.p; d.sums = c->lq_sums.p; d.tile_begin = c->lq_tb.p; d.ng = c->lq_ng.p;
    d.direct = fx->direct.p; d.wcv = fx->wcv.p; d.slot = c->slot; d.n_walkers = fx->nw();
    return d;
}

// Whether a step of this fix reads the FP64 positions x4 beyond the selection's exact
// tests: the near-list build and the CVs whose kernels work on FP64 positions.
bool fx_reads_x4(const mdt_fix* fx) {
    if (fx->ctx->near_on()) return true;
    for (auto* c : fx->cvs)
        if (c->kind == CV_SF || c->kind == CV_LQ || c->kind == CV_SK || c->kind == CV_DIST) return true;
    return false;
}

// x4 brought up to the attached positions after a step that packed the FP32 ones alone
void ensure_x4(mdt_ctx* ctx) {
    if (!ctx->x4_stale) return;
    for (size_t k = 0; k < ctx->w.size(); ++k)
        if (ctx->w[k].x_src) ctx->pack((int)k, ctx->w[k].x_src, true);
}

// The CV values of the step (through the per-atom sums and the direct values) on the
// context stream: positions, near lists, selection and q_lm, and every value kernel.
void enqueue_values(mdt_fix* fx, Timer& T) {
    mdt_ctx* ctx = fx->ctx;
    hipStream_t st = ctx->stream;
    // positions of walkers attached to an engine buffer are read at the start of the step;
    // a step whose only FP64 reader is the selection (which then reads the one walker's
    // engine buffer in place) packs the FP32 positions alone
    const bool x4_readers = fx_reads_x4(fx);
    for (size_t k = 0; k < ctx->w.size(); ++k)
        if (ctx->w[k].x_src) T.run("k_pack_walker", st, [&] { ctx->pack((int)k, ctx->w[k].x_src, x4_readers); });
    if (!fx->envs.empty() && ctx->near_on()) {
        const NearDev nd = ctx->near_dev();
        T.run("k_prune", st, [&] { HIPCHK(launch_prune(ctx->geo(), nd, 0, st)); });
        T.run("k_topology", st, [&] {
            HIPCHK(launch_topology_near(ctx->geo_near(), nd, ctx->eown.p, ctx->revpos.p, 0, st)); });
    }
    // Steinhardt selection and gradient kernels walk the near lists, the rest the bound ones
    const Geo g = ctx->geo_near(), graw = ctx->geo();
    for (auto* e : fx->envs) {
        Geo ge = g;
        ge.group = ctx->group_ptr(e->group);
        T.run("k_env", st, [&] { HIPCHK(launch_env(ge, e->dev, st)); });
    }
    for (auto* c : fx->cvs) {
        if (c->kind == CV_LD && !c->fused) {
            EnvH* e = fx->envs[c->env];
            Geo ge = g;
            ge.group = ctx->group_ptr(e->group);
            LdDev d{};
            d.l = c->l; d.lq = c->lq.p; d.bq = c->bq.p; d.slot = c->psum; d.wslot = c->biased ? kMaxSlotCV + c->slot : -1;
            T.run("k_ld", st, [&] { HIPCHK(launch_ld(ge, e->dev, d, st)); });
        } else if (c->kind == CV_SK) {
            SkDev s{};
            s.nk = (int)(c->kvec.size() / 3);
            s.k = c->krev.p; s.kvec = c->kd.p; s.nchunk = c->nchunk; s.chunk = c->chunk.p;
            s.apart = c->apart.p; s.amp = c->amp.p; s.kpack = c->kpack.p; s.chunk_begin = c->chunk_begin.p;
            s.sterm = c->sterm.p; s.perm = c->perm.p; s.tbeg = c->tbeg.p; s.tref = c->tref.p; s.dr = c->dr.p;
            s.wide = c->sk_wide ? 1 : 0;
            s.direct = fx->direct.p; s.slot = c->slot; s.wslot = c->slot; s.wcv = fx->wcv.p;
            s.n_atoms = c->natoms.p; s.atom_begin = c->abegin.p; s.n_walkers = fx->nw(); s.max_atoms = c->max_atoms;
            s.n_norm = fx->split ? c->nnorm.p : nullptr;
            Geo gk = g;
            gk.group = g.share;             // a share sums its own atoms' amplitudes
            T.run("k_sk_amp", st, [&] { HIPCHK(launch_sk_amp(gk, s, st)); });
        } else if (c->kind == CV_DIST) {
            const DistDev d = dist_dev(fx, c);
            T.run("k_dist_value", st, [&] { HIPCHK(launch_dist_value(g.x4, d, st)); });
        } else if (c->kind == CV_SF) {
            const SfDev d = sf_dev(fx, c);
            Geo gs = graw;
            gs.group = ctx->group_ptr(c->sf_group);
            T.run("k_sf_value", st, [&] { HIPCHK(launch_sf_value(gs, d, st)); });
        } else if (c->kind == CV_LQ) {
            const LqDev d = lq_dev(fx, c);
            Geo gs = graw;
            gs.group = ctx->group_ptr(c->lq_group);
            T.run("k_lq_q", st, [&] { HIPCHK(launch_lq(0, gs, d, nullptr, st)); });
            T.run("k_lq_lq", st, [&] { HIPCHK(launch_lq(1, gs, d, nullptr, st)); });
            T.run("k_lq_sum", st, [&] { HIPCHK(launch_lq(2, gs, d, nullptr, st)); });
        }
    }
}

// The rest of the step: deposit, bias and chain rule, forces into f.
void enqueue_rest(mdt_fix* fx, double* f, Timer& T) {
    mdt_ctx* ctx = fx->ctx;
    hipStream_t st = ctx->stream;
    const Geo g = ctx->geo_near(), graw = ctx->geo();
    T.run("k_reduce", st, [&] { HIPCHK(launch_reduce(fx->red, st)); });
    for (auto* c : fx->cvs) {
        if (c->kind == CV_LD && c->biased) {
            EnvH* e = fx->envs[c->env];
            LdDev d{};
            d.l = c->l; d.lq = c->lq.p; d.bq = c->bq.p; d.slot = c->psum; d.wslot = kMaxSlotCV + c->slot;
            T.run("k_ld_coeff", st, [&] { HIPCHK(launch_ld_coeff(g, e->dev, d, st)); });
        }
    }
    for (auto* e : fx->envs)
        if (e->biased) T.run("k_force", st, [&] { HIPCHK(launch_force(g, e->dev, f, st)); });
    for (auto* c : fx->cvs) {
        if (c->kind == CV_SK && c->biased) {
            SkDev s{};
            s.nk = (int)(c->kvec.size() / 3);
            s.k = c->krev.p; s.kvec = c->kd.p; s.amp = c->amp.p; s.kpack = c->kpack.p;
            s.wslot = c->slot; s.wcv = fx->wcv.p; s.n_atoms = c->natoms.p; s.atom_begin = c->abegin.p;
            s.n_walkers = fx->nw(); s.max_atoms = c->max_atoms; s.perm = c->perm.p;
            s.tbeg = c->tbeg.p; s.tref = c->tref.p; s.dr = c->dr.p; s.wide = c->sk_wide ? 1 : 0;
            s.n_norm = fx->split ? c->nnorm.p : nullptr;
            Geo gk = g;
            gk.group = g.share;             // a share pushes its own atoms
            T.run("k_sk_force", st, [&] { HIPCHK(launch_sk_force(gk, s, f, st)); });
        }
        if (c->kind == CV_DIST && c->biased) {
            const DistDev d = dist_dev(fx, c);
            T.run("k_dist_force", st, [&] { HIPCHK(launch_dist_force(d, f, st)); });
        }
        if (c->kind == CV_SF && c->biased) {
            const SfDev d = sf_dev(fx, c);
            Geo gs = graw;
            gs.group = ctx->group_ptr(c->sf_group);
            T.run("k_sf_force", st, [&] { HIPCHK(launch_sf_force(gs, d, f, st)); });
        }
        if (c->kind == CV_LQ && c->biased) {
            const LqDev d = lq_dev(fx, c);
            Geo gs = graw;
            gs.group = ctx->group_ptr(c->lq_group);
            T.run("k_lq_k", st, [&] { HIPCHK(launch_lq(3, gs, d, f, st)); });
            T.run("k_lq_h", st, [&] { HIPCHK(launch_lq(4, gs, d, f, st)); });
            T.run("k_lq_pair", st, [&] { HIPCHK(launch_lq(5, gs, d, f, st)); });
            T.run("k_lq_force", st, [&] { HIPCHK(launch_lq(6, gs, d, f, st)); });
        }
    }
    // The values of unbiased local-Q CVs fused into the force pass are the fixed-order
    // sums of the tile partials that pass writes: the next step's reduce kernel takes
    // them, and fetch_state launches k_cv_tail for the last step's before a state read.
}

bool has_fused(const mdt_fix* fx) {
    for (auto* c : fx->cvs)
        if (c->fused) return true;
    return false;
}

void enqueue_step(mdt_fix* fx, double* f, Timer* tm = nullptr) {
    static Timer off_timer;
    Timer& T = tm ? *tm : off_timer;
    if (fx->split) throw Err(MDT_ESTATE, "a fix in split mode steps through mdt_fix_partials and mdt_fix_resume");
    enqueue_values(fx, T);
    enqueue_rest(fx, f, T);
}

void use_device(mdt_ctx* ctx) {
    if (ctx->device < 0) throw Err(MDT_ESTATE, "parse-only context (device -1) cannot run steps");
    HIPCHK(hipSetDevice(ctx->device));
}

void ensure_built(mdt_fix* fx) {
    if (fx->ctx->device < 0) throw Err(MDT_ESTATE, "parse-only context (device -1) cannot run steps");
    if (!fx->ctx->committed || !fx->inited || fx->built_gen != fx->ctx->generation || fx->split != fx->built_split)
        fix_build(fx);
}

void set_steps(mdt_fix* fx, int64_t step) {
    const int nw = fx->nw();
    std::vector<int64_t> s(nw, step);
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    HIPCHK(hipMemcpy(fx->step.p, s.data(), sizeof(int64_t) * nw, hipMemcpyHostToDevice));
}

void fetch_state(mdt_fix* fx) {
    const int nw = fx->nw();
    fx->h_state.resize((size_t)nw * kStStride);
    // the fused local-Q values of the last step (a repeated read sums the same partials)
    if (has_fused(fx)) HIPCHK(launch_cv_tail(fx->red, fx->ctx->stream));
    HIPCHK(hipMemcpyAsync(fx->h_state.data(), fx->state.p, sizeof(double) * fx->h_state.size(), hipMemcpyDeviceToHost, fx->ctx->stream));
    int32_t of = 0;
    if (fx->ctx->near_on())
        HIPCHK(hipMemcpyAsync(&of, fx->ctx->noverflow.p, sizeof(int32_t), hipMemcpyDeviceToHost, fx->ctx->stream));
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    if (of)
        throw Err(MDT_ENOTSUP, fmt("an atom came to have more than %d list entries within %.6g of it; at most %d are "
                                   "supported (the results since are not valid)", kMaxList,
                                   std::sqrt(fx->ctx->near_rp2), kMaxList));
    fx->state_valid = true;
}

int fail(mdt_ctx* ctx, const Err& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
}

// temporary storage of a hipcub call: query, allocate, run
static void cub_call(DBuf<uint8_t>& tmp, hipStream_t, const std::function<hipError_t(void*, size_t&)>& f) {
    size_t bytes = 0;
    HIPCHK(f(nullptr, bytes));
    tmp.alloc(std::max<size_t>(bytes, 1));
    HIPCHK(f((void*)tmp.p, bytes));
}

#define API_BEGIN(ctxp) mdt_ctx* ctx__ = (ctxp); try {
#define FIX_BEGIN(fxp) if (!(fxp) || !(fxp)->ctx) return MDT_ESTATE; API_BEGIN((fxp)->ctx)
#define API_END                                              \
    }                                                        \
    catch (const Err& e) { return fail(ctx__, e); }          \
    catch (const std::exception& e) { return fail(ctx__, Err(MDT_EINVAL, e.what())); } \
    return MDT_OK;

}  // namespace

extern "C" {

const char* mdt_version(void) { return "mdt 0.1 gfx950"; }

int mdt_create(mdt_ctx** out, int device, int n_walkers) {
    if (!out) return MDT_EINVAL;
    *out = nullptr;
    if (n_walkers < 1) return MDT_EINVAL;
    if (device == -1) {
        // parse-only context: keyword lists can be validated on a host without a GPU
        mdt_ctx* c = new mdt_ctx();
        c->device = -1;
        c->w.resize(n_walkers);
        *out = c;
        return MDT_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MDT_EHIP;
    if (device < 0 || device >= ndev) return MDT_EINVAL;
    mdt_ctx* c = new mdt_ctx();
    c->device = device;
    c->w.resize(n_walkers);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        scratch_setup(device) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return MDT_EHIP;
    }
    *out = c;
    return MDT_OK;
}

void mdt_destroy(mdt_ctx* ctx) {
    if (!ctx) return;
    if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (mdt_fix* f : ctx->fixes) {
        // a fix whose init failed has no device hill log to flush
        if (ctx->stream && f->inited) { try { flush_hills(f); } catch (...) {} }
        f->ctx = nullptr;
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    (void)hipGetLastError();   // best-effort teardown leaves no error for the next launch to report
}

const char* mdt_last_error(const mdt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mdt_set_box(mdt_ctx* ctx, int walker, const double lo[3], const double hi[3], const int periodic[3]) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    for (int d = 0; d < 3; ++d) {
        if (!(hi[d] > lo[d])) throw Err(MDT_EINVAL, "box hi must exceed lo");
        if (periodic && !periodic[d]) throw Err(MDT_ENOTSUP, "non-periodic boxes: pass ghost-free lists; only periodic boxes are supported");
        a.lo[d] = lo[d]; a.hi[d] = hi[d];
    }
    a.has_box = true;
    ctx->committed = false;
    API_END
}

static void copy_in(void* dst, const void* src, size_t bytes, int on_device) {
    if (!bytes) return;
    if (on_device) HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    else std::memcpy(dst, src, bytes);
}

int mdt_bind_positions(mdt_ctx* ctx, int walker, const double* x, int64_t n_local, int64_t n_all, int on_device) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (n_local < 0 || n_all < n_local || (!x && n_all)) throw Err(MDT_EINVAL, "bad position arguments");
    if (ctx->committed && a.has_x && a.n_local == n_local && a.n_all == n_all) {
        a.x.resize((size_t)n_all * 3);
        copy_in(a.x.data(), x, sizeof(double) * 3 * n_all, on_device);
        return mdt_update_positions(ctx, walker, x, on_device);
    }
    a.n_local = n_local; a.n_all = n_all;
    a.x_src = nullptr;                          // a new layout detaches an engine buffer
    a.x.resize((size_t)n_all * 3);
    copy_in(a.x.data(), x, sizeof(double) * 3 * n_all, on_device);
    a.has_x = true;
    ctx->committed = false;
    API_END
}

int mdt_bind_neighbors(mdt_ctx* ctx, int walker, const int32_t* offsets, const int32_t* idx, int64_t n_center,
                       const int32_t* ghost_owner, int on_device) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions before neighbour lists");
    if (n_center != a.n_local) throw Err(MDT_EINVAL, "neighbour lists must cover every local atom (n_center == n_local)");
    a.off.resize((size_t)n_center + 1);
    copy_in(a.off.data(), offsets, sizeof(int32_t) * (n_center + 1), on_device);
    if (a.off[0] != 0) throw Err(MDT_EINVAL, "offsets[0] must be 0");
    for (int64_t i = 0; i < n_center; ++i)
        if (a.off[i + 1] < a.off[i]) throw Err(MDT_EINVAL, "offsets must be non-decreasing");
    a.nnz = a.off[n_center];
    a.idx.resize((size_t)a.nnz);
    copy_in(a.idx.data(), idx, sizeof(int32_t) * a.nnz, on_device);
    a.owner.resize((size_t)(a.n_all - a.n_local));
    copy_in(a.owner.data(), ghost_owner, sizeof(int32_t) * (a.n_all - a.n_local), on_device);
    a.has_nl = true;
    ctx->committed = false;
    API_END
}

int mdt_bind_group(mdt_ctx* ctx, int walker, const uint8_t* in_group) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions before the group");
    if (in_group) a.groups[std::string()].assign(in_group, in_group + a.n_local);
    else a.groups.erase(std::string());
    ctx->committed = false;
    API_END
}

int mdt_bind_group_named(mdt_ctx* ctx, int walker, const char* name, const uint8_t* in_group) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions before the group");
    const std::string nm = name ? name : "";
    if (nm == "all") throw Err(MDT_EINVAL, "group 'all' is every local atom and cannot be bound");
    if (in_group) a.groups[nm].assign(in_group, in_group + a.n_local);
    else a.groups.erase(nm);
    ctx->committed = false;
    API_END
}

int mdt_bind_types(mdt_ctx* ctx, int walker, const int32_t* type) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions before the atom types");
    if (type) a.types.assign(type, type + a.n_local);
    else a.types.clear();
    ctx->committed = false;
    API_END
}

int mdt_bind_share(mdt_ctx* ctx, int walker, const uint8_t* mine) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions before the share");
    if (mine) a.share.assign(mine, mine + a.n_local);
    else a.share.clear();
    ctx->committed = false;
    API_END
}

int mdt_bind_system(mdt_ctx* ctx, int walker, const mdt_system* sys) {
    API_BEGIN(ctx)
    if (!sys) throw Err(MDT_EINVAL, "null system");
    const int64_t nl = mdt_system_n_local(sys), na = mdt_system_n_all(sys), nz = mdt_system_nnz(sys);
    if (na < nl || na == 0) throw Err(MDT_ESTATE, "system not built (mdt_system_build)");
    double lo[3], hi[3];
    mdt_system_get_box(sys, lo, hi);
    int per[3] = {1, 1, 1};
    int rc = mdt_set_box(ctx, walker, lo, hi, per);
    if (rc) return rc;
    std::vector<double> x((size_t)na * 3);
    std::vector<int32_t> off((size_t)nl + 1), idx((size_t)nz), own((size_t)(na - nl));
    mdt_system_copy(sys, x.data(), off.data(), idx.data(), own.data());
    rc = mdt_bind_positions(ctx, walker, x.data(), nl, na, 0);
    if (rc) return rc;
    rc = mdt_bind_neighbors(ctx, walker, off.data(), idx.data(), nl, own.data(), 0);
    if (rc) return rc;
    API_END
}

int mdt_update_positions(mdt_ctx* ctx, int walker, const double* x, int on_device) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "positions not bound");
    if (!ctx->committed) {
        copy_in(a.x.data(), x, sizeof(double) * 3 * a.n_all, on_device);
        return MDT_OK;
    }
    if (on_device) {
        ctx->pack(walker, x);                   // straight from the caller's device buffer
    } else {
        // staged through the context's FP64 buffer (the walker's slice of it)
        double* st = ctx->x.p + 3 * a.atom_off + 3 * a.ghost_off;
        HIPCHK(hipMemcpyAsync(st, x, sizeof(double) * 3 * a.n_all, hipMemcpyHostToDevice, ctx->stream));
        ctx->pack(walker, st);
        HIPCHK(hipStreamSynchronize(ctx->stream));   // pageable source must stay valid
    }
    API_END
}

int mdt_attach_positions(mdt_ctx* ctx, int walker, const double* x) {
    API_BEGIN(ctx)
    WalkerIn& a = ctx->walker(walker);
    if (!a.has_x) throw Err(MDT_ESTATE, "bind positions (and lists) before attaching a position buffer");
    a.x_src = x;
    ctx->generation_attach++;
    API_END
}

int mdt_fix_create(mdt_ctx* ctx, int argc, const char* const* argv, mdt_fix** out) {
    API_BEGIN(ctx)
    if (!out) throw Err(MDT_EINVAL, "null output");
    *out = nullptr;
    std::unique_ptr<mdt_fix> fx(new mdt_fix());
    fx->ctx = ctx;
    parse_fix(fx.get(), argc, argv);
    fx->device = ctx->device;
    *out = fx.release();
    ctx->fixes.push_back(*out);
    API_END
}

void mdt_fix_destroy(mdt_fix* fx) {
    if (!fx) return;
    if (fx->ctx) {
        if (fx->inited) { try { flush_hills(fx); } catch (...) {} }
        auto& v = fx->ctx->fixes;
        v.erase(std::remove(v.begin(), v.end(), fx), v.end());
    }
    if (fx->device >= 0) (void)hipSetDevice(fx->device);
    delete fx;
    (void)hipGetLastError();
}

int mdt_fix_param(mdt_fix* fx, const char* key, double* out, int cap, int* n) {
    FIX_BEGIN(fx)
    const std::string k = key ? key : "";
    std::vector<double> v;
    if (k == "GAUSSIAN") v = {fx->sigma, fx->height0, fx->biasf};
    else if (k == "PACE") v = {(double)fx->pace};
    else if (k == "CV_dim") v = {(double)fx->cv_dim};
    else if (k == "TEMP") v = {fx->temp};
    else if (k == "BOLTZ") v = {fx->boltz};
    else if (k == "WT") v = {fx->wt ? 1.0 : 0.0};
    else if (k == "METAD_RESTART") v = {fx->restart ? 1.0 : 0.0};
    else if (k == "Gaussian_Hill_type") v = {(double)fx->hill_type};
    else if (k == "SIGMAS") v = fx->sigmas;
    else if (k == "DIM")
        for (int d = 0; d < fx->cv_dim; ++d) { v.push_back(fx->lo[d]); v.push_back(fx->hi[d]); v.push_back((double)fx->nb[d]); }
    else if (k.size() > 5 && k.compare(k.size() - 5, 5, ".KVEC") == 0) {
        // the k vectors of an SK CAL (explicit, (111) and seeded random), kx ky kz each
        const std::string name = k.substr(0, k.size() - 5);
        const CvH* c = nullptr;
        for (auto* q : fx->cvs) if (q->name == name) c = q;
        if (!c || c->kind != CV_SK) throw Err(MDT_EINVAL, fmt("no SK CAL named '%s'", name.c_str()));
        v = c->kvec;
    }
    else throw Err(MDT_EINVAL, fmt("unknown fix parameter '%s'", k.c_str()));
    if (n) *n = (int)v.size();
    if (out) for (int q = 0; q < cap && q < (int)v.size(); ++q) out[q] = v[q];
    API_END
}

int mdt_fix_init(mdt_fix* fx) {
    FIX_BEGIN(fx)
    use_device(fx->ctx);
    ensure_built(fx);
    API_END
}

int mdt_fix_post_force(mdt_fix* fx, int64_t step, double* f, int on_device) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    ensure_hills(fx, 1);
    for (auto& b : fx->hills_bound) b += 1;
    set_steps(fx, step);
    double* fd = ctx->f.p;
    if (f && on_device) fd = f;
    else if (f) HIPCHK(launch_fill(ctx->f.p, ctx->total_local * 3, 0.0, ctx->stream));
    enqueue_step(fx, fd);
    if (fx->hill_type == 3 && (step + 1) % fx->mlp_refit == 0) mlp_refit_all(fx);
    if (f && !on_device) {
        std::vector<double> tmp((size_t)ctx->total_local * 3);
        HIPCHK(hipMemcpyAsync(tmp.data(), ctx->f.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (size_t k = 0; k < tmp.size(); ++k) f[k] += tmp[k];
    }
    fetch_state(fx);
    flush_hills(fx);
    API_END
}

// Layout of a walker's reduction vector: per CV slot in CAL order, STEINH Q/LD (sum over
// the share's group atoms of q_l, atom count), SK (Re A, Im A per k, atom count).
static int64_t split_len(const mdt_fix* fx) {
    int64_t n = 0;
    for (auto* c : fx->cvs) n += c->kind == CV_SK ? 2 * (int64_t)(c->kvec.size() / 3) + 1 : 2;
    return n;
}

int mdt_fix_partials(mdt_fix* fx, int64_t step, double* out, int64_t cap, int64_t* n) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    fx->split = true;
    ensure_built(fx);
    const int nw = fx->nw();
    const int64_t L = split_len(fx);
    if (n) *n = L * nw;
    if (!out) return MDT_OK;
    if (cap < L * nw) throw Err(MDT_EINVAL, fmt("the reduction vector has %lld values", (long long)(L * nw)));
    set_steps(fx, step);
    Timer T;
    enqueue_values(fx, T);
    HIPCHK(launch_cv_sums(fx->red, fx->psum_of.p, fx->psums.p, ctx->stream));
    std::vector<double> ps((size_t)nw * kMaxSlotCV);
    HIPCHK(hipMemcpyAsync(ps.data(), fx->psums.p, sizeof(double) * ps.size(), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<std::vector<double>> amps(fx->cvs.size());
    for (size_t q = 0; q < fx->cvs.size(); ++q) {
        const CvH* c = fx->cvs[q];
        if (c->kind != CV_SK) continue;
        amps[q].resize((size_t)nw * c->kvec.size() / 3 * 2);
        HIPCHK(hipMemcpyAsync(amps[q].data(), c->amp.p, sizeof(double) * amps[q].size(), hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nw; ++k) {
        double* o = out + (size_t)k * L;
        for (size_t q = 0; q < fx->cvs.size(); ++q) {
            const CvH* c = fx->cvs[q];
            if (c->kind == CV_SK) {
                const size_t nk = c->kvec.size() / 3;
                std::memcpy(o, amps[q].data() + (size_t)k * nk * 2, sizeof(double) * nk * 2);
                o += nk * 2;
                *o++ = ctx->share_count("all", k);
            } else {
                *o++ = ps[(size_t)k * kMaxSlotCV + c->slot];
                *o++ = ctx->share_count(fx->envs[c->env]->group, k);
            }
        }
    }
    API_END
}

int mdt_fix_resume(mdt_fix* fx, int64_t step, const double* reduced, double* f, int on_device) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    if (!fx->split || fx->built_split != fx->split || fx->built_gen != ctx->generation)
        throw Err(MDT_ESTATE, "mdt_fix_resume follows mdt_fix_partials of the same step");
    if (!reduced) throw Err(MDT_EINVAL, "null reduction vector");
    const int nw = fx->nw();
    const int64_t L = split_len(fx);
    // the CV values, and 1 / N of the per-atom averages, from the sums over every share
    std::vector<double> direct((size_t)nw * kMaxSlotCV, 0.0), inv((size_t)nw * kMaxSlotCV, 0.0);
    hipStream_t st = ctx->stream;
    for (size_t q = 0; q < fx->cvs.size(); ++q) {
        CvH* c = fx->cvs[q];
        if (c->kind != CV_SK) continue;
        const size_t nk = c->kvec.size() / 3;
        std::vector<double> a((size_t)nw * nk * 2), nn(nw);
        for (int k = 0; k < nw; ++k) {
            const double* o = reduced + (size_t)k * L;
            for (size_t p = 0; p < q; ++p)
                o += fx->cvs[p]->kind == CV_SK ? 2 * (fx->cvs[p]->kvec.size() / 3) + 1 : 2;
            std::memcpy(&a[(size_t)k * nk * 2], o, sizeof(double) * nk * 2);
            nn[k] = o[nk * 2];
        }
        c->amp.upload(a.data(), a.size(), st);
        c->nnorm.upload(nn.data(), nn.size(), st);
    }
    for (int k = 0; k < nw; ++k) {
        const double* o = reduced + (size_t)k * L;
        for (auto* c : fx->cvs) {
            if (c->kind == CV_SK) {
                const size_t nk = c->kvec.size() / 3;
                const double N = o[nk * 2];
                double v = 0.0;
                for (size_t p = 0; p < nk; ++p) v += (o[2 * p] * o[2 * p] + o[2 * p + 1] * o[2 * p + 1]) / N;
                direct[(size_t)k * kMaxSlotCV + c->slot] = N > 0 ? v / (double)nk : 0.0;
                o += nk * 2 + 1;
            } else {
                const double N = o[1];
                direct[(size_t)k * kMaxSlotCV + c->slot] = N > 0 ? o[0] / N : 0.0;
                inv[(size_t)k * kMaxSlotCV + c->slot] = N > 0 ? 1.0 / N : 0.0;
                o += 2;
            }
        }
    }
    fx->direct.upload(direct.data(), direct.size(), st);
    fx->cv_inv.upload(inv.data(), inv.size(), st);
    for (auto* c : fx->cvs) {
        if (c->kind != CV_SK) continue;
        SkDev s{};
        s.nk = (int)(c->kvec.size() / 3);
        s.k = c->krev.p; s.kvec = c->kd.p; s.amp = c->amp.p; s.kpack = c->kpack.p; s.sterm = c->sterm.p;
        s.n_atoms = c->natoms.p; s.n_norm = c->nnorm.p; s.n_walkers = nw;
        HIPCHK(launch_sk_pack(s, st));
    }
    ensure_hills(fx, 1);
    for (auto& b : fx->hills_bound) b += 1;
    double* fd = ctx->f.p;
    if (f && on_device) fd = f;
    else if (f) HIPCHK(launch_fill(ctx->f.p, ctx->total_local * 3, 0.0, st));
    Timer T;
    enqueue_rest(fx, fd, T);
    if (f && !on_device) {
        std::vector<double> tmp((size_t)ctx->total_local * 3);
        HIPCHK(hipMemcpyAsync(tmp.data(), ctx->f.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (size_t k = 0; k < tmp.size(); ++k) f[k] += tmp[k];
    }
    (void)step;
    fetch_state(fx);
    flush_hills(fx);
    API_END
}

constexpr int kRunBatch = 8;   // steps per graph launch in mdt_fix_run

// The graph of k captured steps, built on first use and after a rebuild.
static hipGraphExec_t run_graph(mdt_fix* fx, int k) {
    mdt_ctx* ctx = fx->ctx;
    if (fx->graph_gen != ctx->generation + 1 || fx->graph_attach != ctx->generation_attach) {
        fx->graph_attach = ctx->generation_attach;
        for (auto& kv : fx->run_graphs) (void)hipGraphExecDestroy(kv.second);
        fx->run_graphs.clear();
        fx->graph_gen = ctx->generation + 1;
    }
    auto it = fx->run_graphs.find(k);
    if (it != fx->run_graphs.end()) return it->second;
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int s = 0; s < k; ++s) enqueue_step(fx, ctx->f.p);
    } catch (...) {
        hipGraph_t dummy = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
    }
    HIPCHK(hipStreamEndCapture(ctx->stream, &gr));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) throw Err(MDT_EHIP, fmt("hipGraphInstantiate: %s", hipGetErrorString(e)));
    fx->run_graphs[k] = ex;
    return ex;
}

int mdt_fix_run_steps(mdt_fix* fx, int64_t step0, int64_t nsteps, int steps_per_launch) {
    FIX_BEGIN(fx)
    if (steps_per_launch < 0 || steps_per_launch > 64) throw Err(MDT_EINVAL, "steps_per_launch must be in 0..64");
    if (nsteps < 0) throw Err(MDT_EINVAL, "nsteps must be >= 0");
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    const int64_t ndep = fx->pace > 0 ? nsteps / fx->pace + 1 : 0;
    ensure_hills(fx, ndep);
    for (auto& b : fx->hills_bound) b += ndep;
    set_steps(fx, step0);
    const int K = steps_per_launch;
    if (K == 0) {
        // plain kernel launches, step by step, stopping at refit steps as below
        for (int64_t s = 0; s < nsteps; ++s) {
            enqueue_step(fx, ctx->f.p);
            if (fx->hill_type == 3 && (step0 + s + 1) % fx->mlp_refit == 0) mlp_refit_all(fx);
        }
        fx->state_valid = false;
        return MDT_OK;
    }
    hipGraphExec_t g1 = run_graph(fx, 1);
    hipGraphExec_t gk = K > 1 ? run_graph(fx, K) : g1;
    // whole graphs of K steps, then single steps; with a learned bias each run of
    // launches ends at a refit step
    int64_t s = 0;
    while (s < nsteps) {
        int64_t seg = nsteps - s;
        if (fx->hill_type == 3) seg = std::min<int64_t>(seg, fx->mlp_refit - (step0 + s) % fx->mlp_refit);
        for (; seg >= K; seg -= K, s += K) HIPCHK(hipGraphLaunch(gk, ctx->stream));
        for (; seg > 0; --seg, ++s) HIPCHK(hipGraphLaunch(g1, ctx->stream));
        if (fx->hill_type == 3 && (step0 + s) % fx->mlp_refit == 0) mlp_refit_all(fx);
    }
    fx->state_valid = false;
    API_END
}

int mdt_fix_run(mdt_fix* fx, int64_t step0, int64_t nsteps) {
    return mdt_fix_run_steps(fx, step0, nsteps, kRunBatch);
}

int mdt_fix_profile(mdt_fix* fx, int64_t step0, int64_t nsteps, char* names, int names_cap,
                    double* ms, int ms_cap, int* nkern) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    const int64_t ndep = fx->pace > 0 ? nsteps / fx->pace + 1 : 0;
    ensure_hills(fx, ndep);
    for (auto& b : fx->hills_bound) b += ndep;
    set_steps(fx, step0);
    Timer T;
    T.on = true;
    for (int64_t s = 0; s < nsteps; ++s) {
        enqueue_step(fx, ctx->f.p, &T);
        T.collect();
    }
    std::string all;
    for (size_t k = 0; k < T.names.size(); ++k) {
        if (k) all += ",";
        all += T.names[k];
        if ((int)k < ms_cap) ms[k] = T.count[k] ? T.ms[k] / (double)T.count[k] : 0.0;
    }
    if (names && names_cap > 0) {
        strncpy(names, all.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = 0;
    }
    if (nkern) *nkern = (int)T.names.size();
    fx->state_valid = false;
    API_END
}

int mdt_fix_sync(mdt_fix* fx) {
    FIX_BEGIN(fx)
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    fetch_state(fx);
    flush_hills(fx);
    API_END
}

static const double* state_of(mdt_fix* fx, int w) {
    if (!fx->inited) throw Err(MDT_ESTATE, "no step has run yet");
    if (w < 0 || w >= fx->nw()) throw Err(MDT_EINVAL, "walker out of range");
    if (!fx->state_valid) fetch_state(fx);
    return fx->h_state.data() + (size_t)w * kStStride;
}

int mdt_fix_scalar(mdt_fix* fx, int walker, double* value) {
    FIX_BEGIN(fx)
    *value = state_of(fx, walker)[kStV];
    API_END
}

int mdt_fix_vector(mdt_fix* fx, int walker, int n, double* value) {
    FIX_BEGIN(fx)
    if (n < 0 || n >= fx->D()) { *value = 0.0; return MDT_OK; }
    *value = state_of(fx, walker)[kStS + n];
    API_END
}

static int find_cv(mdt_fix* fx, const char* cal_func) {
    std::string t = cal_func ? cal_func : "";
    size_t dot = t.find('.');
    std::string name = t.substr(0, dot);
    std::string func = dot == std::string::npos ? "AVE" : t.substr(dot + 1);
    for (size_t k = 0; k < fx->cvs.size(); ++k)
        if (fx->cvs[k]->name == name) {
            const CvH* c = fx->cvs[k];
            if (c->kind == CV_LQ) {
                if (dot != std::string::npos && lq_mode(func) != lq_mode(c->func.empty() ? "MEAN_SOLID" : c->func))
                    throw Err(MDT_EINVAL, fmt("We can't find the func %s.", func.c_str()));
                return (int)k;
            }
            if (func != "AVE" && c->kind != CV_DIST && !(c->kind == CV_SF && func == c->func))
                throw Err(MDT_EINVAL, fmt("We can't find the func %s.", func.c_str()));
            return (int)k;
        }
    throw Err(MDT_EINVAL, fmt("no CAL named '%s'", name.c_str()));
}

int mdt_fix_cv_value(mdt_fix* fx, int walker, const char* cal_func, double* value) {
    FIX_BEGIN(fx)
    int c = find_cv(fx, cal_func);
    *value = state_of(fx, walker)[kStCv + fx->cvs[c]->slot];
    API_END
}

int mdt_fix_cv_weight(mdt_fix* fx, int walker, const char* cal_func, double* value) {
    FIX_BEGIN(fx)
    int c = find_cv(fx, cal_func);
    state_of(fx, walker);
    double w = 0;
    HIPCHK(hipMemcpy(&w, fx->wcv.p + (size_t)walker * kWcvStride + fx->cvs[c]->slot, sizeof(double), hipMemcpyDeviceToHost));
    *value = w;
    API_END
}

int mdt_fix_extract_peratom(mdt_fix* fx, int walker, const char* key, double* out) {
    FIX_BEGIN(fx)
    if (!fx->inited) throw Err(MDT_ESTATE, "no step has run yet");
    mdt_ctx* ctx = fx->ctx;
    WalkerIn& a = ctx->walker(walker);
    std::string k = key ? key : "";
    size_t dot = k.find('.');
    if (dot == std::string::npos) throw Err(MDT_EINVAL, "key must be CAL.prop");
    std::string name = k.substr(0, dot), prop = k.substr(dot + 1);
    CvH* c = nullptr;
    for (auto* q : fx->cvs) if (q->name == name) c = q;
    if (!c) throw Err(MDT_EINVAL, fmt("no CAL named '%s'", name.c_str()));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int64_t n = a.n_local, o = a.atom_off;
    if (c->kind == CV_SK || c->kind == CV_DIST || c->kind == CV_SF || c->kind == CV_LQ)
        throw Err(MDT_EINVAL, "SK, DISTANCE, STRU_FACTOR and STEINH L have no per-atom export");
    EnvH* e = fx->envs[c->env];
    if (prop == "stein_q" || (prop == "local_q" && c->kind == CV_LD)) {
        std::vector<float> v((size_t)n);
        if (prop == "stein_q") {
            int li = rec_li(e->lmask, c->l);
            HIPCHK(hipMemcpy(v.data(), e->ql.p + (size_t)li * ctx->total_local + o, sizeof(float) * n, hipMemcpyDeviceToHost));
        } else {
            HIPCHK(hipMemcpy(v.data(), c->lq.p + o, sizeof(float) * n, hipMemcpyDeviceToHost));
        }
        for (int64_t i = 0; i < n; ++i) out[i] = v[i];
    } else if (prop == "nsel" || prop == "count_rc") {
        std::vector<int32_t> v((size_t)n);
        HIPCHK(hipMemcpy(v.data(), (prop == "nsel" ? e->nsel.p : e->cnt.p) + o, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) out[i] = v[i];
    } else if (prop == "kept") {
        // the set the last step's k_env kept (its selection mask), as the neighbour
        // indices of each atom in list order, N per atom, -1 after the last
        // mask positions index the list the step walked: the bound list, or the near list
        // in place at the bound offsets
        const int N = e->N;
        std::vector<uint64_t> m((size_t)n * 4);
        HIPCHK(hipMemcpy(m.data(), e->selmask.p + (size_t)o * 4, sizeof(uint64_t) * m.size(), hipMemcpyDeviceToHost));
        std::vector<int32_t> cnt((size_t)std::max<int64_t>(n, 1)), ni((size_t)std::max<int64_t>(a.nnz, 1));
        if (ctx->near_on()) {
            HIPCHK(hipMemcpy(cnt.data(), ctx->ncnt.p + o, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
            if (a.nnz) HIPCHK(hipMemcpy(ni.data(), ctx->nidx.p + a.nnz_off, sizeof(int32_t) * a.nnz, hipMemcpyDeviceToHost));
        } else {
            for (int64_t i = 0; i < n; ++i) cnt[i] = a.off[i + 1] - a.off[i];
            if (a.nnz) HIPCHK(hipMemcpy(ni.data(), ctx->idx.p + a.nnz_off, sizeof(int32_t) * a.nnz, hipMemcpyDeviceToHost));
        }
        for (int64_t i = 0; i < n; ++i) {
            int k = 0;
            const int64_t c0 = a.off[i];
            for (int64_t p = 0; p < cnt[i]; ++p)
                if ((m[(size_t)i * 4 + (p >> 6)] >> (p & 63)) & 1ull) {
                    if (k >= N) throw Err(MDT_EINVAL, "selection mask holds more than N entries");
                    int32_t j = ni[(size_t)(c0 + p)];      // context index -> the walker's own
                    j = j < ctx->total_local ? (int32_t)(j - a.atom_off)
                                             : (int32_t)(a.n_local + (j - ctx->total_local - a.ghost_off));
                    out[i * N + k++] = j;
                }
            for (; k < N; ++k) out[i * N + k] = -1;
        }
    } else if (prop == "qlm") {
        // q_lm of degree l: 2l+1 floats per atom (Re q_l0, Re/Im q_lm m>=1)
        const int RS = rec_stride(e->lmask), O = rec_off(e->lmask, c->l), W = 2 * c->l + 1;  // q block of l
        std::vector<float> v((size_t)n * RS);
        HIPCHK(hipMemcpy(v.data(), e->rec.p + (size_t)o * RS, sizeof(float) * v.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) for (int q = 0; q < W; ++q) out[i * W + q] = v[(size_t)i * RS + O + q];
    } else {
        throw Err(MDT_EINVAL, fmt("unknown per-atom property '%s'", prop.c_str()));
    }
    API_END
}

int mdt_fix_selection(mdt_fix* fx, int walker, const char* cal_name, int32_t* sel_idx, int32_t* count_rc, int32_t* n_sel) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    WalkerIn& a = ctx->walker(walker);
    CvH* c = nullptr;
    for (auto* q : fx->cvs) if (q->name == (cal_name ? cal_name : "")) c = q;
    if (!c || c->kind == CV_SK || c->kind == CV_DIST || c->kind == CV_SF || c->kind == CV_LQ)
        throw Err(MDT_EINVAL, "selection export needs a STEINH Q or LD CAL name");
    EnvH* e = fx->envs[c->env];
    const int N = e->N;
    const int64_t n = a.n_local, NT = ctx->total_local;
    // the scan and insertion of the step's k_env, run for every atom of every walker
    DBuf<int32_t> si, cr;
    si.alloc((size_t)std::max<int64_t>(NT, 1) * N); cr.alloc(std::max<int64_t>(NT, 1));
    ctx->refresh_near();                 // positions may have moved since the last step
    Geo g = ctx->geo_near();
    g.group = ctx->group_ptr(e->group);
    EnvDev d = e->dev;
    d.sel_out = si.p; d.sel_cnt = cr.p;
    HIPCHK(launch_select_export(g, d, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> cnt((size_t)std::max<int64_t>(n, 1));
    HIPCHK(hipMemcpy(cnt.data(), cr.p + a.atom_off, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (sel_idx) {
        std::vector<int32_t> v((size_t)n * N);
        HIPCHK(hipMemcpy(v.data(), si.p + (size_t)a.atom_off * N, sizeof(int32_t) * v.size(), hipMemcpyDeviceToHost));
        for (auto& j : v) {
            if (j < 0) continue;
            if (j < ctx->total_local) j = (int32_t)(j - a.atom_off);
            else j = (int32_t)(a.n_local + (j - ctx->total_local - a.ghost_off));
        }
        std::memcpy(sel_idx, v.data(), sizeof(int32_t) * v.size());
    }
    if (count_rc) std::memcpy(count_rc, cnt.data(), sizeof(int32_t) * n);
    if (n_sel) for (int64_t i = 0; i < n; ++i) n_sel[i] = std::min(cnt[i], N);
    API_END
}

int mdt_fix_get_forces(mdt_fix* fx, int walker, double* f_host) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    WalkerIn& a = ctx->walker(walker);
    if (!ctx->committed) throw Err(MDT_ESTATE, "no forces yet");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(f_host, ctx->f.p + 3 * a.atom_off, sizeof(double) * 3 * a.n_local, hipMemcpyDeviceToHost));
    API_END
}

int mdt_fix_mlp_refit(mdt_fix* fx, int walker) {
    FIX_BEGIN(fx)
    use_device(fx->ctx);
    ensure_built(fx);
    if (fx->hill_type != 3) throw Err(MDT_EINVAL, "no learned bias (Gaussian_Hill_type 3)");
    fx->ctx->walker(walker);
    mlp_refit(fx, walker);
    fx->state_valid = false;
    API_END
}

int mdt_fix_mlp_status(mdt_fix* fx, int walker, double* err, int32_t* ok, int64_t* fits, int64_t* fallbacks,
                       int64_t* n_fit) {
    FIX_BEGIN(fx)
    use_device(fx->ctx);
    if (fx->hill_type != 3 || !fx->inited) throw Err(MDT_EINVAL, "no learned bias (Gaussian_Hill_type 3)");
    fx->ctx->walker(walker);
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    if (err) HIPCHK(hipMemcpy(err, fx->mlp_err.p + walker, sizeof(double), hipMemcpyDeviceToHost));
    if (ok) HIPCHK(hipMemcpy(ok, fx->mlp_ok.p + walker, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (fits) HIPCHK(hipMemcpy(fits, fx->mlp_fits.p + walker, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (fallbacks) HIPCHK(hipMemcpy(fallbacks, fx->mlp_fallbacks.p + walker, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (n_fit) {
        MlpNet h;
        HIPCHK(hipMemcpy(&h, fx->mlp_stage.p + walker, sizeof(MlpNet), hipMemcpyDeviceToHost));
        *n_fit = h.n_fit;
    }
    API_END
}

int mdt_fix_mlp_params(mdt_fix* fx, int walker, int live, float* params, double* norm) {
    FIX_BEGIN(fx)
    use_device(fx->ctx);
    if (fx->hill_type != 3 || !fx->inited) throw Err(MDT_EINVAL, "no learned bias (Gaussian_Hill_type 3)");
    fx->ctx->walker(walker);
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    MlpNet h;
    HIPCHK(hipMemcpy(&h, (live ? fx->mlp_live.p : fx->mlp_stage.p) + walker, sizeof(MlpNet), hipMemcpyDeviceToHost));
    if (params) std::memcpy(params, h.p, sizeof(float) * kMlpP);
    if (norm) {
        for (int d = 0; d < 3; ++d) { norm[d] = h.in_shift[d]; norm[3 + d] = h.in_scale[d]; }
        norm[6] = h.out_scale; norm[7] = h.out_shift;
    }
    API_END
}

// standalone network: params (4481 floats) + norm[8] -> a device MlpNet
static void upload_net(mdt_ctx* ctx, const float* params, const double* norm, DBuf<MlpNet>& net) {
    MlpNet h{};
    std::memcpy(h.p, params, sizeof(float) * kMlpP);
    auto bf = [](float v) {   // round to nearest even BF16, as the device does
        uint32_t u;
        std::memcpy(&u, &v, 4);
        uint32_t lsb = (u >> 16) & 1u;
        u += 0x7FFFu + lsb;
        return (uint16_t)(u >> 16);
    };
    for (int k = 0; k < kMlpH * kMlpH; ++k) {
        const uint16_t hi = bf(params[kW2 + k]);
        uint32_t hb = (uint32_t)hi << 16;
        float hv;
        std::memcpy(&hv, &hb, 4);
        h.w2bf[k] = (short)hi;
        h.w2lo[k] = (short)bf(params[kW2 + k] - hv);
    }
    for (int d = 0; d < 3; ++d) { h.in_shift[d] = norm[d]; h.in_scale[d] = norm[3 + d]; }
    h.out_scale = norm[6]; h.out_shift = norm[7];
    h.n_fit = 0;
    net.alloc(1);
    net.upload(&h, 1, ctx->stream);
}

static void upload_points(mdt_ctx* ctx, int64_t np, const double* s, DBuf<float4>& ds) {
    std::vector<float4> h((size_t)std::max<int64_t>(np, 1));
    for (int64_t p = 0; p < np; ++p) h[p] = make_float4((float)s[3 * p], (float)s[3 * p + 1], (float)s[3 * p + 2], 0.0f);
    ds.alloc(h.size());
    ds.upload(h.data(), h.size(), ctx->stream);
}

int mdt_neighbor_build(mdt_ctx* ctx, const double* x_local, int64_t n, const double lo[3], const double hi[3],
                       double rg, int on_device, int64_t* n_all, int64_t* nnz) {
    API_BEGIN(ctx)
    if (!ctx) return MDT_EINVAL;
    if (n < 0 || (n > 0 && !x_local) || !lo || !hi) throw Err(MDT_EINVAL, "null positions or box");
    if (!(rg > 0)) throw Err(MDT_EINVAL, "the list cutoff must be positive");
    if (26 * n > (int64_t)INT32_MAX) throw Err(MDT_ENOTSUP, "more than 2^31 / 26 local atoms");
    use_device(ctx);
    hipStream_t st = ctx->stream;
    NlBox b{};
    int64_t nbins = 1;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = lo[d]; b.hi[d] = hi[d]; b.L[d] = hi[d] - lo[d];
        if (!(b.L[d] > 2.0 * rg)) throw Err(MDT_EINVAL, "box edge must exceed twice the list cutoff");
        b.elo[d] = lo[d] - rg;
        const double ext = b.L[d] + 2 * rg;
        b.nb[d] = std::max(1, (int)std::floor(ext / rg));
        b.bw[d] = ext / b.nb[d];
        nbins *= b.nb[d];
    }
    if (nbins + 1 > (int64_t)INT32_MAX) throw Err(MDT_ENOTSUP, "too many bins");
    b.rg = rg;
    ctx->nl_ok = false;
    DBuf<double> dx;
    dx.alloc((size_t)std::max<int64_t>(3 * n, 1));
    if (n > 0) HIPCHK(hipMemcpyAsync(dx.p, x_local, sizeof(double) * 3 * n,
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    DBuf<uint8_t> tmp;
    // ghosts: flags per (shift, atom), compacted in that order
    const int64_t nq = 26 * n;
    DBuf<int32_t> flag, pos;
    flag.alloc((size_t)std::max<int64_t>(nq, 1)); pos.alloc((size_t)std::max<int64_t>(nq, 1));
    int64_t ng = 0;
    if (n > 0) {
        HIPCHK(nl_ghost_flags(dx.p, n, b, flag.p, st));
        cub_call(tmp, st, [&](void* t, size_t& by) { return nl_scan_i32(t, by, flag.p, pos.p, nq, st); });
        int32_t last[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&last[0], pos.p + nq - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&last[1], flag.p + nq - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        ng = (int64_t)last[0] + last[1];
    }
    const int64_t na = n + ng;
    ctx->nl_xall.alloc((size_t)std::max<int64_t>(3 * na, 1));
    ctx->nl_owner.alloc((size_t)std::max<int64_t>(ng, 1));
    HIPCHK(nl_ghost_scatter(dx.p, n, b, flag.p, pos.p, ctx->nl_xall.p, ctx->nl_owner.p, st));
    // bins: (bin, index) pairs sorted stably by bin, so a bin lists its atoms in ascending index
    DBuf<int32_t> bin_of, val, bin_count, bin_start, key_sorted, bin_atoms;
    bin_of.alloc((size_t)std::max<int64_t>(na, 1)); val.alloc((size_t)std::max<int64_t>(na, 1));
    key_sorted.alloc((size_t)std::max<int64_t>(na, 1)); bin_atoms.alloc((size_t)std::max<int64_t>(na, 1));
    bin_count.alloc((size_t)nbins + 1); bin_start.alloc((size_t)nbins + 1);
    HIPCHK(hipMemsetAsync(bin_count.p, 0, sizeof(int32_t) * (nbins + 1), st));
    HIPCHK(nl_bin(ctx->nl_xall.p, na, b, bin_of.p, val.p, bin_count.p, st));
    cub_call(tmp, st, [&](void* t, size_t& by) { return nl_scan_i32(t, by, bin_count.p, bin_start.p, nbins + 1, st); });
    int end_bit = 1;
    while (end_bit < 31 && ((int64_t)1 << end_bit) < nbins) ++end_bit;
    if (na > 0)
        cub_call(tmp, st, [&](void* t, size_t& by) {
            return nl_sort_pairs(t, by, bin_of.p, key_sorted.p, val.p, bin_atoms.p, na, end_bit, st);
        });
    // lists: count, offsets, fill
    DBuf<int64_t> cnt, off64;
    cnt.alloc((size_t)n + 1); off64.alloc((size_t)n + 1);
    HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * (n + 1), st));
    HIPCHK(nl_list_count(ctx->nl_xall.p, n, b, bin_of.p, bin_start.p, bin_atoms.p, cnt.p, st));
    cub_call(tmp, st, [&](void* t, size_t& by) { return nl_scan_i64(t, by, cnt.p, off64.p, n + 1, st); });
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, off64.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (total > (int64_t)INT32_MAX) throw Err(MDT_ENOTSUP, "neighbour list exceeds 2^31 entries");
    ctx->nl_idx.alloc((size_t)std::max<int64_t>(total, 1));
    ctx->nl_off.alloc((size_t)n + 1);
    HIPCHK(nl_list_fill(ctx->nl_xall.p, n, b, bin_of.p, bin_start.p, bin_atoms.p, off64.p, ctx->nl_idx.p,
                        ctx->nl_off.p, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->nl_n = n; ctx->nl_n_all = na; ctx->nl_nnz = total;
    for (int d = 0; d < 3; ++d) { ctx->nl_lo[d] = lo[d]; ctx->nl_hi[d] = hi[d]; }
    ctx->nl_ok = true;
    if (n_all) *n_all = na;
    if (nnz) *nnz = total;
    API_END
}

int mdt_neighbor_fetch(mdt_ctx* ctx, double* x_all, int32_t* offsets, int32_t* idx, int32_t* ghost_owner) {
    API_BEGIN(ctx)
    if (!ctx) return MDT_EINVAL;
    if (!ctx->nl_ok) throw Err(MDT_ESTATE, "no device neighbour build on this context");
    use_device(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (x_all) HIPCHK(hipMemcpy(x_all, ctx->nl_xall.p, sizeof(double) * 3 * ctx->nl_n_all, hipMemcpyDeviceToHost));
    if (offsets) HIPCHK(hipMemcpy(offsets, ctx->nl_off.p, sizeof(int32_t) * (ctx->nl_n + 1), hipMemcpyDeviceToHost));
    if (idx && ctx->nl_nnz) HIPCHK(hipMemcpy(idx, ctx->nl_idx.p, sizeof(int32_t) * ctx->nl_nnz, hipMemcpyDeviceToHost));
    if (ghost_owner && ctx->nl_n_all > ctx->nl_n)
        HIPCHK(hipMemcpy(ghost_owner, ctx->nl_owner.p, sizeof(int32_t) * (ctx->nl_n_all - ctx->nl_n),
                         hipMemcpyDeviceToHost));
    API_END
}

int mdt_neighbor_bind(mdt_ctx* ctx, int walker) {
    if (!ctx) return MDT_EINVAL;
    if (!ctx->nl_ok) { ctx->err = "no device neighbour build on this context"; return MDT_ESTATE; }
    const int periodic[3] = {1, 1, 1};
    int rc = mdt_set_box(ctx, walker, ctx->nl_lo, ctx->nl_hi, periodic);
    if (rc == MDT_OK) rc = mdt_bind_positions(ctx, walker, ctx->nl_xall.p, ctx->nl_n, ctx->nl_n_all, 1);
    if (rc == MDT_OK)
        rc = mdt_bind_neighbors(ctx, walker, ctx->nl_off.p, ctx->nl_idx.p, ctx->nl_n, ctx->nl_owner.p, 1);
    return rc;
}

int mdt_fes_sum(mdt_ctx* ctx, int D, int64_t nh, const double* hills, const double* x, int n0,
                const double* y, int n1, const double* lo, const double* d, double* out) {
    API_BEGIN(ctx)
    if (!ctx) return MDT_EINVAL;
    if (D != 1 && D != 2) throw Err(MDT_ENOTSUP, "FES reconstruction covers 1 and 2 dimensions, as the reference script does");
    if (nh < 0 || n0 < 1 || (D == 2 && n1 < 1)) throw Err(MDT_EINVAL, "empty FES grid or negative hill count");
    if ((nh > 0 && !hills) || !x || (D == 2 && (!y || !lo || !d)) || !out) throw Err(MDT_EINVAL, "null array");
    use_device(ctx);
    const int64_t nout = D == 2 ? (int64_t)n0 * n1 : n0;
    DBuf<double> dh, dx, dy, dout;
    dh.alloc((size_t)std::max<int64_t>(nh, 1) * (D + 3));
    dh.upload(hills, (size_t)nh * (D + 3), ctx->stream);
    dx.alloc(n0); dx.upload(x, n0, ctx->stream);
    dy.alloc(D == 2 ? n1 : 1);
    if (D == 2) dy.upload(y, n1, ctx->stream);
    dout.alloc((size_t)nout);
    HIPCHK(launch_fes_sum(D, nh, dh.p, dx.p, n0, dy.p, D == 2 ? n1 : 1, D == 2 ? lo[0] : 0.0, D == 2 ? lo[1] : 0.0,
                          D == 2 ? d[0] : 1.0, D == 2 ? d[1] : 1.0, dout.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, dout.p, sizeof(double) * nout, hipMemcpyDeviceToHost));
    API_END
}

int mdt_fes_marginal(mdt_ctx* ctx, const double* F, int n0, int n1, int axis, double kT, double bin, double* out) {
    API_BEGIN(ctx)
    if (!ctx) return MDT_EINVAL;
    if (n0 < 1 || n1 < 1 || (axis != 0 && axis != 1) || !F || !out) throw Err(MDT_EINVAL, "bad marginal arguments");
    if (!(kT > 0) || !(bin > 0)) throw Err(MDT_EINVAL, "kT and the bin width must be positive");
    use_device(ctx);
    DBuf<double> dF, dout;
    dF.alloc((size_t)n0 * n1); dF.upload(F, (size_t)n0 * n1, ctx->stream);
    const int nout = axis == 1 ? n0 : n1;
    dout.alloc(nout);
    HIPCHK(launch_fes_marginal(dF.p, n0, n1, axis, kT, bin, dout.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, dout.p, sizeof(double) * nout, hipMemcpyDeviceToHost));
    API_END
}

int mdt_mlp_forward(mdt_ctx* ctx, const float* params, const double* norm, int64_t np, const double* s, double* V) {
    API_BEGIN(ctx)
    use_device(ctx);
    DBuf<MlpNet> net;
    upload_net(ctx, params, norm, net);
    DBuf<float4> ds;
    upload_points(ctx, np, s, ds);
    DBuf<double> dv;
    dv.alloc((size_t)std::max<int64_t>(np, 1));
    HIPCHK(mlp_forward(net.p, np, ds.p, nullptr, dv.p, nullptr, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(V, dv.p, sizeof(double) * np, hipMemcpyDeviceToHost));
    API_END
}

int mdt_mlp_loss_grad(mdt_ctx* ctx, const float* params, const double* norm, int64_t np, const double* s,
                      const double* v, float* grad) {
    API_BEGIN(ctx)
    use_device(ctx);
    DBuf<MlpNet> net;
    upload_net(ctx, params, norm, net);
    DBuf<float4> ds;
    upload_points(ctx, np, s, ds);
    DBuf<double> dv;
    dv.alloc((size_t)std::max<int64_t>(np, 1));
    dv.upload(v, (size_t)np, ctx->stream);
    DBuf<float> gpart, gsum, am, av;
    gpart.alloc((size_t)std::max<int64_t>(mlp_grad_rows(np), 1) * kMlpP);
    gsum.alloc(kMlpP); am.alloc(kMlpP); av.alloc(kMlpP);
    MlpDev m{};
    m.D = 3; m.stage = net.p; m.live = net.p; m.adam_m = am.p; m.adam_v = av.p;
    HIPCHK(mlp_train_step(m, 0, 0, np, np, ds.p, dv.p, gpart.p, 0.0f, gsum.p, 1, 0, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(grad, gsum.p, sizeof(float) * kMlpP, hipMemcpyDeviceToHost));
    API_END
}

int mdt_fix_virial(mdt_fix* fx, int walker, const int32_t* image, double v[6]) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    WalkerIn& a = ctx->walker(walker);
    if (!ctx->committed) throw Err(MDT_ESTATE, "no forces yet");
    const int n = (int)a.n_local;
    const int nb = (n + kBlock - 1) / kBlock;
    DBuf<double> part;
    part.alloc((size_t)std::max(nb, 1) * 6);
    DBuf<int32_t> img;
    if (image) { img.alloc((size_t)std::max(n, 1) * 3); img.upload(image, (size_t)n * 3, ctx->stream); }
    const double L[3] = {a.hi[0] - a.lo[0], a.hi[1] - a.lo[1], a.hi[2] - a.lo[2]};
    ensure_x4(ctx);
    HIPCHK(launch_virial(ctx->x4.p, ctx->f.p, image ? img.p : nullptr, (int)a.atom_off, n, L, part.p, ctx->stream));
    std::vector<double> h((size_t)std::max(nb, 1) * 6);
    HIPCHK(hipMemcpyAsync(h.data(), part.p, sizeof(double) * nb * 6, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < 6; ++k) v[k] = 0.0;
    for (int b = 0; b < nb; ++b) for (int k = 0; k < 6; ++k) v[k] += h[(size_t)b * 6 + k];
    API_END
}

int mdt_fix_virial_noalloc(mdt_fix* fx, int walker, const int32_t* image, double* v) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    WalkerIn& a = ctx->walker(walker);
    if (!ctx->committed) throw Err(MDT_ESTATE, "no forces yet");
    if (!v) throw Err(MDT_EINVAL, "v must be a device pointer to 6 doubles");
    const int64_t n = a.n_local;
    if (n > (int64_t)kScratchBlocks * kBlock)
        throw Err(MDT_EINVAL, fmt("walker %d has %lld local atoms; the allocation-free virial takes at most %lld",
                                  walker, (long long)n, (long long)kScratchBlocks * kBlock));
    double* part = scratch_ptr(ctx->device);
    if (!part) throw Err(MDT_ESTATE, "device scratch was not set up (mdt_create)");
    const int nb = (int)((n + kBlock - 1) / kBlock);
    const double L[3] = {a.hi[0] - a.lo[0], a.hi[1] - a.lo[1], a.hi[2] - a.lo[2]};
    ensure_x4(ctx);
    HIPCHK(launch_virial(ctx->x4.p, ctx->f.p, image, (int)a.atom_off, (int)n, L, part, ctx->stream));
    HIPCHK(launch_virial_sum(part, nb, v, ctx->stream));
    API_END
}

int mdt_fix_zero_forces(mdt_fix* fx) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    if (!ctx->committed) ctx->commit();
    HIPCHK(launch_fill(ctx->f.p, ctx->total_local * 3, 0.0, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    API_END
}

int mdt_fix_bias_eval(mdt_fix* fx, int walker, int64_t np, const double* s, double* V, double* dVds) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    ctx->walker(walker);
    const int D = fx->D();
    DBuf<double> ds, dv, dd;
    ds.alloc((size_t)std::max<int64_t>(np, 1) * D); dv.alloc(std::max<int64_t>(np, 1)); dd.alloc((size_t)std::max<int64_t>(np, 1) * D);
    ds.upload(s, (size_t)np * D, ctx->stream);
    HIPCHK(launch_bias_eval(fx->red, walker, np, ds.p, dv.p, dd.p, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (V) HIPCHK(hipMemcpy(V, dv.p, sizeof(double) * np, hipMemcpyDeviceToHost));
    if (dVds) HIPCHK(hipMemcpy(dVds, dd.p, sizeof(double) * np * D, hipMemcpyDeviceToHost));
    API_END
}

int mdt_fix_bias_add_hills(mdt_fix* fx, int walker, int64_t nh, const double* s, const double* h, double sigma) {
    FIX_BEGIN(fx)
    mdt_ctx* ctx = fx->ctx;
    use_device(ctx);
    ensure_built(fx);
    ctx->walker(walker);
    if (!fx->grid_bias()) { ensure_hills(fx, nh); fx->hills_bound[walker] += nh; }
    const int D = fx->D();
    DBuf<double> ds, dh;
    ds.alloc((size_t)std::max<int64_t>(nh, 1) * D); dh.alloc(std::max<int64_t>(nh, 1));
    ds.upload(s, (size_t)nh * D, ctx->stream);
    dh.upload(h, (size_t)nh, ctx->stream);
    HIPCHK(launch_bias_add(fx->red, walker, nh, ds.p, dh.p, sigma, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (!fx->grid_bias()) {
        int64_t n = 0;
        HIPCHK(hipMemcpy(&n, fx->nhills.p + walker, sizeof(int64_t), hipMemcpyDeviceToHost));
        fx->hills_written[walker] = n;   // externally supplied hills are not echoed to HILLS
    }
    API_END
}

int mdt_fix_grid(mdt_fix* fx, int walker, double* out, int64_t n) {
    FIX_BEGIN(fx)
    ensure_built(fx);
    if (!fx->grid_bias()) throw Err(MDT_EINVAL, "this bias has no grid");
    if (n != fx->red.grid_size) throw Err(MDT_EINVAL, fmt("grid has %lld bins", (long long)fx->red.grid_size));
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    HIPCHK(hipMemcpy(out, fx->grid.p + (size_t)walker * fx->red.grid_size, sizeof(double) * n, hipMemcpyDeviceToHost));
    API_END
}

int mdt_fix_hills(mdt_fix* fx, int walker, int64_t* nh, double* rec, int64_t cap) {
    FIX_BEGIN(fx)
    ensure_built(fx);
    HIPCHK(hipStreamSynchronize(fx->ctx->stream));
    int64_t n = 0;
    HIPCHK(hipMemcpy(&n, fx->nhills.p + walker, sizeof(int64_t), hipMemcpyDeviceToHost));
    *nh = n;
    const int D = fx->D();
    if (rec && cap > 0) {
        int64_t m = std::min(n, cap);
        HIPCHK(hipMemcpy(rec, fx->hills.p + (size_t)walker * fx->hills_cap * (D + 3), sizeof(double) * m * (D + 3), hipMemcpyDeviceToHost));
    }
    API_END
}

}  // extern "C"
