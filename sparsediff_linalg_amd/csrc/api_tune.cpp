// Tuning table of the product plans (include/sdl_device.h, sdl_tune_*): the
// symbolic phase's accumulator thresholds, the row-accumulator refill's row
// cuts and the batched refill's launch shape, recorded per (GPU architecture,
// shape signature) and consulted by every plan creation.  sdl_btda_tune fills
// it by timing candidates on the caller's product; the table saves to and
// loads from a text file closed by an FNV-1a checksum.
//
// The reference has no tuner: its symbolic and numeric phases are one fixed
// CPU algorithm (src/sparse/gustavson.c:20-150).  The thresholds tuned here
// choose between kernels that produce identical structures and values.
#include "../../include/sdl_device.h"
#include "btda_plan.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace sdl {

thread_local const SymTune *g_force_tune = nullptr;

bool tune_valid(const SymTune &t)
{
    const bool pow2 = t.hash_max > 0 && (t.hash_max & (t.hash_max - 1)) == 0;
    return pow2 && t.hash_max >= 1024 && t.hash_max <= 8192 && t.dens >= 1 && t.dens <= 64 &&
           t.rowacc_cut >= 64 && t.rowacc_cut <= 1024 && t.rank_cut > t.rowacc_cut &&
           t.rank_cut <= 65536;
}

namespace {

struct Entry
{
    std::string arch;
    ShapeSig sig;
    SymTune tune;
    RefillCfg refill;  // variant < 0: none recorded
};

std::mutex g_mu;
std::vector<Entry> g_table;

// the current device's architecture without its feature suffixes
// ("gfx950:sramecc+:xnack-" -> "gfx950")
std::string current_arch()
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
        return "unknown";
    std::string a = p.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

bool same_sig(const ShapeSig &a, const ShapeSig &b)
{
    return a.lg_rows == b.lg_rows && a.lg_cols == b.lg_cols && a.lg_tpr == b.lg_tpr;
}

Entry *find(const std::string &arch, const ShapeSig &sig)
{
    for (Entry &e : g_table)
        if (e.arch == arch && same_sig(e.sig, sig)) return &e;
    return nullptr;
}

const char HEADER[] = "sdl-tune-table 1\n";

bool refill_valid(const RefillCfg &c)
{
    if (c.variant < 0) return c.variant == -1;
    for (const RefillCfg &k : refill_candidates())
        if (k.variant == c.variant && k.rchunk == c.rchunk) return true;
    return false;
}

} // namespace

bool tune_lookup(const ShapeSig &sig, SymTune *tune, RefillCfg *refill)
{
    const std::string arch = current_arch();
    std::lock_guard<std::mutex> lk(g_mu);
    const Entry *e = find(arch, sig);
    if (!e) return false;
    if (tune) *tune = e->tune;
    if (refill) *refill = e->refill;
    return true;
}

void tune_record(const ShapeSig &sig, const SymTune *tune, const RefillCfg *refill)
{
    const std::string arch = current_arch();
    std::lock_guard<std::mutex> lk(g_mu);
    Entry *e = find(arch, sig);
    if (!e)
    {
        g_table.push_back(Entry{arch, sig, SymTune(), RefillCfg()});
        e = &g_table.back();
    }
    if (tune) e->tune = *tune;
    if (refill) e->refill = *refill;
}

} // namespace sdl

namespace {

void to_c(const sdl::SymTune &t, sdl_sym_tune *o)
{
    o->hash_max = t.hash_max;
    o->dens = t.dens;
    o->rowacc_cut = t.rowacc_cut;
    o->rank_cut = t.rank_cut;
}

sdl::SymTune from_c(const sdl_sym_tune &o)
{
    sdl::SymTune t;
    t.hash_max = o.hash_max;
    t.dens = o.dens;
    t.rowacc_cut = o.rowacc_cut;
    t.rank_cut = o.rank_cut;
    return t;
}

// Plan creation under forced thresholds, timed on the host clock around a
// device synchronisation (creation synchronises on its own results; the
// plan is destroyed outside the timed region).
int timed_create(int m, int nA, int nB, const int *dAp, const int *dAi, const int *dBp,
                 const int *dBi, int mode, const sdl::SymTune &t, sdl_btda_plan **out,
                 double *ms)
{
    sdl::g_force_tune = &t;
    sdl::sync();
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = sdl_btda_plan_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, out);
    sdl::sync();
    const auto t1 = std::chrono::steady_clock::now();
    sdl::g_force_tune = nullptr;
    *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return rc;
}

// Average device time of reps refills of a plan (events on the library stream).
float time_refill(sdl_btda_plan *pl, const double *dAx, const double *dBx, double *dCx,
                  int reps)
{
    hipStream_t s = sdl::stream();
    hipEvent_t e0, e1;
    SDL_HIP(hipEventCreate(&e0));
    SDL_HIP(hipEventCreate(&e1));
    sdl_btda_refill(pl, dAx, dBx, nullptr, dCx, s);
    SDL_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) sdl_btda_refill(pl, dAx, dBx, nullptr, dCx, s);
    SDL_HIP(hipEventRecord(e1, s));
    SDL_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SDL_HIP(hipEventElapsedTime(&ms, e0, e1));
    SDL_HIP(hipEventDestroy(e0));
    SDL_HIP(hipEventDestroy(e1));
    return ms / (float) reps;
}

} // namespace

extern "C" {
#pragma GCC visibility push(default)

int sdl_btda_plan_create_tuned(int m, int nA, int nB, const int *dAp, const int *dAi,
                               const int *dBp, const int *dBi, int mode,
                               const sdl_sym_tune *tune, sdl_btda_plan **out)
{
    if (!tune) return sdl_btda_plan_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, out);
    const sdl::SymTune t = from_c(*tune);
    if (!sdl::tune_valid(t))
    {
        sdl::set_error("sdl_btda_plan_create_tuned: thresholds out of range");
        return 1;
    }
    sdl::g_force_tune = &t;
    const int rc = sdl_btda_plan_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, out);
    sdl::g_force_tune = nullptr;
    return rc;
}

int sdl_btda_plan_tune(const sdl_btda_plan *plan, sdl_sym_tune *out)
{
    if (!plan || !out) return 1;
    to_c(plan->P.tune, out);
    return 0;
}

int sdl_sym_tune_default(sdl_sym_tune *out)
{
    if (!out) return 1;
    to_c(sdl::SymTune(), out);
    return 0;
}

int sdl_btda_tune(int m, int nA, int nB, const int *dAp, const int *dAi, const int *dBp,
                  const int *dBi, int mode, const double *dAx, const double *dBx, int reps,
                  sdl_sym_tune *best, float *best_ms, float *default_ms)
{
    if (m < 0 || nA < 0 || nB < 0 || !dAp || !dBp || (!dAx) != (!dBx))
    {
        sdl::set_error("sdl_btda_tune: invalid arguments");
        return 1;
    }
    if (reps < 1) reps = 1;
    const sdl::SymTune def;
    // cost of a setting: plan creation plus (with values) one refill, the
    // least of reps (device allocation and release only ever add time, and
    // by widely varying amounts); the structure must not move (checked
    // before any refill)
    long long nnz0 = -1;
    bool rowacc = false;
    sdl::DBuf<double> C;
    auto cost = [&](const sdl::SymTune &t, double *med) -> int {
        std::vector<double> v;
        for (int r = 0; r < reps; r++)
        {
            sdl_btda_plan *pl = nullptr;
            double ms = 0.0;
            if (timed_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, t, &pl, &ms)) return 1;
            const long long nnz = sdl_btda_plan_nnz(pl);
            if (nnz0 < 0)
            {
                nnz0 = nnz;
                rowacc = !pl->P.pull;
                if (dAx) C.alloc((size_t) std::max<long long>(nnz, 1));
            }
            if (nnz != nnz0)
            {
                sdl_btda_plan_destroy(pl);
                sdl::set_error("sdl_btda_tune: thresholds changed the output structure");
                return 1;
            }
            if (dAx) ms += time_refill(pl, dAx, dBx, C.p, 1);
            sdl_btda_plan_destroy(pl);
            v.push_back(ms);
        }
        *med = *std::min_element(v.begin(), v.end());
        return 0;
    };
    {
        // untimed warm-up: kernels loaded, allocator primed
        sdl_btda_plan *pl = nullptr;
        double ms = 0.0;
        if (timed_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, def, &pl, &ms)) return 1;
        sdl_btda_plan_destroy(pl);
    }
    // a candidate replaces the incumbent only when at least 3 % cheaper, so
    // timing noise does not move the table away from the built-in values
    const double MARGIN = 0.97;
    double total_def = 0.0;
    if (cost(def, &total_def)) return 1;
    sdl::SymTune bt = def;
    double total_best = total_def;
    auto consider = [&](const sdl::SymTune &t) -> int {
        double c = 0.0;
        if (cost(t, &c)) return 1;
        if (c < MARGIN * total_best)
        {
            total_best = c;
            bt = t;
        }
        return 0;
    };
    // stage 1: the symbolic thresholds
    for (int hm : {1024, 2048, 4096, 8192})
        for (int dens : {2, 4, 8, 16, 32})
        {
            if (hm == def.hash_max && dens == def.dens) continue;
            sdl::SymTune t = def;
            t.hash_max = hm;
            t.dens = dens;
            if (consider(t)) return 1;
        }
    // stage 2: the refill cuts of a row-accumulator plan, from stage 1's winner
    if (dAx && rowacc)
    {
        const sdl::SymTune base = bt;
        for (int rc : {256, 512, 1024})
            for (int kc : {8192, 16384, 65536})
            {
                if (rc == base.rowacc_cut && kc == base.rank_cut) continue;
                sdl::SymTune t = base;
                t.rowacc_cut = rc;
                t.rank_cut = kc;
                if (consider(t)) return 1;
            }
    }
    // record under this product's signature (a plan with the defaults names it)
    {
        sdl_btda_plan *pl = nullptr;
        double ms = 0.0;
        if (timed_create(m, nA, nB, dAp, dAi, dBp, dBi, mode, def, &pl, &ms)) return 1;
        sdl::ShapeSig sig = pl->P.sig;
        sdl_btda_plan_destroy(pl);
        sdl::RefillCfg keep;
        sdl::tune_lookup(sig, nullptr, &keep);
        sdl::tune_record(sig, &bt, &keep);
    }
    if (best) to_c(bt, best);
    if (best_ms) *best_ms = (float) total_best;
    if (default_ms) *default_ms = (float) total_def;
    return 0;
}

int sdl_btda_tune_record_refill(const sdl_btda_plan *plan, const sdl_refill_cfg *cfg)
{
    if (!plan || !cfg)
    {
        sdl::set_error("sdl_btda_tune_record_refill: invalid arguments");
        return 1;
    }
    sdl::RefillCfg c(cfg->variant, cfg->rchunk);
    if (!sdl::refill_valid(c))
    {
        sdl::set_error("sdl_btda_tune_record_refill: not a refill candidate");
        return 1;
    }
    sdl::SymTune keep = plan->P.tune;
    sdl::tune_lookup(plan->P.sig, &keep, nullptr);
    sdl::tune_record(plan->P.sig, &keep, &c);
    return 0;
}

int sdl_tune_table_size(void)
{
    std::lock_guard<std::mutex> lk(sdl::g_mu);
    return (int) sdl::g_table.size();
}

void sdl_tune_table_clear(void)
{
    std::lock_guard<std::mutex> lk(sdl::g_mu);
    sdl::g_table.clear();
}

int sdl_tune_table_save(const char *path)
{
    if (!path)
    {
        sdl::set_error("sdl_tune_table_save: null path");
        return 1;
    }
    std::string body = sdl::HEADER;
    {
        std::lock_guard<std::mutex> lk(sdl::g_mu);
        for (const sdl::Entry &e : sdl::g_table)
        {
            char line[256];
            std::snprintf(line, sizeof line, "%s %d %d %d %d %d %d %d %d %d\n", e.arch.c_str(),
                          e.sig.lg_rows, e.sig.lg_cols, e.sig.lg_tpr, e.tune.hash_max,
                          e.tune.dens, e.tune.rowacc_cut, e.tune.rank_cut, e.refill.variant,
                          e.refill.rchunk);
            body += line;
        }
    }
    char tail[64];
    std::snprintf(tail, sizeof tail, "checksum %016llx\n",
                  (unsigned long long) sdl::fnv1a64(body.data(), body.size()));
    FILE *f = std::fopen(path, "wb");
    if (!f)
    {
        sdl::set_error("sdl_tune_table_save: cannot open %s", path);
        return 1;
    }
    const bool ok = std::fwrite(body.data(), 1, body.size(), f) == body.size() &&
                    std::fputs(tail, f) >= 0;
    if (std::fclose(f) != 0 || !ok)
    {
        sdl::set_error("sdl_tune_table_save: write to %s failed", path);
        return 1;
    }
    return 0;
}

// Loads a saved table, merging its entries over the in-memory ones (same
// key: the file's wins).  A file with a wrong header, a wrong checksum or an
// out-of-range entry is rejected whole and the table is left unchanged.
int sdl_tune_table_load(const char *path)
{
    if (!path)
    {
        sdl::set_error("sdl_tune_table_load: null path");
        return 1;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f)
    {
        sdl::set_error("sdl_tune_table_load: cannot open %s", path);
        return 1;
    }
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0)
    {
        text.append(buf, got);
        if (text.size() > (1u << 20)) break;
    }
    std::fclose(f);
    auto bad = [&](const char *why) {
        sdl::set_error("sdl_tune_table_load: %s is not a valid tuning table (%s)", path, why);
        return 1;
    };
    const size_t hl = std::strlen(sdl::HEADER);
    if (text.size() > (1u << 20)) return bad("too large");
    if (text.compare(0, hl, sdl::HEADER) != 0) return bad("header");
    const size_t cpos = text.rfind("checksum ");
    if (cpos == std::string::npos || cpos < hl || text[cpos - 1] != '\n')
        return bad("no checksum");
    unsigned long long sum = 0;
    char end = 0;
    if (std::sscanf(text.c_str() + cpos, "checksum %16llx%c", &sum, &end) != 2 || end != '\n' ||
        cpos + 26 != text.size())
        return bad("checksum line");
    if (sum != sdl::fnv1a64(text.data(), cpos)) return bad("checksum mismatch");
    std::vector<sdl::Entry> add;
    size_t pos = hl;
    while (pos < cpos)
    {
        const size_t nl = text.find('\n', pos);
        const std::string line = text.substr(pos, nl - pos);
        pos = nl + 1;
        char arch[64] = {0};
        sdl::Entry e;
        int n = 0;
        if (std::sscanf(line.c_str(), "%63s %d %d %d %d %d %d %d %d %d%n", arch,
                        &e.sig.lg_rows, &e.sig.lg_cols, &e.sig.lg_tpr, &e.tune.hash_max,
                        &e.tune.dens, &e.tune.rowacc_cut, &e.tune.rank_cut, &e.refill.variant,
                        &e.refill.rchunk, &n) != 10 ||
            n != (int) line.size())
            return bad("entry format");
        if (!sdl::tune_valid(e.tune) || !sdl::refill_valid(e.refill) || e.sig.lg_rows < 0 ||
            e.sig.lg_rows > 40 || e.sig.lg_cols < 0 || e.sig.lg_cols > 40 || e.sig.lg_tpr < 0 ||
            e.sig.lg_tpr > 62)
            return bad("entry out of range");
        e.arch = arch;
        add.push_back(e);
    }
    std::lock_guard<std::mutex> lk(sdl::g_mu);
    for (const sdl::Entry &e : add)
    {
        sdl::Entry *x = sdl::find(e.arch, e.sig);
        if (x) *x = e;
        else sdl::g_table.push_back(e);
    }
    return 0;
}

#pragma GCC visibility pop
} // extern "C"
