// Batched randomised range finder (dh_lra_batched) and the autotuner table.
//
// A batch of independent problems (n_i, d_i, l_i, shared q) runs stage by
// stage, each stage one launch over a device table of all problems:
//   prepare  : column means, 16-bit centred points (FP16: scaled by a power of
//              two found on the device), squared norms
//   per pass : pack the operand (Omega or Q) -> persistent 16-bit projection
//              kernels, one per tile class (dpad, column blocks), pulling
//              (problem, row block) items from an atomic queue, longest
//              problem first -> CholeskyQR2 (register-blocked Gram over the
//              row slabs of every problem, one workgroup Cholesky per
//              problem, row-parallel solve), twice
//   copies   : W -> Q after each power iteration
// and the sequence is captured once in a hipGraph and replayed while the
// problem table is unchanged.  Queue counters are cleared by a fill kernel
// inside the graph (no memset nodes).
//
// Breakdown: a problem whose Gram matrix is numerically singular keeps its
// flag in the device info words; after the graph the host redoes such
// problems alone with dh_lra, which falls back to shifted CholeskyQR3.
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "batch.h"

namespace {

constexpr float kLog2e = 1.4426950408889634f;
#define DH_BF16_DEFAULT_WIDE 0      // schedule of a class when tuning is off

// ------------------------------------------------------------------ kernels

struct CopyJob {
    int m, n, lds, ldd;
    const double* src;
    double* dst;
};

__global__ __launch_bounds__(256) void copy_batched_kernel(const CopyJob* __restrict__ jobs) {
    const CopyJob& j = jobs[blockIdx.y];
    const long total = (long)j.m * j.n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        int i = (int)(e % j.m), c = (int)(e / j.m);
        j.dst[i + (size_t)c * j.ldd] = j.src[i + (size_t)c * j.lds];
    }
}

__global__ __launch_bounds__(256) void fill_int_kernel(int* __restrict__ p, int count, int value) {
    for (int e = blockIdx.x * 256 + threadIdx.x; e < count; e += gridDim.x * 256) p[e] = value;
}

// --------------------------------------------------------------- the plan

struct ClassRange {
    DhBf16Class cls;
    int first, count;  // items[first, first + count)
};

struct Plan {
    int prec = -1, q = 0;
    std::vector<dh_lra_problem> probs;
    void* mem = nullptr;
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int* infos = nullptr;   // 2 per problem: status of the latest Q / W CholeskyQR step
    int* sticky = nullptr;  // 1 per problem: some CholeskyQR step broke down
    // Side streams (default priority) and events of the captured fork / join
    // that runs the tile classes of a pass side by side.
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> events;

    ~Plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (mem) (void)hipFree(mem);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
    }
};

struct TuneTable {
    std::map<uint64_t, int> choice;
    std::string arch;  // device tag of the file cache ("" until first needed)
};

// Process-wide file cache of autotuner decisions, shared by all handles:
// one line "<device tag> <key, hex> <choice>" per decision, later lines
// overriding earlier ones.  The device tag is the architecture name and the
// CU count, so a file shared between machines keeps their choices apart.
// Path: DH_TUNE_CACHE if set (empty disables the file), else
// $XDG_CACHE_HOME/darray_hip/tune.txt, else $HOME/.cache/darray_hip/tune.txt.
// A file that cannot be read or written leaves the in-memory tables working.
struct TuneFile {
    std::mutex mu;
    bool loaded = false;
    std::string path;
    std::map<std::pair<std::string, uint64_t>, int> entries;
};

TuneFile& tune_file() {
    static TuneFile f;
    return f;
}

std::string tune_cache_path() {
    if (const char* e = std::getenv("DH_TUNE_CACHE")) return std::string(e);
    std::string base;
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) {
        base = x;
    } else if (const char* hm = std::getenv("HOME"); hm && *hm) {
        base = std::string(hm) + "/.cache";
    } else {
        return "";
    }
    return base + "/darray_hip/tune.txt";
}

void tune_file_load(TuneFile& f) {  // f.mu held
    if (f.loaded) return;
    f.loaded = true;
    f.path = tune_cache_path();
    if (f.path.empty()) return;
    FILE* fp = std::fopen(f.path.c_str(), "r");
    if (!fp) return;
    char line[512], arch[160];
    unsigned long long key = 0;
    int choice = 0;
    while (std::fgets(line, sizeof line, fp))
        if (line[0] != '#' && std::sscanf(line, "%159s %llx %d", arch, &key, &choice) == 3)
            f.entries[{std::string(arch), (uint64_t)key}] = choice;
    std::fclose(fp);
}

void tune_file_append(TuneFile& f, const std::string& arch, uint64_t key, int choice) {
    if (f.path.empty()) return;
    // Parent directories, one level at a time (errors surface at fopen).
    for (size_t pos = f.path.find('/', 1); pos != std::string::npos;
         pos = f.path.find('/', pos + 1))
        (void)mkdir(f.path.substr(0, pos).c_str(), 0755);
    FILE* fp = std::fopen(f.path.c_str(), "a");
    if (!fp) return;
    std::fprintf(fp, "%s %llx %d\n", arch.c_str(), (unsigned long long)key, choice);
    std::fclose(fp);
}

const std::string& tune_arch(dh_handle_t h) {
    auto* t = static_cast<TuneTable*>(h->tune);
    if (t->arch.empty()) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, h->device) == hipSuccess) {
            t->arch = std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount);
            for (char& c : t->arch)
                if (c == ' ') c = '_';
        } else {
            t->arch = "unknown";
        }
    }
    return t->arch;
}

// Bump allocator over one device block (256-byte aligned pieces).
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    }
};

// Field by field: callers that fill the struct member by member leave its
// padding bytes undefined, so a byte comparison would miss equal tables.
bool same_problem(const dh_lra_problem& a, const dh_lra_problem& b) {
    return a.n == b.n && a.d == b.d && a.l == b.l && a.q == b.q && a.X == b.X && a.ldx == b.ldx &&
           a.y == b.y && a.gamma == b.gamma && a.Omega == b.Omega && a.ldo == b.ldo &&
           a.Q == b.Q && a.ldq == b.ldq && a.W == b.W && a.ldw == b.ldw;
}

bool same_problems(const Plan& p, int prec, int count, const dh_lra_problem* probs) {
    if (!p.exec || p.prec != prec || (int)p.probs.size() != count) return false;
    for (int i = 0; i < count; ++i)
        if (!same_problem(p.probs[i], probs[i])) return false;
    return true;
}

}  // namespace

uint64_t dh_tune_key_bf16(int dk, int ncb, int waves, int rows) {
    int lg = 0;
    while ((1 << lg) < rows && lg < 30) ++lg;
    return ((uint64_t)1 << 56) | ((uint64_t)dk << 40) | ((uint64_t)ncb << 32) |
           ((uint64_t)waves << 24) | (uint64_t)lg;
}

// ------------------------------------------------------------ tune table

bool dh_tune_enabled() {
    static int on = -1;
    if (on < 0) {
        const char* e = std::getenv("DH_TUNE");
        on = (e && e[0] == '0') ? 0 : 1;
    }
    return on != 0;
}

int dh_forced_schedule() {
    static int v = -2;
    if (v == -2) {
        const char* e = std::getenv("DH_SCHEDULE");
        v = (e && e[0] >= '0' && e[0] <= '9') ? std::atoi(e) : -1;
    }
    return v;
}

int dh_tune_lookup(dh_handle_t h, uint64_t key) {
    if (!h->tune) h->tune = new TuneTable();
    auto& m = static_cast<TuneTable*>(h->tune)->choice;
    auto it = m.find(key);
    if (it != m.end()) return it->second;
    const std::string& arch = tune_arch(h);
    TuneFile& f = tune_file();
    std::lock_guard<std::mutex> lock(f.mu);
    tune_file_load(f);
    auto fi = f.entries.find({arch, key});
    if (fi == f.entries.end()) return -1;
    m[key] = fi->second;
    return fi->second;
}

void dh_tune_store(dh_handle_t h, uint64_t key, int choice) {
    if (!h->tune) h->tune = new TuneTable();
    static_cast<TuneTable*>(h->tune)->choice[key] = choice;
    const std::string& arch = tune_arch(h);
    TuneFile& f = tune_file();
    std::lock_guard<std::mutex> lock(f.mu);
    tune_file_load(f);
    auto it = f.entries.find({arch, key});
    if (it != f.entries.end() && it->second == choice) return;  // already on file
    f.entries[{arch, key}] = choice;
    tune_file_append(f, arch, key, choice);
}

void dh_tune_release(dh_handle_t h) {
    delete static_cast<TuneTable*>(h->tune);
    h->tune = nullptr;
}

void dh_batch_release(dh_handle_t h) {
    delete static_cast<Plan*>(h->batch);
    h->batch = nullptr;
}

namespace {

// Decide the schedule of each class (8 waves x 32 rows or 4 waves x 64 rows)
// by timing both on the first projection pass (identical results: only the
// wave layout differs).  Runs eagerly, before the capture.
int tune_classes(dh_handle_t h, std::vector<ClassRange>& classes, const DhProjJob* jobs,
                 const DhProjItem* items, int* counters) {
    hipEvent_t e0, e1;
    DH_CHECK_HIP(hipEventCreate(&e0));
    DH_CHECK_HIP(hipEventCreate(&e1));
    int status = 0;
    for (size_t c = 0; c < classes.size() && !status; ++c) {
        ClassRange& cr = classes[c];
        if (cr.cls.dk >= 8) {
            // One schedule only; run it once eagerly so that its one-time
            // launch set-up happens outside the graph capture.
            cr.cls.wide = 0;
            fill_int_kernel<<<1, 256, 0, h->stream>>>(counters, 1, 0);
            status = dh_bf16_project_persistent(h, cr.cls, jobs, items + cr.first, cr.count,
                                                counters);
            continue;
        }
        uint64_t key = dh_tune_key_bf16(cr.cls.dk, cr.cls.ncb, 1, cr.count * 256) ^ ((uint64_t)1 << 60) ^
                       ((uint64_t)cr.cls.f16 << 61);
        int known = dh_tune_lookup(h, key);
        if (known < 0 && !dh_tune_enabled()) known = DH_BF16_DEFAULT_WIDE;
        // Candidates run once eagerly (their launch set-up then happens outside
        // the capture); with no recorded choice the faster one is recorded.
        const int first = known >= 0 ? known : 0, lastc = known >= 0 ? known : 1;
        float best = 3.4e38f;
        int best_wide = first;
        if (first != lastc)  // untimed first launches (code-object loading)
            for (int wide = first; wide <= lastc && !status; ++wide) {
                DhBf16Class k = cr.cls;
                k.wide = wide;
                fill_int_kernel<<<1, 256, 0, h->stream>>>(counters, 1, 0);
                status = dh_bf16_project_persistent(h, k, jobs, items + cr.first, cr.count, counters);
            }
        for (int wide = first; wide <= lastc && !status; ++wide) {
            DhBf16Class k = cr.cls;
            k.wide = wide;
            fill_int_kernel<<<1, 256, 0, h->stream>>>(counters, 1, 0);
            if ((status = hipGetLastError() != hipSuccess ? DH_ERR_HIP : 0)) break;
            (void)hipEventRecord(e0, h->stream);
            status = dh_bf16_project_persistent(h, k, jobs, items + cr.first, cr.count, counters);
            (void)hipEventRecord(e1, h->stream);
            if (!status && first != lastc && hipEventSynchronize(e1) == hipSuccess) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms < best) {
                    best = ms;
                    best_wide = wide;
                }
            }
        }
        cr.cls.wide = best_wide;
        if (first != lastc) dh_tune_store(h, key, best_wide);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return status;
}

int build_plan(dh_handle_t h, Plan& P, int prec, int count, const dh_lra_problem* probs) {
    P.prec = prec;
    P.q = probs[0].q;
    P.probs.assign(probs, probs + count);
    const int passes = P.q + 2;

    // ---- sizes and layout
    struct Inst {
        int dpad, npad, nchunks, ncb, groups, ncb_pack, rows, nslab;
        size_t o_part, o_mean, o_xb, o_a, o_op, o_g, o_gpart, o_dinv;
    };
    std::vector<Inst> in(count);
    Carve cv;
    const size_t o_sets = cv.take(sizeof(DhBf16Set) * count);
    const size_t o_pack = cv.take(sizeof(DhPackJob) * count * 2);
    const size_t o_proj = cv.take(sizeof(DhProjJob) * count * 2);
    const size_t o_gram = cv.take(sizeof(DhGramRegJob) * count * 2);
    const size_t o_spd = cv.take(sizeof(DhSpdJob) * count * 2);
    const size_t o_copy = cv.take(sizeof(CopyJob) * count);
    const size_t o_info = cv.take(sizeof(int) * count * 4);  // statuses, sticky, FP16 maxabs
    int max_n = 0, max_d = 0, max_npad = 0, max_l = 0;
    long max_pack_threads = 0, max_nl = 0;
    size_t nitems = 0;
    // Gram slabs: about four register-blocked workgroups per CU over the batch.
    long total_rows = 0;
    for (int i = 0; i < count; ++i) total_rows += probs[i].n;
    const int gram_rows =
        std::max(256, 16 * (int)dh_cdiv(dh_cdiv(total_rows, 4L * h->num_cu), 16L));
    int ngitems = 0;
    for (int i = 0; i < count; ++i) {
        const dh_lra_problem& p = probs[i];
        Inst& s = in[i];
        s.dpad = dh_bf16_dpad(p.d);
        s.npad = dh_cdiv(p.n, 32) * 32;
        s.nchunks = dh_cdiv(p.n, 32);
        dh_bf16_columns(p.l, &s.ncb, &s.groups, &s.ncb_pack);
        s.rows = dh_cdiv(p.n, dh_bf16_rows(s.dpad));
        s.nslab = dh_cdiv(p.n, gram_rows);
        ngitems += s.nslab;
        s.o_part = cv.take(sizeof(double) * (size_t)dh_cdiv(p.n, 256) * p.d);
        s.o_mean = cv.take(sizeof(double) * p.d);
        s.o_xb = cv.take((size_t)s.npad * s.dpad * 2);
        s.o_a = cv.take(sizeof(float) * s.npad);
        s.o_op = cv.take(dh_bf16_packed_bytes(p.n, s.ncb_pack));
        s.o_g = cv.take(sizeof(double) * (size_t)p.l * p.l);
        s.o_gpart = cv.take(sizeof(double) * (size_t)s.nslab * dh_gram_reg_slab_doubles());
        s.o_dinv = cv.take(sizeof(double) * 256 * (size_t)dh_cdiv(p.l, 16));
        nitems += (size_t)s.rows * s.groups;
        max_n = std::max(max_n, p.n);
        max_d = std::max(max_d, p.d);
        max_npad = std::max(max_npad, s.npad);
        max_l = std::max(max_l, p.l);
        max_pack_threads = std::max(max_pack_threads, (long)s.nchunks * s.ncb_pack * 16);
        max_nl = std::max(max_nl, (long)p.n * p.l);
    }
    const size_t o_items = cv.take(sizeof(DhProjItem) * nitems);
    const size_t o_gitems = cv.take(sizeof(DhGramItem) * ngitems);

    // ---- tile classes and the longest-first work list
    std::vector<DhProjItem> items;
    items.reserve(nitems);
    std::vector<ClassRange> classes;
    {
        std::map<std::pair<int, int>, std::vector<int>> by_class;  // (dk, ncb) -> problems
        for (int i = 0; i < count; ++i) by_class[{in[i].dpad / 32, in[i].ncb}].push_back(i);
        for (auto& kv : by_class) {
            std::vector<int>& ids = kv.second;
            std::stable_sort(ids.begin(), ids.end(),
                             [&](int a, int b) { return probs[a].n > probs[b].n; });
            ClassRange cr;
            cr.cls = {kv.first.first, kv.first.second, dh_bf16_rows(kv.first.first * 32), 0,
                      prec == DH_PREC_F16 ? 1 : 0};
            cr.first = (int)items.size();
            for (int i : ids)
                for (int r = 0; r < in[i].rows; ++r)
                    for (int g = 0; g < in[i].groups; ++g) items.push_back({i, r, g, 0});
            cr.count = (int)items.size() - cr.first;
            classes.push_back(cr);
        }
    }
    const size_t o_counters = cv.take(sizeof(int) * (size_t)passes * classes.size() + 64);

    P.bytes = cv.off;
    DH_CHECK_HIP(hipMalloc(&P.mem, P.bytes));
    char* base = (char*)P.mem;
    auto at = [&](size_t o) { return (void*)(base + o); };

    // ---- tables (host images, uploaded before capture)
    std::vector<DhBf16Set> sets(count);
    std::vector<DhPackJob> pack(2 * count);
    std::vector<DhProjJob> proj(2 * count);
    std::vector<DhGramRegJob> gram(2 * count);
    std::vector<DhSpdJob> spd(2 * count);
    std::vector<CopyJob> copy(count);
    P.infos = (int*)at(o_info);
    P.sticky = P.infos + 2 * count;
    for (int i = 0; i < count; ++i) {
        const dh_lra_problem& p = probs[i];
        const Inst& s = in[i];
        const double neg = -p.gamma * (double)kLog2e;
        sets[i] = {p.n, p.d, s.dpad, s.npad, p.X, p.ldx, neg, (double*)at(s.o_part),
                   (double*)at(s.o_mean), at(s.o_xb), (float*)at(s.o_a), p.gamma, 1.0,
                   (unsigned*)(P.infos + 3 * count + i)};
        for (int t = 0; t < 2; ++t) {
            pack[t * count + i] = {p.n, p.l, s.nchunks, s.ncb_pack, t == 0 ? p.Omega : p.Q,
                                   t == 0 ? p.ldo : p.ldq, p.y, at(s.o_op)};
            const float c2 = (float)(2.0 * p.gamma) * kLog2e;
            proj[t * count + i] = {p.n, s.nchunks, p.l, s.ncb_pack, at(s.o_xb), (float*)at(s.o_a),
                                   c2, at(s.o_op), p.y,
                                   t == 0 ? p.Q : p.W, t == 0 ? p.ldq : p.ldw};
            double* Y = t == 0 ? p.Q : p.W;
            int ldy = t == 0 ? p.ldq : p.ldw;
            gram[t * count + i] = {p.n, p.l, ldy, gram_rows, s.nslab, Y,
                                   (double*)at(s.o_gpart), (double*)at(s.o_g), p.l};
            spd[t * count + i] = {p.n, p.l, ldy, p.l, Y, (double*)at(s.o_g), (double*)at(s.o_dinv),
                                  P.infos + t * count + i, P.sticky + i};
        }
        copy[i] = {p.n, p.l, p.ldw, p.ldq, p.W, p.Q};
    }
    auto up = [&](size_t o, const void* src, size_t bytes) {
        return hipMemcpyAsync(at(o), src, bytes, hipMemcpyHostToDevice, h->stream);
    };
    DH_CHECK_HIP(up(o_sets, sets.data(), sizeof(DhBf16Set) * count));
    DH_CHECK_HIP(up(o_pack, pack.data(), sizeof(DhPackJob) * 2 * count));
    DH_CHECK_HIP(up(o_proj, proj.data(), sizeof(DhProjJob) * 2 * count));
    DH_CHECK_HIP(up(o_gram, gram.data(), sizeof(DhGramRegJob) * 2 * count));
    // Gram items, largest orders first (their slabs cost the most MFMAs).
    std::vector<int> gorder(count);
    for (int i = 0; i < count; ++i) gorder[i] = i;
    std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) {
        return probs[a].l > probs[b].l;
    });
    std::vector<DhGramItem> gitems;
    gitems.reserve(ngitems);
    for (int i : gorder)
        for (int z = 0; z < in[i].nslab; ++z) gitems.push_back({i, z});
    DH_CHECK_HIP(up(o_gitems, gitems.data(), sizeof(DhGramItem) * gitems.size()));
    DH_CHECK_HIP(up(o_spd, spd.data(), sizeof(DhSpdJob) * 2 * count));
    DH_CHECK_HIP(up(o_copy, copy.data(), sizeof(CopyJob) * count));
    DH_CHECK_HIP(up(o_items, items.data(), sizeof(DhProjItem) * items.size()));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));

    const DhBf16Set* d_sets = (const DhBf16Set*)at(o_sets);
    const DhPackJob* d_pack = (const DhPackJob*)at(o_pack);
    DhProjJob* d_proj = (DhProjJob*)at(o_proj);
    const bool f16 = prec == DH_PREC_F16;
    const DhGramRegJob* d_gram = (const DhGramRegJob*)at(o_gram);
    const DhGramItem* d_gitems = (const DhGramItem*)at(o_gitems);
    const DhSpdJob* d_spd = (const DhSpdJob*)at(o_spd);
    const CopyJob* d_copy = (const CopyJob*)at(o_copy);
    const DhProjItem* d_items = (const DhProjItem*)at(o_items);
    int* d_counters = (int*)at(o_counters);
    const int ncounters = passes * (int)classes.size();

    // ---- eager preparation + autotuning of the projection schedules
    DH_CHECK_STATUS(dh_bf16_prepare_batched(h, d_sets, count, max_n, max_d, max_npad, f16, d_proj));
    DH_CHECK_STATUS(dh_bf16_pack_batched(h, d_pack, count, max_pack_threads, f16));
    DH_CHECK_STATUS(tune_classes(h, classes, d_proj, d_items, d_counters));
    // Warm the batched CholeskyQR launcher (it sets a kernel attribute once).
    DH_CHECK_STATUS(dh_cholqr_step_batched(h, d_spd, 0, max_l, max_n, false));
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));

    // ---- capture
    auto orth = [&](int t) -> int {
        const DhGramRegJob* gj = d_gram + t * count;
        const DhSpdJob* sj = d_spd + t * count;
        for (int sweep = 0; sweep < 2; ++sweep) {
            DH_CHECK_STATUS(dh_gram_reg_batched(h, gj, count, d_gitems, ngitems));
            DH_CHECK_STATUS(dh_cholqr_step_batched(h, sj, count, max_l, max_n, false));
        }
        return 0;
    };
    // The classes of a pass are independent (different problems): up to four
    // run side by side, biggest first, on the handle's stream and up to three
    // side streams forked from it and joined back inside the capture.
    std::vector<int> order(classes.size());
    for (size_t c = 0; c < classes.size(); ++c) order[c] = (int)c;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long)classes[a].count * classes[a].cls.dk * classes[a].cls.ncb >
               (long)classes[b].count * classes[b].cls.dk * classes[b].cls.ncb;
    });
    const int lanes = std::min<int>(4, (int)classes.size());
    for (int i = 1; i < lanes; ++i) {
        hipStream_t s = nullptr;
        DH_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        P.side.push_back(s);
    }
    for (int i = 0; i < lanes; ++i) {
        hipEvent_t e = nullptr;
        DH_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        P.events.push_back(e);
    }
    // Projections of one pass: fork, classes round robin over the lanes, join.
    auto project_pass = [&](int pass, int t) -> int {
        hipStream_t main = h->stream;
        if (lanes > 1) {
            if (hipEventRecord(P.events[0], main) != hipSuccess) return DH_ERR_HIP;
            for (hipStream_t s : P.side)
                if (hipStreamWaitEvent(s, P.events[0], 0) != hipSuccess) return DH_ERR_HIP;
        }
        int st = 0;
        for (size_t k = 0; k < order.size() && !st; ++k) {
            const int c = order[k];
            h->stream = (k % lanes == 0) ? main : P.side[k % lanes - 1];
            st = dh_bf16_project_persistent(h, classes[c].cls, d_proj + t * count,
                                            d_items + classes[c].first, classes[c].count,
                                            d_counters + pass * (int)classes.size() + c);
        }
        h->stream = main;
        if (lanes > 1) {
            for (int i = 1; i < lanes; ++i) {
                if (hipEventRecord(P.events[i], P.side[i - 1]) != hipSuccess) return DH_ERR_HIP;
                if (hipStreamWaitEvent(main, P.events[i], 0) != hipSuccess) return DH_ERR_HIP;
            }
        }
        return st;
    };
    DH_CHECK_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int st = 0;
    do {
        fill_int_kernel<<<dh_cdiv(ncounters, 256), 256, 0, h->stream>>>(d_counters, ncounters, 0);
        fill_int_kernel<<<dh_cdiv(3 * count, 256), 256, 0, h->stream>>>(P.infos, 3 * count, 0);
        if ((st = dh_bf16_prepare_batched(h, d_sets, count, max_n, max_d, max_npad, f16, d_proj)))
            break;
        for (int pass = 0; pass < passes && !st; ++pass) {
            const int t = pass == 0 ? 0 : 1;  // operand Omega / Q, output Q / W
            if ((st = dh_bf16_pack_batched(h, d_pack + t * count, count, max_pack_threads, f16)))
                break;
            st = project_pass(pass, t);
            if (st || pass == passes - 1) break;
            if ((st = orth(t))) break;
            if (pass >= 1) {
                copy_batched_kernel<<<dim3(std::min<long>(dh_cdiv(max_nl, 256), 4096), count), 256,
                                      0, h->stream>>>(d_copy);
                if (hipGetLastError() != hipSuccess) st = DH_ERR_HIP;
            }
        }
    } while (0);
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(h->stream, &graph);
    if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (ec != hipSuccess) return dh_fail(DH_ERR_HIP, "dh_lra_batched: capture failed: %s",
                                         hipGetErrorString(ec));
    P.graph = graph;
    DH_CHECK_HIP(hipGraphInstantiate(&P.exec, graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" int dh_lra_batched(dh_handle_t h, int prec, int count, const dh_lra_problem* problems,
                              int* status) {
    const char* name = "dh_lra_batched";
    if (!h) return dh_fail(-1, "%s: NULL handle", name);
    if (prec < DH_PREC_F64 || prec > DH_PREC_FP8) return dh_fail(-2, "%s: bad prec %d", name, prec);
    if (count < 0) return dh_fail(-3, "%s: count < 0", name);
    if (count > 0 && !problems) return dh_fail(-4, "%s: problems is NULL", name);
    if (count == 0) return 0;
    for (int i = 0; i < count; ++i) {
        const dh_lra_problem& p = problems[i];
        if (p.q != problems[0].q || p.q < 0)
            return dh_fail(-4, "%s: problem %d: q must be >= 0 and shared (%d vs %d)", name, i, p.q,
                           problems[0].q);
        if (p.n < 1 || p.d < 1 || p.l < 1 || p.l > p.n || p.ldx < p.d || p.ldo < p.n ||
            p.ldq < p.n || p.ldw < p.n || !p.X || !p.Omega || !p.Q || !p.W)
            return dh_fail(-4, "%s: problem %d: bad sizes or pointers (n=%d d=%d l=%d)", name, i,
                           p.n, p.d, p.l);
        if ((prec == DH_PREC_BF16 || prec == DH_PREC_F16) && (p.d > 256 || p.l > 272))
            return dh_fail(-4, "%s: problem %d: 16-bit batches need d <= 256 and l <= 272", name,
                           i);
    }
    // The batch runs as its staged graph.  A single persistent launch of every
    // stage was tried and hung on the device; it is not available, and asking
    // for it is an error rather than a silent fallback.
    if (const char* e = std::getenv("DH_BATCH_SCHEDULE"); e && std::strcmp(e, "single") == 0)
        return dh_fail(DH_ERR_UNSUPPORTED, "%s: DH_BATCH_SCHEDULE=single is not available",
                       name);
    std::vector<int> st(count, 0);
    int first = 0;
    if (prec != DH_PREC_BF16 && prec != DH_PREC_F16) {
        for (int i = 0; i < count; ++i) {
            const dh_lra_problem& p = problems[i];
            st[i] = dh_lra(h, prec, p.n, p.d, p.l, p.q, p.X, p.ldx, p.y, p.gamma, p.Omega,
                                     p.ldo, p.Q, p.ldq, p.W, p.ldw);
            if (st[i] < 0) return st[i];
        }
    } else {
        DH_CHECK_HIP(hipSetDevice(h->device));
        Plan* P = static_cast<Plan*>(h->batch);
        if (!P || !same_problems(*P, prec, count, problems)) {
            dh_batch_release(h);
            P = new Plan();
            h->batch = P;
            int s = build_plan(h, *P, prec, count, problems);
            if (s) {
                dh_batch_release(h);
                return s;
            }
        }
        DH_CHECK_HIP(hipGraphLaunch(P->exec, h->stream));
        std::vector<int> broke(count);
        DH_CHECK_HIP(hipMemcpyAsync(broke.data(), P->sticky, sizeof(int) * count,
                                    hipMemcpyDeviceToHost, h->stream));
        DH_CHECK_HIP(hipStreamSynchronize(h->stream));
        for (int i = 0; i < count; ++i) {
            if (broke[i] == 0) continue;
            const dh_lra_problem& p = problems[i];
            st[i] = dh_lra(h, prec, p.n, p.d, p.l, p.q, p.X, p.ldx, p.y, p.gamma, p.Omega,
                                     p.ldo, p.Q, p.ldq, p.W, p.ldw);
            if (st[i] < 0) return st[i];
        }
    }
    for (int i = 0; i < count; ++i) {
        if (status) status[i] = st[i];
        if (!first && st[i]) first = st[i];
    }
    DH_CHECK_HIP(hipStreamSynchronize(h->stream));
    return first;
}
