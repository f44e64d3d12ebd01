/*
 * pt_engine.hip -- multi-job executor behind the batched form (pt_batch_*)
 * and the serving form (pt_server_*) of include/pt_api.h.
 *
 * Every job or frame in flight is a lane: the recorded stream work of one
 * synthesis call on its own context (pt_plan.h).  An engine tick advances
 * every lane by its leading closures (copies, uploads, pyramid and advection
 * kernels) and then by exactly one operator with a multi-job form.  The
 * operators of all lanes that share a launch key (operator, kernel shape,
 * patch radius, temporal terms, level-size class) become ONE launch whose
 * grid is the concatenation of every member's tiles; a device table gives
 * each workgroup its job's arguments.  Lanes never wait for each other, and
 * each lane's work stays in its recorded order on one stream, so every job's
 * results are bit-identical to running it alone.
 *
 * The batched form plans all ticks up front, uploads every table once and
 * captures the whole launch sequence in a HIP graph.  The serving form
 * builds ticks on the fly from the frames in flight, staging each tick's
 * tables through a pinned ring into device memory.
 */
#include "pt_plan.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local char g_engine_error[512] = "";

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* level-size classes of the shape policy: < 128^2, < 512^2, >= 512^2 pixels */
int size_class(const PtArgs &a)
{
    int64_t px = (int64_t)a.lw * a.lh;
    return px < 128 * 128 ? 0 : (px < 512 * 512 ? 1 : 2);
}

struct ShapePolicy {
    int shape[2][3]; /* [propagate, scatter][class]; -1 = default policy */
    /* a job step with at least this many tiles fills the GPU alone: it gets
     * a launch of its own (sharing one launch would only spread the gathers
     * over several exemplars at once) */
    int isolate_tiles = 2048;
    /* a launch group of one job uses the single-job kernel (arguments by
     * value, no launch table) when its shape is the default one */
    bool single_direct = true;
    ShapePolicy()
    {
        for (auto &r : shape)
            for (auto &v : r) v = -1;
    }
    int pick(int op, const PtArgs &a) const
    {
        if (op == PT_OP_PROPAGATE || op == PT_OP_SCATTER) {
            int s = shape[op][size_class(a)];
            if (s >= 0 && pt_shape_valid(op, s, a)) return s;
        }
        return pt_default_shape(op, a);
    }
};

struct Lane {
    std::vector<RecOp> ops;
    size_t pos = 0;
    bool done() const { return pos >= ops.size(); }
};

/* One entry of a launch program. */
struct Step {
    int kind = 0; /* 0: closure ops[opi] of lane; 1: multi-job launch; 2: one job's op, default shape */
    int lane = -1;
    size_t opi = 0;
    int op = -1, shape = 0, cls = 0, n = 0, ntiles = 0;
    size_t table_off = 0;
    PtArgs first{};
};

constexpr size_t kTableAlign = 256;

size_t table_bytes(int n)
{
    size_t b = (size_t)n * sizeof(PtArgs) + (size_t)(2 * n + 1) * sizeof(int);
    return (b + kTableAlign - 1) / kTableAlign * kTableAlign;
}

/* Appends one launch table ([PtArgs n][prefix n+1][tiles_x n]) at *off. */
void write_table(std::vector<char> &buf, size_t off, const std::vector<const PtArgs *> &args,
                 const std::vector<int> &tiles, const std::vector<int> &tiles_x)
{
    int n = (int)args.size();
    if (buf.size() < off + table_bytes(n)) buf.resize(off + table_bytes(n));
    char *p = buf.data() + off;
    for (int i = 0; i < n; i++) memcpy(p + i * sizeof(PtArgs), args[i], sizeof(PtArgs));
    int *prefix = (int *)(p + n * sizeof(PtArgs));
    prefix[0] = 0;
    for (int i = 0; i < n; i++) prefix[i + 1] = prefix[i] + tiles[i];
    int *tx = prefix + n + 1;
    for (int i = 0; i < n; i++) tx[i] = tiles_x[i];
}

/*
 * One tick over `lanes` (nullptr entries are skipped): appends the leading
 * closures of every lane, then one multi-job launch per launch key, to
 * `out`; launch tables go to `tables` starting at byte `*toff`.
 */
void build_tick(std::vector<Lane *> &lanes, const ShapePolicy &pol, std::vector<Step> &out, std::vector<char> &tables,
                size_t *toff)
{
    for (size_t li = 0; li < lanes.size(); li++) {
        Lane *L = lanes[li];
        if (!L) continue;
        while (!L->done() && L->ops[L->pos].op < 0) {
            Step s;
            s.kind = 0;
            s.lane = (int)li;
            s.opi = L->pos++;
            out.push_back(s);
        }
    }
    struct Group {
        int op, shape, cls;
        std::vector<int> lanes;
    };
    /* ordered: deterministic launch order; second = lane of an isolated step */
    std::map<std::pair<int, int>, Group> groups;
    for (size_t li = 0; li < lanes.size(); li++) {
        Lane *L = lanes[li];
        if (!L || L->done()) continue;
        const RecOp &o = L->ops[L->pos];
        int shape = pol.pick(o.op, o.a);
        int cls = size_class(o.a);
        int key = pt_shape_key(o.op, shape, o.a) * 4 + cls;
        bool alone = pt_shape_tiles(o.op, shape, o.a, nullptr) >= pol.isolate_tiles;
        Group &g = groups[std::make_pair(key, alone ? (int)li : -1)];
        g.op = o.op;
        g.shape = shape;
        g.cls = cls;
        g.lanes.push_back((int)li);
    }
    for (auto &kv : groups) {
        Group &g = kv.second;
        if (g.lanes.size() == 1 && pol.single_direct) {
            /* one job: the plain single-job kernel with its arguments by
             * value when the policy's shape is the default one (no table) */
            Lane *L = lanes[g.lanes[0]];
            const PtArgs &a = L->ops[L->pos].a;
            if (g.shape == pt_default_shape(g.op, a)) {
                Step s;
                s.kind = 2;
                s.op = g.op;
                s.shape = g.shape;
                s.n = 1;
                s.first = a;
                out.push_back(s);
                L->pos++;
                continue;
            }
        }
        std::vector<const PtArgs *> args;
        std::vector<int> tiles, tx;
        int ntiles = 0;
        for (int li : g.lanes) {
            Lane *L = lanes[li];
            const PtArgs &a = L->ops[L->pos].a;
            int x = 0;
            int t = pt_shape_tiles(g.op, g.shape, a, &x);
            args.push_back(&a);
            tiles.push_back(t);
            tx.push_back(x);
            ntiles += t;
        }
        Step s;
        s.kind = 1;
        s.op = g.op;
        s.shape = g.shape;
        s.cls = g.cls;
        s.n = (int)args.size();
        s.ntiles = ntiles;
        s.first = *args[0];
        s.table_off = *toff;
        write_table(tables, *toff, args, tiles, tx);
        *toff += table_bytes(s.n);
        out.push_back(s);
        for (int li : g.lanes) lanes[li]->pos++;
    }
}

hipError_t issue_step(const Step &s, std::vector<Lane *> &lanes, const char *tables_dev, int xcd, hipStream_t st)
{
    if (s.kind == 0) return lanes[s.lane]->ops[s.opi].fn(st);
    if (s.kind == 2) return pt_run_op(s.op, s.first, st);
    const char *p = tables_dev + s.table_off;
    PtJobTable t;
    t.args = (const PtArgs *)p;
    t.prefix = (const int *)(p + (size_t)s.n * sizeof(PtArgs));
    t.tiles_x = t.prefix + s.n + 1;
    t.n = s.n;
    t.xcd_remap = xcd;
    return pt_launch_multi(s.op, s.shape, s.first, t, s.ntiles, st);
}

} // namespace

/* ------------------------------------------------------------------ */
/* batched form                                                        */
/* ------------------------------------------------------------------ */
struct pt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<pt_ctx *> ctxs;
    std::vector<Lane> lanes;
    std::vector<Lane *> lane_ptrs;
    ShapePolicy pol;
    int use_graph = 1, xcd = 1;
    /* program */
    bool prog_valid = false;
    std::vector<Step> prog;
    std::vector<char> tables;
    char *tables_dev = nullptr;
    size_t tables_cap = 0;
    int64_t ticks = 0, launches = 0, multi = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double last_ms = 0.0, tune_ms = 0.0;
    /* persistent form: every job's program in one launch */
    int persistent = 0;
    int pers_abort = 0; /* the next persistent launch starts with its give-up flag set */
    bool pers_valid = false;
    int pers_grid = 0, pers_max_chunk = 8;
    double pers_timeout_ms = 2000.0;
    int64_t pers_tasks = 0, pers_tiles = 0;
    std::vector<int> pers_order; /* jobs by priority */
    std::vector<unsigned long long> pers_claim0; /* claim words at the start of a launch */
    PtTask *pd_tasks = nullptr;
    int *pd_first = nullptr;
    PtArgs *pd_args = nullptr;
    unsigned long long *pd_claim = nullptr;
    int *pd_done = nullptr;
    int *pd_parts = nullptr;
    unsigned long long *pd_cost = nullptr;
    int *pd_ctl = nullptr;
    /* measured work per job (shader clocks, by job index), used for the
     * placement when set (pt_batch_tune) */
    std::vector<double> pers_measured;
    int pers_xcds = 0;  /* XCDs the persistent grid spans */
    int pers_spans = 0; /* jobs placed on more than one XCD */
};

static int bfail(pt_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    snprintf(g_engine_error, sizeof(g_engine_error), "%s", buf);
    return code;
}

#define B_HIP(b, expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return bfail((b), PT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

static void drop_graph(pt_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

static void free_persistent(pt_batch *b)
{
    (void)hipFree(b->pd_tasks);
    (void)hipFree(b->pd_first);
    (void)hipFree(b->pd_args);
    (void)hipFree(b->pd_claim);
    (void)hipFree(b->pd_done);
    (void)hipFree(b->pd_parts);
    (void)hipFree(b->pd_cost);
    (void)hipFree(b->pd_ctl);
    b->pd_tasks = nullptr;
    b->pd_first = nullptr;
    b->pd_args = nullptr;
    b->pd_claim = nullptr;
    b->pd_done = nullptr;
    b->pd_parts = nullptr;
    b->pd_cost = nullptr;
    b->pd_ctl = nullptr;
    b->pers_valid = false;
}

static const char *op_name(int op)
{
    static const char *const n[] = {"propagate", "scatter", "residual refresh", "sync"};
    return (op >= 0 && op < 4) ? n[op] : "operator";
}

/* Tiles per claim of a task whose bands hold `band` tiles each: a band is
 * shared by its XCD's workgroups, about two claims each for the
 * patch-distance operators (at most max_chunk tiles), one each for the
 * light kinds (copies, fills, syncs, pyramid and set-up kernels), whose
 * tiles take a few microseconds against several for a claim. */
static int task_chunk(const PtTask &t, int band, int wg_per_xcd, int max_chunk)
{
    const bool heavy = t.kind == PT_TASK_PROPAGATE || t.kind == PT_TASK_SCATTER || t.kind == PT_TASK_RESIDUAL;
    int c = heavy ? band / (2 * wg_per_xcd) : (band + wg_per_xcd - 1) / wg_per_xcd;
    const int cap = heavy ? max_chunk : 64;
    return c < 1 ? 1 : (c > cap ? cap : c);
}

/*
 * The persistent form's device program: each job's recorded plan becomes a
 * list of tasks, one per kernel the plain path would launch (operators with
 * a multi-job form directly, closures through a task sink that captures
 * their launches and copies).  Jobs are ordered by their tile count, most
 * first, since the claim scan serves lower indices first: the longest chains
 * of dependent steps start first and the short jobs fill the gaps.
 */
static int build_persistent(pt_batch *b)
{
    free_persistent(b);
    const int n = (int)b->lanes.size();
    std::vector<std::vector<PtTask>> jt(n);
    std::vector<std::vector<PtArgs>> ja(n);
    std::vector<int64_t> work(n, 0);
    std::vector<double> est(n, 0.0); /* estimated work: patch taps of the NNF steps */
    for (int j = 0; j < n; j++) {
        int64_t cum = 0;
        for (const RecOp &o : b->lanes[j].ops) {
            std::vector<PtTask> got;
            std::vector<PtArgs> args;
            if (o.op >= 0) {
                PtTask t;
                const char *why = nullptr;
                int shape = b->pol.pick(o.op, o.a);
                int rc = pt_task_op(o.op, shape, o.a, 0, &t, &why);
                if (rc != PT_OK && shape != pt_default_shape(o.op, o.a))
                    rc = pt_task_op(o.op, pt_default_shape(o.op, o.a), o.a, 0, &t, &why);
                if (rc != PT_OK)
                    return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: job %d, %s: %s", j, op_name(o.op), why);
                if (t.ntiles > 0) {
                    got.push_back(t);
                    args.push_back(o.a);
                }
            } else {
                PtTaskSink sink;
                pt_sink_begin(&sink);
                hipError_t e = o.fn(nullptr);
                pt_sink_end();
                if (sink.unsupported || e != hipSuccess)
                    return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: job %d, step '%s' has no device task (%s)", j,
                                 o.what ? o.what : "?", sink.unsupported ? sink.unsupported : hipGetErrorString(e));
                got = sink.tasks;
                args = sink.args;
            }
            for (PtTask &t : got) {
                if (t.ntiles > PT_PERSIST_MAX_CLAIMS)
                    return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: job %d has a step of %d tiles", j, t.ntiles);
                if (t.arg >= 0) {
                    ja[j].push_back(args[t.arg]);
                    t.arg = (int)ja[j].size() - 1; /* rebased below */
                }
                cum += t.ntiles;
                if (cum > (1ll << 30)) return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: job %d has too many tiles", j);
                if (t.kind <= PT_TASK_RESIDUAL) {
                    const PtArgs &a = ja[j][t.arg];
                    const double taps = (double)a.lw * a.lh * (2 * a.r + 1) * (2 * a.r + 1);
                    est[j] += taps * (t.kind == PT_TASK_PROPAGATE ? 2.0 : (t.kind == PT_TASK_SCATTER ? 0.5 * (a.nscat + 1) : 0.5));
                } else {
                    est[j] += 10.0 * t.ntiles;
                }
                jt[j].push_back(t);
            }
        }
        work[j] = cum;
        if (jt[j].empty()) return bfail(b, PT_ERR_STATE, "persistent form: job %d has no work", j);
        if (jt[j].size() >= (size_t)PT_PERSIST_MAX_STEPS)
            return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: job %d has too many steps", j);
    }
    if ((int)b->pers_measured.size() == n) est = b->pers_measured;
    std::vector<int> order(n);
    for (int j = 0; j < n; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return est[x] > est[y]; });
    b->pers_grid = pt_persistent_grid(b->device);
    if (b->pers_grid < 1) return bfail(b, PT_ERR_HIP, "persistent form: no occupancy for the persistent kernel");
    /* the XCDs that run the persistent grid's workgroups */
    unsigned mask = 0, *dmask = nullptr;
    B_HIP(b, hipMalloc((void **)&dmask, sizeof(unsigned)));
    hipError_t pe = hipMemsetAsync(dmask, 0, sizeof(unsigned), b->stream);
    if (pe == hipSuccess) pe = pt_launch_xcd_probe(b->pers_grid, dmask, b->stream);
    if (pe == hipSuccess) pe = hipStreamSynchronize(b->stream);
    if (pe == hipSuccess) pe = hipMemcpy(&mask, dmask, sizeof(unsigned), hipMemcpyDeviceToHost);
    (void)hipFree(dmask);
    if (pe != hipSuccess) return bfail(b, PT_ERR_HIP, "persistent form: XCD probe failed: %s", hipGetErrorString(pe));
    std::vector<int> xcds;
    for (int x = 0; x < 16; x++)
        if (mask & (1u << x)) xcds.push_back(x);
    if (xcds.empty()) return bfail(b, PT_ERR_HIP, "persistent form: XCD probe found no XCD");
    b->pers_xcds = (int)xcds.size();
    const int wg_per_xcd = b->pers_grid / b->pers_xcds > 0 ? b->pers_grid / b->pers_xcds : 1;
    if (b->pers_xcds > 8) return bfail(b, PT_ERR_UNSUPPORTED, "persistent form: more than 8 XCDs");
    /* longest job first onto the least loaded XCDs; a job with more than
     * one XCD's share of the batch gets ceil(work / share) XCDs */
    const int nx = (int)xcds.size();
    double total = 0.0;
    for (double e : est) total += e;
    const double share = total / nx;
    std::vector<double> load(nx, 0.0);
    std::vector<int> xmask_of(n, 0);
    for (int j : order) {
        int k = share > 0.0 ? (int)std::ceil(est[j] / share - 1e-9) : 1;
        k = k < 1 ? 1 : (k > nx ? nx : k);
        std::vector<int> idx(nx);
        for (int x = 0; x < nx; x++) idx[x] = x;
        std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) { return load[u] < load[v]; });
        for (int q = 0; q < k; q++) {
            load[idx[q]] += est[j] / k;
            xmask_of[j] |= 1 << xcds[idx[q]];
        }
    }
    std::vector<PtTask> tasks;
    std::vector<PtArgs> args;
    std::vector<int> first;
    b->pers_claim0.clear();
    for (int j : order) {
        first.push_back((int)tasks.size());
        const int base = (int)args.size();
        const int nb = __builtin_popcount((unsigned)xmask_of[j]);
        int64_t band_cum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int parts = 0;
        for (PtTask t : jt[j]) {
            if (t.arg >= 0) t.arg += base;
            /* contiguous bands of the tiles, one per XCD of the job */
            for (int q = 0; q <= 8; q++) t.band[q] = q <= nb ? (int)((int64_t)t.ntiles * q / nb) : t.ntiles;
            for (int q = 0; q < 8; q++) {
                const int len = q < nb ? t.band[q + 1] - t.band[q] : 0;
                band_cum[q] += len;
                t.band_end[q] = (int)band_cum[q];
                parts += len > 0;
            }
            t.parts_end = parts;
            t.chunk = task_chunk(t, (t.ntiles + nb - 1) / nb, wg_per_xcd, b->pers_max_chunk);
            tasks.push_back(t);
        }
        const PtTask &t0 = tasks[first.back()];
        for (int q = 0; q < 8; q++)
            b->pers_claim0.push_back(
                q < nb ? pt_claim_word(0, (t0.band[q + 1] - t0.band[q] + t0.chunk - 1) / t0.chunk) : 0ull);
        args.insert(args.end(), ja[j].begin(), ja[j].end());
    }
    first.push_back((int)tasks.size());
    for (int j : order) first.push_back(xmask_of[j]);
    b->pers_spans = 0;
    for (int j = 0; j < n; j++) b->pers_spans += __builtin_popcount((unsigned)xmask_of[j]) > 1;
    if (args.empty()) args.push_back(PtArgs{});
    B_HIP(b, hipMalloc((void **)&b->pd_tasks, tasks.size() * sizeof(PtTask)));
    B_HIP(b, hipMalloc((void **)&b->pd_first, first.size() * sizeof(int)));
    B_HIP(b, hipMalloc((void **)&b->pd_args, args.size() * sizeof(PtArgs)));
    B_HIP(b, hipMalloc((void **)&b->pd_claim, (size_t)n * 8 * sizeof(unsigned long long)));
    B_HIP(b, hipMalloc((void **)&b->pd_done, (size_t)n * 8 * sizeof(int)));
    B_HIP(b, hipMalloc((void **)&b->pd_parts, (size_t)n * sizeof(int)));
    B_HIP(b, hipMalloc((void **)&b->pd_cost, (size_t)n * sizeof(unsigned long long)));
    B_HIP(b, hipMalloc((void **)&b->pd_ctl, 4 * sizeof(int)));
    B_HIP(b, hipMemcpy(b->pd_tasks, tasks.data(), tasks.size() * sizeof(PtTask), hipMemcpyHostToDevice));
    B_HIP(b, hipMemcpy(b->pd_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
    B_HIP(b, hipMemcpy(b->pd_args, args.data(), args.size() * sizeof(PtArgs), hipMemcpyHostToDevice));
    b->pers_tasks = (int64_t)tasks.size();
    b->pers_tiles = 0;
    for (int64_t w : work) b->pers_tiles += w;
    b->pers_order = order;
    b->pers_valid = true;
    return PT_OK;
}

/* One persistent launch over every job; the device-side state is cleared
 * first, and the launch's outcome is read back. */
static int run_persistent(pt_batch *b)
{
    if (!b->pers_valid) {
        int rc = build_persistent(b);
        if (rc != PT_OK) return rc;
    }
    const int n = (int)b->lanes.size();
    B_HIP(b, hipMemcpy(b->pd_claim, b->pers_claim0.data(), b->pers_claim0.size() * sizeof(unsigned long long),
                       hipMemcpyHostToDevice));
    B_HIP(b, hipMemsetAsync(b->pd_done, 0, (size_t)n * 8 * sizeof(int), b->stream));
    B_HIP(b, hipMemsetAsync(b->pd_parts, 0, (size_t)n * sizeof(int), b->stream));
    B_HIP(b, hipMemsetAsync(b->pd_cost, 0, (size_t)n * sizeof(unsigned long long), b->stream));
    const int ctl0[4] = {0, 0, b->pers_abort ? 1 : 0, 0};
    b->pers_abort = 0;
    B_HIP(b, hipMemcpy(b->pd_ctl, ctl0, sizeof(ctl0), hipMemcpyHostToDevice));
    int khz = 0;
    B_HIP(b, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    PtPersist P;
    P.tasks = b->pd_tasks;
    P.first = b->pd_first;
    P.args = b->pd_args;
    P.claim = b->pd_claim;
    P.done = b->pd_done;
    P.parts = b->pd_parts;
    P.cost = b->pd_cost;
    P.ctl = b->pd_ctl;
    P.njobs = n;
    P.timeout_ticks = (long long)((double)(khz > 0 ? khz : 100000) * b->pers_timeout_ms);
    B_HIP(b, pt_launch_persistent(P, b->pers_grid, b->stream));
    B_HIP(b, hipStreamSynchronize(b->stream));
    int ctl[4] = {0, 0, 0, 0};
    B_HIP(b, hipMemcpy(ctl, b->pd_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    if (ctl[2] || ctl[0] != n)
        return bfail(b, PT_ERR_HIP, "persistent launch gave up: %d of %d jobs finished, no progress for %.0f ms", ctl[0],
                     n, b->pers_timeout_ms);
    return PT_OK;
}

static int build_program(pt_batch *b)
{
    drop_graph(b);
    b->prog.clear();
    b->tables.clear();
    b->ticks = 0;
    for (auto &L : b->lanes) L.pos = 0;
    size_t toff = 0;
    while (true) {
        bool any = false;
        for (auto &L : b->lanes) any = any || !L.done();
        if (!any) break;
        build_tick(b->lane_ptrs, b->pol, b->prog, b->tables, &toff);
        b->ticks++;
    }
    b->launches = (int64_t)b->prog.size();
    b->multi = 0;
    for (auto &s : b->prog) b->multi += s.kind == 1;
    if (b->tables.size() > b->tables_cap) {
        if (b->tables_dev) B_HIP(b, hipFree(b->tables_dev));
        b->tables_dev = nullptr;
        b->tables_cap = 0;
        B_HIP(b, hipMalloc((void **)&b->tables_dev, b->tables.size()));
        b->tables_cap = b->tables.size();
    }
    if (!b->tables.empty())
        B_HIP(b, hipMemcpy(b->tables_dev, b->tables.data(), b->tables.size(), hipMemcpyHostToDevice));
    b->prog_valid = true;
    return PT_OK;
}

static int issue_program(pt_batch *b, std::vector<float> *step_ms)
{
    std::vector<hipEvent_t> ev;
    if (step_ms) {
        ev.resize(b->prog.size() + 1);
        for (auto &e : ev) B_HIP(b, hipEventCreate(&e));
    }
    int rc = PT_OK;
    for (size_t i = 0; i < b->prog.size() && rc == PT_OK; i++) {
        if (step_ms && b->prog[i].kind == 1) (void)hipEventRecord(ev[i], b->stream);
        hipError_t e = issue_step(b->prog[i], b->lane_ptrs, b->tables_dev, b->xcd, b->stream);
        if (e != hipSuccess) rc = bfail(b, PT_ERR_HIP, "batch step %zu failed: %s", i, hipGetErrorString(e));
        if (step_ms && b->prog[i].kind == 1) (void)hipEventRecord(ev[i + 1], b->stream);
    }
    if (step_ms) {
        hipError_t e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess && rc == PT_OK) rc = bfail(b, PT_ERR_HIP, "batch run failed: %s", hipGetErrorString(e));
        step_ms->assign(b->prog.size(), 0.f);
        for (size_t i = 0; i < b->prog.size() && rc == PT_OK; i++)
            if (b->prog[i].kind == 1) (void)hipEventElapsedTime(&(*step_ms)[i], ev[i], ev[i + 1]);
        for (auto &x : ev) (void)hipEventDestroy(x);
    }
    return rc;
}

extern "C" int pt_batch_create(pt_batch **out, int device, int num_jobs, const pt_job *jobs)
{
    if (!out || num_jobs < 1 || !jobs) return bfail(nullptr, PT_ERR_INVALID_ARG, "bad batch arguments");
    *out = nullptr;
    pt_batch *b = new pt_batch();
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        bfail(nullptr, PT_ERR_HIP, "batch stream: %s", hipGetErrorString(e));
        delete b;
        return PT_ERR_HIP;
    }
    b->lanes.resize(num_jobs);
    for (int j = 0; j < num_jobs; j++) {
        const pt_job &J = jobs[j];
        pt_ctx *c = nullptr;
        int rc = pt_create(&c, device);
        if (rc == PT_OK) {
            b->ctxs.push_back(c);
            rc = pt_record_synthesis(c, &b->lanes[j].ops, &J.params, J.scheme, J.iterations, J.fixup_passes,
                                     J.image_w, J.image_h, J.num_frames, J.frames, J.style_w, J.style_h, J.style,
                                     J.out_offsets, J.out_canvas_rgba, J.out_residual);
        }
        if (rc != PT_OK) {
            std::string m = std::string("job ") + std::to_string(j) + ": " + (c ? pt_last_error(c) : pt_global_last_error());
            bfail(nullptr, rc, "%s", m.c_str());
            pt_batch_destroy(b);
            return rc;
        }
    }
    for (auto &L : b->lanes) b->lane_ptrs.push_back(&L);
    *out = b;
    return PT_OK;
}

extern "C" int pt_batch_destroy(pt_batch *b)
{
    if (!b) return PT_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    free_persistent(b);
    if (b->tables_dev) (void)hipFree(b->tables_dev);
    for (auto *c : b->ctxs) pt_destroy(c);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PT_OK;
}

extern "C" const char *pt_batch_last_error(pt_batch *b) { return b ? b->err.c_str() : g_engine_error; }

extern "C" int pt_batch_run(pt_batch *b)
{
    if (!b) return bfail(nullptr, PT_ERR_INVALID_ARG, "NULL batch");
    B_HIP(b, hipSetDevice(b->device));
    double t0 = now_ms();
    if (b->persistent) {
        int rc = run_persistent(b);
        if (rc != PT_OK) return rc;
        b->last_ms = now_ms() - t0;
        return PT_OK;
    }
    if (!b->prog_valid) {
        int rc = build_program(b);
        if (rc != PT_OK) return rc;
    }
    if (b->use_graph) {
        if (!b->gexec) {
            B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = issue_program(b, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc != PT_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return bfail(b, PT_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            b->graph = g;
            B_HIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        B_HIP(b, hipGraphLaunch(b->gexec, b->stream));
    } else {
        int rc = issue_program(b, nullptr);
        if (rc != PT_OK) return rc;
    }
    B_HIP(b, hipStreamSynchronize(b->stream));
    b->last_ms = now_ms() - t0;
    return PT_OK;
}

extern "C" int pt_batch_tune(pt_batch *b, int rounds)
{
    if (!b || rounds < 1) return bfail(b, PT_ERR_INVALID_ARG, "bad tune arguments");
    B_HIP(b, hipSetDevice(b->device));
    double t0 = now_ms();
    if (b->persistent) {
        /* the persistent form: first the placement by measured work (one
         * launch with the estimated placement, the shader clocks of each
         * job's claims read back), then the tiles per claim, on whole-batch
         * time */
        {
            b->pers_measured.clear();
            b->pers_valid = false;
            int rc = run_persistent(b);
            if (rc != PT_OK) return rc;
            const int n = (int)b->lanes.size();
            std::vector<unsigned long long> cost(n);
            B_HIP(b, hipMemcpy(cost.data(), b->pd_cost, (size_t)n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            /* pd_cost is in the launch's job order */
            b->pers_measured.assign(n, 0.0);
            for (int k = 0; k < n; k++) b->pers_measured[b->pers_order[k]] = (double)cost[k];
            b->pers_valid = false;
        }
        const int cand[] = {1, 2, 4, 8, 16};
        double best = 0.0;
        int pick = b->pers_max_chunk;
        for (int c : cand) {
            b->pers_max_chunk = c;
            b->pers_valid = false;
            double tt = 0.0;
            for (int r = 0; r < rounds; r++) {
                double t1 = now_ms();
                int rc = run_persistent(b);
                if (rc != PT_OK) return rc;
                tt += now_ms() - t1;
            }
            if (c == cand[0] || tt < best) {
                best = tt;
                pick = c;
            }
        }
        b->pers_max_chunk = pick;
        b->pers_valid = false;
        b->tune_ms += now_ms() - t0;
        return PT_OK;
    }
    ShapePolicy keep = b->pol;
    /* time[op][class][shape], launches counted to check coverage */
    double tm[2][3][PT_SHAPE_COUNT] = {};
    int64_t cnt[2][3][PT_SHAPE_COUNT] = {};
    int64_t units[2][3] = {};
    for (int s = 0; s < PT_SHAPE_COUNT; s++) {
        ShapePolicy p;
        p.isolate_tiles = keep.isolate_tiles;
        for (int op = 0; op < 2; op++)
            for (int c = 0; c < 3; c++) p.shape[op][c] = s;
        b->pol = p;
        int rc = build_program(b);
        if (rc != PT_OK) {
            b->pol = keep;
            return rc;
        }
        for (int r = 0; r < rounds; r++) {
            std::vector<float> ms;
            rc = issue_program(b, &ms);
            if (rc != PT_OK) {
                b->pol = keep;
                b->prog_valid = false;
                return rc;
            }
            for (size_t i = 0; i < b->prog.size(); i++) {
                const Step &st = b->prog[i];
                if (st.kind != 1 || st.op > PT_OP_SCATTER) continue;
                tm[st.op][st.cls][st.shape] += ms[i];
                cnt[st.op][st.cls][st.shape] += st.n;
                if (s == 0 && r == 0) units[st.op][st.cls] += st.n;
            }
        }
    }
    /* a shape qualifies for (op, class) only if it served every job-step
     * of that class in its run (valid everywhere) */
    ShapePolicy best;
    best.isolate_tiles = keep.isolate_tiles;
    for (int op = 0; op < 2; op++)
        for (int c = 0; c < 3; c++) {
            double bt = 0.0;
            int bs = -1;
            for (int s = 0; s < PT_SHAPE_COUNT; s++) {
                if (cnt[op][c][s] != units[op][c] * rounds || units[op][c] == 0) continue;
                if (bs < 0 || tm[op][c][s] < bt) {
                    bt = tm[op][c][s];
                    bs = s;
                }
            }
            best.shape[op][c] = bs;
        }
    b->pol = best;
    /* then the XCD tile mapping, on whole-batch time */
    int rc = build_program(b);
    if (rc != PT_OK) return rc;
    double tx[2] = {0.0, 0.0};
    for (int x = 0; x < 2; x++) {
        b->xcd = x;
        for (int r = 0; r < rounds; r++) {
            double t1 = now_ms();
            rc = issue_program(b, nullptr);
            if (rc == PT_OK) {
                hipError_t e = hipStreamSynchronize(b->stream);
                if (e != hipSuccess) rc = bfail(b, PT_ERR_HIP, "tune run failed: %s", hipGetErrorString(e));
            }
            if (rc != PT_OK) return rc;
            tx[x] += now_ms() - t1;
        }
    }
    b->xcd = tx[1] <= tx[0] ? 1 : 0;
    b->prog_valid = false;
    drop_graph(b);
    b->tune_ms += now_ms() - t0;
    return PT_OK;
}

extern "C" int pt_batch_set(pt_batch *b, int key, int value)
{
    if (!b) return bfail(nullptr, PT_ERR_INVALID_ARG, "NULL batch");
    switch (key) {
    case PT_BATCH_USE_GRAPH: b->use_graph = value != 0; break;
    case PT_BATCH_XCD_REMAP: b->xcd = value != 0; break;
    case PT_BATCH_ISOLATE_TILES:
        if (value < 1) return bfail(b, PT_ERR_INVALID_ARG, "isolate threshold must be >= 1");
        b->pol.isolate_tiles = value;
        break;
    case PT_BATCH_SHAPE: {
        int op = value / 256, cls = (value / 16) % 16, shape = value % 16;
        if (op < 0 || op > 1 || cls > 2 || (shape >= PT_SHAPE_COUNT && shape != 15))
            return bfail(b, PT_ERR_INVALID_ARG, "bad shape setting %d", value);
        b->pol.shape[op][cls] = shape == 15 ? -1 : shape;
        break;
    }
    case PT_BATCH_PERSISTENT: b->persistent = value != 0; return PT_OK;
    case PT_BATCH_PERSIST_CHUNK:
        if (value < 1 || value > 1024) return bfail(b, PT_ERR_INVALID_ARG, "tiles per claim must be 1..1024");
        b->pers_max_chunk = value;
        b->pers_valid = false;
        return PT_OK;
    case PT_BATCH_PERSIST_TIMEOUT_MS:
        if (value < 1) return bfail(b, PT_ERR_INVALID_ARG, "timeout must be >= 1 ms");
        b->pers_timeout_ms = (double)value;
        return PT_OK;
    case PT_BATCH_PERSIST_ABORT: b->pers_abort = value != 0; return PT_OK;
    default: return bfail(b, PT_ERR_INVALID_ARG, "unknown batch key %d", key);
    }
    b->prog_valid = false;
    b->pers_valid = false;
    drop_graph(b);
    return PT_OK;
}

extern "C" int pt_batch_get_stats(pt_batch *b, pt_batch_stats *o)
{
    if (!b || !o) return bfail(b, PT_ERR_INVALID_ARG, "NULL argument");
    if (!b->prog_valid) {
        B_HIP(b, hipSetDevice(b->device));
        int rc = build_program(b);
        if (rc != PT_OK) return rc;
    }
    memset(o, 0, sizeof(*o));
    o->jobs = (int64_t)b->lanes.size();
    for (auto &L : b->lanes) o->ops += (int64_t)L.ops.size();
    o->launches = b->launches;
    o->multi_launches = b->multi;
    o->ticks = b->ticks;
    o->last_run_ms = b->last_ms;
    o->tune_ms = b->tune_ms;
    for (int op = 0; op < 2; op++)
        for (int c = 0; c < 3; c++) o->shape[op][c] = b->pol.shape[op][c];
    o->xcd_remap = b->xcd;
    o->isolate_tiles = b->pol.isolate_tiles;
    o->persistent = b->persistent;
    o->persistent_max_chunk = b->pers_max_chunk;
    if (b->persistent && !b->pers_valid) {
        int rc = build_persistent(b);
        if (rc != PT_OK) return rc;
    }
    o->persistent_grid = b->pers_valid ? b->pers_grid : 0;
    o->persistent_tasks = b->pers_valid ? b->pers_tasks : 0;
    o->persistent_tiles = b->pers_valid ? b->pers_tiles : 0;
    o->persistent_spanned = b->pers_valid ? b->pers_spans : 0;
    return PT_OK;
}

extern "C" int pt_submit_batch(int device, int num_jobs, const pt_job *jobs)
{
    pt_batch *b = nullptr;
    int rc = pt_batch_create(&b, device, num_jobs, jobs);
    if (rc != PT_OK) return rc;
    rc = pt_batch_run(b);
    if (rc != PT_OK) snprintf(g_engine_error, sizeof(g_engine_error), "%s", b->err.c_str());
    pt_batch_destroy(b);
    return rc;
}

/* ------------------------------------------------------------------ */
/* serving form                                                        */
/* ------------------------------------------------------------------ */
namespace {

struct FrameReq {
    uint64_t ticket = 0;
    pt_frame_guides frame{};
    pt_patchxf *out_off = nullptr;
    float *out_canvas = nullptr, *out_res = nullptr;
    double t_submit = 0.0;
};

struct Seq {
    pt_sequence_desc desc{};
    pt_style_guides style{};
    pt_ctx *ctx = nullptr;
    bool open = false, resident = false, busy = false, closing = false;
    size_t bytes = 0;
    std::deque<FrameReq> queue;
};

struct Flight {
    int seq = -1;
    FrameReq req;
    Lane lane;
    bool issued = false; /* every op issued; done_ev recorded after the last */
    hipEvent_t done_ev = nullptr;
};

} // namespace

struct pt_server {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t budget = 0, resident_bytes = 0;
    int max_in_flight = 1;
    std::vector<Seq> seqs;
    std::vector<Flight *> flights;
    uint64_t next_ticket = 1;
    ShapePolicy pol;
    /* table staging ring */
    int nslots = 8;
    size_t slot_bytes = 0;
    char *ring_host = nullptr, *ring_dev = nullptr;
    std::vector<hipEvent_t> slot_ev;
    std::vector<bool> slot_used;
    int slot_next = 0;
    int xcd = 0;
    /* latency policy: a tick advances the in-flight frames oldest first while
     * their steps fit this many workgroup tiles (the oldest always goes);
     * 0 advances every in-flight frame each tick */
    int fill_tiles = 2048;
    /* ticks planned per table upload (one host-to-device copy for all of them) */
    int ticks_ahead = 16;
    std::vector<std::pair<uint64_t, double>> completed;
    int64_t submitted = 0, ncompleted = 0, ticks = 0, launches = 0, multi = 0;
};

static int sfail(pt_server *s, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    snprintf(g_engine_error, sizeof(g_engine_error), "%s", buf);
    return code;
}

#define S_HIP(s, expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return sfail((s), PT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" int pt_server_create(pt_server **out, int device, uint64_t budget, int max_in_flight)
{
    if (!out || max_in_flight < 1 || budget == 0) return sfail(nullptr, PT_ERR_INVALID_ARG, "bad server arguments");
    *out = nullptr;
    pt_server *s = new pt_server();
    s->device = device;
    s->budget = budget;
    s->max_in_flight = max_in_flight;
    /* worst case of one tick: every in-flight frame in a launch group of its own */
    s->slot_bytes = (size_t)s->ticks_ahead * max_in_flight * (sizeof(PtArgs) + 3 * sizeof(int) + kTableAlign);
    s->slot_bytes = (s->slot_bytes + kTableAlign - 1) / kTableAlign * kTableAlign;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->ring_host, s->slot_bytes * s->nslots, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&s->ring_dev, s->slot_bytes * s->nslots);
    s->slot_ev.assign(s->nslots, nullptr);
    s->slot_used.assign(s->nslots, false);
    for (int i = 0; i < s->nslots && e == hipSuccess; i++) e = hipEventCreateWithFlags(&s->slot_ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        sfail(nullptr, PT_ERR_HIP, "server setup: %s", hipGetErrorString(e));
        pt_server_destroy(s);
        return PT_ERR_HIP;
    }
    *out = s;
    return PT_OK;
}

static void release_seq(pt_server *s, Seq &q)
{
    if (q.ctx) pt_destroy(q.ctx);
    q.ctx = nullptr;
    if (q.resident) s->resident_bytes -= q.bytes;
    q.resident = false;
    q.open = false;
}

extern "C" int pt_server_destroy(pt_server *s)
{
    if (!s) return PT_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto *f : s->flights) {
        if (f->done_ev) (void)hipEventDestroy(f->done_ev);
        delete f;
    }
    for (auto &q : s->seqs) release_seq(s, q);
    for (auto e : s->slot_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->ring_host) (void)hipHostFree(s->ring_host);
    if (s->ring_dev) (void)hipFree(s->ring_dev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return PT_OK;
}

extern "C" const char *pt_server_last_error(pt_server *s) { return s ? s->err.c_str() : g_engine_error; }

extern "C" int pt_server_open(pt_server *s, const pt_sequence_desc *d, int32_t *seq)
{
    if (!s || !d || !seq || !d->style) return sfail(s, PT_ERR_INVALID_ARG, "NULL argument");
    if (d->iterations < 1 || d->image_w < 2 || d->image_h < 2 || d->style_w < 2 || d->style_h < 2)
        return sfail(s, PT_ERR_INVALID_ARG, "bad sequence shape");
    Seq q;
    q.desc = *d;
    q.style = *d->style;
    q.desc.style = nullptr;
    q.open = true;
    q.bytes = pt_sequence_bytes(d->image_w, d->image_h, d->style_w, d->style_h, d->params.levels,
                                d->params.residual_window_size);
    if (q.bytes > s->budget) return sfail(s, PT_ERR_BUDGET, "sequence needs %zu bytes, budget is %llu", q.bytes,
                                          (unsigned long long)s->budget);
    s->seqs.push_back(q);
    *seq = (int32_t)s->seqs.size() - 1;
    return PT_OK;
}

extern "C" int pt_server_submit(pt_server *s, int32_t seq, const pt_frame_guides *frame, pt_patchxf *out_off,
                                float *out_canvas, float *out_res, uint64_t *ticket)
{
    if (!s || !frame || !ticket) return sfail(s, PT_ERR_INVALID_ARG, "NULL argument");
    if (seq < 0 || seq >= (int)s->seqs.size() || !s->seqs[seq].open || s->seqs[seq].closing)
        return sfail(s, PT_ERR_INVALID_ARG, "sequence %d is not open", seq);
    FrameReq r;
    r.ticket = s->next_ticket++;
    r.frame = *frame;
    r.out_off = out_off;
    r.out_canvas = out_canvas;
    r.out_res = out_res;
    r.t_submit = now_ms();
    s->seqs[seq].queue.push_back(r);
    s->submitted++;
    *ticket = r.ticket;
    return PT_OK;
}

/* Admits queued frames into free in-flight slots, oldest ticket first; a
 * sequence's first admitted frame makes it resident if the budget allows. */
static int admit(pt_server *s)
{
    while ((int)s->flights.size() < s->max_in_flight) {
        int pick = -1;
        uint64_t oldest = 0;
        for (size_t i = 0; i < s->seqs.size(); i++) {
            Seq &q = s->seqs[i];
            if (q.busy || q.queue.empty()) continue;
            if (!q.resident && s->resident_bytes + q.bytes > s->budget) continue;
            if (pick < 0 || q.queue.front().ticket < oldest) {
                pick = (int)i;
                oldest = q.queue.front().ticket;
            }
        }
        if (pick < 0) return PT_OK;
        Seq &q = s->seqs[pick];
        if (!q.resident) {
            int rc = pt_create(&q.ctx, s->device);
            if (rc != PT_OK) return sfail(s, rc, "sequence %d: %s", pick, pt_global_last_error());
            q.resident = true;
            s->resident_bytes += q.bytes;
        }
        Flight *f = new Flight();
        f->seq = pick;
        f->req = q.queue.front();
        q.queue.pop_front();
        const pt_sequence_desc &d = q.desc;
        int rc = pt_record_next_frame(q.ctx, &f->lane.ops, &d.params, d.iterations, d.image_w, d.image_h,
                                      &f->req.frame, d.style_w, d.style_h, &q.style, f->req.out_off,
                                      f->req.out_canvas, f->req.out_res);
        if (rc != PT_OK) {
            std::string m = pt_last_error(q.ctx);
            delete f;
            return sfail(s, rc, "sequence %d frame: %s", pick, m.c_str());
        }
        hipError_t e = hipEventCreateWithFlags(&f->done_ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            delete f;
            return sfail(s, PT_ERR_HIP, "event: %s", hipGetErrorString(e));
        }
        q.busy = true;
        s->flights.push_back(f);
    }
    return PT_OK;
}

/* Workgroup tiles of the next multi-job step of a lane (its leading
 * closures skipped), 0 when none is left. */
static int next_step_tiles(const Lane &L, const ShapePolicy &pol)
{
    size_t p = L.pos;
    while (p < L.ops.size() && L.ops[p].op < 0) p++;
    if (p >= L.ops.size()) return 0;
    const RecOp &o = L.ops[p];
    return pt_shape_tiles(o.op, pol.pick(o.op, o.a), o.a, nullptr);
}

/* The lanes one tick advances: oldest frame first, its step always;
 * younger frames join while the tiles taken so far are below fill_tiles, so
 * a step that fills the GPU runs alone and the GPU left idle by a small
 * (coarse-level) step is filled with the steps of younger frames.  Empty when
 * no in-flight frame has work left. */
static std::vector<Lane *> tick_lanes(pt_server *s, const std::vector<int> &order)
{
    std::vector<Lane *> lanes(s->flights.size(), nullptr);
    int used = 0;
    bool any = false;
    for (int i : order) {
        Flight *f = s->flights[i];
        if (f->issued || f->lane.done()) continue;
        if (any && s->fill_tiles > 0 && used >= s->fill_tiles) continue;
        lanes[i] = &f->lane;
        used += next_step_tiles(f->lane, s->pol);
        any = true;
    }
    if (!any) lanes.clear();
    return lanes;
}

/* Plans up to min(ticks_ahead, limit) ticks, uploads their launch tables
 * with one copy and issues them; a frame's completion event is recorded
 * right after the tick that issued its last step.  *ran receives the number
 * of ticks issued. */
static int tick(pt_server *s, int limit, int *ran)
{
    *ran = 0;
    std::vector<int> order(s->flights.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return s->flights[a]->req.ticket < s->flights[b]->req.ticket; });
    std::vector<Step> steps;
    std::vector<char> tables;
    std::vector<size_t> tick_end;            /* steps issued by the first k+1 ticks */
    std::vector<int> done_tick(s->flights.size(), -1);
    size_t toff = 0;
    std::vector<Lane *> all(s->flights.size(), nullptr);
    for (size_t i = 0; i < s->flights.size(); i++) all[i] = &s->flights[i]->lane;
    for (int t = 0; t < std::min(s->ticks_ahead, limit); t++) {
        std::vector<Lane *> lanes = tick_lanes(s, order);
        if (lanes.empty()) break;
        build_tick(lanes, s->pol, steps, tables, &toff);
        tick_end.push_back(steps.size());
        for (size_t i = 0; i < s->flights.size(); i++)
            if (done_tick[i] < 0 && !s->flights[i]->issued && s->flights[i]->lane.done()) done_tick[i] = t;
    }
    if (steps.empty()) return PT_OK;
    const char *dev = s->ring_dev;
    int k = -1;
    if (toff > 0) {
        if (toff > s->slot_bytes) return sfail(s, PT_ERR_STATE, "tick tables exceed the staging slot");
        k = s->slot_next;
        s->slot_next = (k + 1) % s->nslots;
        if (s->slot_used[k]) S_HIP(s, hipEventSynchronize(s->slot_ev[k]));
        memcpy(s->ring_host + k * s->slot_bytes, tables.data(), toff);
        S_HIP(s, hipMemcpyAsync(s->ring_dev + k * s->slot_bytes, s->ring_host + k * s->slot_bytes, toff,
                                hipMemcpyHostToDevice, s->stream));
        dev = s->ring_dev + k * s->slot_bytes;
    }
    size_t si = 0;
    for (size_t t = 0; t < tick_end.size(); t++) {
        for (; si < tick_end[t]; si++) {
            hipError_t e = issue_step(steps[si], all, dev, s->xcd, s->stream);
            if (e != hipSuccess) return sfail(s, PT_ERR_HIP, "tick launch failed: %s", hipGetErrorString(e));
            s->launches++;
            s->multi += steps[si].kind == 1;
        }
        s->ticks++;
        (*ran)++;
        for (size_t i = 0; i < s->flights.size(); i++)
            if (done_tick[i] == (int)t) {
                S_HIP(s, hipEventRecord(s->flights[i]->done_ev, s->stream));
                s->flights[i]->issued = true;
            }
    }
    if (k >= 0) {
        S_HIP(s, hipEventRecord(s->slot_ev[k], s->stream));
        s->slot_used[k] = true;
    }
    return PT_OK;
}

static int collect(pt_server *s)
{
    for (size_t i = 0; i < s->flights.size();) {
        Flight *f = s->flights[i];
        if (f->issued) {
            hipError_t e = hipEventQuery(f->done_ev);
            if (e == hipSuccess) {
                s->completed.push_back({f->req.ticket, now_ms() - f->req.t_submit});
                s->ncompleted++;
                Seq &q = s->seqs[f->seq];
                q.busy = false;
                (void)hipEventDestroy(f->done_ev);
                delete f;
                s->flights.erase(s->flights.begin() + i);
                if (q.closing && q.queue.empty()) release_seq(s, q);
                continue;
            }
            if (e != hipErrorNotReady) return sfail(s, PT_ERR_HIP, "frame failed: %s", hipGetErrorString(e));
        }
        i++;
    }
    return PT_OK;
}

extern "C" int pt_server_poll(pt_server *s, int max_ticks, uint64_t *done, double *lat, int max_done, int *n_done)
{
    if (!s || max_ticks < 0 || (max_done > 0 && !done) || !n_done) return sfail(s, PT_ERR_INVALID_ARG, "bad poll");
    S_HIP(s, hipSetDevice(s->device));
    int rc;
    for (int t = 0; t < max_ticks;) {
        if ((rc = collect(s)) != PT_OK) return rc;
        if ((rc = admit(s)) != PT_OK) return rc;
        bool work = false;
        for (auto *f : s->flights) work = work || !f->issued;
        if (!work) break;
        int ran = 0;
        if ((rc = tick(s, max_ticks - t, &ran)) != PT_OK) return rc;
        t += std::max(ran, 1);
    }
    if ((rc = collect(s)) != PT_OK) return rc;
    int n = std::min<int>(max_done, (int)s->completed.size());
    for (int i = 0; i < n; i++) {
        done[i] = s->completed[i].first;
        if (lat) lat[i] = s->completed[i].second;
    }
    s->completed.erase(s->completed.begin(), s->completed.begin() + n);
    *n_done = n;
    return PT_OK;
}

extern "C" int pt_server_drain(pt_server *s)
{
    if (!s) return sfail(nullptr, PT_ERR_INVALID_ARG, "NULL server");
    S_HIP(s, hipSetDevice(s->device));
    while (true) {
        int rc = collect(s);
        if (rc == PT_OK) rc = admit(s);
        if (rc != PT_OK) return rc;
        bool pending = !s->flights.empty();
        for (auto &q : s->seqs) pending = pending || !q.queue.empty();
        if (!pending) return PT_OK;
        bool work = false;
        for (auto *f : s->flights) work = work || !f->issued;
        if (work) {
            int ran = 0;
            rc = tick(s, s->ticks_ahead, &ran);
            if (rc != PT_OK) return rc;
        } else if (!s->flights.empty()) {
            S_HIP(s, hipStreamSynchronize(s->stream));
        } else {
            /* queued frames that no budget can admit */
            return sfail(s, PT_ERR_BUDGET, "queued frames cannot be admitted within the memory budget");
        }
    }
}

extern "C" int pt_server_close(pt_server *s, int32_t seq)
{
    if (!s || seq < 0 || seq >= (int)s->seqs.size() || !s->seqs[seq].open)
        return sfail(s, PT_ERR_INVALID_ARG, "bad sequence");
    Seq &q = s->seqs[seq];
    q.closing = true;
    if (!q.busy && q.queue.empty()) {
        S_HIP(s, hipSetDevice(s->device));
        S_HIP(s, hipStreamSynchronize(s->stream));
        release_seq(s, q);
    }
    return PT_OK;
}

extern "C" int pt_server_set(pt_server *s, int key, int value)
{
    if (!s) return sfail(nullptr, PT_ERR_INVALID_ARG, "NULL server");
    switch (key) {
    case PT_BATCH_XCD_REMAP: s->xcd = value != 0; return PT_OK;
    case PT_SERVER_FILL_TILES:
        if (value < 0) return sfail(s, PT_ERR_INVALID_ARG, "fill threshold must be >= 0");
        s->fill_tiles = value;
        return PT_OK;
    case PT_BATCH_ISOLATE_TILES:
        if (value < 1) return sfail(s, PT_ERR_INVALID_ARG, "isolate threshold must be >= 1");
        s->pol.isolate_tiles = value;
        return PT_OK;
    case PT_BATCH_SHAPE: {
        int op = value / 256, cls = (value / 16) % 16, shape = value % 16;
        if (op < 0 || op > 1 || cls > 2 || (shape >= PT_SHAPE_COUNT && shape != 15))
            return sfail(s, PT_ERR_INVALID_ARG, "bad shape setting %d", value);
        s->pol.shape[op][cls] = shape == 15 ? -1 : shape;
        return PT_OK;
    }
    default: return sfail(s, PT_ERR_INVALID_ARG, "unknown server key %d", key);
    }
}

extern "C" int pt_server_get_stats(pt_server *s, pt_server_stats *o)
{
    if (!s || !o) return sfail(s, PT_ERR_INVALID_ARG, "NULL argument");
    memset(o, 0, sizeof(*o));
    o->submitted = s->submitted;
    o->completed = s->ncompleted;
    o->ticks = s->ticks;
    o->launches = s->launches;
    o->multi_launches = s->multi;
    for (auto &q : s->seqs) {
        o->resident_sequences += q.resident;
        o->waiting_frames += (int64_t)q.queue.size();
    }
    o->in_flight = (int64_t)s->flights.size();
    o->resident_bytes = s->resident_bytes;
    o->budget_bytes = s->budget;
    return PT_OK;
}
