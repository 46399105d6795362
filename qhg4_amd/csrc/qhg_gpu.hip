// Host side of the C ABI declared in include/qhg_gpu.h.
//
// A qhg_pop owns every device buffer of one population and one HIP stream
// (default priority).  The per-step entry points only enqueue kernels; the
// host learns the population size from a status block copied back
// asynchronously, and keeps the capacity ahead of the worst-case growth
// (births <= live/2 per step) without a per-step synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/qhg_gpu.h"
#include "aux_kernels.hpp"
#include "bino.hpp"
#include "step_kernels.hpp"

using namespace qhg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(QHG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum KernelId { KB_BUCKET, KB_PAIR, KB_ACT, KB_NEWBORN, KB_CELLSCAN, KB_HIST, KB_CAPACITY,
                KB_EVAL, KB_COMPACT, KB_COUNT };
const char *kKernelNames[KB_COUNT] = {"k_bucket", "k_pair", "k_act", "k_newborn", "k_cellscan",
                                      "k_hist", "k_capacity", "k_eval", "k_compact"};

constexpr int kRing = kStatusRing;
#ifndef QHG_ACT_GRID_PER_CU
#define QHG_ACT_GRID_PER_CU 8   // k_act blocks per compute unit
#endif

struct AgentBufs {
    AgentPtrs p{};
    int64_t cap = 0;
};

}  // namespace

struct qhg_pop {
    qhg_params P{};
    int32_t ncells = 0, conn = 0;
    int32_t gwords = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 256;

    // device buffers
    AgentBufs bufs[2];
    int cur = 0;
    int32_t *d_nbr = nullptr;
    uint8_t *d_nn = nullptr;
    double *d_K = nullptr, *d_W = nullptr, *d_tmp = nullptr;
    uint8_t *d_ice = nullptr;
    uint32_t *d_cnt[2] = {nullptr, nullptr};
    uint32_t *d_off = nullptr, *d_fillF = nullptr, *d_fillM = nullptr;
    uint64_t *d_bucket = nullptr;
    int32_t *d_brec = nullptr;
    int32_t *d_newidx = nullptr;
    uint32_t *d_tile_cnt = nullptr, *d_tile_off = nullptr, *d_super = nullptr, *d_blk_fert = nullptr;
    uint64_t *d_tiles_cell = nullptr;
    int64_t *d_hw_next = nullptr;
    DevStatus *d_st = nullptr;
    // env on device
    double *d_env[7] = {};   // alt lon lat water npp temp rain
    uint8_t *d_coastal = nullptr;
    std::vector<int64_t> cell_ids;

    Prog prog{};
    bool has_pair = false, has_me = false, has_cap = false;
    bool me_done = false;
    bool acted = false;       // k_act launched for the current step
    bool initialized = false; // initialize phase launched for the current step

    // asynchronous status readback ring
    DevStatus *h_ring = nullptr;     // pinned, kRing entries, written by k_cellscan
    DevStatus *d_ring = nullptr;     // device view of h_ring
    hipEvent_t ring_ev[kRing] = {};
    int64_t ring_step[kRing];
    int64_t steps_enqueued = 0;      // step index of the next step to enqueue
    DevStatus known{};               // newest status known on the host
    int64_t known_step = -1;         // status after step known_step-1 ... (= value of st.step)
    int64_t compactions = 0;
    int64_t last_compact_step = -1;

    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<std::pair<int, size_t>> ev_used;
    double t_sum[KB_COUNT] = {};
    int64_t t_cnt[KB_COUNT] = {};

    Ctx ctx() const {
        Ctx x{};
        x.a = bufs[cur].p;
        x.cap = bufs[cur].cap;
        x.gwords = gwords;
        x.ncells = ncells;
        x.conn = conn;
        x.nbr = d_nbr;
        x.nn = d_nn;
        x.K = d_K;
        x.W = d_W;
        x.ice = d_ice;
        x.cnt[0] = d_cnt[0];
        x.cnt[1] = d_cnt[1];
        x.off = d_off;
        x.fillF = d_fillF;
        x.fillM = d_fillM;
        x.bucket = d_bucket;
        x.brec = d_brec;
        x.tile_cnt = d_tile_cnt;
        x.super_cnt = d_super;
        x.blk_fert = d_blk_fert;
        x.tiles_cell = d_tiles_cell;
        x.ntiles_act_max = bufs[cur].cap / kActTile + 2;
        x.st = d_st;
        x.hring = d_ring;
        x.seed = P.seed;
        return x;
    }
};

namespace {

// k_cellscan runs one block per tile of kCellTile cells
int cell_tiles(const qhg_pop *p) { return (int)((p->ncells + kCellTile - 1) / kCellTile); }

int grid_for(int64_t n, int max_blocks) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (int)b;
}

template <typename... KArgs, typename... Args>
int launch_bs(qhg_pop *p, int kid, void (*kern)(KArgs...), int grid, int block, Args... args);

template <typename... KArgs, typename... Args>
int launch(qhg_pop *p, int kid, void (*kern)(KArgs...), int grid, Args... args) {
    return launch_bs(p, kid, kern, grid, kBlock, args...);
}

template <typename... KArgs, typename... Args>
int launch_bs(qhg_pop *p, int kid, void (*kern)(KArgs...), int grid, int block, Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (p->timing) {
        size_t slot = p->ev_used.size();
        if (slot >= p->ev_pool.size()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            p->ev_pool.push_back({a, b});
        }
        e0 = p->ev_pool[slot].first;
        e1 = p->ev_pool[slot].second;
        p->ev_used.push_back({kid, slot});
        HIPCHK(hipEventRecord(e0, p->stream));
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, p->stream, args...);
    HIPCHK(hipGetLastError());
    if (p->timing) HIPCHK(hipEventRecord(e1, p->stream));
    return 0;
}

int collect_timing(qhg_pop *p) {
    if (p->ev_used.empty()) return 0;
    HIPCHK(hipStreamSynchronize(p->stream));
    for (auto &u : p->ev_used) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p->ev_pool[u.second].first, p->ev_pool[u.second].second));
        p->t_sum[u.first] += ms;
        p->t_cnt[u.first] += 1;
    }
    p->ev_used.clear();
    return 0;
}

void free_bufs(AgentBufs &b) {
    AgentPtrs &a = b.p;
    void *ptrs[] = {a.id, a.cell, a.life, a.gender, a.birth, a.age, a.lastb, a.mate, a.nbabies, a.genome};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    b = AgentBufs{};
}

int alloc_bufs(AgentBufs &b, int64_t cap, int32_t gwords) {
    AgentPtrs &a = b.p;
    a = AgentPtrs{};
    const size_t n = (size_t)cap;
    HIPCHK(hipMalloc(&a.id, n * 8));
    HIPCHK(hipMalloc(&a.cell, n * 4));
    HIPCHK(hipMalloc(&a.life, n));
    HIPCHK(hipMalloc(&a.gender, n));
    HIPCHK(hipMalloc(&a.birth, n * 8));
    HIPCHK(hipMalloc(&a.age, n * 8));
    HIPCHK(hipMalloc(&a.lastb, n * 8));
    HIPCHK(hipMalloc(&a.mate, n * 4));
    HIPCHK(hipMalloc(&a.nbabies, n * 4));
    if (gwords) HIPCHK(hipMalloc(&a.genome, n * (size_t)gwords * 8));
    b.cap = cap;
    return 0;
}

// per-slot scratch that scales with capacity
int alloc_slot_scratch(qhg_pop *p, int64_t cap) {
    if (p->d_bucket) (void)hipFree(p->d_bucket);
    if (p->d_brec) (void)hipFree(p->d_brec);
    if (p->d_newidx) (void)hipFree(p->d_newidx);
    if (p->d_tile_cnt) (void)hipFree(p->d_tile_cnt);
    if (p->d_tile_off) (void)hipFree(p->d_tile_off);
    if (p->d_super) (void)hipFree(p->d_super);
    const size_t nt = (size_t)(cap / kActTile + 2);
    HIPCHK(hipMalloc(&p->d_bucket, (size_t)cap * 8));
    HIPCHK(hipMalloc(&p->d_brec, ((size_t)cap + kActTile) * 12));
    HIPCHK(hipMalloc(&p->d_newidx, (size_t)cap * 4));
    HIPCHK(hipMalloc(&p->d_tile_cnt, nt * 12));
    HIPCHK(hipMalloc(&p->d_tile_off, nt * 4));
    HIPCHK(hipMalloc(&p->d_super, (nt / kSuperTiles + 2) * 12));
    HIPCHK(hipMemsetAsync(p->d_super, 0, (nt / kSuperTiles + 2) * 12, p->stream));
    return 0;
}

template <typename T>
int upload_array(T **dst, const T *src, size_t n, hipStream_t s) {
    if (!src) { *dst = nullptr; return 0; }
    HIPCHK(hipMalloc(dst, n * sizeof(T)));
    HIPCHK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return 0;
}

bool has_kind(const qhg_params &P, int kind) {
    for (int i = 0; i < P.num_actions; i++)
        if (P.action_kind[i] == kind) return true;
    return false;
}

int build_prog(qhg_pop *p) {
    const qhg_params &P = p->P;
    if (P.num_actions < 0 || P.num_actions > QHG_MAX_ACTIONS) return fail(QHG_ERR_ARG, "num_actions out of range");
    std::vector<int> idx(P.num_actions);
    std::iota(idx.begin(), idx.end(), 0);
    // Prioritizer order: prio ascending, then action name
    // (QHG4/core/Prioritizer.cpp:19-30, QHG4/core/SPopulation.cpp:249-257)
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (P.action_prio[a] != P.action_prio[b]) return P.action_prio[a] < P.action_prio[b];
        return std::strncmp(P.action_name[a], P.action_name[b], QHG_NAME_LEN) < 0;
    });
    Prog &g = p->prog;
    g = Prog{};
    g.nops = P.num_actions;
    bool seen_getold = false;
    for (int q = 0; q < P.num_actions; q++) {
        const int k = P.action_kind[idx[q]];
        if (k < QHG_ACT_GETOLD || k > QHG_ACT_GENETICS) return fail(QHG_ERR_ARG, "unknown action kind");
        g.ops[q] = k;
        g.opword |= (uint64_t)k << (4 * q);
        if (k == QHG_ACT_GETOLD) seen_getold = true;
        if (k == QHG_ACT_FERTILITY && !seen_getold) g.load_age = 1;
    }
    g.fert_min = P.fert_min_age;
    g.fert_max = P.fert_max_age;
    g.fert_inter = P.fert_interbirth;
    g.b0 = P.verh_b0;
    g.d0 = P.verh_d0;
    g.theta = P.verh_theta;
    g.Kconst = P.verh_K;
    g.rmove_p = P.rmove_prob;
    g.wmove_p = P.wmove_prob;
    p->has_pair = has_kind(P, QHG_ACT_RANDOMPAIR);
    p->has_me = has_kind(P, QHG_ACT_MULTIEVAL);
    p->has_cap = has_kind(P, QHG_ACT_NPPCAP);
    g.has_pair = p->has_pair;
    for (int q = 0; q < g.nops; q++) {
        switch (g.ops[q]) {
        case QHG_ACT_GETOLD: g.need_birth = 1; break;
        case QHG_ACT_FERTILITY: g.need_gender = 1; g.need_lastb = 1; break;
        case QHG_ACT_VERHULST: g.need_gender = 1; g.need_mate = 1; g.need_count = 1; g.need_act0 = 1; break;
        case QHG_ACT_VERHULSTVARK:
            g.need_gender = 1; g.need_mate = 1; g.need_count = 1; g.need_K = 1; g.need_act0 = 1;
            break;
        case QHG_ACT_RANDOMMOVE: g.need_move = 1; g.need_act1 = 1; break;
        case QHG_ACT_WEIGHTEDMOVE: g.need_move = 1; g.need_act0 = 1; break;
        default: break;
        }
    }
    if (has_kind(P, QHG_ACT_GENETICS)) {
        if (P.gen_genome_size <= 0 || P.gen_genome_size > 64 * 32) return fail(QHG_ERR_ARG, "Genetics_genome_size out of range");
        if (P.gen_num_crossover < -1 || P.gen_num_crossover > kMaxCross) return fail(QHG_ERR_ARG, "Genetics_num_crossover out of range");
        if (!(P.gen_mutation_rate <= 1)) return fail(QHG_ERR_ARG, "Genetics_mutation_rate out of range");
        p->gwords = 2 * qhg_gpu_genome_words(P.gen_genome_size);
        g.gen_size = P.gen_genome_size;
        g.gen_cross = P.gen_num_crossover;
        if (P.gen_mutation_rate > 0) {
            int32_t n = 0;
            // BinomialDist::create(rate, NumParents*GenomeSize, 1e-6), Genetics.cpp:206-213
            if (qhg_gpu_binomial_table(P.gen_mutation_rate, 2 * P.gen_genome_size, 1e-6, g.bino, &n) != 0)
                return fail(QHG_ERR_ARG, "binomial table does not fit");
            g.nbino = n;
        }
    }
    if (p->has_me) {
        if (P.me_num < 1 || P.me_num > QHG_MAX_EVALS) return fail(QHG_ERR_ARG, "MultiEvaluator needs 1..4 evaluators");
        for (int e = 0; e < P.me_num; e++)
            if (P.me_pl_n[e] < 0 || P.me_pl_n[e] > QHG_MAX_PL_POINTS) return fail(QHG_ERR_ARG, "PolyLine size");
    }
    return 0;
}

int run_capacity(qhg_pop *p) {
    EnvPtrs e{p->d_env[0], p->d_env[1], p->d_env[2], p->d_env[3], p->d_env[4], p->d_env[5], p->d_env[6],
              p->d_ice, p->d_coastal};
    CapParams c{p->P.npp_water_factor, p->P.npp_coastal_factor, p->P.npp_coastal_min_lat,
                p->P.npp_coastal_max_lat, p->P.npp_min, p->P.npp_max, p->P.npp_k_min,
                p->P.npp_k_max, p->P.npp_efficiency};
    return launch(p, KB_CAPACITY, k_capacity, grid_for(p->ncells, 4096), p->ctx(), e, c);
}

int run_evaluators(qhg_pop *p) {
    const qhg_params &P = p->P;
    for (int e = 0; e < P.me_num; e++) {
        const double *in = nullptr;
        switch (P.me_input[e]) {
        case QHG_EV_ALTITUDE: in = p->d_env[0]; break;
        case QHG_EV_CAPACITY: in = p->d_K; break;
        case QHG_EV_BASE_NPP: in = p->d_env[4]; break;
        }
        if (!in) return fail(QHG_ERR_STATE, "evaluator input array missing");
        EvalParams ep{};
        ep.mode = P.me_mode;
        ep.cumulate = P.me_cumulate[e];
        ep.npts = P.me_pl_n[e];
        ep.first = (e == 0);
        ep.last = (e == P.me_num - 1);
        ep.weight = P.me_weight[e];
        for (int j = 0; j < ep.npts; j++) { ep.px[j] = P.me_pl_x[e][j]; ep.pv[j] = P.me_pl_v[e][j]; }
        int rc = launch(p, KB_EVAL, k_eval_values, grid_for(p->ncells, 4096), p->ctx(), in, p->d_tmp, ep);
        if (rc) return rc;
        rc = launch(p, KB_EVAL, k_eval_combine, grid_for(p->ncells, 4096), p->ctx(), (const double *)p->d_tmp, ep);
        if (rc) return rc;
    }
    return 0;
}

// counts of live agents per cell + offsets (pre-loop, upload, compaction)
int recount(qhg_pop *p) {
    Ctx x = p->ctx();
    HIPCHK(hipMemsetAsync(p->d_cnt[0], 0, (size_t)p->ncells * 4, p->stream));
    HIPCHK(hipMemsetAsync(p->d_cnt[1], 0, (size_t)p->ncells * 4, p->stream));
    const size_t ntc = (size_t)(p->ncells / kCellTile + 2);
    HIPCHK(hipMemsetAsync(p->d_tiles_cell, 0, ntc * 8, p->stream));
    int rc = launch(p, KB_HIST, k_hist, 2048, x);
    if (rc) return rc;
    return launch(p, KB_CELLSCAN, k_cellscan, cell_tiles(p), x, (int32_t)1);
}

int read_status_sync(qhg_pop *p, DevStatus *out) {
    HIPCHK(hipMemcpyAsync(out, p->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int compact_into(qhg_pop *p, int64_t new_cap) {
    // stable compaction into the other buffer set (reallocated when the
    // capacity changes)
    const int other = 1 - p->cur;
    if (p->bufs[other].cap != new_cap) {
        HIPCHK(hipStreamSynchronize(p->stream));
        free_bufs(p->bufs[other]);
        int rc = alloc_bufs(p->bufs[other], new_cap, p->gwords);
        if (rc) return rc;
    }
    if (new_cap > p->bufs[p->cur].cap) {
        HIPCHK(hipStreamSynchronize(p->stream));
        int32_t *old_newidx = p->d_newidx;
        p->d_newidx = nullptr;
        // scratch for the bigger capacity; newidx needs the old size
        int rc = alloc_slot_scratch(p, new_cap);
        if (rc) return rc;
        (void)old_newidx;
        if (old_newidx) (void)hipFree(old_newidx);
    }
    Ctx x = p->ctx();
    const int g = p->num_cu * 4;
    int rc = launch(p, KB_COMPACT, k_compact_count, g, x, p->d_tile_cnt);
    if (rc) return rc;
    rc = launch_bs(p, KB_COMPACT, k_scan_tiles, 1, 1024, (const uint32_t *)p->d_tile_cnt, p->d_tile_off,
                   (const DevStatus *)p->d_st, p->d_hw_next);
    if (rc) return rc;
    rc = launch(p, KB_COMPACT, k_compact_scatter, g, x, p->bufs[other].p, p->d_newidx,
                (const uint32_t *)p->d_tile_off);
    if (rc) return rc;
    rc = launch(p, KB_COMPACT, k_compact_mate, 2048, x, p->bufs[other].p, (const int32_t *)p->d_newidx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_commit_compact, dim3(1), dim3(64), 0, p->stream, p->d_st, (const int64_t *)p->d_hw_next);
    HIPCHK(hipGetLastError());
    p->cur = other;
    p->compactions++;
    // the status now has hw == n_alive; the host-side knowledge stays an
    // upper bound until the next readback
    return 0;
}

void poll_ring(qhg_pop *p) {
    // newest completed snapshot wins; its own step field says which step it
    // belongs to (a slot may already hold a later step's snapshot)
    for (int k = 0; k < kRing; k++) {
        if (p->ring_step[k] < 0) continue;
        if (hipEventQuery(p->ring_ev[k]) != hipSuccess) continue;
        DevStatus snap;
        std::memcpy(&snap, (const void *)&p->h_ring[k], sizeof(snap));
        if (snap.step > p->known_step) {
            p->known = snap;
            p->known_step = snap.step;
        }
        p->ring_step[k] = -1;
    }
}


// oldest step whose snapshot the host has not consumed (or -1)
int oldest_slot(const qhg_pop *p) {
    int best = -1;
    for (int k = 0; k < kRing; k++)
        if (p->ring_step[k] >= 0 && (best < 0 || p->ring_step[k] < p->ring_step[best])) best = k;
    return best;
}

double hw_bound(const DevStatus &k, int64_t lag) {
    // births per step <= live/2 (one per mated female), live grows <= x1.5
    double hw = (double)k.hw, alive = (double)k.n_alive;
    for (int64_t j = 0; j <= lag; j++) {
        hw += std::ceil(alive / 2) + 1;
        alive = alive * 1.5 + 1;
    }
    return hw;
}

// Keep hw + worst-case births of every step not yet observed below capacity
// (waiting for the oldest in-flight snapshot when the bound is too loose),
// grow the buffers when even the exact status needs it, and compact in
// stream order when holes dominate.
int ensure_capacity(qhg_pop *p) {
    poll_ring(p);
    const int64_t cap = p->bufs[p->cur].cap;
    int64_t lag = std::max<int64_t>(0, p->steps_enqueued - p->known_step);
    while ((hw_bound(p->known, lag) > (double)cap || lag > kRing - 2) && lag > 0) {
        const int k = oldest_slot(p);
        if (k < 0) {
            // nothing in flight: read the status directly
            DevStatus s;
            int rc = read_status_sync(p, &s);
            if (rc) return rc;
            p->known = s;
            p->known_step = s.step;
        } else {
            HIPCHK(hipEventSynchronize(p->ring_ev[k]));
            poll_ring(p);
        }
        lag = std::max<int64_t>(0, p->steps_enqueued - p->known_step);
    }
    if (p->known.error) return fail(QHG_ERR_CAPACITY, "device reported an error (flags 0x" +
                                    std::to_string(p->known.error) + ")");
    if (hw_bound(p->known, 0) > (double)cap) {
        // even the exact status does not leave room for one step
        const DevStatus &s = p->known;
        int64_t need = (int64_t)hw_bound(s, 4);
        int64_t nc = cap;
        if (2 * s.n_alive > cap || need > 2 * cap) {
            while ((double)nc < 1.25 * (double)need) nc *= 2;
        }
        int rc = compact_into(p, nc);
        if (rc) return rc;
        p->last_compact_step = p->steps_enqueued;
        DevStatus s2;
        rc = read_status_sync(p, &s2);
        if (rc) return rc;
        p->known = s2;
        p->known_step = s2.step;
        if (hw_bound(s2, 0) > (double)p->bufs[p->cur].cap) return fail(QHG_ERR_CAPACITY, "capacity growth failed");
        return 0;
    }
    // holes dominate: compact in stream order (no host sync); the known
    // high-water mark stays a valid upper bound until the next snapshot
    const DevStatus &k = p->known;
    if (k.hw - k.n_alive > std::max<int64_t>(k.n_alive, 1 << 16) && p->known_step > p->last_compact_step) {
        int rc = compact_into(p, cap);
        if (rc) return rc;
        p->last_compact_step = p->steps_enqueued;
    }
    return 0;
}

int enqueue_readback(qhg_pop *p) {
    int slot = (int)(p->steps_enqueued % kRing);
    if (p->ring_step[slot] >= 0) {
        HIPCHK(hipEventSynchronize(p->ring_ev[slot]));
        poll_ring(p);
    }
    HIPCHK(hipEventRecord(p->ring_ev[slot], p->stream));
    p->ring_step[slot] = p->steps_enqueued;
    return 0;
}

int phase_initialize(qhg_pop *p) {
    if (p->has_me && !p->me_done) {
        int rc = run_evaluators(p);
        if (rc) return rc;
        p->me_done = true;
    }
    Ctx x = p->ctx();
    if (p->has_pair) {
        int rc = launch(p, KB_BUCKET, k_bucket, kBucketBlocks, x);
        if (rc) return rc;
    }
    const int pgrid = (int)std::min<int64_t>((p->ncells + kPairBlock - 1) / kPairBlock, (int64_t)p->num_cu * 16);
    int rc = launch_bs(p, KB_PAIR, k_pair, std::max(pgrid, 1), kPairBlock, x, (int32_t)p->has_pair);
    if (rc) return rc;
    p->initialized = true;
    return 0;
}

int phase_act(qhg_pop *p) {
    if (!p->initialized) {
        int rc = phase_initialize(p);
        if (rc) return rc;
    }
    Ctx x = p->ctx();
    int rc = launch(p, KB_ACT, k_act, p->num_cu * QHG_ACT_GRID_PER_CU, x, p->prog);
    if (rc) return rc;
    p->acted = true;
    return 0;
}

int phase_finalize(qhg_pop *p) {
    if (!p->acted) {
        int rc = phase_act(p);
        if (rc) return rc;
    }
    Ctx x = p->ctx();
    int rc = launch(p, KB_NEWBORN, k_newborn, p->num_cu * 4, x, p->prog);
    if (rc) return rc;
    rc = launch(p, KB_CELLSCAN, k_cellscan, cell_tiles(p), x, (int32_t)0);
    if (rc) return rc;
    p->acted = false;
    p->initialized = false;
    p->steps_enqueued++;
    return enqueue_readback(p);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int qhg_gpu_abi_version(void) { return QHG_ABI_VERSION; }
const char *qhg_gpu_last_error(void) { return g_err.c_str(); }
int qhg_gpu_genome_words(int32_t genome_size) { return (genome_size + 31) / 32; }

int qhg_gpu_binomial_table(double p, int32_t n, double eps, double *cdf, int32_t *count) {
    int32_t c = bino_table(p, n, eps, cdf, QHG_MAX_BINO);
    if (c < 0) return fail(QHG_ERR_ARG, "binomial table: bad arguments or more than 64 bins");
    if (count) *count = c;
    return 0;
}

int qhg_gpu_create(const qhg_grid_desc *grid, const qhg_env_desc *env, const qhg_params *params,
                   int64_t capacity, qhg_pop **out) {
    if (!grid || !params || !out) return fail(QHG_ERR_ARG, "null argument");
    if (grid->num_cells <= 0 || grid->connectivity <= 0 || grid->connectivity > 6 || !grid->neighbors ||
        !grid->num_neighbors)
        return fail(QHG_ERR_ARG, "bad grid description");
    for (int64_t c = 0; c < (int64_t)grid->num_cells * grid->connectivity; c++)
        if (grid->neighbors[c] < -1 || grid->neighbors[c] >= grid->num_cells)
            return fail(QHG_ERR_ARG, "neighbour index out of range");
    for (int32_t c = 0; c < grid->num_cells; c++)
        if (grid->num_neighbors[c] > grid->connectivity) return fail(QHG_ERR_ARG, "num_neighbors > connectivity");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QHG_ERR_HIP, "no HIP device");
    qhg_pop *p = new qhg_pop();
    p->P = *params;
    p->ncells = grid->num_cells;
    p->conn = grid->connectivity;
    int rc = build_prog(p);
    if (rc) { delete p; return rc; }
    if (p->has_me || p->has_cap) {
        if (!env || !env->altitude) { delete p; return fail(QHG_ERR_ARG, "MultiEvaluator/NPPCapacity need altitude"); }
    }
    auto cleanup_fail = [&](int code) { qhg_gpu_destroy(p); return code; };
    if (hipGetDevice(&p->device) != hipSuccess) return cleanup_fail(fail(QHG_ERR_HIP, "hipGetDevice"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) == hipSuccess && prop.multiProcessorCount > 0)
        p->num_cu = prop.multiProcessorCount;
    {
        // the per-cell look-back scan needs all of its blocks resident at once
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cellscan, kBlock, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        if ((int64_t)cell_tiles(p) > (int64_t)per_cu * p->num_cu)
            return cleanup_fail(fail(QHG_ERR_ARG, "grid too large: more cell tiles than resident blocks"));
    }
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_HIP, "stream"));
    const size_t nc = (size_t)p->ncells;
    if (capacity < 1024) capacity = 1024;
    capacity = (capacity + kActTile - 1) / kActTile * kActTile;
#define CK(x) do { int r_ = (x); if (r_) return cleanup_fail(r_); } while (0)
    CK(alloc_bufs(p->bufs[0], capacity, p->gwords));
    CK(alloc_bufs(p->bufs[1], capacity, p->gwords));
    CK(alloc_slot_scratch(p, capacity));
    CK(upload_array(&p->d_nbr, grid->neighbors, nc * p->conn, p->stream));
    CK(upload_array(&p->d_nn, grid->num_neighbors, nc, p->stream));
    p->cell_ids.resize(nc);
    for (size_t c = 0; c < nc; c++) p->cell_ids[c] = grid->cell_ids ? grid->cell_ids[c] : (int64_t)c;
    if (hipMalloc(&p->d_K, nc * 8) != hipSuccess || hipMalloc(&p->d_W, nc * (p->conn + 1) * 8) != hipSuccess ||
        hipMalloc(&p->d_tmp, nc * 8) != hipSuccess || hipMalloc(&p->d_cnt[0], nc * 4) != hipSuccess ||
        hipMalloc(&p->d_cnt[1], nc * 4) != hipSuccess || hipMalloc(&p->d_off, (nc + 1) * 4) != hipSuccess ||
        hipMalloc(&p->d_fillF, nc * 4) != hipSuccess || hipMalloc(&p->d_fillM, nc * 4) != hipSuccess ||
        hipMalloc(&p->d_tiles_cell, (nc / kCellTile + 2) * 8) != hipSuccess ||
        hipMalloc(&p->d_hw_next, 8) != hipSuccess || hipMalloc(&p->d_st, sizeof(DevStatus)) != hipSuccess ||
        hipMalloc(&p->d_blk_fert, kBucketBlocks * 4) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_NOMEM, "device allocation failed"));
    (void)hipMemsetAsync(p->d_K, 0, nc * 8, p->stream);
    (void)hipMemsetAsync(p->d_W, 0, nc * (p->conn + 1) * 8, p->stream);
    (void)hipMemsetAsync(p->d_cnt[0], 0, nc * 4, p->stream);
    (void)hipMemsetAsync(p->d_cnt[1], 0, nc * 4, p->stream);
    (void)hipMemsetAsync(p->d_off, 0, (nc + 1) * 4, p->stream);
    (void)hipMemsetAsync(p->d_fillF, 0, nc * 4, p->stream);
    (void)hipMemsetAsync(p->d_fillM, 0, nc * 4, p->stream);
    (void)hipMemsetAsync(p->d_st, 0, sizeof(DevStatus), p->stream);
    (void)hipMemsetAsync(p->d_blk_fert, 0, kBucketBlocks * 4, p->stream);
    if (env) {
        const double *src[7] = {env->altitude, env->longitude, env->latitude, env->water,
                                env->base_npp, env->temperature, env->rainfall};
        for (int k = 0; k < 7; k++) CK(upload_array(&p->d_env[k], src[k], nc, p->stream));
        CK(upload_array(&p->d_ice, env->ice, nc, p->stream));
        CK(upload_array(&p->d_coastal, env->coastal, nc, p->stream));
    }
    if (hipHostMalloc((void **)&p->h_ring, sizeof(DevStatus) * kRing, hipHostMallocMapped) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_NOMEM, "pinned status ring"));
    std::memset(p->h_ring, 0, sizeof(DevStatus) * kRing);
    if (hipHostGetDevicePointer((void **)&p->d_ring, p->h_ring, 0) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_HIP, "status ring device pointer"));
    for (int k = 0; k < kRing; k++) {
        if (hipEventCreateWithFlags(&p->ring_ev[k], hipEventDisableTiming) != hipSuccess)
            return cleanup_fail(fail(QHG_ERR_HIP, "event"));
        p->ring_step[k] = -1;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) return cleanup_fail(fail(QHG_ERR_HIP, "create sync"));
#undef CK
    p->known = DevStatus{};
    p->known_step = 0;
    *out = p;
    return 0;
}

int qhg_gpu_destroy(qhg_pop *p) {
    if (!p) return 0;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    free_bufs(p->bufs[0]);
    free_bufs(p->bufs[1]);
    void *ptrs[] = {p->d_nbr, p->d_nn, p->d_K, p->d_W, p->d_tmp, p->d_ice, p->d_cnt[0], p->d_cnt[1],
                    p->d_off, p->d_fillF, p->d_fillM, p->d_bucket, p->d_brec, p->d_newidx,
                    p->d_tile_cnt, p->d_tile_off, p->d_super, p->d_blk_fert, p->d_tiles_cell, p->d_hw_next, p->d_st,
                    p->d_coastal};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (int k = 0; k < 7; k++)
        if (p->d_env[k]) (void)hipFree(p->d_env[k]);
    for (auto &e : p->ev_pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (int k = 0; k < kRing; k++)
        if (p->ring_ev[k]) (void)hipEventDestroy(p->ring_ev[k]);
    if (p->h_ring) (void)hipHostFree(p->h_ring);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int qhg_gpu_upload_agents(qhg_pop *p, const qhg_agents *a) {
    if (!p || !a || a->n < 0 || !a->id || !a->cell || !a->life || !a->gender)
        return fail(QHG_ERR_ARG, "upload: id, cell, life and gender are required");
    if (p->gwords && !a->genome) return fail(QHG_ERR_ARG, "upload: genome required with Genetics");
    std::vector<int64_t> order;
    order.reserve((size_t)a->n);
    for (int64_t i = 0; i < a->n; i++) {
        if (a->life[i] == 0) continue;
        if (a->cell[i] < 0 || a->cell[i] >= p->ncells) return fail(QHG_ERR_ARG, "upload: cell index out of range");
        if (a->id[i] < 0) return fail(QHG_ERR_ARG, "upload: negative agent id");
        order.push_back(i);
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return a->id[x] < a->id[y]; });
    for (size_t k = 1; k < order.size(); k++)
        if (a->id[order[k]] == a->id[order[k - 1]]) return fail(QHG_ERR_ARG, "upload: duplicate agent id");
    const int64_t n = (int64_t)order.size();
    int64_t cap = p->bufs[p->cur].cap;
    if (2 * n + 1024 > cap) {
        int64_t nc = cap;
        while (nc < 2 * n + 1024) nc *= 2;
        HIPCHK(hipStreamSynchronize(p->stream));
        free_bufs(p->bufs[0]);
        free_bufs(p->bufs[1]);
        int rc = alloc_bufs(p->bufs[0], nc, p->gwords);
        if (rc) return rc;
        rc = alloc_bufs(p->bufs[1], nc, p->gwords);
        if (rc) return rc;
        rc = alloc_slot_scratch(p, nc);
        if (rc) return rc;
        p->cur = 0;
        cap = nc;
    }
    std::vector<int64_t> id(n);
    std::vector<int32_t> cell(n), mate(n, -3), nbab(n, 0);
    std::vector<uint8_t> life(n), gender(n);
    std::vector<double> birth(n, 0.0), age(n, 0.0), lastb(n, 0.0);
    std::vector<uint64_t> genome((size_t)n * p->gwords);
    int64_t max_id = 0;
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = order[k];
        id[k] = a->id[i];
        cell[k] = a->cell[i];
        life[k] = a->life[i];
        gender[k] = a->gender[i];
        if (a->birth_time) birth[k] = a->birth_time[i];
        if (a->age) age[k] = a->age[i];
        if (a->last_birth) lastb[k] = a->last_birth[i];
        if (a->num_babies) nbab[k] = a->num_babies[i];
        if (p->gwords)
            std::memcpy(&genome[(size_t)k * p->gwords], a->genome + (size_t)i * p->gwords, (size_t)p->gwords * 8);
        max_id = std::max(max_id, id[k]);
    }
    const AgentPtrs &d = p->bufs[p->cur].p;
    hipStream_t s = p->stream;
    if (n > 0) {
    HIPCHK(hipMemcpyAsync(d.id, id.data(), n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.cell, cell.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.life, life.data(), n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.gender, gender.data(), n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.birth, birth.data(), n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.age, age.data(), n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.lastb, lastb.data(), n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.mate, mate.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.nbabies, nbab.data(), n * 4, hipMemcpyHostToDevice, s));
    if (p->gwords) HIPCHK(hipMemcpyAsync(d.genome, genome.data(), (size_t)n * p->gwords * 8, hipMemcpyHostToDevice, s));
    }
    DevStatus st;
    HIPCHK(hipMemcpyAsync(&st, p->d_st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    st.hw = n;
    st.n_alive = n;
    st.max_id = max_id;
    st.births_cur = st.deaths_cur = st.moves_cur = 0;
    st.tile_ctr_act = st.tile_ctr_cell = st.tile_ctr_cmp = 0;
    st.error = 0;
    HIPCHK(hipMemcpyAsync(p->d_st, &st, sizeof(st), hipMemcpyHostToDevice, s));
    int rc = recount(p);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    p->known = st;
    p->known_step = p->steps_enqueued;
    return 0;
}

int qhg_gpu_num_alive(qhg_pop *p, int64_t *n) {
    if (!p || !n) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    *n = s.n_alive;
    return 0;
}

int qhg_gpu_download_agents(qhg_pop *p, qhg_agents *a) {
    if (!p || !a) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    const int64_t hw = s.hw;
    const AgentPtrs &d = p->bufs[p->cur].p;
    std::vector<int64_t> id(hw);
    std::vector<int32_t> cell(hw), mate(hw), nbab(hw);
    std::vector<uint8_t> life(hw), gender(hw);
    std::vector<double> birth(hw), age(hw), lastb(hw);
    std::vector<uint64_t> genome;
    hipStream_t st = p->stream;
    if (hw > 0) {
        HIPCHK(hipMemcpyAsync(id.data(), d.id, hw * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(cell.data(), d.cell, hw * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(mate.data(), d.mate, hw * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(nbab.data(), d.nbabies, hw * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(life.data(), d.life, hw, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(gender.data(), d.gender, hw, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(birth.data(), d.birth, hw * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(age.data(), d.age, hw * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(lastb.data(), d.lastb, hw * 8, hipMemcpyDeviceToHost, st));
        if (p->gwords && a->genome) {
            genome.resize((size_t)hw * p->gwords);
            HIPCHK(hipMemcpyAsync(genome.data(), d.genome, (size_t)hw * p->gwords * 8, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    }
    int64_t k = 0;
    for (int64_t i = 0; i < hw; i++) {
        if (life[i] == 0) continue;
        if (a->id) a->id[k] = id[i];
        if (a->cell) a->cell[k] = cell[i];
        if (a->life) a->life[k] = life[i];
        if (a->gender) a->gender[k] = gender[i];
        if (a->birth_time) a->birth_time[k] = birth[i];
        if (a->age) a->age[k] = age[i];
        if (a->last_birth) a->last_birth[k] = lastb[i];
        if (a->num_babies) a->num_babies[k] = nbab[i];
        if (a->mate_id) {
            const int32_t m = mate[i];
            a->mate_id[k] = (m >= 0) ? ((m < hw && life[m] != 0) ? id[m] : -1) : m;
        }
        if (a->genome && p->gwords)
            std::memcpy(a->genome + (size_t)k * p->gwords, &genome[(size_t)i * p->gwords], (size_t)p->gwords * 8);
        k++;
    }
    a->n = k;
    return 0;
}

int qhg_gpu_pre_loop(qhg_pop *p, double t) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, p->stream, p->d_st, t);
    HIPCHK(hipGetLastError());
    if (p->has_cap) {
        int rc = run_capacity(p);
        if (rc) return rc;
    }
    int rc = recount(p);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_initialize_step(qhg_pop *p, double t) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    int rc = ensure_capacity(p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, p->stream, p->d_st, t);
    HIPCHK(hipGetLastError());
    return phase_initialize(p);
}

int qhg_gpu_do_actions(qhg_pop *p, uint32_t prio, double t) {
    (void)prio;
    (void)t;
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    // all priority levels run fused in the first call of a step
    if (p->acted) return 0;
    return phase_act(p);
}

int qhg_gpu_finalize_step(qhg_pop *p) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    int rc = phase_finalize(p);
    if (rc) return rc;
    return p->timing ? collect_timing(p) : 0;
}

int qhg_gpu_run(qhg_pop *p, int32_t nsteps) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    for (int32_t s = 0; s < nsteps; s++) {
        int rc = ensure_capacity(p);
        if (rc) return rc;
        rc = phase_initialize(p);
        if (rc) return rc;
        rc = phase_act(p);
        if (rc) return rc;
        rc = phase_finalize(p);
        if (rc) return rc;
        if (p->timing && p->ev_used.size() > 4096) {
            rc = collect_timing(p);
            if (rc) return rc;
        }
    }
    return 0;
}

int qhg_gpu_set_env(qhg_pop *p, const qhg_env_desc *env) {
    if (!p || !env) return fail(QHG_ERR_ARG, "null argument");
    const size_t nc = (size_t)p->ncells;
    HIPCHK(hipStreamSynchronize(p->stream));
    const double *src[7] = {env->altitude, env->longitude, env->latitude, env->water,
                            env->base_npp, env->temperature, env->rainfall};
    for (int k = 0; k < 7; k++) {
        if (!src[k]) continue;
        if (!p->d_env[k]) HIPCHK(hipMalloc(&p->d_env[k], nc * 8));
        HIPCHK(hipMemcpyAsync(p->d_env[k], src[k], nc * 8, hipMemcpyHostToDevice, p->stream));
    }
    if (env->ice) {
        if (!p->d_ice) HIPCHK(hipMalloc(&p->d_ice, nc));
        HIPCHK(hipMemcpyAsync(p->d_ice, env->ice, nc, hipMemcpyHostToDevice, p->stream));
    }
    if (env->coastal) {
        if (!p->d_coastal) HIPCHK(hipMalloc(&p->d_coastal, nc));
        HIPCHK(hipMemcpyAsync(p->d_coastal, env->coastal, nc, hipMemcpyHostToDevice, p->stream));
    }
    if (p->has_cap) {
        int rc = run_capacity(p);
        if (rc) return rc;
    }
    if (p->has_me) {
        int rc = run_evaluators(p);
        if (rc) return rc;
        p->me_done = true;
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_sync(qhg_pop *p) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_stats(qhg_pop *p, qhg_stats *o) {
    if (!p || !o) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    std::memset(o, 0, sizeof(*o));
    o->num_alive = s.n_alive;
    o->high_water = s.hw;
    o->capacity = p->bufs[p->cur].cap;
    o->max_id = s.max_id;
    o->step = s.step;
    o->births_last = s.births_last;
    o->deaths_last = s.deaths_last;
    o->moves_last = s.moves_last;
    o->births_total = s.births_total;
    o->deaths_total = s.deaths_total;
    o->moves_total = s.moves_total;
    o->compactions = p->compactions;
    o->device_error = s.error;
    o->agent_steps = s.agent_steps;
    o->slot_steps = s.slot_steps;
    o->fertile_total = s.fertile_total;
    return 0;
}

int qhg_gpu_cell_counts(qhg_pop *p, uint32_t *out) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    // counts of the step about to run = cnt[step & 1]
    HIPCHK(hipMemcpyAsync(out, p->d_cnt[s.step & 1], (size_t)p->ncells * 4, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_capacities(qhg_pop *p, double *out) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(out, p->d_K, (size_t)p->ncells * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_env_weights(qhg_pop *p, double *out) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(out, p->d_W, (size_t)p->ncells * (p->conn + 1) * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_compact(qhg_pop *p) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    int rc = compact_into(p, p->bufs[p->cur].cap);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_set_timing(qhg_pop *p, int32_t enable) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    int rc = collect_timing(p);
    if (rc) return rc;
    p->timing = enable != 0;
    for (int k = 0; k < KB_COUNT; k++) { p->t_sum[k] = 0; p->t_cnt[k] = 0; }
    return 0;
}

int qhg_gpu_kernel_times(qhg_pop *p, double *avg_ms, int64_t *launches, int32_t max, int32_t *count) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    int rc = collect_timing(p);
    if (rc) return rc;
    int n = std::min<int>(max, KB_COUNT);
    for (int k = 0; k < n; k++) {
        if (avg_ms) avg_ms[k] = p->t_cnt[k] ? p->t_sum[k] / (double)p->t_cnt[k] : 0.0;
        if (launches) launches[k] = p->t_cnt[k];
    }
    if (count) *count = n;
    return 0;
}

const char *qhg_gpu_kernel_name(int32_t index) {
    if (index < 0 || index >= KB_COUNT) return "";
    return kKernelNames[index];
}

}  // extern "C"
