// Host side of the C ABI declared in include/qhg_gpu.h.
//
// A qhg_pop owns every device buffer of one population and one HIP stream
// (default priority).  Agents live cell-sorted in one of two buffers that
// swap every step (the current buffer of step s is buf[s & 1]); the per-step
// entry points only enqueue kernels.  The host learns the population size
// from status snapshots written into pinned memory by the device, and keeps
// the capacity ahead of the worst-case growth (births <= live/2 per step)
// without a per-step synchronisation.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/qhg_gpu.h"
#include "aux_kernels.hpp"
#include "bino.hpp"
#include "cell_kernels.hpp"

using namespace qhg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

// the error hook of the other host sources of the library (config.cpp)
int qhg_internal_fail(int code, const std::string &msg) { return fail(code, msg); }

namespace {

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(QHG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum KernelId { KB_ACT, KB_SCAN, KB_SCATTER, KB_GENOME, KB_CAPACITY, KB_EVAL, KB_GPOOL, KB_ASSORT, KB_COUNT };
const char *kKernelNames[KB_COUNT] = {"k_cell_act", "k_cell_scan",     "k_cell_scatter", "k_genome",
                                      "k_capacity", "k_eval", "k_gpool_compact", "k_assort_pair"};

constexpr int kRing = kStatusRing;
#ifndef QHG_GROUP_AGENTS
#define QHG_GROUP_AGENTS 64.0     // target agents per wave group
#endif
#ifndef QHG_CELL_GRID_PER_CU
#define QHG_CELL_GRID_PER_CU 16   // k_cell_act / k_cell_scatter blocks per compute unit
#endif

void free_agentbuf(AgentBuf &a) {
    void *ptrs[] = {a.id, a.mate_id, a.life, a.gender, a.birth, a.age, a.lastb, a.nbabies, a.ghandle};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    a = AgentBuf{};
}

int alloc_agentbuf(AgentBuf &a, int64_t cap, bool genomes) {
    a = AgentBuf{};
    const size_t n = (size_t)cap;
    HIPCHK(hipMalloc(&a.id, n * 8));
    HIPCHK(hipMalloc(&a.mate_id, n * 8));
    HIPCHK(hipMalloc(&a.life, n));
    HIPCHK(hipMalloc(&a.gender, n));
    HIPCHK(hipMalloc(&a.birth, n * 8));
    HIPCHK(hipMalloc(&a.age, n * 8));
    HIPCHK(hipMalloc(&a.lastb, n * 8));
    HIPCHK(hipMalloc(&a.nbabies, n * 4));
    if (genomes) HIPCHK(hipMalloc(&a.ghandle, n * 4));
    return 0;
}

}  // namespace

struct qhg_pop {
    qhg_params P{};
    int32_t ncells = 0, conn = 0;
    int32_t gwords = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 256;
    int act_grid = 1, scatter_grid = 1, scan_grid = 1, ntiles = 1;
    // launch shapes (results do not depend on them)
    double group_agents = QHG_GROUP_AGENTS;
    int blocks_per_cu = QHG_CELL_GRID_PER_CU;

    // agents: two cell-sorted buffers of cap slots, segment offsets per buffer
    AgentBuf buf[2];
    int64_t cap = 0;
    uint32_t *d_off[2] = {nullptr, nullptr};
    OccCell *d_occ[2] = {nullptr, nullptr};   // occupied-cell lists per buffer
    // per-slot step scratch (cap entries)
    uint8_t *d_dec = nullptr, *d_gflag = nullptr;
    int32_t *d_bfather = nullptr, *d_brec = nullptr;
    int32_t *d_gfu = nullptr, *d_gmu = nullptr, *d_gfs = nullptr, *d_gms = nullptr, *d_gmate = nullptr;
    int64_t *d_gfid = nullptr, *d_gmid = nullptr;
    uint32_t *d_gprot = nullptr;   // AssortativePair: kProtBudget digests per slot
    int32_t *d_gnprot = nullptr;
    // genome pool (gcap handles), a second pool for compaction
    uint64_t *d_gpool[2] = {nullptr, nullptr};
    int gp_cur = 0;
    int64_t gcap = 0;
    // grid and per-cell tables
    int32_t *d_nbr = nullptr;
    uint8_t *d_nn = nullptr;
    int32_t *d_in_e = nullptr;
    int32_t in_w = 1;
    double *d_K = nullptr, *d_W = nullptr, *d_tmp = nullptr;
    uint8_t *d_ice = nullptr, *d_coastal = nullptr;
    double *d_env[7] = {};   // alt lon lat water npp temp rain
    uint32_t *d_cstay = nullptr, *d_cbirth = nullptr, *d_cin = nullptr, *d_carr = nullptr, *d_birth_off = nullptr;
    uint32_t *d_tact = nullptr;   // per scan tile
    uint32_t *d_tick = nullptr;   // the scan's tile tickets
    int32_t *d_outk = nullptr;
    uint32_t *d_bpos = nullptr, *d_rpos = nullptr;
    uint64_t *d_tiles = nullptr, *d_tiles2 = nullptr;
    uint32_t *d_wave_part = nullptr;
    uint64_t *d_scan_part = nullptr;   // 3 per k_cell_scan wave (grid <= ntiles blocks)
    unsigned long long *d_digest = nullptr;   // 2 sums of qhg_gpu_state_digest (allocated at its first call)
    DevStatus *d_st = nullptr;

    // Navigate: the navigation as set (ports in cell order, each port's
    // destinations in cell order), the altitudes the current bridges are
    // chosen by, and the device tables built from them (NavDev)
    bool nav_set = false, nav_preloop_done = false;
    std::vector<int32_t> nav_port, nav_doff, nav_dcell, nav_bridge;
    std::vector<double> nav_ddist, h_alt;
    int32_t nnav = 0;
    int32_t *d_nav_off = nullptr, *d_nav_njump = nullptr, *d_nav_dst = nullptr, *d_nav_bid = nullptr;
    int32_t *d_nav_in_off = nullptr, *d_nav_in_e = nullptr;
    double *d_nav_stay = nullptr, *d_nav_cum = nullptr;
    uint32_t *d_nav_cnt = nullptr, *d_nav_rpos = nullptr;
    uint64_t *d_nav_sel = nullptr;   // per slot (slot scratch)

    Prog prog{};
    bool has_pair = false, has_me = false, has_cap = false, has_assort = false;
    bool has_rare = false;   // OldAgeDeath or ATanDeath: k_cell_act<true, ...>
    bool has_nav = false;    // Navigate is scheduled
    bool has_prot = false;   // the assortative pairing counts protein matches
    bool me_done = false;
    bool acted = false;       // k_cell_act launched for the current step
    bool initialized = false; // initialize phase done for the current step
    double t_next = 0.0;      // time of the next step

    // asynchronous status readback ring
    DevStatus *h_ring = nullptr;     // pinned, kRing entries, written by k_cell_scatter
    DevStatus *d_ring = nullptr;     // device view of h_ring
    int64_t *h_seal = nullptr;       // kRing seal words after the snapshots: a sealed step
    int64_t ring_step[kRing];        // the step of each slot's snapshot not read yet (or -1)
    int64_t steps_enqueued = 0;      // index of the next step to enqueue
    DevStatus known{};               // newest status known on the host
    int64_t known_step = 0;          // its step field: the step it precedes
    int64_t pool_reset_step = -1;    // a pool compaction was enqueued before this step
    int64_t compactions = 0;

    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<std::pair<int, size_t>> ev_used;
    double t_sum[KB_COUNT] = {};
    int64_t t_cnt[KB_COUNT] = {};

    int cur() const { return (int)(steps_enqueued & 1); }

    // batched stepping: the batch this population belongs to, and the
    // stream it replaced while a member
    qhg_batch *batch = nullptr;
    hipStream_t own_stream = nullptr;
    int64_t gen = 0;          // bumped whenever a device buffer moves
    int32_t host_error = 0;   // code with which a batch step of this population failed (it then left its batch)

    Ctx ctx() const { return ctx_for(cur()); }

    Ctx ctx_for(int b) const {
        Ctx x{};
        x.cur = buf[b];
        x.nxt = buf[1 - b];
        x.off_cur = d_off[b];
        x.off_nxt = d_off[1 - b];
        x.occ_cur = d_occ[b];
        x.occ_nxt = d_occ[1 - b];
        x.cap = cap;
        x.ncells = ncells;
        x.conn = conn;
        x.nbr = d_nbr;
        x.nn = d_nn;
        x.K = d_K;
        x.W = d_W;
        x.ice = d_ice;
        x.in_e = d_in_e;
        x.in_w = in_w;
        x.dec = d_dec;
        x.bfather = d_bfather;
        x.brec = d_brec;
        x.gfu = d_gfu;
        x.gmu = d_gmu;
        x.gfs = d_gfs;
        x.gms = d_gms;
        x.gmate = d_gmate;
        x.gfid = d_gfid;
        x.gmid = d_gmid;
        x.gflag = d_gflag;
        x.gprot = d_gprot;
        x.gnprot = d_gnprot;
        x.cstay = d_cstay;
        x.cbirth = d_cbirth;
        x.cin = d_cin;
        x.carr = d_carr;
        x.tact = d_tact;
        x.tick = d_tick;
        x.outk = d_outk;
        x.birth_off = d_birth_off;
        x.bpos = d_bpos;
        x.rpos = d_rpos;
        x.tiles = d_tiles;
        x.tiles2 = d_tiles2;
        x.ntiles = ntiles;
        x.nact_blocks = act_grid;
        x.wave_part = d_wave_part;
        x.scan_part = d_scan_part;
        x.scan_waves = scan_grid * (kScanBlock / kWave);
        x.gpool = d_gpool[gp_cur];
        x.gwords = gwords;
        x.gcap = gcap;
        x.st = d_st;
        x.hring = d_ring;
        x.seed = P.seed;
        x.nav.n = nnav;
        x.nav.off = d_nav_off;
        x.nav.njump = d_nav_njump;
        x.nav.stay = d_nav_stay;
        x.nav.dst = d_nav_dst;
        x.nav.cum = d_nav_cum;
        x.nav.bid = d_nav_bid;
        x.nav.in_off = d_nav_in_off;
        x.nav.in_e = d_nav_in_e;
        x.nav.cnt = d_nav_cnt;
        x.nav.rpos = d_nav_rpos;
        x.nav.sel = d_nav_sel;
        return x;
    }

    StepArgs step_args() const {
        StepArgs a{};
        a.step = (uint32_t)steps_enqueued;
        a.parity = (uint32_t)cur();
        a.t = t_next;
        // cells per wave group from the newest known mean occupancy: about
        // one wave of agents per group (results do not depend on it)
        const int kb = (int)(known_step & 1);
        a.gcells = group_cells(group_agents, known.n_alive[kb], known.n_occ[kb]);
        a.dense = scan_dense(known.n_occ[kb], ncells, 3);
        return a;
    }
};

namespace {

constexpr int kMaxBlocksPerCu = 32;

void set_shape(qhg_pop *p) {
    p->act_grid = (int)std::max<int64_t>(1, std::min<int64_t>(p->ncells, (int64_t)p->num_cu * p->blocks_per_cu));
    p->scatter_grid = p->act_grid;
}

// device memory held by a population: agent buffers and slot scratch (cap
// slots), genome pools, per-cell tables, fixed-size scratch
int64_t device_bytes(const qhg_pop *p) {
    const int64_t g = p->gwords ? 1 : 0;
    const int64_t per_slot = 2 * (46 + 4 * g) + 38 + 12 * g + (p->has_prot ? 16 * kProtBudget + 4 : 0) +
                             (p->has_nav ? 8 : 0);
    int64_t b = p->cap * per_slot;
    for (int k = 0; k < 2; k++)
        if (p->d_gpool[k]) b += p->gcap * p->gwords * 8;
    const int64_t nc = p->ncells, cn = p->conn;
    int64_t per_cell = cn * 4 + 1 + p->in_w * 4 + cn * 4 + 8 + (cn + 1) * 8 + 8 + 2 * 4 + 4 + 4 + p->in_w * 4 + 4 + 4 + 4 +
                       cn * 4 + 2 * (int64_t)sizeof(OccCell);
    for (int k = 0; k < 7; k++)
        if (p->d_env[k]) per_cell += 8;
    if (p->d_ice) per_cell += 1;
    if (p->d_coastal) per_cell += 1;
    b += nc * per_cell + 3 * 4;
    if (p->has_nav) b += nc * (4 + 4 + 8 + 4) + 8 + (int64_t)p->nnav * (4 + 8 + 4 + 4 + 4 + 4);
    b += 2 * (int64_t)p->ntiles * 8 + (int64_t)p->ntiles * 4 + (int64_t)p->num_cu * kMaxBlocksPerCu * kParts * 4 +
         (int64_t)p->ntiles * (kScanBlock / kWave) * 3 * 8 +
         (int64_t)sizeof(DevStatus);
    return b;
}

int check_tuning(const qhg_tuning *t) {
    if (!(t->group_agents >= 1.0 && t->group_agents <= 4096.0))
        return fail(QHG_ERR_ARG, "group_agents must be in [1, 4096]");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > kMaxBlocksPerCu)
        return fail(QHG_ERR_ARG, "blocks_per_cu must be in [1, 32]");
    return 0;
}

struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int init() {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        return 0;
    }
    ~Timer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Launch-shape search on the live workload: every trial runs trial_steps
// real steps (results do not depend on the shape) and is scored in agent-
// steps per ms of device time.  Coordinate search: the group target at the
// starting blocks per CU, then blocks per CU at the best group target.
template <typename Apply, typename Trial>
int tune_search(const qhg_tuning &start, int32_t trial_steps, Apply apply, Trial trial, qhg_tuning *best,
                int32_t *taken) {
    static const double groups[] = {32.0, 48.0, 64.0, 96.0, 128.0, 192.0};
    static const int blocks[] = {4, 8, 16, 32};
    qhg_tuning cur = start, win = start;
    double win_rate = -1.0;
    *taken = 0;
    auto score = [&](const qhg_tuning &t) -> int {
        int rc = apply(t);
        if (rc) return rc;
        double r = 0.0;
        rc = trial(&r);
        if (rc) return rc;
        *taken += trial_steps;
        if (r > win_rate) { win_rate = r; win = t; }
        return 0;
    };
    for (double g : groups) {
        cur.group_agents = g;
        int rc = score(cur);
        if (rc) return rc;
    }
    cur = win;
    for (int bpc : blocks) {
        if (bpc == start.blocks_per_cu) continue;
        cur.blocks_per_cu = bpc;
        int rc = score(cur);
        if (rc) return rc;
    }
    *best = win;
    return apply(win);
}

int grid_for(int64_t n, int max_blocks) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (int)b;
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
        // events attached to the dispatch itself: they time the kernel, not
        // the gap before it
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, p->stream, e0, e1, 0, args...);
        HIPCHK(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, p->stream, args...);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename... KArgs, typename... Args>
int launch(qhg_pop *p, int kid, void (*kern)(KArgs...), int grid, Args... args) {
    return launch_bs(p, kid, kern, grid, kBlock, args...);
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

void free_slot_scratch(qhg_pop *p) {
    void *ptrs[] = {p->d_dec, p->d_gflag, p->d_bfather, p->d_brec, p->d_gfu, p->d_gmu, p->d_gfs, p->d_gms,
                    p->d_gmate, p->d_gfid, p->d_gmid, p->d_gprot, p->d_gnprot, p->d_nav_sel};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p->d_nav_sel = nullptr;
    p->d_dec = p->d_gflag = nullptr;
    p->d_bfather = p->d_brec = p->d_gfu = p->d_gmu = p->d_gfs = p->d_gms = p->d_gmate = nullptr;
    p->d_gfid = p->d_gmid = nullptr;
    p->d_gprot = nullptr;
    p->d_gnprot = nullptr;
}

// per-slot scratch: outcome bytes, birth records, and the list scratch of
// cells too large for LDS (indexed by segment slot, so cells never overlap)
int alloc_slot_scratch(qhg_pop *p, int64_t cap) {
    free_slot_scratch(p);
    const size_t n = (size_t)cap;
    HIPCHK(hipMalloc(&p->d_dec, n));
    HIPCHK(hipMalloc(&p->d_gflag, n));
    HIPCHK(hipMalloc(&p->d_gfu, n * 4));
    HIPCHK(hipMalloc(&p->d_gmu, n * 4));
    HIPCHK(hipMalloc(&p->d_gfs, n * 4));
    HIPCHK(hipMalloc(&p->d_gms, n * 4));
    HIPCHK(hipMalloc(&p->d_gmate, n * 4));
    HIPCHK(hipMalloc(&p->d_gfid, n * 8));
    HIPCHK(hipMalloc(&p->d_gmid, n * 8));
    if (p->gwords) {
        HIPCHK(hipMalloc(&p->d_bfather, n * 4));
        HIPCHK(hipMalloc(&p->d_brec, n * 8));
    }
    if (p->has_prot) {
        HIPCHK(hipMalloc(&p->d_gprot, n * 16 * kProtBudget));
        HIPCHK(hipMalloc(&p->d_gnprot, n * 4));
    }
    if (p->has_nav) HIPCHK(hipMalloc(&p->d_nav_sel, n * 8));
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
        if (k < QHG_ACT_GETOLD || k > QHG_ACT_LAST) return fail(QHG_ERR_ARG, "unknown action kind");
        g.ops[q] = k;
        g.opword |= (uint64_t)k << (4 * q);
        // actions that set the age before Fertility reads it
        if (k == QHG_ACT_GETOLD || k == QHG_ACT_OLDAGEDEATH || k == QHG_ACT_ATANDEATH) seen_getold = true;
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
    g.oad_max = P.oad_max_age;
    g.oad_unc = P.oad_uncertainty;
    g.atan_max = P.atan_max_age;
    g.atan_slope = P.atan_slope;
    if (has_kind(P, QHG_ACT_ATANDEATH)) {
        // ATanDeath::preLoop (QHG4/actions/ATanDeath.cpp:50-51); EPS = 0.001 (ATanDeath.h:14)
        const double q = std::atan(P.atan_slope * P.atan_range);
        if (!(q != 0.0)) return fail(QHG_ERR_ARG, "ATanDeath_slope * ATanDeath_range must not be 0");
        g.atan_scale = (kPi / 2 - 0.001) / q;
    }
    p->has_pair = has_kind(P, QHG_ACT_RANDOMPAIR);
    p->has_prot = has_kind(P, QHG_ACT_ASSORTPROT);
    p->has_assort = has_kind(P, QHG_ACT_ASSORTPAIR) || p->has_prot;
    if ((p->has_pair && p->has_assort) || (p->has_prot && has_kind(P, QHG_ACT_ASSORTPAIR)))
        return fail(QHG_ERR_ARG, "RandomPair, AssortativePairHamming and AssortativePair each pair on their own");
    if (p->has_assort) {
        if (!has_kind(P, QHG_ACT_GENETICS))
            return fail(QHG_ERR_ARG, "AssortativePairHamming and AssortativePair need Genetics");
        // AssortativePairHam::preLoop / findMates (AssortativePairHam.cpp:79-90,
        // :199): calcDist gets the block count where it reads a genome size
        const int32_t blocks = qhg_gpu_genome_words(P.gen_genome_size);
        g.assort_words = qhg_gpu_genome_words(blocks);
        g.assort_cutoff = (float)P.assort_cutoff;
        g.assort_permute = P.assort_permute != 0;
        g.assort_prot = p->has_prot ? 1 : 0;
        g.has_assort = 1;
        g.need_gender = 1;   // k_cell_act counts the FERTILE agents by sex
    }
    p->has_me = has_kind(P, QHG_ACT_MULTIEVAL);
    p->has_cap = has_kind(P, QHG_ACT_NPPCAP);
    p->has_nav = has_kind(P, QHG_ACT_NAVIGATE);
    p->has_rare = has_kind(P, QHG_ACT_OLDAGEDEATH) || has_kind(P, QHG_ACT_ATANDEATH);
    g.nav_bridge_p = P.nav_bridge_prob;
    g.has_pair = p->has_pair;
    for (int q = 0; q < g.nops; q++) {
        switch (g.ops[q]) {
        case QHG_ACT_GETOLD: g.need_birth = 1; break;
        case QHG_ACT_FERTILITY: g.need_gender = 1; g.need_lastb = 1; break;
        case QHG_ACT_VERHULST:
        case QHG_ACT_VERHULSTVARK: g.need_gender = 1; g.need_act0 = 1; break;
        case QHG_ACT_RANDOMMOVE: g.need_act1 = 1; break;
        case QHG_ACT_WEIGHTEDMOVE: g.need_act0 = 1; break;
        case QHG_ACT_OLDAGEDEATH:
        case QHG_ACT_ATANDEATH: g.need_birth = 1; g.need_act1 = 1; break;
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

void free_nav_tables(qhg_pop *p) {
    void *ptrs[] = {p->d_nav_off, p->d_nav_njump, p->d_nav_dst, p->d_nav_bid, p->d_nav_in_off, p->d_nav_in_e,
                    p->d_nav_stay, p->d_nav_cum, p->d_nav_cnt, p->d_nav_rpos};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    p->d_nav_off = p->d_nav_njump = p->d_nav_dst = p->d_nav_bid = p->d_nav_in_off = p->d_nav_in_e = nullptr;
    p->d_nav_stay = p->d_nav_cum = nullptr;
    p->d_nav_cnt = p->d_nav_rpos = nullptr;
    p->nnav = 0;
}

// Navigate::recalculate (QHG4/actions/Navigate.cpp:95-148) into the device
// tables.  A = prob0 / exp(decay * dist0); per port in cell order the
// destinations' A * exp(decay * dist), summed in destination order; a sum
// below 1 gives the cumulative table [1 - sum, ...], a sum of 1 or more
// stops the build (ports after it get none).  The current bridges are those
// with both ends above sea level.  Each cell's edges: its jumps, then one
// per current bridge end at the cell (a bridge from a cell to itself is one
// edge).  Called between steps with the stream idle.
int nav_rebuild(qhg_pop *p) {
    const qhg_params &P = p->P;
    const int32_t nc = p->ncells;
    const double A = P.nav_prob0 / std::exp(P.nav_decay * P.nav_dist0);
    std::vector<int32_t> njump((size_t)nc, 0), nbr_ends((size_t)nc, 0);
    std::vector<double> stay((size_t)nc, 0.0);
    std::vector<int32_t> tab_of((size_t)nc, -1);   // port index with a table
    std::vector<double> cum(p->nav_dcell.size());
    for (size_t k = 0; k < p->nav_port.size(); k++) {
        const int32_t d0 = p->nav_doff[k], n = p->nav_doff[k + 1] - d0;
        double sum = 0;
        for (int32_t j = 0; j < n; j++) {
            const double pr = A * std::exp(P.nav_decay * p->nav_ddist[(size_t)(d0 + j)]);
            sum += pr;
            cum[(size_t)(d0 + j)] = pr;
        }
        if (!(sum < 1)) {
            std::fprintf(stderr, "[Navigate] probabilities for port [%d] add up to %f\n", p->nav_port[k], sum);
            break;
        }
        const int32_t c = p->nav_port[k];
        stay[(size_t)c] = 1 - sum;
        double run = stay[(size_t)c];
        for (int32_t j = 0; j < n; j++) {
            run = cum[(size_t)(d0 + j)] + run;
            cum[(size_t)(d0 + j)] = run;
        }
        njump[(size_t)c] = n;
        tab_of[(size_t)c] = (int32_t)k;
    }
    std::vector<int32_t> cur;   // current bridges, pairs
    for (size_t b = 0; b + 1 < p->nav_bridge.size(); b += 2) {
        const int32_t a0 = p->nav_bridge[b], a1 = p->nav_bridge[b + 1];
        if (!p->h_alt.empty() && p->h_alt[(size_t)a0] > 0 && p->h_alt[(size_t)a1] > 0) {
            cur.push_back(a0);
            cur.push_back(a1);
        }
    }
    for (size_t b = 0; b < cur.size(); b += 2) {
        nbr_ends[(size_t)cur[b]]++;
        if (cur[b + 1] != cur[b]) nbr_ends[(size_t)cur[b + 1]]++;
    }
    std::vector<int32_t> off((size_t)nc + 1, 0);
    for (int32_t c = 0; c < nc; c++) off[(size_t)c + 1] = off[(size_t)c] + njump[(size_t)c] + nbr_ends[(size_t)c];
    const int32_t E = off[(size_t)nc];
    std::vector<int32_t> dst((size_t)E + 1, 0), bid((size_t)E + 1, -1), fill(off.begin(), off.end() - 1);
    std::vector<double> ecum((size_t)E + 1, 0.0);
    for (int32_t c = 0; c < nc; c++) {
        const int32_t k = tab_of[(size_t)c];
        if (k < 0) continue;
        const int32_t d0 = p->nav_doff[(size_t)k];
        for (int32_t j = 0; j < njump[(size_t)c]; j++) {
            const int32_t e = fill[(size_t)c]++;
            dst[(size_t)e] = p->nav_dcell[(size_t)(d0 + j)];
            ecum[(size_t)e] = cum[(size_t)(d0 + j)];
        }
    }
    for (size_t b = 0; b < cur.size(); b += 2) {
        const int32_t a0 = cur[b], a1 = cur[b + 1];
        int32_t e = fill[(size_t)a0]++;
        dst[(size_t)e] = a1;
        bid[(size_t)e] = (int32_t)(b / 2);
        if (a1 != a0) {
            e = fill[(size_t)a1]++;
            dst[(size_t)e] = a0;
            bid[(size_t)e] = (int32_t)(b / 2);
        }
    }
    // edges into each cell, in edge order
    std::vector<int32_t> in_off((size_t)nc + 1, 0), in_e((size_t)E + 1, 0);
    for (int32_t e = 0; e < E; e++) in_off[(size_t)dst[(size_t)e] + 1]++;
    for (int32_t c = 0; c < nc; c++) in_off[(size_t)c + 1] += in_off[(size_t)c];
    {
        std::vector<int32_t> f(in_off.begin(), in_off.end() - 1);
        for (int32_t e = 0; e < E; e++) in_e[(size_t)f[(size_t)dst[(size_t)e]]++] = e;
    }
    free_nav_tables(p);
    const size_t n1 = (size_t)nc + 1, ne = (size_t)E + 1;
    if (hipMalloc(&p->d_nav_off, n1 * 4) != hipSuccess || hipMalloc(&p->d_nav_njump, (size_t)nc * 4) != hipSuccess ||
        hipMalloc(&p->d_nav_stay, (size_t)nc * 8) != hipSuccess || hipMalloc(&p->d_nav_in_off, n1 * 4) != hipSuccess ||
        hipMalloc(&p->d_nav_dst, ne * 4) != hipSuccess || hipMalloc(&p->d_nav_bid, ne * 4) != hipSuccess ||
        hipMalloc(&p->d_nav_cum, ne * 8) != hipSuccess || hipMalloc(&p->d_nav_in_e, ne * 4) != hipSuccess ||
        hipMalloc(&p->d_nav_cnt, ne * 4) != hipSuccess || hipMalloc(&p->d_nav_rpos, ne * 4) != hipSuccess) {
        free_nav_tables(p);
        return fail(QHG_ERR_NOMEM, "Navigate tables");
    }
    HIPCHK(hipMemcpyAsync(p->d_nav_off, off.data(), n1 * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_njump, njump.data(), (size_t)nc * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_stay, stay.data(), (size_t)nc * 8, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_in_off, in_off.data(), n1 * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_dst, dst.data(), ne * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_bid, bid.data(), ne * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_cum, ecum.data(), ne * 8, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_nav_in_e, in_e.data(), ne * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemsetAsync(p->d_nav_cnt, 0, ne * 4, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->nnav = E;
    p->gen++;
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

int read_status_sync(qhg_pop *p, DevStatus *out) {
    HIPCHK(hipMemcpyAsync(out, p->d_st, sizeof(DevStatus), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int refresh_known(qhg_pop *p) {
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    p->known = s;
    p->known_step = p->steps_enqueued;
    return 0;
}

// Genome pool compaction in stream order: the genomes of the current
// buffer's agents move to handles 0..n-1 of the other pool (reallocated to
// new_gcap first when that differs).
int compact_pool(qhg_pop *p, int64_t new_gcap) {
    if (!p->gwords) return 0;
    const int other = 1 - p->gp_cur;
    if (new_gcap != p->gcap || !p->d_gpool[other]) {
        HIPCHK(hipStreamSynchronize(p->stream));
        if (p->d_gpool[other]) (void)hipFree(p->d_gpool[other]);
        p->d_gpool[other] = nullptr;
        HIPCHK(hipMalloc(&p->d_gpool[other], (size_t)new_gcap * p->gwords * 8));
    }
    const int32_t par = p->cur();
    int rc = launch(p, KB_GPOOL, k_gpool_compact, p->num_cu * 8, p->ctx(), p->d_gpool[other], par);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gpool_commit, dim3(p->num_cu * 4), dim3(256), 0, p->stream, p->ctx(), par);
    HIPCHK(hipGetLastError());
    if (new_gcap != p->gcap) {
        // the old pool is no longer referenced once the copy has run
        HIPCHK(hipStreamSynchronize(p->stream));
        (void)hipFree(p->d_gpool[p->gp_cur]);
        p->d_gpool[p->gp_cur] = nullptr;
        p->gcap = new_gcap;
    }
    p->gp_cur = other;
    p->gen++;
    p->pool_reset_step = p->steps_enqueued;
    p->compactions++;
    return 0;
}

// Grow both agent buffers (and the per-slot scratch) to new_cap: the
// current buffer's n live agents are copied over.  Needs an exact status.
int grow_agents(qhg_pop *p, int64_t new_cap, int64_t n) {
    HIPCHK(hipStreamSynchronize(p->stream));
    const int b = p->cur();
    AgentBuf nb[2];
    for (int k = 0; k < 2; k++) {
        int rc = alloc_agentbuf(nb[k], new_cap, p->gwords != 0);
        if (rc) { free_agentbuf(nb[0]); free_agentbuf(nb[1]); return rc; }
    }
    const AgentBuf &o = p->buf[b];
    hipStream_t s = p->stream;
    const size_t m = (size_t)n;
    if (m > 0) {
        HIPCHK(hipMemcpyAsync(nb[b].id, o.id, m * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].mate_id, o.mate_id, m * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].life, o.life, m, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].gender, o.gender, m, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].birth, o.birth, m * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].age, o.age, m * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].lastb, o.lastb, m * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nb[b].nbabies, o.nbabies, m * 4, hipMemcpyDeviceToDevice, s));
        if (p->gwords) HIPCHK(hipMemcpyAsync(nb[b].ghandle, o.ghandle, m * 4, hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    free_agentbuf(p->buf[0]);
    free_agentbuf(p->buf[1]);
    p->buf[0] = nb[0];
    p->buf[1] = nb[1];
    p->cap = new_cap;
    p->gen++;
    return alloc_slot_scratch(p, new_cap);
}

int64_t seal_of(const qhg_pop *p, int k) { return __atomic_load_n(p->h_seal + k, __ATOMIC_ACQUIRE); }

void clear_seals(qhg_pop *p) {
    for (int k = 0; k < kRing; k++) __atomic_store_n(p->h_seal + k, (int64_t)-1, __ATOMIC_RELAXED);
}

void poll_ring(qhg_pop *p) {
    // newest completed snapshot wins; its own step field says which step it
    // precedes.  A slot is read once the step's scatter (k_cell_scatter, or
    // k_batch_scatter for a batch member) has sealed it with its step.
    for (int k = 0; k < kRing; k++) {
        if (p->ring_step[k] < 0) continue;
        if (seal_of(p, k) != p->ring_step[k]) continue;
        DevStatus snap;
        std::memcpy(&snap, (const void *)&p->h_ring[k], sizeof(snap));
        if (snap.step > p->known_step) {
            p->known = snap;
            p->known_step = snap.step;
        }
        p->ring_step[k] = -1;
    }
    // older snapshots not read yet are superseded; their slots are free
    // before the step that rewrites them is enqueued
    for (int k = 0; k < kRing; k++)
        if (p->ring_step[k] >= 0 && p->ring_step[k] <= p->known_step) p->ring_step[k] = -1;
}

// Wait until slot k is sealed.  A step that finds the device error flag set
// writes no snapshot (a batch member's step no longer advances): once the
// stream has drained, the status itself is read.  The stream is asked only every 2 ms of waiting: each query puts a
// marker into the stream, which holds the GPU idle like an event does.
int wait_seal(qhg_pop *p, int k) {
    using clk = std::chrono::steady_clock;
    auto next_query = clk::now() + std::chrono::milliseconds(2);
    for (uint32_t spin = 1;; spin++) {
        if (seal_of(p, k) == p->ring_step[k]) return 0;
        if (spin % 1024 == 0 && clk::now() >= next_query) {
            const hipError_t e = hipStreamQuery(p->stream);
            if (e == hipSuccess) {
                if (seal_of(p, k) == p->ring_step[k]) return 0;
                for (int j = 0; j < kRing; j++) p->ring_step[j] = -1;
                return refresh_known(p);
            }
            if (e != hipErrorNotReady) return fail(QHG_ERR_HIP, std::string("status wait: ") + hipGetErrorString(e));
            next_query = clk::now() + std::chrono::milliseconds(2);
        }
        __builtin_ia32_pause();
    }
}

// oldest step whose snapshot the host has not consumed (or -1)
int oldest_slot(const qhg_pop *p) {
    int best = -1;
    for (int k = 0; k < kRing; k++)
        if (p->ring_step[k] >= 0 && (best < 0 || p->ring_step[k] < p->ring_step[best])) best = k;
    return best;
}

// Upper bounds on the live count and the genome pool high-water mark at the
// start of step steps_enqueued + extra, from the newest known status: every
// step at most adds ceil(live/2) births (one per paired female).
struct Bounds { double live, pool; };
Bounds bounds(const qhg_pop *p, int extra) {
    const int kb = (int)(p->known_step & 1);
    double live = (double)p->known.n_alive[kb], pool = (double)p->known.gpool[kb];
    for (int64_t s = p->known_step; s < p->steps_enqueued + extra; s++) {
        if (s == p->pool_reset_step) pool = live;
        const double b = std::ceil(live / 2);
        live += b;
        pool += b;
    }
    if (p->steps_enqueued + extra == p->pool_reset_step) pool = live;
    return {live, pool};
}

// Before enqueuing a step: make room for its worst case.  A genome pool
// that may overflow is compacted in stream order first (afterwards it holds
// exactly the live agents); while the bounds are too loose the host waits
// for the oldest in-flight snapshot; with the exact status the buffers grow
// when even that leaves no room for one step.
int ensure_capacity(qhg_pop *p) {
    poll_ring(p);
    while (true) {
        const int64_t lag = p->steps_enqueued - p->known_step;
        const Bounds nx = bounds(p, 1);
        const bool live_fits = nx.live <= (double)p->cap;
        bool pool_fits = !p->gwords || nx.pool <= (double)p->gcap;
        if (!pool_fits && nx.live <= (double)p->gcap && p->pool_reset_step != p->steps_enqueued) {
            int rc = compact_pool(p, p->gcap);
            if (rc) return rc;
            continue;
        }
        if ((live_fits && pool_fits && lag <= kRing - 2) || lag <= 0) break;
        const int k = oldest_slot(p);
        if (k < 0) {
            int rc = refresh_known(p);
            if (rc) return rc;
        } else {
            int rc = wait_seal(p, k);
            if (rc) return rc;
            poll_ring(p);
        }
    }
    if (p->known.error & kErrProteins)
        return fail(QHG_ERR_CAPACITY, "an agent has more than " + std::to_string(kProtBudget) +
                                          " proteins (AssortativePair keeps that many digests per agent)");
    if (p->known.error)
        return fail(QHG_ERR_STATE, "device reported an error (flags 0x" + std::to_string(p->known.error) + ")");
    const int kb = (int)(p->known_step & 1);
    if (bounds(p, 1).live > (double)p->cap) {
        // exact status (lag 0) and still no room for one step: grow
        const double need = bounds(p, 4).live;
        int64_t nc = p->cap;
        while ((double)nc < 1.25 * need) nc *= 2;
        int rc = grow_agents(p, nc, p->known.n_alive[kb]);
        if (rc) return rc;
    }
    if (p->gwords) {
        const Bounds now = bounds(p, 0);
        if (bounds(p, 1).pool > (double)p->gcap) {
            // even a compacted pool is too small: compact into a larger one
            const double need = bounds(p, 4).live;
            int64_t ng = p->gcap;
            while ((double)ng < 1.25 * need) ng *= 2;
            int rc = compact_pool(p, ng);
            if (rc) return rc;
        } else if (now.pool > 2.0 * now.live + 65536.0) {
            int rc = compact_pool(p, p->gcap);
            if (rc) return rc;
        }
    }
    return 0;
}

// Every step's scatter writes its status snapshot into the ring and seals
// it every kStatusSealEvery steps; the host reads the sealed ones.  No event
// marks them: an event is a packet of its own between two steps' kernels and
// holds the GPU idle for a few microseconds (the kernel trace shows ~5.6 us
// before a step's first kernel after each one), and a snapshot a few steps
// old bounds the capacity just as well.
constexpr int kRingEvery = kStatusSealEvery;
static_assert(kRing % kRingEvery == 0 && kRing / kRingEvery >= 2, "ring slots for the sealed steps");

int enqueue_readback(qhg_pop *p) {
    if (p->steps_enqueued % kRingEvery != 0) return 0;
    const int slot = (int)(p->steps_enqueued % kRing);
    if (p->ring_step[slot] >= 0) {
        int rc = wait_seal(p, slot);
        if (rc) return rc;
        poll_ring(p);
    }
    p->ring_step[slot] = p->steps_enqueued;
    return 0;
}

int phase_initialize(qhg_pop *p) {
    int rc = ensure_capacity(p);
    if (rc) return rc;
    if (p->has_me && !p->me_done) {
        rc = run_evaluators(p);
        if (rc) return rc;
        p->me_done = true;
    }
    if (p->has_assort) {
        // AssortativePairHam::initialize pairs in the initialize phase
        rc = launch_bs(p, KB_ASSORT, k_assort_pair, p->act_grid, kWave, p->ctx(), p->prog, p->step_args());
        if (rc) return rc;
    }
    p->initialized = true;
    return 0;
}

int phase_act(qhg_pop *p) {
    if (!p->initialized) {
        int rc = phase_initialize(p);
        if (rc) return rc;
    }
    // the age-death branches, the assortative mates and Navigate compile in
    // only where used
    auto kern = p->has_nav ? (p->has_rare ? (p->has_assort ? k_cell_act_nav<true, true> : k_cell_act_nav<true, false>)
                                          : (p->has_assort ? k_cell_act_nav<false, true> : k_cell_act_nav<false, false>))
              : p->has_rare ? (p->has_assort ? k_cell_act<true, true> : k_cell_act<true, false>)
                            : (p->has_assort ? k_cell_act<false, true> : k_cell_act<false, false>);
    int rc = launch_bs(p, KB_ACT, kern, p->act_grid, kCellBlock, p->ctx(), p->prog, p->step_args());
    if (rc) return rc;
    p->acted = true;
    return 0;
}

int phase_finalize(qhg_pop *p) {
    if (!p->acted) {
        int rc = phase_act(p);
        if (rc) return rc;
    }
    const Ctx x = p->ctx();
    const StepArgs sa = p->step_args();
    // a block per tile when they all fit: the straight-line form
    int rc = (p->scan_grid >= p->ntiles) ? launch_bs(p, KB_SCAN, k_cell_scan<true>, p->scan_grid, kScanBlock, x, sa)
                                         : launch_bs(p, KB_SCAN, k_cell_scan<false>, p->scan_grid, kScanBlock, x, sa);
    if (rc) return rc;
    // one more block than the groups use: the step's bookkeeping
    rc = launch_bs(p, KB_SCATTER, k_cell_scatter, p->scatter_grid + 1, kCellBlock, x, sa);
    if (rc) return rc;
    if (p->gwords) {
        rc = launch_bs(p, KB_GENOME, k_genome, p->num_cu * 16, kWave, x, p->prog, sa);
        if (rc) return rc;
    }
    p->acted = false;
    p->initialized = false;
    p->steps_enqueued++;
    p->t_next += 1.0;
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
    const int32_t nc32 = grid->num_cells, conn = grid->connectivity;
    for (int64_t c = 0; c < (int64_t)nc32 * conn; c++)
        if (grid->neighbors[c] < -1 || grid->neighbors[c] >= nc32)
            return fail(QHG_ERR_ARG, "neighbour index out of range");
    for (int32_t c = 0; c < nc32; c++)
        if (grid->num_neighbors[c] > conn) return fail(QHG_ERR_ARG, "num_neighbors > connectivity");
    // incoming runs per cell, padded: src*conn + slot for every slot a move
    // can take (slots below num_neighbors), -1 beyond the cell's in-degree
    std::vector<int32_t> indeg((size_t)nc32, 0);
    for (int32_t s = 0; s < nc32; s++)
        for (int d = 0; d < grid->num_neighbors[s]; d++) {
            const int32_t c = grid->neighbors[(size_t)s * conn + d];
            if (c >= 0) indeg[c]++;
        }
    int32_t in_w = 1;
    for (int32_t c = 0; c < nc32; c++) in_w = std::max(in_w, indeg[c]);
    if (in_w > kMaxIn) return fail(QHG_ERR_ARG, "a cell has more than 8 incoming neighbour slots");
    std::vector<int32_t> in_e((size_t)nc32 * in_w, -1);
    // and the reverse map: the incoming-run index of every (cell, slot)
    std::vector<int32_t> outk((size_t)nc32 * conn, -1);
    {
        std::vector<int32_t> fill((size_t)nc32, 0);
        for (int32_t s = 0; s < nc32; s++)
            for (int d = 0; d < grid->num_neighbors[s]; d++) {
                const int32_t c = grid->neighbors[(size_t)s * conn + d];
                if (c < 0) continue;
                const int32_t k = c * in_w + fill[c]++;
                in_e[(size_t)k] = s * conn + d;
                outk[(size_t)s * conn + d] = k;
            }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QHG_ERR_HIP, "no HIP device");
    qhg_pop *p = new qhg_pop();
    p->P = *params;
    p->ncells = nc32;
    p->conn = conn;
    p->in_w = in_w;
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
    p->ntiles = (int)((p->ncells + kCellTile - 1) / kCellTile);
    {
        // the look-back scan runs at most as many blocks as fit on the device
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cell_scan<false>, kScanBlock, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        p->scan_grid = (int)std::min<int64_t>(p->ntiles, (int64_t)per_cu * p->num_cu);
    }
    set_shape(p);
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_HIP, "stream"));
    const size_t nc = (size_t)p->ncells;
    if (capacity < 1024) capacity = 1024;
    capacity = (capacity + 255) / 256 * 256;
#define CK(x) do { int r_ = (x); if (r_) return cleanup_fail(r_); } while (0)
    p->cap = capacity;
    CK(alloc_agentbuf(p->buf[0], capacity, p->gwords != 0));
    CK(alloc_agentbuf(p->buf[1], capacity, p->gwords != 0));
    CK(alloc_slot_scratch(p, capacity));
    if (p->gwords) {
        p->gcap = capacity;
        if (hipMalloc(&p->d_gpool[0], (size_t)p->gcap * p->gwords * 8) != hipSuccess)
            return cleanup_fail(fail(QHG_ERR_NOMEM, "genome pool"));
    }
    CK(upload_array(&p->d_nbr, grid->neighbors, nc * p->conn, p->stream));
    CK(upload_array(&p->d_nn, grid->num_neighbors, nc, p->stream));
    CK(upload_array(&p->d_in_e, in_e.data(), in_e.size(), p->stream));
    CK(upload_array(&p->d_outk, outk.data(), outk.size(), p->stream));
    if (hipMalloc(&p->d_K, nc * 8) != hipSuccess || hipMalloc(&p->d_W, nc * (p->conn + 1) * 8) != hipSuccess ||
        hipMalloc(&p->d_tmp, nc * 8) != hipSuccess || hipMalloc(&p->d_off[0], (nc + 1) * 4) != hipSuccess ||
        hipMalloc(&p->d_off[1], (nc + 1) * 4) != hipSuccess || hipMalloc(&p->d_cstay, (nc + 2) * 4) != hipSuccess ||
        hipMalloc(&p->d_cbirth, (nc + 2) * 4) != hipSuccess || hipMalloc(&p->d_cin, nc * p->in_w * 4) != hipSuccess ||
        hipMalloc(&p->d_carr, (nc + 2) * 4) != hipSuccess || hipMalloc(&p->d_tact, (size_t)p->ntiles * 4) != hipSuccess ||
        hipMalloc(&p->d_tick, 4) != hipSuccess ||
        hipMalloc(&p->d_birth_off, (nc + 1) * 4) != hipSuccess || hipMalloc(&p->d_bpos, nc * 4) != hipSuccess ||
        hipMalloc(&p->d_rpos, nc * p->conn * 4) != hipSuccess ||
        hipMalloc(&p->d_tiles, (size_t)p->ntiles * 8) != hipSuccess ||
        hipMalloc(&p->d_tiles2, (size_t)p->ntiles * 8) != hipSuccess ||
        hipMalloc(&p->d_occ[0], nc * sizeof(OccCell)) != hipSuccess ||
        hipMalloc(&p->d_occ[1], nc * sizeof(OccCell)) != hipSuccess ||
        hipMalloc(&p->d_wave_part, (size_t)p->num_cu * kMaxBlocksPerCu * kParts * 4) != hipSuccess ||
        hipMalloc(&p->d_scan_part, (size_t)p->ntiles * (kScanBlock / kWave) * 3 * 8) != hipSuccess ||
        hipMalloc(&p->d_st, sizeof(DevStatus)) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_NOMEM, "device allocation failed"));
    (void)hipMemsetAsync(p->d_K, 0, nc * 8, p->stream);
    (void)hipMemsetAsync(p->d_W, 0, nc * (p->conn + 1) * 8, p->stream);
    (void)hipMemsetAsync(p->d_off[0], 0, (nc + 1) * 4, p->stream);
    (void)hipMemsetAsync(p->d_off[1], 0, (nc + 1) * 4, p->stream);
    (void)hipMemsetAsync(p->d_tiles, 0, (size_t)p->ntiles * 8, p->stream);
    (void)hipMemsetAsync(p->d_tiles2, 0, (size_t)p->ntiles * 8, p->stream);
    // per-cell counts start at zero; k_cell_scan restores zeros after use
    (void)hipMemsetAsync(p->d_cstay, 0, (nc + 2) * 4, p->stream);
    (void)hipMemsetAsync(p->d_cbirth, 0, (nc + 2) * 4, p->stream);
    (void)hipMemsetAsync(p->d_cin, 0, nc * p->in_w * 4, p->stream);
    (void)hipMemsetAsync(p->d_carr, 0, (nc + 2) * 4, p->stream);
    (void)hipMemsetAsync(p->d_tact, 0, (size_t)p->ntiles * 4, p->stream);
    (void)hipMemsetAsync(p->d_tick, 0, 4, p->stream);
    (void)hipMemsetAsync(p->d_st, 0, sizeof(DevStatus), p->stream);
    if (env) {
        const double *src[7] = {env->altitude, env->longitude, env->latitude, env->water,
                                env->base_npp, env->temperature, env->rainfall};
        for (int k = 0; k < 7; k++) CK(upload_array(&p->d_env[k], src[k], nc, p->stream));
        if (p->has_nav && env->altitude) p->h_alt.assign(env->altitude, env->altitude + nc);
        CK(upload_array(&p->d_ice, env->ice, nc, p->stream));
        CK(upload_array(&p->d_coastal, env->coastal, nc, p->stream));
    }
    if (hipHostMalloc((void **)&p->h_ring, sizeof(DevStatus) * kRing + sizeof(int64_t) * kRing,
                      hipHostMallocMapped) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_NOMEM, "pinned status ring"));
    std::memset(p->h_ring, 0, sizeof(DevStatus) * kRing);
    p->h_seal = reinterpret_cast<int64_t *>(p->h_ring + kRing);
    clear_seals(p);
    if (hipHostGetDevicePointer((void **)&p->d_ring, p->h_ring, 0) != hipSuccess)
        return cleanup_fail(fail(QHG_ERR_HIP, "status ring device pointer"));
    for (int k = 0; k < kRing; k++) p->ring_step[k] = -1;
    if (hipStreamSynchronize(p->stream) != hipSuccess) return cleanup_fail(fail(QHG_ERR_HIP, "create sync"));
#undef CK
    p->known = DevStatus{};
    p->known_step = 0;
    *out = p;
    return 0;
}

int qhg_gpu_destroy(qhg_pop *p) {
    if (!p) return 0;
    if (p->batch) (void)qhg_batch_remove(p->batch, p);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    free_agentbuf(p->buf[0]);
    free_agentbuf(p->buf[1]);
    free_slot_scratch(p);
    free_nav_tables(p);
    void *ptrs[] = {p->d_off[0], p->d_off[1], p->d_gpool[0], p->d_gpool[1], p->d_nbr, p->d_nn, p->d_in_e,
                    p->d_K, p->d_W, p->d_tmp, p->d_ice, p->d_coastal, p->d_cstay, p->d_cbirth, p->d_cin, p->d_carr, p->d_tact, p->d_tick, p->d_outk,
                    p->d_birth_off, p->d_bpos, p->d_rpos, p->d_tiles, p->d_tiles2, p->d_occ[0], p->d_occ[1],
                    p->d_wave_part, p->d_scan_part, p->d_st, p->d_digest};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (int k = 0; k < 7; k++)
        if (p->d_env[k]) (void)hipFree(p->d_env[k]);
    for (auto &e : p->ev_pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    if (p->h_ring) (void)hipHostFree(p->h_ring);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int qhg_gpu_upload_agents(qhg_pop *p, const qhg_agents *a) {
    if (!p || !a || a->n < 0 || (a->n > 0 && (!a->id || !a->cell || !a->life || !a->gender)))
        return fail(QHG_ERR_ARG, "upload: id, cell, life and gender are required");
    if (p->gwords && a->n > 0 && !a->genome) return fail(QHG_ERR_ARG, "upload: genome required with Genetics");
    std::vector<int64_t> order;
    order.reserve((size_t)a->n);
    for (int64_t i = 0; i < a->n; i++) {
        if (a->life[i] == 0) continue;
        if (a->cell[i] < 0 || a->cell[i] >= p->ncells) return fail(QHG_ERR_ARG, "upload: cell index out of range");
        if (a->id[i] < 0) return fail(QHG_ERR_ARG, "upload: negative agent id");
        order.push_back(i);
    }
    // cell-sorted layout: by cell, then id (input already in that order is kept)
    auto before = [&](int64_t x, int64_t y) {
        if (a->cell[x] != a->cell[y]) return a->cell[x] < a->cell[y];
        return a->id[x] < a->id[y];
    };
    if (!std::is_sorted(order.begin(), order.end(), before)) std::sort(order.begin(), order.end(), before);
    {
        std::vector<int64_t> ids;
        ids.reserve(order.size());
        for (int64_t i : order) ids.push_back(a->id[i]);
        std::sort(ids.begin(), ids.end());
        for (size_t k = 1; k < ids.size(); k++)
            if (ids[k] == ids[k - 1]) return fail(QHG_ERR_ARG, "upload: duplicate agent id");
    }
    const int64_t n = (int64_t)order.size();
    if (n > (int64_t)0x7fffffff / 2) return fail(QHG_ERR_ARG, "upload: too many agents");
    HIPCHK(hipStreamSynchronize(p->stream));
    if (2 * n + 1024 > p->cap) {
        int64_t nc = p->cap;
        while (nc < 2 * n + 1024) nc *= 2;
        free_agentbuf(p->buf[0]);
        free_agentbuf(p->buf[1]);
        int rc = alloc_agentbuf(p->buf[0], nc, p->gwords != 0);
        if (rc) return rc;
        rc = alloc_agentbuf(p->buf[1], nc, p->gwords != 0);
        if (rc) return rc;
        rc = alloc_slot_scratch(p, nc);
        if (rc) return rc;
        p->cap = nc;
    }
    if (p->gwords && 2 * n + 1024 > p->gcap) {
        int64_t ng = p->gcap;
        while (ng < 2 * n + 1024) ng *= 2;
        for (int k = 0; k < 2; k++) {
            if (p->d_gpool[k]) (void)hipFree(p->d_gpool[k]);
            p->d_gpool[k] = nullptr;
        }
        HIPCHK(hipMalloc(&p->d_gpool[0], (size_t)ng * p->gwords * 8));
        p->gp_cur = 0;
        p->gcap = ng;
    }
    std::vector<int64_t> id(n), mate(n, -3);
    std::vector<int32_t> nbab(n, 0), gh(n);
    std::vector<uint32_t> off((size_t)p->ncells + 1, 0);
    std::vector<uint8_t> life(n), gender(n);
    std::vector<double> birth(n, 0.0), age(n, 0.0), lastb(n, 0.0);
    std::vector<uint64_t> genome((size_t)n * p->gwords);
    int64_t max_id = 0;
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = order[k];
        id[k] = a->id[i];
        off[(size_t)a->cell[i] + 1]++;
        life[k] = a->life[i];
        gender[k] = a->gender[i];
        if (a->birth_time) birth[k] = a->birth_time[i];
        if (a->age) age[k] = a->age[i];
        if (a->last_birth) lastb[k] = a->last_birth[i];
        if (a->num_babies) nbab[k] = a->num_babies[i];
        gh[k] = (int32_t)k;
        if (p->gwords)
            std::memcpy(&genome[(size_t)k * p->gwords], a->genome + (size_t)i * p->gwords, (size_t)p->gwords * 8);
        max_id = std::max(max_id, id[k]);
    }
    std::vector<OccCell> occ;
    for (int32_t c = 0; c < p->ncells; c++) {
        if (off[(size_t)c + 1] > 0) occ.push_back(OccCell{c, off[c], off[(size_t)c + 1], 0u});
        off[(size_t)c + 1] += off[c];
    }
    const int b = p->cur();
    const size_t ncz = (size_t)p->ncells;
    HIPCHK(hipMemsetAsync(p->d_cstay, 0, (ncz + 2) * 4, p->stream));
    HIPCHK(hipMemsetAsync(p->d_cbirth, 0, (ncz + 2) * 4, p->stream));
    HIPCHK(hipMemsetAsync(p->d_cin, 0, ncz * p->in_w * 4, p->stream));
    HIPCHK(hipMemsetAsync(p->d_carr, 0, (ncz + 2) * 4, p->stream));
    HIPCHK(hipMemsetAsync(p->d_tact, 0, (size_t)p->ntiles * 4, p->stream));
    if (!occ.empty())
        HIPCHK(hipMemcpyAsync(p->d_occ[b], occ.data(), occ.size() * sizeof(OccCell), hipMemcpyHostToDevice,
                              p->stream));
    const AgentBuf &d = p->buf[b];
    hipStream_t s = p->stream;
    HIPCHK(hipMemcpyAsync(p->d_off[b], off.data(), off.size() * 4, hipMemcpyHostToDevice, s));
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(d.id, id.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.mate_id, mate.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.life, life.data(), n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.gender, gender.data(), n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.birth, birth.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.age, age.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.lastb, lastb.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.nbabies, nbab.data(), n * 4, hipMemcpyHostToDevice, s));
        if (p->gwords) {
            HIPCHK(hipMemcpyAsync(d.ghandle, gh.data(), n * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(p->d_gpool[p->gp_cur], genome.data(), (size_t)n * p->gwords * 8,
                                  hipMemcpyHostToDevice, s));
        }
    }
    DevStatus st;
    HIPCHK(hipMemcpyAsync(&st, p->d_st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    st.n_alive[b] = n;
    st.n_occ[b] = (int64_t)occ.size();
    st.max_id[b] = max_id;
    st.gpool[b] = p->gwords ? n : 0;
    st.step = p->steps_enqueued;
    st.error = 0;
    HIPCHK(hipMemcpyAsync(p->d_st, &st, sizeof(st), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    p->known = st;
    p->known_step = p->steps_enqueued;
    p->pool_reset_step = -1;
    p->gen++;
    return 0;
}

int qhg_gpu_set_clock(qhg_pop *p, int64_t step, double t, int64_t max_id) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    if (step < 0 || step > (int64_t)UINT32_MAX || max_id < 0) return fail(QHG_ERR_ARG, "step or max_id out of range");
    HIPCHK(hipStreamSynchronize(p->stream));
    poll_ring(p);
    for (int k = 0; k < kRing; k++) p->ring_step[k] = -1;
    // the steps may run again from an earlier one: no old seal may match
    clear_seals(p);
    DevStatus st;
    int rc = read_status_sync(p, &st);
    if (rc) return rc;
    const int nb = (int)(step & 1);
    if (nb != p->cur()) {
        // the agents stay where they are: the buffers change roles
        std::swap(p->buf[0], p->buf[1]);
        std::swap(p->d_off[0], p->d_off[1]);
        std::swap(p->d_occ[0], p->d_occ[1]);
        std::swap(st.n_alive[0], st.n_alive[1]);
        std::swap(st.n_occ[0], st.n_occ[1]);
        std::swap(st.max_id[0], st.max_id[1]);
        std::swap(st.gpool[0], st.gpool[1]);
    }
    st.max_id[nb] = std::max(st.max_id[nb], max_id);
    st.step = step;
    HIPCHK(hipMemcpyAsync(p->d_st, &st, sizeof(st), hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->steps_enqueued = step;
    p->t_next = t;
    p->known = st;
    p->known_step = step;
    p->pool_reset_step = -1;
    p->gen++;
    return 0;
}

int qhg_gpu_get_clock(qhg_pop *p, int64_t *step, double *t, int64_t *max_id) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    DevStatus st;
    int rc = read_status_sync(p, &st);
    if (rc) return rc;
    if (step) *step = p->steps_enqueued;
    if (t) *t = p->t_next;
    if (max_id) *max_id = st.max_id[p->cur()];
    return 0;
}

int qhg_gpu_get_params(qhg_pop *p, qhg_params *out) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    *out = p->P;
    return 0;
}

int qhg_gpu_num_alive(qhg_pop *p, int64_t *n) {
    if (!p || !n) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    *n = s.n_alive[p->cur()];
    return 0;
}

int qhg_gpu_download_agents(qhg_pop *p, qhg_agents *a) {
    if (!p || !a) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    const int b = p->cur();
    const int64_t n = s.n_alive[b];
    const int64_t np = p->gwords ? s.gpool[b] : 0;
    const AgentBuf &d = p->buf[b];
    std::vector<int64_t> id(n), mate(n);
    std::vector<int32_t> nbab(n), gh(n);
    std::vector<OccCell> occ((size_t)std::max<int64_t>(0, s.n_occ[b]));
    std::vector<uint8_t> life(n), gender(n);
    std::vector<double> birth(n), age(n), lastb(n);
    std::vector<uint64_t> pool;
    hipStream_t st = p->stream;
    if (!occ.empty())
        HIPCHK(hipMemcpyAsync(occ.data(), p->d_occ[b], occ.size() * sizeof(OccCell), hipMemcpyDeviceToHost, st));
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(id.data(), d.id, n * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(mate.data(), d.mate_id, n * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(nbab.data(), d.nbabies, n * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(life.data(), d.life, n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(gender.data(), d.gender, n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(birth.data(), d.birth, n * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(age.data(), d.age, n * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(lastb.data(), d.lastb, n * 8, hipMemcpyDeviceToHost, st));
        if (p->gwords && a->genome) {
            HIPCHK(hipMemcpyAsync(gh.data(), d.ghandle, n * 4, hipMemcpyDeviceToHost, st));
            pool.resize((size_t)np * p->gwords);
            if (np > 0)
                HIPCHK(hipMemcpyAsync(pool.data(), p->d_gpool[p->gp_cur], (size_t)np * p->gwords * 8,
                                      hipMemcpyDeviceToHost, st));
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    // every agent's cell from the occupied-cell list (cell, first slot, count)
    std::vector<int32_t> cell(n, -1);
    int64_t covered = 0;
    for (const OccCell &e : occ) {
        if (e.c < 0 || e.c >= p->ncells || (int64_t)e.s + e.n > n) return fail(QHG_ERR_STATE, "download: bad cell list");
        for (uint32_t i = e.s; i < e.s + e.n; i++) cell[i] = e.c;
        covered += e.n;
    }
    if (covered != n) return fail(QHG_ERR_STATE, "download: the cell list disagrees with the count");
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return id[x] < id[y]; });
    std::vector<int64_t> sorted_ids(n);
    for (int64_t k = 0; k < n; k++) sorted_ids[k] = id[order[k]];
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = order[k];
        if (a->id) a->id[k] = id[i];
        if (a->cell) a->cell[k] = cell[i];
        if (a->life) a->life[k] = life[i];
        if (a->gender) a->gender[k] = gender[i];
        if (a->birth_time) a->birth_time[k] = birth[i];
        if (a->age) a->age[k] = age[i];
        if (a->last_birth) a->last_birth[k] = lastb[i];
        if (a->num_babies) a->num_babies[k] = nbab[i];
        if (a->mate_id) {
            // the partner's id while it lives, -1 once it has died
            const int64_t m = mate[i];
            a->mate_id[k] = (m >= 0) ? (std::binary_search(sorted_ids.begin(), sorted_ids.end(), m) ? m : -1) : m;
        }
        if (a->genome && p->gwords) {
            const int64_t h = gh[i];
            if (h < 0 || h >= np) return fail(QHG_ERR_STATE, "download: genome handle out of range");
            std::memcpy(a->genome + (size_t)k * p->gwords, &pool[(size_t)h * p->gwords], (size_t)p->gwords * 8);
        }
    }
    a->n = n;
    return 0;
}

int qhg_gpu_state_digest(qhg_pop *p, uint64_t out[3]) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    const int b = p->cur();
    if (!p->d_digest) HIPCHK(hipMalloc(&p->d_digest, 2 * sizeof(unsigned long long)));
    HIPCHK(hipMemsetAsync(p->d_digest, 0, 2 * sizeof(unsigned long long), p->stream));
    const int64_t nocc = s.n_occ[b];
    if (nocc > 0) {
        const int grid = (int)std::min<int64_t>((nocc + kBlock / kWave - 1) / (kBlock / kWave), 8 * p->num_cu);
        hipLaunchKernelGGL(k_state_digest, dim3(grid), dim3(kBlock), 0, p->stream, p->buf[b], p->d_occ[b], nocc,
                           p->gwords ? p->d_gpool[p->gp_cur] : nullptr, p->gwords, p->d_digest);
        HIPCHK(hipGetLastError());
    }
    unsigned long long h[2];
    HIPCHK(hipMemcpyAsync(h, p->d_digest, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    out[0] = (uint64_t)s.n_alive[b];
    out[1] = (uint64_t)h[0];
    out[2] = (uint64_t)h[1];
    return 0;
}

int qhg_gpu_pre_loop(qhg_pop *p, double t) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    p->t_next = t;
    if (p->has_cap) {
        int rc = run_capacity(p);
        if (rc) return rc;
    }
    if (p->has_nav) {
        // Navigate::preLoop (QHG4/actions/Navigate.cpp:155-183): it needs the
        // geography and the navigation, and builds the tables
        if (!p->nav_set) return fail(QHG_ERR_STATE, "Navigate needs the navigation (qhg_gpu_set_navigation)");
        if (p->h_alt.empty()) return fail(QHG_ERR_STATE, "Navigate needs the altitudes");
        HIPCHK(hipStreamSynchronize(p->stream));
        int rc = nav_rebuild(p);
        if (rc) return rc;
        p->nav_preloop_done = true;
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_initialize_step(qhg_pop *p, double t) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    int rc = phase_initialize(p);
    if (rc) return rc;
    p->t_next = t;
    return 0;
}

int qhg_gpu_do_actions(qhg_pop *p, uint32_t prio, double t) {
    (void)prio;
    (void)t;
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    // all priority levels run fused in the first call of a step
    if (p->acted) return 0;
    return phase_act(p);
}

int qhg_gpu_finalize_step(qhg_pop *p) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    int rc = phase_finalize(p);
    if (rc) return rc;
    return p->timing ? collect_timing(p) : 0;
}

int qhg_gpu_run(qhg_pop *p, int32_t nsteps) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    for (int32_t s = 0; s < nsteps; s++) {
        int rc = phase_initialize(p);
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

int qhg_gpu_set_env(qhg_pop *p, const qhg_env_desc *env) { return qhg_gpu_env_event(p, env, 0); }

int qhg_gpu_env_event(qhg_pop *p, const qhg_env_desc *env, int32_t events) {
    if (!p || !env) return fail(QHG_ERR_ARG, "null argument");
    if (events & ~(QHG_EVENT_GEO | QHG_EVENT_CLIMATE | QHG_EVENT_VEG | QHG_EVENT_KILL_WET))
        return fail(QHG_ERR_ARG, "unknown event bits");
    if ((events & QHG_EVENT_KILL_WET) && p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
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
    if ((events & QHG_EVENT_KILL_WET) && (p->d_env[0] || p->d_ice)) {
        poll_ring(p);
        DevStatus st;
        int rc = read_status_sync(p, &st);
        if (rc) return rc;
        const int b = p->cur();
        if (st.n_occ[b] > 0) {
            // d_tmp (ncells doubles) holds the kept cells' old segment starts
            uint32_t *src_s = reinterpret_cast<uint32_t *>(p->d_tmp);
            rc = launch_bs(p, KB_CAPACITY, k_env_kill, 1, kKillBlock, (const OccCell *)p->d_occ[b], st.n_occ[b],
                           (const double *)p->d_env[0], (const uint8_t *)p->d_ice, p->d_occ[1 - b], src_s,
                           p->d_st, b);
            if (rc) return rc;
            rc = launch(p, KB_CAPACITY, k_env_move, grid_for(st.n_occ[b] * kWave, 4096), p->buf[b], p->buf[1 - b],
                        (const OccCell *)p->d_occ[1 - b], (const uint32_t *)src_s, (const DevStatus *)p->d_st,
                        1 - b);
            if (rc) return rc;
            rc = read_status_sync(p, &st);
            if (rc) return rc;
            // the repacked buffer takes the current buffer's role
            std::swap(p->buf[0], p->buf[1]);
            std::swap(p->d_off[0], p->d_off[1]);
            std::swap(p->d_occ[0], p->d_occ[1]);
            std::swap(st.n_alive[0], st.n_alive[1]);
            std::swap(st.n_occ[0], st.n_occ[1]);
            std::swap(st.max_id[0], st.max_id[1]);
            std::swap(st.gpool[0], st.gpool[1]);
            HIPCHK(hipMemcpyAsync(p->d_st, &st, sizeof(st), hipMemcpyHostToDevice, p->stream));
            HIPCHK(hipStreamSynchronize(p->stream));
            p->known = st;
            p->known_step = p->steps_enqueued;
            p->gen++;
        }
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
    if (p->has_nav) {
        // Navigate::notify: a geography event makes an update due, the
        // flush runs it (once the pre-loop has built the first tables)
        if (env->altitude) p->h_alt.assign(env->altitude, env->altitude + nc);
        if (((events & QHG_EVENT_GEO) || env->altitude) && p->nav_preloop_done) {
            HIPCHK(hipStreamSynchronize(p->stream));
            int rc = nav_rebuild(p);
            if (rc) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

int qhg_gpu_set_navigation(qhg_pop *p, const qhg_nav_desc *nav) {
    if (!p || !nav) return fail(QHG_ERR_ARG, "null argument");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    const int32_t np = nav->num_ports, nc = p->ncells;
    if (np < 0 || nav->num_bridges < 0 || nav->num_bridges > 131072) return fail(QHG_ERR_ARG, "navigation sizes");
    if (np && (!nav->ports || !nav->dest_offsets)) return fail(QHG_ERR_ARG, "navigation: null ports");
    if (nav->num_bridges && !nav->bridges) return fail(QHG_ERR_ARG, "navigation: null bridges");
    std::vector<int32_t> order((size_t)np);
    for (int32_t k = 0; k < np; k++) {
        order[(size_t)k] = k;
        if (nav->ports[k] < 0 || nav->ports[k] >= nc) return fail(QHG_ERR_ARG, "navigation: port out of range");
        if (nav->dest_offsets[k] < 0 || nav->dest_offsets[k + 1] < nav->dest_offsets[k])
            return fail(QHG_ERR_ARG, "navigation: destination offsets");
    }
    if (np && nav->dest_offsets[np] > 0 && (!nav->dest_cells || !nav->dest_dists))
        return fail(QHG_ERR_ARG, "navigation: null destinations");
    // the maps' order: ports by cell, each port's destinations by cell
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return nav->ports[a] < nav->ports[b]; });
    std::vector<int32_t> port, doff(1, 0), dcell;
    std::vector<double> ddist;
    for (int32_t i = 0; i < np; i++) {
        const int32_t k = order[(size_t)i];
        if (i > 0 && nav->ports[k] == port.back()) return fail(QHG_ERR_ARG, "navigation: a port is listed twice");
        port.push_back(nav->ports[k]);
        const int32_t a = nav->dest_offsets[k], n = nav->dest_offsets[k + 1] - a;
        std::vector<int32_t> dj((size_t)n);
        std::iota(dj.begin(), dj.end(), a);
        std::stable_sort(dj.begin(), dj.end(), [&](int32_t x, int32_t y) { return nav->dest_cells[x] < nav->dest_cells[y]; });
        for (int32_t j = 0; j < n; j++) {
            const int32_t c = nav->dest_cells[dj[(size_t)j]];
            if (c < 0 || c >= nc) return fail(QHG_ERR_ARG, "navigation: destination out of range");
            if (j > 0 && c == dcell.back()) return fail(QHG_ERR_ARG, "navigation: a destination is listed twice");
            dcell.push_back(c);
            ddist.push_back(nav->dest_dists[dj[(size_t)j]]);
        }
        doff.push_back((int32_t)dcell.size());
    }
    std::vector<int32_t> br((size_t)2 * nav->num_bridges);
    for (size_t b = 0; b < br.size(); b++) {
        br[b] = nav->bridges[b];
        if (br[b] < 0 || br[b] >= nc) return fail(QHG_ERR_ARG, "navigation: bridge end out of range");
    }
    HIPCHK(hipStreamSynchronize(p->stream));
    p->nav_port.swap(port);
    p->nav_doff.swap(doff);
    p->nav_dcell.swap(dcell);
    p->nav_ddist.swap(ddist);
    p->nav_bridge.swap(br);
    p->nav_set = true;
    // before the pre-loop this is the grid's data; afterwards EVENT_ID_NAV
    // and its flush (Navigate::notify, QHG4/actions/Navigate.cpp:79-89)
    if (p->has_nav && p->nav_preloop_done) return nav_rebuild(p);
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
    const int b = p->cur();
    std::memset(o, 0, sizeof(*o));
    o->num_alive = s.n_alive[b];
    o->high_water = s.n_alive[b];
    o->capacity = p->cap;
    o->max_id = s.max_id[b];
    o->step = s.step;
    o->births_last = s.births_last;
    o->deaths_last = s.deaths_last;
    o->moves_last = s.moves_last;
    o->births_total = s.births_total;
    o->deaths_total = s.deaths_total;
    o->moves_total = s.moves_total;
    o->compactions = p->compactions;
    o->device_error = s.error;
    o->host_error = p->host_error;
    o->agent_steps = s.agent_steps;
    o->slot_steps = s.agent_steps;
    o->fertile_total = s.fertile_total;
    o->occupied_cell_steps = s.occupied_total;
    return 0;
}

int qhg_gpu_cell_counts(qhg_pop *p, uint32_t *out) {
    if (!p || !out) return fail(QHG_ERR_ARG, "null argument");
    DevStatus s;
    int rc = read_status_sync(p, &s);
    if (rc) return rc;
    const int b = p->cur();
    std::vector<OccCell> occ((size_t)std::max<int64_t>(0, s.n_occ[b]));
    if (!occ.empty())
        HIPCHK(hipMemcpyAsync(occ.data(), p->d_occ[b], occ.size() * sizeof(OccCell), hipMemcpyDeviceToHost,
                              p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    std::memset(out, 0, (size_t)p->ncells * 4);
    for (const OccCell &e : occ) {
        if (e.c < 0 || e.c >= p->ncells) return fail(QHG_ERR_STATE, "cell counts: bad cell list");
        out[e.c] = e.n;
    }
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
    // the agent buffers never hold holes; the genome pool does
    if (p->gwords) {
        int rc = compact_pool(p, p->gcap);
        if (rc) return rc;
    }
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

int qhg_gpu_modify_attribute(qhg_pop *p, const char *name, double value) {
    if (!p || !name) return fail(QHG_ERR_ARG, "null argument");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    if (std::strncmp(name, "NPPCap_", 7) == 0 || std::strncmp(name, "Multi_weight_", 13) == 0)
        return fail(QHG_ERR_ARG, std::string(name) + " feeds the capacity and weight tables, which change only "
                                 "with environment events");
    qhg_params P = p->P;
    int rc = qhg_params_set_attribute(&P, name, value);
    if (rc) return rc;
    const qhg_params old = p->P;
    const int64_t gw = p->gwords;
    p->P = P;
    rc = build_prog(p);
    if (rc || p->gwords != gw) {
        p->P = old;
        (void)build_prog(p);
        return rc ? rc : fail(QHG_ERR_ARG, "the genome size cannot change");
    }
    p->gen++;   // a batch re-reads the member's program
    return 0;
}

int qhg_gpu_device_bytes(qhg_pop *p, int64_t *bytes) {
    if (!p || !bytes) return fail(QHG_ERR_ARG, "null argument");
    *bytes = device_bytes(p);
    return 0;
}

int qhg_gpu_set_tuning(qhg_pop *p, const qhg_tuning *t) {
    if (!p || !t) return fail(QHG_ERR_ARG, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    p->group_agents = t->group_agents;
    p->blocks_per_cu = t->blocks_per_cu;
    set_shape(p);
    return 0;
}

int qhg_gpu_get_tuning(qhg_pop *p, qhg_tuning *t) {
    if (!p || !t) return fail(QHG_ERR_ARG, "null argument");
    *t = qhg_tuning{p->group_agents, p->blocks_per_cu, 0};
    return 0;
}

int qhg_gpu_autotune(qhg_pop *p, int32_t trial_steps, qhg_tuning *chosen, int32_t *steps_taken) {
    if (!p) return fail(QHG_ERR_ARG, "null handle");
    if (p->batch) return fail(QHG_ERR_STATE, "population is stepped by its batch");
    if (trial_steps < 1) return fail(QHG_ERR_ARG, "trial_steps < 1");
    int rc = qhg_gpu_set_timing(p, 0);
    if (rc) return rc;
    Timer tm;
    rc = tm.init();
    if (rc) return rc;
    auto trial = [&](double *rate) -> int {
        int r = refresh_known(p);
        if (r) return r;
        const int64_t a0 = p->known.agent_steps;
        HIPCHK(hipEventRecord(tm.e0, p->stream));
        r = qhg_gpu_run(p, trial_steps);
        if (r) return r;
        HIPCHK(hipEventRecord(tm.e1, p->stream));
        r = refresh_known(p);
        if (r) return r;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
        *rate = ms > 0 ? (double)(p->known.agent_steps - a0) / ms : 0.0;
        return 0;
    };
    auto apply = [&](const qhg_tuning &t) { return qhg_gpu_set_tuning(p, &t); };
    qhg_tuning t0{p->group_agents, p->blocks_per_cu, 0}, best{};
    int32_t taken = 0;
    rc = tune_search(t0, trial_steps, apply, trial, &best, &taken);
    if (steps_taken) *steps_taken = taken;
    if (rc) return rc;
    if (chosen) *chosen = best;
    return 0;
}

}  // extern "C"

// ===========================================================================
// Batched stepping (many populations per launch)
// ===========================================================================
enum BatchKernelId { BK_ACT, BK_SCAN, BK_SCATTER, BK_GENOME, BK_SCAN_SUMS, BK_COUNT };

struct qhg_batch {
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 256;
    int scan_grid = 1, genome_grid = 1;
    std::vector<qhg_pop *> pops;
    std::vector<int64_t> gens, step0, join;
    std::vector<double> t0;
    bool dirty = true;
    int64_t bstep = 0;
    PopDesc *d_desc = nullptr;
    size_t desc_cap = 0;
    int32_t *d_tile_base = nullptr;
    uint64_t *d_tiles = nullptr, *d_tiles2 = nullptr;
    size_t tiles_cap = 0;
    int32_t total_tiles = 0;
    bool any_genetics = false, any_death = false, any_nav = false;
    int64_t *d_bctr = nullptr;     // the kernels' copies of the batch step index
    // launch shapes (results do not depend on them)
    double group_agents = QHG_GROUP_AGENTS;
    int blocks_per_cu = QHG_CELL_GRID_PER_CU;
    // one step captured as a graph, replayed while the launch set is unchanged
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_key[4] = {-1, -1, -1, -1};
    int64_t table_version = 0;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<std::pair<int, size_t>> ev_used;
    double t_sum[BK_COUNT] = {};
    int64_t t_cnt[BK_COUNT] = {};
};

namespace {

template <typename... KArgs, typename... Args>
int blaunch(qhg_batch *b, int kid, void (*kern)(KArgs...), int grid, int block, Args... args) {
    if (b->timing) {
        const size_t slot = b->ev_used.size();
        if (slot >= b->ev_pool.size()) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            b->ev_pool.push_back({e0, e1});
        }
        b->ev_used.push_back({kid, slot});
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, b->stream, b->ev_pool[slot].first,
                              b->ev_pool[slot].second, 0, args...);
    } else {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, b->stream, args...);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

BatchArgs batch_args(const qhg_batch *b) {
    BatchArgs A{};
    A.pops = b->d_desc;
    A.npops = (int32_t)b->pops.size();
    A.total_tiles = b->total_tiles;
    A.tile_base = b->d_tile_base;
    A.tiles = b->d_tiles;
    A.tiles2 = b->d_tiles2;
    A.bctr = b->d_bctr;
    A.group_agents = b->group_agents;
    A.genome = b->any_genetics ? 1 : 0;
    return A;
}

int batch_act_grid(const qhg_batch *b) { return b->num_cu * b->blocks_per_cu; }

// the kernels of one batch step, in order
int batch_launch_step(qhg_batch *b) {
    const BatchArgs A = batch_args(b);
    const int ag = batch_act_grid(b);
    int rc = b->any_nav     ? blaunch(b, BK_ACT, k_batch_act<true, true>, ag, kCellBlock, A)
             : b->any_death ? blaunch(b, BK_ACT, k_batch_act<true>, ag, kCellBlock, A)
                            : blaunch(b, BK_ACT, k_batch_act<false>, ag, kCellBlock, A);
    if (rc) return rc;
    const int sg = std::max(1, std::min(b->scan_grid, b->total_tiles));
    rc = blaunch(b, BK_SCAN_SUMS, k_batch_scan_sums, sg, kScanBlock, A);
    if (rc) return rc;
    rc = blaunch(b, BK_SCAN, k_batch_scan, sg, kScanBlock, A);
    if (rc) return rc;
    // one more block than the groups use: the members' bookkeeping
    rc = blaunch(b, BK_SCATTER, k_batch_scatter, ag + 1, kCellBlock, A);
    if (rc) return rc;
    if (b->any_genetics) return blaunch(b, BK_GENOME, k_batch_genome, b->genome_grid, kWave, A);
    return 0;
}

void batch_drop_graph(qhg_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    for (auto &k : b->graph_key) k = -1;
}

// one batch step: eagerly (and timed, when timing is on) or as a replay of
// the captured step, recaptured whenever the launch set changes
int batch_step(qhg_batch *b) {
    if (!b->use_graph || b->timing) return batch_launch_step(b);
    const int64_t key[4] = {b->table_version, (int64_t)(b->group_agents * 1024.0), b->blocks_per_cu,
                            (b->any_genetics ? 1 : 0) | (b->any_death ? 2 : 0) | (b->any_nav ? 4 : 0)};
    if (!b->graph_exec || !std::equal(key, key + 4, b->graph_key)) {
        batch_drop_graph(b);
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_launch_step(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(QHG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        std::copy(key, key + 4, b->graph_key);
    }
    HIPCHK(hipGraphLaunch(b->graph_exec, b->stream));
    return 0;
}

int batch_collect_timing(qhg_batch *b) {
    if (b->ev_used.empty()) return 0;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto &u : b->ev_used) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, b->ev_pool[u.second].first, b->ev_pool[u.second].second));
        b->t_sum[u.first] += ms;
        b->t_cnt[u.first] += 1;
    }
    b->ev_used.clear();
    return 0;
}

// device table of the members' contexts (both parities), their scan tiles
// laid out one after another in the batch's look-back arrays
int batch_upload_table(qhg_batch *b) {
    const size_t n = b->pops.size();
    std::vector<int32_t> tile_base(n + 1, 0);
    for (size_t i = 0; i < n; i++) tile_base[i + 1] = tile_base[i] + b->pops[i]->ntiles;
    const size_t tiles = std::max<size_t>(1, (size_t)tile_base[n]);
    HIPCHK(hipStreamSynchronize(b->stream));
    if (tiles > b->tiles_cap) {
        if (b->d_tiles) (void)hipFree(b->d_tiles);
        if (b->d_tiles2) (void)hipFree(b->d_tiles2);
        b->d_tiles = b->d_tiles2 = nullptr;
        HIPCHK(hipMalloc(&b->d_tiles, tiles * 8));
        HIPCHK(hipMalloc(&b->d_tiles2, tiles * 8));
        HIPCHK(hipMemset(b->d_tiles, 0, tiles * 8));
        HIPCHK(hipMemset(b->d_tiles2, 0, tiles * 8));
        b->tiles_cap = tiles;
    }
    if (n + 1 > b->desc_cap) {
        if (b->d_desc) (void)hipFree(b->d_desc);
        if (b->d_tile_base) (void)hipFree(b->d_tile_base);
        b->d_desc = nullptr;
        b->d_tile_base = nullptr;
        const size_t c = std::max<size_t>(16, 2 * (n + 1));
        HIPCHK(hipMalloc(&b->d_desc, c * sizeof(PopDesc)));
        HIPCHK(hipMalloc(&b->d_tile_base, c * sizeof(int32_t)));
        b->desc_cap = c;
    }
    std::vector<PopDesc> desc(std::max<size_t>(1, n));
    b->any_genetics = false;
    b->any_death = false;
    b->any_nav = false;
    for (size_t i = 0; i < n; i++) {
        qhg_pop *p = b->pops[i];
        for (int par = 0; par < 2; par++) {
            desc[i].ctx[par] = p->ctx_for(par);
            desc[i].ctx[par].tiles = b->d_tiles + tile_base[i];
            desc[i].ctx[par].tiles2 = b->d_tiles2 + tile_base[i];
        }
        desc[i].prog = p->prog;
        desc[i].step0 = b->step0[i];
        desc[i].join = b->join[i];
        desc[i].t0 = b->t0[i];
        b->gens[i] = p->gen;
        if (p->gwords) b->any_genetics = true;
        if (p->has_rare) b->any_death = true;
        if (p->has_nav) b->any_nav = true;
    }
    if (n) HIPCHK(hipMemcpy(b->d_desc, desc.data(), n * sizeof(PopDesc), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_tile_base, tile_base.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    b->total_tiles = tile_base[n];
    b->dirty = false;
    b->table_version++;
    return 0;
}

}  // namespace

extern "C" {

int qhg_batch_create(qhg_batch **out) {
    if (!out) return fail(QHG_ERR_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QHG_ERR_HIP, "no HIP device");
    qhg_batch *b = new qhg_batch();
    if (hipGetDevice(&b->device) != hipSuccess) { delete b; return fail(QHG_ERR_HIP, "hipGetDevice"); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(QHG_ERR_HIP, "stream");
    }
    if (hipMalloc(&b->d_bctr, BC_COUNT * sizeof(int64_t)) != hipSuccess ||
        hipMemset(b->d_bctr, 0, BC_COUNT * sizeof(int64_t)) != hipSuccess) {
        qhg_batch_destroy(b);
        return fail(QHG_ERR_NOMEM, "batch counters");
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_scan, kScanBlock, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    b->scan_grid = per_cu * b->num_cu;
    b->genome_grid = b->num_cu * 16;
    *out = b;
    return 0;
}

int qhg_batch_destroy(qhg_batch *b) {
    if (!b) return 0;
    while (!b->pops.empty()) (void)qhg_batch_remove(b, b->pops.back());
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    void *ptrs[] = {b->d_desc, b->d_tile_base, b->d_tiles, b->d_tiles2, b->d_bctr};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (auto &e : b->ev_pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int qhg_batch_add(qhg_batch *b, qhg_pop *p) {
    if (!b || !p) return fail(QHG_ERR_ARG, "null argument");
    if (p->batch) return fail(QHG_ERR_STATE, "population already belongs to a batch");
    if (p->device != b->device) return fail(QHG_ERR_ARG, "population and batch are on different devices");
    if ((int)b->pops.size() >= kMaxPops) return fail(QHG_ERR_CAPACITY, "batch is full (256 populations)");
    if (p->acted || p->initialized) return fail(QHG_ERR_STATE, "population is inside a step");
    if (p->has_assort)
        return fail(QHG_ERR_ARG, "AssortativePairHamming and AssortativePair populations do not run in batches");
    // hand over: everything the population enqueued so far has finished
    HIPCHK(hipStreamSynchronize(p->stream));
    poll_ring(p);
    int rc = refresh_known(p);
    if (rc) return rc;
    if (p->has_me && !p->me_done) {
        rc = run_evaluators(p);
        if (rc) return rc;
        p->me_done = true;
        HIPCHK(hipStreamSynchronize(p->stream));
    }
    p->own_stream = p->stream;
    p->stream = b->stream;
    for (int k = 0; k < kRing; k++) p->ring_step[k] = -1;
    p->batch = b;
    b->pops.push_back(p);
    b->gens.push_back(p->gen);
    b->step0.push_back(p->steps_enqueued);
    b->join.push_back(b->bstep);
    b->t0.push_back(p->t_next);
    b->dirty = true;
    return 0;
}

int qhg_batch_remove(qhg_batch *b, qhg_pop *p) {
    if (!b || !p) return fail(QHG_ERR_ARG, "null argument");
    size_t i = 0;
    while (i < b->pops.size() && b->pops[i] != p) i++;
    if (i == b->pops.size()) return fail(QHG_ERR_ARG, "population is not in this batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    poll_ring(p);
    p->stream = p->own_stream;
    for (int k = 0; k < kRing; k++) p->ring_step[k] = -1;
    p->batch = nullptr;
    int rc = refresh_known(p);
    b->pops.erase(b->pops.begin() + i);
    b->gens.erase(b->gens.begin() + i);
    b->step0.erase(b->step0.begin() + i);
    b->join.erase(b->join.begin() + i);
    b->t0.erase(b->t0.begin() + i);
    b->dirty = true;
    return rc;
}

int qhg_batch_size(qhg_batch *b, int32_t *n) {
    if (!b || !n) return fail(QHG_ERR_ARG, "null argument");
    *n = (int32_t)b->pops.size();
    return 0;
}

int qhg_batch_run(qhg_batch *b, int32_t nsteps) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    int nfailed = 0;
    std::string first_msg;
    for (int32_t s = 0; s < nsteps; s++) {
        if (b->pops.empty()) break;
        // a member whose step cannot be prepared (its device reported an
        // error, its buffers could not grow) leaves the batch with its error
        // recorded; the other members go on
        for (size_t i = 0; i < b->pops.size();) {
            qhg_pop *p = b->pops[i];
            int rc = ensure_capacity(p);
            if (rc == 0) { i++; continue; }
            if (nfailed == 0) first_msg = qhg_gpu_last_error();
            nfailed++;
            p->host_error = rc;
            (void)qhg_batch_remove(b, p);
        }
        if (b->pops.empty()) break;
        bool stale = b->dirty;
        for (size_t i = 0; i < b->pops.size(); i++) stale = stale || b->gens[i] != b->pops[i]->gen;
        if (stale) {
            int rc = batch_upload_table(b);
            if (rc) return rc;
        }
        int rc = batch_step(b);
        if (rc) return rc;
        b->bstep++;
        for (qhg_pop *p : b->pops) {
            p->steps_enqueued++;
            p->t_next += 1.0;
        }
        // each member's snapshots are sealed by k_batch_scatter; no event
        // goes into the stream
        for (qhg_pop *p : b->pops) {
            rc = enqueue_readback(p);
            if (rc) return rc;
        }
        if (b->timing && b->ev_used.size() > 4096) {
            rc = batch_collect_timing(b);
            if (rc) return rc;
        }
    }
    if (nfailed)
        return fail(QHG_ERR_MEMBER, std::to_string(nfailed) + " member(s) failed and left the batch (first: " +
                                        first_msg + ")");
    return 0;
}

int qhg_batch_sync(qhg_batch *b) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(b->stream));
    for (qhg_pop *p : b->pops) poll_ring(p);
    return 0;
}

int qhg_batch_set_graph(qhg_batch *b, int32_t enable) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->use_graph = enable != 0;
    if (!b->use_graph) batch_drop_graph(b);
    return 0;
}

int qhg_batch_set_tuning(qhg_batch *b, const qhg_tuning *t) {
    if (!b || !t) return fail(QHG_ERR_ARG, "null argument");
    int rc = check_tuning(t);
    if (rc) return rc;
    b->group_agents = t->group_agents;
    b->blocks_per_cu = t->blocks_per_cu;
    return 0;
}

int qhg_batch_get_tuning(qhg_batch *b, qhg_tuning *t) {
    if (!b || !t) return fail(QHG_ERR_ARG, "null argument");
    *t = qhg_tuning{b->group_agents, b->blocks_per_cu, 0};
    return 0;
}

int qhg_batch_autotune(qhg_batch *b, int32_t trial_steps, qhg_tuning *chosen, int32_t *steps_taken) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    if (trial_steps < 1) return fail(QHG_ERR_ARG, "trial_steps < 1");
    int rc = qhg_batch_set_timing(b, 0);
    if (rc) return rc;
    Timer tm;
    rc = tm.init();
    if (rc) return rc;
    bool member_failed = false;
    auto agent_steps = [&](int64_t *sum) -> int {
        int r = qhg_batch_sync(b);
        if (r) return r;
        int64_t a = 0;
        for (qhg_pop *p : b->pops) a += p->known.agent_steps;
        *sum = a;
        return 0;
    };
    auto trial = [&](double *rate) -> int {
        int64_t a0 = 0, a1 = 0;
        int r = agent_steps(&a0);
        if (r) return r;
        HIPCHK(hipEventRecord(tm.e0, b->stream));
        r = qhg_batch_run(b, trial_steps);
        if (r == QHG_ERR_MEMBER) {   // the members still in the batch took the trial steps
            member_failed = true;
            r = 0;
        }
        if (r) return r;
        HIPCHK(hipEventRecord(tm.e1, b->stream));
        r = agent_steps(&a1);
        if (r) return r;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, tm.e0, tm.e1));
        *rate = ms > 0 ? (double)(a1 - a0) / ms : 0.0;
        return 0;
    };
    auto apply = [&](const qhg_tuning &t) { return qhg_batch_set_tuning(b, &t); };
    qhg_tuning t0{b->group_agents, b->blocks_per_cu, 0}, best{};
    int32_t taken = 0;
    rc = tune_search(t0, trial_steps, apply, trial, &best, &taken);
    if (steps_taken) *steps_taken = taken;
    if (rc) return rc;
    if (chosen) *chosen = best;
    if (member_failed) return fail(QHG_ERR_MEMBER, "member(s) failed during the tuning steps and left the batch");
    return 0;
}

int qhg_batch_set_timing(qhg_batch *b, int32_t enable) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    int rc = batch_collect_timing(b);
    if (rc) return rc;
    b->timing = enable != 0;
    for (int k = 0; k < BK_COUNT; k++) { b->t_sum[k] = 0; b->t_cnt[k] = 0; }
    return 0;
}

const char *qhg_batch_kernel_name(int32_t index) {
    static const char *names[BK_COUNT] = {"k_batch_act", "k_batch_scan", "k_batch_scatter", "k_batch_genome",
                                          "k_batch_scan_sums"};
    if (index < 0 || index >= BK_COUNT) return "";
    return names[index];
}

int qhg_batch_kernel_times(qhg_batch *b, double *avg_ms, int64_t *launches, int32_t max, int32_t *count) {
    if (!b) return fail(QHG_ERR_ARG, "null handle");
    int rc = batch_collect_timing(b);
    if (rc) return rc;
    const int n = std::min<int>(max, BK_COUNT);
    for (int k = 0; k < n; k++) {
        if (avg_ms) avg_ms[k] = b->t_cnt[k] ? b->t_sum[k] / (double)b->t_cnt[k] : 0.0;
        if (launches) launches[k] = b->t_cnt[k];
    }
    if (count) *count = n;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Environment tables without allocation (qhg_env_tables_noalloc)
// ---------------------------------------------------------------------------
// The evaluator kernels' per-cell scratch (one double per cell, the values
// of the evaluator being combined) lives in this fixed device array instead
// of a population's buffers.  Every device the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in symbol): the address of each
// device's copy is looked up once per device and kept in g_scratch_at.
static __device__ __align__(256) double scratch[QHG_NOALLOC_MAX_CELLS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_at[kMaxDevices] = {};

// the address of the current device's copy of `scratch` (looked up on the
// first call on that device, with the device current)
int scratch_address(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QHG_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(QHG_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_at[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_at[dev];
    return 0;
}

}  // namespace

extern "C" int qhg_env_tables_noalloc(const qhg_grid_desc *grid, const qhg_env_desc *env, const qhg_params *params,
                                      double *K_out, double *W_out, void *stream) {
    if (!grid || !env || !params) return fail(QHG_ERR_ARG, "null argument");
    if (grid->num_cells <= 0 || grid->connectivity < 1 || grid->connectivity > 6)
        return fail(QHG_ERR_ARG, "grid: num_cells must be positive and connectivity in 1..6");
    if ((int64_t)grid->num_cells > (int64_t)QHG_NOALLOC_MAX_CELLS)
        return fail(QHG_ERR_CAPACITY, "grid of " + std::to_string(grid->num_cells) + " cells exceeds the " +
                                          std::to_string(QHG_NOALLOC_MAX_CELLS) +
                                          "-cell scratch of qhg_env_tables_noalloc");
    bool has_cap = false, has_me = false;
    for (int i = 0; i < params->num_actions && i < QHG_MAX_ACTIONS; i++) {
        has_cap |= params->action_kind[i] == QHG_ACT_NPPCAP;
        has_me |= params->action_kind[i] == QHG_ACT_MULTIEVAL;
    }
    if (has_cap && !K_out) return fail(QHG_ERR_ARG, "NPPCapacity is scheduled: K_out is required");
    if (has_me && !W_out) return fail(QHG_ERR_ARG, "MultiEvaluator is scheduled: W_out is required");
    if (has_me && !grid->neighbors) return fail(QHG_ERR_ARG, "MultiEvaluator needs the grid's neighbours");
    if (has_me && (params->me_num < 1 || params->me_num > QHG_MAX_EVALS))
        return fail(QHG_ERR_ARG, "MultiEvaluator needs 1..4 evaluators");
    double *tmp = nullptr;
    int rc = scratch_address(&tmp);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int nb = grid_for(grid->num_cells, 4096);
    Ctx x{};
    x.ncells = grid->num_cells;
    x.conn = grid->connectivity;
    x.nbr = grid->neighbors;
    x.K = K_out;
    x.W = W_out;
    x.ice = env->ice;
    if (has_cap) {
        const qhg_params &P = *params;
        EnvPtrs e{env->altitude, env->longitude, env->latitude, env->water, env->base_npp, env->temperature,
                  env->rainfall, env->ice, env->coastal};
        CapParams c{P.npp_water_factor, P.npp_coastal_factor, P.npp_coastal_min_lat, P.npp_coastal_max_lat,
                    P.npp_min, P.npp_max, P.npp_k_min, P.npp_k_max, P.npp_efficiency};
        hipLaunchKernelGGL(k_capacity, dim3(nb), dim3(kBlock), 0, st, x, e, c);
        HIPCHK(hipGetLastError());
    }
    if (has_me) {
        const qhg_params &P = *params;
        for (int ev = 0; ev < P.me_num; ev++) {
            const double *in = nullptr;
            switch (P.me_input[ev]) {
            case QHG_EV_ALTITUDE: in = env->altitude; break;
            case QHG_EV_CAPACITY: in = K_out; break;
            case QHG_EV_BASE_NPP: in = env->base_npp; break;
            }
            if (!in) return fail(QHG_ERR_ARG, "evaluator input array missing");
            if (P.me_pl_n[ev] < 0 || P.me_pl_n[ev] > QHG_MAX_PL_POINTS)
                return fail(QHG_ERR_ARG, "PolyLine point count out of range");
            EvalParams ep{};
            ep.mode = P.me_mode;
            ep.cumulate = P.me_cumulate[ev];
            ep.npts = P.me_pl_n[ev];
            ep.first = (ev == 0);
            ep.last = (ev == P.me_num - 1);
            ep.weight = P.me_weight[ev];
            for (int j = 0; j < ep.npts; j++) { ep.px[j] = P.me_pl_x[ev][j]; ep.pv[j] = P.me_pl_v[ev][j]; }
            hipLaunchKernelGGL(k_eval_values, dim3(nb), dim3(kBlock), 0, st, x, in, tmp, ep);
            HIPCHK(hipGetLastError());
            hipLaunchKernelGGL(k_eval_combine, dim3(nb), dim3(kBlock), 0, st, x, (const double *)tmp, ep);
            HIPCHK(hipGetLastError());
        }
    }
    return 0;
}
