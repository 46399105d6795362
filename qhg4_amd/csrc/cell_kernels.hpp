// Device side of the population step engine (gfx950, wave64).
//
// Agents are stored by cell: the live agents of cell c occupy the segment
// [off[c], off[c+1]) of one of two structure-of-arrays buffers (current /
// next, swapped every step).  Inside a segment the order is the one the
// previous step wrote (stayers, then arrivals per incoming neighbour slot,
// then newborns); every order-dependent rule sorts by agent id itself, so
// results do not depend on it.  A step is three launches (four with
// Genetics), none with global atomics on agent data:
//
//   k_cell_act      one wave per occupied cell: RandomPair::initialize /
//                   findMates on the cell's segment (QHG4/actions/
//                   RandomPair.cpp:107-279), then every Action<T>::execute of
//                   the step fused in priority order (QHG4/core/
//                   SPopulation.cpp:554-577).  Writes each agent's outcome
//                   (dead / stays / moves to neighbour slot d / gave birth)
//                   and per cell the counts of stayers, movers per slot and
//                   births.
//   k_cell_scan     per cell: next count = stayers + arrivals + births; one
//                   decoupled look-back scan gives the next segment offsets,
//                   the birth offsets (new ids), and the absolute position of
//                   every incoming (source cell, slot) run
//                   (updateNumAgentsPerCell, SPopulation.cpp:509-536).
//   k_cell_scatter  one wave per occupied source cell: stayers and movers
//                   pushed to their next positions, newborns of the cell's
//                   mothers appended in mother-id order (performMoves,
//                   recycleDeadSpaceNew, createAgentAtIndex,
//                   makePopSpecificOffspring; SPopulation.cpp:596-918).
//   k_genome        one thread per birth: Genetics::makeOffspring
//                   (Genetics.cpp:285-337) into the genome pool.
//
// Per-cell tables (count, K, weights, neighbours) are read once per cell.
// Genomes live in a pool addressed by a per-agent handle, so they are only
// touched at births and at occasional pool compactions.
#pragma once
#include <cstddef>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"
#include "detmath.hpp"
#include "protein.hpp"
#include "../../include/qhg_gpu.h"

namespace qhg {

constexpr int kWave = 64;
constexpr int kCellBlock = 64;        // per-cell kernels: one wave per block
constexpr int kLdsCell = 256;         // cells up to this size keep their lists in LDS,
                                      // larger ones in segment-indexed global scratch
constexpr int kGroupSlots = 256;      // agents of one packed group of cells
constexpr int kGroupCells = 16;       // cells of one packed group
constexpr int kGroupCellMax = 64;     // largest cell the packed pairing takes
constexpr int kMaxIn = 8;             // incoming neighbour slots per cell
constexpr int kScanBlock = 256;       // k_cell_scan threads per block
constexpr int kScanPer = 1;           // cells per k_cell_scan thread
constexpr int kCellTile = kScanBlock * kScanPer;   // cells per look-back tile
constexpr int kMaxCross = 32;         // device limit on Genetics_num_crossover
constexpr int kStatusRing = 8;        // host-mapped status snapshots
constexpr int kStatusSealEvery = 4;   // steps whose snapshot k_cell_scatter seals for the host

constexpr int32_t kErrCapacity = 1;   // the next population does not fit
constexpr int32_t kErrLookback = 4;   // a look-back wait exceeded its spin limit
constexpr int32_t kErrGather = 8;     // an internal count disagreed
constexpr uint32_t kLookbackSpins = 1u << 22;
constexpr int kParts = 3;             // per-block partial sums of k_cell_act

// outcome byte of an agent in a step
constexpr uint8_t kDecBirth = 0x80;   // registered a birth
constexpr uint8_t kDecDead = 0x40;    // died
constexpr uint8_t kDecDirMask = 0x07; // 0 stays, d+1 moves to neighbour slot d,
constexpr uint8_t kDecNav = 0x07;     //   kDecNav moves along a Navigate edge
constexpr int kNavDir = kDecNav - 1;  // the move direction of a navigating mover

struct DevStatus {
    int64_t n_alive[2];      // live agents at the start of a step of this parity
    int64_t n_occ[2];        // occupied cells at the start of a step of this parity
    int64_t max_id[2];
    int64_t gpool[2];        // genome pool high-water mark
    int64_t step;            // steps completed
    double t;
    int64_t births_last, deaths_last, moves_last;
    int64_t births_total, deaths_total, moves_total;
    int64_t agent_steps, fertile_total, occupied_total;
    int64_t acc_deaths, acc_moves;   // batched steps: this step's counts so far
    int32_t error, pad;
};

// an occupied cell of a step: index, segment start, agent count
struct OccCell {
    int32_t c;
    uint32_t s;
    uint32_t n;
    uint32_t pad;
};

struct AgentBuf {
    int64_t *id;
    int64_t *mate_id;        // partner id of this step's pairing, -3 when unpaired
    uint8_t *life;
    uint8_t *gender;
    double *birth;
    double *age;
    double *lastb;
    int32_t *nbabies;
    int32_t *ghandle;        // genome pool slot (Genetics only)
};

// Navigate's tables on the device (n == 0: no Navigate, or no edge).  Every
// cell's edges are its sea-way jumps (the port's table, destination order)
// followed by one edge per current bridge end at the cell (bridge order).
struct NavDev {
    int32_t n;                // edges
    const int32_t *off;       // ncells + 1: each cell's edges
    const int32_t *njump;     // ncells: the jump edges among them
    const double *stay;       // ncells: stay-home entry of the port's cumulative table
    const int32_t *dst;       // n: target cell
    const double *cum;        // n: cumulative probability (jump edges)
    const int32_t *bid;       // n: index of the current bridge (bridge edges)
    const int32_t *in_off;    // ncells + 1: edges into each cell
    const int32_t *in_e;      // ... their indices, increasing
    uint32_t *cnt;            // n: movers per edge this step (zero outside a step)
    uint32_t *rpos;           // n: next-buffer slot of each edge's run
    uint64_t *sel;            // per slot: edge | rank << 32 of a navigating mover
};

struct Ctx {
    AgentBuf cur, nxt;
    const uint32_t *off_cur;  // ncells + 1 segment offsets of the current buffer
    uint32_t *off_nxt;
    const OccCell *occ_cur;   // occupied cells of this step, cell order
    OccCell *occ_nxt;         // ... of the next step (written by k_cell_scan)
    int64_t cap;
    int32_t ncells, conn;
    const int32_t *nbr;
    const uint8_t *nn;
    double *K;
    double *W;
    const uint8_t *ice;
    const int32_t *in_e;      // ncells * in_w: incoming runs src*conn+slot, -1 padded
    int32_t in_w;             // <= kMaxIn
    uint8_t *dec;             // per slot of the current buffer: outcome byte
    int32_t *bfather;         // per slot: father slot of a registered birth (Genetics)
    int32_t *brec;            // per birth: mother slot, father slot (Genetics)
    int32_t *gfu, *gmu, *gfs, *gms, *gmate;   // per-slot list scratch of large cells
    int64_t *gfid, *gmid;
    uint8_t *gflag;
    uint32_t *gprot;          // AssortativePair: kProtBudget digests (4 words each) per slot
    int32_t *gnprot;          // ... and the slot's protein count
    uint32_t *cstay, *cbirth;  // per cell: stayers, births (zero outside a step)
    uint32_t *cin;            // ncells * in_w: movers arriving per incoming run
    uint32_t *carr;           // per cell: movers arriving in all (zero outside a step)
    uint32_t *tact;           // per scan tile: cells of it hold agents next step (zero outside a step)
    const int32_t *outk;      // ncells * conn: cin index of each (cell, slot) run,
                              // -1 where the slot takes no moves
    uint32_t *birth_off;      // ncells + 1: births before each cell
    uint32_t *bpos;           // per cell: next-buffer slot of its first newborn
    uint32_t *rpos;           // ncells * conn: next-buffer slot of each (cell, slot) run
    uint64_t *tiles;          // look-back words of k_cell_scan: agents | births,
    uint64_t *tiles2;         //   and occupied cells
    int32_t ntiles;
    uint32_t *tick;           // k_cell_scan's tile tickets (zeroed by k_cell_act)
    int32_t nact_blocks;      // grid of k_cell_act
    uint32_t *wave_part;      // kParts per k_cell_act block: deaths, moves, fertile
    uint64_t *scan_part;      // 3 per k_cell_scan wave: the sums of its slice of wave_part
    int32_t scan_waves;       // waves of k_cell_scan's grid
    uint64_t *gpool;          // genome pool, gwords u64 per handle
    int32_t gwords;
    int64_t gcap;
    DevStatus *st;
    DevStatus *hring;         // host-mapped ring of kStatusRing snapshots
    uint64_t seed;
    NavDev nav;
};

struct StepArgs {
    uint32_t step;            // step index (Philox counter word)
    uint32_t parity;          // step & 1: which status halves are current
    double t;                 // time of the step
    int32_t gcells;           // occupied cells per wave group (1..kGroupCells)
    int32_t dense;            // k_cell_scan reads every cell's incoming runs (many cells occupied)
};

constexpr double kPi = 3.141592653589793;   // Q_PI, QHG4/utils/qhg_consts.h:71

struct Prog {
    int32_t nops;
    int32_t ops[QHG_MAX_ACTIONS];
    uint64_t opword;          // ops packed 4 bits each (op q in bits 4q..4q+3)
    int32_t load_age;         // Fertility runs before GetOld (reads the stored age)
    int32_t has_pair;
    int32_t need_gender, need_birth, need_lastb, need_act0, need_act1;
    double fert_min, fert_max, fert_inter;
    double b0, d0, theta, Kconst;
    double rmove_p, wmove_p;
    double oad_max, oad_unc;                // OldAgeDeath
    double atan_max, atan_slope, atan_scale;  // ATanDeath (scale from ATanDeath::preLoop)
    int32_t gen_size, gen_cross, nbino, pad;
    // AssortativePairHamming: mates come from k_assort_pair (x.gmate)
    int32_t has_assort, assort_permute;
    int32_t assort_words;     // numNucs2Blocks(numBlocks): words per strand position compared
    float assort_cutoff;
    int32_t assort_prot;      // AssortativePair: matches counted over protein digests
    double nav_bridge_p;      // Navigate_bridge_prob
    double bino[QHG_MAX_BINO];
};

// Ordering point between the lanes of a single-wave work-group: LDS (and a
// wave's own global) accesses execute in program order, so only the
// compiler must not move memory operations across; unlike __syncthreads it
// does not wait for loads still in flight.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, k);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)v, d, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

// index of the group cell holding group position q; eb holds the cells'
// first positions in lanes 0..ng-1 (increasing)
__device__ __forceinline__ int group_cell_of(uint32_t q, uint32_t eb, int ng) {
    int k = 0;
    for (int kk = 1; kk < ng; kk++)
        if (q >= readlane_u32(eb, kk)) k = kk;
    return k;
}

// ---------------------------------------------------------------------------
// wave / block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, d, 64);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), d, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t)v, d, 64);
        const uint32_t hi = __shfl_up((uint32_t)(v >> 32), d, 64);
        if (lane >= d) v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// position of the k-th (0-based) set bit of v (k < popcount(v))
__device__ __forceinline__ int select_bit(uint64_t v, int k) {
    int pos = 0;
    int c = __popc((uint32_t)v);
    if (k >= c) { k -= c; v >>= 32; pos += 32; }
    c = __popc((uint32_t)v & 0xFFFFu);
    if (k >= c) { k -= c; v >>= 16; pos += 16; }
    c = __popc((uint32_t)v & 0xFFu);
    if (k >= c) { k -= c; v >>= 8; pos += 8; }
    c = __popc((uint32_t)v & 0xFu);
    if (k >= c) { k -= c; v >>= 4; pos += 4; }
    c = __popc((uint32_t)v & 0x3u);
    if (k >= c) { k -= c; v >>= 2; pos += 2; }
    c = (int)(v & 1u);
    if (k >= c) { pos += 1; }
    return pos;
}

// ---------------------------------------------------------------------------
// decoupled look-back, one 64-bit word per tile: [63:62] flag (1 aggregate,
// 2 inclusive), [61:0] value.  Only the tile words are exchanged between
// blocks, so relaxed agent-scope atomics suffice.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ __forceinline__ void tile_publish(uint64_t *w, uint64_t v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t tile_read(uint64_t *w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back over a pair of words per tile (two independent sums).  Called
// by all 64 lanes of one wave; returns the exclusive prefixes of `tile`.
// Each round trip reads kLookWin windows of 64 predecessors at once, nearest
// first, and sums back to the nearest inclusive word (tiles before the first
// read as inclusive zeros), so up to 64 * kLookWin predecessors resolve in
// one read once their aggregates are out.  A predecessor counts only when
// both of its words carry the same flag (the publisher writes them one after
// the other).  Every wait is bounded: past the spin limit the step is
// flagged as failed.
constexpr int kLookWin = 4;

__device__ void lookback2(uint64_t *t1, uint64_t *t2, uint32_t tile, uint64_t agg1, uint64_t agg2, int32_t *err,
                          uint64_t &ex1, uint64_t &ex2) {
    const int lane = threadIdx.x & 63;
    ex1 = 0;
    ex2 = 0;
    if (tile == 0) {
        if (lane == 0) {
            tile_publish(&t1[0], kFlagInc | agg1);
            tile_publish(&t2[0], kFlagInc | agg2);
        }
        return;
    }
    if (lane == 0) {
        tile_publish(&t1[tile], kFlagAgg | agg1);
        tile_publish(&t2[tile], kFlagAgg | agg2);
    }
    int64_t top = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (true) {
        uint64_t w1[kLookWin], w2[kLookWin];
#pragma unroll
        for (int m = 0; m < kLookWin; m++) {
            const int64_t j = top - lane - 64 * m;
            w1[m] = (j >= 0) ? tile_read(&t1[j]) : kFlagInc;
            w2[m] = (j >= 0) ? tile_read(&t2[j]) : kFlagInc;
        }
        uint64_t s1 = 0, s2 = 0;
        bool wait = false, done = false;
#pragma unroll
        for (int m = 0; m < kLookWin; m++) {
            const uint64_t f1 = w1[m] & ~kValMask, f2 = w2[m] & ~kValMask;
            const bool ready = f1 == f2 && f1 != 0;
            const uint64_t inc = __ballot(ready && f1 == kFlagInc);
            const uint64_t nready = __ballot(!ready);
            const int first = inc ? (__ffsll((unsigned long long)inc) - 1) : 64;
            const uint64_t need = (first >= 63) ? ~0ull : ((2ull << first) - 1ull);
            if (nready & need) {
                wait = true;
                break;
            }
            s1 += wave_sum_u64(lane <= first ? (w1[m] & kValMask) : 0ull);
            s2 += wave_sum_u64(lane <= first ? (w2[m] & kValMask) : 0ull);
            if (first < 64) {
                done = true;
                break;
            }
        }
        if (wait) {
            if (++spins < kLookbackSpins) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            if (lane == 0) atomicOr(err, kErrLookback);
            break;
        }
        ex1 += s1;
        ex2 += s2;
        if (done) break;
        top -= 64 * kLookWin;
    }
    if (lane == 0) {
        tile_publish(&t1[tile], kFlagInc | (ex1 + agg1));
        tile_publish(&t2[tile], kFlagInc | (ex2 + agg2));
    }
}

// ---------------------------------------------------------------------------
// one agent's action program
// ---------------------------------------------------------------------------
struct ActItem {
    uint8_t life, g;
    int64_t id;
    int32_t mate;             // partner's local index in the cell, -1 when unpaired
    double birth, lastb, age;
    bool age_dirty;
    uint32_t bflag;
    int32_t npend;            // registered moves (<= 2), resolved after the program
    int32_t navq;             // 1 + moves registered before Navigate ran, 0: it did not
    int32_t pk0, pk1;         // their action kinds
    uint32_t pu0, pu1;        // their target draws
    uint32_t u[8];            // words 0..7 of the agent's QHG_STREAM_ACT counter
};

struct CellView {             // the per-cell values every agent of the cell reads
    uint32_t N;               // live agents at step start (getNumAgents)
    double K;
    int nn;
};

// Every action of the program in priority order (compute only; the caller
// loaded every field).  A dead agent stops.  Moves are only registered here
// (their decision draws are taken now) and resolved afterwards, as
// performMoves applies them after all actions.
// kRare: the program may hold OldAgeDeath or ATanDeath; kNav: Navigate
// (separate instantiations keep their code out of the common kernel)
template <bool kRare, bool kNav = false, typename CtxT, typename ProgT>
__device__ __forceinline__ void act_program(const CtxT &x, const ProgT &p, const StepArgs &sa, const CellView &cv,
                                            ActItem &a, uint32_t &deaths) {
    const double t = sa.t;
    if (p.need_act0) draw4(x.seed, sa.step, QHG_STREAM_ACT, (uint64_t)a.id, 0, &a.u[0]);
    if (p.need_act1) draw4(x.seed, sa.step, QHG_STREAM_ACT, (uint64_t)a.id, 1, &a.u[4]);
    uint64_t word = p.opword;
    for (int q = 0; q < p.nops; q++, word >>= 4) {
        const int op = (int)(word & 15u);
        if (op == QHG_ACT_GETOLD) {
            // GetOld::execute, QHG4/actions/GetOld.cpp:37-48
            a.age = t - a.birth;
            a.age_dirty = true;
        } else if (op == QHG_ACT_FERTILITY) {
            // Fertility::execute, QHG4/actions/Fertility.cpp:49-74
            if (a.g == 0) {
                a.life = (a.age > p.fert_min && a.age < p.fert_max && (t - a.lastb) > p.fert_inter)
                             ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
            } else {
                a.life = (a.age > p.fert_min) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
            }
        } else if (op == QHG_ACT_VERHULST || op == QHG_ACT_VERHULSTVARK) {
            // LinearBirth::initialize/execute then LinearDeath (Verhulst.cpp:101-115;
            // LinearBirth.cpp:90-168; LinearDeath.cpp:92-153)
            const double N = (double)cv.N;
            double b, d;
            if (op == QHG_ACT_VERHULST) {
                b = p.b0 + (p.theta - p.b0) * (N / p.Kconst);
                d = p.d0 + (p.theta - p.d0) * (N / p.Kconst);
            } else {
                b = (cv.K <= 0) ? 0.0 : p.b0 + (p.theta - p.b0) * (N / cv.K);
                d = (cv.K <= 0) ? 1.0 : p.d0 + (p.theta - p.d0) * (N / cv.K);
            }
            if (b > 0) {
                if (a.g == 0 && a.mate >= 0) {
                    if (wrandd(a.u[QHG_ACTW_LBIRTH]) < b) a.bflag = 1;
                }
            } else if (b < 0) {
                if (wrandd(a.u[QHG_ACTW_LBIRTH]) < -b) a.life = 0;
            }
            if (a.life != 0) {
                if (wrandd(a.u[QHG_ACTW_LDEATH]) < d) a.life = 0;
            }
            if (a.life == 0) { deaths++; break; }
        } else if (op == QHG_ACT_RANDOMMOVE || op == QHG_ACT_WEIGHTEDMOVE) {
            // RandomMove::execute (RandomMove.cpp:65-100) / WeightedMove::execute
            // (WeightedMove.cpp:45-106): decision draw now, target below
            const bool rm = op == QHG_ACT_RANDOMMOVE;
            const uint32_t go = rm ? a.u[QHG_ACTW_RMOVE_GO] : a.u[QHG_ACTW_WMOVE_GO];
            const uint32_t dir = rm ? a.u[QHG_ACTW_RMOVE_DIR] : a.u[QHG_ACTW_WMOVE_DIR];
            if (wrandd(go) < (rm ? p.rmove_p : p.wmove_p)) {
                if (a.npend == 0) { a.pk0 = op; a.pu0 = dir; a.npend = 1; }
                else if (a.npend == 1) { a.pk1 = op; a.pu1 = dir; a.npend = 2; }
            }
        } else if (kRare && op == QHG_ACT_OLDAGEDEATH) {
            // OldAgeDeath::execute: age, then death past max age + U(1-uM, 1+uM)
            a.age = t - a.birth;
            a.age_dirty = true;
            const double r = wrandr(1 - p.oad_unc * p.oad_max, 1 + p.oad_unc * p.oad_max, a.u[QHG_ACTW_OADEATH]);
            if (a.age > p.oad_max + r) { a.life = 0; deaths++; break; }
        } else if (kNav && op == QHG_ACT_NAVIGATE) {
            // Navigate::execute runs once the moves registered so far are
            // known (below): only an agent none of them moves navigates
            a.navq = a.npend + 1;
        } else if (kRare && op == QHG_ACT_ATANDEATH) {
            // ATanDeath::execute (ATanDeath.cpp:66-90): age, then death with
            // probability 1/2 + scale * atan(slope * (age - max age)) / pi
            a.age = t - a.birth;
            a.age_dirty = true;
            const double pd = 0.5 + p.atan_scale * det_atan(p.atan_slope * (a.age - p.atan_max)) / kPi;
            if (wrandd(a.u[QHG_ACTW_ATANDEATH]) < pd) { a.life = 0; deaths++; break; }
        }
        // RandomPair, MultiEvaluator, NPPCapacity, Genetics: no execute
    }
}

// Neighbour slot a registered move lands on, or -1 when the agent stays
// (slot 0, a -1 hole, or ice for WeightedMove).  w: the cell's cumulative
// weight row, nbr: its neighbour row (registers, compile-time indices only).
template <typename CtxT>
__device__ __forceinline__ int move_dir(const CtxT &x, int kind, uint32_t u, int nn, const double (&w)[7],
                                        const int32_t (&nbr)[6]) {
    int idx = -1;
    if (kind == QHG_ACT_RANDOMMOVE) {
        idx = (int)(wrandd(u) * (nn + 1));
    } else {
        double wl = w[0];
#pragma unroll
        for (int q = 1; q < 7; q++) if (q == nn) wl = w[q];
        if (w[0] == wl) {
            idx = (int)wrandi(0, (uint32_t)nn + 1, 1, u);
        } else {
            const double r2 = wrandd(u) * wl;
#pragma unroll
            for (int q = 6; q >= 0; q--) {
                if (q <= nn && r2 < w[q]) idx = q;   // first q with r2 < w[q]
            }
        }
    }
    if (idx <= 0) return -1;
    int32_t nb = -1;
#pragma unroll
    for (int q = 0; q < 6; q++) if (q == idx - 1) nb = nbr[q];
    if (nb < 0) return -1;
    if (kind == QHG_ACT_WEIGHTEDMOVE && x.ice != nullptr && x.ice[nb]) return -1;
    return idx - 1;
}

// Navigate::execute (QHG4/actions/Navigate.cpp:187-244) for an agent at
// cell c, alive and with no move registered: from a port, walk the
// cumulative table while the draw exceeds the entry, at most up to index c
// (the reference bounds the walk by the map key, the port's cell index) and
// at most to the last destination; entry i > 0 registers a move to its cell
// unless that cell is ice.  Then every current bridge with an end at c
// takes a draw and registers a move to the other end when it is below
// Navigate_bridge_prob.  Each registration counts as a move; the last wins.
template <typename CtxT, typename ProgT>
__device__ void nav_execute(const CtxT &x, const ProgT &p, const StepArgs &sa, int64_t id, int32_t c, int32_t e0,
                            int32_t e1, int &dir, int32_t &nave, uint32_t &moves) {
    if (e0 == e1) return;
    const int32_t nj = x.nav.njump[c];
    if (nj > 0) {
        const double r = wrandd(draw(x.seed, sa.step, QHG_STREAM_NAV, (uint64_t)id, 0));
        int32_t i = 0;
        double cum = x.nav.stay[c];
        while (i < c && i < nj && r > cum) {
            cum = x.nav.cum[e0 + i];
            i++;
        }
        if (i > 0) {
            const int32_t e = e0 + i - 1;
            const int32_t d = x.nav.dst[e];
            if (x.ice == nullptr || !x.ice[d]) {
                dir = kNavDir;
                nave = e;
                moves++;
            }
        }
    }
    for (int32_t e = e0 + nj; e < e1; e++) {
        const double r = wrandd(draw(x.seed, sa.step, QHG_STREAM_NAV, (uint64_t)id, 1u + (uint32_t)x.nav.bid[e]));
        if (r < p.nav_bridge_p) {
            dir = kNavDir;
            nave = e;
            moves++;
        }
    }
}

// The registered moves in program order, each counting as a move and the
// last one winning (performMoves applies them in order,
// QHG4/core/SPopulation.cpp:1058-1092).  Navigate takes its place in the
// program and acts only when no move was registered before it
// (LIFE_STATE_MOVING, Navigate.cpp:191).  w, nbr: the cell's weight and
// neighbour rows; ne0, ne1: its Navigate edges.
template <bool kNav, typename CtxT, typename ProgT>
__device__ __forceinline__ void resolve_moves(const CtxT &x, const ProgT &p, const StepArgs &sa, const ActItem &a,
                                              int32_t c, int32_t ne0, int32_t ne1, int nn, const double (&w)[7],
                                              const int32_t (&nbr)[6], int &dir, int32_t &nave, uint32_t &moves) {
    if constexpr (!kNav) {
        int d = move_dir(x, a.pk0, a.pu0, nn, w, nbr);
        if (d >= 0) { dir = d; moves++; }
        if (a.npend > 1) {
            d = move_dir(x, a.pk1, a.pu1, nn, w, nbr);
            if (d >= 0) { dir = d; moves++; }
        }
    } else {
        // the neighbour moves' directions first (each depends on its own
        // draw only), so the weight and neighbour rows are no longer live
        // across Navigate's draws; then the registrations in program order,
        // with one copy of Navigate's code, not one per place it can take
        const int d0 = (a.npend > 0) ? move_dir(x, a.pk0, a.pu0, nn, w, nbr) : -1;
        const int d1 = (a.npend > 1) ? move_dir(x, a.pk1, a.pu1, nn, w, nbr) : -1;
#pragma unroll 1
        for (int q = 0; q <= a.npend; q++) {
            if (a.navq == q + 1 && dir < 0) nav_execute(x, p, sa, a.id, c, ne0, ne1, dir, nave, moves);
            const int d = (q == 0) ? d0 : d1;
            if (q < a.npend && d >= 0) { dir = d; moves++; }
        }
    }
}

// Places one chunk's navigating movers in their edges' runs (called by all
// lanes).  One wave owns every edge of its cells and meets a cell's agents
// in slot order, chunk by chunk, so the movers of an edge rank in slot
// order: after those of earlier chunks (the edge's count so far), then in
// lane order.  Each mover's edge and rank go to nav.sel for k_cell_scatter;
// the edge's count (nav.cnt, read by k_cell_scan) and the target cell's
// arrivals grow by the chunk's movers.
template <typename CtxT>
__device__ __forceinline__ void nav_movers(const CtxT &x, bool mine, int32_t nave, uint32_t slot) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    uint64_t left = __ballot(mine);
    if (!left) return;
    // the movers sharing an edge, each edge led by its first lane: the
    // leaders then take their edges' counts in one atomic instruction (one
    // round trip for the chunk, not one per edge)
    int leader = lane;
    uint32_t rank = 0, cnt = 0;
    while (left) {
        const int l0 = __ffsll((unsigned long long)left) - 1;
        const int32_t e = __shfl(nave, l0, 64);
        const uint64_t same = __ballot(mine && nave == e);
        if ((same >> lane) & 1ull) {
            leader = l0;
            rank = (uint32_t)__popcll(same & lt);
        }
        if (lane == l0) cnt = (uint32_t)__popcll(same);
        left &= ~same;
    }
    uint32_t base = 0;
    if (mine && leader == lane) {
        base = atomicAdd(&x.nav.cnt[nave], cnt);
        const int32_t d = x.nav.dst[nave];
        atomicAdd(&x.carr[d], cnt);
        x.tact[d / kCellTile] = 1u;
    }
    base = (uint32_t)__shfl((int)base, leader, 64);
    if (mine) x.nav.sel[slot] = (uint64_t)(uint32_t)nave | ((uint64_t)(base + rank) << 32);
}

// ---------------------------------------------------------------------------
// k_cell_act: pairing + fused actions, one wave per occupied cell
// ---------------------------------------------------------------------------
struct PairLists {            // per-cell lists, in LDS or in global scratch
    int32_t *fu, *mu;         // local indices of FERTILE females / males, slot order
    int64_t *fid, *mid;       // their ids
    int32_t *fs, *ms;         // id order: positions into fu/mu
    int32_t *mate;            // partner's local index, -1 unpaired
    uint8_t *flag;            // still unpaired (pool side)
};

// rank of each entry of ids[0..n) among them (ids are distinct) -> ord[rank]
__device__ __forceinline__ void id_order(const int64_t *ids, int32_t n, int32_t *ord) {
    const int lane = threadIdx.x;
    for (int32_t j = lane; j < n; j += kWave) {
        const int64_t my = ids[j];
        int32_t r = 0;
        for (int32_t k = 0; k < n; k++) r += ids[k] < my;
        ord[r] = j;
    }
}

template <bool kRare, bool kAssort = false, bool kNav = false, typename CtxT, typename ProgT>
__device__ __forceinline__ void cell_act(const CtxT &x, const ProgT &p, const StepArgs &sa, int32_t c, uint32_t s,
                                         int32_t n, const PairLists &L, uint32_t &my_deaths, uint32_t &my_moves,
                                         uint32_t &my_fert) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int R = x.conn + 1;
    const int nchunks = (n + kWave - 1) / kWave;

    // large cells (rare): fields are loaded per chunk in each pass
    CellView cv;
    cv.N = (uint32_t)n;
    cv.K = x.K[c];
    cv.nn = x.nn[c];

    // ---- RandomPair: FERTILE females and males of the cell
    if (p.has_pair) {
        int32_t F = 0, M = 0;
        for (int ch = 0; ch < nchunks; ch++) {
            const int32_t li = ch * kWave + lane;
            const bool valid = li < n;
            uint8_t l = 0, g = 2;
            int64_t id = 0;
            if (valid) {
                l = x.cur.life[s + li];
                g = x.cur.gender[s + li];
                id = x.cur.id[s + li];
            }
            const bool isF = valid && l == QHG_LIFE_FERTILE && g == 0;
            const bool isM = valid && l == QHG_LIFE_FERTILE && g == 1;
            const uint64_t mF = __ballot(isF), mM = __ballot(isM);
            if (isF) { const int32_t at = F + __popcll(mF & lt); L.fu[at] = li; L.fid[at] = id; }
            if (isM) { const int32_t at = M + __popcll(mM & lt); L.mu[at] = li; L.mid[at] = id; }
            F += __popcll(mF);
            M += __popcll(mM);
            if (valid) L.mate[li] = -1;
        }
        my_fert += (lane == 0) ? (uint32_t)(F + M) : 0u;
        if (F > 0 && M > 0) {
            // findMates, RandomPair.cpp:183-276: the smaller sex chooses in id
            // order, each pick the k-th (id order) still unpaired of the other
            wave_sync();
            id_order(L.fid, F, L.fs);
            id_order(L.mid, M, L.ms);
            const bool fsel = F <= M;
            const int32_t *chs = fsel ? L.fs : L.ms;
            const int32_t *chl = fsel ? L.fu : L.mu;
            const int64_t *chid = fsel ? L.fid : L.mid;
            const int32_t *pool = fsel ? L.ms : L.fs;
            const int32_t *pll = fsel ? L.mu : L.fu;
            const int32_t nch = fsel ? F : M;
            const int32_t np = fsel ? M : F;
            for (int32_t j = lane; j < np; j += kWave) L.flag[j] = 1;
            wave_sync();
            int32_t left = np;
            for (int32_t cb = 0; cb < nch; cb += kWave) {
                const int32_t j = cb + lane;
                uint32_t u = 0;
                if (j < nch) u = draw(x.seed, sa.step, QHG_STREAM_PAIR, (uint64_t)chid[chs[j]], 0);
                const int32_t cnt = (nch - cb < kWave) ? (nch - cb) : kWave;
                for (int32_t q = 0; q < cnt && left > 0; q++) {
                    const uint32_t uq = readlane_u32(u, q);
                    int k = (int)wrandr(0, (double)left, uq);
                    int32_t pos = -1;
                    for (int32_t base = 0; base < np; base += kWave) {
                        const bool f = base + lane < np && L.flag[base + lane];
                        const uint64_t m = __ballot(f);
                        const int pc = __popcll(m);
                        if (k < pc) { pos = base + select_bit(m, k); break; }
                        k -= pc;
                    }
                    if (pos < 0) {   // cannot happen while left counts the set flags
                        if (lane == 0) atomicOr(&x.st->error, kErrGather);
                        break;
                    }
                    const int32_t a = chl[chs[cb + q]], b = pll[pool[pos]];
                    if (lane == 0) {
                        L.flag[pos] = 0;
                        L.mate[a] = b;
                        L.mate[b] = a;
                    }
                    wave_sync();
                    left--;
                }
            }
        }
        wave_sync();
    }

    // ---- actions, 64 agents at a time
    uint32_t stay = 0, births = 0;
    uint32_t mv[6] = {0, 0, 0, 0, 0, 0};
    // the cell's Navigate edges (a population without tables, e.g. another
    // member of a batch, has none: its table pointers are null)
    const bool nav_tab = kNav && x.nav.n != 0;
    const int32_t cne0 = nav_tab ? x.nav.off[c] : 0, cne1 = nav_tab ? x.nav.off[c + 1] : 0;
    for (int ch = 0; ch < nchunks; ch++) {
        const int32_t li = ch * kWave + lane;
        const bool valid = li < n;
        const uint32_t slot = s + (uint32_t)li;
        ActItem a;
        a.life = 0;
        a.g = 0;
        a.id = 0;
        a.birth = 0;
        a.lastb = 0;
        a.age = 0;
        a.mate = -1;
        a.age_dirty = false;
        a.bflag = 0;
        a.npend = 0;
        a.navq = 0;
        a.pk0 = a.pk1 = 0;
        a.pu0 = a.pu1 = 0;
        if (valid) {
            a.life = x.cur.life[slot];
            a.id = x.cur.id[slot];
            if (p.need_gender) a.g = x.cur.gender[slot];
            if (p.need_birth) a.birth = x.cur.birth[slot];
            if (p.need_lastb) a.lastb = x.cur.lastb[slot];
            if (p.load_age) a.age = x.cur.age[slot];
        }
        if (valid && p.has_pair) a.mate = L.mate[li];
        if (kAssort && valid) {
            const int32_t ps = x.gmate[slot];
            a.mate = (ps >= 0) ? (int32_t)((uint32_t)ps - s) : -1;
            my_fert += (a.life == QHG_LIFE_FERTILE && a.g <= 1) ? 1u : 0u;
        }
        const uint8_t lifeS = a.life;
        int dir = -1;
        int32_t nave = 0;
        if (valid && a.life != 0) {
            act_program<kRare, kNav>(x, p, sa, cv, a, my_deaths);
            if (a.npend > 0 || (kNav && a.navq)) {
                // the cell's neighbour and weight rows, for movers only
                int32_t nbr[6];
                double w[7];
#pragma unroll
                for (int qq = 0; qq < 6; qq++) nbr[qq] = (qq < x.conn) ? x.nbr[(int64_t)c * x.conn + qq] : -1;
#pragma unroll
                for (int qq = 0; qq < 7; qq++) w[qq] = (qq < R) ? x.W[(int64_t)c * R + qq] : 0.0;
                resolve_moves<kNav>(x, p, sa, a, c, cne0, cne1, cv.nn, w, nbr, dir, nave, my_moves);
            }
        }
        const bool alive = valid && a.life != 0;
        if (kNav && x.nav.n) nav_movers(x, alive && dir == kNavDir, nave, slot);
        if (valid) {
            uint8_t code = (uint8_t)(dir >= 0 ? dir + 1 : 0);
            if (!alive) code |= kDecDead;
            if (a.bflag) code |= kDecBirth;
            x.dec[slot] = code;
            if (a.life != lifeS) x.cur.life[slot] = a.life;
            if (a.age_dirty) x.cur.age[slot] = a.age;
            if (p.has_pair) x.cur.mate_id[slot] = (a.mate >= 0) ? x.cur.id[s + a.mate] : -3;
            if (a.bflag) {
                x.cur.nbabies[slot] += 1;
                if (x.gwords) x.bfather[slot] = (int32_t)(s + a.mate);
            }
        }
        stay += __popcll(__ballot(alive && dir < 0));
        births += __popcll(__ballot(valid && a.bflag));
#pragma unroll
        for (int d = 0; d < 6; d++) mv[d] += __popcll(__ballot(alive && dir == d));
    }
    if (lane == 0) {
        x.cstay[c] = stay;
        if (births) x.cbirth[c] = births;
        if (stay | births) x.tact[c / kCellTile] = 1u;
    }
#pragma unroll
    for (int d = 0; d < 6; d++)
        if (lane == d && d < x.conn && mv[d]) {
            const int32_t dest = x.nbr[(int64_t)c * x.conn + d];
            x.cin[x.outk[(int64_t)c * x.conn + d]] = mv[d];
            atomicAdd(&x.carr[dest], mv[d]);
            x.tact[dest / kCellTile] = 1u;
        }
    wave_sync();
}

// ---------------------------------------------------------------------------
// packed groups: several small cells in one wave, one lane per agent
// ---------------------------------------------------------------------------
enum { kCntF = 0, kCntM, kCntStay, kCntBirth, kCntMv, kCntN = kCntMv + 6 };

struct GroupActLds {
    // FERTILE lists of cell k at positions [eb_k, eb_k + n_k): females from
    // the front (entry r at eb_k + r), males from the back (at eb_k + n_k-1-r)
    int64_t lid[kGroupSlots];                       // ids of the list entries
    int64_t pid[kGroupSlots];                       // partner id per agent
    double K[kGroupCells];                          // per-cell tables
    double W[kGroupCells][7];
    int32_t nbr[kGroupCells][6];
    int32_t outk[kGroupCells][6];
    uint32_t ud[kGroupSlots];                       // choosers' draws, id rank
    uint32_t cnt[kGroupCells][kCntN];               // per-cell counters
    int16_t list[kGroupSlots];                      // group positions of the entries
    int16_t srt[kGroupSlots];                       // id order: list positions
    int16_t mate[kGroupSlots];                      // partner's group position, -1
    uint8_t nn[kGroupCells];
    int32_t noff[2 * kGroupCells];                  // Navigate: first and end edge per cell
};

struct CellActLds {
    int32_t fu[kLdsCell], mu[kLdsCell], fs[kLdsCell], ms[kLdsCell], mate[kLdsCell];
    int64_t fid[kLdsCell], mid[kLdsCell];
    uint8_t flag[kLdsCell];
};

union ActLds {
    GroupActLds g;
    CellActLds c;
};

// The cells of a group are consecutive occupied cells, so their segments
// form one slot range [S0, S0+T).  Lane k < ng holds cell k's entry e and
// eb = e.s - S0.  Each cell's lists live at its own positions [eb, eb+n).
template <bool kRare, bool kAssort = false, bool kNav = false, typename CtxT, typename ProgT>
__device__ __forceinline__ void group_act(const CtxT &x, const ProgT &p, const StepArgs &sa, int ng, const OccCell &e,
                                          uint32_t eb, uint32_t S0, uint32_t T, GroupActLds &G,
                                          uint32_t &my_deaths, uint32_t &my_moves, uint32_t &my_fert) {
    const int lane = threadIdx.x;
    const int R = x.conn + 1;
    const int nchunks = (int)((T + kWave - 1) / kWave);

    // one round trip: the cells' tables (staged in LDS) and the first
    // chunk's fields
    double tK = 0, tW0 = 0, tW1 = 0;
    int32_t tN0 = -1, tN1 = -1, tO0 = -1, tO1 = -1;
    uint8_t tnn = 0;
    const int wk0 = lane / 7, wj0 = lane % 7, wk1 = (lane + 64) / 7, wj1 = (lane + 64) % 7;
    const int nk0 = lane / 6, nj0 = lane % 6, nk1 = (lane + 64) / 6, nj1 = (lane + 64) % 6;
    const int32_t cw0 = __shfl(e.c, wk0 < ng ? wk0 : 0, 64), cw1 = __shfl(e.c, wk1 < ng ? wk1 : 0, 64);
    const int32_t cn0 = __shfl(e.c, nk0 < ng ? nk0 : 0, 64), cn1 = __shfl(e.c, nk1 < ng ? nk1 : 0, 64);
    int32_t tNo = 0, tNe = 0;
    if (lane < ng) {
        tK = x.K[e.c];
        tnn = x.nn[e.c];
        if (kNav && x.nav.n != 0) {   // no tables (null pointers) when n is 0
            tNo = x.nav.off[e.c];
            tNe = x.nav.off[e.c + 1];
        }
    }
    if (wk0 < ng && wj0 < R) tW0 = x.W[(int64_t)cw0 * R + wj0];
    if (wk1 < ng && wj1 < R) tW1 = x.W[(int64_t)cw1 * R + wj1];
    if (nk0 < ng && nj0 < x.conn) {
        tN0 = x.nbr[(int64_t)cn0 * x.conn + nj0];
        tO0 = x.outk[(int64_t)cn0 * x.conn + nj0];
    }
    if (nk1 < ng && nj1 < x.conn) {
        tN1 = x.nbr[(int64_t)cn1 * x.conn + nj1];
        tO1 = x.outk[(int64_t)cn1 * x.conn + nj1];
    }
    const uint32_t q0 = (uint32_t)lane;
    const bool v0 = q0 < T;
    const int k0 = group_cell_of(q0, eb, ng);
    const uint32_t slot0 = S0 + q0;
    uint8_t life0 = 0, g0 = 2;
    int64_t id0 = 0;
    double birth0 = 0, lastb0 = 0, age0 = 0;
    int32_t nbab0 = 0;
    if (v0) {
        life0 = x.cur.life[slot0];
        g0 = x.cur.gender[slot0];
        id0 = x.cur.id[slot0];
        nbab0 = x.cur.nbabies[slot0];
        if (p.need_birth) birth0 = x.cur.birth[slot0];
        if (p.need_lastb) lastb0 = x.cur.lastb[slot0];
        if (p.load_age) age0 = x.cur.age[slot0];
    }
    // the second chunk's fields in the same round trip (groups hold up to
    // two chunks at the usual target; later chunks load where they are used)
    const uint32_t q1 = (uint32_t)(kWave + lane);
    const bool v1 = q1 < T;
    uint8_t life1 = 0, g1 = 2;
    int64_t id1 = 0;
    double birth1 = 0, lastb1 = 0, age1 = 0;
    int32_t nbab1 = 0;
    if (v1) {
        life1 = x.cur.life[S0 + q1];
        g1 = x.cur.gender[S0 + q1];
        id1 = x.cur.id[S0 + q1];
        nbab1 = x.cur.nbabies[S0 + q1];
        if (p.need_birth) birth1 = x.cur.birth[S0 + q1];
        if (p.need_lastb) lastb1 = x.cur.lastb[S0 + q1];
        if (p.load_age) age1 = x.cur.age[S0 + q1];
    }
    // per-cell tables to LDS: their loads were issued before the fields', so
    // waiting for them leaves the field loads in flight
    if (lane < ng) {
        G.K[lane] = tK;
        G.nn[lane] = tnn;
        if (kNav) {
            G.noff[lane] = tNo;
            G.noff[lane + kGroupCells] = tNe;
        }
#pragma unroll
        for (int q = 0; q < kCntN; q++) G.cnt[lane][q] = 0;
    }
    if (wk0 < ng && wj0 < 7) G.W[wk0][wj0] = tW0;
    if (wk1 < ng && wj1 < 7) G.W[wk1][wj1] = tW1;
    if (nk0 < ng && nj0 < 6) { G.nbr[nk0][nj0] = tN0; G.outk[nk0][nj0] = tO0; }
    if (nk1 < ng && nj1 < 6) { G.nbr[nk1][nj1] = tN1; G.outk[nk1][nj1] = tO1; }
    wave_sync();

    // ---- RandomPair (RandomPair.cpp:107-279), every cell of the group at once
    if (p.has_pair) {
        for (int ch = 0; ch < nchunks; ch++) {
            const uint32_t q = (uint32_t)(ch * kWave + lane);
            const bool valid = q < T;
            const int k = (ch == 0) ? k0 : group_cell_of(q, eb, ng);
            const uint32_t base = (uint32_t)__shfl((int)eb, k, 64);
            const uint32_t nk = (uint32_t)__shfl((int)e.n, k, 64);
            uint8_t l = (ch == 0) ? life0 : life1, g = (ch == 0) ? g0 : g1;
            int64_t id = (ch == 0) ? id0 : id1;
            if (ch > 1 && valid) {
                l = x.cur.life[S0 + q];
                g = x.cur.gender[S0 + q];
                id = x.cur.id[S0 + q];
            }
            if (valid) {
                G.mate[q] = -1;
                if (l == QHG_LIFE_FERTILE && g == 0) {
                    const uint32_t r = atomicAdd(&G.cnt[k][kCntF], 1u);
                    G.list[base + r] = (int16_t)q;
                    G.lid[base + r] = id;
                } else if (l == QHG_LIFE_FERTILE && g == 1) {
                    const uint32_t r = atomicAdd(&G.cnt[k][kCntM], 1u);
                    G.list[base + nk - 1 - r] = (int16_t)q;
                    G.lid[base + nk - 1 - r] = id;
                }
            }
        }
        wave_sync();
        // id order inside each cell's lists; the choosing sex (the smaller,
        // females on a tie) takes its pick draw now
        for (int ch = 0; ch < nchunks; ch++) {
            const uint32_t q = (uint32_t)(ch * kWave + lane);
            const bool valid = q < T;
            const int k = (ch == 0) ? k0 : group_cell_of(q, eb, ng);
            const uint32_t base = (uint32_t)__shfl((int)eb, k, 64);
            const uint32_t nk = (uint32_t)__shfl((int)e.n, k, 64);
            if (!valid) continue;
            const uint32_t Fk = G.cnt[k][kCntF], Mk = G.cnt[k][kCntM];
            const uint32_t t = q - base;
            if (t < Fk) {
                const int64_t my = G.lid[q];
                uint32_t r = 0;
                for (uint32_t u = 0; u < Fk; u++) r += G.lid[base + u] < my;
                G.srt[base + r] = (int16_t)q;
                if (Fk <= Mk) G.ud[base + r] = draw(x.seed, sa.step, QHG_STREAM_PAIR, (uint64_t)my, 0);
            } else if (nk - 1 - t < Mk) {
                const int64_t my = G.lid[q];
                const uint32_t top = base + nk - 1;
                uint32_t r = 0;
                for (uint32_t u = 0; u < Mk; u++) r += G.lid[top - u] < my;
                G.srt[top - r] = (int16_t)q;
                if (Mk < Fk) G.ud[base + r] = draw(x.seed, sa.step, QHG_STREAM_PAIR, (uint64_t)my, 0);
            }
        }
        wave_sync();
        // findMates (RandomPair.cpp:183-276): lane k picks for cell k, each
        // pick the k-th (id order) still unpaired of the other sex
        if (lane < ng) {
            const uint32_t Fk = G.cnt[lane][kCntF], Mk = G.cnt[lane][kCntM];
            my_fert += Fk + Mk;
            if (Fk > 0 && Mk > 0) {
                const bool fsel = Fk <= Mk;
                const uint32_t nchs = fsel ? Fk : Mk, np = fsel ? Mk : Fk;
                const uint32_t top = eb + e.n - 1;
                uint64_t mask = (np >= 64) ? ~0ull : ((1ull << np) - 1ull);
                uint32_t left = np;
                for (uint32_t q = 0; q < nchs; q++) {
                    const int kp = (int)wrandr(0, (double)left, G.ud[eb + q]);
                    const int pos = select_bit(mask, kp);
                    mask &= ~(1ull << pos);
                    const int16_t la = G.srt[fsel ? eb + q : top - q];
                    const int16_t lb = G.srt[fsel ? top - pos : eb + pos];
                    const int16_t ja = G.list[la], jb = G.list[lb];
                    G.mate[ja] = jb;
                    G.mate[jb] = ja;
                    G.pid[ja] = G.lid[lb];
                    G.pid[jb] = G.lid[la];
                    left--;
                }
            }
        }
    }
    wave_sync();

    // ---- actions, one lane per agent
    for (int ch = 0; ch < nchunks; ch++) {
        const uint32_t q = (uint32_t)(ch * kWave + lane);
        const bool valid = q < T;
        const uint32_t slot = S0 + q;
        const int k = (ch == 0) ? k0 : group_cell_of(q, eb, ng);
        const uint32_t nk = (uint32_t)__shfl((int)e.n, k, 64);
        ActItem a;
        a.life = (ch == 0) ? life0 : life1;
        a.g = (ch == 0) ? g0 : g1;
        a.id = (ch == 0) ? id0 : id1;
        a.birth = (ch == 0) ? birth0 : birth1;
        a.lastb = (ch == 0) ? lastb0 : lastb1;
        a.age = (ch == 0) ? age0 : age1;
        a.mate = -1;
        a.age_dirty = false;
        a.bflag = 0;
        a.npend = 0;
        a.navq = 0;
        a.pk0 = a.pk1 = 0;
        a.pu0 = a.pu1 = 0;
        int32_t nbab = (ch == 0) ? nbab0 : nbab1;
        if (ch > 1) {
            a.life = 0;
            if (valid) {
                a.life = x.cur.life[slot];
                a.id = x.cur.id[slot];
                nbab = x.cur.nbabies[slot];
                if (p.need_gender) a.g = x.cur.gender[slot];
                if (p.need_birth) a.birth = x.cur.birth[slot];
                if (p.need_lastb) a.lastb = x.cur.lastb[slot];
                if (p.load_age) a.age = x.cur.age[slot];
            }
        }
        CellView cv;
        cv.N = nk;
        cv.K = G.K[k];
        cv.nn = G.nn[k];
        if (valid && p.has_pair) a.mate = G.mate[q];
        if (kAssort && valid) {
            // the partner's slot from k_assort_pair, as a group position;
            // the FERTILE agents it considered are counted here
            const int32_t ps = x.gmate[slot];
            a.mate = (ps >= 0) ? (int32_t)((uint32_t)ps - S0) : -1;
            my_fert += (a.life == QHG_LIFE_FERTILE && a.g <= 1) ? 1u : 0u;
        }
        const uint8_t lifeS = a.life;
        int dir = -1;
        int32_t nave = 0;
        const int32_t cc = kNav ? __shfl(e.c, k, 64) : 0;   // the agent's cell (Navigate)
        if (valid && a.life != 0) {
            act_program<kRare, kNav>(x, p, sa, cv, a, my_deaths);
            if (a.npend > 0 || (kNav && a.navq)) {
                int32_t nbr[6];
                double w[7];
#pragma unroll
                for (int qq = 0; qq < 6; qq++) nbr[qq] = (qq < x.conn) ? G.nbr[k][qq] : -1;
#pragma unroll
                for (int qq = 0; qq < 7; qq++) w[qq] = (qq < R) ? G.W[k][qq] : 0.0;
                resolve_moves<kNav>(x, p, sa, a, cc, kNav ? G.noff[k] : 0, kNav ? G.noff[k + kGroupCells] : 0,
                                    cv.nn, w, nbr, dir, nave, my_moves);
            }
        }
        const bool alive = valid && a.life != 0;
        if (kNav && x.nav.n) nav_movers(x, alive && dir == kNavDir, nave, slot);
        if (valid) {
            uint8_t code = (uint8_t)(dir >= 0 ? dir + 1 : 0);
            if (!alive) code |= kDecDead;
            if (a.bflag) code |= kDecBirth;
            x.dec[slot] = code;
            if (a.life != lifeS) x.cur.life[slot] = a.life;
            if (a.age_dirty) x.cur.age[slot] = a.age;
            if (p.has_pair) x.cur.mate_id[slot] = (a.mate >= 0) ? G.pid[q] : -3;
            if (a.bflag) {
                x.cur.nbabies[slot] = nbab + 1;
                if (x.gwords) x.bfather[slot] = (int32_t)(S0 + a.mate);
                atomicAdd(&G.cnt[k][kCntBirth], 1u);
            }
            if (alive && (!kNav || dir != kNavDir)) atomicAdd(&G.cnt[k][dir < 0 ? kCntStay : kCntMv + dir], 1u);
        }
    }
    wave_sync();
    if (lane < ng) {
        x.cstay[e.c] = G.cnt[lane][kCntStay];
        if (G.cnt[lane][kCntBirth]) x.cbirth[e.c] = G.cnt[lane][kCntBirth];
        if (G.cnt[lane][kCntStay] | G.cnt[lane][kCntBirth]) x.tact[e.c / kCellTile] = 1u;
        for (int d = 0; d < x.conn; d++) {
            const uint32_t m = G.cnt[lane][kCntMv + d];
            if (m) {
                const int32_t dest = G.nbr[lane][d];
                x.cin[G.outk[lane][d]] = m;
                atomicAdd(&x.carr[dest], m);
                x.tact[dest / kCellTile] = 1u;
            }
        }
    }
    wave_sync();
}

template <bool kRare, bool kAssort, bool kNav>
__device__ __forceinline__ void cell_act_kernel(const Ctx &x, const Prog &p, const StepArgs &sa) {
    __shared__ ActLds lds;
    const int lane = threadIdx.x;
    uint32_t my_deaths = 0, my_moves = 0, my_fert = 0;
    // one round trip: the status and this wave's first group of occupied
    // cells (read ahead of the count check: every index below ncells exists)
    const int32_t err = x.st->error;
    const int64_t nocc = x.st->n_occ[sa.parity];
    const int64_t gc = sa.gcells;
    OccCell e_first{0, 0, 0, 0};
    {
        const int64_t idx = (int64_t)blockIdx.x * gc + lane;
        if (lane < gc && idx < x.ncells) e_first = x.occ_cur[idx];
    }
    // a failed step (capacity, look-back, count) freezes the population:
    // later steps already enqueued touch nothing
    if (err) return;

    // look-back words of this step's k_cell_scan start at zero, and so do
    // the step's death and move counts (k_cell_scan adds them up)
    for (int32_t i = blockIdx.x * kCellBlock + lane; i < x.ntiles; i += gridDim.x * kCellBlock) {
        x.tiles[i] = 0;
        x.tiles2[i] = 0;
    }
    if (blockIdx.x == 0 && lane == 0) {
        x.st->deaths_last = 0;
        x.st->moves_last = 0;
        *x.tick = 0u;
    }

    // groups of gcells consecutive occupied cells, round-robin over the
    // waves; empty cells keep the zero counts the previous scan left
    const int64_t ngroups = (nocc + gc - 1) / gc;
    for (int64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const int64_t i0 = gi * gc;
        const int ng = (int)((nocc - i0 < gc) ? (nocc - i0) : gc);
        OccCell e{0, 0, 0, 0};
        if (lane < ng) e = (gi == blockIdx.x) ? e_first : x.occ_cur[i0 + lane];
        const uint32_t S0 = readlane_u32(e.s, 0);
        const uint32_t T = readlane_u32(e.s, ng - 1) + readlane_u32(e.n, ng - 1) - S0;
        const uint32_t nmax = wave_max_u32(lane < ng ? e.n : 0u);
        if (T <= (uint32_t)kGroupSlots && nmax <= (uint32_t)kGroupCellMax) {
            group_act<kRare, kAssort, kNav>(x, p, sa, ng, e, e.s - S0, S0, T, lds.g, my_deaths, my_moves, my_fert);
            continue;
        }
        for (int k = 0; k < ng; k++) {
            const int32_t c = (int32_t)readlane_u32((uint32_t)e.c, k);
            const uint32_t s = readlane_u32(e.s, k);
            const int32_t n = (int32_t)readlane_u32(e.n, k);
            PairLists L;
            if (n > kLdsCell) {
                L = PairLists{x.gfu + s, x.gmu + s, x.gfid + s, x.gmid + s, x.gfs + s, x.gms + s, x.gmate + s,
                              x.gflag + s};
            } else {
                L = PairLists{lds.c.fu, lds.c.mu, lds.c.fid, lds.c.mid, lds.c.fs, lds.c.ms, lds.c.mate,
                              lds.c.flag};
            }
            cell_act<kRare, kAssort, kNav>(x, p, sa, c, s, n, L, my_deaths, my_moves, my_fert);
        }
    }
    const uint64_t dm = wave_sum_u64(((uint64_t)my_moves << 32) | my_deaths);
    const uint32_t fe = (uint32_t)wave_sum_u64(my_fert);
    if (lane == 0) {
        uint32_t *wp = x.wave_part + kParts * blockIdx.x;
        wp[0] = (uint32_t)dm;
        wp[1] = (uint32_t)(dm >> 32);
        wp[2] = fe;
    }
}

template <bool kRare, bool kAssort>
__global__ void __launch_bounds__(kCellBlock) k_cell_act(Ctx x, Prog p, StepArgs sa) {
    cell_act_kernel<kRare, kAssort, false>(x, p, sa);
}

// with Navigate: its code would take the kernel to 193 VGPRs and 2 waves per
// SIMD; held to 3 waves (a few spilled registers)
template <bool kRare, bool kAssort>
__global__ void __launch_bounds__(kCellBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_cell_act_nav(Ctx x, Prog p, StepArgs sa) {
    cell_act_kernel<kRare, kAssort, true>(x, p, sa);
}

// ---------------------------------------------------------------------------
// k_assort_pair: AssortativePairHamming (AssortativePairHam.cpp:93-238 with
// RankTable.cpp:119-278), one wave per occupied cell, before k_cell_act.
//   - the cell's FERTILE females and males, each in id order;
//   - the match of female i and male j, 1 - calcDist(i, j, numBlocks) /
//     genomeSize as a float (GeneUtils.cpp:414-444), read from the genomes
//     when needed rather than held in a table: a couple's value in the
//     reference's table is that match while both are unpaired and it is at
//     or above the cutoff, else 0;
//   - rounds of RankTable::pairHighestRankers: the largest value, the
//     females that reach it (permuted when asked), each female in turn
//     picking at random among the unpaired males at that value.
// The cell's draws come in the reference's order from QHG_STREAM_ASSORT.
// Output: x.gmate[slot] = partner slot (-1 none) and mate_id; k_cell_act
// reads gmate instead of pairing.  Scratch in the cell's own slot range
// [s, s+n): gfu/gfid, gmu/gmid (gather), gfs/gms (slots in id order),
// gfu/gmu (then their genome handles in id order), gfid (the round's female
// order), gflag (paired: female r at s+r, male r at s+F+r).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bitcount2(uint64_t v) {   // GeneUtils::bitcount2
    return __popcll((v | (v >> 1)) & 0x5555555555555555ull);
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const float o = __shfl_xor(v, d, 64);
        v = (o > v) ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t assort_draw(uint64_t seed, uint32_t step, int32_t c, uint32_t j) {
    return draw(seed, step, QHG_STREAM_ASSORT, (uint64_t)(uint32_t)c | ((uint64_t)(j >> 16) << 32), j & 0xffffu);
}

template <typename CtxT>
__device__ __forceinline__ float assort_match(const CtxT &x, const Prog &p, int32_t hf, int32_t hm) {
    const uint64_t *a = x.gpool + (int64_t)hf * x.gwords;
    const uint64_t *b = x.gpool + (int64_t)hm * x.gwords;
    const int D = p.assort_words;
    int c1 = 0, c2 = 0;
    for (int k = 0; k < 2 * D; k++) c1 += bitcount2(a[k] ^ b[k]);
    for (int k = 0; k < D; k++) c2 += bitcount2(a[k + D] ^ b[k]) + bitcount2(a[k] ^ b[k + D]);
    const int d = (c1 < c2) ? c1 : c2;
    return (float)(1.0 - (1.0 * d) / (double)p.gen_size);
}

template <typename CtxT>
__device__ void assort_cell(const CtxT &x, const Prog &p, const StepArgs &sa, int32_t c, uint32_t s, int32_t n) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    int32_t F = 0, M = 0;
    for (int32_t b = 0; b < n; b += kWave) {
        const int32_t li = b + lane;
        const bool valid = li < n;
        const uint32_t slot = s + (uint32_t)li;
        uint8_t l = 0, g = 2;
        int64_t id = 0;
        if (valid) {
            l = x.cur.life[slot];
            g = x.cur.gender[slot];
            id = x.cur.id[slot];
            x.gmate[slot] = -1;
            x.cur.mate_id[slot] = -3;
        }
        const bool isF = valid && l == QHG_LIFE_FERTILE && g == 0;
        const bool isM = valid && l == QHG_LIFE_FERTILE && g == 1;
        const uint64_t mF = __ballot(isF), mM = __ballot(isM);
        if (isF) { const int32_t at = F + __popcll(mF & lt); x.gfu[s + at] = (int32_t)slot; x.gfid[s + at] = id; }
        if (isM) { const int32_t at = M + __popcll(mM & lt); x.gmu[s + at] = (int32_t)slot; x.gmid[s + at] = id; }
        F += __popcll(mF);
        M += __popcll(mM);
    }
    if (F == 0 || M == 0) return;
    wave_sync();
    // id order (ids are distinct): slots to gfs / gms by rank
    for (int32_t j = lane; j < F; j += kWave) {
        const int64_t my = x.gfid[s + j];
        int32_t r = 0;
        for (int32_t k = 0; k < F; k++) r += x.gfid[s + k] < my;
        x.gfs[s + r] = x.gfu[s + j];
    }
    for (int32_t j = lane; j < M; j += kWave) {
        const int64_t my = x.gmid[s + j];
        int32_t r = 0;
        for (int32_t k = 0; k < M; k++) r += x.gmid[s + k] < my;
        x.gms[s + r] = x.gmu[s + j];
    }
    wave_sync();
    for (int32_t j = lane; j < F; j += kWave) x.gfu[s + j] = x.cur.ghandle[x.gfs[s + j]];
    for (int32_t j = lane; j < M; j += kWave) x.gmu[s + j] = x.cur.ghandle[x.gms[s + j]];
    for (int32_t j = lane; j < F + M; j += kWave) x.gflag[s + j] = 0;
    wave_sync();
    const int32_t *hf = x.gfu + s, *hm = x.gmu + s;
    uint8_t *pf = x.gflag + s, *pm = x.gflag + s + F;
    if (p.assort_prot) {
        // AssortativePair: every listed agent's protein digests into its
        // slot's share of the pool (translateGenomeHash)
        for (int32_t j = lane; j < F + M; j += kWave) {
            const int32_t sl = (j < F) ? x.gfs[s + j] : x.gms[s + (j - F)];
            const int32_t h = (j < F) ? hf[j] : hm[j - F];
            const int32_t np = protein_digests(x.gpool + (int64_t)h * x.gwords, x.gwords, p.gen_size,
                                               x.gprot + (int64_t)sl * (4 * kProtBudget));
            x.gnprot[sl] = np;
            if (np > kProtBudget) atomicOr(&x.st->error, kErrProteins);
        }
        wave_sync();
    }
    // the match of female i and male j (id order)
    const auto match = [&](int32_t i, int32_t j) -> float {
        if (!p.assort_prot) return assort_match(x, p, hf[i], hm[j]);
        const int32_t sf = x.gfs[s + i], sm = x.gms[s + j];
        const int32_t nf = x.gnprot[sf], nm = x.gnprot[sm];
        return (float)protein_matches(x.gprot + (int64_t)sf * (4 * kProtBudget), nf < kProtBudget ? nf : kProtBudget,
                                      x.gprot + (int64_t)sm * (4 * kProtBudget), nm < kProtBudget ? nm : kProtBudget);
    };
    const float cut = p.assort_cutoff;
    uint32_t dj = 0;   // the cell's draws so far
    const int64_t FM = (int64_t)F * M;
    for (int32_t round = 0; round < F; round++) {
        // the largest value at or above the cutoff (0 when none)
        float vmax = 0.f;
        for (int64_t t = lane; t < FM; t += kWave) {
            const int32_t i = (int32_t)(t / M), j = (int32_t)(t - (int64_t)i * M);
            if (pf[i] || pm[j]) continue;
            const float v = match(i, j);
            if (v >= cut && v >= vmax) vmax = v;
        }
        vmax = wave_max_f32(vmax);
        if (!(vmax > 0.f)) break;
        // females reaching it, in id order (a value at vmax > 0 is at or above the cutoff)
        uint32_t np = 0;
        for (int32_t b = 0; b < F; b += kWave) {
            const int32_t i = b + lane;
            bool reach = false;
            if (i < F && !pf[i]) {
                for (int32_t j = 0; j < M && !reach; j++)
                    reach = !pm[j] && match(i, j) >= vmax;
            }
            const uint64_t mk = __ballot(reach);
            if (reach) x.gfid[s + np + (uint32_t)__popcll(mk & lt)] = i;
            np += (uint32_t)__popcll(mk);
        }
        wave_sync();
        int64_t *ord = x.gfid + s;
        if (p.assort_permute && np > 2) {
            // for i = np-1 down to 2: swap entry i with entry wrandi(0, i)
            const uint32_t nperm = np - 2;
            for (uint32_t base = 0; base < nperm; base += kWave) {
                const uint32_t u = (base + lane < nperm) ? assort_draw(x.seed, sa.step, c, dj + base + lane) : 0u;
                const uint32_t cnt = (nperm - base < (uint32_t)kWave) ? nperm - base : (uint32_t)kWave;
                for (uint32_t q = 0; q < cnt; q++) {
                    const uint32_t i = np - 1 - (base + q);
                    const uint32_t k = wrandi(0, i, 1, readlane_u32(u, (int)q));
                    if (lane == 0) {
                        const int64_t tmp = ord[k];
                        ord[k] = ord[i];
                        ord[i] = tmp;
                    }
                }
            }
            dj += nperm;
            wave_sync();
        }
        // each female in turn: a random unpaired male at vmax
        bool paired = false;
        for (uint32_t q = 0; q < np; q++) {
            const int32_t i = (int32_t)ord[q];
            uint32_t cnt = 0;
            for (int32_t b = 0; b < M; b += kWave) {
                const int32_t j = b + lane;
                const bool ok = j < M && !pm[j] && match(i, j) >= vmax;
                cnt += (uint32_t)__popcll(__ballot(ok));
            }
            if (cnt == 0) continue;
            int32_t sel = (int32_t)wrandi(0, cnt, 1, assort_draw(x.seed, sa.step, c, dj++));
            int32_t jm = -1;
            for (int32_t b = 0; b < M; b += kWave) {
                const int32_t j = b + lane;
                const bool ok = j < M && !pm[j] && match(i, j) >= vmax;
                const uint64_t mk = __ballot(ok);
                const int pc = __popcll(mk);
                if (sel < pc) { jm = b + select_bit(mk, sel); break; }
                sel -= pc;
            }
            if (jm < 0) {   // cannot happen: the second pass sees what the first counted
                if (lane == 0) atomicOr(&x.st->error, kErrGather);
                return;
            }
            if (lane == 0) {
                const int32_t sf = x.gfs[s + i], sm = x.gms[s + jm];
                pf[i] = 1;
                pm[jm] = 1;
                x.gmate[sf] = sm;
                x.gmate[sm] = sf;
                x.cur.mate_id[sf] = x.cur.id[sm];
                x.cur.mate_id[sm] = x.cur.id[sf];
            }
            paired = true;
            wave_sync();
        }
        if (!paired) break;   // makeAllPairings stops at a round without pairs
    }
}

// A cell of at most one wave of agents, everything in LDS: the sorted
// lists, the compared genome words and the whole distance table (F + M <=
// 64, so F * M <= 1024), read from HBM once.  A couple's match is the
// block's table value of its distance (d <= 2 * D * 32 = 128), the same
// float as 1 - d / genomeSize.  Same rounds and draws as assort_cell.
constexpr int kAssortMaxDist = 128;
struct AssortLds {
    int64_t fid[kWave], mid[kWave];       // ids, gather order
    int64_t fsid[kWave], msid[kWave];     // ids, id order
    int32_t fslot[kWave], mslot[kWave];   // slots, id order
    int32_t ord[kWave];                   // the round's female order
    uint64_t fw[kWave][4], mw[kWave][4];  // compared genome words (2 * D <= 4), id order
    float val[kAssortMaxDist + 1];        // match of each distance
    uint32_t rmax[kWave];                 // a round's largest value per female (float bits)
    uint8_t T[kWave * kWave / 4];         // distance of female i and male j at i * M + j
    uint8_t pf[kWave], pm[kWave];         // paired
};

__device__ __forceinline__ int assort_dist_w(const uint64_t *a, const uint64_t *b, int D) {
    int c1 = 0, c2 = 0;
    for (int k = 0; k < 2 * D; k++) c1 += bitcount2(a[k] ^ b[k]);
    for (int k = 0; k < D; k++) c2 += bitcount2(a[k + D] ^ b[k]) + bitcount2(a[k] ^ b[k + D]);
    return (c1 < c2) ? c1 : c2;
}

template <typename CtxT>
__device__ void assort_cell_lds(const CtxT &x, const Prog &p, const StepArgs &sa, int32_t c, uint32_t s, int32_t n,
                                AssortLds &L) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const bool valid = lane < n;
    const uint32_t slot = s + (uint32_t)lane;
    uint8_t l = 0, g = 2;
    int64_t id = 0;
    int32_t h = 0;
    if (valid) {
        l = x.cur.life[slot];
        g = x.cur.gender[slot];
        id = x.cur.id[slot];
        h = x.cur.ghandle[slot];
        x.gmate[slot] = -1;
        x.cur.mate_id[slot] = -3;
    }
    const bool isF = valid && l == QHG_LIFE_FERTILE && g == 0;
    const bool isM = valid && l == QHG_LIFE_FERTILE && g == 1;
    const uint64_t mF = __ballot(isF), mM = __ballot(isM);
    const int F = __popcll(mF), M = __popcll(mM);
    if (F == 0 || M == 0) return;
    if (isF) L.fid[__popcll(mF & lt)] = id;
    if (isM) L.mid[__popcll(mM & lt)] = id;
    wave_sync();
    // id order (ids are distinct), then the compared genome words
    int r = 0;
    if (isF) for (int k = 0; k < F; k++) r += L.fid[k] < id;
    if (isM) for (int k = 0; k < M; k++) r += L.mid[k] < id;
    const int D = p.assort_words;
    const uint64_t *gw = x.gpool + (int64_t)h * x.gwords;
    uint64_t w[4] = {0, 0, 0, 0};   // constant indices only: stays in registers
#pragma unroll
    for (int k = 0; k < 4; k++)
        if ((isF || isM) && k < 2 * D) w[k] = gw[k];
    if (isF) {
        L.fslot[r] = (int32_t)slot;
        L.fsid[r] = id;
#pragma unroll
        for (int k = 0; k < 4; k++) L.fw[r][k] = w[k];
        L.pf[r] = 0;
    }
    if (isM) {
        L.mslot[r] = (int32_t)slot;
        L.msid[r] = id;
#pragma unroll
        for (int k = 0; k < 4; k++) L.mw[r][k] = w[k];
        L.pm[r] = 0;
    }
    wave_sync();
    for (int t = lane; t < F * M; t += kWave) {
        const int i = t / M, j = t - i * M;
        L.T[t] = (uint8_t)assort_dist_w(L.fw[i], L.mw[j], D);
    }
    wave_sync();
    const float cut = p.assort_cutoff;
    // the cell's draws, computed 64 at a time and taken in order
    uint32_t dj = 0, ubase = 0;
    uint32_t U = assort_draw(x.seed, sa.step, c, (uint32_t)lane);
    auto next_u = [&]() -> uint32_t {
        if (dj - ubase >= (uint32_t)kWave) {
            ubase = dj;
            U = assort_draw(x.seed, sa.step, c, ubase + (uint32_t)lane);
        }
        return readlane_u32(U, (int)(dj++ - ubase));
    };
    for (int round = 0; round < F; round++) {
        // each unpaired female's largest value at or above the cutoff among
        // the unpaired males (0 when none; positive floats order as their bits)
        if (lane < F) L.rmax[lane] = 0u;
        wave_sync();
        for (int t = lane; t < F * M; t += kWave) {
            const int i = t / M, j = t - i * M;
            if (L.pf[i] || L.pm[j]) continue;
            const float v = L.val[L.T[t]];
            if (v >= cut && v > 0.f) atomicMax(&L.rmax[i], __float_as_uint(v));
        }
        wave_sync();
        const float rm = (lane < F) ? __uint_as_float(L.rmax[lane]) : 0.f;
        const float vmax = wave_max_f32(rm);
        if (!(vmax > 0.f)) break;
        // females reaching it, in id order; lane q holds the round's entry q
        const bool reach = lane < F && rm >= vmax;
        const uint64_t mk = __ballot(reach);
        const uint32_t np = (uint32_t)__popcll(mk);
        if (reach) L.ord[__popcll(mk & lt)] = lane;
        wave_sync();
        int32_t ord = ((uint32_t)lane < np) ? L.ord[lane] : 0;
        if (p.assort_permute && np > 2) {
            // for i = np-1 down to 2: swap entry i with entry wrandi(0, i)
            for (uint32_t i = np - 1; i > 1; i--) {
                const uint32_t k = wrandi(0, i, 1, next_u());
                const int32_t vk = (int32_t)readlane_u32((uint32_t)ord, (int)k);
                const int32_t vi = (int32_t)readlane_u32((uint32_t)ord, (int)i);
                ord = (lane == (int)k) ? vi : ((lane == (int)i) ? vk : ord);
            }
        }
        // each female in turn: a random unpaired male at vmax
        bool paired = false;
        for (uint32_t q = 0; q < np; q++) {
            const int i = (int)readlane_u32((uint32_t)ord, (int)q);
            const bool ok = lane < M && !L.pm[lane] && L.val[L.T[i * M + lane]] >= vmax;
            const uint64_t cm = __ballot(ok);
            const uint32_t cnt = (uint32_t)__popcll(cm);
            if (cnt == 0) continue;
            const int jm = select_bit(cm, (int)wrandi(0, cnt, 1, next_u()));
            if (lane == 0) {
                const int32_t sf = L.fslot[i], sm = L.mslot[jm];
                L.pf[i] = 1;
                L.pm[jm] = 1;
                x.gmate[sf] = sm;
                x.gmate[sm] = sf;
                x.cur.mate_id[sf] = L.msid[jm];
                x.cur.mate_id[sm] = L.fsid[i];
            }
            paired = true;
            wave_sync();
        }
        if (!paired) break;   // makeAllPairings stops at a round without pairs
    }
    wave_sync();
}

__global__ void __launch_bounds__(kWave) k_assort_pair(Ctx x, Prog p, StepArgs sa) {
    __shared__ AssortLds lds;
    if (x.st->error) return;
    const int64_t nocc = x.st->n_occ[sa.parity];
    for (int k = threadIdx.x; k <= kAssortMaxDist; k += kWave)
        lds.val[k] = (float)(1.0 - (1.0 * k) / (double)p.gen_size);
    wave_sync();
    for (int64_t i = blockIdx.x; i < nocc; i += gridDim.x) {
        const OccCell e = x.occ_cur[i];
        // the LDS form tabulates Hamming distances; AssortativePair's
        // protein matches go through the general form
        if (e.n <= (uint32_t)kWave && !p.assort_prot)
            assort_cell_lds(x, p, sa, e.c, e.s, (int32_t)e.n, lds);
        else
            assort_cell(x, p, sa, e.c, e.s, (int32_t)e.n);
    }
}

// ---------------------------------------------------------------------------
// k_cell_scan: next per-cell counts and birth counts, scanned together
// (low 32 bits agents, high bits births), and the occupied cells of the next
// step, with a decoupled look-back over tiles of kCellTile cells, one cell
// per thread.  Blocks take tiles by ticket; the grid fits on the device.
// Also writes the absolute next-buffer position of every incoming run and
// of each cell's newborns.  The block of the last tile closes the step's
// bookkeeping in the other status half.
// The scan is latency-bound (one wave per SIMD at most): every thread's
// loads go out in one round trip, and the code on that path is kept short.
// ---------------------------------------------------------------------------
constexpr int kStageWords = kCellTile * kMaxIn;       // a tile's incoming-run words
constexpr int kStagePer = kStageWords / kScanBlock;   // ... per thread
static_assert(kScanPer == 1, "one cell per scan thread");

struct ScanLds {
    uint64_t w[kScanBlock / 64], w2[kScanBlock / 64];
    uint64_t excl, excl2;
    uint32_t ticket;
    // dense steps: the tile's incoming-run counts and source slots, loaded
    // coalesced (cell-major, in_w words per cell)
    uint32_t cm[kStageWords];
    int32_t ie[kStageWords];
};

// The next tile for this block.  Tiles are handed out in increasing order by
// an atomic counter, so every tile a block waits on in the look-back has
// been taken by a block that is already running: the scan makes progress
// however few of its blocks are resident (other streams' kernels may hold
// compute units meanwhile).
__device__ __forceinline__ uint32_t next_tile(uint32_t *ticket, ScanLds &L) {
    if (threadIdx.x == 0) L.ticket = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)L.ticket);
    __syncthreads();
    return t;
}

// The inputs of one scan tile for this thread's cell: stayers, arrivals,
// births, and next count | births << 32.
struct TileIn {
    uint32_t stay, arr, birth;
    uint64_t v;
    bool staged;              // the tile's incoming runs go to ScanLds (dense step)
    uint32_t sm[kStagePer];   // ... this thread's share of them
    int32_t se[kStagePer];
    uint32_t rst, rbi, rar;   // raw loads, until tile_ready
    bool okc;                 // the cell exists
    int32_t wn;               // staged words of the tile
};

// A tile none of whose cells holds agents next step (the step's last tile
// always runs: it closes the bookkeeping).
template <typename CtxT>
__device__ __forceinline__ bool tile_idle(const CtxT &x, uint32_t tile) {
    const int64_t base = (int64_t)tile * kCellTile;
    return base + kCellTile < x.ncells && x.tact[tile] == 0u;
}

// Reads a tile's inputs into I (the counts stay in place; tile_clear zeroes
// them once the tile is known to be this block's).  With stage and many
// cells occupied (sa.dense), the incoming-run counts and source slots of
// the whole tile are read too, coalesced, for tile_stage to put into LDS;
// otherwise only those of cells that receive movers are read, at the write.
// tile_load only issues the loads (indices past the end are clamped, so no
// branch splits them): tile_ready then drops what lies past the end, so
// that other loads can be issued in between and share the round trip.
template <typename CtxT>
__device__ __forceinline__ void tile_load(const CtxT &x, const StepArgs &sa, uint32_t tile, bool stage, TileIn &I) {
    const int64_t nc = x.ncells;
    const int64_t base = (int64_t)tile * kCellTile;
    const int64_t c = base + (int64_t)threadIdx.x;
    I.okc = c < nc;
    const int64_t cc = I.okc ? c : 0;
    I.rst = x.cstay[cc];
    I.rbi = x.cbirth[cc];
    I.rar = x.carr[cc];
    I.staged = stage && sa.dense != 0;
    const int32_t iw = x.in_w;
    const int64_t w0 = base * iw;
    I.wn = (int32_t)(((nc - base < kCellTile) ? (nc - base) : kCellTile) * iw);
    if (I.staged) {
#pragma unroll
        for (int r = 0; r < kStagePer; r++) {
            const int32_t j = r * kScanBlock + (int32_t)threadIdx.x;
            const int64_t jj = w0 + ((j < I.wn) ? j : 0);
            I.sm[r] = x.cin[jj];
            I.se[r] = x.in_e[jj];
        }
    }
}

__device__ __forceinline__ void tile_ready(TileIn &I) {
    I.stay = I.okc ? I.rst : 0u;
    I.birth = I.okc ? I.rbi : 0u;
    I.arr = I.okc ? I.rar : 0u;
    I.v = (uint64_t)(I.stay + I.arr + I.birth) | ((uint64_t)I.birth << 32);
    if (I.staged) {
#pragma unroll
        for (int r = 0; r < kStagePer; r++) {
            if (r * kScanBlock + (int32_t)threadIdx.x >= I.wn) {
                I.sm[r] = 0u;
                I.se[r] = -1;
            }
        }
    }
}

// The staged incoming runs into LDS (read after the next block barrier).
__device__ __forceinline__ void tile_stage(const TileIn &I, ScanLds &L) {
    if (!I.staged) return;
#pragma unroll
    for (int r = 0; r < kStagePer; r++) {
        const int32_t j = r * kScanBlock + (int32_t)threadIdx.x;
        L.cm[j] = I.sm[r];
        L.ie[j] = I.se[r];
    }
}

// The scan proper owns the tile: the counts it read go back to zero (the
// staged incoming-run counts too; unstaged ones are cleared at the write).
template <typename CtxT>
__device__ __forceinline__ void tile_clear(const CtxT &x, uint32_t tile, const TileIn &I) {
    const int64_t base = (int64_t)tile * kCellTile;
    const int64_t c = base + (int64_t)threadIdx.x;
    if (I.stay) x.cstay[c] = 0u;
    if (I.birth) x.cbirth[c] = 0u;
    if (I.arr) x.carr[c] = 0u;
    if (I.staged) {
        const int64_t w0 = base * x.in_w;
#pragma unroll
        for (int r = 0; r < kStagePer; r++)
            if (I.sm[r] != 0u) x.cin[w0 + r * kScanBlock + (int32_t)threadIdx.x] = 0u;
    }
}

// Block sums of a tile's inputs: the tile totals, and this thread's
// exclusive offsets inside the tile.
struct TileSums {
    uint64_t tot, otot, pre, opre;
};

__device__ __forceinline__ TileSums tile_sums(const TileIn &I, ScanLds &L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t occ = ((uint32_t)I.v != 0u) ? 1u : 0u;
    const uint64_t incl = wave_incl_scan_u64(I.v);
    const uint64_t oincl = wave_incl_scan_u64(occ);
    if (lane == 63) { L.w[wave] = incl; L.w2[wave] = oincl; }
    __syncthreads();
    TileSums S{0, 0, 0, 0};
    uint64_t wpre = 0, owpre = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; w++) {
        if (w < wave) { wpre += L.w[w]; owpre += L.w2[w]; }
        S.tot += L.w[w];
        S.otot += L.w2[w];
    }
    S.pre = wpre + incl - I.v;
    S.opre = owpre + oincl - occ;
    return S;
}

// Writes a tile's outputs from this cell's exclusive prefix (run agents |
// births, orun occupied cells): next offset, run and newborn positions, the
// next occupied-cell list; the step's last tile closes the bookkeeping.
template <typename CtxT>
__device__ __forceinline__ void tile_write(const CtxT &x, const StepArgs &sa, uint32_t tile, const TileIn &I,
                                           const ScanLds &L, uint64_t run, uint64_t orun) {
    const int64_t nc = x.ncells;
    const int64_t base = (int64_t)tile * kCellTile;
    const bool last = base + kCellTile >= nc;
    const int64_t c = base + (int64_t)threadIdx.x;
    const int32_t iw = x.in_w;
    if (I.okc) {
        const uint32_t o = (uint32_t)run;
        x.off_nxt[c] = o;
        uint32_t pos = o + I.stay;
        if (I.arr) {
            // arrivals: one block per incoming run, in run order (the row's
            // counts and slots read at once)
            uint32_t m[kMaxIn];
            int32_t ie[kMaxIn];
            const int32_t lrow = (int32_t)threadIdx.x * iw;
#pragma unroll
            for (int k = 0; k < kMaxIn; k++) {
                const bool in = k < iw;
                m[k] = !in ? 0u : (I.staged ? L.cm[lrow + k] : x.cin[c * iw + k]);
                ie[k] = !in ? -1 : (I.staged ? L.ie[lrow + k] : x.in_e[c * iw + k]);
            }
#pragma unroll
            for (int k = 0; k < kMaxIn; k++) {
                if (m[k]) {
                    x.rpos[ie[k]] = pos;
                    if (!I.staged) x.cin[c * iw + k] = 0;
                    pos += m[k];
                }
            }
            if (x.nav.n) {
                // Navigate's runs follow, in edge order
                const int32_t j1 = x.nav.in_off[c + 1];
                for (int32_t j = x.nav.in_off[c]; j < j1; j++) {
                    const int32_t ne = x.nav.in_e[j];
                    const uint32_t mn = x.nav.cnt[ne];
                    if (mn) {
                        x.nav.rpos[ne] = pos;
                        x.nav.cnt[ne] = 0;
                        pos += mn;
                    }
                }
            }
        }
        if (I.birth) {
            x.bpos[c] = pos;
            x.birth_off[c] = (uint32_t)(run >> 32);
        }
        const uint32_t nn = (uint32_t)I.v;
        if (nn != 0) x.occ_nxt[orun] = OccCell{(int32_t)c, o, nn, 0u};
    }
    if (threadIdx.x == 0) x.tact[tile] = 0u;
    if (last && threadIdx.x == kScanBlock - 1) {
        // last tile: the next step's sizes (the totals the step kernels
        // need; deaths, moves and the host snapshot follow in k_cell_scatter)
        run += I.v;
        orun += ((uint32_t)I.v != 0u) ? 1u : 0u;
        const int64_t total = (int64_t)(uint32_t)run;
        const int64_t B = (int64_t)(run >> 32);
        x.off_nxt[nc] = (uint32_t)total;
        x.birth_off[nc] = (uint32_t)B;
        DevStatus *st = x.st;
        const int p = (int)sa.parity, q = 1 - p;
        if (total > x.cap || (x.gwords && st->gpool[p] + B > x.gcap)) st->error |= kErrCapacity;
        st->agent_steps += st->n_alive[p];
        st->occupied_total += st->n_occ[p];
        st->n_alive[q] = total;
        st->n_occ[q] = (int64_t)orun;
        st->max_id[q] = st->max_id[p] + B;
        st->gpool[q] = st->gpool[p] + (x.gwords ? B : 0);
        st->births_last = B;
        st->births_total += B;
        st->step = (int64_t)sa.step + 1;
        st->t = sa.t + 1.0;
    }
}

// One look-back tile of k_cell_scan, all kScanBlock threads, from its
// inputs already loaded into I (and staged into L).  more: another tile
// follows in this block (the LDS is reused).
template <typename CtxT>
__device__ __forceinline__ void scan_tile(const CtxT &x, const StepArgs &sa, uint32_t tile, uint32_t act,
                                          TileIn &I, ScanLds &L, bool more) {
    if (act == 0u) {
        // no cell of the tile holds agents next step: its sums are zero and
        // none of its per-cell outputs is read (the step totals come from
        // the last tile, which always runs)
        if (threadIdx.x == 0) {
            const uint64_t f = (tile == 0) ? kFlagInc : kFlagAgg;
            tile_publish(&x.tiles[tile], f);
            tile_publish(&x.tiles2[tile], f);
        }
        return;
    }
    tile_clear(x, tile, I);
    const TileSums S = tile_sums(I, L);
    if (threadIdx.x < 64) {
        uint64_t ex, ex2;
        lookback2(x.tiles, x.tiles2, tile, S.tot, S.otot, &x.st->error, ex, ex2);
        if (threadIdx.x == 0) { L.excl = ex; L.excl2 = ex2; }
    }
    __syncthreads();
    tile_write(x, sa, tile, I, L, L.excl + S.pre, L.excl2 + S.opre);
    if (more) __syncthreads();
}

// kOneRound: the grid has a block per tile, so the first round of tickets
// hands out every tile and no block asks again (straight-line code; the
// host launches the looping form otherwise).
template <bool kOneRound>
__global__ void __launch_bounds__(kScanBlock) k_cell_scan(Ctx x, StepArgs sa) {
    __shared__ ScanLds lds;
    const int lane = threadIdx.x & 63;
    const uint32_t nt = (uint32_t)x.ntiles;
    // One round trip for: the status, this thread's first slice of the
    // k_cell_act partials (deaths, moves, FERTILE counts), the inputs of tile
    // blockIdx.x (the ticket this block most likely draws) with its activity
    // flag, and, issued last, the first ticket: waiting for the ticket waits
    // for them all.  The partials' wave sums are stored at the end for
    // k_cell_scatter's bookkeeping block (plain stores: no contended atomics).
    const int32_t err = x.st->error;
    const int32_t j0 = blockIdx.x * kScanBlock + threadIdx.x;
    const int32_t jstride = gridDim.x * kScanBlock;
    const bool pok = j0 < x.nact_blocks;
    const int32_t jc = pok ? j0 : 0;
    const uint32_t p0 = x.wave_part[kParts * jc + 0];
    const uint32_t p1 = x.wave_part[kParts * jc + 1];
    const uint32_t p2 = x.wave_part[kParts * jc + 2];
    const uint32_t spec = (blockIdx.x < nt) ? blockIdx.x : nt - 1;
    const uint32_t ta = x.tact[spec];
    TileIn I;
    tile_load(x, sa, spec, true, I);
    uint32_t tk = 0;
    if (threadIdx.x == 0) tk = atomicAdd(x.tick, 1u);
    tile_ready(I);
    tile_stage(I, lds);
    if (threadIdx.x == 0) lds.ticket = tk;
    __syncthreads();
    const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)lds.ticket);
    if (err) return;
    const auto active = [&](uint32_t t, uint32_t flag) -> uint32_t {
        return ((int64_t)t * kCellTile + kCellTile >= x.ncells || flag != 0u) ? 1u : 0u;
    };
    if (kOneRound) {
        if (first < nt) {
            uint32_t act = active(spec, ta);
            if (first != spec) {
                // not the tile loaded ahead: load this one now
                __syncthreads();   // the LDS staging may still be read
                act = active(first, x.tact[first]);
                tile_load(x, sa, first, true, I);
                tile_ready(I);
                tile_stage(I, lds);
            }
            scan_tile(x, sa, first, act, I, lds, false);
        }
    } else {
        uint32_t tile = spec;
        for (uint32_t t = first; t < nt; t = next_tile(x.tick, lds)) {
            uint32_t act = active(tile, ta);
            if (t != tile) {
                __syncthreads();   // the LDS staging may still be read
                tile = t;
                act = active(tile, x.tact[tile]);
                tile_load(x, sa, tile, true, I);
                tile_ready(I);
                tile_stage(I, lds);
            }
            scan_tile(x, sa, t, act, I, lds, true);
            tile = nt;   // later tiles are loaded on demand
        }
    }
    uint64_t D = pok ? p0 : 0u, Mv = pok ? p1 : 0u, Fe = pok ? p2 : 0u;
    for (int32_t j = j0 + jstride; j < x.nact_blocks; j += jstride) {
        D += x.wave_part[kParts * j + 0];
        Mv += x.wave_part[kParts * j + 1];
        Fe += x.wave_part[kParts * j + 2];
    }
    D = wave_sum_u64(D);
    Mv = wave_sum_u64(Mv);
    Fe = wave_sum_u64(Fe);
    if (lane == 0) {
        uint64_t *sp = x.scan_part + 3 * ((int64_t)blockIdx.x * (kScanBlock / kWave) + (threadIdx.x / kWave));
        sp[0] = D;
        sp[1] = Mv;
        sp[2] = Fe;
    }
}

// ---------------------------------------------------------------------------
// genetics: GeneUtils::makeMultiMask / crossOver / freeReco / mutateNucs
// (GeneUtils.cpp:112-246, :322-360), one genome word per lane
// ---------------------------------------------------------------------------



// ---------------------------------------------------------------------------
// k_cell_scatter: agents of one occupied source cell to the next buffer
// ---------------------------------------------------------------------------
template <typename CtxT>
__device__ __forceinline__ void cell_scatter(const CtxT &x, const StepArgs &sa, int32_t c, uint32_t s, int32_t n,
                                             int32_t *mli, int64_t *mid, int64_t max_id, int64_t gbase) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const uint32_t nbase = x.off_nxt[c];
    const uint32_t bp = x.bpos[c];
    const uint32_t boff = x.birth_off[c];
    uint32_t rp[6];
#pragma unroll
    for (int d = 0; d < 6; d++) rp[d] = (d < x.conn) ? x.rpos[(int64_t)c * x.conn + d] : 0u;
    uint32_t nstay = 0, nmoth = 0;
    uint32_t nrun[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t b = 0; b < n; b += kWave) {
        const int32_t li = b + lane;
        const bool valid = li < n;
        const uint32_t from = s + (uint32_t)li;
        uint8_t code = kDecDead;
        int64_t id = 0, mate = 0;
        uint8_t life = 0, g = 0;
        double birth = 0, age = 0, lastb = 0;
        int32_t nbab = 0, gh = 0;
        if (valid) {
            code = x.dec[from];
            id = x.cur.id[from];
            mate = x.cur.mate_id[from];
            life = x.cur.life[from];
            g = x.cur.gender[from];
            birth = x.cur.birth[from];
            age = x.cur.age[from];
            lastb = x.cur.lastb[from];
            nbab = x.cur.nbabies[from];
            if (x.gwords) gh = x.cur.ghandle[from];
        }
        const bool alive = valid && !(code & kDecDead);
        const uint32_t dir = code & kDecDirMask;
        int64_t to = -1;
        const uint64_t ms = __ballot(alive && dir == 0);
        if (alive && dir == 0) to = (int64_t)nbase + nstay + __popcll(ms & lt);
        nstay += __popcll(ms);
#pragma unroll
        for (int d = 0; d < 6; d++) {
            const bool mine = alive && dir == (uint32_t)(d + 1);
            const uint64_t mm = __ballot(mine);
            if (mine) to = (int64_t)rp[d] + nrun[d] + __popcll(mm & lt);
            nrun[d] += __popcll(mm);
        }
        if (alive && dir == kDecNav) {
            // a Navigate edge's run: its slot and the rank k_cell_act gave it
            const uint64_t ns = x.nav.sel[from];
            to = (int64_t)x.nav.rpos[(uint32_t)ns] + (uint32_t)(ns >> 32);
        }
        if (to >= 0) {
            x.nxt.id[to] = id;
            x.nxt.mate_id[to] = mate;
            x.nxt.life[to] = life;
            x.nxt.gender[to] = g;
            x.nxt.birth[to] = birth;
            x.nxt.age[to] = age;
            x.nxt.lastb[to] = lastb;
            x.nxt.nbabies[to] = nbab;
            if (x.gwords) x.nxt.ghandle[to] = gh;
        }
        const bool isb = valid && (code & kDecBirth);
        const uint64_t mb = __ballot(isb);
        if (isb) {
            const uint32_t at = nmoth + __popcll(mb & lt);
            mli[at] = li;
            mid[at] = id;
        }
        nmoth += __popcll(mb);
    }
    if (nmoth == 0) return;
    wave_sync();
    // newborns in mother-id order: createAgentAtIndex + makePopSpecificOffspring
    // (SPopulation.cpp:880-918); new ids rank births by (cell, mother id)
    for (uint32_t j = lane; j < nmoth; j += kWave) {
        const int64_t my = mid[j];
        uint32_t r = 0;
        for (uint32_t k = 0; k < nmoth; k++) r += mid[k] < my;
        const int64_t to = (int64_t)bp + r;
        const int64_t rg = (int64_t)boff + r;
        const int64_t nid = max_id + 1 + rg;
        x.nxt.id[to] = nid;
        x.nxt.mate_id[to] = -3;
        x.nxt.birth[to] = sa.t;
        x.nxt.age[to] = 0.0;
        x.nxt.lastb[to] = 0.0;
        x.nxt.nbabies[to] = 0;
        const uint32_t u = draw(x.seed, sa.step, QHG_STREAM_NEWBORN, (uint64_t)nid, 0);
        const uint8_t gg = (uint8_t)(2 * wrandd(u));
        x.nxt.gender[to] = gg;
        x.nxt.life[to] = (gg == 0) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        if (x.gwords) {
            const uint32_t mslot = s + (uint32_t)mli[j];
            x.nxt.ghandle[to] = (int32_t)(gbase + rg);
            x.brec[2 * rg] = (int32_t)mslot;
            x.brec[2 * rg + 1] = x.bfather[mslot];
        }
    }
    wave_sync();
}

enum { kScStay = 0, kScMv = 1, kScMoth = 7, kScN = 8 };

struct GroupScatLds {
    int64_t mid[kGroupSlots];                 // ids of the mothers, per cell region
    uint32_t cnt[kGroupCells][kScN];          // per-cell running counts
    int16_t mj[kGroupSlots];                  // group positions of the mothers
};

struct CellScatLds {
    int32_t mli[kLdsCell];
    int64_t mid[kLdsCell];
};

union ScatLds {
    GroupScatLds g;
    CellScatLds c;
};

// all cells of a group (see group_act): stayers and movers ranked per cell
// and run by masked ballots, running counts across chunks in LDS
template <typename CtxT>
__device__ __forceinline__ void group_scatter(const CtxT &x, const StepArgs &sa, int ng, const OccCell &e,
                                              uint32_t eb, uint32_t S0, uint32_t T, GroupScatLds &G,
                                              int64_t max_id, int64_t gbase) {
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int nchunks = (int)((T + kWave - 1) / kWave);
    // the cells' newborn positions and birth ranks, read now so the newborn
    // pass below needs no round trip of its own (values of cells without
    // births are stale and unused)
    uint32_t cb_pos = 0, cb_off = 0;
    if (lane < ng) {
#pragma unroll
        for (int q = 0; q < kScN; q++) G.cnt[lane][q] = 0;
        cb_pos = x.bpos[e.c];
        cb_off = x.birth_off[e.c];
    }
    wave_sync();
    for (int ch = 0; ch < nchunks; ch++) {
        const uint32_t q = (uint32_t)(ch * kWave + lane);
        const bool valid = q < T;
        const uint32_t from = S0 + q;
        const int k = group_cell_of(q, eb, ng);
        const int32_t c = __shfl(e.c, k, 64);
        const uint32_t base = (uint32_t)__shfl((int)eb, k, 64);
        const uint32_t nk = (uint32_t)__shfl((int)e.n, k, 64);
        uint8_t code = kDecDead;
        int64_t id = 0, mate = 0;
        uint8_t life = 0, g = 0;
        double birth = 0, age = 0, lastb = 0;
        int32_t nbab = 0, gh = 0;
        uint32_t nbase = 0;
        uint32_t rp[6] = {0, 0, 0, 0, 0, 0};
        if (valid) {
            code = x.dec[from];
            id = x.cur.id[from];
            mate = x.cur.mate_id[from];
            life = x.cur.life[from];
            g = x.cur.gender[from];
            birth = x.cur.birth[from];
            age = x.cur.age[from];
            lastb = x.cur.lastb[from];
            nbab = x.cur.nbabies[from];
            if (x.gwords) gh = x.cur.ghandle[from];
            nbase = x.off_nxt[c];
#pragma unroll
            for (int d = 0; d < 6; d++)
                if (d < x.conn) rp[d] = x.rpos[(int64_t)c * x.conn + d];
        }
        // lanes of this lane's cell inside the chunk
        const int lo = (int)base - ch * kWave;
        const int hi = (int)(base + nk) - ch * kWave;
        const uint64_t cm = (valid ? ((hi >= 64) ? ~0ull : ((1ull << hi) - 1ull)) : 0ull) &
                            ~((lo <= 0) ? 0ull : ((1ull << lo) - 1ull));
        const bool last_of_cell = valid && ((int)lane == ((hi >= 64) ? 63 : hi - 1));
        const bool alive = valid && !(code & kDecDead);
        const uint32_t dir = code & kDecDirMask;
        int64_t to = -1;
        const uint64_t ms = __ballot(alive && dir == 0);
        const uint32_t c_stay = valid ? G.cnt[k][kScStay] : 0u;
        if (alive && dir == 0) to = (int64_t)nbase + c_stay + __popcll(ms & lt & cm);
        uint32_t add[7];
        add[0] = __popcll(ms & cm);
#pragma unroll
        for (int d = 0; d < 6; d++) {
            const bool mine = alive && dir == (uint32_t)(d + 1);
            const uint64_t mm = __ballot(mine);
            if (mine) to = (int64_t)rp[d] + G.cnt[k][kScMv + d] + __popcll(mm & lt & cm);
            add[1 + d] = __popcll(mm & cm);
        }
        if (alive && dir == kDecNav) {
            const uint64_t ns = x.nav.sel[from];
            to = (int64_t)x.nav.rpos[(uint32_t)ns] + (uint32_t)(ns >> 32);
        }
        if (to >= 0) {
            x.nxt.id[to] = id;
            x.nxt.mate_id[to] = mate;
            x.nxt.life[to] = life;
            x.nxt.gender[to] = g;
            x.nxt.birth[to] = birth;
            x.nxt.age[to] = age;
            x.nxt.lastb[to] = lastb;
            x.nxt.nbabies[to] = nbab;
            if (x.gwords) x.nxt.ghandle[to] = gh;
        }
        if (valid && (code & kDecBirth)) {
            const uint32_t r = atomicAdd(&G.cnt[k][kScMoth], 1u);
            G.mj[base + r] = (int16_t)q;
            G.mid[base + r] = id;
        }
        wave_sync();
        if (last_of_cell) {
#pragma unroll
            for (int d = 0; d < 7; d++) G.cnt[k][kScStay + d] += add[d];
        }
        wave_sync();
    }
    // newborns in mother-id order per cell: createAgentAtIndex +
    // makePopSpecificOffspring (SPopulation.cpp:880-918); new ids rank births
    // by (cell, mother id)
    for (int ch = 0; ch < nchunks; ch++) {
        const uint32_t q = (uint32_t)(ch * kWave + lane);
        const bool valid = q < T;
        const int k = group_cell_of(q, eb, ng);
        const uint32_t base = (uint32_t)__shfl((int)eb, k, 64);
        const uint32_t kpos = (uint32_t)__shfl((int)cb_pos, k, 64);
        const uint32_t koff = (uint32_t)__shfl((int)cb_off, k, 64);
        if (!valid) continue;
        const uint32_t nm = G.cnt[k][kScMoth];
        const uint32_t t = q - base;
        if (t >= nm) continue;
        const int64_t my = G.mid[q];
        uint32_t r = 0;
        for (uint32_t u = 0; u < nm; u++) r += G.mid[base + u] < my;
        const int64_t to = (int64_t)kpos + r;
        const int64_t rg = (int64_t)koff + r;
        const int64_t nid = max_id + 1 + rg;
        x.nxt.id[to] = nid;
        x.nxt.mate_id[to] = -3;
        x.nxt.birth[to] = sa.t;
        x.nxt.age[to] = 0.0;
        x.nxt.lastb[to] = 0.0;
        x.nxt.nbabies[to] = 0;
        const uint32_t u = draw(x.seed, sa.step, QHG_STREAM_NEWBORN, (uint64_t)nid, 0);
        const uint8_t gg = (uint8_t)(2 * wrandd(u));
        x.nxt.gender[to] = gg;
        x.nxt.life[to] = (gg == 0) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        if (x.gwords) {
            const uint32_t mslot = S0 + (uint32_t)G.mj[q];
            x.nxt.ghandle[to] = (int32_t)(gbase + rg);
            x.brec[2 * rg] = (int32_t)mslot;
            x.brec[2 * rg + 1] = x.bfather[mslot];
        }
    }
    wave_sync();
}

__global__ void __launch_bounds__(kCellBlock) k_cell_scatter(Ctx x, StepArgs sa) {
    __shared__ ScatLds lds;
    const int lane = threadIdx.x;
    const int pr = (int)sa.parity;
    const int64_t G = gridDim.x - 1;
    const int64_t b0 = (int64_t)blockIdx.x - 1;
    // one round trip: the status and this wave's first group (read ahead of
    // the count check: every index below ncells exists)
    const int32_t err = x.st->error;
    const int64_t max_id = x.st->max_id[pr];
    const int64_t gbase = x.st->gpool[pr];
    const int64_t nocc = x.st->n_occ[pr];
    const int64_t gc = sa.gcells;
    OccCell e_first{0, 0, 0, 0};
    {
        const int64_t idx = b0 * gc + lane;
        if (b0 >= 0 && lane < gc && idx < x.ncells) e_first = x.occ_cur[idx];
    }
    if (err) return;
    if (blockIdx.x == 0) {
        // the step's deaths, moves and FERTILE counts (k_cell_scan's per-wave
        // sums) go into the status and the host's snapshot; the status words
        // and the sums load in the same round trip, one word per lane
        constexpr int kWords = (int)(sizeof(DevStatus) / 8);
        static_assert(kWords <= kWave, "one status word per lane");
        uint64_t *sw = reinterpret_cast<uint64_t *>(x.st);
        uint64_t w = (lane < kWords) ? sw[lane] : 0ull;
        uint64_t D = 0, Mv = 0, Fe = 0;
        for (int32_t j = lane; j < x.scan_waves; j += kWave) {
            D += x.scan_part[3 * j + 0];
            Mv += x.scan_part[3 * j + 1];
            Fe += x.scan_part[3 * j + 2];
        }
        D = wave_sum_u64(D);
        Mv = wave_sum_u64(Mv);
        Fe = wave_sum_u64(Fe);
        bool upd = true;
        if (lane == (int)(offsetof(DevStatus, deaths_last) / 8)) w = D;
        else if (lane == (int)(offsetof(DevStatus, deaths_total) / 8)) w += D;
        else if (lane == (int)(offsetof(DevStatus, moves_last) / 8)) w = Mv;
        else if (lane == (int)(offsetof(DevStatus, moves_total) / 8)) w += Mv;
        else if (lane == (int)(offsetof(DevStatus, fertile_total) / 8)) w += Fe;
        else upd = false;
        if (upd) sw[lane] = w;
        if (x.hring) {
            constexpr int kStep = (int)(offsetof(DevStatus, step) / 8);
            const uint64_t step = ((uint64_t)readlane_u32((uint32_t)(w >> 32), kStep) << 32) |
                                  readlane_u32((uint32_t)w, kStep);
            const int64_t slot = (int64_t)(step % (uint64_t)kStatusRing);
            uint64_t *dst = reinterpret_cast<uint64_t *>(x.hring + slot);
            if (lane < kWords) dst[lane] = w;
            if (step % (uint64_t)kStatusSealEvery == 0) {
                // the seal words follow the kStatusRing snapshots: the step
                // goes there once the snapshot is visible to the host (a
                // system-scope release), and the host polls it instead of
                // waiting on an event between two steps' kernels
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                int64_t *seal = reinterpret_cast<int64_t *>(x.hring + kStatusRing) + slot;
                if (lane == 0) __hip_atomic_store(seal, (int64_t)step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        return;
    }
    const int64_t ngroups = (nocc + gc - 1) / gc;
    for (int64_t gi = b0; gi < ngroups; gi += G) {
        const int64_t i0 = gi * gc;
        const int ng = (int)((nocc - i0 < gc) ? (nocc - i0) : gc);
        OccCell e{0, 0, 0, 0};
        if (lane < ng) e = (gi == b0) ? e_first : x.occ_cur[i0 + lane];
        const uint32_t S0 = readlane_u32(e.s, 0);
        const uint32_t T = readlane_u32(e.s, ng - 1) + readlane_u32(e.n, ng - 1) - S0;
        if (T <= (uint32_t)kGroupSlots) {
            group_scatter(x, sa, ng, e, e.s - S0, S0, T, lds.g, max_id, gbase);
            continue;
        }
        for (int k = 0; k < ng; k++) {
            const int32_t c = (int32_t)readlane_u32((uint32_t)e.c, k);
            const uint32_t s = readlane_u32(e.s, k);
            const int32_t n = (int32_t)readlane_u32(e.n, k);
            // mothers of a large cell go to the cell's own range of global scratch
            int32_t *mli = (n > kLdsCell) ? x.gfu + s : lds.c.mli;
            int64_t *mid = (n > kLdsCell) ? x.gfid + s : lds.c.mid;
            cell_scatter(x, sa, c, s, n, mli, mid, max_id, gbase);
        }
    }
}

// bits lo..hi (inclusive, lo <= hi <= 63)
__device__ __forceinline__ uint64_t bit_range(int lo, int hi) {
    const uint64_t up = (hi >= 63) ? ~0ull : ((1ull << (hi + 1)) - 1ull);
    return up & ~((1ull << lo) - 1ull);
}

// Word w of a child genome (Genetics::makeOffspring, Genetics.cpp:285-337):
// strand s = w / nbk comes from parent s (0 mother, 1 father), block
// b = w % nbk.  The draws follow makeOffspring's order: i1, i2, the mother's
// crossover draws, the father's, the mutation count, then (position, mask)
// pairs; mutations that hit this word are applied here.  Draws 0-3 and the
// block holding the mutation count are computed once.
template <typename CtxT, typename ProgT>
__device__ uint64_t child_word(const CtxT &x, const ProgT &p, uint32_t step, uint64_t nid, const uint64_t *par,
                               int s, int b, int nbk, int w) {
    uint32_t d0[4];
    draw4(x.seed, step, QHG_STREAM_GENETICS, nid, 0, d0);
    auto D = [&](uint32_t n) -> uint32_t {
        if (n < 4) {
            uint32_t v = d0[0];
#pragma unroll
            for (int q = 1; q < 4; q++) if ((uint32_t)q == n) v = d0[q];
            return v;
        }
        return draw(x.seed, step, QHG_STREAM_GENETICS, nid, n);
    };
    const int sel = (int)(2 * 1.0 * wrandd(s ? d0[1] : d0[0]));   // i1 for the mother, i2 for the father
    const int C = p.gen_cross;
    uint64_t word;
    uint32_t nmut;
    if (C > 0) {
        // GeneUtils::crossOver (GeneUtils.cpp:184-246) restricted to block b:
        // this parent's breaks are draws 2 + s*C + q, q < C
        const uint32_t nbits = (uint32_t)nbk * 64u;
        int before = 0, in = 0;
        for (int q = 0; q < C; q++) {
            const uint32_t bp = wrandi(0, nbits, 2, D(2 + (uint32_t)(s * C + q)));
            before += (int)(bp >> 6) < b;
            in += (int)(bp >> 6) == b;
        }
        const int cur = before & 1;
        const uint64_t gc = par[cur * nbk + b], go = par[(1 - cur) * nbk + b];
        uint64_t L = ~0ull;
        if (in) {
            // makeMultiMask (GeneUtils.cpp:156-175) over the block's breaks in
            // draw order, walked from the last: bit 63-j takes the current
            // value; a break at bit v flips it at j = 64-v, but only if that j
            // is still ahead, otherwise the walk stops there (unsorted breaks,
            // repeats and breaks at bit 0 included)
            uint32_t k = 1u - (uint32_t)(in % 2);
            int jc = 0;
            uint64_t m = 0;
            for (int q = C - 1; q >= 0; q--) {
                const uint32_t bp = wrandi(0, nbits, 2, D(2 + (uint32_t)(s * C + q)));
                if ((int)(bp >> 6) != b) continue;
                const int t = 64 - (int)(bp & 63u);
                if (t > 63 || t <= jc) break;
                if (k) m |= bit_range(64 - t, 63 - jc);
                k = 1u - k;
                jc = t;
            }
            if (k) m |= bit_range(0, 63 - jc);
            L = m;
        }
        word = (sel == 0) ? ((L & gc) | (~L & go)) : ((L & go) | (~L & gc));
        nmut = 2 + 2 * (uint32_t)C;
    } else if (C == -1) {
        // freeReco + makeFreeMask (GeneUtils.cpp:322-360): two draws per block
        const uint32_t n0 = 2 + (uint32_t)(s * 2 * nbk + 2 * b);
        uint64_t v = ((uint64_t)D(n0) << 32) + D(n0 + 1);
        v &= 0x5555555555555555ull;
        v += v << 1;
        const uint64_t a0 = par[b], a1 = par[nbk + b];
        word = (sel == 0) ? ((v & a0) | (~v & a1)) : ((v & a1) | (~v & a0));
        nmut = 2 + 4 * (uint32_t)nbk;
    } else {
        word = par[sel * nbk + b];
        nmut = 2;
    }
    if (p.nbino > 0) {
        // BinomialDist::getN then GeneUtils::mutateNucs (GeneUtils.cpp:112-141)
        uint32_t dm[4];
        const uint32_t mb = nmut >> 2;
        draw4(x.seed, step, QHG_STREAM_GENETICS, nid, mb, dm);
        auto M = [&](uint32_t n) -> uint32_t {
            if ((n >> 2) == mb) {
                uint32_t v = dm[0];
#pragma unroll
                for (int q = 1; q < 4; q++) if ((uint32_t)q == (n & 3u)) v = dm[q];
                return v;
            }
            return draw(x.seed, step, QHG_STREAM_GENETICS, nid, n);
        };
        const double um = wrandd(M(nmut));
        int k = 0;
        while (k < p.nbino && um > p.bino[k]) k++;
        const uint32_t nbits_all = 2u * 2u * (uint32_t)p.gen_size;
        for (int j = 0; j < k; j++) {
            const uint32_t bp = wrandi(0, nbits_all, 2, M(nmut + 1 + 2 * (uint32_t)j));
            if ((int)(bp >> 6) == w) {
                const uint64_t mk = (uint64_t)wrandi(1, 4, 1, M(nmut + 2 + 2 * (uint32_t)j));
                word ^= mk << (bp & 63u);
            }
        }
    }
    return word;
}

// k_genome: the step's births, one lane per genome word (64 / gwords births
// per wave when a genome has at most 64 words)
__global__ void __launch_bounds__(kWave) k_genome(Ctx x, Prog p, StepArgs sa) {
    if (x.st->error) return;
    const int lane = threadIdx.x;
    const int pr = (int)sa.parity;
    const int64_t B = x.st->births_last;
    const int64_t max_id = x.st->max_id[pr];
    const int64_t gbase = x.st->gpool[pr];
    const int gw = x.gwords, nbk = gw / 2;
    const int per = (gw <= kWave) ? kWave / gw : 1;
    const int pb = (gw <= kWave) ? lane / gw : 0;
    for (int64_t base = (int64_t)blockIdx.x * per; base < B; base += (int64_t)gridDim.x * per) {
        const int64_t rg = base + pb;
        if (pb >= per || rg >= B) continue;
        const int32_t ms = x.brec[2 * rg], fs = x.brec[2 * rg + 1];
        const int64_t hm = x.cur.ghandle[ms], hf = x.cur.ghandle[fs];
        const uint64_t nid = (uint64_t)(max_id + 1 + rg);
        uint64_t *out = x.gpool + (gbase + rg) * gw;
        for (int w = (gw <= kWave) ? lane % gw : lane; w < gw; w += (gw <= kWave) ? gw : kWave) {
            const int s = w / nbk, b = w % nbk;
            const uint64_t *par = x.gpool + (s ? hf : hm) * gw;
            out[w] = child_word(x, p, sa.step, nid, par, s, b, nbk, w);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched steps: many independent populations per launch.  Each member keeps
// its own buffers, status and parameters; a device table holds its context
// for both step parities.  Work items (groups of occupied cells, scan tiles,
// births) are numbered across the members by prefixes computed in every block
// from the members' own status, so one launch per phase steps every member
// exactly as it would step alone.
// ---------------------------------------------------------------------------
constexpr int kMaxPops = 256;          // members per batch (4 per lane)

struct PopDesc {
    Ctx ctx[2];               // by the member's step parity
    Prog prog;
    int64_t step0;            // the member's step and time when it joined,
    int64_t join;             // at batch step `join`
    double t0;
    int64_t pad;
};

// The batch step index lives in device memory, one copy per kernel of the
// step (act, scan, scatter, genome), so the launch arguments stay the same
// from step to step and one step can be captured in a graph and replayed.
// Every kernel reads its own copy; block 0 sets the next kernel's copy, which
// no running kernel reads, and the last kernel of the step advances the act
// kernel's copy for the next step.
enum BatchCounter { BC_ACT, BC_SCAN, BC_SCATTER, BC_GENOME, BC_COUNT };

struct BatchArgs {
    const PopDesc *pops;
    int32_t npops;
    int32_t total_tiles;      // scan tiles of all members
    const int32_t *tile_base; // npops + 1: first global tile of each member
    uint64_t *tiles;          // per tile of every member: agents | births totals
    uint64_t *tiles2;         //   ... and occupied-cell totals (k_batch_scan_sums)
    int64_t *bctr;            // BC_COUNT copies of the batch step index
    double group_agents;      // target agents per wave group
    int32_t genome;           // k_batch_genome runs in this step
    int32_t pad;
};

// values every lane holds alike (ballot and shuffle results), as such
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    return (T *)(uintptr_t)uniform_i64((int64_t)(uintptr_t)p);
}

__device__ __forceinline__ int64_t batch_step_in(const BatchArgs &B, int k, int next) {
    const int64_t b = uniform_i64(B.bctr[k]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (next >= 0) B.bctr[next] = b;
        if (next < 0 || (k == BC_SCATTER && !B.genome)) B.bctr[BC_ACT] = b + 1;
    }
    return b;
}

// the scan loads every cell's incoming runs at once when at least one cell
// in 2^shift is occupied: worth it where the scan is latency-bound (one
// population: shift 3), less so where many tiles keep it bandwidth-bound
// (a batch: shift 1).  Results do not depend on it.
__host__ __device__ __forceinline__ int32_t scan_dense(int64_t occ, int64_t ncells, int shift) {
    return ((occ << shift) >= ncells) ? 1 : 0;
}

// cells per wave group: about `target` agents per group from the mean
// occupancy (results do not depend on it)
__host__ __device__ __forceinline__ int32_t group_cells(double target, int64_t alive, int64_t occ) {
    const double o = (double)(occ > 0 ? occ : 1);
    double mean = (double)alive / o;
    mean = mean > 1.0 ? mean : 1.0;
    const double g = floor(target / mean);
    return (int32_t)(g < 1.0 ? 1.0 : (g > (double)kGroupCells ? (double)kGroupCells : g));
}

// The descriptor table is written by the host before a launch and by no
// kernel: read through the constant address space, its fields go to scalar
// registers (and stay there across the kernels' own stores).
typedef __attribute__((address_space(4))) const PopDesc ConstPopDesc;
typedef __attribute__((address_space(4))) const Ctx ConstCtx;
typedef __attribute__((address_space(4))) const Prog ConstProg;

__device__ __forceinline__ ConstPopDesc &pop_desc(const BatchArgs &B, int r) {
    return ((ConstPopDesc *)B.pops)[r];
}

// A member's step arguments: step and time advance by one per batch step
// from where the member joined (fixed for the whole batch step, whatever the
// kernels write into its status); the group size follows its occupancy.
template <typename Desc>
__device__ __forceinline__ StepArgs pop_step_args(const Desc &d, int64_t bstep, double group_agents) {
    StepArgs a{};
    const int64_t step = d.step0 + (bstep - d.join);
    a.step = (uint32_t)step;
    a.parity = (uint32_t)(step & 1);
    a.t = d.t0 + (double)(bstep - d.join);
    const DevStatus *st = d.ctx[0].st;
    a.gcells = group_cells(group_agents, st->n_alive[a.parity], st->n_occ[a.parity]);
    a.dense = scan_dense(st->n_occ[a.parity], d.ctx[0].ncells, 1);
    return a;
}

// the same for a member every lane of the wave works on (its values as such)
template <typename Desc>
__device__ __forceinline__ StepArgs pop_step_args_uniform(const Desc &d, int64_t bstep, double group_agents) {
    StepArgs a = pop_step_args(d, bstep, group_agents);
    a.gcells = uniform_i32(a.gcells);
    a.dense = uniform_i32(a.dense);
    return a;
}

// Exclusive prefix over the members of per-member counts (lane l holds
// members l, l+64, l+128, l+192); returns the total.
__device__ __forceinline__ int64_t pop_prefix(const int64_t (&cnt)[4], int64_t (&base)[4]) {
    int64_t carry = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t incl = wave_incl_scan_u64((uint64_t)cnt[k]);
        base[k] = carry + (int64_t)incl - cnt[k];
        carry += (int64_t)__shfl((long long)incl, 63, 64);
    }
    return carry;
}

// the member owning global item g: the last member whose base is <= g
__device__ __forceinline__ int pop_of(const int64_t (&base)[4], int npops, int64_t g) {
    const int lane = threadIdx.x & 63;
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) n += __popcll(__ballot(lane + 64 * k < npops && base[k] <= g));
    return uniform_i32(n - 1);
}

__device__ __forceinline__ int64_t lane_value(const int64_t (&v)[4], int r) {
    const int k = r >> 6, l = r & 63;
    int64_t out = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (q == k) out = (int64_t)__shfl((long long)v[q], l, 64);
    return uniform_i64(out);
}

// per-member group counts of the current step
__device__ __forceinline__ void batch_groups(const BatchArgs &B, int64_t bstep, int64_t (&cnt)[4]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = lane + 64 * k;
        cnt[k] = 0;
        if (r < B.npops) {
            const DevStatus *st = pop_desc(B, r).ctx[0].st;
            if (!st->error) {
                const StepArgs a = pop_step_args(pop_desc(B, r), bstep, B.group_agents);
                const int64_t nocc = st->n_occ[a.parity];
                cnt[k] = (nocc + a.gcells - 1) / a.gcells;
            }
        }
    }
}

template <bool kRare, bool kNav = false>
__global__ void __launch_bounds__(kCellBlock) k_batch_act(BatchArgs B) {
    __shared__ ActLds lds;
    const int lane = threadIdx.x;
    const int64_t bstep = batch_step_in(B, BC_ACT, BC_SCAN);
    int64_t cnt[4], base[4];
    batch_groups(B, bstep, cnt);
    const int64_t total = pop_prefix(cnt, base);
    // a contiguous run of groups per block, so a block stays with one member
    // for long stretches and adds its counts to that member once per stretch
    const int64_t per = (total + gridDim.x - 1) / gridDim.x;
    const int64_t g0 = (int64_t)blockIdx.x * per;
    const int64_t g1 = (g0 + per < total) ? g0 + per : total;
    // one stretch per member: its context is fixed for the inner loop
    for (int64_t g = g0; g < g1;) {
        const int r = pop_of(base, B.npops, g);
        const int64_t rbase = lane_value(base, r);
        const int64_t rend0 = rbase + lane_value(cnt, r);
        const int64_t rend = rend0 < g1 ? rend0 : g1;
        const StepArgs sa = pop_step_args_uniform(pop_desc(B, r), bstep, B.group_agents);
        ConstCtx &x = *uniform_ptr(&pop_desc(B, r).ctx[sa.parity]);
        ConstProg &pg = *uniform_ptr(&pop_desc(B, r).prog);
        const int64_t nocc = uniform_i64(x.st->n_occ[sa.parity]);
        const int64_t gc = sa.gcells;
        uint32_t deaths = 0, moves = 0, fert = 0;
        for (; g < rend; g++) {
            const int64_t i0 = (g - rbase) * gc;
            const int ng = (int)((nocc - i0 < gc) ? (nocc - i0) : gc);
            OccCell e{0, 0, 0, 0};
            if (lane < ng) e = x.occ_cur[i0 + lane];
            const uint32_t S0 = readlane_u32(e.s, 0);
            const uint32_t T = readlane_u32(e.s, ng - 1) + readlane_u32(e.n, ng - 1) - S0;
            const uint32_t nmax = wave_max_u32(lane < ng ? e.n : 0u);
            if (T <= (uint32_t)kGroupSlots && nmax <= (uint32_t)kGroupCellMax) {
                group_act<kRare, false, kNav>(x, pg, sa, ng, e, e.s - S0, S0, T, lds.g, deaths, moves, fert);
                continue;
            }
            for (int k = 0; k < ng; k++) {
                const int32_t c = (int32_t)readlane_u32((uint32_t)e.c, k);
                const uint32_t s = readlane_u32(e.s, k);
                const int32_t n = (int32_t)readlane_u32(e.n, k);
                PairLists L;
                if (n > kLdsCell) {
                    L = PairLists{x.gfu + s, x.gmu + s, x.gfid + s, x.gmid + s, x.gfs + s, x.gms + s,
                                  x.gmate + s, x.gflag + s};
                } else {
                    L = PairLists{lds.c.fu, lds.c.mu, lds.c.fid, lds.c.mid, lds.c.fs, lds.c.ms, lds.c.mate,
                                  lds.c.flag};
                }
                cell_act<kRare, false, kNav>(x, pg, sa, c, s, n, L, deaths, moves, fert);
            }
        }
        const uint64_t dm = wave_sum_u64(((uint64_t)moves << 32) | deaths);
        const uint32_t fe = (uint32_t)wave_sum_u64(fert);
        if (lane == 0) {
            atomicAdd((unsigned long long *)&x.st->acc_deaths, (unsigned long long)(uint32_t)dm);
            atomicAdd((unsigned long long *)&x.st->acc_moves, (unsigned long long)(dm >> 32));
            atomicAdd((unsigned long long *)&x.st->fertile_total, (unsigned long long)fe);
        }
    }
}

// The batched scan runs in two passes, so no block ever waits on another
// (a look-back across thousands of member tiles would serialise on its tile
// handout, and static tile ownership could wait on blocks that other
// streams keep from running).  Pass 1: every tile's totals (agents |
// births, occupied cells) into the member's tile words; idle tiles are zero.
__global__ void __launch_bounds__(kScanBlock) k_batch_scan_sums(BatchArgs B) {
    __shared__ ScanLds lds;
    const int lane = threadIdx.x & 63;
    const int64_t bstep = uniform_i64(B.bctr[BC_SCAN]);
    int64_t tb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = lane + 64 * k;
        tb[k] = (r < B.npops) ? B.tile_base[r] : 0;
    }
    for (int64_t t = blockIdx.x; t < B.total_tiles; t += gridDim.x) {
        const int r = pop_of(tb, B.npops, t);
        const int64_t lt = t - lane_value(tb, r);
        const DevStatus *st = pop_desc(B, r).ctx[0].st;
        if (st->error) continue;
        const StepArgs sa = pop_step_args_uniform(pop_desc(B, r), bstep, B.group_agents);
        const ConstCtx &x = *uniform_ptr(&pop_desc(B, r).ctx[sa.parity]);
        uint64_t tot = 0, otot = 0;
        if (!tile_idle(x, (uint32_t)lt)) {
            TileIn I;
            tile_load(x, sa, (uint32_t)lt, false, I);
            tile_ready(I);
            const TileSums S = tile_sums(I, lds);
            tot = S.tot;
            otot = S.otot;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            B.tiles[t] = tot;
            B.tiles2[t] = otot;
        }
    }
}

// Pass 2: each busy tile sums its member's earlier tile totals for its
// exclusive prefix, then scans and writes as k_cell_scan does.
__global__ void __launch_bounds__(kScanBlock) k_batch_scan(BatchArgs B) {
    __shared__ ScanLds lds;
    const int lane = threadIdx.x & 63;
    const int64_t bstep = batch_step_in(B, BC_SCAN, BC_SCATTER);
    int64_t tb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = lane + 64 * k;
        tb[k] = (r < B.npops) ? B.tile_base[r] : 0;
    }
    for (int64_t t = blockIdx.x; t < B.total_tiles; t += gridDim.x) {
        const int r = pop_of(tb, B.npops, t);
        const int64_t first = lane_value(tb, r);
        const int64_t lt = t - first;
        const DevStatus *st = pop_desc(B, r).ctx[0].st;
        if (st->error) continue;
        const StepArgs sa = pop_step_args_uniform(pop_desc(B, r), bstep, B.group_agents);
        const ConstCtx &x = *uniform_ptr(&pop_desc(B, r).ctx[sa.parity]);
        if (tile_idle(x, (uint32_t)lt)) continue;
        if (threadIdx.x < 64) {
            uint64_t a = 0, o = 0;
            for (int64_t u = first + lane; u < t; u += 64) {
                a += B.tiles[u];
                o += B.tiles2[u];
            }
            a = wave_sum_u64(a);
            o = wave_sum_u64(o);
            if (lane == 0) {
                lds.excl = a;
                lds.excl2 = o;
            }
        }
        TileIn I;
        tile_load(x, sa, (uint32_t)lt, true, I);
        tile_ready(I);
        tile_stage(I, lds);
        tile_clear(x, (uint32_t)lt, I);
        const TileSums S = tile_sums(I, lds);
        tile_write(x, sa, (uint32_t)lt, I, lds, lds.excl + S.pre, lds.excl2 + S.opre);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kCellBlock) k_batch_scatter(BatchArgs B) {
    __shared__ ScatLds lds;
    const int lane = threadIdx.x;
    const int64_t bstep = batch_step_in(B, BC_SCATTER, BC_GENOME);
    if (blockIdx.x == 0) {
        // close every member's step and copy its status into host memory,
        // one lane per member
        for (int r = lane; r < B.npops; r += kWave) {
            DevStatus *st = pop_desc(B, r).ctx[0].st;
            if (!st->error) {
                const int64_t dd = st->acc_deaths, dm = st->acc_moves;
                st->deaths_last = dd;
                st->moves_last = dm;
                st->deaths_total += dd;
                st->moves_total += dm;
                st->acc_deaths = 0;
                st->acc_moves = 0;
            }
            DevStatus *hr = pop_desc(B, r).ctx[0].hring;
            if (hr) {
                const uint64_t *srcw = reinterpret_cast<const uint64_t *>(st);
                uint64_t *dstw = reinterpret_cast<uint64_t *>(hr + (st->step % kStatusRing));
                for (int w = 0; w < (int)(sizeof(DevStatus) / 8); w++) dstw[w] = srcw[w];
            }
        }
        // every member's snapshot of a step divisible by kStatusSealEvery is
        // sealed with its step, as k_cell_scatter does, once all of them are
        // visible to the host (one system-scope release for the wave)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        for (int r = lane; r < B.npops; r += kWave) {
            const int64_t step = pop_desc(B, r).ctx[0].st->step;
            DevStatus *hr = pop_desc(B, r).ctx[0].hring;
            if (hr && step % kStatusSealEvery == 0) {
                int64_t *seal = reinterpret_cast<int64_t *>(hr + kStatusRing) + step % kStatusRing;
                __hip_atomic_store(seal, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        return;
    }
    // the same group numbering as k_batch_act
    int64_t cnt[4], base[4];
    batch_groups(B, bstep, cnt);
    const int64_t total = pop_prefix(cnt, base);
    const int64_t G = gridDim.x - 1;
    const int64_t per = (total + G - 1) / G;
    const int64_t g0 = (int64_t)(blockIdx.x - 1) * per;
    const int64_t g1 = (g0 + per < total) ? g0 + per : total;
    // one stretch per member: its context is fixed for the inner loop
    for (int64_t g = g0; g < g1;) {
        const int r = pop_of(base, B.npops, g);
        const int64_t rbase = lane_value(base, r);
        const int64_t rend0 = rbase + lane_value(cnt, r);
        const int64_t rend = rend0 < g1 ? rend0 : g1;
        const StepArgs sa = pop_step_args_uniform(pop_desc(B, r), bstep, B.group_agents);
        ConstCtx &x = *uniform_ptr(&pop_desc(B, r).ctx[sa.parity]);
        const DevStatus *st = x.st;
        const int64_t nocc = uniform_i64(st->n_occ[sa.parity]);
        const int64_t max_id = uniform_i64(st->max_id[sa.parity]);
        const int64_t gbase = uniform_i64(st->gpool[sa.parity]);
        const int64_t gc = sa.gcells;
        for (; g < rend; g++) {
            const int64_t i0 = (g - rbase) * gc;
            const int ng = (int)((nocc - i0 < gc) ? (nocc - i0) : gc);
            OccCell e{0, 0, 0, 0};
            if (lane < ng) e = x.occ_cur[i0 + lane];
            const uint32_t S0 = readlane_u32(e.s, 0);
            const uint32_t T = readlane_u32(e.s, ng - 1) + readlane_u32(e.n, ng - 1) - S0;
            if (T <= (uint32_t)kGroupSlots) {
                group_scatter(x, sa, ng, e, e.s - S0, S0, T, lds.g, max_id, gbase);
                continue;
            }
            for (int k = 0; k < ng; k++) {
                const int32_t c = (int32_t)readlane_u32((uint32_t)e.c, k);
                const uint32_t s = readlane_u32(e.s, k);
                const int32_t n = (int32_t)readlane_u32(e.n, k);
                int32_t *mli = (n > kLdsCell) ? x.gfu + s : lds.c.mli;
                int64_t *mid = (n > kLdsCell) ? x.gfid + s : lds.c.mid;
                cell_scatter(x, sa, c, s, n, mli, mid, max_id, gbase);
            }
        }
    }
}

// births of every member with Genetics, one lane per genome word
__global__ void __launch_bounds__(kWave) k_batch_genome(BatchArgs B) {
    const int lane = threadIdx.x;
    const int64_t bstep = batch_step_in(B, BC_GENOME, -1);
    int64_t cnt[4], base[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int r = lane + 64 * k;
        cnt[k] = 0;
        if (r < B.npops) {
            ConstCtx &x0 = pop_desc(B, r).ctx[0];
            if (x0.gwords && !x0.st->error) {
                const int per = (x0.gwords <= kWave) ? kWave / x0.gwords : 1;
                cnt[k] = (x0.st->births_last + per - 1) / per;   // births of the step just scanned
            }
        }
    }
    const int64_t total = pop_prefix(cnt, base);
    for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
        const int r = pop_of(base, B.npops, g);
        const int64_t gl = g - lane_value(base, r);
        const DevStatus *st = pop_desc(B, r).ctx[0].st;
        const StepArgs sa = pop_step_args_uniform(pop_desc(B, r), bstep, B.group_agents);
        const int par = (int)sa.parity;
        ConstCtx &x = *uniform_ptr(&pop_desc(B, r).ctx[par]);
        ConstProg &pg = *uniform_ptr(&pop_desc(B, r).prog);
        const int gw = x.gwords, nbk = gw / 2;
        const int per = (gw <= kWave) ? kWave / gw : 1;
        const int pb = (gw <= kWave) ? lane / gw : 0;
        const int64_t rg = gl * per + pb;
        if (pb >= per || rg >= st->births_last) continue;
        const int64_t max_id = st->max_id[par], gbase = st->gpool[par];
        const int32_t ms = x.brec[2 * rg], fs = x.brec[2 * rg + 1];
        const int64_t hm = x.cur.ghandle[ms], hf = x.cur.ghandle[fs];
        const uint64_t nid = (uint64_t)(max_id + 1 + rg);
        uint64_t *out = x.gpool + (gbase + rg) * gw;
        for (int w = (gw <= kWave) ? lane % gw : lane; w < gw; w += (gw <= kWave) ? gw : kWave) {
            const int s = w / nbk, b = w % nbk;
            const uint64_t *par_g = x.gpool + (s ? hf : hm) * gw;
            out[w] = child_word(x, pg, sa.step, nid, par_g, s, b, nbk, w);
        }
    }
}

// ---------------------------------------------------------------------------
// genome pool compaction: the live agents' genomes to handles 0..n-1 in slot
// order (into a second pool); the handles are rewritten
// ---------------------------------------------------------------------------
// Genome pool compaction: the genomes of the buffer's n live agents move to
// rows 0..n-1 of dst.  Lanes take consecutive words of a row (several rows
// per wave when a row is shorter than a wave), so every row moves as one
// coalesced read and write; each wave keeps kPoolUnroll row groups in flight.
// The handles are renumbered afterwards, by k_gpool_commit.
constexpr int kPoolUnroll = 4;

__global__ void __launch_bounds__(256) k_gpool_compact(Ctx x, uint64_t *dst, int32_t parity) {
    const int64_t n = x.st->n_alive[parity];
    const int64_t gw = x.gwords;
    const int lane = threadIdx.x & 63;
    const int rpw = gw <= 64 ? (int)(64 / gw) : 1;          // rows per wave pass
    const int r = gw <= 64 ? lane / (int)gw : 0;            // this lane's row
    const int q0 = gw <= 64 ? lane - r * (int)gw : lane;    // ... and first word
    const bool on = r < rpw;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
    const int64_t span = (int64_t)rpw * kPoolUnroll;
    for (int64_t base = wave * span; base < n; base += nwaves * span) {
        int64_t h[kPoolUnroll];
#pragma unroll
        for (int u = 0; u < kPoolUnroll; u++) {
            const int64_t i = base + (int64_t)u * rpw + r;
            h[u] = (on && i < n) ? (int64_t)x.cur.ghandle[i] : -1;
        }
        for (int64_t q = q0; q < gw; q += 64) {
            uint64_t v[kPoolUnroll];
#pragma unroll
            for (int u = 0; u < kPoolUnroll; u++) v[u] = (h[u] >= 0) ? x.gpool[h[u] * gw + q] : 0ull;
#pragma unroll
            for (int u = 0; u < kPoolUnroll; u++)
                if (h[u] >= 0) dst[(base + (int64_t)u * rpw + r) * gw + q] = v[u];
        }
    }
}

// after the copy: agent i's genome is row i, and the pool holds n rows
__global__ void __launch_bounds__(256) k_gpool_commit(Ctx x, int32_t parity) {
    const int64_t n = x.st->n_alive[parity];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x.cur.ghandle[i] = (int32_t)i;
    if (blockIdx.x == 0 && threadIdx.x == 0) x.st->gpool[parity] = n;
}

}  // namespace qhg
