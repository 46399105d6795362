// Device side of the population step engine (gfx950, wave64).
//
// One step = five launches on the handle's stream, all with fixed grids so a
// step can be captured once into a hipGraph:
//   k_bucket   per agent slot   RandomPair::initialize (mate := -3) and the
//                               per-cell bucketing of FERTILE agents
//                               (QHG4/actions/RandomPair.cpp:107-181)
//   k_pair     per cell         sort candidates by id, pick the k-th
//                               unpaired partner (RandomPair.cpp:183-276)
//   k_act      per agent slot   every Action<T>::execute of the step fused in
//                               priority order (QHG4/core/SPopulation.cpp:554-577)
//                               plus a decoupled look-back scan that writes
//                               the birth records in mother-id order
//   k_newborn  per birth        createAgentAtIndex + makePopSpecificOffspring
//                               + Genetics::makeOffspring
//                               (SPopulation.cpp:838-918, Genetics.cpp:285-337)
//   k_cellscan per cell         updateNumAgentsPerCell result -> per-cell
//                               offsets for the next pairing; step bookkeeping
//                               (SPopulation.cpp:509-544)
//
// Agents live in structure-of-arrays slots [0, hw) in id order: births are
// appended at hw with increasing ids and compaction is stable, so slot order
// is id order.  Dead agents stay as holes (life == 0) until compaction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"
#include "../../include/qhg_gpu.h"

namespace qhg {

constexpr int kBlock = 256;          // threads per block (4 waves)
#ifndef QHG_ACT_ITEMS
#define QHG_ACT_ITEMS 1
#endif
#ifndef QHG_PAIR_LDS
#define QHG_PAIR_LDS 32
#endif
constexpr int kActItems = QHG_ACT_ITEMS;   // agent slots per thread per tile
constexpr int kActTile = kBlock * kActItems;   // agent slots per look-back tile
constexpr int kCellTile = 1024;      // cells per look-back tile (4 per thread)
constexpr int kMaxOps = QHG_MAX_ACTIONS;
constexpr int kMaxCross = 32;        // device limit on Genetics_num_crossover
constexpr int kStatusRing = 8;       // host-mapped status snapshots, one per step
constexpr int kBucketBlocks = 512;   // fixed grid of k_bucket (per-block fertile counts)
constexpr int kSuperTiles = 64;      // act tiles per supertile (birth offsets without a scan kernel)

struct DevStatus {
    int64_t hw, n_alive, max_id, step;
    double t;
    int64_t births_cur, deaths_cur, moves_cur;
    int64_t births_last, deaths_last, moves_last;
    int64_t births_total, deaths_total, moves_total;
    int64_t agent_steps;     // sum over steps of the live count at step start
    int64_t slot_steps;      // sum over steps of the high-water mark at step start
    int64_t fertile_total;   // agents bucketed for pairing, all steps
    uint32_t tile_ctr_act, tile_ctr_cell, tile_ctr_cmp, pad0;
    int32_t error, pad1;
};

// device error bits
constexpr int32_t kErrCapacity = 1;   // births did not fit
constexpr int32_t kErrBucket = 2;     // fertile agents exceed the cell's live count

struct AgentPtrs {
    int64_t *id;
    int32_t *cell;
    uint8_t *life;
    uint8_t *gender;
    double *birth;
    double *age;
    double *lastb;
    int32_t *mate;
    int32_t *nbabies;
    uint64_t *genome;
};

struct Ctx {
    AgentPtrs a;
    int64_t cap;
    int32_t gwords;        // u64 per genome (both strands), 0 = no genetics
    int32_t ncells, conn;
    const int32_t *nbr;
    const uint8_t *nn;
    double *K;
    double *W;
    const uint8_t *ice;
    uint32_t *cnt[2];
    uint32_t *off;         // ncells + 1
    uint32_t *fillF, *fillM;
    uint64_t *bucket;      // cap entries: slot << 32 | the agent's pairing draw
    int32_t *brec;         // 3 * (cap + kActTile): birth records staged per act tile
    uint32_t *tile_cnt;    // 3 per act tile: births, deaths, moves
    uint32_t *super_cnt;   // 3 per group of kSuperTiles act tiles (sums of tile_cnt)
    uint32_t *blk_fert;    // fertile agents bucketed per k_bucket block
    uint64_t *tiles_cell;
    int64_t ntiles_act_max;
    DevStatus *st;
    DevStatus *hring;      // host-mapped ring of kStatusRing snapshots (may be null)
    uint64_t seed;
};

struct Prog {
    int32_t nops;
    int32_t ops[kMaxOps];
    uint64_t opword;         // ops packed 4 bits each (op q in bits 4q..4q+3)
    int32_t load_age;        // Fertility runs before GetOld (reads stored age)
    int32_t has_pair;
    // fields the program reads (set by the host from the action list)
    int32_t need_gender, need_birth, need_lastb, need_mate, need_count, need_K, need_move;
    int32_t need_act0, need_act1;   // Philox blocks 0 / 1 of the ACT stream
    double fert_min, fert_max, fert_inter;
    double b0, d0, theta, Kconst;
    double rmove_p, wmove_p;
    int32_t gen_size, gen_cross, nbino, pad;
    double bino[QHG_MAX_BINO];
};

// ---------------------------------------------------------------------------
// decoupled look-back (single pass scan), one word per tile:
// [63:62] flag (1 aggregate, 2 inclusive), [61:0] value.  Tile words are
// zeroed by the kernel launched before the scanning kernel.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

// The look-back only exchanges the tile words themselves (no other data
// written in the same launch is read by another block), so relaxed
// agent-scope atomics suffice and no cache write-back is needed.
__device__ __forceinline__ void tile_publish(uint64_t *w, uint64_t v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t tile_read(uint64_t *w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Called by all 64 lanes of one wave.  Publishes the tile's aggregate, then
// reads 64 predecessor tiles per probe: the nearest inclusive prefix ends the
// walk; aggregates on the way are summed.  Returns the exclusive prefix of
// `tile` (same value in every lane) and publishes the tile's inclusive value.
constexpr int32_t kErrLookback = 4;   // a look-back wait exceeded its spin limit
constexpr uint32_t kLookbackSpins = 1u << 22;

__device__ uint64_t lookback(uint64_t *tiles, uint32_t tile, uint64_t agg, int32_t *err) {
    const int lane = threadIdx.x & 63;
    if (tile == 0) {
        if (lane == 0) tile_publish(&tiles[0], kFlagInc | agg);
        return 0;
    }
    if (lane == 0) tile_publish(&tiles[tile], kFlagAgg | agg);
    uint64_t excl = 0;
    int64_t top = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (true) {
        const int64_t j = top - lane;
        const uint64_t w = (j >= 0) ? tile_read(&tiles[j]) : kFlagInc;
        const uint64_t flag = w & ~kValMask;
        const uint64_t inc = __ballot(flag == kFlagInc);
        const uint64_t nready = __ballot(flag == 0);
        const int first = inc ? (__ffsll((unsigned long long)inc) - 1) : 64;
        const uint64_t need = (first >= 63) ? ~0ull : ((2ull << first) - 1ull);
        if (nready & need) {
            // every wave exits: after the spin limit the step is flagged as
            // failed instead of waiting on
            if (++spins < kLookbackSpins) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            if (lane == 0) atomicOr(err, kErrLookback);
            break;
        }
        excl += wave_sum_u64(lane <= first ? (w & kValMask) : 0ull);
        if (first < 64) break;
        top -= 64;
    }
    if (lane == 0) tile_publish(&tiles[tile], kFlagInc | (excl + agg));
    return excl;
}

// exclusive block scan of one u32 per thread (256 threads, wave64)
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
        uint32_t s = s_wave[w];
        if (w < wave) wpre += s;
        tot += s;
    }
    total = tot;
    return wpre + x - v;
}

// ---------------------------------------------------------------------------
// k_bucket: RandomPair::initialize -- mate := -3 for all, FERTILE agents
// go into their cell's region [off[c], off[c+1]): females from the front,
// males from the back (region size = live count of the cell >= F + M).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_bucket(Ctx x) {
    __shared__ uint32_t s_fert;
    if (threadIdx.x == 0) s_fert = 0;
    __syncthreads();
    uint32_t my_fert = 0;
    const int64_t hw = x.st->hw;
    const uint32_t step = (uint32_t)x.st->step;
    constexpr int I = 4;
    for (int64_t b0 = (int64_t)blockIdx.x * kBlock * I; b0 < hw; b0 += (int64_t)gridDim.x * kBlock * I) {
        uint8_t l[I];
        int32_t c[I];
        uint8_t g[I];
        int64_t id[I];
#pragma unroll
        for (int k = 0; k < I; k++) {
            const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
            l[k] = (i < hw) ? x.a.life[i] : (uint8_t)0;
        }
#pragma unroll
        for (int k = 0; k < I; k++) {
            const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
            if (l[k] == 0) continue;
            x.a.mate[i] = -3;
            if (l[k] == QHG_LIFE_FERTILE) { c[k] = x.a.cell[i]; g[k] = x.a.gender[i]; id[k] = x.a.id[i]; }
        }
        uint32_t pos[I], lo[I], hi[I];
#pragma unroll
        for (int k = 0; k < I; k++) {
            if (l[k] != QHG_LIFE_FERTILE || g[k] > 1) continue;
            lo[k] = x.off[c[k]];
            hi[k] = x.off[c[k] + 1];
            pos[k] = atomicAdd(g[k] == 0 ? &x.fillF[c[k]] : &x.fillM[c[k]], 1u);
        }
#pragma unroll
        for (int k = 0; k < I; k++) {
            if (l[k] != QHG_LIFE_FERTILE || g[k] > 1) continue;
            const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
            my_fert++;
            // the draw this agent uses if it chooses a partner (RandomPair.cpp:212, :249)
            const uint32_t u = draw(x.seed, step, QHG_STREAM_PAIR, (uint64_t)id[k], 0);
            const uint64_t e = ((uint64_t)i << 32) | u;
            if (pos[k] < hi[k] - lo[k]) x.bucket[g[k] == 0 ? lo[k] + pos[k] : hi[k] - 1 - pos[k]] = e;
            else atomicOr(&x.st->error, kErrBucket);
        }
    }
    if (my_fert) atomicAdd(&s_fert, my_fert);
    __syncthreads();
    if (threadIdx.x == 0) x.blk_fert[blockIdx.x] = s_fert;
}

__device__ void sort_u64(uint64_t *v, uint32_t n) {
    if (n < 2) return;
    // shell sort (Ciura gaps, extended by x2.25 above 1750)
    const uint32_t gaps[] = {1750, 701, 301, 132, 57, 23, 10, 4, 1};
    uint32_t g0 = 1750;
    while (g0 * 9 / 4 < n) g0 = g0 * 9 / 4;
    for (uint32_t gap = g0; gap > 1750; gap = gap * 4 / 9) {
        for (uint32_t i = gap; i < n; i++) {
            const uint64_t k = v[i];
            uint32_t j = i;
            while (j >= gap && v[j - gap] > k) { v[j] = v[j - gap]; j -= gap; }
            v[j] = k;
        }
    }
    for (int gi = 0; gi < 9; gi++) {
        const uint32_t gap = gaps[gi];
        if (gap >= n) continue;
        for (uint32_t i = gap; i < n; i++) {
            const uint64_t k = v[i];
            uint32_t j = i;
            while (j >= gap && v[j - gap] > k) { v[j] = v[j - gap]; j -= gap; }
            v[j] = k;
        }
    }
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
// k_pair: RandomPair::findMates per cell (RandomPair.cpp:183-276).  The
// smaller sex chooses in id order; each choice takes the k-th still
// unpaired member of the other sex, k = (int)wrandr(0, left).  One thread
// per cell; candidates (slots, i.e. id order) are sorted in LDS when a sex
// has at most kPairLds of them, in the bucket array otherwise.
// ---------------------------------------------------------------------------
constexpr int kPairBlock = 64;
constexpr int kPairLds = QHG_PAIR_LDS;

// One cell's selection loop.  chooser/pool entries are slot << 32 | draw
// (both sorted by slot = id); the chooser's draw picks k = (int)wrandr(0,
// left), the k-th still unpaired pool member.
__device__ __forceinline__ void pair_select(const Ctx &x, const uint64_t *chooser, uint32_t nch,
                                            uint64_t *pool, uint32_t np) {
    uint32_t left = np;
    if (np <= 64) {
        uint64_t avail = (np == 64) ? ~0ull : ((1ull << np) - 1ull);
        for (uint32_t j = 0; j < nch && left > 0; j++) {
            const uint64_t ea = chooser[j];
            const int k = (int)wrandr(0, (double)(int32_t)left, (uint32_t)ea);
            const int pos = select_bit(avail, k);
            avail &= ~(1ull << pos);
            const uint32_t a = (uint32_t)(ea >> 32), b = (uint32_t)(pool[pos] >> 32);
            x.a.mate[a] = (int32_t)b;
            x.a.mate[b] = (int32_t)a;
            left--;
        }
    } else {
        // more than 64 candidates: mark taken partners with bit 63 (slots < 2^31)
        for (uint32_t j = 0; j < nch && left > 0; j++) {
            const uint64_t ea = chooser[j];
            const int k = (int)wrandr(0, (double)(int32_t)left, (uint32_t)ea);
            int found = -1;
            uint32_t m = 0;
            for (m = 0; m < np; m++) {
                if (!(pool[m] >> 63)) { found++; if (found == k) break; }
            }
            const uint32_t a = (uint32_t)(ea >> 32), b = (uint32_t)(pool[m] >> 32);
            pool[m] |= 1ull << 63;
            x.a.mate[a] = (int32_t)b;
            x.a.mate[b] = (int32_t)a;
            left--;
        }
    }
}

__device__ __forceinline__ void lds_insertion_sort(uint64_t *v, uint32_t n) {
    for (uint32_t i = 1; i < n; i++) {
        const uint64_t k = v[i];
        int32_t j = (int32_t)i - 1;
        while (j >= 0 && v[j] > k) { v[j + 1] = v[j]; j--; }
        v[j + 1] = k;
    }
}

__global__ void __launch_bounds__(kPairBlock) k_pair(Ctx x, int32_t do_pair) {
    __shared__ uint64_t s_buf[kPairBlock * 2 * kPairLds];
    {
        const int64_t ntiles = (x.st->hw + kActTile - 1) / kActTile;
        const int64_t ns = 3 * ((ntiles + kSuperTiles - 1) / kSuperTiles);
        for (int64_t j = (int64_t)blockIdx.x * kPairBlock + threadIdx.x; j < ns; j += (int64_t)gridDim.x * kPairBlock)
            x.super_cnt[j] = 0;
    }
    if (!do_pair) return;
    if (blockIdx.x == 0) {
        // fold the per-block fertile counts of k_bucket (diagnostics)
        __shared__ uint32_t s_sum;
        if (threadIdx.x == 0) s_sum = 0;
        __syncthreads();
        uint32_t v = 0;
        for (int j = threadIdx.x; j < kBucketBlocks; j += kPairBlock) v += x.blk_fert[j];
        if (v) atomicAdd(&s_sum, v);
        __syncthreads();
        if (threadIdx.x == 0) x.st->fertile_total += s_sum;
    }
    uint64_t *lf = s_buf + threadIdx.x * 2 * kPairLds;
    uint64_t *lm = lf + kPairLds;
    for (int64_t c = (int64_t)blockIdx.x * kPairBlock + threadIdx.x; c < x.ncells;
         c += (int64_t)gridDim.x * kPairBlock) {
        const uint32_t F = x.fillF[c], M = x.fillM[c];
        if (F == 0 || M == 0) continue;
        const uint32_t lo = x.off[c], hi = x.off[c + 1];
        if (F + M > hi - lo) { atomicOr(&x.st->error, kErrBucket); continue; }
        uint64_t *fb = x.bucket + lo;
        uint64_t *mb = x.bucket + hi - M;
        if (F <= kPairLds && M <= kPairLds) {
            for (uint32_t j = 0; j < F; j++) lf[j] = fb[j];
            for (uint32_t j = 0; j < M; j++) lm[j] = mb[j];
            lds_insertion_sort(lf, F);
            lds_insertion_sort(lm, M);
            if (F <= M) pair_select(x, lf, F, lm, M);
            else pair_select(x, lm, M, lf, F);
        } else {
            sort_u64(fb, F);
            sort_u64(mb, M);
            if (F <= M) pair_select(x, fb, F, mb, M);
            else pair_select(x, mb, M, fb, F);
        }
    }
}

struct StepView {
    uint32_t step;
    double t;
    const uint32_t *cnt_prev;
    uint32_t *cnt_next;
};

// Per-item state of one tile, kept in registers (loops over k unroll).
struct ActItem {
    uint8_t life, life0, g;
    int32_t c, mate, nn;
    int64_t id;
    double birth, lastb, age, K;
    uint32_t N;
    bool age_dirty;
    uint32_t bflag;
    int32_t npend;            // registered moves (<= 2), resolved after the program
    int32_t pk0, pk1;         // their action kinds
    uint32_t pu0, pu1;        // their target draws
    uint32_t u[8];            // words 0..7 of the agent's QHG_STREAM_ACT counter
};

// The action program of one agent (compute only; every field it reads was
// loaded by the caller).  A dead agent stops.  Moves are only registered
// here (RandomMove / WeightedMove draw their decision now) and resolved
// afterwards, as performMoves applies them after all actions.
__device__ __forceinline__ void act_program(const Ctx &x, const Prog &p, const StepView &v, ActItem &a,
                                            uint32_t &deaths) {
    const uint32_t step = v.step;
    const double t = v.t;
    if (p.need_act0) draw4(x.seed, step, QHG_STREAM_ACT, (uint64_t)a.id, 0, &a.u[0]);
    if (p.need_act1) draw4(x.seed, step, QHG_STREAM_ACT, (uint64_t)a.id, 1, &a.u[4]);
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
            const double N = (double)a.N;
            double b, d;
            if (op == QHG_ACT_VERHULST) {
                b = p.b0 + (p.theta - p.b0) * (N / p.Kconst);
                d = p.d0 + (p.theta - p.d0) * (N / p.Kconst);
            } else {
                b = (a.K <= 0) ? 0.0 : p.b0 + (p.theta - p.b0) * (N / a.K);
                d = (a.K <= 0) ? 1.0 : p.d0 + (p.theta - p.d0) * (N / a.K);
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
        }
        // RandomPair, MultiEvaluator, NPPCapacity, Genetics: no execute
    }
}

// Target cell of a registered move, -1 when it stays (slot 0 or a -1 hole).
// w (conn+1 cumulative weights) and nbr (conn neighbours) are register rows,
// indexed with compile-time indices only.
__device__ __forceinline__ int32_t move_target(const Ctx &x, int kind, uint32_t u, int nn, const double (&w)[7],
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
    return nb;
}

// exclusive ranks of flags over a tile laid out as kActItems sub-tiles of
// kBlock slots (slot = base + k*kBlock + tid): rank[k] for this thread's
// item k, tile total.  s_cnt has kActItems * waves entries.
__device__ __forceinline__ void tile_ranks(const uint32_t (&flag)[kActItems], uint32_t (&rank)[kActItems],
                                           uint32_t *s_cnt, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int W = kBlock / 64;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t inwave[kActItems];
#pragma unroll
    for (int k = 0; k < kActItems; k++) {
        const uint64_t m = __ballot(flag[k] != 0);
        inwave[k] = (uint32_t)__popcll(m & lt);
        if (lane == 0) s_cnt[k * W + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t run = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < kActItems * W; j++) {
        const uint32_t c = s_cnt[j];
        tot += c;
    }
#pragma unroll
    for (int k = 0; k < kActItems; k++) {
        run = 0;
        for (int j = 0; j < k * W + wave; j++) run += s_cnt[j];
        rank[k] = run + inwave[k];
    }
    total = tot;
}

// ---------------------------------------------------------------------------
// k_act: fused doActions.  Each agent runs the step's actions in priority
// order; a dead agent stops (SPopulation.cpp:566-571).  Moves are applied at
// the end of the agent's program (performMoves, SPopulation.cpp:1058-1092):
// actions of this step see the pre-move cell, births are registered in it.
// Tiles of kActTile slots are assigned statically (tile = block, block +
// grid, ...).  A tile stages its birth records in slot (= mother id) order
// in its own region of brec and writes its birth/death/move counts, also
// summed per supertile; k_newborn derives the global birth ranks from them.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_act(Ctx x, Prog p) {
    __shared__ uint32_t s_cnt[kActItems * (kBlock / 64)];
    __shared__ uint32_t s_red[2];
    const int64_t hw = x.st->hw;
    StepView v;
    v.step = (uint32_t)x.st->step;
    v.t = x.st->t;
    v.cnt_prev = x.cnt[v.step & 1];
    v.cnt_next = x.cnt[(v.step + 1) & 1];
    const int64_t ntiles = (hw + kActTile - 1) / kActTile;
    const int R = x.conn + 1;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t base = tile * kActTile;
        uint32_t my_deaths = 0, my_moves = 0;
        if (threadIdx.x == 0) { s_red[0] = 0; s_red[1] = 0; }
        ActItem a[kActItems];
        // stage A: life of every slot of the tile
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
            a[k].life = (i < hw) ? x.a.life[i] : (uint8_t)0;
            a[k].life0 = a[k].life;
            a[k].age_dirty = false;
            a[k].bflag = 0;
            a[k].npend = 0;
            a[k].g = 0;
            a[k].mate = -1;
            a[k].nn = 0;
            a[k].N = 0;
            a[k].K = 0;
            a[k].birth = a[k].lastb = a[k].age = 0;
        }
        // stage B: per-agent fields of the live slots (independent loads)
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            if (!a[k].life) continue;
            const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
            a[k].c = x.a.cell[i];
            a[k].id = x.a.id[i];
            if (p.need_gender) a[k].g = x.a.gender[i];
            if (p.need_birth) a[k].birth = x.a.birth[i];
            if (p.need_lastb) a[k].lastb = x.a.lastb[i];
            if (p.need_mate) a[k].mate = x.a.mate[i];
            if (p.load_age) a[k].age = x.a.age[i];
        }
        // stage C: per-cell fields
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            if (!a[k].life) continue;
            if (p.need_count) a[k].N = v.cnt_prev[a[k].c];
            if (p.need_K) a[k].K = x.K[a[k].c];
            if (p.need_move) a[k].nn = x.nn[a[k].c];
        }
        // stage D: the action program (compute only)
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            if (a[k].life) act_program(x, p, v, a[k], my_deaths);
        }
        // stage E: resolve registered moves (weight and neighbour rows);
        // the last registered move that lands wins
        int32_t dest[kActItems];
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            dest[k] = -1;
            if (a[k].npend == 0) continue;
            const int32_t c = a[k].c;
            int32_t nbr[6];
            double w[7];
#pragma unroll
            for (int q = 0; q < 6; q++) nbr[q] = (q < x.conn) ? x.nbr[(int64_t)c * x.conn + q] : -1;
            const bool need_w = a[k].pk0 == QHG_ACT_WEIGHTEDMOVE ||
                                (a[k].npend > 1 && a[k].pk1 == QHG_ACT_WEIGHTEDMOVE);
#pragma unroll
            for (int q = 0; q < 7; q++) w[q] = (need_w && q < R) ? x.W[(int64_t)c * R + q] : 0.0;
            int32_t nb = move_target(x, a[k].pk0, a[k].pu0, a[k].nn, w, nbr);
            if (nb >= 0) dest[k] = nb;
            if (a[k].npend > 1) {
                nb = move_target(x, a[k].pk1, a[k].pu1, a[k].nn, w, nbr);
                if (nb >= 0) dest[k] = nb;
            }
            if (dest[k] >= 0) my_moves++;
        }
        // stage F: write back
        uint32_t bflag[kActItems];
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            bflag[k] = a[k].bflag;
            if (!a[k].life0) continue;
            const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
            if (a[k].life != a[k].life0) x.a.life[i] = a[k].life;
            if (a[k].age_dirty) x.a.age[i] = a[k].age;
            if (a[k].bflag) x.a.nbabies[i] += 1;
            // the next step's counts start as a copy of this step's: only
            // deaths and moves of survivors change them
            if (dest[k] >= 0) x.a.cell[i] = dest[k];
            if (a[k].life == 0) {
                atomicSub(&v.cnt_next[a[k].c], 1u);
            } else if (dest[k] >= 0) {
                atomicSub(&v.cnt_next[a[k].c], 1u);
                atomicAdd(&v.cnt_next[dest[k]], 1u);
            }
        }
        uint32_t rank[kActItems], tile_total;
        tile_ranks(bflag, rank, s_cnt, tile_total);   // contains a barrier
        int32_t *rec = x.brec + 3 * base;
#pragma unroll
        for (int k = 0; k < kActItems; k++) {
            if (bflag[k]) {
                const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
                rec[3 * rank[k] + 0] = (int32_t)i;
                rec[3 * rank[k] + 1] = a[k].mate;
                rec[3 * rank[k] + 2] = a[k].c;
            }
        }
        if (my_deaths) atomicAdd(&s_red[0], my_deaths);
        if (my_moves) atomicAdd(&s_red[1], my_moves);
        __syncthreads();
        if (threadIdx.x == 0) {
            x.tile_cnt[3 * tile + 0] = tile_total;
            x.tile_cnt[3 * tile + 1] = s_red[0];
            x.tile_cnt[3 * tile + 2] = s_red[1];
            uint32_t *sc = x.super_cnt + 3 * (tile / kSuperTiles);
            if (tile_total) atomicAdd(&sc[0], tile_total);
            if (s_red[0]) atomicAdd(&sc[1], s_red[0]);
            if (s_red[1]) atomicAdd(&sc[2], s_red[1]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// genetics helpers: GeneUtils::makeMultiMask / crossOver (GeneUtils.cpp:156-246)
// ---------------------------------------------------------------------------
__device__ uint64_t multimask_block(const uint32_t *pos, int n, uint32_t blk) {
    // breaks of this block, in draw order, as makeMultiMask sees them
    int cnt = 0;
    for (int q = 0; q < n; q++) cnt += (pos[q] >> 6) == blk;
    uint32_t k = 1 - (uint32_t)(cnt % 2);
    uint64_t out = 0;
    // walk the block's breaks from the back of the vector
    int i = n - 1;
    while (i >= 0 && (pos[i] >> 6) != blk) i--;
    for (uint32_t j = 0; j < 64; j++) {
        if (i >= 0 && j == 64 - (pos[i] & 63u)) {
            i--;
            while (i >= 0 && (pos[i] >> 6) != blk) i--;
            k = 1 - k;
        }
        out = (out << 1) + k;
    }
    return out;
}

// one strand (sel) of crossOver(parent) -> out[0..nb)
__device__ void crossover_strand(uint64_t *out, const uint64_t *par, int nb, const uint32_t *pos,
                                 int n, int sel) {
    for (int b = 0; b < nb; b++) {
        int before = 0, in = 0;
        for (int q = 0; q < n; q++) {
            const uint32_t blk = pos[q] >> 6;
            before += (int)blk < b;
            in += (int)blk == b;
        }
        const int cur = before & 1;
        const uint64_t gc = par[cur * nb + b], go = par[(1 - cur) * nb + b];
        const uint64_t L = in ? multimask_block(pos, n, (uint32_t)b) : ~0ull;
        out[b] = (sel == 0) ? ((L & gc) | (~L & go)) : ((L & go) | (~L & gc));
    }
}

// ---------------------------------------------------------------------------
// k_newborn: one thread per birth record (rank r): slot hw + r, id
// max_id + 1 + r.  Also clears the look-back tiles / counter of k_cellscan.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t *s_red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = wave_sum_u64(v);
    __syncthreads();
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; w++) t += s_red[w];
    return t;
}

__global__ void __launch_bounds__(kBlock) k_newborn(Ctx x, Prog p) {
    __shared__ uint64_t s_red[kBlock / 64];
    const int64_t gtid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t gsize = (int64_t)gridDim.x * kBlock;
    {
        const int64_t nt = (x.ncells + kCellTile - 1) / kCellTile + 1;
        for (int64_t j = gtid; j < nt; j += gsize) x.tiles_cell[j] = 0;
    }
    const int64_t hw = x.st->hw;
    const int64_t max_id = x.st->max_id;
    const uint32_t step = (uint32_t)x.st->step;
    const double t = x.st->t;
    uint32_t *cnt_next = x.cnt[(step + 1) & 1];
    const int64_t ntiles = (hw + kActTile - 1) / kActTile;
    const int64_t nsuper = (ntiles + kSuperTiles - 1) / kSuperTiles;
    // step totals from the supertile sums (every block needs the birth total)
    uint64_t b = 0, d = 0, m = 0;
    for (int64_t sidx = threadIdx.x; sidx < nsuper; sidx += kBlock) {
        b += x.super_cnt[3 * sidx];
        d += x.super_cnt[3 * sidx + 1];
        m += x.super_cnt[3 * sidx + 2];
    }
    const int64_t nbt = (int64_t)block_sum_u64(b, s_red);
    const bool fits = hw + nbt <= x.cap;
    if (blockIdx.x == 0) {
        const int64_t dt = (int64_t)block_sum_u64(d, s_red);
        const int64_t mt = (int64_t)block_sum_u64(m, s_red);
        if (threadIdx.x == 0) {
            x.st->deaths_cur = dt;
            x.st->moves_cur = mt;
            // no slot for the births: drop them and flag the step as failed
            x.st->births_cur = fits ? nbt : 0;
            if (!fits) x.st->error |= kErrCapacity;
        }
    }
    if (nbt == 0 || !fits) return;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      const uint32_t nbirth = x.tile_cnt[3 * tile];
      // births before this tile: earlier supertiles + earlier tiles of its supertile
      const int64_t s0 = tile / kSuperTiles;
      uint64_t pre = 0;
      for (int64_t q = threadIdx.x; q < s0; q += kBlock) pre += x.super_cnt[3 * q];
      for (int64_t q = s0 * kSuperTiles + threadIdx.x; q < tile; q += kBlock) pre += x.tile_cnt[3 * q];
      const int64_t off = (int64_t)block_sum_u64(pre, s_red);
      if (nbirth == 0) continue;
      const int32_t *rec = x.brec + 3 * tile * kActTile;
      for (uint32_t j = threadIdx.x; j < nbirth; j += kBlock) {
        const int64_t r = off + j;
        const int32_t mother = rec[3 * j + 0];
        const int32_t father = rec[3 * j + 1];
        const int32_t c = rec[3 * j + 2];
        const int64_t s = hw + r;
        const int64_t nid = max_id + 1 + r;
        x.a.id[s] = nid;
        x.a.cell[s] = c;
        x.a.birth[s] = t;
        x.a.age[s] = 0.0;
        x.a.lastb[s] = 0.0;
        x.a.mate[s] = -3;
        x.a.nbabies[s] = 0;
        const uint32_t u = draw(x.seed, step, QHG_STREAM_NEWBORN, (uint64_t)nid, 0);
        const uint8_t g = (uint8_t)(2 * wrandd(u));
        x.a.gender[s] = g;
        x.a.life[s] = (g == 0) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        atomicAdd(&cnt_next[c], 1u);
        if (x.gwords) {
            // Genetics::makeOffspring, QHG4/actions/Genetics.cpp:285-337
            const int nbk = x.gwords / 2;
            const uint64_t *g1 = x.a.genome + (int64_t)mother * x.gwords;
            const uint64_t *g2 = x.a.genome + (int64_t)father * x.gwords;
            uint64_t *out = x.a.genome + s * x.gwords;
            uint32_t n = 0;
#define GD() draw(x.seed, step, QHG_STREAM_GENETICS, (uint64_t)nid, n++)
            const int i1 = (int)(2 * 1.0 * wrandd(GD()));
            const int i2 = (int)(2 * 1.0 * wrandd(GD()));
            const int C = p.gen_cross;
            if (C > 0) {
                uint32_t pos[kMaxCross];
                const uint32_t nbits = (uint32_t)nbk * 64u;
                for (int q = 0; q < C; q++) pos[q] = wrandi(0, nbits, 2, GD());
                crossover_strand(out, g1, nbk, pos, C, i1);
                for (int q = 0; q < C; q++) pos[q] = wrandi(0, nbits, 2, GD());
                crossover_strand(out + nbk, g2, nbk, pos, C, i2);
            } else if (C == -1) {
                // freeReco + makeFreeMask, GeneUtils.cpp:322-360
                for (int par = 0; par < 2; par++) {
                    const uint64_t *gp = par ? g2 : g1;
                    const int sel = par ? i2 : i1;
                    for (int b = 0; b < nbk; b++) {
                        const uint32_t hi = GD();
                        const uint32_t lo = GD();
                        uint64_t v = ((uint64_t)hi << 32) + lo;
                        v &= 0x5555555555555555ull;
                        v += v << 1;
                        const uint64_t a0 = gp[b], a1 = gp[nbk + b];
                        out[par * nbk + b] = (sel == 0) ? ((v & a0) | (~v & a1)) : ((v & a1) | (~v & a0));
                    }
                }
            } else {
                for (int b = 0; b < nbk; b++) out[b] = g1[i1 * nbk + b];
                for (int b = 0; b < nbk; b++) out[nbk + b] = g2[i2 * nbk + b];
            }
            if (p.nbino > 0) {
                // BinomialDist::getN then GeneUtils::mutateNucs (GeneUtils.cpp:112-141)
                const double um = wrandd(GD());
                int k = 0;
                while (k < p.nbino && um > p.bino[k]) k++;
                const uint32_t nbits_all = 2u * 2u * (uint32_t)p.gen_size;
                for (int j = 0; j < k; j++) {
                    const uint32_t bp = wrandi(0, nbits_all, 2, GD());
                    const uint64_t mk = (uint64_t)wrandi(1, 4, 1, GD());
                    out[bp >> 6] ^= mk << (bp & 63u);
                }
            }
#undef GD
        }
      }
    }
}

// ---------------------------------------------------------------------------
// k_cellscan: exclusive scan of the new per-cell counts -> off[] for the
// next pairing; clears the fill counters and copies the counts into the
// buffer that the next step updates incrementally.  The last tile closes the step's bookkeeping.
// mode 0: end of step (src = cnt[(step+1)&1]); mode 1: pre-loop (src =
// cnt[step&1], no bookkeeping).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_cellscan(Ctx x, int32_t mode) {
    __shared__ uint32_t s_tile, s_step;
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint64_t s_excl;
    const int64_t nc = x.ncells;
    constexpr int per = kCellTile / kBlock;
    // One tile per block; the grid is exactly the number of tiles and small
    // enough that every block is resident, so the look-back always drains.
    // The block of the last tile advances st->step only after every other
    // block has published its aggregate, i.e. after every block read the step.
    {
        if (threadIdx.x == 0) s_step = (uint32_t)x.st->step;
        __syncthreads();
        const uint32_t step = s_step;
        const uint32_t *src = mode ? x.cnt[step & 1] : x.cnt[(step + 1) & 1];
        uint32_t *clr = mode ? x.cnt[(step + 1) & 1] : x.cnt[step & 1];
        const uint32_t tile = blockIdx.x;
        const int64_t base = (int64_t)tile * kCellTile;
        if (base >= nc) return;
        uint32_t v[per];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < per; q++) {
            const int64_t c = base + (int64_t)threadIdx.x * per + q;
            v[q] = (c < nc) ? src[c] : 0u;
            sum += v[q];
        }
        uint32_t tot;
        const uint32_t pre = block_excl_scan(sum, s_wave, tot);
        if (threadIdx.x < 64) {
            const uint64_t ex = lookback(x.tiles_cell, tile, tot, &x.st->error);
            if (threadIdx.x == 0) s_excl = ex;
        }
        __syncthreads();
        uint64_t run = s_excl + pre;
#pragma unroll
        for (int q = 0; q < per; q++) {
            const int64_t c = base + (int64_t)threadIdx.x * per + q;
            if (c < nc) {
                x.off[c] = (uint32_t)run;
                x.fillF[c] = 0;
                x.fillM[c] = 0;
                clr[c] = v[q];
            }
            run += v[q];
        }
        if (base + kCellTile >= nc && threadIdx.x == kBlock - 1) {
            x.off[nc] = (uint32_t)run;
            DevStatus *st = x.st;
            if (mode == 0) {
                st->agent_steps += st->n_alive;
                st->slot_steps += st->hw;
            }
            st->n_alive = (int64_t)run;
            if (mode == 0) {
                const int64_t b = st->births_cur;
                st->hw += b;
                st->max_id += b;
                st->births_last = b;
                st->deaths_last = st->deaths_cur;
                st->moves_last = st->moves_cur;
                st->births_total += b;
                st->deaths_total += st->deaths_cur;
                st->moves_total += st->moves_cur;
                st->births_cur = 0;
                st->deaths_cur = 0;
                st->moves_cur = 0;
                st->step += 1;
                st->t += 1.0;
                if (x.hring) {
                    // snapshot for the host, slot = new step mod ring size
                    DevStatus *dst = x.hring + (st->step % kStatusRing);
                    const uint64_t *srcw = reinterpret_cast<const uint64_t *>(st);
                    uint64_t *dstw = reinterpret_cast<uint64_t *>(dst);
                    for (int q = 0; q < (int)(sizeof(DevStatus) / 8); q++) dstw[q] = srcw[q];
                }
            }
        }
        __syncthreads();
    }
}

__global__ void k_set_time(DevStatus *st, double t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->t = t;
}

// live-agent histogram (pre-loop / after upload or compaction)
__global__ void __launch_bounds__(kBlock) k_hist(Ctx x) {
    const int64_t hw = x.st->hw;
    uint32_t *cnt = x.cnt[x.st->step & 1];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < hw;
         i += (int64_t)gridDim.x * kBlock) {
        if (x.a.life[i] != 0) atomicAdd(&cnt[x.a.cell[i]], 1u);
    }
}

}  // namespace qhg
