/*
 * qhg_oracle.c -- CPU oracle of the QHG4 per-step population loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; never part of
 * the product).  A restatement of the reference step loop, action by action,
 * with Philox4x32-10 draws instead of per-thread WELL512 streams.
 *
 * Reference functions restated here (paths relative to the reference root):
 *   step driver        QHG4/core/PopLooper.cpp:166-202
 *   initializeStep     QHG4/core/SPopulation.cpp:394-417
 *   doActions          QHG4/core/SPopulation.cpp:554-577
 *   finalizeStep       QHG4/core/SPopulation.cpp:439-477
 *   births/deaths      QHG4/core/SPopulation.cpp:596-724, 838-938
 *   moves              QHG4/core/SPopulation.cpp:1014-1092
 *   counts             QHG4/core/SPopulation.cpp:509-536
 *   actions            QHG4/actions/{GetOld,Fertility,RandomPair,Verhulst,
 *                      AssortativePairHam (+ QHG4/utils/RankTable.cpp),
 *                      VerhulstVarK,LinearBirth,LinearDeath,RandomMove,
 *                      WeightedMove,SingleEvaluator,MultiEvaluator,
 *                      NPPCapacity,Genetics,Navigate}.cpp
 *   genome ops         QHG4/genes/GeneUtils.cpp:112-360
 *   binomial table     QHG4/utils/BinomialDist.cpp:60-104, bino_tools.cpp
 *   PolyLine           QHG4/utils/PolyLine.cpp:60-89
 *   WELL conversions   QHG4/utils/WELL512.h:33-39
 *
 * Schedule-independent replacements (documented deviations):
 *   - draws: Philox4x32-10 keyed (seed, step, stream, agent id, draw#);
 *   - per-cell pairing orders candidates by agent id (reference: slot index);
 *   - new ids: max_id + 1 + rank of the birth in (birth cell, mother id)
 *     order among this step's births (reference: thread-strided IDGen);
 *   - time fields are FP64 (reference: float).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#include <omp.h>

#include "qhg_oracle.h"

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; Random123)        */
/* ------------------------------------------------------------------ */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += PH_W0; k1 += PH_W1; }
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint32_t orc_draw(uint64_t seed, uint32_t step, uint32_t stream, uint64_t id, uint32_t n) {
    uint32_t ctr[4] = { step, (stream << 16) | (n >> 2), (uint32_t)id, (uint32_t)(id >> 32) };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[n & 3];
}

/* WELL512 conversion formulas, QHG4/utils/WELL512.h:33-39 (MAX_WRAND = 2^32-1) */
double orc_wrandd(uint32_t u) {
    return (1.0 * u) / (1 + 4294967295.0);
}
double orc_wrandr(double a, double b, uint32_t u) {
    return a + ((b - a) * u) / (1 + 4294967295.0);
}
uint32_t orc_wrandi(uint32_t a, uint32_t b, uint32_t s, uint32_t u) {
    uint32_t r = b - a;
    r += (r % s);
    return a + s * (uint32_t)((1.0 * r * u) / (s * (1.0 + 4294967295.0)));
}

/* ------------------------------------------------------------------ */
/* Binomial lookup table: QHG4/utils/BinomialDist.cpp:60-104 with the  */
/* Numerical-Recipes incomplete beta of QHG4/utils/bino_tools.cpp      */
/* ------------------------------------------------------------------ */
static double orc_gammaln(double xx) {
    static const double cof[6] = { 76.18009172947146, -86.50532032941677,
                                   24.01409824083091, -1.231739572450155,
                                   0.1208650973866179e-2, -0.5395239384953e-5 };
    double ser = 1.000000000190015;
    double x = xx, y = xx + 1, tmp = x + 5.5;
    tmp -= (x + 0.5) * log(tmp);
    for (int k = 0; k <= 5; k++) { ser += cof[k] / y; y++; }
    return -tmp + log(2.5066282746310005 * ser / x);
}

static double orc_betacf(double a, double b, double x) {
    const int MAXIT = 100; const double EPS = 3.0e-7; const float FPMIN = 1.0e-30f;
    double qab = a + b, qap = a + 1.0, qam = a - 1.0, c = 1.0;
    double d = 1.0 - qab * x / qap;
    if (fabs(d) < FPMIN) d = FPMIN;
    d = 1.0 / d;
    double h = d;
    int m;
    for (m = 1; m <= MAXIT; m++) {
        int m2 = 2 * m;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
        c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d; h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d; if (fabs(d) < FPMIN) d = FPMIN;
        c = 1.0 + aa / c; if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < EPS) break;
    }
    if (m > MAXIT) h = -1;
    return h;
}

double orc_betai(double a, double b, double x) {
    if (x < 0.0 || x > 1.0) return -1;
    double bt = (x == 0.0 || x == 1.0) ? 0.0
        : exp(orc_gammaln(a + b) - orc_gammaln(a) - orc_gammaln(b) + a * log(x) + b * log(1.0 - x));
    if (x < (a + 1.0) / (a + b + 2.0)) return bt * orc_betacf(a, b, x) / a;
    return 1.0 - bt * orc_betacf(b, a, 1.0 - x) / b;
}

int32_t orc_bino_table(double p, int32_t n, double eps, double *cdf, int32_t max) {
    if (!(p >= 0 && p < 1 && n > 0)) return -1;
    int32_t cnt = 0;
    int k = 1;
    double d2 = orc_betai(k, n - k + 1, p);
    while (d2 > eps && k < n) {
        if (cnt >= max) return -1;
        cdf[cnt++] = 1 - d2;
        k++;
        d2 = orc_betai(k, n - k + 1, p);
    }
    if (cnt >= max) return -1;
    cdf[cnt++] = 1 - d2;
    return cnt;
}

int32_t orc_bino_getn(const double *cdf, int32_t nbins, double u) {
    if (!(u >= 0)) return -1;
    int32_t i = 0;
    while (i < nbins && u > cdf[i]) i++;
    return i;
}

/* PolyLine::getVal, QHG4/utils/PolyLine.cpp:60-89 (npts = segments + 1) */
double orc_polyline(int32_t npts, const double *x, const double *v, double X) {
    if (npts < 2) return X;
    int32_t ns = npts - 1;
    if (X >= x[ns]) return v[ns];
    int32_t i = 0;
    while (i <= ns && X > x[i]) i++;
    if (i == 0) return v[0];
    if (i <= ns) {
        double a = (v[i] - v[i - 1]) / (x[i] - x[i - 1]);
        return v[i - 1] + a * (X - x[i - 1]);
    }
    return v[ns];
}

/* ------------------------------------------------------------------ */
/* 2-bit genome operations, QHG4/genes/GeneUtils.cpp                   */
/* ------------------------------------------------------------------ */
int32_t orc_genome_words(int32_t genome_size) { return (genome_size + 31) / 32; }

/* makeMultiMask, GeneUtils.cpp:156-175: walks bits from the MSB, consuming
 * breaks from the back of the vector (in draw order, not sorted). */
uint64_t orc_multimask(const uint32_t *breaks, int32_t n) {
    uint32_t k = 1 - (uint32_t)(n % 2);
    uint64_t out = 0;
    int32_t i = n - 1;
    for (uint32_t j = 0; j < 64; j++) {
        if (i >= 0 && j == 64 - breaks[i]) { i--; k = 1 - k; }
        out = (out << 1) + k;
    }
    return out;
}

/* crossOver, GeneUtils.cpp:184-246, with the break positions given
 * (in draw order).  out and parent hold two strands of nb words each. */
void orc_crossover(uint64_t *out, const uint64_t *parent, int32_t genome_size,
                   const uint32_t *break_pos, int32_t nbreaks) {
    int32_t nb = orc_genome_words(genome_size);
    /* group breaks by block in draw order (std::map<block, vector>) */
    uint32_t bb[64][64];
    int32_t bn[64];
    memset(bn, 0, sizeof(bn));
    for (int32_t i = 0; i < nbreaks; i++) {
        uint32_t blk = break_pos[i] / 64, bit = break_pos[i] % 64;
        bb[blk][bn[blk]++] = bit;
    }
    int32_t last = 0;
    uint32_t s = 0;
    for (int32_t blk = 0; blk < nb; blk++) {
        if (bn[blk] == 0) continue;
        int32_t cur = (int32_t)(s % 2);
        for (int32_t q = last; q < blk; q++) {
            out[q] = parent[cur * nb + q];
            out[nb + q] = parent[(1 - cur) * nb + q];
        }
        uint64_t L = orc_multimask(bb[blk], bn[blk]);
        uint64_t R = ~L;
        out[blk] = (L & parent[cur * nb + blk]) | (R & parent[(1 - cur) * nb + blk]);
        out[blk + nb] = (L & parent[(1 - cur) * nb + blk]) | (R & parent[cur * nb + blk]);
        s += (uint32_t)bn[blk];
        last = blk + 1;
    }
    if (nb > last) {
        int32_t cur = (int32_t)(s % 2);
        for (int32_t q = last; q < nb; q++) {
            out[q] = parent[cur * nb + q];
            out[nb + q] = parent[(1 - cur) * nb + q];
        }
    }
}

/* freeReco + makeFreeMask, GeneUtils.cpp:322-360 (mask = hi:lo draws) */
void orc_free_reco(uint64_t *out, const uint64_t *parent, int32_t nb,
                   const uint32_t *draws) {
    for (int32_t b = 0; b < nb; b++) {
        uint64_t v = ((uint64_t)draws[2 * b] << 32) + draws[2 * b + 1];
        v &= 0x5555555555555555ull;
        v += v << 1;
        uint64_t L = v, R = ~v;
        out[b] = (L & parent[b]) | (R & parent[nb + b]);
        out[nb + b] = (L & parent[nb + b]) | (R & parent[b]);
    }
}

/* mutateNucs, GeneUtils.cpp:112-141: genome_size = loci over both strands */
void orc_mutate_nucs(uint64_t *genome, int32_t genome_size, int32_t nmut,
                     const uint32_t *pos_draws, const uint32_t *mask_draws) {
    uint32_t nbits = 2 * (uint32_t)genome_size;
    for (int32_t i = 0; i < nmut; i++) {
        uint32_t pos = orc_wrandi(0, nbits, 2, pos_draws[i]);
        uint64_t mask = (uint64_t)orc_wrandi(1, 4, 1, mask_draws[i]);
        genome[pos / 64] ^= mask << (pos % 64);
    }
}

/* ------------------------------------------------------------------ */
/* population state                                                    */
/* ------------------------------------------------------------------ */
struct orc {
    qhg_params P;
    int32_t  ncells, conn;
    int32_t *nbr;
    uint8_t *nn;
    int64_t *cell_ids;
    /* environment */
    double  *alt, *lon, *lat, *water, *npp, *temp, *rain;
    uint8_t *ice, *coastal;
    /* per-cell */
    double  *K;         /* capacities (NPPCapacity output)   */
    double  *W;         /* env weights ncells*(conn+1)       */
    double  *single;    /* SingleEvaluator scratch           */
    double  *bc, *dc;   /* LinearBirth / LinearDeath per cell */
    uint32_t *count;
    int      me_first;
    /* action order (prio, then name) */
    int32_t  nact;
    int32_t  act[QHG_MAX_ACTIONS];
    /* agents, id order */
    int64_t  n, cap;
    int64_t *id;
    int32_t *cell, *mate, *nbabies, *dest;
    uint8_t *life, *gender, *bflag;
    double  *birth, *age, *lastb;
    uint64_t *genome;
    int32_t  gw;         /* u64 words per genome (both strands) */
    /* genetics */
    double   bino[QHG_MAX_BINO];
    int32_t  nbino;
    /* counters */
    int64_t  max_id;
    uint32_t step;
    int64_t  births_last, deaths_last, moves_last;
    int64_t  births_total, deaths_total, moves_total;
    int64_t  agent_steps;
    /* pairing scratch */
    int32_t *pc_off, *pc_buf;
    /* navigation: Navigation::m_mDestinations (ports and each port's
     * destinations in increasing cell order) and m_vBridges */
    int32_t  nav_set, nav_nports, nav_nbridges;
    int32_t *nav_port, *nav_doff, *nav_dcell, *nav_bridge;
    double  *nav_ddist;
    /* Navigate's state: m_mJumpProbs (a table per port, reached through the
     * port's cell) and m_vCurBridges */
    int32_t  nav_need_update, nav_preloop_done;
    int32_t *jp_index;            /* ncells: the port's table, -1 for none */
    int32_t *jp_n, *jp_off;       /* per table: destinations, first entry  */
    int32_t *jp_targ;             /* entries: n + 1 per table, stay first  */
    double  *jp_prob;
    int32_t  njp, njp_entries;
    int32_t *cur_bridge;          /* 2 per current bridge */
    int32_t  ncur_bridges;
};

static int g_threads = 0;
void orc_set_threads(int32_t n) { g_threads = n; if (n > 0) omp_set_num_threads(n); }
int32_t orc_get_threads(void) { return omp_get_max_threads(); }

static void *xdup(const void *src, size_t bytes) {
    if (!src) return NULL;
    void *p = malloc(bytes);
    if (p) memcpy(p, src, bytes);
    return p;
}

static int has_action(const orc *o, int kind) {
    for (int i = 0; i < o->nact; i++) if (o->act[i] == kind) return 1;
    return 0;
}

static void sort_actions(orc *o) {
    /* Prioritizer order: prio ascending, then action name (std::map order),
     * QHG4/core/Prioritizer.cpp:19-30, QHG4/core/SPopulation.cpp:249-257 */
    int32_t idx[QHG_MAX_ACTIONS];
    int32_t n = o->P.num_actions;
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        int32_t v = idx[i]; int j = i - 1;
        while (j >= 0) {
            int32_t a = idx[j];
            int less = (o->P.action_prio[v] < o->P.action_prio[a]) ||
                       (o->P.action_prio[v] == o->P.action_prio[a] &&
                        strcmp(o->P.action_name[v], o->P.action_name[a]) < 0);
            if (!less) break;
            idx[j + 1] = idx[j]; j--;
        }
        idx[j + 1] = v;
    }
    o->nact = n;
    for (int i = 0; i < n; i++) o->act[i] = o->P.action_kind[idx[i]];
}

int orc_create(const qhg_grid_desc *g, const qhg_env_desc *env, const qhg_params *p, orc **out) {
    if (!g || !p || !out || g->num_cells <= 0 || g->connectivity <= 0) return QHG_ERR_ARG;
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->P = *p;
    o->ncells = g->num_cells;
    o->conn = g->connectivity;
    size_t nc = (size_t)o->ncells;
    o->nbr = (int32_t *)xdup(g->neighbors, nc * o->conn * sizeof(int32_t));
    o->nn = (uint8_t *)xdup(g->num_neighbors, nc);
    o->cell_ids = (int64_t *)malloc(nc * sizeof(int64_t));
    for (size_t c = 0; c < nc; c++) o->cell_ids[c] = g->cell_ids ? g->cell_ids[c] : (int64_t)c;
    if (env) {
        o->alt = (double *)xdup(env->altitude, nc * 8);
        o->lon = (double *)xdup(env->longitude, nc * 8);
        o->lat = (double *)xdup(env->latitude, nc * 8);
        o->water = (double *)xdup(env->water, nc * 8);
        o->npp = (double *)xdup(env->base_npp, nc * 8);
        o->temp = (double *)xdup(env->temperature, nc * 8);
        o->rain = (double *)xdup(env->rainfall, nc * 8);
        o->ice = (uint8_t *)xdup(env->ice, nc);
        o->coastal = (uint8_t *)xdup(env->coastal, nc);
    }
    o->K = (double *)calloc(nc, 8);
    o->W = (double *)calloc(nc * (o->conn + 1), 8);
    o->single = (double *)calloc(nc * (o->conn + 1), 8);
    o->bc = (double *)calloc(nc, 8);
    o->dc = (double *)calloc(nc, 8);
    o->count = (uint32_t *)calloc(nc, 4);
    o->pc_off = (int32_t *)calloc(nc + 1, 4);
    o->me_first = 1;
    o->nav_need_update = 1;
    o->jp_index = (int32_t *)malloc(nc * sizeof(int32_t));
    for (size_t c = 0; c < nc; c++) o->jp_index[c] = -1;
    sort_actions(o);
    o->gw = 0;
    if (has_action(o, QHG_ACT_GENETICS)) {
        o->gw = 2 * orc_genome_words(p->gen_genome_size);
        if (p->gen_mutation_rate > 0) {
            o->nbino = orc_bino_table(p->gen_mutation_rate, 2 * p->gen_genome_size, 1e-6,
                                      o->bino, QHG_MAX_BINO);
            if (o->nbino < 0) { orc_destroy(o); return QHG_ERR_ARG; }
        }
    }
    *out = o;
    return 0;
}

void orc_destroy(orc *o) {
    if (!o) return;
    void *ptrs[] = { o->nbr, o->nn, o->cell_ids, o->alt, o->lon, o->lat, o->water, o->npp,
                     o->temp, o->rain, o->ice, o->coastal, o->K, o->W, o->single, o->bc,
                     o->dc, o->count, o->pc_off, o->pc_buf, o->id, o->cell, o->mate,
                     o->nbabies, o->dest, o->life, o->gender, o->bflag, o->birth, o->age,
                     o->lastb, o->genome, o->nav_port, o->nav_doff, o->nav_dcell, o->nav_bridge,
                     o->nav_ddist, o->jp_index, o->jp_n, o->jp_off, o->jp_targ, o->jp_prob, o->cur_bridge };
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(o);
}

static void reserve(orc *o, int64_t need) {
    if (need <= o->cap) return;
    int64_t nc = o->cap ? o->cap : 1024;
    while (nc < need) nc *= 2;
#define GROW(f, T) o->f = (T *)realloc(o->f, (size_t)nc * sizeof(T))
    GROW(id, int64_t); GROW(cell, int32_t); GROW(mate, int32_t); GROW(nbabies, int32_t);
    GROW(dest, int32_t); GROW(life, uint8_t); GROW(gender, uint8_t); GROW(bflag, uint8_t);
    GROW(birth, double); GROW(age, double); GROW(lastb, double); GROW(pc_buf, int32_t);
#undef GROW
    if (o->gw) o->genome = (uint64_t *)realloc(o->genome, (size_t)nc * o->gw * 8);
    o->cap = nc;
}

static void update_counts(orc *o) {
    memset(o->count, 0, (size_t)o->ncells * 4);
    for (int64_t i = 0; i < o->n; i++)
        if (o->life[i] > 0) o->count[o->cell[i]]++;
}

int orc_set_agents(orc *o, const qhg_agents *a) {
    if (!a || a->n < 0 || !a->id || !a->cell || !a->life || !a->gender) return QHG_ERR_ARG;
    reserve(o, a->n + 1);
    /* id order */
    int64_t *perm = (int64_t *)malloc((size_t)(a->n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < a->n; i++) perm[i] = i;
    /* simple merge sort by id */
    int64_t *tmp = (int64_t *)malloc((size_t)(a->n + 1) * sizeof(int64_t));
    for (int64_t w = 1; w < a->n; w *= 2) {
        for (int64_t lo = 0; lo < a->n; lo += 2 * w) {
            int64_t mid = lo + w < a->n ? lo + w : a->n, hi = lo + 2 * w < a->n ? lo + 2 * w : a->n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = (a->id[perm[i]] <= a->id[perm[j]]) ? perm[i++] : perm[j++];
            while (i < mid) tmp[k++] = perm[i++];
            while (j < hi) tmp[k++] = perm[j++];
        }
        memcpy(perm, tmp, (size_t)a->n * sizeof(int64_t));
    }
    free(tmp);
    o->n = 0;
    o->max_id = 0;
    for (int64_t q = 0; q < a->n; q++) {
        int64_t i = perm[q];
        if (a->life[i] == 0) continue;
        int64_t k = o->n++;
        o->id[k] = a->id[i];
        o->cell[k] = a->cell[i];
        o->life[k] = a->life[i];
        o->gender[k] = a->gender[i];
        o->birth[k] = a->birth_time ? a->birth_time[i] : 0.0;
        o->age[k] = a->age ? a->age[i] : 0.0;
        o->lastb[k] = a->last_birth ? a->last_birth[i] : 0.0;
        o->mate[k] = -3;
        o->nbabies[k] = a->num_babies ? a->num_babies[i] : 0;
        if (o->gw) {
            if (a->genome) memcpy(o->genome + k * o->gw, a->genome + i * o->gw, (size_t)o->gw * 8);
            else memset(o->genome + k * o->gw, 0, (size_t)o->gw * 8);
        }
        if (a->id[i] > o->max_id) o->max_id = a->id[i];
    }
    free(perm);
    update_counts(o);
    return 0;
}

int64_t orc_num_alive(orc *o) {
    int64_t c = 0;
    for (int64_t i = 0; i < o->n; i++) c += o->life[i] > 0;
    return c;
}

int orc_get_agents(orc *o, qhg_agents *a) {
    int64_t k = 0;
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        if (a->id) a->id[k] = o->id[i];
        if (a->cell) a->cell[k] = o->cell[i];
        if (a->life) a->life[k] = o->life[i];
        if (a->gender) a->gender[k] = o->gender[i];
        if (a->birth_time) a->birth_time[k] = o->birth[i];
        if (a->age) a->age[k] = o->age[i];
        if (a->last_birth) a->last_birth[k] = o->lastb[i];
        if (a->mate_id) {
            int32_t m = o->mate[i];
            a->mate_id[k] = (m >= 0) ? (o->life[m] > 0 ? o->id[m] : -1) : m;
        }
        if (a->num_babies) a->num_babies[k] = o->nbabies[i];
        if (a->genome && o->gw) memcpy(a->genome + k * o->gw, o->genome + i * o->gw, (size_t)o->gw * 8);
        k++;
    }
    a->n = k;
    return 0;
}

/* An exponential built from IEEE +,-,*,/ and ldexp only, so that the
 * restatement and the device agree to the bit where the reference calls
 * libm exp (the Miami NPP below).  Cody-Waite reduction x = k ln2 + r with
 * ln2 = head (32 significant bits, so k * head is exact) + tail, then
 * e^r = 1 + r(1 + r/2(1 + r/3(... (1 + r/13)))), then 2^k e^r.  Below
 * 2^-1022 the result is 0, above 2^1023 it is +inf.  Pinned against libm exp
 * to a few ulp by the tests. */
double orc_exp(double x) {
    if (isnan(x)) return x;
    if (x > 709.0) return INFINITY;
    if (x < -708.0) return 0.0;
    double kr = x * 1.44269504088896338700e+00;
    int k = (int)(kr + (kr >= 0.0 ? 0.5 : -0.5));
    double r = (x - (double)k * 6.93147180369123816490e-01) - (double)k * 1.90821492927058770002e-10;
    double s = 1.0;
    for (int n = 13; n >= 2; n--) s = 1.0 + (r / (double)n) * s;
    s = 1.0 + r * s;
    return ldexp(s, k);
}

/* ------------------------------------------------------------------ */
/* NPPCapacity::recalculate, QHG4/actions/NPPCapacity.cpp:138-217;     */
/* Miami NPP, QHG4/core/NPPCalcMiami.cpp:28-41 (exp via orc_exp)       */
/* ------------------------------------------------------------------ */
static void npp_capacity(orc *o) {
    const qhg_params *P = &o->P;
    for (int32_t i = 0; i < o->ncells; i++) {
        double miami = 0;
        if (o->temp && o->rain) {
            double nT = 3000.0 / (1 + orc_exp(1.315 - 0.119 * o->temp[i]));
            double nP = 3000.0 * (1 - orc_exp(-0.000664 * o->rain[i]));
            miami = 0.000475 * ((nT < nP) ? nT : nP);
        }
        double cap;
        double a = o->alt ? o->alt[i] : 0.0;
        if (a > 0) {
            float A0 = 1500, A1 = 2500;
            double af = (a < A0) ? 1 : ((A1 - a) / (A1 - A0));
            if (af < 0) af = 0;
            double tn = o->npp ? o->npp[i] : 0.0;
            if (o->lon && o->lat && o->lon[i] > 115.0 && o->lat[i] > -12.0 &&
                o->lon[i] < 150.0 && o->lat[i] < 1.0) {
                if ((o->npp ? o->npp[i] : 0.0) < P->npp_min) tn = miami;
            }
            tn *= af;
            if (tn < P->npp_min) cap = P->npp_k_min;
            else if (tn > P->npp_max) cap = P->npp_k_max;
            else cap = P->npp_k_min + tn * P->npp_k_max / (P->npp_max - P->npp_min);
            if (o->coastal && o->lat && o->coastal[i] && o->lat[i] > P->npp_coastal_min_lat &&
                o->lat[i] < P->npp_coastal_max_lat)
                cap += P->npp_coastal_factor * P->npp_k_max;
            if (o->water) cap += o->water[i] * P->npp_water_factor * P->npp_k_max;
            if (cap > P->npp_k_max) cap = P->npp_k_max;
        } else {
            cap = 0;
        }
        o->K[i] = cap * P->npp_efficiency;
    }
}

/* SingleEvaluator::calcValues + exchangeAndCumulate,
 * QHG4/actions/SingleEvaluator.cpp:174-243 */
static void single_eval(orc *o, int e, double *out) {
    const qhg_params *P = &o->P;
    int32_t R = o->conn + 1;
    const double *in = NULL;
    switch (P->me_input[e]) {
    case QHG_EV_ALTITUDE: in = o->alt; break;
    case QHG_EV_CAPACITY: in = o->K; break;
    case QHG_EV_BASE_NPP: in = o->npp; break;
    }
    memset(out, 0, (size_t)o->ncells * R * 8);
    for (int32_t c = 0; c < o->ncells; c++) {
        if (!o->ice || !o->ice[c]) {
            double v;
            if (P->me_pl_n[e] >= 2)
                v = orc_polyline(P->me_pl_n[e], P->me_pl_x[e], P->me_pl_v[e], (double)(float)in[c]);
            else
                v = in[c];
            out[(size_t)c * R] = (v > 0) ? v : 0;
        } else {
            out[(size_t)c * R] = 0;
        }
    }
    for (int32_t c = 0; c < o->ncells; c++) {
        double w = out[(size_t)c * R];
        for (int32_t i = 0; i < o->conn; i++) {
            double cw = 0;
            int32_t nb = o->nbr[(size_t)c * o->conn + i];
            if (nb >= 0) cw = out[(size_t)nb * R];
            cw = (cw > 0) ? cw : 0;
            w = P->me_cumulate[e] ? w + cw : cw;
            out[(size_t)c * R + i + 1] = w;
        }
    }
}

/* MultiEvaluator::initialize, QHG4/actions/MultiEvaluator.cpp:142-478 */
static void multi_eval(orc *o) {
    const qhg_params *P = &o->P;
    int32_t R = o->conn + 1;
    size_t N = (size_t)o->ncells * R;
    double init = 0;
    if (P->me_mode == QHG_ME_MUL_SIMPLE) init = 1.0;
    else if (P->me_mode == QHG_ME_MAX_SIMPLE) init = -INFINITY;
    else if (P->me_mode == QHG_ME_MIN_SIMPLE) init = INFINITY;
    for (size_t i = 0; i < N; i++) o->W[i] = init;
    for (int e = 0; e < P->me_num; e++) {
        single_eval(o, e, o->single);
        double w = P->me_weight[e];
        for (size_t i = 0; i < N; i++) {
            double s = o->single[i];
            switch (P->me_mode) {
            case QHG_ME_ADD_SIMPLE: o->W[i] += s * w; break;
            case QHG_ME_MUL_SIMPLE: o->W[i] *= s * w; break;
            case QHG_ME_MAX_SIMPLE: { double nv = s * w; if (o->W[i] < nv) o->W[i] = nv; } break;
            case QHG_ME_MIN_SIMPLE: { double nv = s * w; if (o->W[i] > nv) o->W[i] = nv; } break;
            }
        }
    }
    if (P->me_mode != QHG_ME_MAX_SIMPLE) {
        for (size_t r = 0; r < N; r += R)
            for (size_t c = r + 1; c < r + R; c++) o->W[c] += o->W[c - 1];
    }
}

/* PopBase::modifyAttributes restated for the tests: the parameter block is
 * replaced between steps (same actions; per-step attributes only) */
int orc_set_params(orc *o, const qhg_params *p) {
    if (p->num_actions != o->P.num_actions) return -1;
    o->P = *p;
    sort_actions(o);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Navigate (QHG4/actions/Navigate.cpp) over the sea-ways and bridges  */
/* of Navigation (QHG4/core/Navigation.h:12-40)                        */
/* ------------------------------------------------------------------ */

/* Navigate::cleanup (Navigate.cpp:63-72): drop the tables and bridges */
static void nav_cleanup(orc *o) {
    for (int32_t c = 0; c < o->ncells; c++) o->jp_index[c] = -1;
    o->njp = 0;
    o->njp_entries = 0;
    o->ncur_bridges = 0;
}

/* Navigate::recalculate (Navigate.cpp:95-148), run when an update is due:
 * A = prob0 / exp(decay * dist0); per port (increasing cell index, the
 * distancemap's order) the probabilities A * exp(decay * dist) of its
 * destinations (increasing cell index, the distlist's order), summed in that
 * order; with a sum below 1 the table is [stay = 1 - sum, destinations...]
 * made cumulative, otherwise the build stops at that port.  The current
 * bridges are those with both ends above sea level. */
static void nav_recalculate(orc *o) {
    if (!o->nav_need_update) return;
    const qhg_params *P = &o->P;
    const double A = P->nav_prob0 / exp(P->nav_decay * P->nav_dist0);
    nav_cleanup(o);
    int32_t total = o->nav_nports + (o->nav_nports ? o->nav_doff[o->nav_nports] : 0);
    o->jp_n = (int32_t *)realloc(o->jp_n, (size_t)(o->nav_nports + 1) * sizeof(int32_t));
    o->jp_off = (int32_t *)realloc(o->jp_off, (size_t)(o->nav_nports + 1) * sizeof(int32_t));
    o->jp_targ = (int32_t *)realloc(o->jp_targ, (size_t)(total + 1) * sizeof(int32_t));
    o->jp_prob = (double *)realloc(o->jp_prob, (size_t)(total + 1) * sizeof(double));
    for (int32_t k = 0; k < o->nav_nports; k++) {
        const int32_t d0 = o->nav_doff[k], n = o->nav_doff[k + 1] - d0;
        int32_t *targ = o->jp_targ + o->njp_entries;
        double *prob = o->jp_prob + o->njp_entries;
        double sum = 0;
        for (int32_t j = 0; j < n; j++) {
            double pr = A * exp(P->nav_decay * o->nav_ddist[d0 + j]);
            sum += pr;
            targ[j + 1] = o->nav_dcell[d0 + j];
            prob[j + 1] = pr;
        }
        if (sum < 1) {
            targ[0] = -1;
            prob[0] = 1 - sum;
            for (int32_t j = 1; j < n + 1; j++) prob[j] += prob[j - 1];
            o->jp_n[o->njp] = n;
            o->jp_off[o->njp] = o->njp_entries;
            o->jp_index[o->nav_port[k]] = o->njp;
            o->njp++;
            o->njp_entries += n + 1;
        } else {
            fprintf(stderr, "[Navigate] probabilities for port [%d] add up to %f\n", o->nav_port[k], sum);
            break;
        }
    }
    o->cur_bridge = (int32_t *)realloc(o->cur_bridge, (size_t)(2 * o->nav_nbridges + 2) * sizeof(int32_t));
    for (int32_t b = 0; b < o->nav_nbridges; b++) {
        const int32_t a0 = o->nav_bridge[2 * b], a1 = o->nav_bridge[2 * b + 1];
        if (o->alt && o->alt[a0] > 0 && o->alt[a1] > 0) {
            o->cur_bridge[2 * o->ncur_bridges] = a0;
            o->cur_bridge[2 * o->ncur_bridges + 1] = a1;
            o->ncur_bridges++;
        }
    }
    o->nav_need_update = 0;
}

/* Navigation::setData / setBridges (QHG4/core/Navigation.cpp:24-60) as
 * NavGroupReader fills them: ports and destinations sorted as the maps
 * hold them; a repeated port or destination fails (NavGroupReader's size
 * check).  Before the population's pre-loop this is the grid's data;
 * afterwards it is EVENT_ID_NAV with its flush (Navigate::notify,
 * Navigate.cpp:79-89): the tables are rebuilt at once. */
int orc_set_navigation(orc *o, const qhg_nav_desc *nav) {
    if (!nav || nav->num_ports < 0 || nav->num_bridges < 0 || nav->num_bridges > 131072) return QHG_ERR_ARG;
    const int32_t np = nav->num_ports;
    if (np && (!nav->ports || !nav->dest_offsets)) return QHG_ERR_ARG;
    const int32_t nd = np ? nav->dest_offsets[np] : 0;
    if (nd && (!nav->dest_cells || !nav->dest_dists)) return QHG_ERR_ARG;
    if (nav->num_bridges && !nav->bridges) return QHG_ERR_ARG;
    int64_t *key = (int64_t *)malloc((size_t)(np + 1) * sizeof(int64_t));
    for (int32_t k = 0; k < np; k++) {
        if (nav->ports[k] < 0 || nav->ports[k] >= o->ncells || nav->dest_offsets[k] < 0 ||
            nav->dest_offsets[k + 1] < nav->dest_offsets[k]) { free(key); return QHG_ERR_ARG; }
        key[k] = ((int64_t)nav->ports[k] << 32) | k;
    }
    /* ports in cell order (insertion sort of the keys) */
    for (int32_t i = 1; i < np; i++) {
        int64_t v = key[i]; int32_t j = i - 1;
        while (j >= 0 && key[j] > v) { key[j + 1] = key[j]; j--; }
        key[j + 1] = v;
    }
    for (int32_t k = 1; k < np; k++)
        if ((key[k] >> 32) == (key[k - 1] >> 32)) { free(key); return QHG_ERR_ARG; }
    int32_t *port = (int32_t *)malloc((size_t)(np + 1) * sizeof(int32_t));
    int32_t *doff = (int32_t *)malloc((size_t)(np + 1) * sizeof(int32_t));
    int32_t *dcell = (int32_t *)malloc((size_t)(nd + 1) * sizeof(int32_t));
    double *ddist = (double *)malloc((size_t)(nd + 1) * sizeof(double));
    int32_t w = 0, bad = 0;
    for (int32_t i = 0; i < np && !bad; i++) {
        const int32_t k = (int32_t)(key[i] & 0xffffffff);
        port[i] = nav->ports[k];
        doff[i] = w;
        const int32_t a = nav->dest_offsets[k], n = nav->dest_offsets[k + 1] - a;
        /* destinations in cell order: selection of the next larger cell */
        int32_t last = -1;
        for (int32_t j = 0; j < n && !bad; j++) {
            int32_t best = -1;
            for (int32_t q = 0; q < n; q++) {
                const int32_t c = nav->dest_cells[a + q];
                if (c > last && (best < 0 || c < nav->dest_cells[a + best])) best = q;
            }
            if (best < 0) { bad = 1; break; }    /* a destination listed twice */
            const int32_t c = nav->dest_cells[a + best];
            if (c < 0 || c >= o->ncells) { bad = 1; break; }
            dcell[w] = c;
            ddist[w] = nav->dest_dists[a + best];
            w++;
            last = c;
        }
    }
    doff[np] = w;
    for (int32_t b = 0; b < 2 * nav->num_bridges && !bad; b++)
        if (nav->bridges[b] < 0 || nav->bridges[b] >= o->ncells) bad = 1;
    free(key);
    if (bad) { free(port); free(doff); free(dcell); free(ddist); return QHG_ERR_ARG; }
    free(o->nav_port); free(o->nav_doff); free(o->nav_dcell); free(o->nav_ddist); free(o->nav_bridge);
    o->nav_port = port;
    o->nav_doff = doff;
    o->nav_dcell = dcell;
    o->nav_ddist = ddist;
    o->nav_nports = np;
    o->nav_nbridges = nav->num_bridges;
    o->nav_bridge = (int32_t *)malloc((size_t)(2 * nav->num_bridges + 2) * sizeof(int32_t));
    if (nav->num_bridges) memcpy(o->nav_bridge, nav->bridges, (size_t)nav->num_bridges * 2 * sizeof(int32_t));
    o->nav_set = 1;
    o->nav_need_update = 1;
    if (o->nav_preloop_done) nav_recalculate(o);
    return 0;
}

/* Navigate::execute (Navigate.cpp:187-244) for an agent at cell c that is
 * alive and has no move registered: from a port, draw dR and walk the
 * cumulative table while dR exceeds the entry, at most up to index c (the
 * reference bounds the walk by the map key, the port's cell index, where
 * the destination count was meant) and at most to the last destination;
 * entry i > 0 registers a move to its cell unless that cell is ice.  Then
 * for every current bridge with an end at c, in list order: one draw, and a
 * move to the other end when it is below Navigate_bridge_prob.  draw(ctx, b)
 * gives the jump draw (b = -1) or the draw of current bridge b.  Writes the
 * first max registered destinations to dst and the last one to *last;
 * returns the number of registrations. */
static int32_t nav_execute(const orc *o, int32_t c, double (*draw)(void *, int32_t), void *ctx, int32_t *last,
                           int32_t *dst, int32_t max) {
    int32_t nreg = 0;
    const int32_t t = o->jp_index[c];
    if (t >= 0) {
        const int32_t n = o->jp_n[t];
        const double *cum = o->jp_prob + o->jp_off[t];
        const double dR = draw(ctx, -1);
        int32_t i = 0;
        while (i < c && i < n && dR > cum[i]) i++;
        if (i > 0) {
            const int32_t nc = o->jp_targ[o->jp_off[t] + i];
            if (!o->ice || !o->ice[nc]) {
                if (nreg < max) dst[nreg] = nc;
                nreg++;
                *last = nc;
            }
        }
    }
    for (int32_t b = 0; b < o->ncur_bridges; b++) {
        const int32_t a0 = o->cur_bridge[2 * b], a1 = o->cur_bridge[2 * b + 1];
        int32_t nc = -1;
        if (a0 == c) nc = a1;
        else if (a1 == c) nc = a0;
        if (nc >= 0) {
            const double dR = draw(ctx, b);
            if (dR < o->P.nav_bridge_prob) {
                if (nreg < max) dst[nreg] = nc;
                nreg++;
                *last = nc;
            }
        }
    }
    return nreg;
}

typedef struct { uint64_t seed; uint32_t step; uint64_t id; } nav_philox;

static double nav_draw_philox(void *ctx, int32_t b) {
    const nav_philox *q = (const nav_philox *)ctx;
    return orc_wrandd(orc_draw(q->seed, q->step, QHG_STREAM_NAV, q->id, (uint32_t)(b + 1)));
}

typedef struct { const double *seq; int32_t n, used; } nav_seq;

static double nav_draw_seq(void *ctx, int32_t b) {
    nav_seq *q = (nav_seq *)ctx;
    (void)b;
    return (q->used < q->n) ? q->seq[q->used++] : 2.0;
}

static void act_navigate(orc *o) {
    int64_t moves = 0;
#pragma omp parallel for schedule(static) reduction(+:moves)
    for (int64_t i = 0; i < o->n; i++) {
        /* alive and not already moving (LIFE_STATE_MOVING: a move registered) */
        if (o->life[i] == 0 || o->dest[i] >= 0) continue;
        nav_philox q = { o->P.seed, o->step, (uint64_t)o->id[i] };
        int32_t last = -1;
        const int32_t k = nav_execute(o, o->cell[i], nav_draw_philox, &q, &last, NULL, 0);
        if (k) { o->dest[i] = last; moves += k; }
    }
    o->moves_last += moves;
}

/* the pieces above, for the tests */
int32_t orc_nav_table(orc *o, int32_t cell, int32_t *targ, double *cum, int32_t max) {
    if (cell < 0 || cell >= o->ncells || o->jp_index[cell] < 0) return -1;
    const int32_t t = o->jp_index[cell], n = o->jp_n[t] + 1;
    for (int32_t j = 0; j < n && j < max; j++) {
        targ[j] = o->jp_targ[o->jp_off[t] + j];
        cum[j] = o->jp_prob[o->jp_off[t] + j];
    }
    return n;
}

int32_t orc_nav_bridges(orc *o, int32_t *pairs, int32_t max) {
    for (int32_t b = 0; b < o->ncur_bridges && b < max; b++) {
        pairs[2 * b] = o->cur_bridge[2 * b];
        pairs[2 * b + 1] = o->cur_bridge[2 * b + 1];
    }
    return o->ncur_bridges;
}

int32_t orc_nav_execute(orc *o, int32_t cell, const double *draws, int32_t ndraws, int32_t *used, int32_t *dst,
                        int32_t max) {
    nav_seq q = { draws, ndraws, 0 };
    int32_t last = -1;
    const int32_t k = nav_execute(o, cell, nav_draw_seq, &q, &last, dst, max);
    *used = q.used;
    return k;
}

/* Environment events between steps: replace the given fields; with
 * QHG_EVENT_KILL_WET remove every agent in a cell with altitude < 0 or ice
 * (OoAPop::updateEvent, QHG4/populations/OoAPop.cpp:164-196: registerDeath
 * then performDeaths before the next step; not counted as step deaths);
 * then recompute capacities and evaluator weights (the flush,
 * QHG4/actions/NPPCapacity.cpp:121-131, MultiEvaluator.cpp:142-210). */
static void replace_field(double **dst, const double *src, size_t n) {
    if (!src) return;
    if (!*dst) *dst = (double *)malloc(n * 8);
    memcpy(*dst, src, n * 8);
}

static void replace_flags(uint8_t **dst, const uint8_t *src, size_t n) {
    if (!src) return;
    if (!*dst) *dst = (uint8_t *)malloc(n);
    memcpy(*dst, src, n);
}

int orc_env_event(orc *o, const qhg_env_desc *env, int32_t events) {
    if (!env) return QHG_ERR_ARG;
    size_t nc = (size_t)o->ncells;
    replace_field(&o->alt, env->altitude, nc);
    replace_field(&o->lon, env->longitude, nc);
    replace_field(&o->lat, env->latitude, nc);
    replace_field(&o->water, env->water, nc);
    replace_field(&o->npp, env->base_npp, nc);
    replace_field(&o->temp, env->temperature, nc);
    replace_field(&o->rain, env->rainfall, nc);
    replace_flags(&o->ice, env->ice, nc);
    replace_flags(&o->coastal, env->coastal, nc);
    if (events & QHG_EVENT_KILL_WET) {
        int64_t w = 0;
        for (int64_t i = 0; i < o->n; i++) {
            int32_t c = o->cell[i];
            int wet = (o->alt && o->alt[c] < 0) || (o->ice && o->ice[c] > 0);
            if (wet || o->life[i] == 0) continue;
            if (w != i) {
                o->id[w] = o->id[i]; o->cell[w] = o->cell[i]; o->mate[w] = o->mate[i];
                o->nbabies[w] = o->nbabies[i]; o->life[w] = o->life[i]; o->gender[w] = o->gender[i];
                o->birth[w] = o->birth[i]; o->age[w] = o->age[i]; o->lastb[w] = o->lastb[i];
                if (o->gw) memmove(o->genome + w * o->gw, o->genome + i * o->gw, (size_t)o->gw * 8);
            }
            w++;
        }
        o->n = w;
        /* mate indices of the last pairing no longer hold: read back as
         * unpaired until the next step pairs again */
        for (int64_t i = 0; i < o->n; i++) o->mate[i] = -3;
        update_counts(o);
    }
    if (has_action(o, QHG_ACT_NPPCAP)) npp_capacity(o);
    if (has_action(o, QHG_ACT_MULTIEVAL)) { multi_eval(o); o->me_first = 0; }
    /* Navigate::notify: a geography event makes an update due, the flush
     * runs it (once the pre-loop has built the first tables) */
    if ((events & QHG_EVENT_GEO) || env->altitude) {
        o->nav_need_update = 1;
        if (o->nav_preloop_done) nav_recalculate(o);
    }
    return 0;
}

int orc_pre_loop(orc *o) {
    if (has_action(o, QHG_ACT_NPPCAP)) npp_capacity(o);
    if (has_action(o, QHG_ACT_NAVIGATE)) {
        /* Navigate::preLoop (Navigate.cpp:155-183): needs the geography and
         * the navigation; the tables are built here */
        if (!o->nav_set || !o->alt) return QHG_ERR_ARG;
        o->nav_need_update = 1;
        nav_recalculate(o);
    }
    o->nav_preloop_done = 1;
    update_counts(o);
    return 0;
}

/* ------------------------------------------------------------------ */
/* initialize phase                                                    */
/* ------------------------------------------------------------------ */

/* The pairing rule of RandomPair::findMates (QHG4/actions/RandomPair.cpp
 * :212-235 and, with the sexes swapped, :249-272): the side with fewer
 * candidates chooses in list order; chooser j takes
 * k = (int)wrandr(0, left, u[j]), left = candidates still unpaired, and is
 * matched with the k-th unpaired candidate (0-based) in list order.
 * pick[j] receives the chosen candidate's list index; taken (ncand bytes)
 * is scratch.  Requires nchoose <= ncand. */
void orc_pair_cell(int32_t nchoose, int32_t ncand, const uint32_t *u, uint8_t *taken, int32_t *pick) {
    memset(taken, 0, (size_t)ncand);
    int32_t left = ncand;
    for (int32_t j = 0; j < nchoose && left > 0; j++) {
        int32_t choose = (int32_t)orc_wrandr(0, left, u[j]);
        int32_t found = -1, m = -1;
        while (found < choose) { m++; if (!taken[m]) found++; }
        taken[m] = 1;
        pick[j] = m;
        left--;
    }
}

/* RandomPair::initialize + findMates, QHG4/actions/RandomPair.cpp:107-279 */
static void random_pair(orc *o) {
    int64_t n = o->n;
    int32_t nc = o->ncells;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) o->mate[i] = -3;
    /* bucket FERTILE agents by cell in id order: counting sort */
    int32_t *off = o->pc_off;
    memset(off, 0, (size_t)(nc + 1) * 4);
    for (int64_t i = 0; i < n; i++)
        if (o->life[i] == QHG_LIFE_FERTILE && o->gender[i] <= 1) off[o->cell[i] + 1]++;
    for (int32_t c = 0; c < nc; c++) off[c + 1] += off[c];
    int32_t *fill = (int32_t *)malloc((size_t)nc * 4);
    memcpy(fill, off, (size_t)nc * 4);
    for (int64_t i = 0; i < n; i++)
        if (o->life[i] == QHG_LIFE_FERTILE && o->gender[i] <= 1) o->pc_buf[fill[o->cell[i]]++] = (int32_t)i;
    free(fill);
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
#pragma omp parallel
    {
        int32_t tcap = 0;
        int32_t *fem = NULL, *mal = NULL, *pick = NULL;
        uint32_t *u = NULL;
        uint8_t *taken = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int32_t c = 0; c < nc; c++) {
            int32_t s = off[c], e = off[c + 1];
            if (e - s < 2) continue;
            if (e - s > tcap) {
                tcap = e - s;
                fem = (int32_t *)realloc(fem, (size_t)tcap * 4);
                mal = (int32_t *)realloc(mal, (size_t)tcap * 4);
                pick = (int32_t *)realloc(pick, (size_t)tcap * 4);
                u = (uint32_t *)realloc(u, (size_t)tcap * 4);
                taken = (uint8_t *)realloc(taken, (size_t)tcap);
            }
            int32_t F = 0, M = 0;
            for (int32_t q = s; q < e; q++) {
                int32_t a = o->pc_buf[q];
                if (o->gender[a] == 0) fem[F++] = a; else mal[M++] = a;
            }
            if (F == 0 || M == 0) continue;
            /* females choose among males when F <= M (:212-235), else males
             * among females (:249-272); each chooser's draw is keyed by its id */
            int32_t *chooser = (F <= M) ? fem : mal, *cand = (F <= M) ? mal : fem;
            int32_t nch = (F <= M) ? F : M, nca = (F <= M) ? M : F;
            for (int32_t j = 0; j < nch; j++)
                u[j] = orc_draw(seed, step, QHG_STREAM_PAIR, (uint64_t)o->id[chooser[j]], 0);
            orc_pair_cell(nch, nca, u, taken, pick);
            for (int32_t j = 0; j < nch; j++) {
                o->mate[chooser[j]] = cand[pick[j]];
                o->mate[cand[pick[j]]] = chooser[j];
            }
        }
        free(fem); free(mal); free(pick); free(u); free(taken);
    }
}

/* GeneUtils::bitcount2 (QHG4/genes/GeneUtils.cpp:63-68): the 2-bit
 * nucleotides of x that are not zero */
static inline int orc_bitcount2(uint64_t x) {
    return __builtin_popcountll((x | (x >> 1)) & 0x5555555555555555ull);
}

/* GeneUtils::calcDist (QHG4/genes/GeneUtils.cpp:414-444).  `size_arg` is
 * read as a genome size in nucleotides: the distance covers
 * numNucs2Blocks(size_arg) = ceil(size_arg / 32) words per strand position. */
int32_t orc_calc_dist(const uint64_t *g1, const uint64_t *g2, int32_t size_arg) {
    const int32_t nb = (int32_t)(((uint32_t)size_arg + 31u) / 32u);
    int32_t c1 = 0, c2 = 0;
    for (int32_t k = 0; k < 2 * nb; k++) c1 += orc_bitcount2(g1[k] ^ g2[k]);
    for (int32_t k = 0; k < nb; k++) c2 += orc_bitcount2(g1[k + nb] ^ g2[k]) + orc_bitcount2(g1[k] ^ g2[k + nb]);
    return c1 < c2 ? c1 : c2;
}

/* RankTable::makeAllPairings, pairHighestRankers, randomPair, pair and
 * clearCross (QHG4/utils/RankTable.cpp:119-278) over the nf x nm table R
 * (row-major, changed in place).  next(ctx) gives the draws in order.
 * perm (nf) and midx (nm) are scratch; the pairs go to pairs[2k] (row) and
 * pairs[2k+1] (column).  Returns the number of pairs. */
typedef uint32_t (*orc_next_fn)(void *ctx);
static int32_t rank_pairs(int32_t nf, int32_t nm, float *R, float cutoff, int32_t permute, orc_next_fn next,
                          void *ctx, int32_t *perm, int32_t *midx, int32_t *pairs) {
    int32_t npairs = 0;
    for (;;) {
        /* current maximum at or above the cutoff; entries below it become 0 */
        float fmax = 0;
        for (int32_t i = 0; i < nf; i++)
            for (int32_t j = 0; j < nm; j++) {
                float *r = &R[(size_t)i * nm + j];
                if (*r < cutoff) *r = 0;
                else if (*r >= fmax) fmax = *r;
            }
        if (!(fmax > 0)) break;
        /* rows that reach the maximum, in row order */
        uint32_t np = 0;
        for (int32_t i = 0; i < nf; i++)
            for (int32_t j = 0; j < nm; j++)
                if (R[(size_t)i * nm + j] >= fmax) { perm[np++] = i; break; }
        if (permute) {
            for (uint32_t i = np - 1; i > 1; i--) {
                uint32_t k = orc_wrandi(0, i, 1, next(ctx));
                int32_t t = perm[k]; perm[k] = perm[i]; perm[i] = t;
            }
        }
        int have = 0;
        for (uint32_t q = 0; q < np; q++) {
            int32_t i = perm[q], cnt = 0;
            for (int32_t j = 0; j < nm; j++)
                if (R[(size_t)i * nm + j] >= fmax) midx[cnt++] = j;
            if (cnt > 0) {
                int32_t j = midx[orc_wrandi(0, (uint32_t)cnt, 1, next(ctx))];
                pairs[2 * npairs] = i;
                pairs[2 * npairs + 1] = j;
                npairs++;
                for (int32_t jj = 0; jj < nm; jj++) R[(size_t)i * nm + jj] = 0;
                for (int32_t ii = 0; ii < nf; ii++) R[(size_t)ii * nm + j] = 0;
                have = 1;
            }
        }
        if (!have) break;
    }
    return npairs;
}

typedef struct { const uint32_t *u; int32_t n, used; } orc_seq;
static uint32_t seq_next(void *ctx) {
    orc_seq *s = (orc_seq *)ctx;
    uint32_t v = (s->used < s->n) ? s->u[s->used] : 0u;
    s->used++;
    return v;
}

int32_t orc_rank_pairs(int32_t nf, int32_t nm, float *ranks, float cutoff, int32_t permute, const uint32_t *u,
                       int32_t nu, int32_t *pairs, int32_t *used) {
    int32_t *perm = (int32_t *)malloc((size_t)(nf + 1) * 4), *midx = (int32_t *)malloc((size_t)(nm + 1) * 4);
    orc_seq s = { u, nu, 0 };
    int32_t n = rank_pairs(nf, nm, ranks, cutoff, permute, seq_next, &s, perm, midx, pairs);
    free(perm); free(midx);
    if (used) *used = s.used;
    return (s.used > nu) ? -1 : n;
}

/* MD5 (RFC 1321 sections 3.1-3.5): message padded with 0x80 and zeros to
 * 56 mod 64 bytes, then the bit length as a little-endian 64-bit word;
 * 64-byte blocks as 16 little-endian words; four rounds of 16 operations
 * with the functions F, G, H, I, the sine table T and the per-round shifts;
 * the digest is A, B, C, D in little-endian byte order.  The table is
 * computed here from its definition (T[i] = floor(2^32 |sin(i + 1)|)). */
static uint32_t md5_T[64];
static int md5_T_ready = 0;

static void md5_init_table(void) {
    if (md5_T_ready) return;
    for (int i = 0; i < 64; i++) md5_T[i] = (uint32_t)(fabs(sin((double)(i + 1))) * 4294967296.0);
    md5_T_ready = 1;
}

static uint32_t md5_rotl(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

static void md5_compress(uint32_t st[4], const uint8_t blk[64]) {
    static const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    uint32_t X[16];
    for (int k = 0; k < 16; k++)
        X[k] = (uint32_t)blk[4 * k] | ((uint32_t)blk[4 * k + 1] << 8) | ((uint32_t)blk[4 * k + 2] << 16) |
               ((uint32_t)blk[4 * k + 3] << 24);
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        int k;
        switch (i / 16) {
        case 0: f = (b & c) | (~b & d); k = i; break;
        case 1: f = (b & d) | (c & ~d); k = (1 + 5 * i) % 16; break;
        case 2: f = b ^ c ^ d; k = (5 + 3 * i) % 16; break;
        default: f = c ^ (b | ~d); k = (7 * i) % 16; break;
        }
        uint32_t na = b + md5_rotl(a + f + X[k] + md5_T[i], S[i / 16][i % 4]);
        a = d; d = c; c = b; b = na;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

void orc_md5(const uint8_t *msg, int64_t len, uint8_t out[16]) {
    md5_init_table();
    uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
    int64_t done = 0;
    for (; len - done >= 64; done += 64) md5_compress(st, msg + done);
    uint8_t tail[128];
    int64_t r = len - done;
    memset(tail, 0, sizeof tail);
    memcpy(tail, msg + done, (size_t)r);
    tail[r] = 0x80;
    int64_t tl = (r + 1 + 8 <= 64) ? 64 : 128;
    uint64_t bits = (uint64_t)len * 8u;
    for (int k = 0; k < 8; k++) tail[tl - 8 + k] = (uint8_t)(bits >> (8 * k));
    md5_compress(st, tail);
    if (tl == 128) md5_compress(st, tail + 64);
    for (int k = 0; k < 4; k++)
        for (int b = 0; b < 4; b++) out[4 * k + b] = (uint8_t)(st[k] >> (8 * b));
}

/* The genetic code of ProteinBuilder (aTranslationTable,
 * QHG4/genes/ProteinBuilder.h:64-78): codon X*16 + Y*4 + Z (A=0, C=1, G=2,
 * T=3) to the amino-acid number (Met = 10, stop = 20). */
static const uint8_t orc_code[64] = {
    8, 11, 8, 11, 16, 16, 16, 16, 14, 15, 14, 15, 7, 7, 10, 7,
    13, 6, 13, 6, 12, 12, 12, 12, 14, 14, 14, 14, 9, 9, 9, 9,
    3, 2, 3, 2, 0, 0, 0, 0, 5, 5, 5, 5, 17, 17, 17, 17,
    20, 19, 20, 19, 15, 15, 15, 15, 20, 1, 18, 1, 9, 4, 9, 4,
};
enum { AA_MET = 10, AA_STP = 20, AA_END_TOTAL = 33, AA_END_ALLELE = 34 };

/* ProteinBuilder's reader (getNextNuc / getNextAA,
 * QHG4/genes/ProteinBuilder.cpp:231-289): nucleotides strand by strand,
 * 2 bits each from the low end of 32-nucleotide words; a strand's end ends
 * the codon being read (its nucleotides are dropped) */
typedef struct {
    const uint64_t *g;
    int32_t G, nb, pos, strand;
    uint64_t cur;
} orc_reader;

static int reader_nuc(orc_reader *r) {   /* 0..3, or -1 strand end, -2 genome end */
    r->pos++;
    if (r->pos >= r->G) {
        if (r->strand < 1) { r->strand++; r->pos = -1; return -1; }
        return -2;
    }
    if (r->pos % 32 == 0) r->cur = r->g[(size_t)r->strand * r->nb + r->pos / 32];
    int v = (int)(r->cur & 3u);
    r->cur >>= 2;
    return v;
}

static int reader_aa(orc_reader *r) {
    int idx = 0;
    for (int i = 0; i < 3; i++) {
        int v = reader_nuc(r);
        if (v == -2) return AA_END_TOTAL;
        if (v == -1) return AA_END_ALLELE;
        idx = idx * 4 + v;
    }
    return orc_code[idx];
}

/* ProteinBuilder::translateGenomeHash (ProteinBuilder.cpp:163-220): the MD5
 * digests of the agent's proteins, in order (the reference sorts its
 * digest pointers by address, an order no result depends on).  Writes at
 * most max digests; returns the number of proteins. */
int32_t orc_proteome(const uint64_t *genome, int32_t gwords, int32_t genome_size, uint8_t *digests, int32_t max) {
    if (genome_size <= 3) return 0;   /* ProteinBuilder::init refuses it */
    orc_reader r = { genome, genome_size, gwords / 2, -1, 0, 0 };
    uint8_t *buf = (uint8_t *)malloc((size_t)genome_size);
    int32_t np = 0, n = 0, building = 0;
    int aa = -1;
    while (aa != AA_END_TOTAL) {
        while (!building && aa != AA_END_TOTAL) {
            aa = reader_aa(&r);
            if (aa == AA_MET) { building = 1; n = 0; }
        }
        while (building && aa != AA_END_TOTAL) {
            aa = reader_aa(&r);
            if (aa == AA_STP || aa == AA_END_TOTAL || aa == AA_END_ALLELE) {
                if (np < max) orc_md5(buf, n, digests + 16 * (size_t)np);
                np++;
                building = 0;
            } else {
                buf[n++] = (uint8_t)aa;
            }
        }
    }
    free(buf);
    return np;
}

/* ProteomeComparator::countProtHashMatches (ProteomeComparator.cpp:47-56):
 * every pair of equal digests counts */
int32_t orc_prot_matches(const uint8_t *a, int32_t na, const uint8_t *b, int32_t nb) {
    int32_t n = 0;
    for (int32_t i = 0; i < na; i++)
        for (int32_t j = 0; j < nb; j++) n += memcmp(a + 16 * (size_t)i, b + 16 * (size_t)j, 16) == 0;
    return n;
}

/* a cell's draws of AssortativePairHamming, in order */
typedef struct { uint64_t seed; uint32_t step; int32_t cell; uint32_t j; } orc_cell_draws;
static uint32_t cell_next(void *ctx) {
    orc_cell_draws *d = (orc_cell_draws *)ctx;
    uint32_t j = d->j++;
    return orc_draw(d->seed, d->step, QHG_STREAM_ASSORT, (uint64_t)(uint32_t)d->cell | ((uint64_t)(j >> 16) << 32),
                    j & 0xffffu);
}

/* AssortativePairHam::initialize + findMates
 * (QHG4/actions/AssortativePairHam.cpp:93-238): per cell, the FERTILE
 * females and males (id order), the match table
 * 1 - calcDist(female, male, numBlocks) / genomeSize as a float, and the
 * RankTable pairing with the cutoff and permute attributes.  The reference
 * passes the block count where calcDist reads a genome size; that call is
 * kept.  AssortativePair (QHG4/actions/AssortativePair.cpp:94-271) differs
 * only in the table: the number of equal protein digests of the two
 * agents (ProteinBuilder + ProteomeComparator). */
static void assort_pair(orc *o) {
    int64_t n = o->n;
    int32_t nc = o->ncells;
    for (int64_t i = 0; i < n; i++) o->mate[i] = -3;
    int32_t *off = o->pc_off;
    memset(off, 0, (size_t)(nc + 1) * 4);
    for (int64_t i = 0; i < n; i++)
        if (o->life[i] == QHG_LIFE_FERTILE && o->gender[i] <= 1) off[o->cell[i] + 1]++;
    for (int32_t c = 0; c < nc; c++) off[c + 1] += off[c];
    int32_t *fill = (int32_t *)malloc((size_t)nc * 4);
    memcpy(fill, off, (size_t)nc * 4);
    for (int64_t i = 0; i < n; i++)
        if (o->life[i] == QHG_LIFE_FERTILE && o->gender[i] <= 1) o->pc_buf[fill[o->cell[i]]++] = (int32_t)i;
    free(fill);
    const int32_t gsize = o->P.gen_genome_size;
    const int32_t nblocks = (int32_t)(((uint32_t)gsize + 31u) / 32u);
    const float cutoff = (float)o->P.assort_cutoff;
    const int32_t permute = o->P.assort_permute != 0;
    const int prot = has_action(o, QHG_ACT_ASSORTPROT);
    /* the most proteins a genome can hold: one per two codons per strand */
    const int32_t pmax = 2 * ((gsize / 3 + 1) / 2 + 1);
    md5_init_table();   /* before the threads share it */
#pragma omp parallel
    {
        int32_t tcap = 0;
        int32_t *fem = NULL, *mal = NULL, *perm = NULL, *midx = NULL, *pairs = NULL;
        float *R = NULL;
        size_t rcap = 0;
#pragma omp for schedule(dynamic, 16)
        for (int32_t c = 0; c < nc; c++) {
            int32_t s = off[c], e = off[c + 1];
            if (e - s < 2) continue;
            if (e - s > tcap) {
                tcap = e - s;
                fem = (int32_t *)realloc(fem, (size_t)tcap * 4);
                mal = (int32_t *)realloc(mal, (size_t)tcap * 4);
                perm = (int32_t *)realloc(perm, (size_t)tcap * 4);
                midx = (int32_t *)realloc(midx, (size_t)tcap * 4);
                pairs = (int32_t *)realloc(pairs, (size_t)tcap * 8);
            }
            int32_t F = 0, M = 0;
            for (int32_t q = s; q < e; q++) {
                int32_t a = o->pc_buf[q];
                if (o->gender[a] == 0) fem[F++] = a; else mal[M++] = a;
            }
            if (F == 0 || M == 0) continue;
            if ((size_t)F * M > rcap) { rcap = (size_t)F * M; R = (float *)realloc(R, rcap * 4); }
            if (prot) {
                /* every listed agent's digests, then the match counts */
                int32_t *np = (int32_t *)malloc((size_t)(F + M) * 4);
                uint8_t *dg = (uint8_t *)malloc((size_t)(F + M) * (size_t)pmax * 16);
                for (int32_t k = 0; k < F + M; k++) {
                    int32_t a = (k < F) ? fem[k] : mal[k - F];
                    np[k] = orc_proteome(o->genome + (size_t)a * o->gw, o->gw, gsize, dg + (size_t)k * pmax * 16, pmax);
                }
                for (int32_t f = 0; f < F; f++)
                    for (int32_t m = 0; m < M; m++)
                        R[(size_t)f * M + m] = (float)orc_prot_matches(dg + (size_t)f * pmax * 16, np[f],
                                                                       dg + (size_t)(F + m) * pmax * 16, np[F + m]);
                free(np); free(dg);
            } else {
                for (int32_t f = 0; f < F; f++)
                    for (int32_t m = 0; m < M; m++) {
                        int32_t d = 0;
                        if (o->gw) d = orc_calc_dist(o->genome + (size_t)fem[f] * o->gw,
                                                     o->genome + (size_t)mal[m] * o->gw, nblocks);
                        R[(size_t)f * M + m] = (float)(1.0 - (1.0 * d) / gsize);
                    }
            }
            orc_cell_draws dr = { o->P.seed, o->step, c, 0 };
            int32_t np = rank_pairs(F, M, R, cutoff, permute, cell_next, &dr, perm, midx, pairs);
            for (int32_t k = 0; k < np; k++) {
                int32_t a = fem[pairs[2 * k]], b = mal[pairs[2 * k + 1]];
                o->mate[a] = b;
                o->mate[b] = a;
            }
        }
        free(fem); free(mal); free(perm); free(midx); free(pairs); free(R);
    }
}

/* LinearBirth::initialize / LinearDeath::initialize,
 * QHG4/actions/LinearBirth.cpp:90-115, LinearDeath.cpp:92-123 */
static void verhulst_init(orc *o, int varK) {
    const qhg_params *P = &o->P;
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o->ncells; c++) {
        double N = (double)o->count[c];
        if (!varK) {
            o->bc[c] = P->verh_b0 + (P->verh_theta - P->verh_b0) * (N / P->verh_K);
            o->dc[c] = P->verh_d0 + (P->verh_theta - P->verh_d0) * (N / P->verh_K);
        } else {
            double K = o->K[c];
            o->bc[c] = (K <= 0) ? 0 : P->verh_b0 + (P->verh_theta - P->verh_b0) * (N / K);
            o->dc[c] = (K <= 0) ? 1 : P->verh_d0 + (P->verh_theta - P->verh_d0) * (N / K);
        }
    }
}

/* ------------------------------------------------------------------ */
/* execute phase: one loop over all agents per action (doActions)      */
/* ------------------------------------------------------------------ */
static void act_getold(orc *o, double t) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < o->n; i++)
        if (o->life[i] > 0) o->age[i] = t - o->birth[i];
}

static void act_fertility(orc *o, double t) {
    const qhg_params *P = &o->P;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        if (o->gender[i] == 0) {
            o->life[i] = (o->age[i] > P->fert_min_age && o->age[i] < P->fert_max_age &&
                          (t - o->lastb[i]) > P->fert_interbirth) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        } else {
            o->life[i] = (o->age[i] > P->fert_min_age) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        }
    }
}

/* Verhulst::execute = LinearBirth::execute then LinearDeath::execute */
static void act_verhulst(orc *o) {
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
    int64_t deaths = 0;
#pragma omp parallel for schedule(static) reduction(+:deaths)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        int32_t c = o->cell[i];
        double b = o->bc[c];
        if (b > 0) {
            if (o->gender[i] == 0 && o->mate[i] >= 0) {
                double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_LBIRTH));
                if (r < b) o->bflag[i] = 1;
            }
        } else if (b < 0) {
            double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_LBIRTH));
            if (r < -b) { o->life[i] = 0; deaths++; }
        }
        if (o->life[i] > 0) {
            double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_LDEATH));
            if (r < o->dc[c]) { o->life[i] = 0; deaths++; }
        }
    }
    o->deaths_last += deaths;
}

/* A deterministic arctangent built from IEEE +,-,*,/ only, so that a
 * restatement and the device agree to the bit (libm's atan differs between
 * implementations in the last place, which can flip a death decision):
 * odd symmetry, atan(x) = pi/2 - atan(1/x) above 1, then
 * atan(x) = atan(k/8) + atan((x - k/8) / (1 + x k/8)) with k/8 nearest x,
 * and the Taylor series of the small remainder through its 17th power.
 * The atan(k/8) values are the doubles nearest the true values (50-digit
 * evaluation).  Pinned against libm atan to a few ulp by the tests. */
double orc_atan(double x) {
    static const double at8[9] = {0.0, 0.12435499454676144, 0.24497866312686414, 0.35877067027057225,
                                  0.4636476090008061, 0.5585993153435624, 0.6435011087932844,
                                  0.7188299996216245, 0.7853981633974483};
    int neg = x < 0.0;
    double a = neg ? -x : x;
    int inv = a > 1.0;
    if (inv) a = 1.0 / a;
    int k = (int)(a * 8.0 + 0.5);
    double c = (double)k * 0.125;
    double t = (a - c) / (1.0 + a * c);
    double z = t * t;
    /* Horner from the t^17 term down */
    double s = 1.0 / 17.0;
    s = -1.0 / 15.0 + z * s;
    s = 1.0 / 13.0 + z * s;
    s = -1.0 / 11.0 + z * s;
    s = 1.0 / 9.0 + z * s;
    s = -1.0 / 7.0 + z * s;
    s = 1.0 / 5.0 + z * s;
    s = -1.0 / 3.0 + z * s;
    s = 1.0 + z * s;
    double r = at8[k] + t * s;
    if (inv) r = 1.5707963267948966 - r;
    return neg ? -r : r;
}

/* OldAgeDeath::execute (QHG4/actions/OldAgeDeath.cpp, execute): the age is
 * set, then the agent dies when older than max_age + U(1 - u*max_age,
 * 1 + u*max_age) */
static void act_oldagedeath(orc *o, double t) {
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
    double M = o->P.oad_max_age, u = o->P.oad_uncertainty;
    int64_t deaths = 0;
#pragma omp parallel for schedule(static) reduction(+:deaths)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        o->age[i] = t - o->birth[i];
        double r = orc_wrandr(1 - u * M, 1 + u * M,
                              orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_OADEATH));
        if (o->age[i] > M + r) { o->life[i] = 0; deaths++; }
    }
    o->deaths_last += deaths;
}

/* ATanDeath::execute (QHG4/actions/ATanDeath.cpp:66-90) with the scale of
 * ATanDeath::preLoop (:50-51): death with probability
 * 1/2 + scale * atan(slope * (age - max_age)) / pi */
static void act_atandeath(orc *o, double t) {
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
    const double PI = 3.141592653589793;
    double scale = (PI / 2 - 0.001) / atan(o->P.atan_slope * o->P.atan_range);
    int64_t deaths = 0;
#pragma omp parallel for schedule(static) reduction(+:deaths)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        o->age[i] = t - o->birth[i];
        double pd = 0.5 + scale * orc_atan(o->P.atan_slope * (o->age[i] - o->P.atan_max_age)) / PI;
        double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_ATANDEATH));
        if (r < pd) { o->life[i] = 0; deaths++; }
    }
    o->deaths_last += deaths;
}

static void act_randommove(orc *o) {
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
    double p = o->P.rmove_prob;
    int64_t moves = 0;
#pragma omp parallel for schedule(static) reduction(+:moves)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        int32_t c = o->cell[i];
        double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_RMOVE_GO));
        if (r < p) {
            double r2 = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_RMOVE_DIR));
            int32_t slot = (int32_t)(r2 * (o->nn[c] + 1));
            if (slot > 0) {
                int32_t nb = o->nbr[(size_t)c * o->conn + slot - 1];
                if (nb >= 0) { o->dest[i] = nb; moves++; }
            }
        }
    }
    o->moves_last += moves;
}

static void act_weightedmove(orc *o) {
    uint32_t step = o->step;
    uint64_t seed = o->P.seed;
    double p = o->P.wmove_prob;
    int32_t R = o->conn + 1;
    int64_t moves = 0;
#pragma omp parallel for schedule(static) reduction(+:moves)
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        int32_t c = o->cell[i];
        double r = orc_wrandd(orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_WMOVE_GO));
        if (r < p) {
            int32_t nn = o->nn[c];
            const double *w = o->W + (size_t)c * R;
            int32_t idx = -1;
            uint32_t u2 = orc_draw(seed, step, QHG_STREAM_ACT, (uint64_t)o->id[i], QHG_ACTW_WMOVE_DIR);
            if (w[0] == w[nn]) {
                idx = (int32_t)orc_wrandi(0, (uint32_t)nn + 1, 1, u2);
            } else {
                double r2 = orc_wrandd(u2) * w[nn];
                for (int32_t q = 0; q < nn + 1; q++) if (r2 < w[q]) { idx = q; break; }
            }
            if (idx > 0) {
                int32_t nb = o->nbr[(size_t)c * o->conn + idx - 1];
                if (nb >= 0 && (!o->ice || !o->ice[nb])) { o->dest[i] = nb; moves++; }
            }
        }
    }
    o->moves_last += moves;
}

/* ------------------------------------------------------------------ */
/* finalize: births (recycleDeadSpaceNew + createAgentAtIndex +        */
/* makePopSpecificOffspring + Genetics::makeOffspring), moves, counts  */
/* ------------------------------------------------------------------ */
static void make_genome(orc *o, int64_t baby, int64_t mother, int64_t father, uint64_t nid) {
    const qhg_params *P = &o->P;
    int32_t nb = o->gw / 2;
    uint64_t *out = o->genome + baby * o->gw;
    const uint64_t *g1 = o->genome + mother * o->gw;
    const uint64_t *g2 = o->genome + father * o->gw;
    uint64_t seed = P->seed;
    uint32_t step = o->step;
    uint32_t n = 0;
#define GD() orc_draw(seed, step, QHG_STREAM_GENETICS, nid, n++)
    int i1 = (int)(2 * 1.0 * orc_wrandd(GD()));
    int i2 = (int)(2 * 1.0 * orc_wrandd(GD()));
    uint64_t t1[128], t2[128];
    int32_t C = P->gen_num_crossover;
    if (C > 0) {
        uint32_t br[256];
        uint32_t nbits = (uint32_t)nb * 64;
        for (int32_t k = 0; k < C; k++) br[k] = orc_wrandi(0, nbits, 2, GD());
        orc_crossover(t1, g1, P->gen_genome_size, br, C);
        for (int32_t k = 0; k < C; k++) br[k] = orc_wrandi(0, nbits, 2, GD());
        orc_crossover(t2, g2, P->gen_genome_size, br, C);
        memcpy(out, t1 + i1 * nb, (size_t)nb * 8);
        memcpy(out + nb, t2 + i2 * nb, (size_t)nb * 8);
    } else if (C == -1) {
        uint32_t d[256];
        for (int32_t k = 0; k < 2 * nb; k++) d[k] = GD();
        orc_free_reco(t1, g1, nb, d);
        for (int32_t k = 0; k < 2 * nb; k++) d[k] = GD();
        orc_free_reco(t2, g2, nb, d);
        memcpy(out, t1 + i1 * nb, (size_t)nb * 8);
        memcpy(out + nb, t2 + i2 * nb, (size_t)nb * 8);
    } else {
        memcpy(out, g1 + i1 * nb, (size_t)nb * 8);
        memcpy(out + nb, g2 + i2 * nb, (size_t)nb * 8);
    }
    if (P->gen_mutation_rate > 0) {
        int32_t k = orc_bino_getn(o->bino, o->nbino, orc_wrandd(GD()));
        for (int32_t j = 0; j < k; j++) {
            uint32_t pd = GD();
            uint32_t md = GD();
            orc_mutate_nucs(out, 2 * P->gen_genome_size, 1, &pd, &md);
        }
    }
#undef GD
}

static void finalize(orc *o, double t) {
    /* births ranked by (birth cell, mother id): a stable counting sort of the
     * mothers (id order) by the cell they registered the birth in */
    int64_t nb = 0;
    for (int64_t i = 0; i < o->n; i++) nb += o->bflag[i];
    reserve(o, o->n + nb + 1);
    int64_t base = o->n;
    int64_t *morder = (int64_t *)malloc((size_t)(nb + 1) * sizeof(int64_t));
    {
        int64_t *cstart = (int64_t *)calloc((size_t)o->ncells + 1, sizeof(int64_t));
        for (int64_t i = 0; i < base; i++) if (o->bflag[i]) cstart[o->cell[i] + 1]++;
        for (int32_t c = 0; c < o->ncells; c++) cstart[c + 1] += cstart[c];
        for (int64_t i = 0; i < base; i++) if (o->bflag[i]) morder[cstart[o->cell[i]]++] = i;
        free(cstart);
    }
    for (int64_t k = 0; k < nb; k++) {
        const int64_t i = morder[k];
        int64_t s = base + k;
        int64_t nid = o->max_id + 1 + k;
        o->id[s] = nid;
        o->cell[s] = o->cell[i];                 /* registered cell (pre-move) */
        o->birth[s] = t;
        uint32_t u = orc_draw(o->P.seed, o->step, QHG_STREAM_NEWBORN, (uint64_t)nid, 0);
        o->gender[s] = (uint8_t)(2 * orc_wrandd(u));
        o->life[s] = (o->gender[s] == 0) ? QHG_LIFE_FERTILE : QHG_LIFE_ALIVE;
        o->age[s] = 0.0;
        o->lastb[s] = 0.0;
        o->mate[s] = -3;
        o->nbabies[s] = 0;
        o->dest[s] = -1;
        o->bflag[s] = 0;
        o->nbabies[i]++;
        if (o->gw) make_genome(o, s, i, o->mate[i], (uint64_t)nid);
    }
    free(morder);
    o->n = base + nb;
    o->max_id += nb;
    o->births_last = nb;
    /* performMoves */
    for (int64_t i = 0; i < base; i++) {
        if (o->dest[i] >= 0) { o->cell[i] = o->dest[i]; o->dest[i] = -1; }
        o->bflag[i] = 0;
    }
    /* drop the dead (slot recycling is invisible to parity) and remap mates */
    int32_t *newidx = (int32_t *)malloc((size_t)o->n * 4);
    int64_t w = 0;
    for (int64_t i = 0; i < o->n; i++) newidx[i] = (o->life[i] > 0) ? (int32_t)w++ : -1;
    w = 0;
    for (int64_t i = 0; i < o->n; i++) {
        if (o->life[i] == 0) continue;
        int32_t m = o->mate[i];
        o->id[w] = o->id[i]; o->cell[w] = o->cell[i]; o->life[w] = o->life[i];
        o->gender[w] = o->gender[i]; o->birth[w] = o->birth[i]; o->age[w] = o->age[i];
        o->lastb[w] = o->lastb[i]; o->nbabies[w] = o->nbabies[i];
        o->mate[w] = (m >= 0) ? newidx[m] : m;
        o->dest[w] = -1; o->bflag[w] = 0;
        if (o->gw && w != i) memcpy(o->genome + w * o->gw, o->genome + i * o->gw, (size_t)o->gw * 8);
        w++;
    }
    free(newidx);
    o->n = w;
    update_counts(o);
}

int orc_step(orc *o) {
    double t = (double)o->step;
    o->agent_steps += orc_num_alive(o);
    o->births_last = o->deaths_last = o->moves_last = 0;
    for (int64_t i = 0; i < o->n; i++) { o->dest[i] = -1; o->bflag[i] = 0; }
    /* initializeStep: every action's initialize, in prio order */
    for (int a = 0; a < o->nact; a++) {
        switch (o->act[a]) {
        case QHG_ACT_RANDOMPAIR: random_pair(o); break;
        case QHG_ACT_ASSORTPAIR:
        case QHG_ACT_ASSORTPROT: assort_pair(o); break;
        case QHG_ACT_VERHULST: verhulst_init(o, 0); break;
        case QHG_ACT_VERHULSTVARK: verhulst_init(o, 1); break;
        case QHG_ACT_MULTIEVAL: if (o->me_first) { multi_eval(o); o->me_first = 0; } break;
        default: break;
        }
    }
    /* doActions for every prio level */
    for (int a = 0; a < o->nact; a++) {
        switch (o->act[a]) {
        case QHG_ACT_GETOLD: act_getold(o, t); break;
        case QHG_ACT_FERTILITY: act_fertility(o, t); break;
        case QHG_ACT_VERHULST:
        case QHG_ACT_VERHULSTVARK: act_verhulst(o); break;
        case QHG_ACT_RANDOMMOVE: act_randommove(o); break;
        case QHG_ACT_WEIGHTEDMOVE: act_weightedmove(o); break;
        case QHG_ACT_OLDAGEDEATH: act_oldagedeath(o, t); break;
        case QHG_ACT_ATANDEATH: act_atandeath(o, t); break;
        case QHG_ACT_NAVIGATE: act_navigate(o); break;
        default: break;
        }
    }
    finalize(o, t);
    o->births_total += o->births_last;
    o->deaths_total += o->deaths_last;
    o->moves_total += o->moves_last;
    o->step++;
    return 0;
}

int orc_run(orc *o, int32_t nsteps) {
    for (int32_t s = 0; s < nsteps; s++) orc_step(o);
    return 0;
}

void orc_stats(orc *o, qhg_stats *s) {
    memset(s, 0, sizeof(*s));
    s->num_alive = orc_num_alive(o);
    s->high_water = o->n;
    s->capacity = o->cap;
    s->max_id = o->max_id;
    s->step = o->step;
    s->births_last = o->births_last; s->deaths_last = o->deaths_last; s->moves_last = o->moves_last;
    s->births_total = o->births_total; s->deaths_total = o->deaths_total; s->moves_total = o->moves_total;
    s->agent_steps = o->agent_steps;
}

void orc_cell_counts(orc *o, uint32_t *out) { memcpy(out, o->count, (size_t)o->ncells * 4); }
void orc_capacities(orc *o, double *out) { memcpy(out, o->K, (size_t)o->ncells * 8); }
void orc_env_weights(orc *o, double *out) { memcpy(out, o->W, (size_t)o->ncells * (o->conn + 1) * 8); }
