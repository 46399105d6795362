// Test infrastructure: C entry points over the reference's own utility code
// (WELL512 conversions, GeneUtils mutation/recombination, BinomialDist,
// bino_tools, RankTable, ProteinBuilder), compiled from the reference sources by `make -C oracle ref`
// into oracle/_ref/libqhg_ref.so.  The tests feed these and the oracle the
// same raw 32-bit draws and compare the results bit for bit; nothing outside
// tests/ loads this library.
//
// The reference functions draw from a WELL512 object themselves; every entry
// point here starts two generators from the same state: one is handed to the
// reference function, the other reports the raw wrand() words that the first
// consumed (each conversion takes exactly one word, QHG4/utils/WELL512.h:33-39).
#include <cstdint>
#include <cstring>
#include <vector>

#include "BinomialDist.h"
#include "GeneUtils.h"
#include "ProteinBuilder.h"
#include "RankTable.h"
#include "WELL512.h"
#include "bino_tools.h"

double gammaln(double xx);   // QHG4/utils/bino_tools.cpp:13

namespace {

// makeMultiMask is protected (QHG4/genes/GeneUtils.h:75)
struct GeneProbe : public GeneUtils {
    static ulong multimask(const std::vector<uint> &v) { return makeMultiMask(v); }
};

void raw_words(const uint32_t *state, uint32_t index, int n, uint32_t *out) {
    uint32_t st[16];
    std::memcpy(st, state, sizeof(st));
    WELL512 w(st, index);
    for (int i = 0; i < n; i++) out[i] = w.wrand();
}

}  // namespace

extern "C" {

// n draws of each conversion from the same state: raw, wrandd, wrandr(a,b),
// wrandi(ia, ib, is)   (QHG4/utils/WELL512.h:30-39)
void ref_well_draws(const uint32_t *state, uint32_t index, int n, uint32_t *raw, double *d, double *r,
                    uint32_t *iv, double a, double b, uint32_t ia, uint32_t ib, uint32_t is) {
    raw_words(state, index, n, raw);
    uint32_t s1[16], s2[16], s3[16];
    std::memcpy(s1, state, sizeof(s1));
    std::memcpy(s2, state, sizeof(s2));
    std::memcpy(s3, state, sizeof(s3));
    WELL512 w1(s1, index), w2(s2, index), w3(s3, index);
    for (int i = 0; i < n; i++) {
        d[i] = w1.wrandd();
        r[i] = w2.wrandr(a, b);
        iv[i] = w3.wrandi(ia, ib, is);
    }
}

// GeneUtils::makeMultiMask (QHG4/genes/GeneUtils.cpp:156-175)
uint64_t ref_multimask(const uint32_t *breaks, int n) {
    std::vector<uint> v(breaks, breaks + n);
    return GeneProbe::multimask(v);
}

// GeneUtils::crossOver (QHG4/genes/GeneUtils.cpp:184-246): out gets both
// strands; raw gets the nbreaks words it drew
void ref_crossover(uint64_t *out, const uint64_t *parent, int genome_size, int nbreaks, const uint32_t *state,
                   uint32_t *raw) {
    raw_words(state, 0, nbreaks, raw);
    uint32_t st[16];
    std::memcpy(st, state, sizeof(st));
    WELL512 w(st, 0);
    GeneUtils::crossOver((ulong *)out, (const ulong *)parent, genome_size, nbreaks, &w);
}

// GeneUtils::freeReco (QHG4/genes/GeneUtils.cpp:322-360): 2 words per block
void ref_free_reco(uint64_t *out, uint64_t *parent, int nblocks, const uint32_t *state, uint32_t *raw) {
    raw_words(state, 0, 2 * nblocks, raw);
    uint32_t st[16];
    std::memcpy(st, state, sizeof(st));
    WELL512 w(st, 0);
    GeneUtils::freeReco((ulong *)out, (ulong *)parent, nblocks, &w);
}

// GeneUtils::mutateNucs (QHG4/genes/GeneUtils.cpp:112-141): 2 words per
// mutation (position, mask)
void ref_mutate_nucs(uint64_t *genome, int genome_size, int nmut, const uint32_t *state, uint32_t *raw) {
    raw_words(state, 0, 2 * nmut, raw);
    uint32_t st[16];
    std::memcpy(st, state, sizeof(st));
    WELL512 w(st, 0);
    GeneUtils::mutateNucs((ulong *)genome, genome_size, nmut, &w);
}

// BinomialDist::create + getN (QHG4/utils/BinomialDist.cpp:10-104)
int ref_bino_getn(double p, int n, double eps, const double *u, int m, int32_t *out) {
    BinomialDist *bd = BinomialDist::create(p, n, eps);
    if (!bd) return -1;
    for (int i = 0; i < m; i++) out[i] = bd->getN(u[i]);
    delete bd;
    return 0;
}

// bino_tools.cpp:13 and :94
double ref_gammaln(double x) { return gammaln(x); }
double ref_beta(double a, double b, double x) { return beta(a, b, x); }

// GeneUtils::calcDist (QHG4/genes/GeneUtils.cpp:414-444)
int ref_calc_dist(uint64_t *g1, uint64_t *g2, int size_arg) {
    return GeneUtils::calcDist((ulong *)g1, (ulong *)g2, size_arg);
}

// RankTable::createInstance + makeAllPairings (QHG4/utils/RankTable.cpp:17-150)
// over a row-major nf x nm table, with one WELL512 generator (thread 0's).
// Returns the number of pairs; *nraw = raw words consumed (copied to raw).
int ref_rank_pairs(int nf, int nm, const float *ranks, float cutoff, int permute, const uint32_t *state,
                   int32_t *pairs, uint32_t *raw, int maxraw, int *nraw) {
    uint32_t st[16];
    std::memcpy(st, state, sizeof(st));
    WELL512 w(st, 0);
    WELL512 *aw[1] = {&w};
    std::vector<float *> rows((size_t)nf);
    for (int i = 0; i < nf; i++) rows[(size_t)i] = const_cast<float *>(ranks + (size_t)i * nm);
    RankTable *rt = RankTable::createInstance((uint)nf, (uint)nm, aw, rows.data());
    if (!rt) return -1;
    rt->makeAllPairings(cutoff, permute != 0);
    const couples &c = rt->getPairs();
    for (size_t k = 0; k < c.size(); k++) {
        pairs[2 * k] = (int32_t)c[k].first;
        pairs[2 * k + 1] = (int32_t)c[k].second;
    }
    const int n = (int)c.size();
    delete rt;
    *nraw = (int)w.getCount();
    if (*nraw <= maxraw) raw_words(state, 0, *nraw, raw);
    return n;
}

// ProteinBuilder::translateGenomeHash on a genome (QHG4/genes/ProteinBuilder.cpp:
// 163-220): the MD5 digests of its proteins, 16 bytes each, at most max of
// them into out (in the builder's order: it sorts the digest pointers);
// returns the number of proteins, -1 when the builder refuses the genome
int ref_protein_hashes(const uint64_t *genome, int genome_size, uint8_t *out, int max) {
    ProteinBuilder *pb = ProteinBuilder::createInstance(reinterpret_cast<const ulong *>(genome), genome_size);
    if (!pb) return -1;
    int n = -1;
    if (pb->translateGenomeHash() == 0) {
        const prothash &h = pb->getProtHash();
        n = (int)h.size();
        for (int i = 0; i < n && i < max; i++) std::memcpy(out + 16 * i, h[i], 16);
    }
    delete pb;
    return n;
}

}  // extern "C"
