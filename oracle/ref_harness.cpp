// ref_harness -- runs the reference's own population code, unmodified and
// compiled from the reference tree, for the statistical comparison with this
// build (TEST INFRASTRUCTURE ONLY: nothing in the product links or runs it).
//
// The population is QHG4's tut_SexualPop (QHG4/populations/tut_SexualPop.cpp,
// with SPopulation, Prioritizer and the GetOld / ATanDeath / RandomMove /
// Fertility / Verhulst / RandomPair actions, LBController and WELL512), on a
// non-periodic W x H lattice built as GridFactory::createCellsRect builds it
// (QHG4/tools_io/GridFactory.cpp:966-1068).  The species parameters and the
// initial agents are read from a QDF file written by this build's QDF
// library, through the reference's PopReader (QHG4/io/PopReader.cpp:133-183),
// so the reference also checks that it reads those files.  The step loop is
// PopLooper::doStep's (QHG4/core/PopLooper.cpp:166-202).  Every draw comes
// from the reference's per-thread WELL512 streams, seeded from `seed`.
//
// usage: ref_harness <pop.qdf> <species> <width> <height> <seed> <steps>
// prints one line per step: step, live agents, occupied cells (0 0 once the
// population has died out).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>

#include <omp.h>

#include "IDGen.h"
#include "PopReader.h"
#include "SCellGrid.h"
#include "WELL512.h"
#include "tut_SexualPop.h"

namespace {

// the lattice of GridFactory::createCellsRectPeriodic (slot 0 +x, 1 +y,
// 2 -x, 3 -y) with the non-periodic border links set to -1 and the
// neighbour counts decremented (createCellsRect)
SCellGrid *lattice(int W, int H) {
    stringmap surf;
    surf[SURF_TYPE] = SURF_LATTICE;
    surf[SURF_LTC_LINKS] = "4";
    SCellGrid *cg = new SCellGrid(0, (uint)(W * H), surf);
    cg->m_aCells = new SCell[W * H];
    cg->m_iMaxNeighbors = 4;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const int i = y * W + x;
            SCell &c = cg->m_aCells[i];
            c.m_iGlobalID = i;
            c.m_iNumNeighbors = 4;
            c.m_aNeighbors[0] = y * W + (x + 1) % W;
            c.m_aNeighbors[1] = ((y + 1) % H) * W + x;
            c.m_aNeighbors[2] = y * W + (x - 1 + W) % W;
            c.m_aNeighbors[3] = ((y - 1 + H) % H) * W + x;
            c.m_aNeighbors[4] = -1;
            c.m_aNeighbors[5] = -1;
            cg->m_mIDIndexes[i] = i;
        }
    }
    for (int y = 0; y < H; y++) {
        SCell &r = cg->m_aCells[y * W + W - 1];
        r.m_aNeighbors[0] = -1;
        r.m_iNumNeighbors--;
        SCell &l = cg->m_aCells[y * W];
        l.m_aNeighbors[2] = -1;
        l.m_iNumNeighbors--;
    }
    for (int x = 0; x < W; x++) {
        SCell &t = cg->m_aCells[(H - 1) * W + x];
        t.m_aNeighbors[1] = -1;
        t.m_iNumNeighbors--;
        SCell &b = cg->m_aCells[x];
        b.m_aNeighbors[3] = -1;
        b.m_iNumNeighbors--;
    }
    return cg;
}

// WELL512 states from the seed: SplitMix64 outputs (any non-zero state is a
// valid WELL512 state)
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 7) {
        fprintf(stderr, "usage: %s <pop.qdf> <species> <width> <height> <seed> <steps>\n", argv[0]);
        return 2;
    }
    const std::string path = argv[1], species = argv[2];
    const int W = atoi(argv[3]), H = atoi(argv[4]);
    uint64_t s = strtoull(argv[5], nullptr, 10);
    const int steps = atoi(argv[6]);

    SCellGrid *cg = lattice(W, H);
    const int T = omp_get_max_threads();
    IDGen **idg = new IDGen *[T];
    for (int t = 0; t < T; t++) idg[t] = new IDGen(0, t, T);   // SimParams.cpp:134-141
    uint32_t state[STATE_SIZE];
    for (unsigned k = 0; k < STATE_SIZE; k++) state[k] = (uint32_t)splitmix(s) | 1u;
    uint seeds[NUM_SEEDS];
    for (int k = 0; k < NUM_SEEDS; k++) seeds[k] = (uint)splitmix(s);

    tut_SexualPop *pop = new tut_SexualPop(cg, nullptr, 65536, idg, state, seeds);
    pop->setAgentDataType();   // as SimParams::setPops does before reading (SimParams.cpp:1483)
    PopReader *pr = PopReader::create(path);
    if (!pr) {
        fprintf(stderr, "cannot open %s\n", path.c_str());
        return 1;
    }
    int rc = pr->read(pop, species, W * H, false);
    delete pr;
    if (rc != 0) {
        fprintf(stderr, "PopReader::read failed (%d)\n", rc);
        return 1;
    }
    // PopLooper::addPop (PopLooper.cpp:96-113) and the id bases
    if (pop->setPrioList() != 0) {
        fprintf(stderr, "setPrioList failed\n");
        return 1;
    }
    std::set<uint> prios;
    static_cast<PopBase *>(pop)->getPrios(prios);   // public in PopBase
    const idtype maxid = pop->getMaxLoadedID();
    for (int t = 0; t < T; t++) idg[t]->setBase(maxid + 1);
    if (pop->preLoop() != 0) {
        fprintf(stderr, "preLoop failed\n");
        return 1;
    }
    for (int step = 0; step < steps; step++) {
        if (pop->getNumAgentsEffective() == 0) {
            // extinct: the reference's step does not handle an empty
            // population, and nothing can change any more
            printf("%d 0 0\n", step + 1);
            continue;
        }
        const float t = (float)step;
        rc = pop->initializeStep(t);
        for (uint p : prios) rc += pop->doActions(p, t);
        rc += pop->finalizeStep();
        if (rc != 0) {
            fprintf(stderr, "step %d failed\n", step);
            return 1;
        }
        pop->updateNumAgentsPerCell();
        ulong occupied = 0;
        for (int c = 0; c < W * H; c++) occupied += pop->getNumAgents(c) > 0 ? 1 : 0;
        printf("%d %lu %lu\n", step + 1, pop->getNumAgentsEffective(), occupied);
    }
    delete pop;
    return 0;
}
