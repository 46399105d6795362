/*
 * qhg_oracle.h -- CPU oracle of the QHG4 per-step population loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X path and the CPU baseline that bench.py times beside it.  Nothing
 * in the product (qhg4_amd/) links, loads or calls it.
 *
 * It restates the reference algorithm (cited per function in the .c file)
 * with every WELL512 draw replaced by a counter-based Philox4x32-10 draw
 * keyed by (seed, step, stream, agent id, draw#).  Agents are kept in id
 * order; slot recycling of the reference allocator is not modelled (slot
 * indices are not part of parity).
 */
#ifndef QHG_ORACLE_H
#define QHG_ORACLE_H

#include "../include/qhg_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

int      orc_create(const qhg_grid_desc *grid, const qhg_env_desc *env,
                    const qhg_params *params, orc **out);
void     orc_destroy(orc *o);
int      orc_set_agents(orc *o, const qhg_agents *a);
int64_t  orc_num_alive(orc *o);
int      orc_get_agents(orc *o, qhg_agents *a);
int      orc_pre_loop(orc *o);
int      orc_set_params(orc *o, const qhg_params *p);
int      orc_env_event(orc *o, const qhg_env_desc *env, int32_t events);
int      orc_set_navigation(orc *o, const qhg_nav_desc *nav);
int      orc_step(orc *o);
int      orc_run(orc *o, int32_t nsteps);
void     orc_stats(orc *o, qhg_stats *s);
void     orc_cell_counts(orc *o, uint32_t *out);
void     orc_capacities(orc *o, double *out);
void     orc_env_weights(orc *o, double *out);
void     orc_set_threads(int32_t n);
int32_t  orc_get_threads(void);

/* building blocks, exported so tests can pin them one by one */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_draw(uint64_t seed, uint32_t step, uint32_t stream, uint64_t id, uint32_t n);
double   orc_wrandd(uint32_t u);
double   orc_wrandr(double a, double b, uint32_t u);
uint32_t orc_wrandi(uint32_t a, uint32_t b, uint32_t s, uint32_t u);
/* RandomPair's k-th-unpaired rule for one cell: nchoose choosers (the
 * smaller sex, id order) with draws u[], ncand candidates; pick[j] = the
 * list index chooser j takes.  taken is ncand bytes of scratch. */
void     orc_pair_cell(int32_t nchoose, int32_t ncand, const uint32_t *u, uint8_t *taken, int32_t *pick);
/* AssortativePairHamming pieces: GeneUtils::calcDist and the RankTable
 * pairing over a row-major nf x nm table with the draws u[0..nu) */
int32_t  orc_calc_dist(const uint64_t *g1, const uint64_t *g2, int32_t size_arg);
/* MD5 of len bytes (RFC 1321) */
void     orc_md5(const uint8_t *msg, int64_t len, uint8_t out[16]);
/* the MD5 digests of a genome's proteins (ProteinBuilder::translateGenomeHash):
 * at most max written (16 bytes each); returns the number of proteins */
int32_t  orc_proteome(const uint64_t *genome, int32_t gwords, int32_t genome_size, uint8_t *digests, int32_t max);
/* equal digest pairs of two lists (ProteomeComparator::countProtHashMatches) */
int32_t  orc_prot_matches(const uint8_t *a, int32_t na, const uint8_t *b, int32_t nb);
int32_t  orc_rank_pairs(int32_t nf, int32_t nm, float *ranks, float cutoff, int32_t permute, const uint32_t *u,
                        int32_t nu, int32_t *pairs, int32_t *used);
uint64_t orc_multimask(const uint32_t *breaks, int32_t n);
void     orc_crossover(uint64_t *out, const uint64_t *parent, int32_t genome_size,
                       const uint32_t *break_pos, int32_t nbreaks);
/* GeneUtils::freeReco (GeneUtils.cpp:322-360): 2 draws per block, hi then lo */
void     orc_free_reco(uint64_t *out, const uint64_t *parent, int32_t nb, const uint32_t *draws);
void     orc_mutate_nucs(uint64_t *genome, int32_t genome_size, int32_t nmut,
                         const uint32_t *pos_draws, const uint32_t *mask_draws);
int32_t  orc_bino_table(double p, int32_t n, double eps, double *cdf, int32_t max);
int32_t  orc_bino_getn(const double *cdf, int32_t nbins, double u);
double   orc_betai(double a, double b, double x);
double   orc_polyline(int32_t npts, const double *x, const double *v, double X);
double   orc_atan(double x);
double   orc_exp(double x);
int32_t  orc_genome_words(int32_t genome_size);
/* Navigate pieces: the jump table of the port at cell (n + 1 entries, the
 * stay-home entry first, cumulative; -1 when the cell has none), the
 * current bridges (pairs), and Navigate::execute for an agent at cell over
 * the given draws in the order it takes them (jump first, then one per
 * bridge at the cell): returns the number of registered moves, their
 * destinations in dst (at most max) and the draws taken in *used */
int32_t  orc_nav_table(orc *o, int32_t cell, int32_t *targ, double *cum, int32_t max);
int32_t  orc_nav_bridges(orc *o, int32_t *pairs, int32_t max);
int32_t  orc_nav_execute(orc *o, int32_t cell, const double *draws, int32_t ndraws, int32_t *used, int32_t *dst,
                         int32_t max);

#ifdef __cplusplus
}
#endif
#endif
