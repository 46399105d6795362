/*
 * bytes_model.h -- algorithmic HBM bytes of each step kernel.
 *
 * Compulsory traffic: every field a kernel must read is read once, every
 * field it must write is written once; per-cell rows are counted once per
 * occupied cell per launch.  Caches may make a kernel move less than this or
 * more (re-reads); rocprof FETCH_SIZE/WRITE_SIZE give the measured traffic.
 * bench.py reads these constants from this file.
 *
 * Units: bytes.  "occupied" = cells holding agents at step start, "live" =
 * agents alive at step start, "survivor" = live agents that do not die in the
 * step, "birth" = new agents, "conn" = the grid connectivity, "gwords" =
 * 64-bit words of one genome (both strands).
 */
#ifndef QHG_BYTES_MODEL_H
#define QHG_BYTES_MODEL_H

/* k_cell_act (RandomPair::initialize/findMates + fused doActions).  Only
   occupied cells are visited: empty cells keep the zero counts the last
   scan left, so no term runs over every cell. */
#define QHG_B_ACT_OCC           29  /* occupied-cell entry r 16 + K r 8 + neighbour count
                                       r 1 + stayer count w 4                               */
#define QHG_B_ACT_OCC_PER_CONN  16  /* neighbour index r 4 + run slot r 4 + weight r 8     */
#define QHG_B_ACT_OCC_W_EXTRA    8  /*   ... the extra weight of the row (conn+1 weights)   */
#define QHG_B_ACT_LIVE_BASE     10  /* id r 8 (Philox key) + life r 1 + outcome w 1          */
#define QHG_B_ACT_LIVE_PAIR     10  /* gender r 1 + life r 1 (pairing pass) + partner id w 8 */
#define QHG_B_ACT_LIVE_GETOLD   16  /* birth time r 8 + age w 8                              */
#define QHG_B_ACT_LIVE_FERT     10  /* gender r 1 + last birth r 8 + life w 1                */
#define QHG_B_ACT_BIRTH         12  /* mother's NumBabies r/w 8 + father slot w 4            */

/* k_cell_scan (next counts -> segment and birth offsets).  Counted per
   occupied cell only: the scan also reads the empty cells of tiles that
   hold agents, and the incoming-run rows of cells that receive movers
   (count r 4 + zero w 4 + source r 4 + position w 4 per moving run) are
   not counted: a lower bound. */
#define QHG_B_SCAN_OCC          20  /* stayers r 4 + births r 4 + arrivals r 4 + next
                                       offset w 4 + occupied-cell entry share 4             */
/* k_batch_scan_sums (the batched scan's first pass): the same three counts
   read to sum each tile, per occupied cell (a lower bound as above); its
   second pass is counted as k_cell_scan */
#define QHG_B_SCAN_SUMS_OCC     12  /* stayers r 4 + births r 4 + arrivals r 4              */

/* k_cell_scatter (performMoves + createAgentAtIndex + makeOffspring) */
#define QHG_B_SCAT_OCC          20  /* occupied-cell entry r 16 + next offset r 4           */
#define QHG_B_SCAT_LIVE          1  /* outcome byte r of every agent at step start           */
#define QHG_B_SCAT_SURVIVOR     93  /* record r 46 + w 46 (id, partner, life, gender, birth,
                                       age, last birth, NumBabies) + outcome r 1             */
#define QHG_B_SCAT_SURVIVOR_GEN  8  /* genome handle r 4 + w 4                               */
#define QHG_B_SCAT_BIRTH        54  /* newborn record w 46 + birth offset r 4 + father r 4   */
#define QHG_B_SCAT_BIRTH_GEN    12  /* handle w 4 + birth record w 8                         */

/* k_assort_pair (AssortativePairHamming), per live agent and per FERTILE
   agent of a cell holding both sexes; each couple's genome words are read
   through the caches, counted once per agent (a lower bound) */
#define QHG_B_ASSORT_LIVE       22  /* life r 1 + gender r 1 + id r 8 + partner slot w 4
                                       + mate id w 8                                         */
#define QHG_B_ASSORT_FERTILE     4  /* genome handle r 4, + 2 x D x 8 genome words r          */

/* k_genome (Genetics::makeOffspring), per birth */
#define QHG_B_SCAT_GENOME_WORDS  3  /* x gwords x 8: two parent genomes r + child w          */

#endif
