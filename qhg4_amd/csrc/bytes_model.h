/*
 * bytes_model.h -- algorithmic HBM bytes of each step kernel.
 *
 * Compulsory traffic: every field a kernel must read is read once, every
 * field it must write is written once; per-cell tables are counted once per
 * cell per launch.  Caches may make a kernel move less than this (hits on
 * per-cell rows) or more (re-reads); rocprof FETCH_SIZE/WRITE_SIZE give the
 * measured traffic.  bench.py reads these constants from this file.
 *
 * Units: bytes.  "live" = agents alive at step start, "slot" = every slot in
 * [0, high-water) including holes, "mover" = agents that move, "birth" = new
 * agents, "fertile" = FERTILE agents bucketed for pairing.
 */
#ifndef QHG_BYTES_MODEL_H
#define QHG_BYTES_MODEL_H

/* k_act (fused doActions) */
#define QHG_B_ACT_SLOT          1   /* life byte of every slot                          */
#define QHG_B_ACT_LIVE_BASE    13   /* cell 4 + id 8 (Philox key) + life 1              */
#define QHG_B_ACT_LIVE_GETOLD  16   /* birth time r 8 + age w 8                         */
#define QHG_B_ACT_LIVE_FERT    10   /* gender 1 + last birth 8 + life w 1               */
#define QHG_B_ACT_LIVE_VERH     4   /* mate 4                                           */
#define QHG_B_ACT_CELL_VERH    12   /* count 4 + K 8 (VarK) / count 4 (+8 unused)       */
#define QHG_B_ACT_CELL_COUNT    4   /* next-step count written                          */
#define QHG_B_ACT_CELL_WMOVE   81   /* weight row 7*8 + neighbours 6*4 + count byte 1   */
#define QHG_B_ACT_CELL_RMOVE   25   /* neighbours 6*4 + count byte 1                    */
#define QHG_B_ACT_MOVER         4   /* new cell written                                 */
#define QHG_B_ACT_BIRTH        20   /* birth record 12 + mother's NumBabies r/w 8       */

/* k_newborn (createAgentAtIndex + Genetics::makeOffspring) */
#define QHG_B_NEWBORN_BIRTH    62   /* record r 12 + agent w 46 + count 4               */
#define QHG_B_NEWBORN_GENOME  384   /* 2 parent genomes r 2*128 + child w 128 (2x256 loci) */

/* k_bucket (RandomPair::initialize bucketing) */
#define QHG_B_BUCKET_SLOT       1   /* life byte                                        */
#define QHG_B_BUCKET_LIVE       4   /* mate := -3                                       */
#define QHG_B_BUCKET_FERTILE   17   /* cell 4 + gender 1 + bucket w 4 + fill 4 + off 4  */

/* k_pair (per-cell pairing) */
#define QHG_B_PAIR_CELL        16   /* fill F/M 8 + offsets 8                           */
#define QHG_B_PAIR_FERTILE     20   /* bucket r+w 8 + id 8 + mate w 4                   */

/* k_cellscan (counts -> offsets) */
#define QHG_B_SCAN_CELL        20   /* count r 4 + offset w 4 + fills w 8 + clear w 4   */

#endif
