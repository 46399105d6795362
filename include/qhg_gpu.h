/*
 * qhg_gpu.h -- C ABI of the MI355X population step engine.
 *
 * One handle = one population on one GPU (one HIP stream).  The entry
 * points below replace the per-step population interface of QHG4:
 *
 *   PopBase::preLoop              QHG4/core/PopBase.h:16-121, SPopulation.cpp:273
 *   SPopulation::initializeStep   QHG4/core/SPopulation.cpp:394-417
 *   SPopulation::doActions        QHG4/core/SPopulation.cpp:554-577
 *   SPopulation::finalizeStep     QHG4/core/SPopulation.cpp:439-477
 *   PopLooper::doStep             QHG4/core/PopLooper.cpp:166-202
 *   SPopulation::getNumAgentsArray / getNumAgentsEffective
 *                                 QHG4/core/SPopulation.h:143-148
 *   SPopulation::addAgent / writeAgentDataQDFSafe (agent upload/download)
 *                                 QHG4/core/SPopulation.cpp:1465-1567
 *
 * Conventions (follow the reference's error convention,
 * e.g. QHG4/actions/Verhulst.cpp:80-91): every function returns 0 on
 * success and a negative code on error; qhg_gpu_last_error() returns a
 * thread-local message for the last failing call.  Host arrays are borrowed
 * only for the duration of a call.  Calls on one handle are not re-entrant;
 * different handles may be used concurrently.
 *
 * Agent time fields (BirthTime, Age, LastBirth) are FP64.
 */
#ifndef QHG_GPU_H
#define QHG_GPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QHG_ABI_VERSION 8

/* error codes */
#define QHG_OK              0
#define QHG_ERR_ARG        -1
#define QHG_ERR_HIP        -2
#define QHG_ERR_CAPACITY   -3
#define QHG_ERR_STATE      -4
#define QHG_ERR_NOMEM      -5
#define QHG_ERR_MEMBER     -6   /* batch: member(s) failed and left the batch; the others completed the call */

/* life states, QHG4/core/SPopulation.h:70-74 */
#define QHG_LIFE_DEAD     0
#define QHG_LIFE_ALIVE    1
#define QHG_LIFE_FERTILE  5

/* action kinds (one per supported Action<T> subclass) */
#define QHG_ACT_GETOLD        1   /* QHG4/actions/GetOld.cpp:37-48            */
#define QHG_ACT_FERTILITY     2   /* QHG4/actions/Fertility.cpp:49-74         */
#define QHG_ACT_RANDOMPAIR    3   /* QHG4/actions/RandomPair.cpp:107-279      */
#define QHG_ACT_VERHULST      4   /* QHG4/actions/Verhulst.cpp:70-115         */
#define QHG_ACT_VERHULSTVARK  5   /* QHG4/actions/VerhulstVarK.cpp:69-115     */
#define QHG_ACT_RANDOMMOVE    6   /* QHG4/actions/RandomMove.cpp:65-100       */
#define QHG_ACT_WEIGHTEDMOVE  7   /* QHG4/actions/WeightedMove.cpp:45-106     */
#define QHG_ACT_MULTIEVAL     8   /* QHG4/actions/MultiEvaluator.cpp:142-478  */
#define QHG_ACT_NPPCAP        9   /* QHG4/actions/NPPCapacity.cpp:138-217     */
#define QHG_ACT_GENETICS     10   /* QHG4/actions/Genetics.cpp:285-337        */
#define QHG_ACT_OLDAGEDEATH  11   /* QHG4/actions/OldAgeDeath.cpp (execute)   */
#define QHG_ACT_ATANDEATH    12   /* QHG4/actions/ATanDeath.cpp:48-90         */
#define QHG_ACT_ASSORTPAIR   13   /* QHG4/actions/AssortativePairHam.cpp:93-238
                                     + QHG4/utils/RankTable.cpp:119-278        */
#define QHG_ACT_ASSORTPROT   14   /* QHG4/actions/AssortativePair.cpp:94-271:
                                     matches counted over the proteins of
                                     QHG4/genes/ProteinBuilder.cpp:163-220, 231-289
                                     (MD5 per protein) and
                                     QHG4/genes/ProteomeComparator.cpp:25-60,
                                     then the same RankTable pairing       */
#define QHG_ACT_NAVIGATE     15   /* QHG4/actions/Navigate.cpp:99-244: sea-way
                                     jumps from ports and bridge crossings,
                                     over the tables qhg_gpu_set_navigation
                                     gives                                  */
#define QHG_ACT_LAST         QHG_ACT_NAVIGATE

/* MultiEvaluator combination modes, QHG4/actions/MultiEvaluator.h:20-25 */
#define QHG_ME_ADD_SIMPLE 0
#define QHG_ME_MUL_SIMPLE 2
#define QHG_ME_MAX_SIMPLE 3
#define QHG_ME_MIN_SIMPLE 5

/* inputs a SingleEvaluator can read */
#define QHG_EV_ALTITUDE   0
#define QHG_EV_CAPACITY   1   /* NPPCapacity output (K per cell) */
#define QHG_EV_BASE_NPP   2

#define QHG_MAX_ACTIONS   16
#define QHG_MAX_EVALS     4
#define QHG_MAX_PL_POINTS 64
#define QHG_MAX_BINO      64
#define QHG_NAME_LEN      48

/* Philox stream tags: the 16-bit stream field of the counter.  Each agent
 * draws from its own (step, stream, id) counters: one pairing word, one
 * block of action words, and for a newborn its gender and genetics words. */
#define QHG_STREAM_PAIR     1   /* word 0: the chooser's pick in RandomPair     */
#define QHG_STREAM_ACT      2   /* per-agent action words, QHG_ACTW_*           */
#define QHG_STREAM_NEWBORN  6   /* word 0: gender of the newborn                */
#define QHG_STREAM_GENETICS 7   /* Genetics::makeOffspring draws, in order      */
#define QHG_STREAM_ASSORT   8   /* AssortativePair(Hamming): one cell's draws in order
                                   (key: cell index | (draw >> 16) << 32, word: draw & 0xffff) */
#define QHG_STREAM_NAV      9   /* Navigate: word 0 the jump draw, word 1 + b the
                                   draw of current bridge b                     */

/* words of QHG_STREAM_ACT (one Philox block covers words 0-3) */
#define QHG_ACTW_WMOVE_GO   0   /* WeightedMove: move?                          */
#define QHG_ACTW_WMOVE_DIR  1   /* WeightedMove: target slot                    */
#define QHG_ACTW_LBIRTH     2   /* LinearBirth                                  */
#define QHG_ACTW_LDEATH     3   /* LinearDeath                                  */
#define QHG_ACTW_RMOVE_GO   4   /* RandomMove: move?                            */
#define QHG_ACTW_RMOVE_DIR  5   /* RandomMove: target slot                      */
#define QHG_ACTW_OADEATH    6   /* OldAgeDeath: age threshold jitter            */
#define QHG_ACTW_ATANDEATH  7   /* ATanDeath: death draw                        */

typedef struct qhg_grid_desc {
    int32_t        num_cells;
    int32_t        connectivity;   /* max neighbours per cell: 4 lattice, 6 ico */
    const int32_t *neighbors;      /* num_cells*connectivity, -1 padded        */
    const uint8_t *num_neighbors;  /* num_cells                                */
    const int64_t *cell_ids;       /* num_cells global ids, NULL => index      */
} qhg_grid_desc;

/* Geography / climate / vegetation arrays (any may be NULL). */
typedef struct qhg_env_desc {
    const double  *altitude;
    const double  *longitude;
    const double  *latitude;
    const double  *water;
    const uint8_t *ice;
    const uint8_t *coastal;
    const double  *base_npp;
    const double  *temperature;   /* annual mean temp (Miami NPP) */
    const double  *rainfall;      /* annual rainfall  (Miami NPP) */
} qhg_env_desc;

/* Parameter block.  Field names follow the reference's attribute strings
 * (e.g. Verhulst_b0 -> verh_b0); the host-side action classes fill it. */
typedef struct qhg_params {
    uint64_t seed;
    int32_t  num_actions;
    int32_t  action_kind[QHG_MAX_ACTIONS];
    int32_t  action_prio[QHG_MAX_ACTIONS];
    char     action_name[QHG_MAX_ACTIONS][QHG_NAME_LEN];

    /* Fertility_min_age, Fertility_max_age, Fertility_interbirth */
    double fert_min_age, fert_max_age, fert_interbirth;
    /* Verhulst_b0, Verhulst_d0, Verhulst_theta, Verhulst_K */
    double verh_b0, verh_d0, verh_theta, verh_K;
    /* RandomMove_prob, WeightedMove_prob */
    double rmove_prob, wmove_prob;

    /* MultiEvaluator: Multi_weight_* and the SingleEvaluator PolyLines */
    int32_t me_mode;
    int32_t me_num;
    int32_t me_input[QHG_MAX_EVALS];
    int32_t me_cumulate[QHG_MAX_EVALS];
    double  me_weight[QHG_MAX_EVALS];
    int32_t me_pl_n[QHG_MAX_EVALS];          /* points; 0 = identity       */
    double  me_pl_x[QHG_MAX_EVALS][QHG_MAX_PL_POINTS];
    double  me_pl_v[QHG_MAX_EVALS][QHG_MAX_PL_POINTS];

    /* NPPCap_water_factor, NPPCap_coastal_factor, NPPCap_coastal_min_latitude,
       NPPCap_coastal_max_latitude, NPPCap_NPP_min, NPPCap_NPP_max,
       NPPCap_K_min, NPPCap_K_max, NPPCap_efficiency */
    double npp_water_factor, npp_coastal_factor, npp_coastal_min_lat, npp_coastal_max_lat;
    double npp_min, npp_max, npp_k_min, npp_k_max, npp_efficiency;

    /* Genetics_genome_size, Genetics_num_crossover, Genetics_mutation_rate */
    int32_t gen_genome_size;        /* loci per strand                       */
    int32_t gen_num_crossover;      /* >0 fixed count, 0 none, -1 free reco   */
    double  gen_mutation_rate;

    /* OAD_max_age, OAD_uncertainty */
    double oad_max_age, oad_uncertainty;
    /* ATanDeath_max_age, ATanDeath_range, ATanDeath_slope */
    double atan_max_age, atan_range, atan_slope;
    /* AssortativePairHamming_cutoff / AssortativePair_cutoff (a float in
       the reference), ..._permute (0 / 1): a class has one of the two */
    double  assort_cutoff;
    int32_t assort_permute;
    int32_t assort_pad;
    /* Navigate_decay, Navigate_dist0, Navigate_prob0, Navigate_min_dens
       (read, not used by the action, as in the reference),
       Navigate_bridge_prob (0 when a QDF species group lacks it) */
    double nav_decay, nav_dist0, nav_prob0, nav_min_dens, nav_bridge_prob;
} qhg_params;

/* Agent records as structure-of-arrays.  Arrays the caller does not need may
 * be NULL on download.  mate_id is the partner's agent id, -3 when unpaired
 * this step, -1 when the partner has since died. */
typedef struct qhg_agents {
    int64_t   n;
    int64_t  *id;
    int32_t  *cell;
    uint8_t  *life;
    uint8_t  *gender;
    double   *birth_time;
    double   *age;
    double   *last_birth;
    int64_t  *mate_id;
    int32_t  *num_babies;
    uint64_t *genome;      /* n * 2 * genome_words(gen_genome_size) */
} qhg_agents;

typedef struct qhg_stats {
    int64_t  num_alive;
    int64_t  high_water;     /* slots in use (the buffers hold no holes: = num_alive) */
    int64_t  capacity;
    int64_t  max_id;
    int64_t  step;
    int64_t  births_last, deaths_last, moves_last;
    int64_t  births_total, deaths_total, moves_total;
    int64_t  compactions;
    int32_t  device_error;
    int32_t  host_error;     /* code with which a batch step of this population failed, 0 if none */
    int64_t  agent_steps;    /* sum over steps of the live count at step start */
    int64_t  slot_steps;     /* sum over steps of the slots in use at step start */
    int64_t  fertile_total;  /* FERTILE agents considered for pairing, all steps */
    int64_t  occupied_cell_steps;  /* sum over steps of the cells holding agents */
} qhg_stats;

typedef struct qhg_pop qhg_pop;

/* version / diagnostics */
int         qhg_gpu_abi_version(void);
const char *qhg_gpu_last_error(void);
int         qhg_gpu_genome_words(int32_t genome_size);   /* u64 per strand */

/* Binomial mutation-count table (QHG4/utils/BinomialDist.cpp:60-104):
 * fills cdf[0..*n) and returns 0.  Shared by the device path. */
int qhg_gpu_binomial_table(double p, int32_t n, double eps, double *cdf, int32_t *count);

/* lifecycle: SPopulation ctor/dtor (QHG4/core/SPopulation.cpp:46-140) */
int qhg_gpu_create(const qhg_grid_desc *grid, const qhg_env_desc *env,
                   const qhg_params *params, int64_t capacity, qhg_pop **out);
int qhg_gpu_destroy(qhg_pop *pop);

/* agents: SPopulation::addAgent / reading agents (QHG4/core/SPopulation.cpp
 * :1190-1290) and writeAgentDataQDFSafe (:1465-1567).  Upload replaces the
 * population; agents must have distinct ids. Download returns live agents
 * in id order; qhg_gpu_num_alive() gives the size to allocate. */
int qhg_gpu_upload_agents(qhg_pop *pop, const qhg_agents *agents);
int qhg_gpu_num_alive(qhg_pop *pop, int64_t *n);
int qhg_gpu_download_agents(qhg_pop *pop, qhg_agents *agents);

/* step phases, PopLooper::doStep order (QHG4/core/PopLooper.cpp:166-202) */
int qhg_gpu_pre_loop(qhg_pop *pop, double t);
int qhg_gpu_initialize_step(qhg_pop *pop, double t);
int qhg_gpu_do_actions(qhg_pop *pop, uint32_t prio, double t);
int qhg_gpu_finalize_step(qhg_pop *pop);
/* nsteps whole steps at t = step, step+1, ... (no host sync per step) */
int qhg_gpu_run(qhg_pop *pop, int32_t nsteps);

/* environment change (EVENT_ID_GEO/VEG/CLIMATE + FLUSH): replace env arrays
 * and recompute capacities and evaluator weights (QHG4/populations/OoAPop.cpp
 * :164-209, NPPCapacity.cpp:121-131) */
int qhg_gpu_set_env(qhg_pop *pop, const qhg_env_desc *env);

/* Environment events between steps (Simulator::runLoop's checkEvents ->
 * updateEvent -> flushEvents, QHG4/app/Simulator.cpp:335-374): the fields
 * the env descriptor holds (non-NULL) replace the current ones; with
 * QHG_EVENT_KILL_WET every agent in a cell with altitude < 0 or ice is
 * removed at once, the geography-event rule of OoAPop::updateEvent
 * (QHG4/populations/OoAPop.cpp:164-196), OoANavGen2bitPop::updateEvent
 * (QHG4/populations/OoANavGen2bitPop.cpp:177-210) and
 * tut_EnvironCapAltPop::updateEvent (QHG4/populations/tut_EnvironCapAltPop.cpp
 * :120-150); then capacities and evaluator weights are recomputed (the
 * flush, NPPCapacity.cpp:121-131, MultiEvaluator.cpp:142-210).  The removed
 * agents are not step deaths (stats count none).  Event bits follow
 * QHG4/utils_qhg/EventConsts.h (EVENT_ID_GEO 2, EVENT_ID_CLIMATE 3, EVENT_ID_VEG 4
 * map to the bits below); qhg_params_event_flags adds KILL_WET for the
 * classes whose updateEvent applies it. */
#define QHG_EVENT_GEO       0x01
#define QHG_EVENT_CLIMATE   0x02
#define QHG_EVENT_VEG       0x04
#define QHG_EVENT_KILL_WET  0x100
int qhg_gpu_env_event(qhg_pop *pop, const qhg_env_desc *env, int32_t events);

/* Sea-ways and bridges for Navigate: the grid's navigation group as
 * NavGroupReader::readData / readBridges load it into Navigation::setData /
 * setBridges (QHG4/io/NavGroupReader.cpp:96-205, QHG4/core/Navigation.cpp
 * :24-60).  Each port lists its destination cells with their distances;
 * a bridge joins two cells.  Setting them is EVENT_ID_NAV with its flush:
 * Navigate's jump tables are rebuilt at once (Navigate::recalculate,
 * QHG4/actions/Navigate.cpp:99-150), as are its current bridges (both
 * ends above sea level) after every geography event.  Ports and each
 * port's destinations may come in any order (the reference keeps them in
 * ordered maps); a port or a port's destination listed twice is an error,
 * as NavGroupReader's size check makes it.  A port whose destination
 * probabilities add up to 1 or more stops the table build there, as in the
 * reference: ports after it (in cell order) get no table.  The population
 * keeps its own copy of the arrays.  qhg_gpu_pre_loop fails when Navigate
 * is scheduled and no navigation was set (Navigate::preLoop,
 * QHG4/actions/Navigate.cpp:157-183). */
typedef struct qhg_nav_desc {
    int32_t        num_ports;
    const int32_t *ports;          /* num_ports cell indices                  */
    const int32_t *dest_offsets;   /* num_ports + 1: port k's destinations are
                                      entries dest_offsets[k] .. [k+1] - 1   */
    const int32_t *dest_cells;     /* destination cell indices                */
    const double  *dest_dists;     /* their distances                         */
    double         sample_dist;    /* Navigation::m_dSampleDist (carried)     */
    int32_t        num_bridges;    /* at most 131072                          */
    const int32_t *bridges;        /* num_bridges * 2 cell indices            */
} qhg_nav_desc;
int qhg_gpu_set_navigation(qhg_pop *pop, const qhg_nav_desc *nav);

/* observables */
int qhg_gpu_sync(qhg_pop *pop);
int qhg_gpu_stats(qhg_pop *pop, qhg_stats *out);
int qhg_gpu_cell_counts(qhg_pop *pop, uint32_t *out);        /* num_cells */
int qhg_gpu_capacities(qhg_pop *pop, double *out);           /* num_cells */
int qhg_gpu_env_weights(qhg_pop *pop, double *out);          /* num_cells*(conn+1) */
int qhg_gpu_compact(qhg_pop *pop);   /* LBController::compactData: here the genome pool */
/* Digest of the whole population state: out[0] = live agents, out[1..2] =
 * two order-independent sums of a 64-bit hash of every live agent's fields
 * (id, cell, life state, gender, birth time, age, last birth, stored partner
 * id, number of babies, genome words).  Two populations in the same state
 * have the same digest whatever order their buffers hold the agents in.  No
 * counterpart in the reference; it lets ensembles and servers compare full
 * states without downloading them. */
int qhg_gpu_state_digest(qhg_pop *pop, uint64_t out[3]);

/* profiling: average device time (ms) per launch of each kernel since the
 * last reset, measured with HIP events on the handle's stream. */
int qhg_gpu_set_timing(qhg_pop *pop, int32_t enable);
int qhg_gpu_kernel_times(qhg_pop *pop, double *avg_ms, int64_t *launches, int32_t max, int32_t *count);
const char *qhg_gpu_kernel_name(int32_t index);

/* Device memory held by the population (agent buffers at the current
 * capacity, genome pools, per-cell tables, scratch), in bytes.  A capacity
 * growth briefly holds the old and the new agent buffers.  No counterpart
 * in the reference; for callers that admit work against a memory budget. */
int qhg_gpu_device_bytes(qhg_pop *pop, int64_t *bytes);

/* Launch shapes of the step kernels.  group_agents: target agents per wave
 * group of occupied cells (1..4096); blocks_per_cu: act and scatter blocks
 * per compute unit (1..32).  Results never depend on them.  No counterpart
 * in the reference (OpenMP schedules its loops itself). */
typedef struct qhg_tuning {
    double group_agents;
    int32_t blocks_per_cu;
    int32_t pad;
} qhg_tuning;
int qhg_gpu_set_tuning(qhg_pop *pop, const qhg_tuning *t);
int qhg_gpu_get_tuning(qhg_pop *pop, qhg_tuning *t);
/* Autotuner: times candidate shapes on the population's own steps
 * (trial_steps each; the population advances by *steps_taken steps), keeps
 * the one with the most agent-steps per unit of device time and returns it. */
int qhg_gpu_autotune(qhg_pop *pop, int32_t trial_steps, qhg_tuning *chosen, int32_t *steps_taken);

/* The population's clock: the index of its next step, that step's time
 * and the highest agent id so far (new ids continue from it; the
 * reference's IDGen state, QHG4/core/IDGen.h:28).  set_clock restores a
 * clock read from a dump (SPopulation::restoreSpeciesDataQDF,
 * QHG4/core/SPopulation.cpp:2138-2207) before stepping resumes; max_id
 * never decreases below the uploaded agents' ids. */
int qhg_gpu_set_clock(qhg_pop *pop, int64_t step, double t, int64_t max_id);
int qhg_gpu_get_clock(qhg_pop *pop, int64_t *step, double *t, int64_t *max_id);
int qhg_gpu_get_params(qhg_pop *pop, qhg_params *out);

/* Species configuration from the reference's XML.  A species block names a
 * population class and lists a <module> per action (nested modules for a
 * MultiEvaluator's SingleEvaluators) with <param name value> entries, and a
 * <priorities> map.  qhg_params_from_modules applies the reference's rules
 * (SPopulation::readSpeciesData, QHG4/core/SPopulation.cpp:1108-1142;
 * Prioritizer::getActionAttributes, QHG4/core/Prioritizer.cpp:440-462;
 * Action::checkAttributes, QHG4/actions/Action.cpp:54-75): the class fixes
 * its actions and evaluator wiring ("tut_SexualPop", "tut_EnvironCapAltPop",
 * "OoAPop", "OoANavGen2bitPop", ...); every action's module must hold
 * exactly its attribute names; actions without a priority are not
 * scheduled; a priority for an action the class lacks is an error.  Module
 * names are those of the
 * Prioritizer: "Verhulst", "MultiEvaluator[Alt+NPP]", "SingleEvaluator[Alt]". */
typedef struct qhg_module_attrs {
    const char *name;
    int32_t num_params;
    const char *const *keys;
    const char *const *values;
} qhg_module_attrs;
int qhg_params_from_modules(const char *pop_class, const qhg_module_attrs *mods, int32_t nmods,
                            const char *const *prio_names, const int32_t *prio_values, int32_t nprios,
                            uint64_t seed, qhg_params *out);
/* The same from one flat name -> value map (the attributes of a species
 * group of a QDF file, Prioritizer::writeActionParamsQDF): each name goes to
 * the class action that owns it; names no action of the class owns are
 * ignored.  PolyLine values are strings "x0 v0 x1 v1 ...". */
int qhg_params_from_attribute_map(const char *pop_class, int32_t n, const char *const *keys,
                                  const char *const *values, const char *const *prio_names,
                                  const int32_t *prio_values, int32_t nprios, uint64_t seed, qhg_params *out);
/* Attribute names: of an action kind by index (NULL past the last), and of
 * a class's MultiEvaluator weights and SingleEvaluator PolyLines by
 * evaluator index (NULL when none). */
const char *qhg_params_attribute_name(int32_t kind, int32_t index);
const char *qhg_params_weight_name(const char *pop_class, int32_t eval);
const char *qhg_params_polyline_name(const char *pop_class, int32_t eval);
/* The reference's names of a class's actions, as its Prioritizer keys them
 * (Action::getActionName, QHG4/actions/Action.cpp:11-19, e.g. "Verhulst",
 * "MultiEvaluator[Alt+NPP]"): the index-th (NULL past the last), its kind in
 * *kind; and
 * of the MultiEvaluator's evaluators ("SingleEvaluator[Alt]").  The QDF
 * writer names the per-action groups of a species group after them. */
const char *qhg_params_class_action(const char *pop_class, int32_t index, int32_t *kind);
const char *qhg_params_evaluator_name(const char *pop_class, int32_t eval);
/* A numeric attribute by its reference name (e.g. "Verhulst_b0"). */
int qhg_params_get_attribute(const qhg_params *p, const char *name, double *value);
int qhg_params_set_attribute(qhg_params *p, const char *name, double value);
/* PopBase::modifyAttributes (QHG4/core/PopBase.h:75; e.g.
 * QHG4/actions/Verhulst.cpp:235-254) for attributes read every step; the
 * capacity and evaluator attributes are refused. */
int qhg_gpu_modify_attribute(qhg_pop *pop, const char *name, double value);
/* The event bits a class reacts to: events | QHG_EVENT_KILL_WET when events
 * holds QHG_EVENT_GEO and the class drowns agents on geography events
 * (tut_EnvironCapAltPop, OoAPop, OoANavGen2bitPop); events unchanged for
 * the others (SPopulation::updateEvent does nothing,
 * QHG4/core/SPopulation.h:116); negative for an unknown class. */
int qhg_params_event_flags(const char *pop_class, int32_t events);

/* Batched stepping.  No counterpart in the reference (QHG4 steps one
 * population per process, PopLooper.cpp:166-202); it serves callers that run
 * many independent scenarios on one GPU (ensembles, a scenario server).  A
 * batch step advances every member by one step (the same as its
 * initialize/doActions/finalize) with one launch per phase for all members;
 * each member keeps its own buffers, parameters, seed and step count, and
 * ends in exactly the state it would reach stepped alone.  While a member,
 * a population takes no step calls of its own (they return QHG_ERR_STATE);
 * its observables (stats, download, cell counts) stay available.  At most
 * 256 members; all on the batch's device.  Destroying a member removes it. */
typedef struct qhg_batch qhg_batch;
int qhg_batch_create(qhg_batch **out);
int qhg_batch_destroy(qhg_batch *batch);
int qhg_batch_add(qhg_batch *batch, qhg_pop *pop);
int qhg_batch_remove(qhg_batch *batch, qhg_pop *pop);
int qhg_batch_size(qhg_batch *batch, int32_t *n);
/* A member whose step cannot be prepared (its device reported an error,
 * its buffers could not grow) leaves the batch with its code in
 * qhg_stats.host_error; the other members complete the call's steps and the
 * call returns QHG_ERR_MEMBER (qhg_batch_autotune likewise). */
int qhg_batch_run(qhg_batch *batch, int32_t nsteps);
int qhg_batch_sync(qhg_batch *batch);
int qhg_batch_set_timing(qhg_batch *batch, int32_t enable);
int qhg_batch_kernel_times(qhg_batch *batch, double *avg_ms, int64_t *launches, int32_t max, int32_t *count);
const char *qhg_batch_kernel_name(int32_t index);
/* Replay each batch step from a captured HIP graph (recaptured when the
 * members or shapes change; not while timing is on). */
int qhg_batch_set_graph(qhg_batch *batch, int32_t enable);
int qhg_batch_set_tuning(qhg_batch *batch, const qhg_tuning *t);
int qhg_batch_get_tuning(qhg_batch *batch, qhg_tuning *t);
int qhg_batch_autotune(qhg_batch *batch, int32_t trial_steps, qhg_tuning *chosen, int32_t *steps_taken);

/* Capacities (NPPCapacity) and evaluator weights (MultiEvaluator) of a grid
 * from its environment arrays, as qhg_gpu_set_env recomputes them, with no
 * population and no allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Every pointer in
 * grid and env is DEVICE memory (neighbors needed with MultiEvaluator);
 * K_out (num_cells doubles) is written when params schedule NPPCapacity and
 * read as the capacity input of evaluators otherwise; W_out (num_cells *
 * (connectivity + 1) doubles) is written when they schedule MultiEvaluator.
 * The kernels' per-cell scratch is a fixed device array of
 * QHG_NOALLOC_MAX_CELLS doubles (8 MiB held by every GPU the library runs
 * on until the process ends); larger grids are refused with
 * QHG_ERR_CAPACITY.  The work is enqueued on `stream` (a hipStream_t, NULL
 * for the null stream) and the call returns without waiting.  Calls that
 * overlap on one GPU share its scratch array and race: serialise them (one
 * stream, or an event between them). */
#define QHG_NOALLOC_MAX_CELLS (1 << 20)
int qhg_env_tables_noalloc(const qhg_grid_desc *grid, const qhg_env_desc *env, const qhg_params *params,
                           double *K_out, double *W_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
