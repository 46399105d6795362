/*
 * qhg_qdf.h -- QDF (HDF5) files of populations and grids, for the MI355X
 * population step engine (libqhg4_qdf.so, host only, links libhdf5).
 *
 * Replaces the population dump / restore and grid reading of QHG4:
 *
 *   SPopulation::writeSpeciesDataQDF / writeAgentDataQDF / dumpAdditionalDataQDF
 *                              QHG4/core/SPopulation.cpp:1748-1778, 1465-1567, 2138-2207
 *   SPopulation::readSpeciesDataQDF / readAgentDataQDF / restoreAdditionalDataQDF
 *                              QHG4/core/SPopulation.cpp:1784-1830, 2138-2207
 *   createAgentDataTypeQDF     QHG4/core/SPopulation.cpp:1356-1372
 *   qdf_createFile, group and dataset names   QHG4/io/QDFUtils.h:13-131, QDFUtils.cpp:70-110
 *   grid / geography / climate / vegetation groups (GridReader, GeoReader, ...)
 *
 * Layout written (names of the reference; a reference reader selects the
 * compound members it knows by name):
 *   /                          QHG, Step, StartTime (string attributes)
 *   /Populations/<species>     ClassName, SpeciesName (strings), NumCells,
 *                              PrioInfo {FuncName[128], FuncPrio} per action,
 *                              every action attribute by its reference name
 *                              (doubles; PolyLines as {X, V, slopes}),
 *                              IDGenState (the last id handed out), and this
 *                              engine's PhiloxSeed, StepCount, NextTime,
 *                              MultiEvaluatorMode
 *     <action name>            one group per action of the class, named as its
 *                              Prioritizer keys it ("Verhulst",
 *                              "MultiEvaluator[Alt+NPP]"), holding that
 *                              action's attributes, as the reference's
 *                              Prioritizer::writeActionParamsQDF writes them
 *                              and its readers look for them
 *                              (QHG4/core/Prioritizer.cpp:217-246); a
 *                              MultiEvaluator group holds its weights and
 *                              one "SingleEvaluator[<name>]" group per
 *                              evaluator with its PolyLine
 *     AgentDataSet             {LifeState, CellIdx, CellID, AgentID, BirthTime,
 *                              Gender, Age, LastBirth, MateID, NumBabies}
 *                              (times FP64)
 *     Genome                   num_agents x genome_words uint64 [Genetics]
 *     DeadSpaces               int32 slots awaiting reuse (empty: the device
 *                              buffers hold no holes)
 *
 * Every function returns 0 or a negative QHG_ERR_* code; qhg_qdf_last_error()
 * gives the message.
 */
#ifndef QHG_QDF_H
#define QHG_QDF_H

#include "qhg_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qhg_qdf_info {
    int64_t  num_agents;
    int32_t  genome_words;    /* per agent (both strands); 0: no Genome dataset */
    int32_t  num_cells;
    int64_t  step;            /* index of the next step */
    double   time;            /* its time */
    int64_t  max_id;          /* highest agent id so far */
    uint64_t seed;
    int32_t  has_params;      /* ClassName and PrioInfo present: parameters can be rebuilt */
    int32_t  pad;
    char     class_name[64];
    char     species_name[64];
} qhg_qdf_info;

const char *qhg_qdf_last_error(void);

/* Write agents and parameters of a species (create != 0: new file, else the
 * file is opened and the species group replaced).  cell_ids may be NULL
 * (CellID = cell index). */
int qhg_qdf_write_agents(const char *path, int32_t create, const char *species, const char *pop_class,
                         const qhg_params *params, int32_t num_cells, const int64_t *cell_ids, int64_t step,
                         double time, int64_t max_id, const qhg_agents *agents, int32_t genome_words);
/* The same for a population on the device (its live agents, clock and
 * parameters). */
int qhg_qdf_write_population(qhg_pop *pop, const char *path, int32_t create, const char *species,
                             const char *pop_class, int32_t num_cells, const int64_t *cell_ids);

int qhg_qdf_read_info(const char *path, const char *species, qhg_qdf_info *info);
/* The parameter block, rebuilt from the class name, PrioInfo and the action
 * attributes (files this library or the reference wrote); the key of the
 * draws from PhiloxSeed, or `seed` when the file has none. */
int qhg_qdf_read_params(const char *path, const char *species, uint64_t seed, qhg_params *out);
/* Agents into caller arrays of info.num_agents entries (genome:
 * num_agents * genome_words); any array may be NULL except id. */
int qhg_qdf_read_agents(const char *path, const char *species, qhg_agents *a);
/* Upload a dump into a population created with the same grid and
 * parameters, and restore its clock: stepping continues exactly where the
 * dumped run stopped. */
int qhg_qdf_restore_population(const char *path, const char *species, qhg_pop *pop);

/* The same pieces on groups and datasets the caller opened (HDF5 hid_t
 * handles passed as int64_t), for a population inside the reference's
 * writer and reader, which open the species group and create the
 * AgentDataSet themselves (PopWriter::write, QHG4/io/PopWriter.cpp:57-130;
 * PopReader::read, QHG4/io/PopReader.h:34):
 *   qhg_qdf_agent_type           createAgentDataTypeQDF (SPopulation.cpp:1356-1372);
 *                                the caller closes the type
 *   qhg_qdf_write_species_group  writeSpeciesDataQDF / dumpSpeciesDataQDF
 *                                (SPopulation.cpp:1748-1778)
 *   qhg_qdf_read_species_group   readSpeciesDataQDF / restoreSpeciesDataQDF
 *                                (SPopulation.cpp:1784-1830); info may be NULL
 *                                (its step/time/max_id come from the group only)
 *   qhg_qdf_write_agent_dataset  writeAgentDataQDF / dumpAgentDataQDF: every row
 *                                of the dataset (its extent must equal a->n)
 *   qhg_qdf_read_agent_dataset   readAgentDataQDF / restoreAgentDataQDF
 *   qhg_qdf_write_genome / read_genome
 *                                writeAdditionalDataQDF / readAdditionalDataQDF of
 *                                Genetics (QHG4/actions/Genetics.cpp, Genome
 *                                dataset n x words; words < 0 on read: any) */
int64_t qhg_qdf_agent_type(void);
int qhg_qdf_write_species_group(int64_t group, const char *species, const char *pop_class, const qhg_params *params,
                                int32_t num_cells, int64_t step, double time, int64_t max_id);
int qhg_qdf_read_species_group(int64_t group, uint64_t seed, qhg_params *out, qhg_qdf_info *info);
int qhg_qdf_dataset_rows(int64_t dataset, int64_t *n);
int qhg_qdf_write_agent_dataset(int64_t dataset, const qhg_agents *agents, const int64_t *cell_ids);
int qhg_qdf_read_agent_dataset(int64_t dataset, qhg_agents *agents);
int qhg_qdf_write_genome(int64_t group, const uint64_t *genome, int64_t n, int32_t words);
int qhg_qdf_read_genome(int64_t group, uint64_t *genome, int64_t n, int32_t words);

/* Grids (the reference's Grid/CellDataSet {CellID, NumNeighbors,
 * Neighbors[max]}, Geography, Climate and Vegetation groups). */
typedef struct qhg_qdf_grid_info {
    int32_t num_cells;
    int32_t max_neighbors;
    int32_t has_geography, has_climate, has_vegetation;
    int32_t pad;
} qhg_qdf_grid_info;
int qhg_qdf_read_grid_info(const char *path, qhg_qdf_grid_info *info);
/* Arrays of num_cells (neighbors: num_cells * max_neighbors); NULL skips a
 * field; a field the file lacks is left untouched and flagged in *missing
 * (bit k for the k-th of lon, lat, alt, water, ice, coastal, base_npp,
 * temperature, rainfall). */
int qhg_qdf_read_grid(const char *path, int32_t *cell_ids, uint8_t *num_neighbors, int32_t *neighbors,
                      double *lon, double *lat, double *alt, double *water, uint8_t *ice, uint8_t *coastal,
                      double *base_npp, double *temperature, double *rainfall, int32_t *missing);

#ifdef __cplusplus
}
#endif
#endif
