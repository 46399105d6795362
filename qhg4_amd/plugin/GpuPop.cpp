// GpuPop: PopBase over the MI355X population engine (see GpuPop.h).
#include "GpuPop.h"

#include <cstdio>
#include <cstring>
#include <functional>

#include "ParamProvider2.h"
#include "SCellGrid.h"
#include "qhg_qdf.h"

static_assert(sizeof(bool) == 1, "Geography's bool arrays are read as bytes");

GpuPop::GpuPop(const std::string &sClassName, SCellGrid *pCG, PopFinder *, int, IDGen **, uint32_t *aulState)
    : m_sClassName(sClassName),
      m_pCG(pCG),
      m_ulSeed(0),
      m_grid{},
      m_env{},
      m_bRestored(false),
      m_iRestoreStep(0),
      m_dRestoreTime(0),
      m_iRestoreMaxID(0),
      m_bSnapValid(false),
      m_hAgentDataType(H5P_DEFAULT),
      m_iStaged(0) {
    // the Philox key: the application's WELL seed state (SimParams's
    // m_aulState, 16 words); the seed array of createPop is not passed by
    // DynPopFactory (QHG4/populations/DynPopFactory.cpp:18) and is not read
    if (aulState) m_ulSeed = ((uint64_t)aulState[1] << 32) | aulState[0];
    const int nc = (int)pCG->m_iNumCells;
    const int conn = (pCG->m_iConnectivity > 0) ? pCG->m_iConnectivity : pCG->m_iMaxNeighbors;
    m_vNeighbors.resize((size_t)nc * conn);
    m_vNumNeighbors.resize((size_t)nc);
    m_vCellIDs.resize((size_t)nc);
    for (int c = 0; c < nc; c++) {
        const SCell &cell = pCG->m_aCells[c];
        // every direction slot as the grid holds it (lattices keep -1 holes
        // at fixed directions, QHG4/tools_io/GridFactory.cpp:979-1015)
        for (int j = 0; j < conn; j++) m_vNeighbors[(size_t)c * conn + j] = (j < MAX_NEIGH) ? cell.m_aNeighbors[j] : -1;
        m_vNumNeighbors[(size_t)c] = cell.m_iNumNeighbors;
        m_vCellIDs[(size_t)c] = cell.m_iGlobalID;
    }
    m_grid = qhg_grid_desc{nc, conn, m_vNeighbors.data(), m_vNumNeighbors.data(), m_vCellIDs.data()};
    if (Geography *g = pCG->m_pGeography) {
        m_env.altitude = g->m_adAltitude;
        m_env.longitude = g->m_adLongitude;
        m_env.latitude = g->m_adLatitude;
        m_env.water = g->m_adWater;
        m_env.ice = reinterpret_cast<const uint8_t *>(g->m_abIce);
        m_env.coastal = reinterpret_cast<const uint8_t *>(g->m_abCoastal);
    }
    if (Climate *c = pCG->m_pClimate) {
        m_env.temperature = c->m_adAnnualMeanTemp;
        m_env.rainfall = c->m_adAnnualRainfall;
    }
    if (Vegetation *v = pCG->m_pVegetation) m_env.base_npp = v->m_adBaseANPP;
}

// PopWriter closes the agent type it was handed (QHG4/io/PopWriter.cpp:26-31),
// so the handle may be gone by now: close it only while it is still valid
GpuPop::~GpuPop() {
    if (m_hAgentDataType != H5P_DEFAULT && H5Iis_valid(m_hAgentDataType) > 0) H5Tclose(m_hAgentDataType);
}

int GpuPop::report(int iResult, const char *pWhere) {
    if (iResult != 0)
        std::fprintf(stderr, "[GpuPop::%s] %s\n", pWhere,
                     m_pPop ? m_pPop->lastError().c_str() : "no device population (readSpeciesData first)");
    return iResult;
}

// -- species, priorities, actions ------------------------------------------

int GpuPop::readSpeciesData(ParamProvider2 *pPP) {
    m_sClassName = pPP->getSelected();
    m_sSpeciesName = pPP->getSpeciesName();
    const classinfo *ci = pPP->getClassInfo();
    if (!ci) return report(-1, "readSpeciesData");
    qhg::ModuleList mods;
    std::function<void(const modulemap &)> walk = [&](const modulemap &mm) {
        for (const auto &m : mm) {
            mods.push_back({m.first, m.second->getAttributes()});
            walk(m.second->getSubModules());
        }
    };
    walk(ci->mods);
    m_pPop.reset(new qhg::Population(m_grid, m_env, m_sClassName, m_sSpeciesName));
    return report(m_pPop->readSpeciesData(mods, ci->prios, m_ulSeed), "readSpeciesData");
}

int GpuPop::setPrioList() { return m_pPop ? m_pPop->setPrioList() : report(-1, "setPrioList"); }

uint GpuPop::getPrios(std::set<uint> &vPrios) {
    std::set<unsigned> s;
    const unsigned n = m_pPop ? m_pPop->getPrios(s) : 0;
    vPrios.insert(s.begin(), s.end());
    return n;
}

// the action set is compiled into the device program at readSpeciesData
int GpuPop::removeAction(std::string name) {
    std::fprintf(stderr, "[GpuPop::removeAction] %s: the device population's actions are fixed\n", name.c_str());
    return -1;
}
int GpuPop::disableAction(std::string name) { return removeAction(name); }
int GpuPop::enableAction(std::string name) { return hasAction(name) ? 0 : removeAction(name); }

bool GpuPop::hasAction(const std::string sAction) {
    if (!m_pPop) return false;
    const qhg_params &p = m_pPop->params();
    for (int i = 0; i < p.num_actions; i++)
        if (sAction == p.action_name[i]) return true;
    return false;
}

bool GpuPop::hasParam(const std::string sParam) {
    double v = 0;
    return m_pPop && qhg_params_get_attribute(&m_pPop->params(), sParam.c_str(), &v) == 0;
}

int GpuPop::modifyAttributes(const std::string sAttrName, double dValue) {
    return m_pPop ? report(m_pPop->modifyAttributes(sAttrName, dValue), "modifyAttributes")
                  : report(-1, "modifyAttributes");
}

int GpuPop::setParams(const std::string) { return 0; }

// -- the loop ----------------------------------------------------------------

int GpuPop::uploadLoaded() {
    if (m_loaded.size() == 0) return 0;
    int iResult = m_pPop->addAgents(m_loaded);
    m_loaded = qhg::Agents{};
    return report(iResult, "addAgents");
}

// Navigation::m_mDestinations (port -> destination -> distance, both maps
// in cell order) and m_vBridges (QHG4/core/Navigation.h:12-40) as the
// device population takes them; cell ids are used as cell indices, as
// Navigate does (QHG4/actions/Navigate.cpp:195-230)
int GpuPop::passNavigation() {
    const Navigation *nav = m_pCG->m_pNavigation;
    if (!nav) return 0;
    std::vector<int32_t> ports, off(1, 0), cells, bridges;
    std::vector<double> dists;
    for (const auto &p : nav->m_mDestinations) {
        ports.push_back(p.first);
        for (const auto &d : p.second) {
            cells.push_back(d.first);
            dists.push_back(d.second);
        }
        off.push_back((int32_t)cells.size());
    }
    for (const auto &b : nav->m_vBridges) {
        bridges.push_back(b.first);
        bridges.push_back(b.second);
    }
    const qhg_nav_desc d{(int32_t)ports.size(), ports.data(), off.data(), cells.data(), dists.data(),
                         nav->m_dSampleDist, (int32_t)nav->m_vBridges.size(), bridges.data()};
    return report(m_pPop->setNavigation(d), "setNavigation");
}

int GpuPop::preLoop() {
    if (!m_pPop) return report(-1, "preLoop");
    int iResult = uploadLoaded();
    if (iResult == 0) iResult = passNavigation();
    if (iResult == 0) iResult = report(m_pPop->preLoop(), "preLoop");
    if (iResult == 0 && m_bRestored) {
        iResult = report(m_pPop->setClock(m_iRestoreStep, m_dRestoreTime, m_iRestoreMaxID), "preLoop (clock)");
        m_bRestored = false;
    }
    m_bSnapValid = false;
    return iResult;
}

int GpuPop::postLoop() { return m_pPop ? report(m_pPop->postLoop(), "postLoop") : 0; }

int GpuPop::initializeStep(float fTime) {
    m_bSnapValid = false;
    return report(m_pPop ? m_pPop->initializeStep(fTime) : -1, "initializeStep");
}

int GpuPop::doActions(uint iPrio, float fTime) { return report(m_pPop->doActions(iPrio, fTime), "doActions"); }

int GpuPop::finalizeStep() { return report(m_pPop->finalizeStep(), "finalizeStep"); }

int GpuPop::preWrite(float fTime) { return m_pPop ? report(m_pPop->preWrite(fTime), "preWrite") : 0; }

void GpuPop::registerMove(int, int, int) {
    std::fprintf(stderr, "[GpuPop::registerMove] moves are decided on the device\n");
}
void GpuPop::registerBirth(int, int, int) {
    std::fprintf(stderr, "[GpuPop::registerBirth] births are decided on the device\n");
}
void GpuPop::registerDeath(int, int) {
    std::fprintf(stderr, "[GpuPop::registerDeath] deaths are decided on the device\n");
}

int GpuPop::reserveAgentSpace(int iNumAgents) {
    return m_pPop ? m_pPop->reserveAgentSpace(iNumAgents) : 0;
}

// -- counts --------------------------------------------------------------------

void GpuPop::showAgents() {
    const qhg::Agents &a = snapshot();
    for (size_t i = 0; i < a.size(); i++)
        std::printf("%zu: id %ld cell %d life %u gender %u birth %f\n", i, (long)a.id[i], a.cell[i], a.life[i],
                    a.gender[i], a.birth_time[i]);
}
ulong GpuPop::getNumAgents(int iCellIndex) { return m_pPop ? m_pPop->getNumAgents(iCellIndex) : 0; }
ulong GpuPop::getNumAgentsTotal() { return m_pPop ? m_pPop->getNumAgentsTotal() : 0; }
ulong GpuPop::getNumAgentsEffective() { return m_pPop ? m_pPop->getNumAgentsEffective() : 0; }
ulong GpuPop::getNumAgentsMax() { return m_pPop ? m_pPop->getNumAgentsMax() : 0; }
void GpuPop::updateTotal() {
    if (m_pPop) m_pPop->updateTotal();
}
void GpuPop::updateNumAgentsPerCell() {
    if (m_pPop) m_pPop->updateNumAgentsPerCell();
}
int GpuPop::getNumCells() { return (int)m_pCG->m_iNumCells; }

idtype GpuPop::getMaxLoadedID() { return m_pPop ? (idtype)m_pPop->getMaxLoadedID() : -1; }
// new ids come from the device (max id + 1 + rank of the birth)
idtype GpuPop::getUID() { return getMaxLoadedID() + 1; }
void GpuPop::compactData() {
    if (m_pPop) m_pPop->compactData();
}
int GpuPop::checkLists() { return 0; }
// the Philox streams are keyed by (seed, step, agent id): nothing to advance
void GpuPop::randomize(int) {}
int GpuPop::flushDeadSpace() { return m_pPop ? m_pPop->flushDeadSpace() : 0; }

WELL512 **GpuPop::getWELL() { return nullptr; }

// -- events ------------------------------------------------------------------

int GpuPop::updateEvent(int EventID, char *, float fT) { return m_pPop ? m_pPop->updateEvent(EventID, fT) : 0; }

void GpuPop::flushEvents(float fT) {
    if (!m_pPop) return;
    // the simulator may have replaced the grid's environment objects
    qhg_env_desc env = m_env;
    if (Geography *g = m_pCG->m_pGeography) {
        env.altitude = g->m_adAltitude;
        env.ice = reinterpret_cast<const uint8_t *>(g->m_abIce);
        env.water = g->m_adWater;
        env.coastal = reinterpret_cast<const uint8_t *>(g->m_abCoastal);
    }
    if (Climate *c = m_pCG->m_pClimate) {
        env.temperature = c->m_adAnnualMeanTemp;
        env.rainfall = c->m_adAnnualRainfall;
    }
    if (Vegetation *v = m_pCG->m_pVegetation) env.base_npp = v->m_adBaseANPP;
    m_env = env;
    m_pPop->setEnvironment(env);
    if (m_pPop->takeNavEvent()) passNavigation();
    report(m_pPop->flushEvents(fT), "flushEvents");
    m_bSnapValid = false;
}

// -- agents in and out -------------------------------------------------------

size_t GpuPop::agentDataSize() const { return agentRealSizeQDF(); }

int GpuPop::addAgent(int iCellIndex, char *pData, bool) {
    if (!m_pPop) return report(-1, "addAgent");
    m_iStaged++;
    m_bSnapValid = false;
    return report(m_pPop->addAgent(iCellIndex, pData), "addAgent");
}

int GpuPop::addAgentData(int iCellIndex, int, char **ppData) {
    return (ppData && *ppData) ? addAgent(iCellIndex, *ppData, false) : -1;
}

int GpuPop::addPopSpecificAgentData(int, char **) { return 0; }   // Age, LastBirth: read by addAgent

int GpuPop::createAgentAtIndex(int, int) {
    std::fprintf(stderr, "[GpuPop::createAgentAtIndex] agents are created on the device\n");
    return -1;
}
void GpuPop::setAgentBasic(int, void *) {
    std::fprintf(stderr, "[GpuPop::setAgentBasic] agents live on the device\n");
}
int GpuPop::getFreeIndex() { return (int)(getNumAgentsTotal() + m_iStaged); }

const qhg::Agents &GpuPop::snapshot() {
    if (!m_bSnapValid && m_pPop) {
        m_snap = qhg::Agents{};
        if (m_pPop->download(m_snap) == 0) m_bSnapValid = true;
    }
    return m_snap;
}

int GpuPop::getFirstAgentIndex() { return snapshot().size() ? 0 : -1; }
int GpuPop::getLastAgentIndex() { return (int)snapshot().size() - 1; }
uint GpuPop::getAgentLifeState(int i) { return snapshot().life.at((size_t)i); }
int GpuPop::getAgentCellIndex(int i) { return snapshot().cell.at((size_t)i); }
idtype GpuPop::getAgentID(int i) { return (idtype)snapshot().id.at((size_t)i); }
gridtype GpuPop::getAgentCellID(int i) { return m_pCG->m_aCells[snapshot().cell.at((size_t)i)].m_iGlobalID; }
float GpuPop::getAgentBirthTime(int i) { return (float)snapshot().birth_time.at((size_t)i); }
uchar GpuPop::getAgentGender(int i) { return snapshot().gender.at((size_t)i); }

int GpuPop::mergePop(PopBase *pOther) {
    GpuPop *p = dynamic_cast<GpuPop *>(pOther);
    if (!p || !m_pPop || !p->m_pPop || p->m_sClassName != m_sClassName || p->m_sSpeciesName != m_sSpeciesName) {
        std::fprintf(stderr, "[GpuPop::mergePop] class or species differ: not compatible\n");
        return -1;
    }
    qhg::Agents a;
    int iResult = report(p->m_pPop->download(a), "mergePop");
    if (iResult == 0) iResult = report(m_pPop->addAgents(a), "mergePop");
    m_bSnapValid = false;
    return iResult;
}

// -- QDF -----------------------------------------------------------------------

hid_t GpuPop::createAgentDataTypeQDF() { return (hid_t)qhg_qdf_agent_type(); }

void GpuPop::setAgentDataType() {
    if (m_hAgentDataType != H5P_DEFAULT && H5Iis_valid(m_hAgentDataType) > 0) H5Tclose(m_hAgentDataType);
    m_hAgentDataType = createAgentDataTypeQDF();
}

// a fresh type when none was made yet or a writer has closed the last one
hid_t GpuPop::getAgentQDFDataType() {
    if (m_hAgentDataType == H5P_DEFAULT || H5Iis_valid(m_hAgentDataType) <= 0) setAgentDataType();
    return m_hAgentDataType;
}

size_t GpuPop::agentRealSizeQDF() const {
    const hid_t t = (hid_t)qhg_qdf_agent_type();
    const size_t s = H5Tget_size(t);
    H5Tclose(t);
    return s;
}

int GpuPop::writeSpeciesDataQDF(hid_t hSpeciesGroup) {
    if (!m_pPop) return report(-1, "writeSpeciesDataQDF");
    int64_t step = 0, max_id = 0;
    double t = 0;
    if (qhg_gpu_get_clock(m_pPop->handle(), &step, &t, &max_id)) return report(-1, "writeSpeciesDataQDF");
    const int rc = qhg_qdf_write_species_group((int64_t)hSpeciesGroup, m_sSpeciesName.c_str(), m_sClassName.c_str(),
                                               &m_pPop->params(), m_grid.num_cells, step, t, max_id);
    if (rc) std::fprintf(stderr, "[GpuPop::writeSpeciesDataQDF] %s\n", qhg_qdf_last_error());
    return rc;
}

int GpuPop::readSpeciesDataQDF(hid_t hSpeciesGroup) {
    qhg_params p;
    qhg_qdf_info info;
    const int rc = qhg_qdf_read_species_group((int64_t)hSpeciesGroup, m_ulSeed, &p, &info);
    if (rc) {
        std::fprintf(stderr, "[GpuPop::readSpeciesDataQDF] %s\n", qhg_qdf_last_error());
        return rc;
    }
    if (info.species_name[0]) m_sSpeciesName = info.species_name;
    if (!m_pPop) {
        m_pPop.reset(new qhg::Population(m_grid, m_env, m_sClassName, m_sSpeciesName));
        if (int e = m_pPop->setParams(p)) return report(e, "readSpeciesDataQDF");
    }
    return 0;
}

int GpuPop::writeAgentDataQDF(hid_t, hid_t hDataSet, hid_t) {
    if (!m_pPop) return report(-1, "writeAgentDataQDF");
    qhg::Agents a;
    if (int rc = m_pPop->download(a)) return report(rc, "writeAgentDataQDF");
    qhg_agents v = a.view();
    const int rc = qhg_qdf_write_agent_dataset((int64_t)hDataSet, &v, m_vCellIDs.data());
    if (rc) std::fprintf(stderr, "[GpuPop::writeAgentDataQDF] %s\n", qhg_qdf_last_error());
    return rc;
}

int GpuPop::readAgentDataQDF(hid_t, hid_t hDataSet, hid_t) {
    int64_t n = 0;
    if (qhg_qdf_dataset_rows((int64_t)hDataSet, &n)) return -1;
    m_loaded = qhg::Agents{};
    m_loaded.resize((size_t)n, m_pPop ? m_pPop->genomeWords() : 0);
    qhg_agents v = m_loaded.view();
    v.genome = nullptr;   // the Genome dataset comes with the additional data
    const int rc = qhg_qdf_read_agent_dataset((int64_t)hDataSet, &v);
    if (rc) std::fprintf(stderr, "[GpuPop::readAgentDataQDF] %s\n", qhg_qdf_last_error());
    return rc;
}

int GpuPop::writeAdditionalDataQDF(hid_t hSpeciesGroup) {
    if (!m_pPop || m_pPop->genomeWords() == 0) return 0;
    qhg::Agents a;
    if (int rc = m_pPop->download(a)) return report(rc, "writeAdditionalDataQDF");
    const int rc = qhg_qdf_write_genome((int64_t)hSpeciesGroup, a.genome.data(), (int64_t)a.size(), a.genome_words);
    if (rc) std::fprintf(stderr, "[GpuPop::writeAdditionalDataQDF] %s\n", qhg_qdf_last_error());
    return rc;
}

int GpuPop::readAdditionalDataQDF(hid_t hSpeciesGroup) {
    if (!m_pPop || m_pPop->genomeWords() == 0 || m_loaded.size() == 0) return 0;
    const int rc = qhg_qdf_read_genome((int64_t)hSpeciesGroup, m_loaded.genome.data(), (int64_t)m_loaded.size(),
                                       m_pPop->genomeWords());
    if (rc) std::fprintf(stderr, "[GpuPop::readAdditionalDataQDF] %s\n", qhg_qdf_last_error());
    return rc;
}

// a dump is a write with this engine's state (clock, max id, seed and the
// parameter block are written with every species group)
int GpuPop::dumpSpeciesDataQDF(hid_t hSpeciesGroup, int) { return writeSpeciesDataQDF(hSpeciesGroup); }
int GpuPop::dumpAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) {
    return writeAgentDataQDF(hDataSpace, hDataSet, hAgentType);
}
int GpuPop::dumpAdditionalDataQDF(hid_t hSpeciesGroup) { return writeAdditionalDataQDF(hSpeciesGroup); }

int GpuPop::restoreSpeciesDataQDF(hid_t hSpeciesGroup) {
    qhg_qdf_info info;
    int rc = readSpeciesDataQDF(hSpeciesGroup);
    if (rc == 0) rc = qhg_qdf_read_species_group((int64_t)hSpeciesGroup, m_ulSeed, nullptr, &info);
    if (rc == 0) {
        m_bRestored = true;
        m_iRestoreStep = info.step;
        m_dRestoreTime = info.time;
        m_iRestoreMaxID = info.max_id;
    }
    return rc;
}
int GpuPop::restoreAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) {
    return readAgentDataQDF(hDataSpace, hDataSet, hAgentType);
}
int GpuPop::restoreAdditionalDataQDF(hid_t hSpeciesGroup) { return readAdditionalDataQDF(hSpeciesGroup); }

const std::string GpuPop::getClassName() { return m_sClassName; }
const std::string GpuPop::getSpeciesName() { return m_sSpeciesName; }
