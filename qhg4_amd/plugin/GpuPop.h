// GpuPop: a QHG4 population whose step runs on the MI355X engine.
//
// It is a PopBase (QHG4/core/PopBase.h:16-121), so the reference's driver
// loads it like any population plugin: DynPopFactory dlopens
// Gpu<Class>Wrapper.so, getInfo() names the class, createPop() builds it
// (QHG4/populations/DynPopFactory.cpp:80-160,
// QHG4/dynpops/WrapperTemplate.cpp.tmp), readSpeciesData() takes the
// species XML, and PopLooper steps it (QHG4/core/PopLooper.cpp:166-202).
// Every call is forwarded to qhg::Population (include/qhg_population.hpp)
// over the C ABI; QDF groups and datasets go through libqhg4_qdf.
//
// Per-agent host hooks (registerMove/Birth/Death, createAgentAtIndex,
// setAgentBasic) have no meaning here: the device decides births, deaths
// and moves inside the step; they report an error.  Agent accessors by
// index (getAgentID etc.) read a host snapshot of the live agents in id
// order, taken when first asked after a step.
#ifndef GPUPOP_H
#define GPUPOP_H

#include <memory>
#include <set>
#include <string>

#include "PopBase.h"
#include "qhg_population.hpp"

class PopFinder;
class IDGen;

class GpuPop : public PopBase {
public:
    GpuPop(const std::string &sClassName, SCellGrid *pCG, PopFinder *pPopFinder, int iLayerSize, IDGen **apIDG,
           uint32_t *aulState);
    virtual ~GpuPop();

    int setPrioList() override;
    uint getPrios(std::set<uint> &vPrios) override;
    int removeAction(std::string name) override;
    int disableAction(std::string name) override;
    int enableAction(std::string name) override;

    int preLoop() override;
    int postLoop() override;
    int doActions(uint iPrio, float fTime) override;
    int initializeStep(float fTime) override;
    int finalizeStep() override;
    int preWrite(float fTime) override;

    void registerMove(int iCellIndexFrom, int iAgentIndex, int iCellIndexTo) override;
    void registerBirth(int iCellIndex, int iMotherIndex, int iFatherIndex = -1) override;
    void registerDeath(int iCellIndex, int iAgentIndex) override;
    int reserveAgentSpace(int iNumAgents) override;

    void showAgents() override;
    ulong getNumAgents(int iCellIndex) override;
    ulong getNumAgentsTotal() override;
    ulong getNumAgentsEffective() override;
    ulong getNumAgentsMax() override;
    void updateTotal() override;
    void updateNumAgentsPerCell() override;

    int readSpeciesData(ParamProvider2 *pPP) override;
    const std::string getClassName() override;
    const std::string getSpeciesName() override;
    size_t agentDataSize() const override;
    int addAgent(int iCellIndex, char *pData, bool bUpdateCounts) override;
    int addAgentData(int iCellIndex, int iAgentIndex, char **ppData) override;
    int createAgentAtIndex(int iAgentIndex, int iCellIndex) override;
    void setAgentBasic(int iAgentIndex, void *pBasic) override;
    int getFreeIndex() override;
    size_t agentRealSizeQDF() const override;

    void setAgentDataType() override;
    hid_t createAgentDataTypeQDF() override;
    int writeSpeciesDataQDF(hid_t hSpeciesGroup) override;
    int readSpeciesDataQDF(hid_t hSpeciesGroup) override;
    int writeAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) override;
    int readAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) override;
    int writeAdditionalDataQDF(hid_t hSpeciesGroup) override;
    int readAdditionalDataQDF(hid_t hSpeciesGroup) override;

    int dumpSpeciesDataQDF(hid_t hSpeciesGroup, int iDumpMode) override;
    int restoreSpeciesDataQDF(hid_t hSpeciesGroup) override;
    int dumpAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) override;
    int restoreAgentDataQDF(hid_t hDataSpace, hid_t hDataSet, hid_t hAgentType) override;
    int dumpAdditionalDataQDF(hid_t hSpeciesGroup) override;
    int restoreAdditionalDataQDF(hid_t hSpeciesGroup) override;

    int mergePop(PopBase *pPop) override;
    int modifyAttributes(const std::string sAttrName, double dValue) override;

    idtype getMaxLoadedID() override;
    idtype getUID() override;
    void compactData() override;
    int checkLists() override;

    int getNumCells() override;
    void randomize(int i) override;

    int getFirstAgentIndex() override;
    int getLastAgentIndex() override;

    int setParams(const std::string sParams) override;
    int flushDeadSpace() override;

    int updateEvent(int EventID, char *pData, float fT) override;
    void flushEvents(float fT) override;

    uint getAgentLifeState(int iAgentIndex) override;
    int getAgentCellIndex(int iAgentIndex) override;
    idtype getAgentID(int iAgentIndex) override;
    gridtype getAgentCellID(int iAgentIndex) override;
    float getAgentBirthTime(int iAgentIndex) override;
    uchar getAgentGender(int iAgentIndex) override;

    hid_t getAgentQDFDataType() override;

    bool hasAction(const std::string sAction) override;
    bool hasParam(const std::string sParam) override;

    WELL512 **getWELL() override;

    int addPopSpecificAgentData(int iAgentIndex, char **ppData) override;

    qhg::Population *engine() { return m_pPop.get(); }

private:
    int report(int iResult, const char *pWhere);
    // the grid's Navigation (sea-ways and bridges) to the device population
    int passNavigation();
    const qhg::Agents &snapshot();
    int uploadLoaded();

    std::string m_sClassName;
    std::string m_sSpeciesName;
    SCellGrid *m_pCG;
    uint64_t m_ulSeed;
    // grid and environment in the engine's layout (the environment arrays
    // are the grid's own, read in place)
    std::vector<int32_t> m_vNeighbors;
    std::vector<uint8_t> m_vNumNeighbors;
    std::vector<int64_t> m_vCellIDs;
    qhg_grid_desc m_grid;
    qhg_env_desc m_env;
    std::unique_ptr<qhg::Population> m_pPop;
    // agents read from a QDF file, uploaded at preLoop
    qhg::Agents m_loaded;
    bool m_bRestored;
    int64_t m_iRestoreStep;
    double m_dRestoreTime;
    int64_t m_iRestoreMaxID;
    // host snapshot for the per-index accessors
    qhg::Agents m_snap;
    bool m_bSnapValid;
    hid_t m_hAgentDataType;
    int64_t m_iStaged;
};

#endif
