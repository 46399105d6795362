// QDF (HDF5) files of populations and grids (host only; libqhg4_qdf.so).
//
// The layout follows the reference's names (QHG4/io/QDFUtils.h:13-131;
// QHG4/core/SPopulation.cpp:1356-1372 for the agent compound), so either
// side can read the other's population groups by member and attribute name.
// Times are written as FP64 (this engine's agent time type); a reader with
// FP32 members gets them converted by HDF5.
#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qhg_gpu.h"
#include "../../include/qhg_qdf.h"

namespace {

thread_local std::string g_err;

int qfail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// closes an HDF5 handle when it leaves scope
struct Hid {
    hid_t id;
    herr_t (*close)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0) close(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    operator hid_t() const { return id; }
    bool ok() const { return id >= 0; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// agent record in memory, one per AgentDataSet row
struct QAgent {
    int32_t life, cell_idx, cell_id, pad0;
    int64_t agent_id;
    double birth;
    uint8_t gender, pad1[7];
    double age, lastb;
    int64_t mate;
    int32_t nbabies, pad2;
};

struct PrioRec {
    char name[128];
    int32_t prio;
};

int put_str(hid_t loc, const char *name, const std::string &v) {
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t, v.size() + 1);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    hsize_t d = 1;
    Hid sp(H5Screate_simple(1, &d, nullptr), H5Sclose);
    if (H5Aexists(loc, name) > 0) H5Adelete(loc, name);
    Hid a(H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Awrite(a, t, v.c_str()) < 0) return qfail(QHG_ERR_ARG, std::string("writing attribute ") + name);
    return 0;
}

int put_num(hid_t loc, const char *name, hid_t type, const void *v, hsize_t n) {
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (H5Aexists(loc, name) > 0) H5Adelete(loc, name);
    Hid a(H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Awrite(a, type, v) < 0) return qfail(QHG_ERR_ARG, std::string("writing attribute ") + name);
    return 0;
}

// PolyLine as qdf_writePolyLine does (QHG4/io/QDFUtils.cpp:859-900): X, V, slopes
int put_polyline(hid_t loc, const char *name, int n, const double *x, const double *v) {
    std::vector<double> xs(x, x + n), vs(v, v + n), as(n > 1 ? n - 1 : 0);
    for (int k = 0; k + 1 < n; k++) as[k] = (vs[k + 1] - vs[k]) / (xs[k + 1] - xs[k]);
    hvl_t d[3];
    d[0].len = xs.size();
    d[0].p = xs.data();
    d[1].len = vs.size();
    d[1].p = vs.data();
    d[2].len = as.size();
    d[2].p = as.data();
    Hid t(H5Tvlen_create(H5T_NATIVE_DOUBLE), H5Tclose);
    return put_num(loc, name, t, d, 3);
}

bool get_str(hid_t loc, const char *name, std::string &out) {
    if (H5Aexists(loc, name) <= 0) return false;
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    Hid ft(H5Aget_type(a), H5Tclose);
    if (H5Tget_class(ft) != H5T_STRING || H5Tis_variable_str(ft) > 0) return false;
    const size_t sz = H5Tget_size(ft);
    std::vector<char> buf(sz + 1, 0);
    Hid mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt, sz);
    if (H5Aread(a, mt, buf.data()) < 0) return false;
    out.assign(buf.data());
    return true;
}

bool get_num(hid_t loc, const char *name, hid_t mtype, void *out, hsize_t n) {
    if (H5Aexists(loc, name) <= 0) return false;
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    Hid sp(H5Aget_space(a), H5Sclose);
    if ((hsize_t)H5Sget_simple_extent_npoints(sp) != n) return false;
    return H5Aread(a, mtype, out) >= 0;
}

hid_t agent_type(hid_t file_type = -1) {
    // the members the file has (all of them when writing)
    hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(QAgent));
    struct M {
        const char *name;
        size_t off;
        hid_t type;
    } ms[] = {{"LifeState", HOFFSET(QAgent, life), H5T_NATIVE_INT32},
              {"CellIdx", HOFFSET(QAgent, cell_idx), H5T_NATIVE_INT32},
              {"CellID", HOFFSET(QAgent, cell_id), H5T_NATIVE_INT32},
              {"AgentID", HOFFSET(QAgent, agent_id), H5T_NATIVE_INT64},
              {"BirthTime", HOFFSET(QAgent, birth), H5T_NATIVE_DOUBLE},
              {"Gender", HOFFSET(QAgent, gender), H5T_NATIVE_UINT8},
              {"Age", HOFFSET(QAgent, age), H5T_NATIVE_DOUBLE},
              {"LastBirth", HOFFSET(QAgent, lastb), H5T_NATIVE_DOUBLE},
              {"MateID", HOFFSET(QAgent, mate), H5T_NATIVE_INT64},
              {"NumBabies", HOFFSET(QAgent, nbabies), H5T_NATIVE_INT32}};
    for (const M &m : ms)
        if (file_type < 0 || H5Tget_member_index(file_type, m.name) >= 0) H5Tinsert(t, m.name, m.off, m.type);
    return t;
}

bool has_kind(const qhg_params *p, int kind) {
    for (int i = 0; i < p->num_actions; i++)
        if (p->action_kind[i] == kind) return true;
    return false;
}

// the numeric attributes of one action kind, by their reference names
int put_kind_attrs(hid_t loc, const qhg_params *P, int kind) {
    for (int i = 0;; i++) {
        const char *n = qhg_params_attribute_name(kind, i);
        if (!n) return 0;
        double v = 0;
        if (qhg_params_get_attribute(P, n, &v) != 0) {
            // attributes the reference reads that steer only how genomes
            // are made at the start (Genetics.cpp): in a dump the genomes
            // are in the file, 2 bits per nucleotide
            if (std::strcmp(n, "Genetics_initial_muts") == 0 || std::strcmp(n, "Genetics_create_new_genome") == 0)
                v = 0;
            else if (std::strcmp(n, "Genetics_bits_per_nuc") == 0)
                v = 2;
            else
                continue;   // non-numeric attribute
        }
        int rc = put_num(loc, n, H5T_NATIVE_DOUBLE, &v, 1);
        if (rc) return rc;
    }
}

// The reference's current layout (Prioritizer::writeActionParamsQDF,
// QHG4/core/Prioritizer.cpp:340-375): one group per action of the class,
// named as its Prioritizer names it, holding the action's attributes; the
// MultiEvaluator's group holds its weights and one group per evaluator with
// the evaluator's PolyLine (MultiEvaluator::writeAttributesQDF,
// QHG4/actions/MultiEvaluator.cpp:515-540).
int put_action_groups(hid_t g, const char *pop_class, const qhg_params *P) {
    for (int i = 0;; i++) {
        int32_t kind = 0;
        const char *an = qhg_params_class_action(pop_class, i, &kind);
        if (!an) return 0;
        if (kind == 0) continue;
        Hid ag(H5Gcreate2(g, an, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!ag.ok()) return qfail(QHG_ERR_ARG, std::string("cannot create action group ") + an);
        int rc = put_kind_attrs(ag, P, kind);
        if (rc) return rc;
        if (kind != QHG_ACT_MULTIEVAL) continue;
        for (int e = 0; e < P->me_num; e++) {
            const char *w = qhg_params_weight_name(pop_class, e);
            if (w && (rc = put_num(ag, w, H5T_NATIVE_DOUBLE, &P->me_weight[e], 1))) return rc;
            const char *en = qhg_params_evaluator_name(pop_class, e);
            if (!en) continue;
            Hid eg(H5Gcreate2(ag, en, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
            if (!eg.ok()) return qfail(QHG_ERR_ARG, std::string("cannot create evaluator group ") + en);
            const char *pl = qhg_params_polyline_name(pop_class, e);
            if (pl && P->me_pl_n[e] > 0 && (rc = put_polyline(eg, pl, P->me_pl_n[e], P->me_pl_x[e], P->me_pl_v[e])))
                return rc;
        }
    }
}

int write_species_attrs(hid_t g, const char *species, const char *pop_class, const qhg_params *P, int32_t num_cells,
                        int64_t step, double time, int64_t max_id) {
    int rc = put_str(g, "ClassName", pop_class ? pop_class : "");
    if (!rc) rc = put_str(g, "SpeciesName", species);
    if (!rc) rc = put_num(g, "NumCells", H5T_NATIVE_INT32, &num_cells, 1);
    if (rc) return rc;
    // PrioInfo: {FuncName, FuncPrio} per scheduled action (insertPrioDataAttribute)
    {
        std::vector<PrioRec> pr((size_t)std::max(1, P->num_actions));
        for (int i = 0; i < P->num_actions; i++) {
            std::memset(pr[i].name, 0, sizeof(pr[i].name));
            std::strncpy(pr[i].name, P->action_name[i], sizeof(pr[i].name) - 1);
            pr[i].prio = P->action_prio[i];
        }
        Hid st(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(st, 128);
        Hid ct(H5Tcreate(H5T_COMPOUND, sizeof(PrioRec)), H5Tclose);
        H5Tinsert(ct, "FuncName", HOFFSET(PrioRec, name), st);
        H5Tinsert(ct, "FuncPrio", HOFFSET(PrioRec, prio), H5T_NATIVE_INT32);
        rc = put_num(g, "PrioInfo", ct, pr.data(), (hsize_t)P->num_actions);
        if (rc) return rc;
    }
    // the action attributes by their reference names (writeActionParamsQDF),
    // for every action kind, scheduled or not: a class builds all of its
    // actions and each reads its attributes (Prioritizer::getActionAttributes,
    // QHG4/core/Prioritizer.cpp:440-462); a reader ignores the names its
    // class does not own
    for (int kind = QHG_ACT_GETOLD; kind <= QHG_ACT_LAST && !rc; kind++) rc = put_kind_attrs(g, P, kind);
    if (rc) return rc;
    if (has_kind(P, QHG_ACT_MULTIEVAL) && pop_class) {
        for (int e = 0; e < P->me_num; e++) {
            const char *w = qhg_params_weight_name(pop_class, e);
            if (w && (rc = put_num(g, w, H5T_NATIVE_DOUBLE, &P->me_weight[e], 1))) return rc;
            const char *pl = qhg_params_polyline_name(pop_class, e);
            if (pl && P->me_pl_n[e] > 0 && (rc = put_polyline(g, pl, P->me_pl_n[e], P->me_pl_x[e], P->me_pl_v[e])))
                return rc;
        }
    }
    // ... and the same per action group (the layout the reference reads
    // now; the flat attributes above serve readers of the older layout)
    if (pop_class && *pop_class && (rc = put_action_groups(g, pop_class, P))) return rc;
    // the id generator's state under the reference's name (the last id
    // handed out; the reference writes one entry per thread,
    // SPopulation::dumpSpeciesDataQDF, QHG4/core/SPopulation.cpp:2184-2191)
    if (!rc) rc = put_num(g, "IDGenState", H5T_NATIVE_INT64, &max_id, 1);
    // this engine's own state: the draws' key and the clock, and the
    // MultiEvaluator's combination mode (the reference classes fix it in
    // their constructors, e.g. MODE_ADD_SIMPLE at QHG4/populations/OoAPop.cpp:56;
    // this engine takes it from the parameters)
    if (!rc && has_kind(P, QHG_ACT_MULTIEVAL)) rc = put_num(g, "MultiEvaluatorMode", H5T_NATIVE_INT32, &P->me_mode, 1);
    if (!rc) rc = put_num(g, "PhiloxSeed", H5T_NATIVE_UINT64, &P->seed, 1);
    if (!rc) rc = put_num(g, "StepCount", H5T_NATIVE_INT64, &step, 1);
    if (!rc) rc = put_num(g, "NextTime", H5T_NATIVE_DOUBLE, &time, 1);
    // DeadSpaces (SPopulation::dumpDeadSpaces, SPopulation.cpp:2456-2505):
    // the slots the previous step's deaths left for recycling.  The device
    // buffers hold no holes, so the list is empty.
    if (!rc) {
        const hsize_t zero = 0;
        Hid sp(H5Screate_simple(1, &zero, nullptr), H5Sclose);
        Hid d(H5Dcreate2(g, "DeadSpaces", H5T_NATIVE_INT32, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!d.ok()) rc = qfail(QHG_ERR_ARG, "cannot create DeadSpaces");
    }
    return rc;
}

hid_t open_species(hid_t f, const char *species) {
    const std::string path = std::string("/Populations/") + species;
    if (H5Lexists(f, "/Populations", H5P_DEFAULT) <= 0 || H5Lexists(f, path.c_str(), H5P_DEFAULT) <= 0) return -1;
    return H5Gopen2(f, path.c_str(), H5P_DEFAULT);
}

// collects the numeric scalar and PolyLine attributes of a species group as strings
herr_t collect_attr(hid_t loc, const char *name, const H5A_info_t *, void *op) {
    auto *out = static_cast<std::map<std::string, std::string> *>(op);
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    Hid ft(H5Aget_type(a), H5Tclose);
    Hid sp(H5Aget_space(a), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    const H5T_class_t cls = H5Tget_class(ft);
    char buf[64];
    if ((cls == H5T_FLOAT || cls == H5T_INTEGER) && n == 1) {
        double v = 0;
        if (H5Aread(a, H5T_NATIVE_DOUBLE, &v) >= 0) {
            std::snprintf(buf, sizeof(buf), "%.17g", v);
            (*out)[name] = buf;
        }
    } else if (cls == H5T_VLEN && n == 3) {
        Hid mt(H5Tvlen_create(H5T_NATIVE_DOUBLE), H5Tclose);
        hvl_t d[3];
        if (H5Aread(a, mt, d) >= 0) {
            std::string s;
            const double *x = (const double *)d[0].p, *v = (const double *)d[1].p;
            for (size_t k = 0; k < d[0].len && k < d[1].len; k++) {
                std::snprintf(buf, sizeof(buf), "%s%.17g %.17g", k ? " " : "", x[k], v[k]);
                s += buf;
            }
            (*out)[name] = s;
            H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, d);
        }
    }
    return 0;
}

// AgentDataSet rows from agent arrays (all rows of the dataset at once)
int write_records(hid_t ds, const qhg_agents *agents, const int64_t *cell_ids) {
    const int64_t n = agents->n;
    Hid sp(H5Dget_space(ds), H5Sclose);
    if (!sp.ok() || (int64_t)H5Sget_simple_extent_npoints(sp) != n)
        return qfail(QHG_ERR_ARG, "AgentDataSet rows differ from the agent count");
    std::vector<QAgent> rec((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        QAgent &r = rec[(size_t)i];
        std::memset(&r, 0, sizeof(r));
        r.life = agents->life ? agents->life[i] : QHG_LIFE_ALIVE;
        r.cell_idx = agents->cell ? agents->cell[i] : 0;
        r.cell_id = (cell_ids && agents->cell) ? (int32_t)cell_ids[agents->cell[i]] : r.cell_idx;
        r.agent_id = agents->id[i];
        r.birth = agents->birth_time ? agents->birth_time[i] : 0.0;
        r.gender = agents->gender ? agents->gender[i] : 0;
        r.age = agents->age ? agents->age[i] : 0.0;
        r.lastb = agents->last_birth ? agents->last_birth[i] : -1.0;
        r.mate = agents->mate_id ? agents->mate_id[i] : -3;
        r.nbabies = agents->num_babies ? agents->num_babies[i] : 0;
    }
    Hid t(agent_type(), H5Tclose);
    if (n > 0 && H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, rec.data()) < 0)
        return qfail(QHG_ERR_ARG, "writing AgentDataSet");
    return 0;
}

// agent arrays from AgentDataSet rows: the members the file has, defaults
// for the others
int read_records(hid_t ds, qhg_agents *a) {
    Hid sp(H5Dget_space(ds), H5Sclose);
    if (!sp.ok()) return qfail(QHG_ERR_ARG, "no dataspace");
    const int64_t n = (int64_t)H5Sget_simple_extent_npoints(sp);
    if (n != a->n) return qfail(QHG_ERR_ARG, "agent count differs from the arrays given");
    Hid ft(H5Dget_type(ds), H5Tclose);
    for (const char *need : {"LifeState", "CellIdx", "AgentID", "BirthTime", "Gender"})
        if (H5Tget_member_index(ft, need) < 0) return qfail(QHG_ERR_ARG, std::string("AgentDataSet lacks ") + need);
    std::vector<QAgent> rec((size_t)n);
    for (QAgent &r : rec) {
        std::memset(&r, 0, sizeof(r));
        r.lastb = -1.0;
        r.mate = -3;
    }
    const bool has_age = H5Tget_member_index(ft, "Age") >= 0;
    Hid mt(agent_type(ft), H5Tclose);
    if (n > 0 && H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, rec.data()) < 0)
        return qfail(QHG_ERR_ARG, "reading AgentDataSet");
    for (int64_t i = 0; i < n; i++) {
        const QAgent &r = rec[(size_t)i];
        a->id[i] = r.agent_id;
        if (a->cell) a->cell[i] = r.cell_idx;
        if (a->life) a->life[i] = (uint8_t)r.life;
        if (a->gender) a->gender[i] = r.gender;
        if (a->birth_time) a->birth_time[i] = r.birth;
        if (a->age) a->age[i] = has_age ? r.age : 0.0;
        if (a->last_birth) a->last_birth[i] = r.lastb;
        if (a->mate_id) a->mate_id[i] = r.mate;
        if (a->num_babies) a->num_babies[i] = r.nbabies;
    }
    return 0;
}

int write_genome(hid_t g, const uint64_t *genome, int64_t n, int32_t words) {
    if (H5Lexists(g, "Genome", H5P_DEFAULT) > 0) H5Ldelete(g, "Genome", H5P_DEFAULT);
    hsize_t d[2] = {(hsize_t)n, (hsize_t)words};
    Hid sp(H5Screate_simple(2, d, nullptr), H5Sclose);
    Hid ds(H5Dcreate2(g, "Genome", H5T_NATIVE_UINT64, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok() || (n > 0 && H5Dwrite(ds, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, genome) < 0))
        return qfail(QHG_ERR_ARG, "writing Genome");
    return 0;
}

int read_genome(hid_t g, uint64_t *genome, int64_t n, int32_t words) {
    if (H5Lexists(g, "Genome", H5P_DEFAULT) <= 0) return qfail(QHG_ERR_ARG, "no Genome dataset");
    Hid gd(H5Dopen2(g, "Genome", H5P_DEFAULT), H5Dclose);
    Hid gs(H5Dget_space(gd), H5Sclose);
    hsize_t d[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(gs) != 2) return qfail(QHG_ERR_ARG, "Genome is not two-dimensional");
    H5Sget_simple_extent_dims(gs, d, nullptr);
    if ((int64_t)d[0] != n) return qfail(QHG_ERR_ARG, "Genome rows differ from the agent count");
    if (words >= 0 && (int64_t)d[1] != words) return qfail(QHG_ERR_ARG, "Genome words differ from the population's");
    if (n > 0 && H5Dread(gd, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, genome) < 0)
        return qfail(QHG_ERR_ARG, "reading Genome");
    return 0;
}

// the attributes of every subgroup of `loc`, down to `depth` levels
void collect_groups(hid_t loc, std::map<std::string, std::string> &kv, int depth) {
    if (depth <= 0) return;
    H5G_info_t gi;
    if (H5Gget_info(loc, &gi) < 0) return;
    for (hsize_t k = 0; k < gi.nlinks; k++) {
        char name[256];
        if (H5Lget_name_by_idx(loc, ".", H5_INDEX_NAME, H5_ITER_INC, k, name, sizeof(name), H5P_DEFAULT) < 0) continue;
        H5O_info_t oi;
        if (H5Oget_info_by_name(loc, name, &oi, H5P_DEFAULT) < 0 || oi.type != H5O_TYPE_GROUP) continue;
        Hid sub(H5Gopen2(loc, name, H5P_DEFAULT), H5Gclose);
        if (!sub.ok()) continue;
        hsize_t idx = 0;
        H5Aiterate2(sub, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_attr, &kv);
        collect_groups(sub, kv, depth - 1);
    }
}

// every entry of a 1-D integer attribute (e.g. IDGenState: one per thread
// of the run that wrote it)
bool get_i64s(hid_t loc, const char *name, std::vector<int64_t> &out) {
    if (H5Aexists(loc, name) <= 0) return false;
    Hid a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
    Hid sp(H5Aget_space(a), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    if (n <= 0) return false;
    out.resize((size_t)n);
    return H5Aread(a, H5T_NATIVE_INT64, out.data()) >= 0;
}

// the parameter block of a species group, rebuilt from ClassName, PrioInfo
// and the action attributes; the draws' key from PhiloxSeed, or `seed` when
// the group has none (a file the reference wrote)
int params_from_group(hid_t g, uint64_t seed, qhg_params *out) {
    get_num(g, "PhiloxSeed", H5T_NATIVE_UINT64, &seed, 1);
    std::string cls;
    if (!get_str(g, "ClassName", cls)) return qfail(QHG_ERR_ARG, "no ClassName attribute");
    std::vector<std::string> pn;
    std::vector<int32_t> pv;
    if (H5Aexists(g, "PrioInfo") > 0) {
        Hid a(H5Aopen(g, "PrioInfo", H5P_DEFAULT), H5Aclose);
        Hid sp(H5Aget_space(a), H5Sclose);
        const hssize_t n = H5Sget_simple_extent_npoints(sp);
        std::vector<PrioRec> pr((size_t)std::max<hssize_t>(n, 1));
        Hid st(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(st, 128);
        Hid ct(H5Tcreate(H5T_COMPOUND, sizeof(PrioRec)), H5Tclose);
        H5Tinsert(ct, "FuncName", HOFFSET(PrioRec, name), st);
        H5Tinsert(ct, "FuncPrio", HOFFSET(PrioRec, prio), H5T_NATIVE_INT32);
        if (H5Aread(a, ct, pr.data()) < 0) return qfail(QHG_ERR_ARG, "reading PrioInfo");
        for (hssize_t k = 0; k < n; k++) {
            pr[k].name[127] = 0;
            pn.push_back(pr[k].name);
            pv.push_back(pr[k].prio);
        }
    }
    std::map<std::string, std::string> kv;
    hsize_t idx = 0;
    H5Aiterate2(g, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_attr, &kv);
    // per-action groups (and a MultiEvaluator's evaluator groups) of the
    // reference's current layout
    collect_groups(g, kv, 2);
    std::vector<const char *> k, v, p;
    for (auto &e : kv) {
        k.push_back(e.first.c_str());
        v.push_back(e.second.c_str());
    }
    for (auto &str : pn) p.push_back(str.c_str());
    int rc = qhg_params_from_attribute_map(cls.c_str(), (int32_t)k.size(), k.data(), v.data(), p.data(), pv.data(),
                                           (int32_t)p.size(), seed, out);
    if (rc) return qfail(rc, qhg_gpu_last_error());
    int32_t mode = 0;
    if (get_num(g, "MultiEvaluatorMode", H5T_NATIVE_INT32, &mode, 1)) {
        if (mode != QHG_ME_ADD_SIMPLE && mode != QHG_ME_MUL_SIMPLE && mode != QHG_ME_MAX_SIMPLE &&
            mode != QHG_ME_MIN_SIMPLE)
            return qfail(QHG_ERR_ARG, "unknown MultiEvaluatorMode");
        out->me_mode = mode;
    }
    return 0;
}

// what a species group says of itself; root (or -1) supplies the file's
// Step / StartTime when the group has no clock of its own
void info_from_group(hid_t g, hid_t root, qhg_qdf_info *info) {
    qhg_qdf_info I;
    std::memset(&I, 0, sizeof(I));
    hid_t ds = -1;
    if (H5Lexists(g, "AgentDataSet", H5P_DEFAULT) > 0) {
        ds = H5Dopen2(g, "AgentDataSet", H5P_DEFAULT);
        Hid sp(H5Dget_space(ds), H5Sclose);
        I.num_agents = (int64_t)H5Sget_simple_extent_npoints(sp);
    }
    Hid dsh(ds, H5Dclose);
    if (H5Lexists(g, "Genome", H5P_DEFAULT) > 0) {
        Hid gd(H5Dopen2(g, "Genome", H5P_DEFAULT), H5Dclose);
        Hid gs(H5Dget_space(gd), H5Sclose);
        hsize_t d[2] = {0, 0};
        if (H5Sget_simple_extent_ndims(gs) == 2) {
            H5Sget_simple_extent_dims(gs, d, nullptr);
            I.genome_words = (int32_t)d[1];
        }
    }
    std::string s;
    if (get_str(g, "ClassName", s)) std::strncpy(I.class_name, s.c_str(), sizeof(I.class_name) - 1);
    if (get_str(g, "SpeciesName", s)) std::strncpy(I.species_name, s.c_str(), sizeof(I.species_name) - 1);
    get_num(g, "NumCells", H5T_NATIVE_INT32, &I.num_cells, 1);
    if (!get_num(g, "StepCount", H5T_NATIVE_INT64, &I.step, 1) && root >= 0 && get_str(root, "Step", s))
        I.step = std::strtoll(s.c_str(), nullptr, 10);
    if (!get_num(g, "NextTime", H5T_NATIVE_DOUBLE, &I.time, 1) && root >= 0 && get_str(root, "StartTime", s))
        I.time = std::strtod(s.c_str(), nullptr);
    get_num(g, "PhiloxSeed", H5T_NATIVE_UINT64, &I.seed, 1);
    I.has_params = H5Aexists(g, "ClassName") > 0 && H5Aexists(g, "PrioInfo") > 0;
    // the highest id handed out: the largest IDGenState entry, else the
    // largest agent id in the file
    std::vector<int64_t> idg;
    if (get_i64s(g, "IDGenState", idg)) {
        I.max_id = *std::max_element(idg.begin(), idg.end());
    } else if (I.num_agents > 0) {
        std::vector<int64_t> ids((size_t)I.num_agents);
        Hid t(H5Tcreate(H5T_COMPOUND, sizeof(int64_t)), H5Tclose);
        H5Tinsert(t, "AgentID", 0, H5T_NATIVE_INT64);
        if (H5Dread(dsh, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, ids.data()) >= 0)
            I.max_id = *std::max_element(ids.begin(), ids.end());
    }
    *info = I;
}

static_assert(sizeof(hid_t) == sizeof(int64_t), "HDF5 handles travel as int64_t");

}  // namespace

extern "C" {

const char *qhg_qdf_last_error(void) { return g_err.c_str(); }

int qhg_qdf_write_agents(const char *path, int32_t create, const char *species, const char *pop_class,
                         const qhg_params *params, int32_t num_cells, const int64_t *cell_ids, int64_t step,
                         double time, int64_t max_id, const qhg_agents *agents, int32_t genome_words) {
    if (!path || !species || !params || !agents || agents->n < 0 || (agents->n > 0 && !agents->id))
        return qfail(QHG_ERR_ARG, "null argument");
    if (genome_words < 0 || (genome_words > 0 && agents->n > 0 && !agents->genome))
        return qfail(QHG_ERR_ARG, "genome words without genomes");
    quiet();
    Hid f(create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT) : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT),
          H5Fclose);
    if (!f.ok()) return qfail(QHG_ERR_ARG, std::string("cannot open ") + path);
    {
        Hid root(H5Gopen2(f, "/", H5P_DEFAULT), H5Gclose);
        char buf[64];
        int rc = create ? put_str(root, "QHG", "QHG4") : 0;
        std::snprintf(buf, sizeof(buf), "%lld", (long long)step);
        if (!rc) rc = put_str(root, "Step", buf);
        std::snprintf(buf, sizeof(buf), "%.17g", time);
        if (!rc) rc = put_str(root, "StartTime", buf);
        if (rc) return rc;
    }
    if (H5Lexists(f, "/Populations", H5P_DEFAULT) <= 0) {
        Hid g(H5Gcreate2(f, "/Populations", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return qfail(QHG_ERR_ARG, "cannot create /Populations");
    }
    Hid pops(H5Gopen2(f, "/Populations", H5P_DEFAULT), H5Gclose);
    if (H5Lexists(pops, species, H5P_DEFAULT) > 0) H5Ldelete(pops, species, H5P_DEFAULT);
    Hid g(H5Gcreate2(pops, species, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return qfail(QHG_ERR_ARG, std::string("cannot create species group ") + species);
    int rc = write_species_attrs(g, species, pop_class, params, num_cells, step, time, max_id);
    if (rc) return rc;
    const int64_t n = agents->n;
    {
        Hid t(agent_type(), H5Tclose);
        hsize_t d = (hsize_t)n;
        Hid sp(H5Screate_simple(1, &d, nullptr), H5Sclose);
        Hid ds(H5Dcreate2(g, "AgentDataSet", t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return qfail(QHG_ERR_ARG, "creating AgentDataSet");
        rc = write_records(ds, agents, cell_ids);
        if (rc) return rc;
    }
    if (genome_words > 0) return write_genome(g, agents->genome, n, genome_words);
    return 0;
}

int qhg_qdf_write_population(qhg_pop *pop, const char *path, int32_t create, const char *species,
                             const char *pop_class, int32_t num_cells, const int64_t *cell_ids) {
    if (!pop || !path || !species) return qfail(QHG_ERR_ARG, "null argument");
    qhg_params P;
    int64_t step = 0, max_id = 0, n = 0;
    double t = 0;
    if (qhg_gpu_get_params(pop, &P) || qhg_gpu_get_clock(pop, &step, &t, &max_id) || qhg_gpu_num_alive(pop, &n))
        return qfail(QHG_ERR_STATE, qhg_gpu_last_error());
    const int32_t gw = has_kind(&P, QHG_ACT_GENETICS) ? 2 * qhg_gpu_genome_words(P.gen_genome_size) : 0;
    std::vector<int64_t> id(n), mate(n);
    std::vector<int32_t> cell(n), nb(n);
    std::vector<uint8_t> life(n), gender(n);
    std::vector<double> birth(n), age(n), lastb(n);
    std::vector<uint64_t> genome((size_t)n * gw);
    qhg_agents a{n,          id.data(),    cell.data(),  life.data(), gender.data(), birth.data(),
                 age.data(), lastb.data(), mate.data(),  nb.data(),   gw ? genome.data() : nullptr};
    if (qhg_gpu_download_agents(pop, &a)) return qfail(QHG_ERR_STATE, qhg_gpu_last_error());
    return qhg_qdf_write_agents(path, create, species, pop_class, &P, num_cells, cell_ids, step, t, max_id, &a, gw);
}

int qhg_qdf_read_info(const char *path, const char *species, qhg_qdf_info *info) {
    if (!path || !species || !info) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return qfail(QHG_ERR_ARG, std::string("cannot open ") + path);
    Hid g(open_species(f, species), H5Gclose);
    if (!g.ok()) return qfail(QHG_ERR_ARG, std::string("no population group ") + species);
    if (H5Lexists(g, "AgentDataSet", H5P_DEFAULT) <= 0) return qfail(QHG_ERR_ARG, "no AgentDataSet");
    Hid root(H5Gopen2(f, "/", H5P_DEFAULT), H5Gclose);
    info_from_group(g, root, info);
    return 0;
}

int qhg_qdf_read_params(const char *path, const char *species, uint64_t seed, qhg_params *out) {
    if (!path || !species || !out) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return qfail(QHG_ERR_ARG, std::string("cannot open ") + path);
    Hid g(open_species(f, species), H5Gclose);
    if (!g.ok()) return qfail(QHG_ERR_ARG, std::string("no population group ") + species);
    return params_from_group(g, seed, out);
}

int qhg_qdf_read_agents(const char *path, const char *species, qhg_agents *a) {
    if (!path || !species || !a || (!a->id && a->n > 0)) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return qfail(QHG_ERR_ARG, std::string("cannot open ") + path);
    Hid g(open_species(f, species), H5Gclose);
    if (!g.ok()) return qfail(QHG_ERR_ARG, std::string("no population group ") + species);
    Hid ds(H5Dopen2(g, "AgentDataSet", H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return qfail(QHG_ERR_ARG, "no AgentDataSet");
    int rc = read_records(ds, a);
    if (rc) return rc;
    if (a->genome && H5Lexists(g, "Genome", H5P_DEFAULT) > 0) return read_genome(g, a->genome, a->n, -1);
    return 0;
}

int qhg_qdf_restore_population(const char *path, const char *species, qhg_pop *pop) {
    if (!pop) return qfail(QHG_ERR_ARG, "null handle");
    qhg_qdf_info I;
    int rc = qhg_qdf_read_info(path, species, &I);
    if (rc) return rc;
    qhg_params P;
    if (qhg_gpu_get_params(pop, &P)) return qfail(QHG_ERR_STATE, qhg_gpu_last_error());
    const int32_t gw = has_kind(&P, QHG_ACT_GENETICS) ? 2 * qhg_gpu_genome_words(P.gen_genome_size) : 0;
    if (gw != I.genome_words) return qfail(QHG_ERR_ARG, "the dump's genomes do not match the population's Genetics");
    const int64_t n = I.num_agents;
    std::vector<int64_t> id(n), mate(n);
    std::vector<int32_t> cell(n), nb(n);
    std::vector<uint8_t> life(n), gender(n);
    std::vector<double> birth(n), age(n), lastb(n);
    std::vector<uint64_t> genome((size_t)n * gw);
    qhg_agents a{n,          id.data(),    cell.data(), life.data(), gender.data(), birth.data(),
                 age.data(), lastb.data(), mate.data(), nb.data(),   gw ? genome.data() : nullptr};
    rc = qhg_qdf_read_agents(path, species, &a);
    if (rc) return rc;
    if (qhg_gpu_upload_agents(pop, &a) || qhg_gpu_set_clock(pop, I.step, I.time, I.max_id))
        return qfail(QHG_ERR_STATE, qhg_gpu_last_error());
    return 0;
}

// -- species groups and datasets opened by the caller ---------------------

int64_t qhg_qdf_agent_type(void) { return (int64_t)agent_type(); }

int qhg_qdf_write_species_group(int64_t group, const char *species, const char *pop_class, const qhg_params *params,
                                int32_t num_cells, int64_t step, double time, int64_t max_id) {
    if (group < 0 || !species || !params) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    return write_species_attrs((hid_t)group, species, pop_class, params, num_cells, step, time, max_id);
}

int qhg_qdf_read_species_group(int64_t group, uint64_t seed, qhg_params *out, qhg_qdf_info *info) {
    if (group < 0) return qfail(QHG_ERR_ARG, "invalid group");
    quiet();
    if (info) info_from_group((hid_t)group, -1, info);
    return out ? params_from_group((hid_t)group, seed, out) : 0;
}

int qhg_qdf_dataset_rows(int64_t dataset, int64_t *n) {
    if (dataset < 0 || !n) return qfail(QHG_ERR_ARG, "null argument");
    Hid sp(H5Dget_space((hid_t)dataset), H5Sclose);
    if (!sp.ok()) return qfail(QHG_ERR_ARG, "no dataspace");
    *n = (int64_t)H5Sget_simple_extent_npoints(sp);
    return 0;
}

int qhg_qdf_write_agent_dataset(int64_t dataset, const qhg_agents *agents, const int64_t *cell_ids) {
    if (dataset < 0 || !agents || agents->n < 0 || (agents->n > 0 && !agents->id))
        return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    return write_records((hid_t)dataset, agents, cell_ids);
}

int qhg_qdf_read_agent_dataset(int64_t dataset, qhg_agents *agents) {
    if (dataset < 0 || !agents || (!agents->id && agents->n > 0)) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    return read_records((hid_t)dataset, agents);
}

int qhg_qdf_write_genome(int64_t group, const uint64_t *genome, int64_t n, int32_t words) {
    if (group < 0 || n < 0 || words <= 0 || (n > 0 && !genome)) return qfail(QHG_ERR_ARG, "bad argument");
    quiet();
    return write_genome((hid_t)group, genome, n, words);
}

int qhg_qdf_read_genome(int64_t group, uint64_t *genome, int64_t n, int32_t words) {
    if (group < 0 || n < 0 || (n > 0 && !genome)) return qfail(QHG_ERR_ARG, "bad argument");
    quiet();
    return read_genome((hid_t)group, genome, n, words);
}

int qhg_qdf_read_grid_info(const char *path, qhg_qdf_grid_info *info) {
    if (!path || !info) return qfail(QHG_ERR_ARG, "null argument");
    quiet();
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return qfail(QHG_ERR_ARG, std::string("cannot open ") + path);
    if (H5Lexists(f, "/Grid", H5P_DEFAULT) <= 0 || H5Lexists(f, "/Grid/CellDataSet", H5P_DEFAULT) <= 0)
        return qfail(QHG_ERR_ARG, "no /Grid/CellDataSet");
    Hid ds(H5Dopen2(f, "/Grid/CellDataSet", H5P_DEFAULT), H5Dclose);
    Hid sp(H5Dget_space(ds), H5Sclose);
    Hid ft(H5Dget_type(ds), H5Tclose);
    const int mi = H5Tget_member_index(ft, "Neighbors");
    if (mi < 0 || H5Tget_member_index(ft, "CellID") < 0 || H5Tget_member_index(ft, "NumNeighbors") < 0)
        return qfail(QHG_ERR_ARG, "CellDataSet lacks CellID, NumNeighbors or Neighbors");
    Hid nt(H5Tget_member_type(ft, (unsigned)mi), H5Tclose);
    hsize_t dims[1] = {0};
    if (H5Tget_class(nt) != H5T_ARRAY || H5Tget_array_ndims(nt) != 1) return qfail(QHG_ERR_ARG, "Neighbors is not an array");
    H5Tget_array_dims2(nt, dims);
    qhg_qdf_grid_info I{};
    I.num_cells = (int32_t)H5Sget_simple_extent_npoints(sp);
    I.max_neighbors = (int32_t)dims[0];
    I.has_geography = H5Lexists(f, "/Geography", H5P_DEFAULT) > 0;
    I.has_climate = H5Lexists(f, "/Climate", H5P_DEFAULT) > 0;
    I.has_vegetation = H5Lexists(f, "/Vegetation", H5P_DEFAULT) > 0;
    *info = I;
    return 0;
}

int qhg_qdf_read_grid(const char *path, int32_t *cell_ids, uint8_t *num_neighbors, int32_t *neighbors, double *lon,
                      double *lat, double *alt, double *water, uint8_t *ice, uint8_t *coastal, double *base_npp,
                      double *temperature, double *rainfall, int32_t *missing) {
    qhg_qdf_grid_info I;
    int rc = qhg_qdf_read_grid_info(path, &I);
    if (rc) return rc;
    Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    const int32_t nc = I.num_cells, mx = I.max_neighbors;
    // the cell records with the neighbour ids, then ids -> indices
    const size_t rec = 8 + 4 * (size_t)mx;
    std::vector<uint8_t> buf(rec * (size_t)nc);
    {
        Hid ds(H5Dopen2(f, "/Grid/CellDataSet", H5P_DEFAULT), H5Dclose);
        hsize_t d = (hsize_t)mx;
        Hid at(H5Tarray_create2(H5T_NATIVE_INT32, 1, &d), H5Tclose);
        Hid mt(H5Tcreate(H5T_COMPOUND, rec), H5Tclose);
        H5Tinsert(mt, "CellID", 0, H5T_NATIVE_INT32);
        H5Tinsert(mt, "NumNeighbors", 4, H5T_NATIVE_UINT8);
        H5Tinsert(mt, "Neighbors", 8, at);
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
            return qfail(QHG_ERR_ARG, "reading CellDataSet");
    }
    std::unordered_map<int32_t, int32_t> index;
    index.reserve((size_t)nc * 2);
    for (int32_t c = 0; c < nc; c++) {
        int32_t id;
        std::memcpy(&id, &buf[rec * c], 4);
        index[id] = c;
        if (cell_ids) cell_ids[c] = id;
        const uint8_t k = buf[rec * c + 4];
        if (k > mx) return qfail(QHG_ERR_ARG, "NumNeighbors above the array size");
        if (num_neighbors) num_neighbors[c] = k;
    }
    if (neighbors) {
        for (int32_t c = 0; c < nc; c++) {
            const uint8_t k = buf[rec * c + 4];
            for (int32_t j = 0; j < mx; j++) {
                int32_t nid;
                std::memcpy(&nid, &buf[rec * c + 8 + 4 * j], 4);
                auto it = index.find(nid);
                neighbors[(size_t)c * mx + j] = (j < k && it != index.end()) ? it->second : -1;
            }
        }
    }
    struct Field {
        const char *path;
        void *out;
        hid_t type;
    } fields[] = {{"/Geography/Longitude", lon, H5T_NATIVE_DOUBLE},
                  {"/Geography/Latitude", lat, H5T_NATIVE_DOUBLE},
                  {"/Geography/Altitude", alt, H5T_NATIVE_DOUBLE},
                  {"/Geography/Water", water, H5T_NATIVE_DOUBLE},
                  {"/Geography/IceCover", ice, H5T_NATIVE_UINT8},
                  {"/Geography/Coastal", coastal, H5T_NATIVE_UINT8},
                  {"/Vegetation/BaseNPP", base_npp, H5T_NATIVE_DOUBLE},
                  {"/Climate/AnnualMeanTemp", temperature, H5T_NATIVE_DOUBLE},
                  {"/Climate/AnnualRainfall", rainfall, H5T_NATIVE_DOUBLE}};
    int32_t miss = 0;
    for (int k = 0; k < 9; k++) {
        if (!fields[k].out) continue;
        const std::string p = fields[k].path;
        const std::string grp = p.substr(0, p.find('/', 1));
        if (H5Lexists(f, grp.c_str(), H5P_DEFAULT) <= 0 || H5Lexists(f, p.c_str(), H5P_DEFAULT) <= 0) {
            miss |= 1 << k;
            continue;
        }
        Hid ds(H5Dopen2(f, p.c_str(), H5P_DEFAULT), H5Dclose);
        Hid sp(H5Dget_space(ds), H5Sclose);
        if (H5Sget_simple_extent_npoints(sp) != nc) return qfail(QHG_ERR_ARG, p + " has a different cell count");
        if (H5Dread(ds, fields[k].type, H5S_ALL, H5S_ALL, H5P_DEFAULT, fields[k].out) < 0)
            return qfail(QHG_ERR_ARG, "reading " + p);
    }
    if (missing) *missing = miss;
    return 0;
}

}  // extern "C"
