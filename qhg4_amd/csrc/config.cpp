// Species configuration from the reference's XML attributes (host only).
//
// A QHG4 species block names a population class and lists one <module> per
// action (nested <module>s for the evaluators of a MultiEvaluator) with
// <param name value> entries, plus a <priorities> map from action names to
// priority levels.  SPopulation::readSpeciesData
// (QHG4/core/SPopulation.cpp:1108-1142) hands the modules to every action of
// the class (Prioritizer::getActionAttributes, QHG4/core/Prioritizer.cpp:
// 440-462) and the priorities to the Prioritizer (SPopulation::setPrioList,
// :249-257).  This file applies the same rules to fill a qhg_params:
//   - the class fixes the actions, their ids and the evaluator wiring
//     (QHG4/populations/tut_SexualPop.cpp:24-43, tut_EnvironCapAltPop.cpp,
//     OoAPop.cpp:32-88, OoANavGen2bitPop.cpp:30-100,
//     OoANavGen2bitAssHamPop.cpp:31-93, OoANavGen2bitAssPop.cpp:33-95);
//   - an action's module must hold exactly its attribute names
//     (Action::checkAttributes, QHG4/actions/Action.cpp:54-75 and
//     QHG4/core/clsutils.cpp:75-103): a missing or unknown name is an error;
//   - an action that has attributes but no module is an error;
//   - a priority for an action the class does not have is an error
//     (Prioritizer::setPrio, QHG4/core/Prioritizer.cpp:19-30); an action
//     without a priority is never scheduled;
//   - a scheduled action that has no device implementation is an error.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/qhg_gpu.h"

int qhg_internal_fail(int code, const std::string &msg);   // sets qhg_gpu_last_error (qhg_gpu.hip)

namespace {

struct EvalSpec {
    const char *id;        // SingleEvaluator id: module "SingleEvaluator[id]"
    int input;             // QHG_EV_*
    const char *pl_attr;   // PolyLine attribute name, "" = values used directly
    const char *weight;    // the MultiEvaluator's weight attribute for it
};

struct ClassSpec {
    const char *name;
    std::vector<int> kinds;               // actions of the class on the GPU path
    const char *me_id;                    // MultiEvaluator id ("" = none)
    int me_mode;
    std::vector<EvalSpec> evals;
    std::vector<const char *> other;      // actions of the class with no device implementation
};

const std::vector<ClassSpec> &classes() {
    static const std::vector<ClassSpec> c = {
        {"tut_SexualPop",
         {QHG_ACT_GETOLD, QHG_ACT_ATANDEATH, QHG_ACT_RANDOMMOVE, QHG_ACT_FERTILITY, QHG_ACT_VERHULST,
          QHG_ACT_RANDOMPAIR},
         "", 0, {}, {}},
        {"tut_EnvironCapAltPop",
         {QHG_ACT_GETOLD, QHG_ACT_ATANDEATH, QHG_ACT_FERTILITY, QHG_ACT_VERHULSTVARK, QHG_ACT_RANDOMPAIR,
          QHG_ACT_MULTIEVAL, QHG_ACT_WEIGHTEDMOVE, QHG_ACT_NPPCAP},
         "NPP+Alt", QHG_ME_ADD_SIMPLE,
         {{"Alt", QHG_EV_ALTITUDE, "AltPref", "Multi_weight_alt"}, {"NPP", QHG_EV_CAPACITY, "", "Multi_weight_npp"}},
         {}},
        {"OoAPop",
         {QHG_ACT_MULTIEVAL, QHG_ACT_WEIGHTEDMOVE, QHG_ACT_VERHULSTVARK, QHG_ACT_RANDOMPAIR, QHG_ACT_GETOLD,
          QHG_ACT_OLDAGEDEATH, QHG_ACT_FERTILITY, QHG_ACT_NPPCAP},
         "Alt+NPP", QHG_ME_ADD_SIMPLE,
         {{"Alt", QHG_EV_ALTITUDE, "AltCapPref", "Multi_weight_alt"}, {"NPP", QHG_EV_CAPACITY, "", "Multi_weight_npp"}},
         {}},
        {"OoANavGen2bitPop",
         {QHG_ACT_MULTIEVAL, QHG_ACT_WEIGHTEDMOVE, QHG_ACT_VERHULSTVARK, QHG_ACT_RANDOMPAIR, QHG_ACT_GETOLD,
          QHG_ACT_OLDAGEDEATH, QHG_ACT_FERTILITY, QHG_ACT_NPPCAP, QHG_ACT_GENETICS, QHG_ACT_NAVIGATE},
         "Alt+NPP", QHG_ME_ADD_SIMPLE,
         {{"Alt", QHG_EV_ALTITUDE, "AltCapPref", "Multi_weight_alt"},
          {"NPP", QHG_EV_CAPACITY, "NPPPref", "Multi_weight_npp"}},
         {}},
        {"OoANavGen2bitAssHamPop",
         {QHG_ACT_MULTIEVAL, QHG_ACT_WEIGHTEDMOVE, QHG_ACT_VERHULSTVARK, QHG_ACT_ASSORTPAIR, QHG_ACT_GETOLD,
          QHG_ACT_OLDAGEDEATH, QHG_ACT_FERTILITY, QHG_ACT_NPPCAP, QHG_ACT_GENETICS, QHG_ACT_NAVIGATE},
         "Alt+NPP", QHG_ME_ADD_SIMPLE,
         {{"Alt", QHG_EV_ALTITUDE, "AltCapPref", "Multi_weight_alt"},
          {"NPP", QHG_EV_CAPACITY, "NPPPref", "Multi_weight_npp"}},
         {}},
        {"OoANavGen2bitAssPop",
         {QHG_ACT_MULTIEVAL, QHG_ACT_WEIGHTEDMOVE, QHG_ACT_VERHULSTVARK, QHG_ACT_ASSORTPROT, QHG_ACT_GETOLD,
          QHG_ACT_OLDAGEDEATH, QHG_ACT_FERTILITY, QHG_ACT_NPPCAP, QHG_ACT_GENETICS, QHG_ACT_NAVIGATE},
         "Alt+NPP", QHG_ME_ADD_SIMPLE,
         {{"Alt", QHG_EV_ALTITUDE, "AltCapPref", "Multi_weight_alt"},
          {"NPP", QHG_EV_CAPACITY, "NPPPref", "Multi_weight_npp"}},
         {}},
    };
    return c;
}

const char *base_name(int kind) {
    switch (kind) {
    case QHG_ACT_GETOLD: return "GetOld";
    case QHG_ACT_FERTILITY: return "Fertility";
    case QHG_ACT_RANDOMPAIR: return "RandomPair";
    case QHG_ACT_VERHULST: return "Verhulst";
    case QHG_ACT_VERHULSTVARK: return "VerhulstVarK";
    case QHG_ACT_RANDOMMOVE: return "RandomMove";
    case QHG_ACT_WEIGHTEDMOVE: return "WeightedMove";
    case QHG_ACT_MULTIEVAL: return "MultiEvaluator";
    case QHG_ACT_NPPCAP: return "NPPCapacity";
    case QHG_ACT_GENETICS: return "Genetics";
    case QHG_ACT_OLDAGEDEATH: return "OldAgeDeath";
    case QHG_ACT_ATANDEATH: return "ATanDeath";
    case QHG_ACT_ASSORTPAIR: return "AssortativePairHamming";
    case QHG_ACT_ASSORTPROT: return "AssortativePair";
    case QHG_ACT_NAVIGATE: return "Navigate";
    default: return "";
    }
}

// An attribute: its name and where its value goes.
struct Attr {
    const char *name;
    double qhg_params::*d;    // a double field, or
    int32_t qhg_params::*i;   // an int field, or neither (read and checked only)
};

// the attribute names of each action (the asNames of its .cpp), in order
const std::vector<Attr> &attrs(int kind) {
    static const std::vector<Attr> none;
    static const std::vector<Attr> fert = {{"Fertility_min_age", &qhg_params::fert_min_age, nullptr},
                                           {"Fertility_max_age", &qhg_params::fert_max_age, nullptr},
                                           {"Fertility_interbirth", &qhg_params::fert_interbirth, nullptr}};
    static const std::vector<Attr> verh = {{"Verhulst_b0", &qhg_params::verh_b0, nullptr},
                                           {"Verhulst_d0", &qhg_params::verh_d0, nullptr},
                                           {"Verhulst_theta", &qhg_params::verh_theta, nullptr},
                                           {"Verhulst_K", &qhg_params::verh_K, nullptr}};
    static const std::vector<Attr> verhk = {{"Verhulst_b0", &qhg_params::verh_b0, nullptr},
                                            {"Verhulst_d0", &qhg_params::verh_d0, nullptr},
                                            {"Verhulst_theta", &qhg_params::verh_theta, nullptr}};
    static const std::vector<Attr> rmove = {{"RandomMove_prob", &qhg_params::rmove_prob, nullptr}};
    static const std::vector<Attr> wmove = {{"WeightedMove_prob", &qhg_params::wmove_prob, nullptr}};
    static const std::vector<Attr> npp = {{"NPPCap_water_factor", &qhg_params::npp_water_factor, nullptr},
                                          {"NPPCap_coastal_factor", &qhg_params::npp_coastal_factor, nullptr},
                                          {"NPPCap_coastal_min_latitude", &qhg_params::npp_coastal_min_lat, nullptr},
                                          {"NPPCap_coastal_max_latitude", &qhg_params::npp_coastal_max_lat, nullptr},
                                          {"NPPCap_NPP_min", &qhg_params::npp_min, nullptr},
                                          {"NPPCap_NPP_max", &qhg_params::npp_max, nullptr},
                                          {"NPPCap_K_min", &qhg_params::npp_k_min, nullptr},
                                          {"NPPCap_K_max", &qhg_params::npp_k_max, nullptr},
                                          {"NPPCap_efficiency", &qhg_params::npp_efficiency, nullptr}};
    static const std::vector<Attr> oad = {{"OAD_max_age", &qhg_params::oad_max_age, nullptr},
                                          {"OAD_uncertainty", &qhg_params::oad_uncertainty, nullptr}};
    static const std::vector<Attr> atd = {{"ATanDeath_max_age", &qhg_params::atan_max_age, nullptr},
                                          {"ATanDeath_range", &qhg_params::atan_range, nullptr},
                                          {"ATanDeath_slope", &qhg_params::atan_slope, nullptr}};
    static const std::vector<Attr> gen = {{"Genetics_genome_size", nullptr, &qhg_params::gen_genome_size},
                                          {"Genetics_num_crossover", nullptr, &qhg_params::gen_num_crossover},
                                          {"Genetics_mutation_rate", &qhg_params::gen_mutation_rate, nullptr},
                                          {"Genetics_initial_muts", nullptr, nullptr},
                                          {"Genetics_create_new_genome", nullptr, nullptr},
                                          {"Genetics_bits_per_nuc", nullptr, nullptr}};
    // AssortativePairHam.h:11-13 (the cutoff is read as a float)
    static const std::vector<Attr> asp = {{"AssortativePairHamming_cutoff", &qhg_params::assort_cutoff, nullptr},
                                          {"AssortativePairHamming_permute", nullptr, &qhg_params::assort_permute}};
    // AssortativePair.h:11-13 (the same two values)
    static const std::vector<Attr> aspp = {{"AssortativePair_cutoff", &qhg_params::assort_cutoff, nullptr},
                                           {"AssortativePair_permute", nullptr, &qhg_params::assort_permute}};
    // Navigate.cpp:18-24 (tryGetAttributes requires all five through
    // checkAttributes; a QDF file may lack the bridge probability, 0 then,
    // extractAttributesQDF, Navigate.cpp:263-302)
    static const std::vector<Attr> nav = {{"Navigate_decay", &qhg_params::nav_decay, nullptr},
                                          {"Navigate_dist0", &qhg_params::nav_dist0, nullptr},
                                          {"Navigate_prob0", &qhg_params::nav_prob0, nullptr},
                                          {"Navigate_min_dens", &qhg_params::nav_min_dens, nullptr},
                                          {"Navigate_bridge_prob", &qhg_params::nav_bridge_prob, nullptr}};
    switch (kind) {
    case QHG_ACT_FERTILITY: return fert;
    case QHG_ACT_VERHULST: return verh;
    case QHG_ACT_VERHULSTVARK: return verhk;
    case QHG_ACT_RANDOMMOVE: return rmove;
    case QHG_ACT_WEIGHTEDMOVE: return wmove;
    case QHG_ACT_NPPCAP: return npp;
    case QHG_ACT_OLDAGEDEATH: return oad;
    case QHG_ACT_ATANDEATH: return atd;
    case QHG_ACT_GENETICS: return gen;
    case QHG_ACT_ASSORTPAIR: return asp;
    case QHG_ACT_ASSORTPROT: return aspp;
    case QHG_ACT_NAVIGATE: return nav;
    default: return none;
    }
}

// strToNum (QHG4/utils/strutils.cpp): the whole string is a number
bool to_double(const char *s, double *out) {
    if (!s || !*s) return false;
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(s, &end);
    while (end && (*end == ' ' || *end == '\t')) end++;
    if (errno != 0 || !end || *end != '\0') return false;
    *out = v;
    return true;
}

bool to_int(const char *s, int32_t *out) {
    if (!s || !*s) return false;
    char *end = nullptr;
    errno = 0;
    const long v = std::strtol(s, &end, 10);
    while (end && (*end == ' ' || *end == '\t')) end++;
    if (errno != 0 || !end || *end != '\0' || v < INT32_MIN || v > INT32_MAX) return false;
    *out = (int32_t)v;
    return true;
}

// PolyLine::readFromString (QHG4/utils/PolyLine.cpp:92-127): an even number
// (at least 4) of numbers, x v pairs
bool read_polyline(const char *s, std::vector<double> &xs, std::vector<double> &vs) {
    std::vector<double> all;
    const char *p = s;
    while (*p) {
        while (*p == ' ' || *p == '\t') p++;
        if (!*p) break;
        const char *q = p;
        while (*q && *q != ' ' && *q != '\t') q++;
        const std::string tok(p, q);
        double v = 0;
        if (!to_double(tok.c_str(), &v)) return false;
        all.push_back(v);
        p = q;
    }
    if (all.size() < 4 || all.size() % 2) return false;
    xs.clear();
    vs.clear();
    for (size_t k = 0; k < all.size(); k += 2) {
        xs.push_back(all[k]);
        vs.push_back(all[k + 1]);
    }
    return true;
}

using ParamMap = std::map<std::string, std::string>;

std::string action_name(const ClassSpec &c, int kind) {
    if (kind == QHG_ACT_MULTIEVAL) return std::string("MultiEvaluator[") + c.me_id + "]";
    return base_name(kind);
}

// The reference's names of a class's actions (Action::getActionName,
// QHG4/actions/Action.cpp:11-19, with the ids the class constructors give
// them), built once; then its evaluators' names.
const std::vector<std::string> &class_names(const ClassSpec &c, bool evaluators) {
    static const auto *tab = [] {
        auto *t = new std::map<std::string, std::pair<std::vector<std::string>, std::vector<std::string>>>();
        for (const ClassSpec &k : classes()) {
            auto &e = (*t)[k.name];
            for (int kind : k.kinds) e.first.push_back(action_name(k, kind));
            for (const char *o : k.other) e.first.push_back(o);
            for (const EvalSpec &ev : k.evals) e.second.push_back(std::string("SingleEvaluator[") + ev.id + "]");
        }
        return t;
    }();
    const auto &e = tab->at(c.name);
    return evaluators ? e.second : e.first;
}

}  // namespace

extern "C" {

int qhg_params_from_modules(const char *pop_class, const qhg_module_attrs *mods, int32_t nmods,
                            const char *const *prio_names, const int32_t *prio_values, int32_t nprios, uint64_t seed,
                            qhg_params *out) {
    if (!pop_class || !out || (nmods > 0 && !mods) || (nprios > 0 && (!prio_names || !prio_values)))
        return qhg_internal_fail(QHG_ERR_ARG, "null argument");
    const ClassSpec *cs = nullptr;
    for (const ClassSpec &c : classes())
        if (std::strcmp(c.name, pop_class) == 0) cs = &c;
    if (!cs) return qhg_internal_fail(QHG_ERR_ARG, std::string("unknown population class ") + pop_class);
    std::map<std::string, ParamMap> modules;
    for (int32_t m = 0; m < nmods; m++) {
        if (!mods[m].name || (mods[m].num_params > 0 && (!mods[m].keys || !mods[m].values)))
            return qhg_internal_fail(QHG_ERR_ARG, "module without a name or parameter arrays");
        ParamMap &pm = modules[mods[m].name];
        for (int32_t k = 0; k < mods[m].num_params; k++) {
            if (!mods[m].keys[k] || !mods[m].values[k]) return qhg_internal_fail(QHG_ERR_ARG, "null parameter");
            pm[mods[m].keys[k]] = mods[m].values[k];
        }
    }
    qhg_params P;
    std::memset(&P, 0, sizeof(P));
    P.seed = seed;
    P.npp_efficiency = 1.0;
    // attributes of every action of the class (Prioritizer::getActionAttributes)
    for (int kind : cs->kinds) {
        const std::string an = action_name(*cs, kind);
        auto it = modules.find(an);
        if (kind == QHG_ACT_MULTIEVAL) {
            if (it == modules.end()) return qhg_internal_fail(QHG_ERR_ARG, "no module for " + an);
            // MultiEvaluator::tryGetAttributes (MultiEvaluator.cpp:550-571):
            // its weights, then every evaluator from its sub-module
            P.me_mode = cs->me_mode;
            P.me_num = (int32_t)cs->evals.size();
            for (size_t e = 0; e < cs->evals.size(); e++) {
                const EvalSpec &ev = cs->evals[e];
                auto w = it->second.find(ev.weight);
                if (w == it->second.end() || !to_double(w->second.c_str(), &P.me_weight[e]))
                    return qhg_internal_fail(QHG_ERR_ARG, an + ": missing or bad " + ev.weight);
                P.me_input[e] = ev.input;
                P.me_cumulate[e] = 0;
                P.me_pl_n[e] = 0;
                if (*ev.pl_attr) {
                    const std::string sub = std::string("SingleEvaluator[") + ev.id + "]";
                    auto sm = modules.find(sub);
                    if (sm == modules.end()) return qhg_internal_fail(QHG_ERR_ARG, "no module for " + sub);
                    auto pl = sm->second.find(ev.pl_attr);
                    std::vector<double> xs, vs;
                    if (pl == sm->second.end() || !read_polyline(pl->second.c_str(), xs, vs))
                        return qhg_internal_fail(QHG_ERR_ARG, sub + ": missing or bad PolyLine " + ev.pl_attr);
                    if ((int)xs.size() > QHG_MAX_PL_POINTS)
                        return qhg_internal_fail(QHG_ERR_ARG, sub + ": PolyLine has more than 64 points");
                    P.me_pl_n[e] = (int32_t)xs.size();
                    for (size_t j = 0; j < xs.size(); j++) {
                        P.me_pl_x[e][j] = xs[j];
                        P.me_pl_v[e][j] = vs[j];
                    }
                }
            }
            for (const auto &kv : it->second) {
                bool known = false;
                for (const EvalSpec &ev : cs->evals) known = known || kv.first == ev.weight;
                if (!known) return qhg_internal_fail(QHG_ERR_ARG, an + ": unknown parameter " + kv.first);
            }
            continue;
        }
        const std::vector<Attr> &names = attrs(kind);
        if (it == modules.end()) {
            if (!names.empty()) return qhg_internal_fail(QHG_ERR_ARG, "no module for " + an);
            continue;
        }
        for (const auto &kv : it->second) {
            bool known = false;
            for (const Attr &a : names) known = known || kv.first == a.name;
            if (!known) return qhg_internal_fail(QHG_ERR_ARG, an + ": unknown parameter " + kv.first);
        }
        for (const Attr &a : names) {
            auto v = it->second.find(a.name);
            if (v == it->second.end()) return qhg_internal_fail(QHG_ERR_ARG, an + ": missing parameter " + a.name);
            bool ok = true;
            if (a.d) ok = to_double(v->second.c_str(), &(P.*(a.d)));
            else if (a.i) ok = to_int(v->second.c_str(), &(P.*(a.i)));
            if (!ok) return qhg_internal_fail(QHG_ERR_ARG, an + ": bad value for " + a.name);
            if (std::strcmp(a.name, "Genetics_bits_per_nuc") == 0) {
                int32_t b = 0;
                if (!to_int(v->second.c_str(), &b) || b != 2)
                    return qhg_internal_fail(QHG_ERR_ARG, "Genetics_bits_per_nuc: only 2-bit genomes run on the GPU");
            }
        }
    }
    // priorities (SPopulation::setPrioList + Prioritizer::setPrio)
    std::set<std::string> seen;
    for (int32_t k = 0; k < nprios; k++) {
        const std::string n = prio_names[k];
        if (!seen.insert(n).second) return qhg_internal_fail(QHG_ERR_ARG, "priority given twice for " + n);
        int kind = 0;
        for (int kk : cs->kinds)
            if (action_name(*cs, kk) == n) kind = kk;
        if (!kind) {
            for (const char *o : cs->other)
                if (n == o) return qhg_internal_fail(QHG_ERR_ARG, "action " + n + " has no GPU implementation");
            return qhg_internal_fail(QHG_ERR_ARG, "priority for " + n + ", which " + cs->name + " does not have");
        }
        if (P.num_actions >= QHG_MAX_ACTIONS) return qhg_internal_fail(QHG_ERR_ARG, "too many actions");
        const int q = P.num_actions++;
        P.action_kind[q] = kind;
        P.action_prio[q] = prio_values[k];
        std::strncpy(P.action_name[q], n.c_str(), QHG_NAME_LEN - 1);
    }
    *out = P;
    return 0;
}

const char *qhg_params_attribute_name(int32_t kind, int32_t index) {
    const std::vector<Attr> &a = attrs(kind);
    return (index >= 0 && index < (int32_t)a.size()) ? a[index].name : nullptr;
}

const char *qhg_params_weight_name(const char *pop_class, int32_t eval) {
    for (const ClassSpec &c : classes())
        if (pop_class && std::strcmp(c.name, pop_class) == 0)
            return (eval >= 0 && eval < (int32_t)c.evals.size()) ? c.evals[eval].weight : nullptr;
    return nullptr;
}

const char *qhg_params_polyline_name(const char *pop_class, int32_t eval) {
    for (const ClassSpec &c : classes())
        if (pop_class && std::strcmp(c.name, pop_class) == 0)
            return (eval >= 0 && eval < (int32_t)c.evals.size() && *c.evals[eval].pl_attr) ? c.evals[eval].pl_attr
                                                                                              : nullptr;
    return nullptr;
}

const char *qhg_params_class_action(const char *pop_class, int32_t index, int32_t *kind) {
    for (const ClassSpec &c : classes()) {
        if (!pop_class || std::strcmp(c.name, pop_class) != 0) continue;
        const std::vector<std::string> &names = class_names(c, false);
        if (index < 0 || index >= (int32_t)names.size()) return nullptr;
        if (kind) *kind = (index < (int32_t)c.kinds.size()) ? c.kinds[index] : 0;
        return names[index].c_str();
    }
    return nullptr;
}

const char *qhg_params_evaluator_name(const char *pop_class, int32_t eval) {
    for (const ClassSpec &c : classes()) {
        if (!pop_class || std::strcmp(c.name, pop_class) != 0) continue;
        const std::vector<std::string> &names = class_names(c, true);
        return (eval >= 0 && eval < (int32_t)names.size()) ? names[eval].c_str() : nullptr;
    }
    return nullptr;
}

// A species' attributes as one flat map (the attributes of a QDF species
// group, written by Prioritizer::writeActionParamsQDF): every name goes to
// the module of the class's action that owns it; names no action of the
// class owns (ClassName, NumCells, other actions' attributes) are ignored.
int qhg_params_from_attribute_map(const char *pop_class, int32_t n, const char *const *keys,
                                  const char *const *values, const char *const *prio_names,
                                  const int32_t *prio_values, int32_t nprios, uint64_t seed, qhg_params *out) {
    if (!pop_class || (n > 0 && (!keys || !values))) return qhg_internal_fail(QHG_ERR_ARG, "null argument");
    const ClassSpec *cs = nullptr;
    for (const ClassSpec &c : classes())
        if (std::strcmp(c.name, pop_class) == 0) cs = &c;
    if (!cs) return qhg_internal_fail(QHG_ERR_ARG, std::string("unknown population class ") + pop_class);
    std::map<std::string, std::string> kv;
    for (int32_t k = 0; k < n; k++)
        if (keys[k] && values[k]) kv[keys[k]] = values[k];
    if (kv.count("Navigate_decay") && !kv.count("Navigate_bridge_prob")) kv["Navigate_bridge_prob"] = "0";
    std::vector<std::pair<std::string, std::vector<std::pair<std::string, std::string>>>> mods;
    auto add = [&](const std::string &module, const char *key) {
        auto it = kv.find(key);
        if (it == kv.end()) return;
        for (auto &m : mods)
            if (m.first == module) { m.second.push_back({key, it->second}); return; }
        mods.push_back({module, {{key, it->second}}});
    };
    for (int kind : cs->kinds) {
        const std::string an = action_name(*cs, kind);
        if (kind == QHG_ACT_MULTIEVAL) {
            mods.push_back({an, {}});
            for (const EvalSpec &ev : cs->evals) {
                add(an, ev.weight);
                if (*ev.pl_attr) add(std::string("SingleEvaluator[") + ev.id + "]", ev.pl_attr);
            }
            continue;
        }
        for (const Attr &a : attrs(kind)) add(an, a.name);
    }
    std::vector<qhg_module_attrs> ma(mods.size());
    std::vector<std::vector<const char *>> kp(mods.size()), vp(mods.size());
    for (size_t m = 0; m < mods.size(); m++) {
        for (auto &e : mods[m].second) {
            kp[m].push_back(e.first.c_str());
            vp[m].push_back(e.second.c_str());
        }
        ma[m] = qhg_module_attrs{mods[m].first.c_str(), (int32_t)kp[m].size(), kp[m].data(), vp[m].data()};
    }
    return qhg_params_from_modules(pop_class, ma.data(), (int32_t)ma.size(), prio_names, prio_values, nprios, seed,
                                   out);
}

int qhg_params_get_attribute(const qhg_params *p, const char *name, double *value) {
    if (!p || !name || !value) return qhg_internal_fail(QHG_ERR_ARG, "null argument");
    for (int kind = QHG_ACT_GETOLD; kind <= QHG_ACT_LAST; kind++)
        for (const Attr &a : attrs(kind)) {
            if (std::strcmp(a.name, name) != 0) continue;
            if (a.d) { *value = p->*(a.d); return 0; }
            if (a.i) { *value = (double)(p->*(a.i)); return 0; }
        }
    static const char *weights[] = {"Multi_weight_alt", "Multi_weight_npp"};
    for (int e = 0; e < 2 && e < p->me_num; e++)
        if (std::strcmp(name, weights[e]) == 0) { *value = p->me_weight[e]; return 0; }
    return qhg_internal_fail(QHG_ERR_ARG, std::string("no numeric attribute ") + name);
}

// modifyAttributes of the actions (e.g. Verhulst.cpp:235-254): the named
// double attribute takes the value; int attributes and unknown names fail
int qhg_params_set_attribute(qhg_params *p, const char *name, double value) {
    if (!p || !name) return qhg_internal_fail(QHG_ERR_ARG, "null argument");
    for (int kind = QHG_ACT_GETOLD; kind <= QHG_ACT_LAST; kind++)
        for (const Attr &a : attrs(kind))
            if (std::strcmp(a.name, name) == 0 && a.d) {
                p->*(a.d) = value;
                return 0;
            }
    static const char *weights[] = {"Multi_weight_alt", "Multi_weight_npp"};
    for (int e = 0; e < 2 && e < p->me_num; e++)
        if (std::strcmp(name, weights[e]) == 0) { p->me_weight[e] = value; return 0; }
    return qhg_internal_fail(QHG_ERR_ARG, std::string("cannot modify attribute ") + name);
}

// The updateEvent rules of the classes: the three environment-driven ones
// drown agents on geography events (OoAPop.cpp:164-196,
// OoANavGen2bitPop.cpp:177-210, tut_EnvironCapAltPop.cpp:120-150);
// tut_SexualPop keeps SPopulation's no-op (SPopulation.h:116).
int qhg_params_event_flags(const char *pop_class, int32_t events) {
    if (!pop_class) return qhg_internal_fail(QHG_ERR_ARG, "null class name");
    for (const ClassSpec &c : classes()) {
        if (std::strcmp(c.name, pop_class) != 0) continue;
        const bool drowns = std::strcmp(c.name, "tut_SexualPop") != 0;
        return (drowns && (events & QHG_EVENT_GEO)) ? (events | QHG_EVENT_KILL_WET) : events;
    }
    return qhg_internal_fail(QHG_ERR_ARG, std::string("unknown population class ") + pop_class);
}

}  // extern "C"
