// Parameter dictionary (see params.hpp).
#include "params.hpp"

#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

namespace xamg {

std::string ParamList::get_string(const std::string &k) const {
    auto it = kv.find(k);
    if (it == kv.end())
        throw std::runtime_error("xamg: parameter '" + k + "' is not set");
    return it->second;
}

long ParamList::get_int(const std::string &k) const { return std::strtol(get_string(k).c_str(), nullptr, 10); }

double ParamList::get_float(const std::string &k) const {
    return std::strtod(get_string(k).c_str(), nullptr);
}

bool ParamList::get_bool(const std::string &k) const {
    std::string s = get_string(k);
    return s == "1" || s == "true" || s == "True" || s == "yes";
}

ParamList &GlobalParams::get(const std::string &role) { return roles[role]; }

const ParamList &GlobalParams::get(const std::string &role) const {
    auto it = roles.find(role);
    if (it == roles.end())
        throw std::runtime_error("xamg: no parameter block for role '" + role + "'");
    return it->second;
}

ParamList GlobalParams::get_level(const std::string &role, int lev) const {
    ParamList l = get(role);
    auto it = overrides.find(role);
    if (it != overrides.end())
        for (const auto &o : it->second)
            if (lev >= o.lev_lo && lev <= o.lev_hi)
                l.set(o.key, o.value);
    return l;
}

static void split(const std::string &s, char d, std::set<std::string> &out) {
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, d))
        if (!t.empty())
            out.insert(t);
}

static std::string num(double v) {
    char b[64];
    snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

void set_family_defaults(ParamList &list, const std::string &method, const std::string &role) {
    if (method == "Identity" || method == "Direct")
        return;
    if (method == "IterativeRefinement") {
        // xamg_params_details.h:180-185 returns here: tolerances and max_iters of the
        // meta solver come from the caller (the solver falls back to 0 / 20).
        list.set_if_missing("subsolver_reduced_precision", "0");
        return;
    }
    if (method == "ResidualReplacement") {
        list.set_if_missing("RR_deviation", num(2.0));
        list.set_if_missing("RR_stepping", num(0.1));
        return;
    }
    if (method == "Chebyshev") {
        list.set_if_missing("convergence_details", "nocheck,noprint");
        list.set_if_missing("polynomial_order", "2");
        list.set_if_missing("spectrum_fraction", num(0.3));
        list.set_if_missing("abs_tolerance", num(0));
        list.set_if_missing("rel_tolerance", num(0));
        return;
    }
    long iters = 0;
    if (role == "solver" || role == "meta_solver") {
        list.set_if_missing("abs_tolerance", num(1e-8));
        list.set_if_missing("rel_tolerance", num(1e-8));
        list.set_if_missing("convergence_details", "check-L2,print");
    } else {
        list.set_if_missing("abs_tolerance", num(0));
        list.set_if_missing("rel_tolerance", num(0));
        list.set_if_missing("convergence_details", "nocheck,noprint");
    }
    if (method == "Jacobi" || method == "HSGS") {
        iters = (role == "solver") ? 100 : 2;
        list.set_if_missing("relax_factor", num(0.99));
    } else if (method == "MultiGrid") {
        iters = (role == "solver") ? 30 : 1;
        list.set_if_missing("hypre_log", "1");
        list.set_if_missing("hypre_per_level_hierarchy", "0");
        list.set_if_missing("mg_reduced_precision", "0");
        list.set_if_missing("mg_cycle", "0");
        list.set_if_missing("mg_max_levels", "25");
        list.set_if_missing("mg_coarse_matrix_size", "100");
        list.set_if_missing("mg_strength_threshold", num(0.5));
        list.set_if_missing("mg_trunc_factor", num(0.25));
        list.set_if_missing("mg_num_paths", "3");
        list.set_if_missing("mg_coarsening_type", "6");
        list.set_if_missing("mg_interpolation_type", "0");
        list.set_if_missing("mg_max_row_sum", num(0.9));
        list.set_if_missing("mg_Pmax_elements", "4");
        list.set_if_missing("mg_nonGalerkin_tol", num(0.0));
        list.set_if_missing("mg_agg_num_levels", "10");
        list.set_if_missing("mg_agg_interpolation_type", "4");
        list.set_if_missing("mg_agg_trunc_factor", num(0.25));
        list.set_if_missing("mg_agg_P12_trunc_factor", num(0.0));
        list.set_if_missing("mg_agg_Pmax_elements", "4");
        list.set_if_missing("mg_agg_P12max_elements", "4");
    } else if (method == "IterativeRefinement") {
        iters = 20;
    } else if (method == "PCG" || method == "BiCGStab" || method == "MergedBiCGStab" ||
               method == "PBiCGStab" || method == "MergedPBiCGStab" || method == "RBiCGStab" ||
               method == "MergedRBiCGStab") {
        iters = (role == "solver") ? 50 : 1;
    } else {
        throw std::runtime_error("xamg: method '" + method + "' not found");
    }
    list.set_if_missing("max_iters", std::to_string(iters));
}

// xamg_params_details.h:142-169 dictionary defaults, then every list's family defaults.
void GlobalParams::set_defaults() {
    bool mg = false;
    for (auto &r : roles)
        if (r.second.has("method") && r.second.get_string("method") == "MultiGrid")
            mg = true;
    if (mg) {
        if (!has("pre_smoother"))
            roles["pre_smoother"].set("method", "Jacobi");
        if (!has("post_smoother"))
            roles["post_smoother"].set("method", "Jacobi");
        if (!has("coarse_grid_solver"))
            roles["coarse_grid_solver"].set("method", "Direct");
    }
    if (has("meta_solver"))
        roles["meta_solver"].set_if_missing("convergence_details", "check-L2,print");
    if (has("solver")) {
        auto &l = roles["solver"];
        l.set_if_missing("convergence_details", "check-L2,print");
        std::set<std::string> cd;
        split(l.get_string("convergence_details"), ',', cd);
        if (cd.count("RR") && !has("residual_control"))
            throw std::runtime_error(
                "xamg: residual_control parameters block must present if RR convergence mode is used");
    }
    for (auto &r : roles) {
        if (!r.second.has("method"))
            continue;
        set_family_defaults(r.second, r.second.get_string("method"), r.first);
    }
}

void GlobalParams::print(FILE *f) const {
    for (const auto &r : roles) {
        fprintf(f, "%s:\n", r.first.c_str());
        for (const auto &kv : r.second.kv)
            fprintf(f, "  %s: %s\n", kv.first.c_str(), kv.second.c_str());
        auto it = overrides.find(r.first);
        if (it != overrides.end())
            for (const auto &o : it->second)
                fprintf(f, "  %s: %s@lev%d-%d\n", o.key.c_str(), o.value.c_str(), o.lev_lo,
                        o.lev_hi);
    }
}

Override parse_override(const std::string &key, const std::string &val) {
    Override o;
    o.key = key;
    auto at = val.find("@lev");
    if (at == std::string::npos) {
        o.value = val;
        return o;
    }
    o.value = val.substr(0, at);
    std::string r = val.substr(at + 4);
    auto dash = r.find('-');
    if (dash == std::string::npos) {
        o.lev_lo = o.lev_hi = std::atoi(r.c_str());
    } else {
        o.lev_lo = std::atoi(r.substr(0, dash).c_str());
        o.lev_hi = std::atoi(r.substr(dash + 1).c_str());
    }
    return o;
}

} // namespace xamg
