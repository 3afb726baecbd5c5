// Parameter dictionary behind XAMG_param_* (src/param/params.h + xamg_params_details.h).
// Role -> method -> key=value with the reference's per-role, per-method defaults and
// per-MG-level overrides ("value@levX-Y").
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>

namespace xamg {

struct ParamList {
    std::map<std::string, std::string> kv;
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    std::string get_string(const std::string &k) const;
    long get_int(const std::string &k) const;
    double get_float(const std::string &k) const;
    bool get_bool(const std::string &k) const;
    void set(const std::string &k, const std::string &v) { kv[k] = v; }
    void set_if_missing(const std::string &k, const std::string &v) {
        if (!has(k))
            kv[k] = v;
    }
};

struct Override {
    std::string key, value;
    int lev_lo = 0, lev_hi = 1 << 30;
};

struct GlobalParams {
    std::map<std::string, ParamList> roles;
    std::map<std::string, std::vector<Override>> overrides; // role -> per-level overrides
    bool has(const std::string &role) const { return roles.count(role) != 0; }
    ParamList &get(const std::string &role);
    const ParamList &get(const std::string &role) const;
    // role list with the overrides of MG level `lev` applied (params.get(role, lev))
    ParamList get_level(const std::string &role, int lev) const;
    void set_defaults();
    void print(FILE *f) const;
};

// Family defaults of one list (xamg_params_details.h:171-289).
void set_family_defaults(ParamList &list, const std::string &method, const std::string &role);

// Parse "val@levA-B" / "val@levA" override syntax (cmdline.h:102-120).
Override parse_override(const std::string &key, const std::string &val);

} // namespace xamg
