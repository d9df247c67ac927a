// config.cpp — CLIP's config.txt surface (host only).
//
// Reads the same `key = value` file as the reference and fills clip_config:
//   * scalar / array keys          InputData::read_value / read_array (InputData.cu:144-320)
//   * derived parameters           InputData::read_config           (InputData.cu:57-115)
//   * `boundary = [ {...} ]` list  Boundary::readBoundaries          (Boundary.cu:146-279)
//   * `geometry = [ {...} ]` list  Geometry::readGeometries          (Geometry.cu:72-182)
//
// Behaviour kept from the reference: first matching key wins, '#' starts a
// comment line, scalar values are read with operator>> (so trailing comments
// are ignored), arrays lose their brackets and are split on ',' with missing
// entries set to 0, block lines have all whitespace removed.
// Deliberate differences (each turns reference UB or a silent no-op into a
// defined result): a missing required key is an error instead of an
// uninitialised value; multi-word boundary types ("free convect", "slip wall",
// "do nothing") are accepted although the reference's whitespace stripping
// makes them parse as Unknown (Boundary.cu:282-288 vs 44-60); an unknown
// boundary type is an error; single-line blocks `{ side = "x-", type = ... }`
// (the form shown in the reference README) are accepted; a side's type is
// bound to its side at the end of the block, whatever the key order.
#include "clip_amd.h"
#include "common.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace clipamd {

namespace {

std::string trim_ends(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return std::string();
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

std::string strip_all_space(const std::string& s) {
    std::string r;
    r.reserve(s.size());
    for (char c : s)
        if (c != ' ' && c != '\t' && c != '\r') r.push_back(c);
    return r;
}

std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct KeyValueFile {
    std::vector<std::string> lines;

    // First line `key = value` whose trimmed key equals `key` (InputData.cu:294-317).
    bool find(const std::string& key, std::string& value) const {
        for (const auto& line : lines) {
            if (line.empty() || line[0] == '#') continue;
            size_t pos = line.find('=');
            if (pos == std::string::npos) continue;
            if (trim_ends(line.substr(0, pos)) == key) {
                value = trim_ends(line.substr(pos + 1));
                return true;
            }
        }
        return false;
    }
    bool scalar(const std::string& key, double& out) const {
        std::string v;
        if (!find(key, v)) return false;
        std::istringstream is(v);
        double x;
        if (!(is >> x)) throw ConfigError("cannot read a number for key '" + key + "': " + v);
        out = x;
        return true;
    }
    bool scalar(const std::string& key, unsigned& out) const {
        double x;
        if (!scalar(key, x)) return false;
        if (x < 0) throw ConfigError("key '" + key + "' must be non-negative");
        out = (unsigned)x;
        return true;
    }
    bool word(const std::string& key, std::string& out) const {
        std::string v;
        if (!find(key, v)) return false;
        std::istringstream is(v);
        is >> out;
        return true;
    }
    // Bracketed list, up to n entries, the rest set to 0 (InputData.cu:168-205).
    bool array(const std::string& key, double* out, int n) const {
        std::string v;
        if (!find(key, v)) return false;
        for (int i = 0; i < n; ++i) out[i] = 0.0;
        v.erase(std::remove(v.begin(), v.end(), '['), v.end());
        v.erase(std::remove(v.begin(), v.end(), ']'), v.end());
        std::stringstream ss(v);
        std::string tok;
        int count = 0;
        while (count < n && std::getline(ss, tok, ',')) {
            tok = trim_ends(tok);
            if (tok.empty()) continue;
            try {
                out[count++] = std::stod(tok);
            } catch (...) {
                throw ConfigError("cannot read entry " + std::to_string(count) + " of '" + key + "'");
            }
        }
        return true;
    }
};

int case_from_string(const std::string& s) {
    std::string l = lower(s);
    if (l == "drop") return CLIP_CASE_DROP;
    if (l == "bubble") return CLIP_CASE_BUBBLE;
    if (l == "jet") return CLIP_CASE_JET;
    if (l == "rti") return CLIP_CASE_RTI;
    throw ConfigError("Unknown case type: " + s);
}

int side_from_string(const std::string& s) {
    std::string l = lower(s);
    if (l == "x-") return CLIP_X_MINUS;
    if (l == "x+") return CLIP_X_PLUS;
    if (l == "y-") return CLIP_Y_MINUS;
    if (l == "y+") return CLIP_Y_PLUS;
    if (l == "z-") return CLIP_Z_MINUS;
    if (l == "z+") return CLIP_Z_PLUS;
    return -1;
}

int bc_from_string(const std::string& s) {
    std::string l = strip_all_space(lower(s));
    if (l == "wall") return CLIP_BC_WALL;
    if (l == "slipwall") return CLIP_BC_SLIP_WALL;
    if (l == "freeconvect") return CLIP_BC_FREE_CONVECT;
    if (l == "periodic") return CLIP_BC_PERIODIC;
    if (l == "neumann") return CLIP_BC_NEUMANN;
    if (l == "velocity") return CLIP_BC_VELOCITY;
    if (l == "donothing") return CLIP_BC_DO_NOTHING;
    return CLIP_BC_UNKNOWN;
}

int geom_from_string(const std::string& s) {
    std::string l = lower(s);
    if (l == "circle") return CLIP_GEOM_CIRCLE;
    if (l == "sphere") return CLIP_GEOM_SPHERE;
    if (l == "square") return CLIP_GEOM_SQUARE;
    if (l == "cube") return CLIP_GEOM_CUBE;
    if (l == "perturbation") return CLIP_GEOM_PERTURBATION;
    return CLIP_GEOM_UNKNOWN;
}

void read_vec3(std::string v, double out[3]) {
    if (!v.empty() && v.front() == '[') v.erase(0, 1);
    if (!v.empty() && v.back() == ']') v.pop_back();
    std::stringstream ss(v);
    std::string tok;
    int d = 0;
    while (d < 3 && std::getline(ss, tok, ',')) out[d++] = std::stod(tok);
}

// Splits a whitespace-free block body `k1=v1,k2=[a,b,c],k3=v3` at top-level commas.
std::vector<std::string> split_block_items(const std::string& body) {
    std::vector<std::string> items;
    std::string cur;
    int depth = 0;
    for (char c : body) {
        if (c == '[') depth++;
        if (c == ']') depth--;
        if (c == ',' && depth == 0) {
            if (!cur.empty()) items.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) items.push_back(cur);
    return items;
}

// Collects the key=value items of each `{ ... }` block of list `name`
// (Boundary.cu:167-233, Geometry.cu:86-176).  Lines are stripped of all
// whitespace; a block may span lines or sit on one line.
std::vector<std::vector<std::pair<std::string, std::string>>> read_blocks(
    const std::vector<std::string>& raw_lines, const std::string& name) {
    std::vector<std::vector<std::pair<std::string, std::string>>> blocks;
    bool in_list = false, in_block = false;
    std::string body;
    for (const auto& raw : raw_lines) {
        std::string line = strip_all_space(raw);
        if (line.empty() || line[0] == '#') continue;
        if (!in_list) {
            size_t eq = line.find('=');
            if (eq != std::string::npos && line.substr(0, eq) == name) {
                in_list = true;
                line = line.substr(eq + 1);
                if (!line.empty() && line[0] == '[') line = line.substr(1);
                if (line.empty()) continue;
            } else {
                continue;
            }
        }
        size_t pos = 0;
        while (pos < line.size()) {
            char c = line[pos];
            if (!in_block) {
                if (c == '{') {
                    in_block = true;
                    body.clear();
                } else if (c == ']') {
                    return blocks;
                }
                pos++;
                continue;
            }
            if (c == '}') {
                std::vector<std::pair<std::string, std::string>> kv;
                for (const auto& item : split_block_items(body)) {
                    size_t eq = item.find('=');
                    if (eq == std::string::npos) continue;
                    std::string k = item.substr(0, eq), v = item.substr(eq + 1);
                    v.erase(std::remove(v.begin(), v.end(), '"'), v.end());
                    kv.emplace_back(k, v);
                }
                blocks.push_back(kv);
                in_block = false;
                pos++;
                continue;
            }
            body.push_back(c);
            pos++;
        }
        // A line break inside a block separates items like a comma.
        if (in_block && !body.empty() && body.back() != ',' && body.back() != '{') body.push_back(',');
    }
    if (in_block) throw ConfigError("unterminated block in list '" + name + "'");
    return blocks;
}

void parse(const std::vector<std::string>& lines, int dim_override, clip_config* cfg) {
    std::memset(cfg, 0, sizeof(*cfg));
    KeyValueFile kv{lines};

    std::string cs;
    if (!kv.word("case", cs)) throw ConfigError("missing key 'case'");
    cfg->case_type = case_from_string(cs);

    kv.scalar("tFinal", cfg->t_final);
    kv.scalar("finalStep", cfg->final_step);
    kv.scalar("outputInterval", cfg->output_interval);
    kv.scalar("checkpointInterval", cfg->checkpoint_interval);
    kv.scalar("checkpointCopy", cfg->checkpoint_copy);
    kv.scalar("reportInterval", cfg->report_interval);
    if (!kv.array("N", cfg->N, 3)) throw ConfigError("missing key 'N'");
    if (!kv.scalar("referenceLength", cfg->reference_length)) throw ConfigError("missing key 'referenceLength'");
    if (!kv.scalar("interfaceWidth", cfg->interface_width)) throw ConfigError("missing key 'interfaceWidth'");
    kv.array("gravity", cfg->gravity, 3);
    bool has_mobility = kv.scalar("mobility", cfg->mobility);
    if (!kv.scalar("muRatio", cfg->mu_ratio)) throw ConfigError("missing key 'muRatio'");
    if (!kv.scalar("rhoRatio", cfg->rho_ratio)) throw ConfigError("missing key 'rhoRatio'");
    kv.scalar("amplitude", cfg->amplitude);
    kv.scalar("referenceVelocity", cfg->reference_velocity);
    kv.scalar("Bo", cfg->Bo);
    kv.scalar("Mo", cfg->Mo);

    switch (cfg->case_type) {
        case CLIP_CASE_BUBBLE:
        case CLIP_CASE_DROP:
            if (!kv.scalar("We", cfg->We)) throw ConfigError("missing key 'We'");
            if (!kv.scalar("Re", cfg->Re)) throw ConfigError("missing key 'Re'");
            if (!has_mobility) throw ConfigError("missing key 'mobility'");
            break;
        case CLIP_CASE_RTI:
            if (!kv.scalar("Ca", cfg->Ca)) throw ConfigError("missing key 'Ca'");
            if (!kv.scalar("Pe", cfg->Pe)) throw ConfigError("missing key 'Pe'");
            if (!kv.scalar("Re", cfg->Re)) throw ConfigError("missing key 'Re'");
            break;
        case CLIP_CASE_JET:
            if (!kv.scalar("We", cfg->We)) throw ConfigError("missing key 'We'");
            if (!kv.scalar("Re", cfg->Re)) throw ConfigError("missing key 'Re'");
            if (!kv.scalar("referenceVelocity", cfg->reference_velocity))
                throw ConfigError("missing key 'referenceVelocity'");
            if (!has_mobility) throw ConfigError("missing key 'mobility'");
            break;
    }

    // Explicit overrides (no reference counterpart): let a config state sigma
    // and mu_H directly, or give the |g| used only for the derivation.
    cfg->has_sigma = kv.scalar("sigma", cfg->sigma_override);
    cfg->has_mu_h = kv.scalar("muH", cfg->mu_h_override);
    cfg->has_gravity_scale = kv.scalar("gravityScale", cfg->gravity_scale);

    if (dim_override == 2 || dim_override == 3) {
        cfg->dim = dim_override;
    } else {
        double d = 0;
        if (kv.scalar("dim", d)) cfg->dim = (int)d;
        else cfg->dim = (cfg->N[2] <= 1.0) ? 2 : 3;
    }
    if (cfg->dim != 2 && cfg->dim != 3) throw ConfigError("dim must be 2 or 3");

    // Inlet modulation extension.
    double nm = 0;
    if (kv.scalar("inletModes", nm)) {
        cfg->inlet_modes = (int)nm;
        if (cfg->inlet_modes < 0 || cfg->inlet_modes > CLIP_MAX_INLET_MODES)
            throw ConfigError("inletModes must be in [0, 8]");
        kv.array("inletModeAmplitude", cfg->inlet_mode_amp, CLIP_MAX_INLET_MODES);
        kv.array("inletModePhase", cfg->inlet_mode_phase, CLIP_MAX_INLET_MODES);
    }

    // Slip-side behaviour (extension): the reference's (default) or specular.
    std::string slip;
    if (kv.word("slipWall", slip)) {
        const std::string l = lower(slip);
        if (l == "specular") cfg->slip_specular = 1;
        else if (l != "reference") throw ConfigError("slipWall must be 'reference' or 'specular'");
    }

    // Boundary blocks.
    for (int s = 0; s < CLIP_NUM_SIDES; ++s) cfg->bc_type[s] = CLIP_BC_UNKNOWN;
    auto bblocks = read_blocks(lines, "boundary");
    for (const auto& blk : bblocks) {
        int side = -1, type = CLIP_BC_UNKNOWN;
        double val[3] = {0, 0, 0};
        bool has_type = false;
        for (const auto& it : blk) {
            if (it.first == "side") side = side_from_string(it.second);
            else if (it.first == "type") { type = bc_from_string(it.second); has_type = true; }
            else if (it.first == "value") read_vec3(it.second, val);
        }
        if (side < 0) throw ConfigError("boundary block without a valid side");
        if (!has_type || type == CLIP_BC_UNKNOWN)
            throw ConfigError("boundary block for side " + std::to_string(side) + " has an unknown type");
        cfg->bc_type[side] = type;
        cfg->bc_present[side] = 1;
        for (int d = 0; d < 3; ++d) cfg->bc_value[side][d] = val[d];
    }
    static const char* names[6] = {"x-", "x+", "y-", "y+", "z-", "z+"};
    std::string missing;
    for (int s = 0; s < 2 * cfg->dim; ++s)
        if (!cfg->bc_present[s]) missing += (missing.empty() ? "" : ", ") + std::string(names[s]);
    if (!missing.empty()) throw ConfigError("Missing boundary conditions for sides: " + missing);
    // Sides beyond the dimension carry no condition.
    for (int s = 2 * cfg->dim; s < 6; ++s) cfg->bc_type[s] = CLIP_BC_UNKNOWN;

    // Geometry blocks.
    auto gblocks = read_blocks(lines, "geometry");
    if ((int)gblocks.size() > CLIP_MAX_GEOMETRIES) throw ConfigError("more than 16 geometry objects");
    cfg->num_geometries = (int)gblocks.size();
    for (size_t g = 0; g < gblocks.size(); ++g) {
        cfg->geom_type[g] = CLIP_GEOM_UNKNOWN;
        cfg->geom_id[g] = -1;
        for (const auto& it : gblocks[g]) {
            if (it.first == "type") cfg->geom_type[g] = geom_from_string(it.second);
            else if (it.first == "center") read_vec3(it.second, cfg->geom_center[g]);
            else if (it.first == "length") read_vec3(it.second, cfg->geom_length[g]);
            else if (it.first == "radius") cfg->geom_radius[g] = std::stod(it.second);
            else if (it.first == "amplitude") cfg->geom_amplitude[g] = std::stod(it.second);
            else if (it.first == "id") cfg->geom_id[g] = std::stoi(it.second);
        }
        if (cfg->geom_type[g] == CLIP_GEOM_UNKNOWN)
            throw ConfigError("Encountered a geometry with unknown type during parsing.");
    }
}

}  // namespace

void derive(clip_config* c) {
    if (c->dim != 2 && c->dim != 3) throw ConfigError("dim must be 2 or 3");
    if (!(c->rho_ratio > 0)) throw ConfigError("rhoRatio must be > 0");
    if (!(c->mu_ratio > 0)) throw ConfigError("muRatio must be > 0");
    if (!(c->interface_width > 0)) throw ConfigError("interfaceWidth must be > 0");
    for (int d = 0; d < c->dim; ++d)
        if (!(c->N[d] >= (c->dim == 2 ? 1.0 : 3.0)))
            throw ConfigError("N is too small along dimension " + std::to_string(d));
    if (c->checkpoint_copy != 0 && c->checkpoint_interval == 0)
        throw ConfigError("checkpointInterval must be > 0 when checkpointCopy != 0");

    // InputData.cu:76-114.
    c->rho_h = 1.0;
    c->rho_l = c->rho_h / c->rho_ratio;
    const double L = c->reference_length;
    const double gy = c->has_gravity_scale ? std::fabs(c->gravity_scale) : std::fabs(c->gravity[1]);
    if (c->case_type == CLIP_CASE_BUBBLE || c->case_type == CLIP_CASE_DROP) {
        c->sigma = (c->rho_h * gy * L * L) / c->We;
        c->mu_h = std::sqrt(gy * c->rho_h * (c->rho_h - c->rho_l) * L * L * L) / c->Re;
    } else if (c->case_type == CLIP_CASE_RTI) {
        c->mu_h = (c->rho_h * std::sqrt(gy * L) * L) / c->Re;
        c->sigma = (c->mu_h * std::sqrt(gy * L)) / c->Ca;
        c->mobility = (std::sqrt(gy * L) * L) / c->Pe;
    } else if (c->case_type == CLIP_CASE_JET) {
        c->sigma = (c->rho_h * c->reference_velocity * c->reference_velocity * L) / c->We;
        c->mu_h = (c->rho_h * c->reference_velocity * L) / c->Re;
    }
    if (c->has_mobility_override) c->mobility = c->mobility_override;
    if (c->has_sigma) c->sigma = c->sigma_override;
    if (c->has_mu_h) c->mu_h = c->mu_h_override;
    c->mu_l = c->mu_h / c->mu_ratio;
    c->tau_h = 3.0 * (c->mu_h / c->rho_h);
    c->tau_l = 3.0 * (c->mu_l / c->rho_l);
    c->kappa = 1.50 * c->sigma * c->interface_width;
    c->beta = 8.0 * c->sigma / c->interface_width;
    if (!std::isfinite(c->sigma) || !std::isfinite(c->mu_h) || !std::isfinite(c->mobility))
        throw ConfigError("derived parameters are not finite (check We/Re/Ca/Pe and gravity)");
    for (int s = 0; s < 2 * c->dim; ++s)
        if (c->bc_type[s] < 0 || c->bc_type[s] >= CLIP_BC_UNKNOWN)
            throw ConfigError("side " + std::to_string(s) + " has no valid boundary type");
}

void parse_text(const std::string& text, int dim_override, clip_config* cfg) {
    std::vector<std::string> lines;
    std::stringstream ss(text);
    std::string line;
    while (std::getline(ss, line)) lines.push_back(line);
    parse(lines, dim_override, cfg);
    derive(cfg);
}

}  // namespace clipamd

extern "C" int clip_config_from_string(const char* text, int dim_override, clip_config* out) {
    if (!text || !out) return clipamd::set_error(CLIP_ERR_ARG, "null argument");
    try {
        clipamd::parse_text(text, dim_override, out);
    } catch (const clipamd::ConfigError& e) {
        return clipamd::set_error(CLIP_ERR_PARSE, e.what());
    } catch (const std::exception& e) {
        return clipamd::set_error(CLIP_ERR_PARSE, std::string("config: ") + e.what());
    }
    return CLIP_OK;
}

extern "C" int clip_config_from_file(const char* path, int dim_override, clip_config* out) {
    if (!path || !out) return clipamd::set_error(CLIP_ERR_ARG, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) return clipamd::set_error(CLIP_ERR_IO, std::string("Error opening config file: ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return clip_config_from_string(ss.str().c_str(), dim_override, out);
}

extern "C" int clip_config_derive(clip_config* cfg) {
    if (!cfg) return clipamd::set_error(CLIP_ERR_ARG, "null argument");
    try {
        clipamd::derive(cfg);
    } catch (const std::exception& e) {
        return clipamd::set_error(CLIP_ERR_ARG, e.what());
    }
    return CLIP_OK;
}
