// config.cpp -- config-file parsing with the reference's grammar.
//
// GameConfigData::setConfigDataFromFile (game_configdata.cpp:22-54): every
// space is removed from each line, empty lines are skipped, a line is
// KEY=VALUE split at the first '='; WATER= and OBJECT= take
// path;scale;x;y;z (at least five fields, else an error).  Lines without '='
// are ignored.  getInt/getFloat (game_configdata.cpp:56-83) are std::stoi /
// std::stof of the stored string; a missing key is an error.  The SimConfig
// fields are read as in Simulation::initDeviceParams (Simulation.cpp:56-77).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mpm.h"

namespace mpm {
void set_last_error(const std::string& msg);
}

namespace {

struct Parsed {
    std::unordered_map<std::string, std::string> kv;
    std::vector<mpm_config_object> waters, objects;
};

std::vector<std::string> split(const std::string& s, char delim) {
    std::vector<std::string> out;
    std::string tok;
    std::istringstream is(s);
    while (std::getline(is, tok, delim)) out.push_back(tok);
    return out;
}

bool parse_text(const std::string& text, Parsed& p, std::string& err) {
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) {
        line.erase(std::remove(line.begin(), line.end(), ' '), line.end());
        if (line.empty()) continue;
        const size_t pos = line.find('=');
        if (pos == std::string::npos) continue;
        const std::string key = line.substr(0, pos), value = line.substr(pos + 1);
        if (key == "WATER" || key == "OBJECT") {
            std::vector<std::string> parts = split(value, ';');
            if (parts.size() < 5) {
                err = "malformed " + key + " entry: " + value;
                return false;
            }
            mpm_config_object o{};
            std::strncpy(o.model_path, parts[0].c_str(), sizeof(o.model_path) - 1);
            try {
                o.scale = std::stof(parts[1]);
                o.x = std::stof(parts[2]);
                o.y = std::stof(parts[3]);
                o.z = std::stof(parts[4]);
            } catch (const std::exception&) {
                err = "malformed number in " + key + " entry: " + value;
                return false;
            }
            (key == "WATER" ? p.waters : p.objects).push_back(o);
        } else {
            p.kv[key] = value;
        }
    }
    return true;
}

bool get_int(const Parsed& p, const char* key, int64_t& out, std::string& err) {
    auto it = p.kv.find(key);
    if (it == p.kv.end()) {
        err = std::string("missing config key ") + key;
        return false;
    }
    try {
        out = std::stoi(it->second);
    } catch (const std::exception&) {
        err = std::string("config key ") + key + " is not an integer: " + it->second;
        return false;
    }
    return true;
}

bool get_float(const Parsed& p, const char* key, float& out, std::string& err) {
    auto it = p.kv.find(key);
    if (it == p.kv.end()) {
        err = std::string("missing config key ") + key;
        return false;
    }
    try {
        out = std::stof(it->second);
    } catch (const std::exception&) {
        err = std::string("config key ") + key + " is not a number: " + it->second;
        return false;
    }
    return true;
}

int fill_config(const Parsed& p, mpm_sim_config* cfg, mpm_config_object* waters, int* n_waters,
                mpm_config_object* objects, int* n_objects, int cap) {
    std::string err;
    int64_t v;
    mpm_sim_config c{};
    const char* ints[] = {"SIZE_X", "SIZE_Y", "SIZE_Z", "PADDING"};
    uint32_t* int_dst[] = {&c.SIZE_X, &c.SIZE_Y, &c.SIZE_Z, &c.PADDING};
    for (int i = 0; i < 4; ++i) {
        if (!get_int(p, ints[i], v, err)) { mpm::set_last_error(err); return MPM_EIO; }
        *int_dst[i] = (uint32_t)v;
    }
    const char* floats[] = {"GRAVITY", "DT", "GAMMA", "COMPRESSION", "RESOLUTION"};
    float* flt_dst[] = {&c.GRAVITY, &c.DT, &c.GAMMA, &c.COMPRESSION, &c.RESOLUTION};
    for (int i = 0; i < 5; ++i)
        if (!get_float(p, floats[i], *flt_dst[i], err)) { mpm::set_last_error(err); return MPM_EIO; }
    if (!get_int(p, "SUBSTEPS", v, err)) { mpm::set_last_error(err); return MPM_EIO; }
    c.SUBSTEPS = (uint32_t)v;
    if (!get_int(p, "SDF_RESOLUTION", v, err)) { mpm::set_last_error(err); return MPM_EIO; }
    c.SDF_RESOLUTION = (uint32_t)v;
    mpm_config_derive(&c);
    if (cfg) *cfg = c;
    if (n_waters) *n_waters = (int)p.waters.size();
    if (n_objects) *n_objects = (int)p.objects.size();
    for (int i = 0; waters && i < (int)p.waters.size() && i < cap; ++i) waters[i] = p.waters[i];
    for (int i = 0; objects && i < (int)p.objects.size() && i < cap; ++i) objects[i] = p.objects[i];
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_config_parse_string(const char* text, mpm_sim_config* cfg, mpm_config_object* waters, int* n_waters,
                            mpm_config_object* objects, int* n_objects, int cap) {
    if (!text) { mpm::set_last_error("null text"); return MPM_EINVAL; }
    Parsed p;
    std::string err;
    if (!parse_text(text, p, err)) { mpm::set_last_error(err); return MPM_EIO; }
    return fill_config(p, cfg, waters, n_waters, objects, n_objects, cap);
}

int mpm_config_parse_file(const char* path, mpm_sim_config* cfg, mpm_config_object* waters, int* n_waters,
                          mpm_config_object* objects, int* n_objects, int cap) {
    if (!path) { mpm::set_last_error("null path"); return MPM_EINVAL; }
    std::ifstream f(path);
    if (!f) { mpm::set_last_error(std::string("cannot open ") + path); return MPM_EIO; }
    std::stringstream ss;
    ss << f.rdbuf();
    return mpm_config_parse_string(ss.str().c_str(), cfg, waters, n_waters, objects, n_objects, cap);
}

int mpm_config_get_int(const char* text, const char* key, int64_t* value) {
    if (!text || !key || !value) { mpm::set_last_error("null argument"); return MPM_EINVAL; }
    Parsed p;
    std::string err;
    if (!parse_text(text, p, err) || !get_int(p, key, *value, err)) { mpm::set_last_error(err); return MPM_EIO; }
    return MPM_OK;
}

}  // extern "C"
