// load_config (reference interface: src/cpp/core/config.h:53, defaults as in
// config.cpp:12-40).  Parses every key of config/default.yaml.
#include "core/config.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace seismic {

namespace {

std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    return s.substr(b, s.find_last_not_of(ws) - b + 1);
}

double to_d(const std::string& v) { return std::stod(v); }
int to_i(const std::string& v) { return std::stoi(v); }

TomographyConfig defaults() {
    TomographyConfig c;
    c.solver = SolverConfig{2.0, 0.001, 5000, 4000.0, 0.05, 20, 0.015, 100, "./output", 0};
    c.model = ModelConfig{1450.0, 8500.0, 0.0, 5000.0, 1000.0, 3300.0, 1500.0, 0.0, 2200.0};
    c.inversion = InversionConfig{50, 0.01, 2.0, 1e-4, 1e-6, "cross_correlation", 1, 10, 0.5};
    return c;
}

std::vector<double> parse_list(const std::string& v) {
    std::string s = v;
    if (!s.empty() && s.front() == '[') s.erase(0, 1);
    if (!s.empty() && s.back() == ']') s.pop_back();
    std::vector<double> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!trim(item).empty()) out.push_back(to_d(trim(item)));
    return out;
}

bool set_solver(SolverConfig& s, const std::string& k, const std::string& v) {
    if (k == "frequency") s.frequency = to_d(v);
    else if (k == "dt") s.dt = to_d(v);
    else if (k == "nt") s.nt = to_i(v);
    else if (k == "water_depth") s.water_depth = to_d(v);
    else if (k == "seafloor_gradient") s.seafloor_gradient = to_d(v);
    else if (k == "sponge_width") s.sponge_width = to_i(v);
    else if (k == "sponge_damping") s.sponge_damping = to_d(v);
    else if (k == "snapshot_interval") s.snapshot_interval = to_i(v);
    else if (k == "output_dir") s.output_dir = v;
    else if (k == "fd_order") s.fd_order = to_i(v);
    else return false;
    return true;
}

bool set_model(ModelConfig& m, const std::string& k, const std::string& v) {
    double* f = k == "vp_min" ? &m.vp_min : k == "vp_max" ? &m.vp_max : k == "vs_min" ? &m.vs_min
              : k == "vs_max" ? &m.vs_max : k == "rho_min" ? &m.rho_min : k == "rho_max" ? &m.rho_max
              : k == "initial_vp" ? &m.initial_vp : k == "initial_vs" ? &m.initial_vs
              : k == "initial_rho" ? &m.initial_rho : nullptr;
    if (!f) return false;
    *f = to_d(v);
    return true;
}

bool set_inversion(InversionConfig& c, const std::string& k, const std::string& v) {
    if (k == "max_iterations") c.max_iterations = to_i(v);
    else if (k == "step_length") c.step_length = to_d(v);
    else if (k == "gradient_smoothing_sigma") c.gradient_smoothing_sigma = to_d(v);
    else if (k == "regularization_lambda") c.regularization_lambda = to_d(v);
    else if (k == "misfit_tolerance") c.misfit_tolerance = to_d(v);
    else if (k == "misfit_type") c.misfit_type = v;
    else if (k == "gradient_precondition") c.gradient_precondition = to_i(v);
    else if (k == "line_search_max") c.line_search_max = to_i(v);
    else if (k == "line_search_factor") c.line_search_factor = to_d(v);
    else return false;
    return true;
}

}  // namespace

TomographyConfig load_config(const std::string& yaml_path) {
    TomographyConfig cfg = defaults();
    std::ifstream in(yaml_path);
    if (!in) return cfg;
    std::string raw, section;
    while (std::getline(in, raw)) {
        std::string line = trim(raw);
        if (line.empty() || line[0] == '#') continue;
        const size_t hash = line.find(" #");
        if (hash != std::string::npos) line = trim(line.substr(0, hash));
        if (line.back() == ':') {
            section = trim(line.substr(0, line.size() - 1));
            continue;
        }
        const size_t colon = line.find(':');
        if (colon == std::string::npos) continue;
        const std::string key = trim(line.substr(0, colon));
        const std::string val = trim(line.substr(colon + 1));
        if (key == "source_positions") {
            cfg.source_positions = parse_list(val);
        } else if (section == "solver") {
            set_solver(cfg.solver, key, val);
        } else if (section == "model") {
            set_model(cfg.model, key, val);
        } else if (section == "inversion") {
            set_inversion(cfg.inversion, key, val);
        } else if (section == "obs_stations") {
            cfg.obs_stations[key] = val;
        }
    }
    return cfg;
}

}  // namespace seismic
