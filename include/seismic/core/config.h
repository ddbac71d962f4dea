// Run parameters and the YAML loader (drop-in for src/cpp/core/config.h:9-51).
// Every key of config/default.yaml is parsed, including the ones the
// reference's load_config skips; new keys have defaults that leave old files
// meaning what they meant.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace seismic {

struct SolverConfig {         // config.h:9-19
    double frequency;
    double dt;
    int nt;
    double water_depth;
    double seafloor_gradient;
    int sponge_width;
    double sponge_damping;
    int snapshot_interval;
    std::string output_dir;
    // new: spatial order of the FD stencil (2 = the reference's 7-point
    // stencil, 8 = radius-4 star); 0 = 8 when the grid's halo width is >= 4,
    // else 2
    int fd_order = 0;
};

struct ModelConfig {          // config.h:21-31
    double vp_min;
    double vp_max;
    double vs_min;
    double vs_max;
    double rho_min;
    double rho_max;
    double initial_vp;
    double initial_vs;
    double initial_rho;
};

struct InversionConfig {      // config.h:33-43
    int max_iterations;
    double step_length;
    double gradient_smoothing_sigma;
    double regularization_lambda;
    double misfit_tolerance;
    std::string misfit_type;
    int gradient_precondition;
    int line_search_max;
    double line_search_factor;
};

struct TomographyConfig {     // config.h:45-51
    SolverConfig solver;
    ModelConfig model;
    InversionConfig inversion;
    std::map<std::string, std::string> obs_stations;
    std::vector<double> source_positions;
};

// Defaults of config.cpp:12-40 when the file cannot be opened.  Lines are
// "section:" or "key: value"; blank lines and lines starting with '#' are
// skipped, a '#' after a value starts a comment, unknown keys are ignored.
// Numbers that do not parse throw std::invalid_argument, as the reference's
// std::stod / std::stoi do.
TomographyConfig load_config(const std::string& yaml_path);

}  // namespace seismic
