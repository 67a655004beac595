// sim_info.cpp -- INI reading and readInifile (SimInfo.h:169-305).
#include "sim_info.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace fv3d {

static std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

static std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
  return s;
}

static std::string strip_inline_comment(const std::string& v) {
  for (size_t i = 1; i < v.size(); ++i)
    if (v[i] == ';' && std::isspace((unsigned char)v[i - 1])) return trim(v.substr(0, i));
  return v;
}

std::string IniReader::key(const std::string& section, const std::string& name) {
  return lower(section) + "=" + lower(name);
}

IniReader::IniReader(const std::string& filename) {
  std::ifstream f(filename);
  if (!f) {
    error_ = -1;
    return;
  }
  std::string line, section, prev;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const std::string s = trim(line);
    // a line starting with white space continues the previous value
    if (!prev.empty() && !line.empty() && std::isspace((unsigned char)line[0]) && !s.empty() &&
        s[0] != ';' && s[0] != '#') {
      values_[prev] += "\n" + strip_inline_comment(s);
      continue;
    }
    if (s.empty() || s[0] == ';' || s[0] == '#') continue;
    if (s[0] == '[') {
      const size_t end = s.find(']');
      if (end == std::string::npos) {
        if (!error_) error_ = lineno;
        continue;
      }
      section = trim(s.substr(1, end - 1));
      prev.clear();
      continue;
    }
    const size_t eq = s.find_first_of("=:");
    if (eq == std::string::npos) {
      if (!error_) error_ = lineno;
      continue;
    }
    const std::string name = trim(s.substr(0, eq));
    const std::string value = strip_inline_comment(trim(s.substr(eq + 1)));
    const std::string k = key(section, name);
    auto it = values_.find(k);
    if (it != values_.end() && !it->second.empty()) it->second += "\n" + value;
    else values_[k] = value;
    prev = k;
  }
}

std::string IniReader::Get(const std::string& section, const std::string& name,
                           const std::string& default_value) const {
  auto it = values_.find(key(section, name));
  return it == values_.end() ? default_value : it->second;
}

long IniReader::GetInteger(const std::string& section, const std::string& name,
                           long default_value) const {
  const std::string v = Get(section, name, "");
  const char* s = v.c_str();
  char* end = nullptr;
  errno = 0;
  const long n = std::strtol(s, &end, 0);
  return (end > s && *end == '\0' && errno == 0) ? n : default_value;
}

double IniReader::GetFloat(const std::string& section, const std::string& name,
                           double default_value) const {
  const std::string v = Get(section, name, "");
  const char* s = v.c_str();
  char* end = nullptr;
  const double d = std::strtod(s, &end);
  return (end > s && *end == '\0') ? d : default_value;
}

bool IniReader::GetBoolean(const std::string& section, const std::string& name,
                           bool default_value) const {
  const std::string v = lower(Get(section, name, ""));
  if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
  if (v == "false" || v == "no" || v == "off" || v == "0") return false;
  return default_value;
}

// std::map::operator[] in the reference: an unknown string becomes 0
template <typename M>
static int lookup(const M& m, const std::string& v, const char* what, RunConfig& cfg) {
  auto it = m.find(v);
  if (it == m.end()) {
    cfg.warnings.push_back(std::string("unknown ") + what + " '" + v +
                           "': the reference maps it to 0");
    return 0;
  }
  return it->second;
}

RunConfig readInifile(const std::string& filename) {
  IniReader reader(filename);
  RunConfig cfg;
  if (reader.ParseError())
    cfg.warnings.push_back("ini parse error " + std::to_string(reader.ParseError()) + " in '" +
                           filename + "': defaults used for what was not read");
  fv3d_params& res = cfg.params;
  fv3d_params_defaults(&res);

  res.Nx = (int)reader.GetInteger("mesh", "Nx", 32);
  res.Ny = (int)reader.GetInteger("mesh", "Ny", 32);
  res.Nz = (int)reader.GetInteger("mesh", "Nz", 32);
  res.Ng = (int)reader.GetInteger("mesh", "Nghosts", 2);
  res.xmin = reader.GetFloat("mesh", "xmin", 0.0);
  res.xmax = reader.GetFloat("mesh", "xmax", 1.0);
  res.ymin = reader.GetFloat("mesh", "ymin", 0.0);
  res.ymax = reader.GetFloat("mesh", "ymax", 1.0);
  res.zmin = reader.GetFloat("mesh", "zmin", 0.0);
  res.zmax = reader.GetFloat("mesh", "zmax", 1.0);
  fv3d_params_finalize(&res);

  res.tend = reader.GetFloat("run", "tend", 1.0);
  res.save_freq = reader.GetFloat("run", "save_freq", 1.0e-1);
  cfg.filename_out = reader.Get("run", "output_filename", "run");

  const std::map<std::string, int> bc_map{{"reflecting", FV3D_BC_REFLECTING},
                                          {"absorbing", FV3D_BC_ABSORBING},
                                          {"periodic", FV3D_BC_PERIODIC},
                                          {"C91", FV3D_BC_C91}};
  res.boundary_x = lookup(bc_map, reader.Get("run", "boundaries_x", "reflecting"), "boundaries_x", cfg);
  res.boundary_y = lookup(bc_map, reader.Get("run", "boundaries_y", "reflecting"), "boundaries_y", cfg);
  res.boundary_z = lookup(bc_map, reader.Get("run", "boundaries_z", "reflecting"), "boundaries_z", cfg);
  const std::map<std::string, int> recons_map{{"pcm", FV3D_PCM}, {"pcm_wb", FV3D_PCM_WB}, {"plm", FV3D_PLM}};
  res.reconstruction = lookup(recons_map, reader.Get("solvers", "reconstruction", "pcm"), "reconstruction", cfg);
  const std::map<std::string, int> riemann_map{{"hll", FV3D_HLL}, {"hllc", FV3D_HLLC}};
  res.riemann_solver = lookup(riemann_map, reader.Get("solvers", "riemann_solver", "hllc"), "riemann_solver", cfg);
  const std::map<std::string, int> ts_map{{"euler", FV3D_TS_EULER}, {"RK2", FV3D_TS_RK2}};
  res.time_stepping = lookup(ts_map, reader.Get("solvers", "time_stepping", "euler"), "time_stepping", cfg);
  res.CFL = reader.GetFloat("solvers", "CFL", 0.8);

  res.epsilon = reader.GetFloat("misc", "epsilon", 1.0e-6);
  res.gamma0 = reader.GetFloat("physics", "gamma0", 5.0 / 3.0);
  res.gravity = reader.GetBoolean("physics", "gravity", false);
  res.g = reader.GetFloat("physics", "g", 0.0);
  res.m1 = reader.GetFloat("polytrope", "m1", 1.0);
  res.theta1 = reader.GetFloat("polytrope", "theta1", 10.0);
  res.m2 = reader.GetFloat("polytrope", "m2", 1.0);
  res.theta2 = reader.GetFloat("polytrope", "theta2", 10.0);
  cfg.problem = reader.Get("physics", "problem", "blast");
  res.well_balanced_flux_at_z_bc = reader.GetBoolean("physics", "well_balanced_flux_at_z_bc", false);

  res.thermal_conductivity_active = reader.GetBoolean("thermal_conduction", "active", false);
  const std::map<std::string, int> tcm_map{{"constant", FV3D_TCM_CONSTANT}, {"B02", FV3D_TCM_B02}};
  res.thermal_conductivity_mode =
      lookup(tcm_map, reader.Get("thermal_conduction", "conductivity_mode", "constant"), "conductivity_mode", cfg);
  res.kappa = reader.GetFloat("thermal_conduction", "kappa", 0.0);
  const std::map<std::string, int> bctc_map{{"none", FV3D_BCTC_NONE},
                                            {"fixed_temperature", FV3D_BCTC_FIXED_TEMPERATURE},
                                            {"fixed_gradient", FV3D_BCTC_FIXED_GRADIENT}};
  res.bctc_zmin = lookup(bctc_map, reader.Get("thermal_conduction", "bc_zmin", "none"), "bc_zmin", cfg);
  res.bctc_zmax = lookup(bctc_map, reader.Get("thermal_conduction", "bc_zmax", "none"), "bc_zmax", cfg);
  res.bctc_zmin_value = reader.GetFloat("thermal_conduction", "bc_zmin_value", 1.0);
  res.bctc_zmax_value = reader.GetFloat("thermal_conduction", "bc_zmax_value", 1.0);

  res.viscosity_active = reader.GetBoolean("viscosity", "active", false);
  const std::map<std::string, int> visc_map{{"constant", FV3D_VSC_CONSTANT}};
  res.viscosity_mode = lookup(visc_map, reader.Get("viscosity", "viscosity_mode", "constant"), "viscosity_mode", cfg);
  res.mu = reader.GetFloat("viscosity", "mu", 0.0);

  res.h84_pert = reader.GetFloat("H84", "perturbation", 1.0e-4);
  res.c91_pert = reader.GetFloat("C91", "perturbation", 1.0e-3);
  res.seed = (int)reader.GetInteger("misc", "seed", 12345);
  res.log_frequency = (int)reader.GetInteger("misc", "log_frequency", 10);

  cfg.blast_radius = reader.GetFloat("blast", "radius", 0.2);
  cfg.blast_rho_in = reader.GetFloat("blast", "rho_in", 1.0);
  cfg.blast_p_in = reader.GetFloat("blast", "p_in", 10.0);
  cfg.blast_rho_out = reader.GetFloat("blast", "rho_out", 1.2);
  cfg.blast_p_out = reader.GetFloat("blast", "p_out", 0.1);
  cfg.max_steps = reader.GetInteger("run", "max_steps", 0);
  return cfg;
}

}  // namespace fv3d
