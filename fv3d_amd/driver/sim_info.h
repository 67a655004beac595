// sim_info.h -- host-side run configuration of the fv3d driver.
//
// readInifile (SimInfo.h:169-305) over an INI reader with inih INIReader
// semantics: case-insensitive sections and keys, ';'/'#' comment lines, a
// ';' preceded by white space starts an inline comment, repeated keys are
// joined with '\n', numbers must parse completely (strtol base 0 / strtod)
// or the default is used, and an unreadable file yields every default.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "fv3d_params.h"

namespace fv3d {

class IniReader {
 public:
  explicit IniReader(const std::string& filename);
  int ParseError() const { return error_; }
  std::string Get(const std::string& section, const std::string& name,
                  const std::string& default_value) const;
  long GetInteger(const std::string& section, const std::string& name, long default_value) const;
  double GetFloat(const std::string& section, const std::string& name, double default_value) const;
  bool GetBoolean(const std::string& section, const std::string& name, bool default_value) const;

 private:
  int error_ = 0;
  std::map<std::string, std::string> values_;
  static std::string key(const std::string& section, const std::string& name);
};

// fv3d_params plus the members that stay on the host side of the C ABI.
struct RunConfig {
  fv3d_params params;
  std::string problem = "blast";
  std::string filename_out = "run";
  // extension keys ([blast] radius, rho_in, p_in, rho_out, p_out); the
  // defaults are the reference's constants (Init.h:88-95)
  double blast_radius = 0.2, blast_rho_in = 1.0, blast_p_in = 10.0;
  double blast_rho_out = 1.2, blast_p_out = 0.1;
  // extension key [run] max_steps: stop after that many steps (0: no limit)
  long max_steps = 0;
  std::vector<std::string> warnings;
};

RunConfig readInifile(const std::string& filename);

}  // namespace fv3d
