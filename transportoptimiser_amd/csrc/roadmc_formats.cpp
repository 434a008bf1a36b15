// On-disk formats: the surrogate training data file and the key-value input
// files of the reference's data directory.  Host code only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "roadmc.h"

namespace {

// Utility::getDictionary (Utility.cpp:365-381): split on a delimiter, trim
// the ends and collapse runs of spaces in each field.
std::vector<std::string> fields(const std::string& line, char del) {
  std::vector<std::string> out;
  std::string cur;
  auto flush = [&]() {
    std::string t;
    bool space = false;
    for (char ch : cur) {
      if (ch == ' ' || ch == '\t' || ch == '\r') {
        space = !t.empty();
      } else {
        if (space) t.push_back(' ');
        space = false;
        t.push_back(ch);
      }
    }
    out.push_back(t);
    cur.clear();
  };
  for (char ch : line) {
    if (ch == del)
      flush();
    else
      cur.push_back(ch);
  }
  flush();
  return out;
}

// "Key : value" lines of a file (comment and banner lines skipped); the first
// occurrence of a key wins.
bool read_dict(const char* path, std::map<std::string, std::string>& kv,
               std::vector<std::vector<std::string>>* tables) {
  std::ifstream in(path);
  if (!in.good()) return false;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    const std::vector<std::string> f = fields(line, ':');
    if (f.size() < 2) continue;
    if (f.size() > 2 && tables) {
      tables->push_back(f);
      continue;
    }
    if (!kv.count(f[0])) kv[f[0]] = f[1];
  }
  return true;
}

bool num(const std::map<std::string, std::string>& kv, const char* key, double* out) {
  auto it = kv.find(key);
  if (it == kv.end()) return false;
  char* end = nullptr;
  const double v = std::strtod(it->second.c_str(), &end);
  if (end == it->second.c_str()) return false;
  *out = v;
  return true;
}

void copy_name(char* dst, const std::string& s, size_t n) {
  std::strncpy(dst, s.c_str(), n - 1);
  dst[n - 1] = '\0';
}

// The stream formatting of the reference's writer: width 20, default
// precision (6 significant digits, %g).
std::string w20(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%20.6g", v);
  return buf;
}

}  // namespace

extern "C" int roadmc_surrogate_data_write(const char* path, int32_t kind, int32_t n_species,
                                           int32_t n_rows, const double* rows, double max_benefit,
                                           int32_t* n_written) {
  if (!path || n_species < 1 || n_rows < 0 || (n_rows > 0 && !rows) ||
      (kind != ROADMC_SURR_MTE && kind != ROADMC_SURR_CONTROLLED))
    return ROADMC_ERR_INVALID;
  bool exists;
  {
    std::ifstream chk(path);
    exists = chk.good();
  }
  std::ofstream out(path, exists ? std::ios::app : std::ios::out);
  if (!out.good()) return ROADMC_ERR_INVALID;
  if (!exists) {
    const std::string bar(100, '#');
    out << bar << "\n"
        << "######################################## SURROGATE RESULTS "
           "#########################################\n"
        << bar << "\n";
    char buf[128];
    if (kind == ROADMC_SURR_MTE) {
      for (int s = 0; s < n_species; ++s) {
        std::snprintf(buf, sizeof buf, "IAR_%03d                  END_POP_%017dEND_POP_SD%015d",
                      s, s, s);
        out << buf;
      }
    } else {
      for (int s = 0; s < n_species; ++s) {
        std::snprintf(buf, sizeof buf, "IAR_%3d                  ", s);
        out << buf;
      }
      out << "INITIAL_UNIT_PROFIT      AVERAGE_VALUE            VALUE_SD";
    }
    out << "\n";
  }
  int written = 0;
  const int width = kind == ROADMC_SURR_MTE ? 3 * n_species : n_species + 3;
  for (int r = 0; r < n_rows; ++r) {
    const double* row = rows + (size_t)r * width;
    if (kind == ROADMC_SURR_MTE) {
      for (int s = 0; s < n_species; ++s)
        out << w20(row[3 * s]) << "     " << w20(row[3 * s + 1]) << "     " << w20(row[3 * s + 2]);
    } else {
      // RoadGA.cpp:3040-3057: only samples whose value and SD stay within the
      // maximum benefit.
      const double v = row[n_species + 1], sd = row[n_species + 2];
      if (!std::isfinite(max_benefit) || !(std::fabs(v) <= std::fabs(max_benefit)) ||
          !(std::fabs(sd) <= std::fabs(max_benefit)))
        continue;
      for (int c = 0; c < width; ++c) out << w20(row[c]);
    }
    out << "\n";
    ++written;
  }
  out.flush();
  if (!out.good()) return ROADMC_ERR_INVALID;
  if (n_written) *n_written = written;
  return ROADMC_OK;
}

extern "C" int roadmc_surrogate_data_read(const char* path, int32_t kind, int32_t n_species,
                                          int32_t max_rows, double* rows, int32_t* n_rows) {
  if (!path || !n_rows || n_species < 1 || max_rows < 0 || (max_rows > 0 && !rows) ||
      (kind != ROADMC_SURR_MTE && kind != ROADMC_SURR_CONTROLLED))
    return ROADMC_ERR_INVALID;
  std::ifstream in(path);
  if (!in.good()) return ROADMC_ERR_INVALID;
  const int width = kind == ROADMC_SURR_MTE ? 3 * n_species : n_species + 3;
  std::string line;
  for (int i = 0; i < 4; ++i)
    if (!std::getline(in, line)) {
      *n_rows = 0;
      return ROADMC_OK;
    }
  int count = 0;
  std::vector<double> v(width);
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    int got = 0;
    while (got < width && (iss >> v[got])) ++got;
    if (got == 0) continue;  // blank line
    if (got < width) return ROADMC_ERR_NUMERIC;
    if (count < max_rows)
      for (int c = 0; c < width; ++c) rows[(size_t)count * width + c] = v[c];
    ++count;
  }
  *n_rows = count;
  return ROADMC_OK;
}

extern "C" int roadmc_read_commodity(const char* path, roadmc_commodity_file* o) {
  if (!path || !o) return ROADMC_ERR_INVALID;
  std::map<std::string, std::string> kv;
  if (!read_dict(path, kv, nullptr)) return ROADMC_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  if (!kv.count("Name")) return ROADMC_ERR_INVALID;
  copy_name(o->name, kv["Name"], ROADMC_NAME_LEN);
  const bool ok = num(kv, "Current Value", &o->current) && num(kv, "Mean Value", &o->mean) &&
                  num(kv, "Standard Deviation", &o->sd) && num(kv, "Reversion", &o->reversion) &&
                  num(kv, "Jump Probability", &o->jump_prob) &&
                  num(kv, "Poisson Parameter", &o->poisson_jump) &&
                  num(kv, "Mean Contribution to Transported Ore", &o->ore_mean) &&
                  num(kv, "Standard Deviation of Contribution", &o->ore_sd);
  return ok ? ROADMC_OK : ROADMC_ERR_INVALID;
}

extern "C" int roadmc_read_vehicle(const char* path, roadmc_vehicle_file* o) {
  if (!path || !o) return ROADMC_ERR_INVALID;
  std::map<std::string, std::string> kv;
  if (!read_dict(path, kv, nullptr)) return ROADMC_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  if (!kv.count("Name")) return ROADMC_ERR_INVALID;
  copy_name(o->name, kv["Name"], ROADMC_NAME_LEN);
  double fuel = 0.0;
  const bool ok = num(kv, "Width (m)", &o->width) && num(kv, "Length (m)", &o->length) &&
                  num(kv, "Traffic Proportion", &o->proportion) &&
                  num(kv, "Load (tonnes)", &o->load) && num(kv, "Fuel", &fuel) &&
                  num(kv, "a", &o->a) && num(kv, "agr", &o->agr) && num(kv, "av", &o->av) &&
                  num(kv, "avsq", &o->avsq) && num(kv, "Travel Cost per Hour ($)", &o->travel_cost);
  o->fuel = (int32_t)fuel;
  return ok ? ROADMC_OK : ROADMC_ERR_INVALID;
}

extern "C" int roadmc_read_species(const char* path, roadmc_species_file* o) {
  if (!path || !o) return ROADMC_ERR_INVALID;
  std::map<std::string, std::string> kv;
  std::vector<std::vector<std::string>> tab;
  if (!read_dict(path, kv, &tab)) return ROADMC_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  if (!kv.count("Name")) return ROADMC_ERR_INVALID;
  copy_name(o->name, kv["Name"], ROADMC_NAME_LEN);
  // HABITAT TYPES rows: TYPE : max pop : cost : preference mean : SD : vegetation
  // (RoadGA.cpp:4288-4311).
  for (const auto& row : tab) {
    if (o->n_habitat_types >= ROADMC_MAX_HABITAT_TYPES || row.size() < 5) continue;
    const int i = o->n_habitat_types++;
    copy_name(o->habitat_name[i], row[0], 16);
    o->max_pop[i] = std::strtod(row[1].c_str(), nullptr);
    o->cost_per_m2[i] = std::strtod(row[2].c_str(), nullptr);
    o->pref_mean[i] = std::strtod(row[3].c_str(), nullptr);
    o->pref_sd[i] = std::strtod(row[4].c_str(), nullptr);
  }
  // GROWTH RATE PARAMETERS (RoadGA.cpp:4361-4382).
  const bool ok = num(kv, "Current", &o->growth_current) && num(kv, "Mean", &o->growth_mean) &&
                  num(kv, "Standard Deviation", &o->growth_sd) &&
                  num(kv, "Reversion", &o->growth_reversion) &&
                  num(kv, "Jump Probability", &o->growth_jump_prob) &&
                  num(kv, "Jump Poisson Parameter", &o->growth_poisson_jump);
  return ok ? ROADMC_OK : ROADMC_ERR_INVALID;
}

namespace {

// A comma-separated list of numbers (the multiplier and flow lists).
int parse_list(const std::string& s, double* out, int cap) {
  int n = 0;
  for (const std::string& f : fields(s, ',')) {
    if (f.empty()) continue;
    char* end = nullptr;
    const double v = std::strtod(f.c_str(), &end);
    if (end == f.c_str()) return -1;
    if (n >= cap) return -1;
    out[n++] = v;
  }
  return n;
}

int int_or(const std::map<std::string, std::string>& kv, const char* key, int dflt) {
  double v;
  return num(kv, key, &v) ? (int)v : dflt;
}

double num_or_nan(const std::map<std::string, std::string>& kv, const char* key) {
  double v;
  return num(kv, key, &v) ? v : std::numeric_limits<double>::quiet_NaN();
}

int flag_or(const std::map<std::string, std::string>& kv, const char* key) {
  auto it = kv.find(key);
  if (it == kv.end()) return -1;
  return it->second == "TRUE" ? 1 : 0;
}

bool list_of(const std::map<std::string, std::string>& kv, const char* key, roadmc_list* l) {
  l->n = 0;
  auto it = kv.find(key);
  if (it == kv.end()) return true;
  const int n = parse_list(it->second, l->v, ROADMC_MAX_LIST);
  if (n < 0) return false;
  l->n = n;
  return true;
}

// "<prefix> 1", "<prefix> 2", ... in order, up to the first gap.
int file_list(const std::map<std::string, std::string>& kv, const char* prefix,
              char (*out)[ROADMC_PATH_LEN]) {
  int n = 0;
  while (n < ROADMC_MAX_FILES) {
    auto it = kv.find(std::string(prefix) + " " + std::to_string(n + 1));
    if (it == kv.end()) break;
    copy_name(out[n], it->second, ROADMC_PATH_LEN);
    ++n;
  }
  return n;
}

}  // namespace

// The simulation subset of the main input file, read by key where the
// reference reads by line position (RoadGA::initialiseFromTextInput,
// RoadGA.cpp:3464-4500), so that both the older layout of tests/Input_Data
// and the current one of the experiment directories are accepted.
extern "C" int roadmc_read_run_settings(const char* path, roadmc_run_settings* o) {
  if (!path || !o) return ROADMC_ERR_INVALID;
  std::map<std::string, std::string> kv;
  std::vector<std::vector<std::string>> tables;
  if (!read_dict(path, kv, &tables)) return ROADMC_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  o->confidence_interval = num_or_nan(kv, "Confidence Interval");
  o->use_gpu = flag_or(kv, "Use GPU");
  o->surrogate_dim_res = int_or(kv, "Surrogate Dimension Resolution", -1);
  o->interpolation = int_or(kv, "Interpolation Routine", -1);
  o->traffic_growth = num_or_nan(kv, "Traffic Growth Rate (p.a.)");
  o->rrr = num_or_nan(kv, "Required Rate of Return (per cen per annum)");
  o->n_years = int_or(kv, "Number of Years", -1);
  o->time_step = num_or_nan(kv, "Step Size (Years)");
  o->n_paths = int_or(kv, "Number of Simulation Paths", -1);
  o->rov_dim_res = int_or(kv, "ROV Dimension Resolution", -1);
  o->optimisation_method = int_or(kv, "Optimisation Method", -1);
  o->rov_method = int_or(kv, "ROV Method", -1);
  const struct {
    const char* key;
    roadmc_list* l;
  } lists[] = {
      {"Animal Bridges to Build", &o->animal_bridges},
      {"Habitat Preference SD Multipliers", &o->hab_pref_sd},
      {"Movement Propensity SD Multipliers", &o->movement_sd},
      {"Ranging Coefficient SD Multipliers", &o->ranging_sd},
      {"Population Growth Rate Mean Multipliers", &o->growth_mean},
      {"Population Growth Rate SD Multipliers", &o->growth_sd},
      {"Commodity Price Mean Multipliers", &o->price_mean},
      {"Commodity Price SD Multipliers", &o->price_sd},
      {"Ore Composition SD Multipliers", &o->ore_sd},
      {"Minimum Population Levels (proportion of min)", &o->pop_level},
      {"Comparison Road Variable Cost Multipliers", &o->comparison_cost},
      {"Fuel Consumption Per Journey", &o->fuel_per_journey},
  };
  for (const auto& l : lists)
    if (!list_of(kv, l.key, l.l)) return ROADMC_ERR_NUMERIC;
  // Traffic programs: "Program k : flows : switching [: bridge]" in file order.
  for (const std::vector<std::string>& t : tables) {
    if (t[0].rfind("Program ", 0) != 0) continue;
    if (o->n_programs >= ROADMC_MAX_PROGRAMS) return ROADMC_ERR_INVALID;
    const int k = o->n_programs;
    const int nf = parse_list(t[1], o->flows[k], ROADMC_MAX_CONTROLS);
    const int ns = parse_list(t[2], o->switching[k], 1 << ROADMC_MAX_CONTROLS);
    if (nf < 0 || ns < 0) return ROADMC_ERR_NUMERIC;
    o->n_flows[k] = nf;
    o->n_switching[k] = ns;
    o->bridge[k] = t.size() > 3 && t[3] == "TRUE" ? 1 : 0;
    ++o->n_programs;
  }
  const int declared = int_or(kv, "Programs", -1);
  if (declared >= 0 && declared != o->n_programs) return ROADMC_ERR_INVALID;
  o->n_fuels = int_or(kv, "Number of Fuels", -1);
  o->n_transported = int_or(kv, "Number of Transported Commodities",
                            int_or(kv, "Number of Commodities", -1));
  o->n_commodities = file_list(kv, "Commodity", o->commodity);
  o->n_vehicles = file_list(kv, "Vehicle", o->vehicle);
  o->n_species = file_list(kv, "Species", o->species);
  o->road_exists = flag_or(kv, "Road Exists");
  o->travel_time = num_or_nan(kv, "Travel Time (hours)");
  o->road_length = num_or_nan(kv, "Length (m)");
  o->accident_cost = num_or_nan(kv, "Accident Cost per Unit Traffic");
  return ROADMC_OK;
}
