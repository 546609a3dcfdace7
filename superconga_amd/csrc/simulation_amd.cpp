// simulation_amd.cpp -- the reference's `simulation` binary
// (src/simulation.cu:64-530), called the same way
// (frontend/common/cpp.py:16-39: `simulation_amd <config.json>`), on the
// C++ adaptors of include/superconga_amd.hpp.
//
// It reads simulation_config.json, warm-starts from misc.load_path
// (order_parameter.csv, domain.csv, vector_potential.csv;
// src/simulation.cu:296-375), runs the burn-in and the iterations with the
// config's limits (runSimulation, run_simulation.h:124-187;
// runIterations, run_iteration.h:85-141) and writes misc.save_path the way
// the reference's CSV writers do (io/SimulationWriter.h:50-236,
// io/CSVGridWriter.h:30-242, io/CSVSimulationIterationWriter.h:20-101):
// order_parameter.csv, current_density.csv, vector_potential.csv,
// magnetic_flux_density.csv, domain.csv, simulation_results.json,
// internal_parameters.json, simulation_convergence.csv and a copy of the
// config.  HDF5 output (data_format "h5") is refused with an error.
#include "../../include/superconga_amd.hpp"
#include "sc_json.hpp"

#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

using Ctx = superconga_amd::Context<double>;
namespace json = scamd::json;

// order-parameter components in the order src/simulation.cu:201-213 adds them
const char *const kComponents[] = {"dx2-y2", "dxy", "g", "s"};
constexpr int kGrainMargin = 3; // defines.h

bool readFile(const std::string &path, std::string &out) {
  std::ifstream f(path);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

bool exists(const std::string &path) {
  struct stat st;
  return ::stat(path.c_str(), &st) == 0;
}

std::string join(const std::string &dir, const std::string &name) {
  if (dir.empty() || dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

void makeDirs(const std::string &path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty() && !exists(cur) && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
        throw std::runtime_error("could not create directory " + cur);
    }
    if (i < path.size()) cur += path[i];
  }
}

// One value the way std::scientific with precision 8 prints it.
std::string fmt(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.8e", v);
  return b;
}

// A JSON number that reads back to the same double.
std::string jnum(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}

// CSVReader::read (io/CSVReader.h:30-191): columns by header name.
struct Csv {
  std::vector<std::string> header;
  std::vector<std::vector<double>> columns;
  const std::vector<double> *column(const std::string &name) const {
    for (size_t i = 0; i < header.size(); ++i)
      if (header[i] == name) return &columns[i];
    return nullptr;
  }
};

Csv readCsv(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("-- ERROR! Could not open file " + path);
  Csv c;
  std::string line;
  if (!std::getline(f, line)) return c;
  {
    std::stringstream ss(line);
    std::string name;
    while (std::getline(ss, name, ',')) {
      while (!name.empty() && (name.back() == '\r' || name.back() == ' ')) name.pop_back();
      c.header.push_back(name);
    }
  }
  c.columns.assign(c.header.size(), {});
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    const char *p = line.c_str();
    for (size_t i = 0; i < c.header.size(); ++i) {
      char *end = nullptr;
      const double v = std::strtod(p, &end);
      if (end == p) throw std::runtime_error("-- ERROR! Malformed CSV row in " + path);
      c.columns[i].push_back(v);
      p = end;
      if (*p == ',') ++p;
    }
  }
  return c;
}

const std::vector<double> &squareColumn(const Csv &c, const std::string &name, int n,
                                        const std::string &what) {
  const std::vector<double> *col = c.column(name);
  if (!col) throw std::runtime_error("the file does not contain the column '" + name + "'");
  if (col->size() != size_t(n) * n)
    throw std::runtime_error(what + ": the data is not on this " + std::to_string(n) + "x" +
                             std::to_string(n) + " lattice (resampling is not supported)");
  return *col;
}

// CSVGridWriter::write: "x_idx,y_idx,<titles>", one row per site (x fastest).
void writeGridCsv(const std::string &path, const std::vector<std::string> &titles,
                  const std::vector<const std::vector<double> *> &columns, int n,
                  bool integer = false) {
  std::FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("-- ERROR! Could not open file " + path);
  std::string line = "x_idx,y_idx";
  for (const std::string &t : titles) line += "," + t;
  std::fprintf(f, "%s\n", line.c_str());
  for (int q = 0; q < n * n; ++q) {
    std::fprintf(f, "%d,%d", q % n, q / n);
    for (const std::vector<double> *c : columns) {
      if (integer)
        std::fprintf(f, ",%d", static_cast<int>((*c)[size_t(q)]));
      else
        std::fprintf(f, ",%s", fmt((*c)[size_t(q)]).c_str());
    }
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

// curl_kernel (GridFunctions_GPU.h:326-353): one-sided at the edges.
std::vector<double> curl(const std::vector<double> &ax, const std::vector<double> &ay, int n,
                         double h) {
  std::vector<double> b(size_t(n) * n);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const int xn = std::min(x + 1, n - 1), xp = std::max(x - 1, 0);
      const int yn = std::min(y + 1, n - 1), yp = std::max(y - 1, 0);
      const double dAydx =
          (ay[size_t(y) * n + xn] - ay[size_t(y) * n + xp]) / double(xn - xp);
      const double dAxdy =
          (ax[size_t(yn) * n + x] - ax[size_t(yp) * n + x]) / double(yn - yp);
      b[size_t(y) * n + x] = (dAydx - dAxdy) / h;
    }
  return b;
}

struct Run {
  json::Value cfg;
  std::string text;
  std::vector<std::string> types; // order-parameter components, reference order
  bool field = false;             // finite penetration depth
  std::string savePath;
  bool save = false;
  bool verbose = false;
};

void loadState(Ctx &ctx, const Run &r, const std::string &loadPath) {
  const int n = ctx.dims().n0;
  const std::string opPath = join(loadPath, "order_parameter.csv");
  const std::string domPath = join(loadPath, "domain.csv");
  const std::string aPath = join(loadPath, "vector_potential.csv");
  for (const std::string &p : {opPath, domPath, aPath})
    if (!exists(p)) throw std::runtime_error("-- ERROR! Could not open file " + p);
  const Csv op = readCsv(opPath);
  for (size_t c = 0; c < r.types.size(); ++c) {
    const std::vector<double> &re = squareColumn(op, r.types[c] + "_re", n, "order parameter");
    const std::vector<double> &im = squareColumn(op, r.types[c] + "_im", n, "order parameter");
    std::vector<std::complex<double>> v(re.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = {re[i], im[i]};
    ctx.setOrderParameter(static_cast<int>(c), v);
  }
  const Csv a = readCsv(aPath);
  std::vector<double> ax = squareColumn(a, "a_x", n, "vector potential");
  std::vector<double> ay = squareColumn(a, "a_y", n, "vector potential");
  if (!r.field) {
    std::fill(ax.begin(), ax.end(), 0.0);
    std::fill(ay.begin(), ay.end(), 0.0);
  }
  ctx.setVectorPotential(ax, ay);
}

// SimulationWriter::write (io/SimulationWriter.h:50-236), CSV format.
void saveState(Ctx &ctx, const Run &r) {
  makeDirs(r.savePath);
  const sc_dims &d = ctx.dims();
  const int n = d.n0;
  const size_t n0 = size_t(n) * n;
  const double h = d.h;
  const sc_results res = ctx.results();
  const std::vector<char> domain = ctx.getDomain();
  std::vector<double> dom(n0);
  for (size_t i = 0; i < n0; ++i) dom[i] = domain[i];
  writeGridCsv(join(r.savePath, "domain.csv"), {"domain"}, {&dom}, n, true);
  // ComputeOrderParameter::getResult: the last gap equation's result
  std::vector<std::vector<double>> planes;
  std::vector<std::string> titles;
  double sumAbsDelta = 0.0;
  for (size_t c = 0; c < r.types.size(); ++c) {
    const std::vector<std::complex<double>> v = ctx.getOrderParameter(static_cast<int>(c), 1);
    std::vector<double> re(n0), im(n0);
    for (size_t i = 0; i < n0; ++i) {
      re[i] = v[i].real();
      im[i] = v[i].imag();
    }
    planes.push_back(std::move(re));
    planes.push_back(std::move(im));
    titles.push_back(r.types[c] + "_re");
    titles.push_back(r.types[c] + "_im");
  }
  // |Delta| summed over components (numpy's sum over the whole array)
  for (size_t c = 0; c < r.types.size(); ++c)
    for (size_t i = 0; i < n0; ++i)
      sumAbsDelta += std::hypot(planes[2 * c][i], planes[2 * c + 1][i]);
  std::vector<const std::vector<double> *> cols;
  for (const auto &p : planes) cols.push_back(&p);
  writeGridCsv(join(r.savePath, "order_parameter.csv"), titles, cols, n);
  std::vector<double> jx(n0), jy(n0);
  superconga_amd::check(sc_get_current(ctx.handle(), jx.data(), jy.data()), "sc_get_current");
  writeGridCsv(join(r.savePath, "current_density.csv"), {"j_x", "j_y"}, {&jx, &jy}, n);
  std::vector<double> ax, ay;
  ctx.getVectorPotential(ax, ay, 0); // the unscaled induced potential
  if (!r.field) {
    std::fill(ax.begin(), ax.end(), 0.0);
    std::fill(ay.begin(), ay.end(), 0.0);
  }
  writeGridCsv(join(r.savePath, "vector_potential.csv"), {"a_x", "a_y"}, {&ax, &ay}, n);
  const std::vector<double> bz = curl(ax, ay, n, h);
  writeGridCsv(join(r.savePath, "magnetic_flux_density.csv"), {"b_z"}, {&bz}, n);
  // the scalar results (SimulationWriter.h:50-130)
  const int ng = d.num_grain_sites;
  const double area = h * h * ng;
  double sumJ = 0.0, sumB = 0.0, sumA = 0.0, moment = 0.0;
  const double centre = 0.5 * (n - 1);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const size_t q = size_t(y) * n + x;
      sumJ += std::hypot(jx[q], jy[q]);
      if (domain[q]) {
        sumB += bz[q];
        sumA += std::hypot(ax[q], ay[q]);
      }
      // GridFunctions::moment (GridFunctions_CPU.h:291-320): 0.5 h^3 sum r x j
      moment += (x - centre) * jy[q] - (y - centre) * jx[q];
    }
  moment *= 0.5 * h * h * h;
  std::FILE *f = std::fopen(join(r.savePath, "simulation_results.json").c_str(), "w");
  if (!f) throw std::runtime_error("-- ERROR! Could not write simulation_results.json");
  std::fprintf(f,
               "{\n    \"free_energy_per_unit_area\": %s,\n"
               "    \"magnetic_free_energy_per_unit_area\": %s,\n"
               "    \"order_parameter_average_magnitude\": %s,\n"
               "    \"current_density_average_magnitude\": %s,\n"
               "    \"magnetic_moment_per_unit_area\": %s,\n"
               "    \"flux_density_average\": %s,\n"
               "    \"vector_potential_average_magnitude\": %s\n}",
               jnum(res.free_energy_per_area).c_str(),
               jnum(res.magnetic_free_energy_per_area).c_str(), jnum(sumAbsDelta / ng).c_str(),
               jnum(sumJ / ng).c_str(), jnum(moment / area).c_str(), jnum(sumB / ng).c_str(),
               jnum(sumA / ng).c_str());
  std::fclose(f);
  f = std::fopen(join(r.savePath, "internal_parameters.json").c_str(), "w");
  if (!f) throw std::runtime_error("-- ERROR! Could not write internal_parameters.json");
  std::fprintf(f,
               "{\n    \"area_compute\": %s,\n    \"internal_lattice_points\": %d,\n"
               "    \"grain_margin\": %d,\n    \"num_lattice_points\": %d\n}",
               jnum(area).c_str(), ng, kGrainMargin, n * n);
  std::fclose(f);
}

// CSVSimulationIterationWriter: one row per iteration, appended to the file
// (which is never removed, CSVSimulationIterationWriter.h:24-35).
void writeConvergence(const std::string &path, const sc_results &r) {
  const bool fresh = !exists(path);
  std::FILE *f = std::fopen(path.c_str(), "a");
  if (!f) throw std::runtime_error("-- ERROR! Could not open file " + path);
  if (fresh)
    std::fprintf(f, "iteration,boundary_coherence_residual,order_parameter_residual,"
                    "current_density_residual,vector_potential_residual,"
                    "free_energy_residual,free_energy_per_unit_area\n");
  std::fprintf(f, "%d,%s,%s,%s,%s,%s,%s\n", r.iterations, fmt(r.residual_boundary).c_str(),
               fmt(r.residual_order_parameter).c_str(), fmt(r.residual_current).c_str(),
               fmt(r.residual_vector_potential).c_str(), fmt(r.residual_free_energy).c_str(),
               fmt(r.free_energy_per_area).c_str());
  std::fclose(f);
}

// runIterations (run_iteration.h:85-141): until converged with the minimum
// reached, or to the maximum (< 0: none); a burn-in converges on the
// boundary residual.  Every iteration writes a convergence row.
bool runIterations(Ctx &ctx, const Run &r, bool compute, int nMin, int nMax, double criterion) {
  if (nMin == 0 && nMax == 0) return false;
  const std::string convPath = join(r.savePath, "simulation_convergence.csv");
  for (int i = 1;; ++i) {
    if (compute)
      ctx.compute();
    else
      ctx.burnIn();
    const sc_results res = ctx.results();
    const bool conv = compute ? res.converged != 0 : res.residual_boundary < criterion;
    if (r.verbose)
      std::printf("-- iteration %d residual %.3e free energy %.10f\n", res.iterations,
                  res.residual, res.free_energy_per_area);
    if (r.save) writeConvergence(convPath, res);
    if ((conv && i >= nMin) || (nMax >= 0 && i >= nMax)) return conv;
  }
}

int simulate(const std::string &configPath) {
  Run r;
  if (!readFile(configPath, r.text)) {
    std::fprintf(stderr, "-- ERROR! Could not find/open config file '%s'.\n",
                 configPath.c_str());
    return EXIT_FAILURE;
  }
  r.cfg = json::Parser(r.text).parse();
  const json::Value &misc = r.cfg["misc"];
  const json::Value &nu = r.cfg["numerics"];
  const std::string format = misc["data_format"].asString("csv");
  r.savePath = misc["save_path"].asString("");
  const int saveFreq = misc["save_frequency"].asInt(-1);
  r.save = !r.savePath.empty() && saveFreq != 0;
  r.verbose = misc["verbose"].asBool(false);
  if (format != "csv" && format != "h5")
    throw std::runtime_error("misc.data_format must be 'csv' or 'h5', got '" + format + "'");
  if (format == "h5" && r.save)
    throw std::runtime_error("misc.data_format 'h5' (HDF5 output) is not supported by this "
                             "build: set misc.data_format to 'csv' (the reference writes the "
                             "same grids as CSV, io/CSVGridWriter.h)");
  const json::Value &op = r.cfg["physics"]["order_parameter"];
  for (const char *name : kComponents)
    if (op.has(name)) r.types.push_back(name);
  r.field = r.cfg["physics"]["penetration_depth"].asDouble(-1.0) > 0.0;
  const double criterion = nu["convergence_criterion"].asDouble(1e-6);

  Ctx ctx(r.text);
  const std::string loadPath = misc["load_path"].asString("");
  if (!loadPath.empty()) loadState(ctx, r, loadPath);
  if (r.save) {
    makeDirs(r.savePath);
    std::ofstream(join(r.savePath, "simulation_config.json")) << r.text;
  }
  const int nBurn = nu["num_iterations_burnin"].asInt(0);
  const int nMin = nu["num_iterations_min"].asInt(0);
  const int nMax = nu["num_iterations_max"].asInt(-1);
  runIterations(ctx, r, false, nBurn, nBurn, criterion);
  bool converged = false;
  if (saveFreq > 0) {
    // one iteration at a time; the limits and the save schedule count the
    // burn-in iterations too (Context::getNumIterations)
    for (;;) {
      converged = runIterations(ctx, r, true, 0, 1, criterion);
      const int it = ctx.getNumIterations();
      if ((converged && it >= nMin) || (nMax >= 0 && it >= nMax)) break;
      if (r.save && it % saveFreq == 0) saveState(ctx, r);
    }
  } else {
    converged = runIterations(ctx, r, true, nMin, nMax, criterion);
  }
  if (r.save) saveState(ctx, r);
  const sc_results res = ctx.results();
  std::printf("-- %s after %d iterations, residual %.3e, free energy per unit area %.10f\n",
              converged ? "Converged" : "Not converged", res.iterations, res.residual,
              res.free_energy_per_area);
  return EXIT_SUCCESS;
}

} // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "-- ERROR! Please provide a config file.\n");
    return EXIT_FAILURE;
  }
  try {
    return simulate(argv[1]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "-- ERROR! %s\n", e.what());
    return EXIT_FAILURE;
  }
}
