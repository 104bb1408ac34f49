// cdfci_omp-compatible driver: `cdfci_omp <input.json>`.
//
// Reproduces the reference's command surface (src/main.cpp:19-36,
// src/run_wrapper.h:22-241, src/solver.h:672-893): the same JSON keys and
// defaults, the same validation messages and exit codes (0 ok, 1 on an
// exception with "ERROR: <what>" on stderr, 255 on the exit(-1) paths), the
// same report rows, with the solver running on the HIP device through the
// C ABI of include/mcdfci.h.  Extension key: solver.cdfci.precision
// ("fp64" default, or "mixed").

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcdfci.h"
#include "json.h"

using mcjson::Value;

namespace {

void print_header() {
  std::cout << R"(  ______  _______   _______   ______  __  )" << std::endl;
  std::cout << R"( /      ||       \ |   ____| /      ||  | )" << std::endl;
  std::cout << R"(|  ,----'|  .--.  ||  |__   |  ,----'|  | )" << std::endl;
  std::cout << R"(|  |     |  |  |  ||   __|  |  |     |  | )" << std::endl;
  std::cout << R"(|  `----.|  '--'  ||  |     |  `----.|  | )" << std::endl;
  std::cout << R"( \______||_______/ |__|      \______||__|  )" << "0.1.0" << std::endl;
  std::cout << std::endl;
  std::cout << "Compiled on " << __DATE__ << " at " << __TIME__ << std::endl;
  std::cout << std::endl;
  std::cout << "MI355X (gfx950) HIP implementation of coordinate-descent FCI" << std::endl;
  std::cout << "=======================================================" << std::endl;
  std::cout << "HIP: device solver (one GPU)" << std::endl;
  std::cout << std::endl;
}

std::string check_argument(int argc, char* argv[]) {
  if (argc < 2)
    throw std::invalid_argument(
        "No input file specified. Please use the path of the input file as the first argument.");
  return argv[1];
}

Value option_from_file(const std::string& path) {
  std::ifstream f(path);
  if (f.fail()) throw std::invalid_argument("Can not read the input file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  try {
    return mcjson::parse(ss.str());
  } catch (...) {
    throw std::invalid_argument("Can not read the input file " + path);
  }
}

Value read_input(const std::string& input_file) {
  Value cd = Value::object();
  cd["num_iterations"] = Value::of_uint(10);
  cd["report_interval"] = Value::of_uint(1);
  cd["coordinate_pick"] = Value::of_str("gcd_grad");
  cd["coordinate_update"] = Value::of_str("ls");
  cd["ref_det_occ"] = Value::array();
  cd["z_threshold"] = Value::of(0.0);
  cd["z_threshold_search"] = Value::of_bool(false);
  cd["max_wavefunction_size"] = Value::of_uint(1);
  cd["max_memory"] = Value::of(0.0);
  cd["num_of_coordinate"] = Value::of_uint(1);
  cd["estimate"] = Value::of_uint(100);
  Value hm = Value::object();
  hm["type"] = Value::of_str("molecule");
  hm["threshold"] = Value::of(1e-13);
  Value option = option_from_file(input_file);
  const Value* sol = option.find("solver");
  const Value* cdp = sol ? sol->find("cdfci") : nullptr;
  if (cdp) mcjson::merge_patch(cd, *cdp);
  option["solver"]["cdfci"] = cd;
  const Value* hp = option.find("hamiltonian");
  if (hp) mcjson::merge_patch(hm, *hp);
  option["hamiltonian"] = hm;
  if (option["hamiltonian"]["type"].kind != Value::String || option["hamiltonian"]["type"].s != "molecule")
    throw std::invalid_argument(
        "hamiltonian:type is invalid. Currently the only supported Hamiltonian is \"molecule\".");
  double max_memory = option["solver"]["cdfci"]["max_memory"].num();
  if (max_memory <= 0.0)
    throw std::invalid_argument(
        "solver:cdfci:max_memory is invalid. It is the maximum memory (GB) allowed for the wavefunction. "
        "Please set a positive number.");
  bool zts = option["solver"]["cdfci"]["z_threshold_search"].boolean();
  double zt = option["solver"]["cdfci"]["z_threshold"].num();
  long report_interval = (long)option["solver"]["cdfci"]["report_interval"].integer();
  if (zts) {
    std::cout << "Warning: z cut-off threshold auto search (z_threshold_search) is enabled. ";
    std::cout << "This feature is experimental and may not always work." << std::endl << std::endl;
    const double start = 1e-10;
    if (zt < start) {
      std::cout << "Warning: The initial z_threshold may be too small. Set z_threshold = " << std::scientific
                << std::setprecision(2) << start << std::endl
                << std::endl;
    }
  }
  if (zts && report_interval < 1000)
    throw std::invalid_argument("Please set report_interval >= 1000 when z_threshold_search = true.");
  return option;
}

struct Ctx {
  mcdfci_ctx* c = nullptr;
  ~Ctx() {
    if (c) mcdfci_destroy(c);
  }
};

void check(mcdfci_ctx* c, int rc) {
  if (rc == MCDFCI_OK) return;
  std::string msg = mcdfci_last_error(c);
  if (rc == MCDFCI_E_TABLE_FULL) throw std::overflow_error(msg);
  throw std::runtime_error(msg);
}

double run(Value& option) {
  Value& cdo = option["solver"]["cdfci"];
  const std::string path = option["hamiltonian"]["fcidump_path"].str();
  mcdfci_opts o{};
  o.num_of_coordinate = (int)cdo["num_of_coordinate"].integer();
  o.estimate = (int)cdo["estimate"].integer();
  o.z_threshold = cdo["z_threshold"].num();
  o.max_memory = cdo["max_memory"].num();
  o.threshold = option["hamiltonian"]["threshold"].num();
  o.precision = MCDFCI_PREC_FP64;
  if (const Value* pr = cdo.find("precision")) {
    if (pr->str() == "mixed") o.precision = MCDFCI_PREC_MIXED;
    else if (pr->str() != "fp64") throw std::invalid_argument("solver:cdfci:precision must be \"fp64\" or \"mixed\".");
  }
  o.device = 0;
  if (const char* dv = std::getenv("MCDFCI_DEVICE")) o.device = std::atoi(dv);
  // The update functor is validated below, where the reference does it;
  // the context is created with the line search when it is named.
  o.coordinate_update = (cdo["coordinate_update"].kind == Value::String && cdo["coordinate_update"].s == "ls")
                            ? MCDFCI_UPDATE_LS
                            : MCDFCI_UPDATE_EIG;
  // Validate the functor choices before touching the device (src/run_wrapper.h:136-156).
  Ctx ctx;
  int rc = mcdfci_create_from_fcidump(path.c_str(), &o, &ctx.c);
  if (rc == MCDFCI_E_INPUT) throw std::invalid_argument(mcdfci_last_error(nullptr));
  check(ctx.c, rc);
  mcdfci_ctx* c = ctx.c;
  int64_t info[8] = {0};
  check(c, mcdfci_info(c, info));
  const int nw = (int)info[0];
  const int nso = (int)info[1];
  std::cout << "FCIDUMP information" << std::endl;
  std::cout << "-------------------" << std::endl;
  std::cout << "Number of elections: " << info[2] << std::endl;
  std::cout << "Number of spin-orbitals: " << nso << std::endl;
  std::cout << "Ms2: " << info[3] << std::endl;
  std::cout << "Hartree Fock: " << (info[7] ? "unrestricted" : "restricted") << std::endl << std::endl;
  std::cout << "Machine information" << std::endl;
  std::cout << "-------------------" << std::endl;
  std::cout << "sizeof(size_t): 8 bytes or 64 bits" << std::endl;
  std::cout << "Number of size_t to represent one Slater determinant: " << nw << std::endl;
  std::cout << "Extended precision scalars: double-double; z: 128-bit fixed point" << std::endl << std::endl;

  mcdfci_report r0{};
  check(c, mcdfci_report_get(c, &r0));
  const int bytes_per_det = nw * 8 + 16;
  const size_t slots = (size_t)r0.table_slots;
  cdo["max_wavefunction_size"] = Value::of_uint((unsigned long long)(0.79 * (double)slots));
  std::cout << "Input option" << std::endl;
  std::cout << "------------" << std::endl;
  std::cout << mcjson::dump(option, 4) << std::endl;
  std::cout << "Note:" << std::endl;
  std::cout << "In the wavefunction, each determinant needs " << bytes_per_det << " bytes." << std::endl;
  std::cout << "The size of the wavefunction is 2^" << (int)std::log2((double)slots) << " = " << slots
            << ", which can store " << (size_t)(0.79 * (double)slots) << " determinants at most." << std::endl;
  std::cout << "The wavefunction uses about " << std::fixed << std::setprecision(3)
            << (double)slots * bytes_per_det / 1073741824 << " GB memory." << std::endl
            << std::endl;

  if (cdo["coordinate_pick"].kind != Value::String || cdo["coordinate_pick"].s != "gcd_grad") {
    std::cout << "Error: the coordinate pick method " << mcjson::dump(cdo["coordinate_pick"], 4)
              << " is not implemented." << std::endl;
    std::exit(-1);
  }
  if (cdo["coordinate_update"].kind != Value::String ||
      (cdo["coordinate_update"].s != "eig" && cdo["coordinate_update"].s != "ls")) {
    std::cout << "Error: the coordinate update method " << mcjson::dump(cdo["coordinate_update"], 4)
              << " is not implemented." << std::endl;
    std::exit(-1);
  }

  const int k = o.num_of_coordinate;
  const long num_iter = (long)cdo["num_iterations"].integer();
  const long report_interval = (long)cdo["report_interval"].integer();
  const bool zts = cdo["z_threshold_search"].boolean();
  double z_threshold = cdo["z_threshold"].num();

  std::cout << "CDFCI calculation" << std::endl;
  std::cout << "-----------------" << std::endl;
  std::cout << "N = " << nw << std::endl;
  uint64_t ref[2] = {0, 0};
  const Value& occ = cdo["ref_det_occ"];
  if (occ.kind != Value::Array || occ.a.empty()) {
    if (!info[6]) std::cout << "Warning: no orbital energy found. Assume orbitals are sorted." << std::endl << std::endl;
    check(c, mcdfci_hartree_fock(c, ref));
  } else {
    int count = 0;
    for (const Value& v : occ.a) {
      long orb = (long)v.integer();
      if (orb >= 0 && orb < nso) {
        uint64_t bit = 1ULL << (orb & 63);
        if (!(ref[orb >> 6] & bit)) ++count;
        ref[orb >> 6] |= bit;
      }
    }
    if (count != (int)info[2])
      throw std::invalid_argument(
          "Reference determinant is invalid. The number of valid orbitals does not match the number of "
          "electrons. Please provide " + std::to_string(info[2]) + " orbitals from 0 to " +
          std::to_string(nso - 1) + " or remove the argument.");
  }
  double hf_energy;
  check(c, mcdfci_submatrix(c, 1, ref, &hf_energy));
  std::cout << "Hartree Fock determinant occupied spin-orbitals:" << std::endl;
  for (int s = 0; s < nso; ++s)
    if ((ref[s >> 6] >> (s & 63)) & 1) std::cout << s << "  ";
  std::cout << std::endl;
  std::cout << "Hartree Fock energy: " << std::fixed << std::setprecision(10) << hf_energy << std::endl
            << std::endl;

  const bool timing = std::getenv("MCDFCI_TIMING") != nullptr;
  double energy = 0.0;
  while (true) {
    check(c, mcdfci_set_z_threshold(c, z_threshold));
    check(c, mcdfci_init(c, ref));
    check(c, mcdfci_set_timing(c, timing ? 1 : 0));
    // Report blocks replay graphs of captured iterations (no effect on results).
    check(c, mcdfci_set_graph(c, (int)std::min<long>(report_interval, 32)));
    std::cout << std::setw(13) << std::left << "Iteration";
    std::cout << std::setw(18) << std::right << "Energy";
    std::cout << std::setw(18) << "|dx|_max";
    std::cout << std::setw(12) << "|x|_0";
    std::cout << std::setw(12) << "|z|_0";
    std::cout << std::setw(10) << "|H[I]|_0";
    std::cout << std::setw(10) << "upds";
    std::cout << std::setw(10) << "Time";
    std::cout << std::setw(10) << "scale";
    std::cout << std::endl;
    auto t0 = std::chrono::high_resolution_clock::now();
    long long last_z = 0;
    bool restart = false;
    for (long i = 0; i < num_iter / report_interval; ++i) {
      check(c, mcdfci_step(c, report_interval));
      mcdfci_report r{};
      int rc2 = mcdfci_report_get(c, &r);
      check(c, rc2);
      energy = r.energy;
      double el = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
      std::cout << std::setw(13) << std::left << (i + 1) * report_interval;
      std::cout << std::setw(18) << std::right << std::fixed << std::setprecision(10) << energy;
      std::cout << std::setw(18) << std::scientific << std::setprecision(4) << r.dx_max;
      std::cout << std::setw(12) << r.x_size;
      std::cout << std::setw(12) << r.z_size;
      std::cout << std::setw(10) << r.cols_size;
      std::cout << std::setw(10) << r.updates;
      std::cout << std::setw(10) << std::fixed << std::setprecision(2) << el;
      std::cout << std::setw(10) << std::fixed << std::setprecision(6) << r.scale;
      std::cout << std::endl;
      printf(
          "[LOG] max residual: %e, max eigenvector entry: %e, number of iterations provoking inv iters: %d, "
          "avg number of inv iters: %.1f, curr number of inv iters: %d, avg diff inv iter makes: %e, curr diff "
          "inv iter makes: %e, max diff inv iter makes: %e.\n",
          r.max_residual, r.curr_max_abs_evec, (int)r.sum_iterations,
          ((double)r.sum_inv_iter) / (double)r.sum_iterations, (int)r.curr_inv_iter,
          r.sum_diff_inv_iter / (double)r.sum_iterations, r.curr_diff_inv_iter, r.max_diff_inv_iter);
      if (r.eig_unconverged > 0)
        printf("[LOG] subspace solves used without passing the residual test: %lld.\n",
               (long long)r.eig_unconverged);
      fflush(stdout);
      if (zts) {
        double inc = 1000.0 * (double)(r.z_size - last_z) / (double)r.candidates / (double)report_interval;
        if (inc >= 1 && (double)r.z_size > 0.7 * r.max_size) {
          restart = true;
          break;
        }
      }
      last_z = r.z_size;
      if ((double)r.z_size > 0.79 * r.max_size)
        throw std::overflow_error("The hash table is full. Please increase max_memory or z_threshold.");
    }
    if (restart) {
      z_threshold *= 2;
      std::cout << "z_threshold_search: z_threshold is too small. Increase z_threhold to ";
      std::cout << std::scientific << std::setprecision(2) << z_threshold << " and restart." << std::endl;
      continue;
    }
    break;
  }
  double t6[6];
  check(c, mcdfci_timers(c, t6));
  std::cout << "[LOG] step timer: ";
  for (int i = 0; i < 6; ++i) printf("%.2f ", t6[i]);
  std::cout << std::endl;
  (void)k;
  return energy;
}

}  // namespace

int main(int argc, char* argv[]) {
  try {
    print_header();
    std::string input_file = check_argument(argc, argv);
    Value option = read_input(input_file);
    run(option);
    return 0;
  } catch (const std::exception& e) {
    std::cout.flush();
    std::cerr << std::endl << "ERROR: " << e.what() << std::endl;
    return 1;
  }
}
