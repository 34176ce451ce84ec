// actflow <dir> -- drop-in replacement for the reference executable
// `ActFlow_V3-2 <dir>` (ActFlow_V3-2.cu:214-877): same argv forms, same
// specifications.txt, same snapshot and energy files, same messages and exit
// codes; the time step runs on the GPU through libactflow.so.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/actflow.h"
#include "file_writer.hpp"

namespace {

using af_cli::Writer;

std::string snap_name(const std::string& dir, long long j) {
  return dir + "/ActTurb_snap" + std::to_string(j) + ".bin";
}

bool exists(const std::string& path) {
  std::ifstream f(path);
  return f.good();
}

void print_help() {
  std::cout << "\n--------------------------------------------------\n"
            << "actflow: active-turbulence vorticity solver (James-Wilczek model with the mean-flow ODE),\n"
            << "time step on an AMD MI355X GPU.\n\n"
            << "usage:\n"
            << "  ./actflow <dir>   read <dir>/specifications.txt and run; snapshots and energy.txt go to\n"
            << "                    save_directory (default <dir>)\n"
            << "  ./actflow spec    write specifications.txt with the default values here\n"
            << "  ./actflow math    write theory.tex with the equations that are solved\n"
            << "  ./actflow help    print this message\n\n"
            << "initial_condition: random [n] | blank | bandlimited seed amp kmax | continue [T [snap]]\n"
            << "extension keys: Gamma0: Gamma2: precision: dealias: mean_scheme: snapshot_precision:\n"
            << "--------------------------------------------------" << std::endl;
}

int write_math() {
  std::ofstream file("theory.tex", std::ios::out);
  if (!file.is_open()) {
    std::cerr << "Error: unable to create theory file" << std::endl;
    return 1;
  }
  file << "\\documentclass{article}\n\\begin{document}\n"
       << "Vorticity equation solved pseudo-spectrally with an integrating factor:\n"
       << "\\begin{equation}\n\\partial_t\\omega + \\lambda(\\vec v\\cdot\\nabla)\\omega = "
       << "\\sigma(-\\Delta)\\,\\omega - \\beta\\,\\hat z\\cdot\\nabla\\times(|\\vec v|^2\\vec v),\\qquad "
       << "\\sigma(k) = c_0 + c_2 k^2 + c_4 k^4\n\\end{equation}\n"
       << "with $c_0=-1-\\alpha$, $c_2=2$, $c_4=-1$ for $-(1+\\Delta)^2-\\alpha$, and the mean velocity\n"
       << "\\begin{equation}\n\\partial_t\\langle\\vec v\\rangle = c_0\\langle\\vec v\\rangle "
       << "-\\beta\\langle|\\vec v|^2\\vec v\\rangle\n\\end{equation}\n\\end{document}" << std::endl;
  std::cout << "a Latex file \"theory.tex\" has been created" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  af_spec spec;
  af_spec_defaults(&spec);
  std::string directory = "./";
  if (argc == 2) {
    if (!std::strcmp(argv[1], "help") || !std::strcmp(argv[1], "h")) {
      print_help();
      return 0;
    } else if (!std::strcmp(argv[1], "spec")) {
      if (af_spec_write_example("specifications.txt", &spec) != AF_OK) {
        std::cerr << "Error: unable to create example file" << std::endl;
        return 1;
      }
      std::cout << "example has been created" << std::endl;
      return 0;
    } else if (!std::strcmp(argv[1], "math")) {
      return write_math();
    }
    directory = argv[1];
  } else {
    std::cerr << "More than one argument is no longer supported. See the change log in the beginning of the "
                 "source file." << std::endl;
    return 1;
  }

  std::cout << "reading the specifications (path=" << directory << "/specifications.txt)" << std::endl;
  char err[2048] = {0};
  if (af_spec_read(directory.c_str(), &spec, err, sizeof err) != AF_OK) {
    std::cerr << err << std::endl;
    return 1;
  }
  directory = spec.directory;
  int M = spec.M;
  const int snap = spec.snap;
  const float dt = (float)spec.dt;
  std::cout << "--------------------------------------------------\n"
            << "T = " << (float)spec.T << "\ndt = " << dt << "\nM = " << M << "\nLx = " << (float)spec.Lx
            << "\nLy = " << (float)spec.Ly << "\nNx = " << spec.Nx << "\nNy = " << spec.Ny
            << "\ndx = " << (float)spec.Lx / spec.Nx << "\ndy = " << (float)spec.Ly / spec.Ny
            << "\nalpha = " << (float)spec.alpha << "\nbeta = " << (float)spec.beta
            << "\nlambda = " << (float)spec.lambda << "\ninitial condition: " << spec.initial
            << "\ntotal amount of snapshots: " << (snap > 0 ? M / snap : 0)
            << "\n--------------------------------------------------" << std::endl;
  if (snap <= 0) {
    std::cerr << "Error: dt_snap must be at least one time step" << std::endl;
    return 1;
  }

  af_params p;
  if (spec.has_gamma)
    af_params_jw(&p, spec.Nx, spec.Ny, spec.Lx, spec.Ly, spec.dt, spec.lambda, spec.alpha, spec.beta, spec.gamma0,
                 spec.gamma2);
  else
    af_params_reference(&p, spec.Nx, spec.Ny, spec.Lx, spec.Ly, spec.dt, spec.lambda, spec.alpha, spec.beta);
  p.precision = spec.precision;
  p.dealias = spec.dealias;
  p.mean_scheme = spec.mean_scheme;

  const size_t nc = (size_t)spec.Ny * (spec.Nx / 2 + 1);
  std::vector<double> w(2 * nc, 0.0);
  double Mv[2] = {0.0, 0.0};
  long long m = 1;
  const std::string initial = spec.initial;

  if (initial.substr(0, 6) == "random") {
    std::cout << "generating random initial condition" << std::endl;
    long long seed = (long long)std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    if (initial.find(' ') != std::string::npos) seed /= std::stoi(initial.substr(initial.find(' ') + 1));
    if (af_ic_random(spec.Nx, spec.Ny, (unsigned long long)seed, p.dealias, w.data()) != AF_OK) {
      std::cerr << "Error: " << af_last_error() << std::endl;
      return 1;
    }
  } else if (initial == "blank") {
    std::cout << "generating blank initial condition" << std::endl;
  } else if (initial.substr(0, 11) == "bandlimited") {
    std::cout << "generating band-limited initial condition" << std::endl;
    unsigned long long seed = 1;
    double amp = 1e-3, kmax = 10;
    std::sscanf(initial.c_str() + 11, "%llu %lf %lf", &seed, &amp, &kmax);
    if (af_ic_bandlimited(spec.Nx, spec.Ny, seed, amp, kmax, w.data()) != AF_OK) {
      std::cerr << "Error: " << af_last_error() << std::endl;
      return 1;
    }
  } else if (initial.substr(0, 8) == "continue") {
    int cont_snap = -1;
    for (int i = 0; i < 1000000; ++i) {
      if (exists(snap_name(directory, i))) cont_snap = i; else break;
    }
    if (cont_snap == -1) {
      std::cerr << "Error: Tried to continue from previous simulation, but no previous snapshot was found."
                << std::endl;
      return 1;
    }
    if (initial.find(' ') != std::string::npos) {
      const std::string args = initial.substr(initial.find(' ') + 1);
      if (args.find(' ') != std::string::npos) {
        cont_snap = std::stoi(args.substr(args.find(' ') + 1));
        M = (int)(cont_snap * snap + std::round(std::stof(args.substr(0, args.find(' '))) / dt));
      } else {
        M = (int)(cont_snap * snap + std::round(std::stof(args) / dt));
      }
    }
    std::cout << "continue execution from snapshot " << cont_snap << std::endl;
    m = (long long)cont_snap * snap;
    if (af_snapshot_read(snap_name(directory, cont_snap).c_str(), spec.Nx, spec.Ny, w.data(), Mv) != AF_OK) {
      std::cerr << "Error: unable to open previous simulation" << std::endl;
      return 1;
    }
    // The kx = Nx/2 column is not represented by the GPU backend (the
    // library refuses a state that carries it): drop it here, visibly.
    const size_t nxh = (size_t)spec.Nx / 2 + 1;
    bool nyq = false;
    for (int y = 0; y < spec.Ny; ++y) {
      double* v = w.data() + 2 * ((size_t)y * nxh + nxh - 1);
      nyq = nyq || v[0] != 0.0 || v[1] != 0.0;
      v[0] = v[1] = 0.0;
    }
    if (nyq)
      std::cerr << "Warning: snapshot " << cont_snap << " has a nonzero kx = Nx/2 column; it is set to zero "
                << "(not represented on this backend)" << std::endl;
  } else {
    std::cerr << "Error: unknown specification for the initial condition" << std::endl;
    return 1;
  }
  if (m == 1) {
    // written before the run is created, as the reference does (.cu:682-693)
    std::cout << "processing the initial condition" << std::endl;
    w[0] = 0;
    w[1] = 0;
    if (af_snapshot_write(snap_name(directory, 0).c_str(), spec.Nx, spec.Ny, w.data(), Mv,
                          spec.snapshot_precision) != AF_OK) {
      std::cerr << "Error: unable to create snapshot file" << std::endl;
      return 1;
    }
  }
  Writer writer;   // later files are written in order on its thread while the GPU steps

  af_run* run = nullptr;
  af_status st = af_run_create(&p, &run);
  if (st == AF_OK) st = af_run_set_state(run, w.data(), Mv);
  if (st != AF_OK) {
    std::cerr << "Error: " << af_status_str(st) << ": " << af_last_error() << std::endl;
    af_run_destroy(run);
    return 1;
  }
  // a failed write ends the run with the reference's message for that file
  auto write_failed = [&](std::string msg) {
    if (msg.empty()) return false;
    std::cerr << msg << std::endl;
    af_run_destroy(run);
    return true;
  };
  std::cout << "initializing the time stepping" << std::endl;
  const auto t0 = std::chrono::high_resolution_clock::now();
  const std::string energy_path = directory + "/energy.txt";
  std::vector<double> energy;
  while (m <= M) {
    if (write_failed(writer.failure())) return 1;
    // chunk up to the next snapshot step (inclusive) or the end
    long long next = ((m + snap - 1) / snap) * snap;
    if (next < m) next = m;
    if (next > M) next = M;
    const int k = (int)(next - m + 1);
    energy.assign(k, 0.0);
    long long nan_step = -1;
    st = af_run_step(run, k, energy.data(), &nan_step);
    const int good = (st == AF_ENAN) ? (int)(nan_step - (af_run_steps_done(run) - k)) : k;
    writer.energy(energy_path, m, dt, energy.data(), good < k ? good : k);
    if (st == AF_ENAN) {
      if (write_failed(writer.finish())) return 1;
      std::cerr << "\nError: array contains nan values after " << (m + good - 1) << " timesteps" << std::endl;
      af_run_destroy(run);
      return 1;
    }
    if (st != AF_OK) {
      std::cerr << "Error: " << af_status_str(st) << ": " << af_last_error() << std::endl;
      af_run_destroy(run);
      return 1;
    }
    m += k;
    const long long last = m - 1;
    if (last % snap == 0) {
      if (af_run_get_state(run, w.data(), Mv) != AF_OK) {
        writer.finish();
        std::cerr << "Error: unable to create snapshot file" << std::endl;
        af_run_destroy(run);
        return 1;
      }
      writer.snapshot(snap_name(directory, last / snap), spec.Nx, spec.Ny, w, Mv, spec.snapshot_precision);
      const std::chrono::duration<double> el = std::chrono::high_resolution_clock::now() - t0;
      std::cout << "\r" << last << " / " << M << " timesteps done in " << el.count() << "s, w[0]=("
                << (float)w[0] << "," << (float)w[1] << ")    " << std::flush;
    }
  }
  if (write_failed(writer.finish())) return 1;
  af_run_destroy(run);
  const std::chrono::duration<double> el = std::chrono::high_resolution_clock::now() - t0;
  std::cout << "\nprogram finished successfully after " << el.count() << "s.\n" << std::endl;
  return 0;
}
