// cbns3d_amd: drop-in driver for the reference's main (src/cbns3d.cpp:27-308).
//
//   cbns3d_amd <dimX> <dimY> <dimZ> [--fused] [--mu-sidecar] [--async-io]
//
// Reads ./input.txt exactly as the reference does, splits the domain into
// dimX x dimY x dimZ blocks (held on one GPU instead of one per MPI rank),
// and runs the same loop with the same output: the per-step "n_t, dt, t_cur"
// line, checkpoint files checkpointNNN.dat per block, VTK output directories
// outputNNN with one .vtr per block and a .vtm, and the wall/CPU time lines.
// --fused runs each step as one device call instead of the reference's
// method-by-method sequence (same arithmetic).  --mu-sidecar writes each
// checkpoint's mu beside it (Block3d::set_mu_sidecar), so that a restart
// from it continues bit for bit as the uninterrupted run.  --async-io writes
// the checkpoint and VTK files from a device snapshot on writer threads
// while the next steps run (Block3d::set_async_io; the same files); the
// wall time then includes waiting for the writes still pending at the end.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "Block3d.h"
#include "output_vtk_xml.h"
#include "read_input.h"

int main(int argc, char** argv) {
  if (4 > argc) {
    std::cout << "Usage: program <dimX> <dimY> <dimZ>\n";
    return 1;
  }
  ProcessInfo proc_info{};
  proc_info.dims[0] = std::stoi(argv[1]);
  proc_info.dims[1] = std::stoi(argv[2]);
  proc_info.dims[2] = std::stoi(argv[3]);
  bool fused = false, sidecar = false, async_io = false;
  for (int a = 4; a < argc; a++) {
    if (std::strcmp(argv[a], "--fused") == 0)
      fused = true;
    else if (std::strcmp(argv[a], "--mu-sidecar") == 0)
      sidecar = true;
    else if (std::strcmp(argv[a], "--async-io") == 0)
      async_io = true;
    else {
      std::cout << "unknown option " << argv[a] << "\n";
      return 1;
    }
  }
  const int nblocks = proc_info.dims[0] * proc_info.dims[1] * proc_info.dims[2];

  SimulationParams simParam;
  if (0 != read_input(simParam)) return 1;

  auto create_dir = [](const char* path) {
    if (!std::filesystem::exists(path) && !std::filesystem::create_directory(path)) {
      std::cerr << "Failed to create directory: " << path << std::endl;
      std::exit(1);
    }
  };
  char dirName[20];
  for (size_type i = 0; i <= simParam.num_output; i++) {
    std::snprintf(dirName, 20, "output%.3i", i);
    create_dir(dirName);
  }

  // block 0's coordinates and periodic neighbours (src/cbns3d.cpp:83-97)
  proc_info.rank = 0;
  proc_info.coords[0] = proc_info.coords[1] = proc_info.coords[2] = 0;
  auto id = [&](int x, int y, int z) {
    auto w = [](int a, int n) { return ((a % n) + n) % n; };
    return (w(x, proc_info.dims[0]) * proc_info.dims[1] + w(y, proc_info.dims[1])) * proc_info.dims[2] +
           w(z, proc_info.dims[2]);
  };
  proc_info.left = id(-1, 0, 0);
  proc_info.right = id(1, 0, 0);
  proc_info.down = id(0, -1, 0);
  proc_info.up = id(0, 1, 0);
  proc_info.back = id(0, 0, -1);
  proc_info.front = id(0, 0, 1);

  Block3d block(&proc_info, &simParam, simParam.IM, simParam.JM, simParam.KM);
  block.gen_mesh(0, 0, 0);
  block.set_mu_sidecar(sidecar);
  block.set_async_io(async_io);

  // src/cbns3d.cpp:132-134: the checkpoint name of this process's rank; the
  // facade derives every block's file from it
  std::ostringstream oss;
  oss << std::setw(3) << std::setfill('0') << proc_info.rank;
  simParam.checkpoint_fname += oss.str() + ".dat";

  // src/cbns3d.cpp:136-144 (is_restart = 1 means a fresh Taylor-Green start)
  if (simParam.is_restart) {
    block.initial_condition();
    block.calc_conservative(block.get_Q());
  } else {
    if (0 != block.read_bin(simParam.checkpoint_fname)) {
      std::cerr << "Unable to open the checkpoint file by process " << 0 << "\n";
      return 1;
    }
  }
  boundary_condition(&block, block.get_Q());
  block.calc_primitive(block.get_Q());
  block.init_Q_p();
  if (block.last_status() != CBNS_OK) {
    std::cerr << "device setup failed: " << cbns_strerror(block.last_status()) << "\n";
    return 1;
  }

  size_type cnt_out = 0;
  oss.str("");
  oss << std::setw(3) << std::setfill('0') << cnt_out;
  if (0 != output_vtk_xml(&block, 0, nblocks, "output" + oss.str(), "output")) return 1;

  value_type dt_out = simParam.t_end / simParam.num_output;
  value_type t_out = ((int)(simParam.t_cur / dt_out) + 1) * dt_out;

  auto start = std::chrono::high_resolution_clock::now();
  for (size_type n_t = 0; n_t < simParam.nstep_max; n_t++) {
    if (block.is_finished()) {
      std::cout << "Finished!\n";
      break;
    }
    value_type dt = 0.0;
    if (fused) {
      if (block.step(&dt) != CBNS_OK) {
        std::cerr << "step failed: " << cbns_strerror(block.last_status()) << "\n";
        return 1;
      }
    } else {
      dt = block.calc_dt();
      if (simParam.t_end - simParam.t_cur < dt) dt = simParam.t_end - simParam.t_cur;
      simParam.t_cur += dt;
    }
    std::cout << n_t << ", " << dt << ", " << simParam.t_cur << std::endl;
    if (!fused) {
      for (size_type stage = 1; stage <= 3; stage++) {
        value_type* in = stage == 1 ? block.get_Q() : block.get_Q_p();
        value_type* out = stage == 3 ? block.get_Q() : block.get_Q_p();
        block.rec_riemann_x(in);
        block.rec_riemann_y(in);
        block.rec_riemann_z(in);
        block.calc_viscous_flux_contribution();
        block.update_rk3(dt, stage);
        boundary_condition(&block, out);
        block.calc_primitive(out);
      }
      if (block.last_status() != CBNS_OK) {
        std::cerr << "step failed: " << cbns_strerror(block.last_status()) << "\n";
        return 1;
      }
    }
    if (0 == (n_t + 1) % simParam.checkpoint_freq) {
      if (0 != block.output_bin(simParam.checkpoint_fname)) {
        std::cerr << "Unable to save the checkpoint file by process 0\n";
        return 1;
      }
    }
    if (t_out < simParam.t_cur) {
      cnt_out++;
      oss.str("");
      oss << std::setw(3) << std::setfill('0') << cnt_out;
      if (0 != output_vtk_xml(&block, 0, nblocks, "output" + oss.str(), "output")) return 1;
      t_out += dt_out;
    }
  }
  cbns_sync(block.context());
  if (0 != block.io_wait()) {
    std::cerr << "Unable to write an output file\n";
    return 1;
  }
  auto end = std::chrono::high_resolution_clock::now();
  double wall_time = std::chrono::duration_cast<std::chrono::duration<double>>(end - start).count();
  std::cout << "  Wall time: " << wall_time << " seconds\n"
            << "  CPU time: " << wall_time << " seconds\n";

  cnt_out++;
  oss.str("");
  oss << std::setw(3) << std::setfill('0') << cnt_out;
  if (0 != output_vtk_xml(&block, 0, nblocks, "output" + oss.str(), "output")) return 1;
  if (0 != block.io_wait()) {
    std::cerr << "Unable to write an output file\n";
    return 1;
  }
  return 0;
}
