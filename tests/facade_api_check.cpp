// Compile-only check (tests/test_frontend_cpu.py): the reference driver's
// calls on the Block3d surface (src/cbns3d.cpp:122-262), written as that
// driver makes them, compile against the facade.  In particular the
// checkpoint name gets the rank suffix before read_bin / output_bin see it
// (src/cbns3d.cpp:132-134), and output_vtr takes a file name only
// (include/Block3d.h:141).  Never linked or run.
#include <iomanip>
#include <sstream>

#include "Block3d.h"
#include "output_vtk_xml.h"
#include "read_input.h"

int facade_surface(ProcessInfo& proc_info, SimulationParams& simParam, size_type IM, size_type JM, size_type KM) {
  Block3d block(&proc_info, &simParam, IM, JM, KM);
  block.gen_mesh(proc_info.coords[0] * IM, proc_info.coords[1] * JM, proc_info.coords[2] * KM);

  std::ostringstream oss;
  oss << std::setw(3) << std::setfill('0') << proc_info.rank;
  simParam.checkpoint_fname += oss.str() + ".dat";

  if (simParam.is_restart) {
    block.initial_condition();
    block.calc_conservative(block.get_Q());
  } else if (0 != block.read_bin(simParam.checkpoint_fname)) {
    return 1;
  }
  boundary_condition(&block, block.get_Q());
  block.calc_primitive(block.get_Q());
  block.init_Q_p();
  if (0 != output_vtk_xml(&block, proc_info.rank, 1, "output000", "output")) return 1;

  for (size_type n_t = 0; n_t < simParam.nstep_max; n_t++) {
    if (block.is_finished()) break;
    value_type dt = block.calc_dt();
    simParam.t_cur += dt;
    block.rec_riemann_x(block.get_Q());
    block.rec_riemann_y(block.get_Q());
    block.rec_riemann_z(block.get_Q());
    block.calc_viscous_terms();
    block.calc_viscous_flux();
    block.calc_viscous_flux_contribution();
    block.update_rk3(dt, 1);
    boundary_condition(&block, block.get_Q_p());
    block.calc_primitive(block.get_Q_p());
    if (0 == (n_t + 1) % simParam.checkpoint_freq && 0 != block.output_bin(simParam.checkpoint_fname)) return 1;
  }
  if (0 != block.output_vtr("output000/output000.vtr")) return 1;
  return block.output_vtm("output000/output.vtm", "output", 1);
}

int facade_read(SimulationParams& simParam) { return read_input(simParam); }
