// read_input: src/read_input.cpp:15-103 restated for the facade.
#include "read_input.h"

#include <fstream>
#include <iostream>

int read_input(SimulationParams& params) { return read_input(params, "input.txt"); }

int read_input(SimulationParams& params, const std::string& path) {
  std::ifstream fh(path);
  if (!fh) {
    std::cout << "Exception opening/reading file\n" << path << std::endl;
    return -1;
  }
  value_type T_inf{0.0};
  int is_restart = 0;
  // src/read_input.cpp:24-72: fixed order, one value per token
  auto rd = [&](const char* name, auto& v) {
    fh >> v;
    std::cout << name << " : " << v << std::endl;
  };
  rd("checkpoint_fname", params.checkpoint_fname);
  rd("CFL", params.CFL);
  rd("t_cur", params.t_cur);
  rd("t_end", params.t_end);
  rd("Pr", params.Pr);
  rd("gamma", params.gamma);
  rd("Mach", params.Mach);
  rd("Re", params.Re);
  rd("T_inf", T_inf);
  rd("L", params.L);
  rd("IM", params.IM);
  rd("JM", params.JM);
  rd("KM", params.KM);
  rd("nstep_max", params.nstep_max);
  rd("checkpoint_freq", params.checkpoint_freq);
  rd("num_output", params.num_output);
  rd("is_restart", is_restart);
  if (!fh) {
    std::cout << "Exception opening/reading file\n" << path << std::endl;
    return -1;
  }
  params.is_restart = is_restart != 0;
  params.T_inf = T_inf;
  cbns_params p = to_cbns_params(params);
  cbns_params_derive(&p);  // src/read_input.cpp:78-90
  params.Mach2 = p.Mach2;
  params.Re_inv = p.Re_inv;
  params.Pr_inv = p.Pr_inv;
  params.gam1 = p.gam1;
  params.gM2 = p.gM2;
  params.gam1_inv = p.gam1_inv;
  params.p_inf = p.p_inf;
  params.C_T_inf = p.C_T_inf;
  params.C_dt_v = p.C_dt_v;
  return 0;
}

cbns_params to_cbns_params(const SimulationParams& s) {
  cbns_params p{};
  p.CFL = s.CFL;
  p.t_cur = s.t_cur;
  p.t_end = s.t_end;
  p.Pr = s.Pr;
  p.gamma = s.gamma;
  p.Mach = s.Mach;
  p.Re = s.Re;
  p.T_inf = s.T_inf;
  p.L = s.L;
  p.IM = s.IM;
  p.JM = s.JM;
  p.KM = s.KM;
  p.nstep_max = s.nstep_max;
  p.checkpoint_freq = s.checkpoint_freq;
  p.num_output = s.num_output;
  p.is_restart = s.is_restart ? 1 : 0;
  p.Mach2 = s.Mach2;
  p.gM2 = s.gM2;
  p.Re_inv = s.Re_inv;
  p.p_inf = s.p_inf;
  p.gam1 = s.gam1;
  p.gam1_inv = s.gam1_inv;
  p.Pr_inv = s.Pr_inv;
  p.C_T_inf = s.C_T_inf;
  p.C_dt_v = s.C_dt_v;
  return p;
}
