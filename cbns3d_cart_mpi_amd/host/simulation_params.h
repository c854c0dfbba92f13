// Host facade types of the cbns3d driver surface.
//
// Same names and meanings as the reference's include/data_type.h:4-6,
// include/constants.h:11-21 and include/simulation_params.h:15-53, so code
// written against that surface compiles unchanged against this one.
#ifndef CBNS_HOST_SIMULATION_PARAMS_H
#define CBNS_HOST_SIMULATION_PARAMS_H

#include <string>

#include "cbns3d_amd.h"

typedef double value_type;
typedef unsigned int size_type;
typedef int index_type;

namespace constant {
const size_type NEQ = CBNS_NEQ;
const size_type NGI = 3;
const size_type NGV = CBNS_NGV;
const size_type NG = CBNS_NG;
const value_type PI = 3.14159265358979323846;
}  // namespace constant

struct SimulationParams {
  std::string checkpoint_fname;

  size_type IM;
  size_type JM;
  size_type KM;

  value_type CFL;

  value_type t_cur;
  value_type t_end;

  size_type nstep_max;
  size_type checkpoint_freq;
  size_type num_output;

  bool is_restart;

  value_type L;

  value_type Mach;
  value_type Re;
  value_type Mach2;
  value_type gM2;
  value_type Re_inv;

  value_type p_inf;

  value_type gamma;
  value_type Pr;
  value_type gam1;
  value_type gam1_inv;
  value_type Pr_inv;

  value_type C_T_inf;
  value_type C_dt_v;

  // T_inf itself is not a field of the reference's struct (only C_T_inf is);
  // the facade keeps it so the C ABI can derive C_T_inf the same way.
  value_type T_inf;
};

// SimulationParams -> the C ABI's POD parameters.
cbns_params to_cbns_params(const SimulationParams& s);

#endif
