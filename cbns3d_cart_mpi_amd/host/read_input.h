// read_input (reference: include/read_input.h:14, src/read_input.cpp:15-103).
#ifndef CBNS_HOST_READ_INPUT_H
#define CBNS_HOST_READ_INPUT_H

#include "simulation_params.h"

// Reads the 17 positional values of ./input.txt (the order python/parser_ini.py
// writes them in), echoes each as the reference does, and derives Mach2,
// Re_inv, Pr_inv, gam1, gM2, gam1_inv, p_inf, C_T_inf and C_dt_v.
// Returns 0 on success and -1 if the file cannot be opened or read (the
// reference returns 0 with garbage values in that case; see the design notes).
int read_input(SimulationParams& params);
int read_input(SimulationParams& params, const std::string& path);

#endif
