// output_vtk_xml: src/output_vtk_xml.cpp:22-48 for on-device blocks.
#include "output_vtk_xml.h"

#include <cstdio>
#include <iostream>

int output_vtk_xml(Block3d* blk, const int rank, const size_type num_blocks, const std::string dir_path,
                   const std::string fname) {
  const std::string output_fname = dir_path + "/" + fname;
  for (size_type b = 0; b < num_blocks; b++) {
    char nm[16];
    std::snprintf(nm, sizeof nm, "%03d", (int)b);
    if (0 != blk->output_vtr(output_fname + nm + ".vtr", (int)b)) {
      std::cerr << "Unable to open the .vtr file by process " << rank << "\n";
      return -1;
    }
  }
  if (0 == rank) {
    if (0 != blk->output_vtm(output_fname + ".vtm", fname, num_blocks)) {
      std::cerr << "Unable to open the .vtm file\n";
      return -1;
    }
  }
  return 0;
}
