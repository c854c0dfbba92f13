// output_vtk_xml (reference: include/output_vtk_xml.h:23-26, src/output_vtk_xml.cpp:22-48).
#ifndef CBNS_HOST_OUTPUT_VTK_XML_H
#define CBNS_HOST_OUTPUT_VTK_XML_H

#include <string>

#include "Block3d.h"

// Writes dir_path/fname%03d.vtr for every on-device block and, for rank 0,
// dir_path/fname.vtm listing them.  Returns 0 on success, -1 on failure.
int output_vtk_xml(Block3d* blk, const int rank, const size_type num_blocks, const std::string dir_path,
                   const std::string fname);

#endif
