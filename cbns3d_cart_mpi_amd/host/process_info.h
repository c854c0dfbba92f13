// ProcessInfo (reference: include/process_info.h:12-26) without MPI: the
// Cartesian decomposition now describes on-device blocks of one process.
#ifndef CBNS_HOST_PROCESS_INFO_H
#define CBNS_HOST_PROCESS_INFO_H

struct ProcessInfo {
  int rank;       // this process (always 0: one process drives one GPU)
  int coords[3];  // of block 0
  int dims[3];    // on-device blocks per axis (the reference's MPI dims)
  // periodic neighbours of block 0, as MPI_Cart_shift would give them
  int left, right, down, up, back, front;
};

#endif
