// Drop-in facade of the reference's include/BCMTools.h (face/subface
// enums, exit codes, Exit macro, assert with NDEBUG unless DEBUG, the MPI
// header) for drivers built against libbcm_hip.
#ifndef BCM_FACADE_BCMTOOLS_H
#define BCM_FACADE_BCMTOOLS_H

#include <cstdio>
#include <cstdlib>

#include "bcm_hip.h"
#include "mpi.h"

// As the reference: assertions are off unless DEBUG is defined.
#ifndef DEBUG
#ifndef NDEBUG
#define NDEBUG
#endif
#endif
#include <cassert>

// Exit(x): print the call site and leave (include/BCMTools.h:36-37).
#define Exit(x) ((void)std::printf("exit at %s:%u\n", __FILE__, __LINE__), std::exit((x)))

enum Face { X_M = BCM_X_M, X_P, Y_M, Y_P, Z_M, Z_P, NUM_FACE };
enum Subface { SF_00 = BCM_SF_00, SF_01, SF_10, SF_11, NUM_SUBFACE };

enum ExitStatus {
  EX_SUCCESS = 0,
  EX_USAGE = 16,
  EX_MEMORY = 17,
  EX_OPEN_FILE = 18,
  EX_READ_CONFIG = 19,
  EX_READ_DATA = 20,
  EX_WRITE_DATA = 21,
  EX_FAILURE = 1,
};

#endif
