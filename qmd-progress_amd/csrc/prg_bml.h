// The library's own dense matrix handle standing in for BML's bml_matrix_t.
#pragma once

#include "prg_internal.h"
#include "../../include/progress_gpu.h"

struct bml_matrix_t {
  int N = 0;
  int M = 0;
  int type = 0;
  double* host = nullptr;  // N*N column-major
  double* dev = nullptr;   // ld*ld column-major, zero padded
  int ld = 0;
  bool host_valid = true;
  bool dev_valid = false;
};

namespace prg {
// Device view of A with the host contents uploaded if they are newer.
int bml_dev_read(bml_matrix_t* A, DMat* out);
// Device view of A for overwriting; marks the device copy as current.
int bml_dev_write(bml_matrix_t* A, DMat* out);
// Bring the host copy up to date.
int bml_host_sync(bml_matrix_t* A);
bml_matrix_t* bml_new(int N, int M, int type);
// Convenience check used by the C-ABI wrappers.
int check_pair(bml_matrix_t* a, bml_matrix_t* b, const char* who);
}  // namespace prg
