// Dense stand-in for the BML handle (lanl/bml is not available to this
// build).  The host copy is what bml_* accessors see; prg_* routines work on
// an HBM mirror that is synchronised lazily in either direction, so chained
// solver calls on the same handles never leave the GPU.
#include "prg_bml.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace prg {

bml_matrix_t* bml_new(int N, int M, int type) {
  if (N <= 0) {
    set_error(ERR_ARG, "bml: N must be positive");
    return nullptr;
  }
  bml_matrix_t* A = new bml_matrix_t;
  A->N = N;
  A->M = M;
  A->type = type;
  A->host = (double*)std::calloc((size_t)N * N, sizeof(double));
  if (!A->host) {
    delete A;
    set_error(ERR_ARG, "bml: host allocation failed");
    return nullptr;
  }
  A->ld = padded(N);
  A->host_valid = true;
  A->dev_valid = false;
  return A;
}

static int ensure_dev_alloc(bml_matrix_t* A) {
  if (A->dev) return OK;
  int st = ensure_device();
  if (st) return st;
  const size_t bytes = sizeof(double) * (size_t)A->ld * A->ld;
  PRG_HIP_CHECK(hipMalloc(&A->dev, bytes));
  PRG_HIP_CHECK(hipMemsetAsync(A->dev, 0, bytes, stream()));
  return OK;
}

// Host <-> device copy of the N x N block through the pinned staging buffer
// (pageable host memory is never handed to the copy engine), in column
// blocks of at most 64 MiB, each complete before the buffer is reused.
static int copy_columns(bml_matrix_t* A, bool to_device) {
  const size_t col_bytes = sizeof(double) * (size_t)A->N;
  const int per = (int)std::max<size_t>(1, ((size_t)64 << 20) / col_bytes);
  int st = OK;
  double* stage = (double*)host_staging(col_bytes * std::min(per, A->N), &st);
  if (st) return st;
  for (int j0 = 0; j0 < A->N; j0 += per) {
    const int nc = std::min(per, A->N - j0);
    double* h = A->host + (size_t)j0 * A->N;
    double* d = A->dev + (size_t)j0 * A->ld;
    if (to_device) {
      std::memcpy(stage, h, col_bytes * nc);
      PRG_HIP_CHECK(hipMemcpy2DAsync(d, sizeof(double) * A->ld, stage, col_bytes, col_bytes, nc,
                                     hipMemcpyHostToDevice, stream()));
      PRG_HIP_CHECK(hipStreamSynchronize(stream()));
    } else {
      PRG_HIP_CHECK(hipMemcpy2DAsync(stage, col_bytes, d, sizeof(double) * A->ld, col_bytes, nc,
                                     hipMemcpyDeviceToHost, stream()));
      PRG_HIP_CHECK(hipStreamSynchronize(stream()));
      std::memcpy(h, stage, col_bytes * nc);
    }
  }
  return OK;
}

int bml_dev_read(bml_matrix_t* A, DMat* out) {
  if (!A) {
    set_error(ERR_ARG, "null matrix handle");
    return ERR_ARG;
  }
  int st = ensure_dev_alloc(A);
  if (st) return st;
  if (!A->dev_valid) {
    if ((st = copy_columns(A, true))) return st;
    A->dev_valid = true;
  }
  out->n = A->N;
  out->ld = A->ld;
  out->p = A->dev;
  return OK;
}

int bml_dev_write(bml_matrix_t* A, DMat* out) {
  if (!A) {
    set_error(ERR_ARG, "null matrix handle");
    return ERR_ARG;
  }
  int st = ensure_dev_alloc(A);
  if (st) return st;
  A->dev_valid = true;
  A->host_valid = false;
  out->n = A->N;
  out->ld = A->ld;
  out->p = A->dev;
  return OK;
}

int bml_host_sync(bml_matrix_t* A) {
  if (!A) {
    set_error(ERR_ARG, "null matrix handle");
    return ERR_ARG;
  }
  if (A->host_valid) return OK;
  const int st = copy_columns(A, false);
  if (st) return st;
  A->host_valid = true;
  return OK;
}

static void host_modified(bml_matrix_t* A) {
  A->host_valid = true;
  A->dev_valid = false;
}

int check_pair(bml_matrix_t* a, bml_matrix_t* b, const char* who) {
  if (!a || !b) {
    set_error(ERR_ARG, std::string(who) + ": null matrix handle");
    return ERR_ARG;
  }
  if (a->N != b->N) {
    set_error(ERR_ARG, std::string(who) + ": matrix sizes differ");
    return ERR_ARG;
  }
  return OK;
}

}  // namespace prg

using namespace prg;

extern "C" {

bml_matrix_t* bml_zero_matrix(bml_matrix_type_t matrix_type, bml_matrix_precision_t precision,
                              int N, int M, bml_distribution_mode_t distrib_mode) {
  (void)distrib_mode;
  clear_error();
  if (precision != double_real) {
    set_error(ERR_UNSUPPORTED, "bml_zero_matrix: only double_real is supported");
    return nullptr;
  }
  return bml_new(N, M, (int)matrix_type);
}

bml_matrix_t* bml_identity_matrix(bml_matrix_type_t matrix_type, bml_matrix_precision_t precision,
                                  int N, int M, bml_distribution_mode_t distrib_mode) {
  bml_matrix_t* A = bml_zero_matrix(matrix_type, precision, N, M, distrib_mode);
  if (!A) return nullptr;
  for (int i = 0; i < N; ++i) A->host[i + (size_t)i * N] = 1.0;
  return A;
}

bml_matrix_t* bml_import_from_dense(bml_matrix_type_t matrix_type,
                                    bml_matrix_precision_t precision, bml_dense_order_t order,
                                    int N, int M, void* Av, double threshold,
                                    bml_distribution_mode_t distrib_mode) {
  bml_matrix_t* A = bml_zero_matrix(matrix_type, precision, N, M, distrib_mode);
  if (!A) return nullptr;
  const double* src = (const double*)Av;
  if (!src) {
    set_error(ERR_ARG, "bml_import_from_dense: null data");
    bml_deallocate(&A);
    return nullptr;
  }
  if (order == dense_column_major) {
    std::memcpy(A->host, src, sizeof(double) * (size_t)N * N);
  } else {
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) A->host[i + (size_t)j * N] = src[(size_t)i * N + j];
  }
  // The sparse BML formats keep only the entries above the threshold
  // (|a| > threshold); stored dense here, the others become zeros.
  if (matrix_type != dense && threshold > 0.0) {
    for (size_t k = 0; k < (size_t)N * N; ++k)
      if (!(std::fabs(A->host[k]) > threshold)) A->host[k] = 0.0;
  }
  return A;
}

void* bml_export_to_dense(bml_matrix_t* A, bml_dense_order_t order) {
  clear_error();
  if (bml_host_sync(A)) return nullptr;
  const int N = A->N;
  double* out = (double*)std::malloc(sizeof(double) * (size_t)N * N);
  if (!out) return nullptr;
  if (order == dense_column_major) {
    std::memcpy(out, A->host, sizeof(double) * (size_t)N * N);
  } else {
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) out[(size_t)i * N + j] = A->host[i + (size_t)j * N];
  }
  return out;
}

void bml_copy(bml_matrix_t* A, bml_matrix_t* B) {
  clear_error();
  if (check_pair(A, B, "bml_copy")) return;
  if (A->dev_valid && !A->host_valid) {
    DMat a, b;
    if (bml_dev_read(A, &a) || bml_dev_write(B, &b)) return;
    copy_block(a, b);
    return;
  }
  std::memcpy(B->host, A->host, sizeof(double) * (size_t)A->N * A->N);
  host_modified(B);
}

bml_matrix_t* bml_copy_new(bml_matrix_t* A) {
  clear_error();
  if (!A) return nullptr;
  bml_matrix_t* B = bml_new(A->N, A->M, A->type);
  if (!B) return nullptr;
  bml_copy(A, B);
  return B;
}

bml_matrix_t* bml_transpose_new(bml_matrix_t* A) {
  clear_error();
  if (bml_host_sync(A)) return nullptr;
  bml_matrix_t* B = bml_new(A->N, A->M, A->type);
  if (!B) return nullptr;
  const int N = A->N;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) B->host[j + (size_t)i * N] = A->host[i + (size_t)j * N];
  return B;
}

void bml_add(bml_matrix_t* A, bml_matrix_t* B, double alpha, double beta, double threshold) {
  (void)threshold;
  clear_error();
  if (check_pair(A, B, "bml_add")) return;
  DMat a, b;
  if (bml_dev_read(A, &a) || bml_dev_read(B, &b)) return;
  A->host_valid = false;  // device copy becomes the current one
  axpby(a, alpha, b, beta, 0.0, a, nullptr, nullptr);
}

void bml_scale(void* scale_factor, bml_matrix_t* A, bml_matrix_t* B) {
  clear_error();
  if (check_pair(A, B, "bml_scale")) return;
  const double s = *(const double*)scale_factor;
  DMat a, b;
  if (bml_dev_read(A, &a)) return;
  if (A == B) {
    A->host_valid = false;
    affine(a, s, 0.0, a, nullptr, nullptr);
    return;
  }
  if (bml_dev_write(B, &b)) return;
  affine(a, s, 0.0, b, nullptr, nullptr);
}

double bml_trace(bml_matrix_t* A) {
  clear_error();
  if (bml_host_sync(A)) return NAN;
  double s = 0.0;
  for (int i = 0; i < A->N; ++i) s += A->host[i + (size_t)i * A->N];
  return s;
}

// BML's thresholding (entries with |a| <= threshold become zero), as the
// graph construction of examples/gptest/gptest.F90:100 uses it.
void bml_threshold(bml_matrix_t* A, double threshold) {
  clear_error();
  if (bml_host_sync(A)) return;
  for (size_t k = 0; k < (size_t)A->N * A->N; ++k)
    if (!(std::fabs(A->host[k]) > threshold)) A->host[k] = 0.0;
  host_modified(A);
}

// Tr(A B) = sum_ij A_ij B_ji (examples/gptest/gptest.F90:68,131: band energy).
double bml_trace_mult(bml_matrix_t* A, bml_matrix_t* B) {
  clear_error();
  if (check_pair(A, B, "bml_trace_mult") || bml_host_sync(A) || bml_host_sync(B)) return NAN;
  const int N = A->N;
  double s = 0.0;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) s += A->host[i + (size_t)j * N] * B->host[j + (size_t)i * N];
  return s;
}

// ||A - B||_F (examples/gptest/gptest.F90:144).
double bml_fnorm2(bml_matrix_t* A, bml_matrix_t* B) {
  clear_error();
  if (check_pair(A, B, "bml_fnorm2") || bml_host_sync(A) || bml_host_sync(B)) return NAN;
  double s = 0.0;
  for (size_t k = 0; k < (size_t)A->N * A->N; ++k) {
    const double d = A->host[k] - B->host[k];
    s += d * d;
  }
  return std::sqrt(s);
}

// Gershgorin interval as BML returns it: a malloc'ed pair {min, max} the
// caller frees (examples/gptest/gptest.F90:82).
void* bml_gershgorin(bml_matrix_t* A) {
  clear_error();
  double* g = (double*)std::malloc(2 * sizeof(double));
  if (!g) return nullptr;
  prg_gershgorin(A, g);
  if (last_error_code() != OK) {
    std::free(g);
    return nullptr;
  }
  return g;
}

int bml_get_N(bml_matrix_t* A) { return A ? A->N : -1; }
int bml_get_M(bml_matrix_t* A) { return A ? A->M : -1; }

double bml_get_element(bml_matrix_t* A, int i, int j) {
  clear_error();
  if (bml_host_sync(A)) return NAN;
  if (i < 0 || j < 0 || i >= A->N || j >= A->N) {
    set_error(ERR_ARG, "bml_get_element: index out of range");
    return NAN;
  }
  return A->host[i + (size_t)j * A->N];
}

void bml_read_bml_matrix(bml_matrix_t* A, char* filename) {
  clear_error();
  if (!A || !filename) {
    set_error(ERR_ARG, "bml_read_bml_matrix: null argument");
    return;
  }
  FILE* f = std::fopen(filename, "r");
  if (!f) {
    set_error(ERR_ARG, std::string("bml_read_bml_matrix: cannot open ") + filename);
    return;
  }
  char line[4096];
  if (!std::fgets(line, sizeof line, f) || !std::strstr(line, "MatrixMarket")) {
    std::fclose(f);
    set_error(ERR_ARG, "bml_read_bml_matrix: not a Matrix Market file");
    return;
  }
  const bool symmetric = std::strstr(line, "symmetric") != nullptr;
  do {
    if (!std::fgets(line, sizeof line, f)) {
      std::fclose(f);
      set_error(ERR_ARG, "bml_read_bml_matrix: truncated header");
      return;
    }
  } while (line[0] == '%');
  long nr = 0, nc = 0, nnz = 0;
  if (std::sscanf(line, "%ld %ld %ld", &nr, &nc, &nnz) != 3 || nr != A->N || nc != A->N) {
    std::fclose(f);
    set_error(ERR_ARG, "bml_read_bml_matrix: size does not match the handle");
    return;
  }
  std::memset(A->host, 0, sizeof(double) * (size_t)A->N * A->N);
  for (long e = 0; e < nnz; ++e) {
    long i, j;
    double v;
    if (std::fscanf(f, "%ld %ld %lf", &i, &j, &v) != 3 || i < 1 || j < 1 || i > nr || j > nc) {
      std::fclose(f);
      set_error(ERR_ARG, "bml_read_bml_matrix: bad entry");
      return;
    }
    A->host[(i - 1) + (size_t)(j - 1) * A->N] = v;
    if (symmetric) A->host[(j - 1) + (size_t)(i - 1) * A->N] = v;
  }
  std::fclose(f);
  host_modified(A);
}

void bml_write_bml_matrix(bml_matrix_t* A, char* filename) {
  clear_error();
  if (bml_host_sync(A)) return;
  FILE* f = std::fopen(filename, "w");
  if (!f) {
    set_error(ERR_ARG, std::string("bml_write_bml_matrix: cannot open ") + filename);
    return;
  }
  const int N = A->N;
  long nnz = 0;
  for (size_t k = 0; k < (size_t)N * N; ++k) nnz += A->host[k] != 0.0;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %ld\n", N, N, nnz);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double v = A->host[i + (size_t)j * N];
      if (v != 0.0) std::fprintf(f, "%d %d %.17g\n", i + 1, j + 1, v);
    }
  std::fclose(f);
}

void bml_deallocate(bml_matrix_t** A) {
  if (!A || !*A) return;
  if ((*A)->dev) {
    hipStreamSynchronize(stream());
    hipFree((*A)->dev);
  }
  std::free((*A)->host);
  delete *A;
  *A = nullptr;
}

}  // extern "C"
