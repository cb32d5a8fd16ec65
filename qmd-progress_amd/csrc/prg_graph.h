// Graph-partitioned SP2: packed batch plan shared by the host driver and the
// batched kernels.
#pragma once

#include "prg_internal.h"
#include "../../include/progress_gpu.h"

namespace prg {

struct PartDesc {
  size_t offset;   // first double of the part's padded m x m block in the packed buffer
  int m;           // padded size (multiple of 64)
  int n;           // core + halo size
  int core;        // core rows (first `core` indices)
  int idx_offset;  // first entry of the part's index list in the packed index array
};

struct BatchPlan {
  int nparts = 0;
  int ntiles = 0;
  int wide = 0;              // 1: tiles of 128-wide panels (see batched_sp2_kernel)
  int depth = 1;             // K slabs prefetched ahead (1 or 2)
  int kslab = 16;            // K slab depth (16 or 32)
  int grid = 0;              // workgroups of the persistent batched SP2 launch
  int split = 0;             // 0 FP64 products; 1 FP16 / 2 BF16 hi+lo split products
  void* planes[4] = {nullptr, nullptr, nullptr, nullptr};  // hi0, lo0, hi1, lo1 (split)
  size_t total = 0;          // doubles in one packed buffer
  PartDesc* d_parts = nullptr;
  int4* d_tiles = nullptr;
  int* d_queue_off = nullptr;  // 9 offsets: the tiles of XCD queue q
  int* d_idx = nullptr;
};

// Core+halo lists of every part from the device graph G (nonzero off-diagonal
// entries are edges), in the reference's discovery order; with H (may be
// null) the core nodes' neighbours in H join the halo before the double jump.
// Fills core_halo_index, lsize and llsize of gp (prg_halo.hip).
int halo_indices_device(const DMat& G, prg_graph_partitioning_t* gp, int djflag,
                        const DMat* H = nullptr);

int batched_gather(const DMat& H, const BatchPlan& plan, double* x, double mineval,
                   double maxeval);
// One SP2 iteration of every part: xn from x, per-tile vv partials in
// d_part; d_tickets: 8 queue counters, zero before the launch.  Split plans:
// parity selects the planes of x (planes[2 parity], [2 parity + 1]); those of
// xn are the others.
int batched_step(const BatchPlan& plan, const double* x, double* xn, int sigma, double* d_part,
                 int* d_tickets, int parity = 0);
// hi / lo planes 0 and 1 of a split plan from the packed FP64 x
int batched_split_prepare(const BatchPlan& plan, const double* x);
// whether a kernel is compiled for plan's wide / depth / kslab
bool batched_config_valid(const BatchPlan& plan);
// plan->grid = resident workgroups of the kernel that wide/depth/kslab select
int batched_slots(BatchPlan* plan);
// d_tickets[0, count) = 0 by a kernel (safe inside a captured graph)
int batched_zero_tickets(int* d_tickets, int count);
// vv[it] += sum of d_part[it * ntiles + t] over the tiles, it < iters.
int batched_vv(const BatchPlan& plan, const double* d_part, int iters, double* d_vv);
// rho[core rows, core+halo columns] = scale x (rho zeroed first unless zero
// is false)
int batched_scatter(const BatchPlan& plan, const double* x, DMat rho, double scale,
                    bool zero = true);
// X_p = H[idx, idx], padding decoupled with `pad` on its diagonal
int batched_gather_pad(const DMat& H, const BatchPlan& plan, double* x, double pad);
// U_p = V_p diag(2 f(w_p)), dnocc[p] = sum 2 df/def(w_p)  (finite-T occupations)
int batched_fermi_scale(const BatchPlan& plan, const double* v, const double* w,
                        const size_t* d_w_offset, double* u, double ef, double kbt,
                        double* d_dnocc);

// U_p = V_p diag(w_p^-1/2), wmin[p] = min w_p  (Loewdin factor of an overlap)
int batched_invsqrt_scale(const BatchPlan& plan, const double* v, const double* w,
                          const size_t* d_w_offset, double* u, double* d_wmin);

}  // namespace prg
