// Graph partitioning and the subgraph SP2 loop
// (src/prg_graph_mod.F90, src/prg_subgraphloop_mod.F90).
//
// The partition structure is a flat C struct (include/progress_gpu.h) standing
// in for the reference's Fortran graph_partitioning_t.  Core+halo index lists
// are extracted on the host from the graph matrix (as the reference's OpenMP
// loop does); all SP2 work -- gather and normalisation of every part, the
// lock-step purification of all parts, and the scatter of the core rows into
// the density -- runs on the GPU as batched kernels.
#include "prg_graph.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "prg_bml.h"

namespace prg {
namespace {

void free_parts(prg_graph_partitioning_t* gp) {
  if (!gp) return;
  if (gp->nodeInPart)
    for (int i = 0; i < gp->totalParts; ++i) std::free(gp->nodeInPart[i]);
  if (gp->core_halo_index)
    for (int i = 0; i < gp->totalParts; ++i) std::free(gp->core_halo_index[i]);
  std::free(gp->nodeInPart);
  std::free(gp->core_halo_index);
  std::free(gp->nnodesInPart);
  std::free(gp->lsize);
  std::free(gp->llsize);
  std::memset(gp, 0, sizeof(*gp));
}

int init_parts(prg_graph_partitioning_t* gp, int nparts, int nnodes) {
  free_parts(gp);
  gp->totalParts = nparts;
  gp->totalNodes = nnodes;
  gp->nnodesInPart = (int*)std::calloc(nparts, sizeof(int));
  gp->nodeInPart = (int**)std::calloc(nparts, sizeof(int*));
  gp->lsize = (int*)std::calloc(nparts, sizeof(int));
  gp->llsize = (int*)std::calloc(nparts, sizeof(int));
  gp->core_halo_index = (int**)std::calloc(nparts, sizeof(int*));
  if (!gp->nnodesInPart || !gp->nodeInPart || !gp->lsize || !gp->llsize ||
      !gp->core_halo_index) {
    set_error(ERR_ARG, "graph partitioning: allocation failed");
    return ERR_ARG;
  }
  return OK;
}

float g_ms[3] = {0.f, 0.f, 0.f};

}  // namespace

// bml_matrix2submatrix_index as called at src/prg_subgraphloop_mod.F90:60-63:
// core nodes first (given order), then every node adjacent in G (nonzero,
// off-diagonal entry of the core node's row) in order of discovery, then with
// the double-jump flag the neighbours of those halo nodes.
int halo_indices(bml_matrix_t* g_bml, prg_graph_partitioning_t* gp, int djflag) {
  int st;
  if ((st = bml_host_sync(g_bml))) return st;
  const int N = g_bml->N;
  const double* G = g_bml->host;
  // row adjacency lists (ascending column order), built column by column
  std::vector<std::vector<int>> adj(N);
  for (int j = 0; j < N; ++j) {
    const double* col = G + (size_t)j * N;
    for (int i = 0; i < N; ++i)
      if (i != j && col[i] != 0.0) adj[i].push_back(j);
  }
  std::vector<char> seen(N, 0);
  std::vector<int> list;
  for (int p = 0; p < gp->totalParts; ++p) {
    list.clear();
    const int nc = gp->nnodesInPart[p];
    for (int k = 0; k < nc; ++k) {
      const int c = gp->nodeInPart[p][k];
      if (c < 0 || c >= N) {
        set_error(ERR_ARG, "graph partitioning: node index out of range");
        return ERR_ARG;
      }
      if (!seen[c]) {
        seen[c] = 1;
        list.push_back(c);
      }
    }
    const size_t ncore = list.size();
    for (size_t k = 0; k < ncore; ++k)
      for (int nb : adj[list[k]])
        if (!seen[nb]) {
          seen[nb] = 1;
          list.push_back(nb);
        }
    if (djflag) {
      const size_t nh = list.size();
      for (size_t k = ncore; k < nh; ++k)
        for (int nb : adj[list[k]])
          if (!seen[nb]) {
            seen[nb] = 1;
            list.push_back(nb);
          }
    }
    for (int v : list) seen[v] = 0;
    std::free(gp->core_halo_index[p]);
    gp->core_halo_index[p] = (int*)std::malloc(sizeof(int) * std::max<size_t>(1, list.size()));
    std::memcpy(gp->core_halo_index[p], list.data(), sizeof(int) * list.size());
    gp->lsize[p] = (int)list.size();
    gp->llsize[p] = (int)ncore;
  }
  return OK;
}

int subgraph_sp2_loop(const DMat& H, DMat rho, prg_graph_partitioning_t* gp) {
  int st = ensure_device();
  if (st) return st;
  const int np = gp->totalParts;
  if (np <= 0 || gp->maxIter < 0 || gp->maxIter > 100) {
    set_error(ERR_ARG, "prg_subgraphSP2Loop: bad partitioning (parts / maxIter)");
    return ERR_ARG;
  }
  if (!(gp->maxeval - gp->mineval != 0.0)) {
    set_error(ERR_NUMERIC, "prg_subgraphSP2Loop: mineval == maxeval");
    return ERR_NUMERIC;
  }
  // plan: parts by padded size (largest first), tiles per part
  std::vector<int> order(np);
  for (int p = 0; p < np; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return gp->lsize[a] > gp->lsize[b]; });
  std::vector<PartDesc> parts(np);
  std::vector<int> idx;
  std::vector<int4> tiles;
  size_t off = 0;
  for (int q = 0; q < np; ++q) {
    const int p = order[q];
    PartDesc d;
    d.n = gp->lsize[p];
    d.core = gp->llsize[p];
    d.m = std::max(64, ((d.n + 63) / 64) * 64);
    d.offset = off;
    d.idx_offset = (int)idx.size();
    off += (size_t)d.m * d.m;
    idx.insert(idx.end(), gp->core_halo_index[p], gp->core_halo_index[p] + d.n);
    parts[q] = d;
    const int nt = d.m / 64;
    for (int J = 0; J < nt; ++J)
      for (int I = 0; I <= J; ++I) tiles.push_back(int4{q, I, J, 0});
  }
  BatchPlan plan;
  plan.nparts = np;
  plan.ntiles = (int)tiles.size();
  plan.total = off;
  plan.d_parts = (PartDesc*)workspace("gp_parts", sizeof(PartDesc) * np, &st);
  plan.d_tiles = (int4*)workspace("gp_tiles", sizeof(int4) * tiles.size(), &st);
  plan.d_idx = (int*)workspace("gp_idx", sizeof(int) * std::max<size_t>(1, idx.size()), &st);
  double* x0 = (double*)workspace("gp_x0", sizeof(double) * off, &st);
  double* x1 = (double*)workspace("gp_x1", sizeof(double) * off, &st);
  double* d_part = (double*)workspace("gp_part", sizeof(double) * tiles.size(), &st);
  double* d_vv = (double*)workspace("gp_vv", sizeof(double) * 100, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(plan.d_parts, parts.data(), sizeof(PartDesc) * np,
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(plan.d_tiles, tiles.data(), sizeof(int4) * tiles.size(),
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(plan.d_idx, idx.data(), sizeof(int) * idx.size(),
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(d_vv, gp->vv, sizeof(double) * 100, hipMemcpyHostToDevice,
                               stream()));
  hipEvent_t ev[4];
  for (auto& e : ev) PRG_HIP_CHECK(hipEventCreate(&e));
  PRG_HIP_CHECK(hipEventRecord(ev[0], stream()));
  if ((st = batched_gather(H, plan, x0, gp->mineval, gp->maxeval))) return st;
  PRG_HIP_CHECK(hipEventRecord(ev[1], stream()));
  double* cur = x0;
  double* nxt = x1;
  for (int it = 0; it < gp->maxIter; ++it) {
    if ((st = batched_step(plan, cur, nxt, gp->pp[it] != 0, d_part, d_vv, it))) return st;
    std::swap(cur, nxt);
  }
  PRG_HIP_CHECK(hipEventRecord(ev[2], stream()));
  if ((st = batched_scatter(plan, cur, rho))) return st;
  PRG_HIP_CHECK(hipEventRecord(ev[3], stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(gp->vv, d_vv, sizeof(double) * 100, hipMemcpyDeviceToHost,
                               stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  for (int k = 0; k < 3; ++k) PRG_HIP_CHECK(hipEventElapsedTime(&g_ms[k], ev[k], ev[k + 1]));
  for (auto& e : ev) PRG_HIP_CHECK(hipEventDestroy(e));
  return OK;
}

}  // namespace prg

using namespace prg;

extern "C" {

// src/prg_graph_mod.F90:359-399
void prg_equalPartition(prg_graph_partitioning_t* gp, int nodesPerPart, int nnodes) {
  clear_error();
  if (!gp || nodesPerPart <= 0 || nnodes <= 0) {
    set_error(ERR_ARG, "prg_equalPartition: bad arguments");
    return;
  }
  const int np = (nnodes <= nodesPerPart) ? 1 : (nnodes + nodesPerPart - 1) / nodesPerPart;
  if (init_parts(gp, np, nnodes)) return;
  for (int i = 1; i <= np; ++i) {
    const int psize = (i * nodesPerPart <= nnodes) ? nodesPerPart : nnodes - nodesPerPart * (i - 1);
    gp->nnodesInPart[i - 1] = psize;
    gp->nodeInPart[i - 1] = (int*)std::malloc(sizeof(int) * std::max(1, psize));
    for (int j = 0; j < psize; ++j) gp->nodeInPart[i - 1][j] = (i - 1) * nodesPerPart + j;
  }
}

// src/prg_graph_mod.F90:880-929 (node ids in the file are 0-based; kept 0-based here)
void prg_filePartition(prg_graph_partitioning_t* gp, const char* partFile) {
  clear_error();
  std::ifstream f(partFile ? partFile : "");
  if (!gp || !f) {
    set_error(ERR_ARG, "prg_filePartition: cannot open the part file");
    return;
  }
  std::string name;
  int total_nodes = 0, total_parts = 0;
  if (!(f >> name >> total_nodes >> total_parts) || total_parts <= 0) {
    set_error(ERR_ARG, "prg_filePartition: bad header");
    return;
  }
  if (init_parts(gp, total_parts, total_nodes)) return;
  for (int i = 0; i < total_parts; ++i) f >> gp->nnodesInPart[i];
  for (int i = 0; i < total_parts; ++i) {
    int pid = 0;
    f >> pid;
    const int sz = gp->nnodesInPart[i];
    gp->nodeInPart[i] = (int*)std::malloc(sizeof(int) * std::max(1, sz));
    for (int j = 0; j < sz; ++j) f >> gp->nodeInPart[i][j];
  }
  if (!f) set_error(ERR_ARG, "prg_filePartition: truncated file");
}

// Partition given as flat arrays (a partitioner's output): sizes[nparts],
// nodes[sum sizes] 0-based.  No reference counterpart (the reference fills
// graph_partitioning_t from METIS or its own partitioners).
void prg_setPartition(prg_graph_partitioning_t* gp, int nparts, int nnodes, const int* sizes,
                      const int* nodes) {
  clear_error();
  if (!gp || nparts <= 0 || !sizes || !nodes) {
    set_error(ERR_ARG, "prg_setPartition: bad arguments");
    return;
  }
  if (init_parts(gp, nparts, nnodes)) return;
  size_t pos = 0;
  for (int i = 0; i < nparts; ++i) {
    gp->nnodesInPart[i] = sizes[i];
    gp->nodeInPart[i] = (int*)std::malloc(sizeof(int) * std::max(1, sizes[i]));
    std::memcpy(gp->nodeInPart[i], nodes + pos, sizeof(int) * sizes[i]);
    pos += sizes[i];
  }
}

void prg_destroyGraphPartitioning(prg_graph_partitioning_t* gp) { free_parts(gp); }

// src/prg_subgraphloop_mod.F90:336-368 (core+halo lists from the graph only)
void prg_getPartitionHalosFromGraph(prg_graph_partitioning_t* gp, bml_matrix_t* g_bml,
                                    int djflag) {
  clear_error();
  if (!gp || !g_bml) {
    set_error(ERR_ARG, "prg_getPartitionHalosFromGraph: null argument");
    return;
  }
  halo_indices(g_bml, gp, djflag);
}

// src/prg_graph_mod.F90:933-971 (one rank): vv <- sqrt(vv) for the first maxIter
void prg_fnormGraph(prg_graph_partitioning_t* gp) {
  clear_error();
  if (!gp) return;
  for (int i = 0; i < gp->maxIter && i < 100; ++i) gp->vv[i] = std::sqrt(gp->vv[i]);
}

// src/prg_subgraphloop_mod.F90:36-126 on one GPU: core+halo extraction with
// the double-jump flag set (as the reference calls it), lock-step SP2 of all
// parts with gp->pp / gp->maxIter and the (mineval, maxeval) normalisation,
// core rows scattered into rho, vv accumulated then square-rooted by
// prg_fnormGraph.  The threshold is not applied (dense storage).
void prg_subgraphSP2Loop(bml_matrix_t* h_bml, bml_matrix_t* g_bml, bml_matrix_t* rho_bml,
                         prg_graph_partitioning_t* gp, double threshold) {
  (void)threshold;
  clear_error();
  if (!gp || check_pair(h_bml, g_bml, "prg_subgraphSP2Loop") ||
      check_pair(h_bml, rho_bml, "prg_subgraphSP2Loop"))
    return;
  if (halo_indices(g_bml, gp, 1)) return;
  DMat H, R;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(rho_bml, &R)) return;
  if (subgraph_sp2_loop(H, R, gp)) return;
  prg_fnormGraph(gp);
}

// Device time (ms) of the last prg_subgraphSP2Loop: gather + normalise, the
// purification loop, the core-row scatter.
void prg_gpu_subgraph_timing(double* ms_gather, double* ms_loop, double* ms_scatter) {
  if (ms_gather) *ms_gather = g_ms[0];
  if (ms_loop) *ms_loop = g_ms[1];
  if (ms_scatter) *ms_scatter = g_ms[2];
}

}  // extern "C"
