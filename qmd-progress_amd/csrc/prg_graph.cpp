// Graph partitioning and the subgraph SP2 loop
// (src/prg_graph_mod.F90, src/prg_subgraphloop_mod.F90).
//
// The partition structure is a flat C struct (include/progress_gpu.h) standing
// in for the reference's Fortran graph_partitioning_t.  Core+halo index lists
// are extracted on the GPU from the graph matrix (prg_halo.hip); all SP2 work
// -- gather and normalisation of every part, the lock-step purification of
// all parts, and the scatter of the core rows into the density -- runs on the
// GPU as batched kernels.
#include "prg_graph.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "prg_bml.h"
#include "prg_eigen.h"
#include "prg_ops.h"
#include "prg_sp2.h"

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

// bml_matrix2submatrix_index (lanl/bml; restated, BML is absent): core nodes
// first (given order), then every node adjacent in G (nonzero, off-diagonal
// entry of the core node's row) in order of discovery; with h_bml -- passed by
// the subgraph SP2 loop, src/prg_subgraphloop_mod.F90:60-63 -- the core
// nodes' neighbours in H next; then with the double-jump flag the graph
// neighbours of those halo nodes.  Graph-only callers:
// src/prg_subgraphloop_mod.F90:355-358, src/prg_graphsolver_mod.F90:112-115.
// The lists are built on the GPU from the device copies (halo_indices_device).
int halo_indices(bml_matrix_t* g_bml, prg_graph_partitioning_t* gp, int djflag,
                 bml_matrix_t* h_bml = nullptr) {
  int st = ensure_device();
  if (st) return st;
  DMat G, H;
  if ((st = bml_dev_read(g_bml, &G))) return st;
  if (h_bml && (st = bml_dev_read(h_bml, &H))) return st;
  return halo_indices_device(G, gp, djflag, h_bml ? &H : nullptr);
}

// Packed batch of all parts: largest padded size first, each part m x m with
// m = ceil64(core + halo).
struct HostPlan {
  std::vector<int> order;  // packed position -> part id
  std::vector<PartDesc> parts;
  std::vector<int> idx;
};

int make_plan(const prg_graph_partitioning_t* gp, HostPlan* hp, BatchPlan* plan) {
  int st = OK;
  const int np = gp->totalParts;
  hp->order.resize(np);
  for (int p = 0; p < np; ++p) hp->order[p] = p;
  std::stable_sort(hp->order.begin(), hp->order.end(),
                   [&](int a, int b) { return gp->lsize[a] > gp->lsize[b]; });
  hp->parts.assign(np, PartDesc{});
  hp->idx.clear();
  size_t off = 0;
  for (int q = 0; q < np; ++q) {
    const int p = hp->order[q];
    PartDesc d;
    d.n = gp->lsize[p];
    d.core = gp->llsize[p];
    d.m = std::max(64, ((d.n + 63) / 64) * 64);
    d.offset = off;
    d.idx_offset = (int)hp->idx.size();
    off += (size_t)d.m * d.m;
    hp->idx.insert(hp->idx.end(), gp->core_halo_index[p], gp->core_halo_index[p] + d.n);
    hp->parts[q] = d;
  }
  plan->nparts = np;
  plan->total = off;
  plan->d_parts = (PartDesc*)workspace("gp_parts", sizeof(PartDesc) * np, &st);
  plan->d_idx = (int*)workspace("gp_idx", sizeof(int) * std::max<size_t>(1, hp->idx.size()), &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(plan->d_parts, hp->parts.data(), sizeof(PartDesc) * np,
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(plan->d_idx, hp->idx.data(), sizeof(int) * hp->idx.size(),
                               hipMemcpyHostToDevice, stream()));
  return OK;
}

// Upper-triangle tiles (part, first row, first column, shape) of every part
// over panels of width `panel` (64 or 128); the last panel of a part is 64
// wide when 128 does not divide its size, so the shapes are 128 x 128
// (code 3), 128 x 64 (1) and 64 x 64 (0).  The parts are dealt to 8 queues,
// one per XCD, largest work first to the least loaded queue (work of a tile:
// rows x columns x m); a queue lists its parts' tiles part by part, column
// by column.  off[q] is the first tile of queue q, off[8] the total.
struct TileQueues {
  std::vector<int4> tiles;
  std::vector<int> off;
};

TileQueues make_tiles(const HostPlan& hp, int panel) {
  const int np = (int)hp.parts.size();
  std::vector<std::vector<int4>> per(np);
  std::vector<double> work(np, 0.0);
  std::vector<int> start, width;
  for (int q = 0; q < np; ++q) {
    const int m = hp.parts[q].m;
    start.clear();
    width.clear();
    for (int s = 0; s < m; s += panel) {
      start.push_back(s);
      width.push_back(std::min(panel, m - s));
    }
    for (size_t b = 0; b < start.size(); ++b)
      for (size_t a = 0; a <= b; ++a) {
        per[q].push_back(int4{q, start[a], start[b], (width[a] == 128 ? 1 : 0) | (width[b] == 128 ? 2 : 0)});
        work[q] += (double)width[a] * width[b] * m;
      }
  }
  std::vector<int> order(np);
  for (int q = 0; q < np; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  std::vector<std::vector<int>> bins(8);
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q : order) {
    const int b = (int)(std::min_element(load, load + 8) - load);
    bins[b].push_back(q);
    load[b] += work[q];
  }
  TileQueues out;
  out.off.push_back(0);
  for (int b = 0; b < 8; ++b) {
    for (int q : bins[b]) out.tiles.insert(out.tiles.end(), per[q].begin(), per[q].end());
    out.off.push_back((int)out.tiles.size());
  }
  return out;
}

int upload_tiles(const TileQueues& tq, int panel, BatchPlan* plan) {
  int st = OK;
  plan->wide = panel == 128 ? 1 : 0;
  plan->ntiles = (int)tq.tiles.size();
  plan->d_tiles = (int4*)workspace(panel == 128 ? "gp_tiles128" : "gp_tiles64",
                                   sizeof(int4) * std::max<size_t>(1, tq.tiles.size()), &st);
  plan->d_queue_off =
      (int*)workspace(panel == 128 ? "gp_qoff128" : "gp_qoff64", sizeof(int) * 9, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(plan->d_tiles, tq.tiles.data(), sizeof(int4) * tq.tiles.size(),
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(plan->d_queue_off, tq.off.data(), sizeof(int) * 9,
                               hipMemcpyHostToDevice, stream()));
  return OK;
}

// Kernel configuration per batch: the tile panel width (64 or 128), the
// number of K slabs prefetched ahead (1 or 2) and the slab depth (16 or 32);
// the compiled combinations are the candidates below.  PRG_BATCH_TILE,
// PRG_BATCH_DEPTH and PRG_BATCH_KSLAB restrict them; the remaining ones are
// timed on the batch itself (one untimed and three timed iterations each on
// the gathered first iterate, results discarded) and the fastest is kept for
// every later call with the same padded part sizes and restrictions.
struct BatchConfig {
  int panel = 0, depth = 0, kslab = 0;
};
constexpr BatchConfig kCandidates[] = {{64, 1, 16}, {64, 2, 16}, {64, 1, 32}, {128, 1, 16}};
// split-precision products: 64 x 64 tiles (64-deep slabs), 128-wide panels
// (32-deep slabs through registers), or 128-wide panels with the slabs copied
// into a 4-slot LDS ring (depth 4)
constexpr BatchConfig kSplitCandidates[] = {{64, 1, 64}, {128, 1, 32}, {128, 4, 32}};

std::map<std::pair<uint64_t, int>, BatchConfig>& config_cache() {
  static std::map<std::pair<uint64_t, int>, BatchConfig> c;
  return c;
}

uint64_t size_signature(const HostPlan& hp) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the padded sizes
  for (const PartDesc& d : hp.parts) {
    h ^= (uint64_t)d.m;
    h *= 1099511628211ull;
  }
  return h;
}

// Device time of one batched iteration (min of three, after one untimed).
int time_step(const BatchPlan& plan, const double* x, double* xn, double* part, int* tickets,
              float* ms) {
  hipEvent_t a, b;
  PRG_HIP_CHECK(hipEventCreate(&a));
  PRG_HIP_CHECK(hipEventCreate(&b));
  int st = OK;
  for (int r = 0; r < 4 && !st; ++r) {
    if ((st = batched_zero_tickets(tickets, 8))) break;
    PRG_HIP_CHECK(hipEventRecord(a, stream()));
    if ((st = batched_step(plan, x, xn, 1, part, tickets))) break;
    PRG_HIP_CHECK(hipEventRecord(b, stream()));
    PRG_HIP_CHECK(hipEventSynchronize(b));
    float t = 0.f;
    PRG_HIP_CHECK(hipEventElapsedTime(&t, a, b));
    if (r == 1 || (r > 1 && t < *ms)) *ms = t;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return st;
}

// The captured iteration launches of the last subgraph loop, replayed while
// the buffers, the tile list and the branch sequence stay the same.
struct LoopGraph {
  std::vector<uint64_t> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
};
LoopGraph& loop_graph() {
  static LoopGraph* g = new LoopGraph;  // never destroyed: outlives the HIP runtime at exit
  return *g;
}

// panel width, loop launch (0 eager, 1 captured, 2 replayed), prefetch depth, slab depth
int g_cfg[4] = {0, 0, 0, 0};

// iters SP2 iterations of every part from x0, ping-ponging with x1 (the
// result is in x0 for even iters), then vv[it] += the iteration's core-row
// idempotency sums.  d_part holds iters x ntiles partials.
int enqueue_loop(const BatchPlan& plan, double* x0, double* x1, const int* pp, int iters,
                 double* d_part, int* d_tickets, double* d_vv) {
  int st = batched_zero_tickets(d_tickets, 8 * iters);
  const double* cur = x0;
  double* nxt = x1;
  for (int it = 0; it < iters && !st; ++it) {
    st = batched_step(plan, cur, nxt, pp[it] != 0, d_part + (size_t)it * plan.ntiles,
                      d_tickets + 8 * it, it & 1);
    const double* t = cur;
    cur = nxt;
    nxt = const_cast<double*>(t);
  }
  if (!st) st = batched_vv(plan, d_part, iters, d_vv);
  return st;
}

int run_loop(const BatchPlan& plan, double* x0, double* x1, const int* pp, int iters,
             double* d_part, int* d_tickets, double* d_vv) {
  const char* e = std::getenv("PRG_BATCH_GRAPH");
  if (e && std::atoi(e) == 0) {
    g_cfg[1] = 0;
    return enqueue_loop(plan, x0, x1, pp, iters, d_part, d_tickets, d_vv);
  }
  std::vector<uint64_t> key = {(uint64_t)(uintptr_t)x0,           (uint64_t)(uintptr_t)x1,
                               (uint64_t)(uintptr_t)d_part,       (uint64_t)(uintptr_t)d_vv,
                               (uint64_t)(uintptr_t)d_tickets,    (uint64_t)(uintptr_t)plan.d_parts,
                               (uint64_t)(uintptr_t)plan.d_tiles, (uint64_t)(uintptr_t)plan.d_queue_off,
                               (uint64_t)plan.ntiles,             (uint64_t)plan.wide,
                               (uint64_t)plan.depth,              (uint64_t)plan.grid,
                               (uint64_t)plan.kslab,
                               (uint64_t)plan.split,
                               (uint64_t)(uintptr_t)plan.planes[0],
                               (uint64_t)(uintptr_t)plan.planes[2],
                               (uint64_t)iters,
                               (uint64_t)workspace_generation()};
  for (int it = 0; it < iters; ++it) key.push_back(pp[it] != 0);
  LoopGraph& lg = loop_graph();
  if (lg.exec && lg.key == key) {
    g_cfg[1] = 2;
    PRG_HIP_CHECK(hipGraphLaunch(lg.exec, stream()));
    return OK;
  }
  lg.reset();
  PRG_HIP_CHECK(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
  int st = enqueue_loop(plan, x0, x1, pp, iters, d_part, d_tickets, d_vv);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(stream(), &graph);
  if (st || ec != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    if (!st) set_error(ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    return st ? st : ERR_HIP;
  }
  lg.graph = graph;
  const hipError_t ei = hipGraphInstantiate(&lg.exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    lg.exec = nullptr;
    lg.reset();
    set_error(ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    return ERR_HIP;
  }
  lg.key = key;
  g_cfg[1] = 1;
  PRG_HIP_CHECK(hipGraphLaunch(lg.exec, stream()));
  return OK;
}

int run_batch(const DMat& H, DMat rho, prg_graph_partitioning_t* gp, const BatchPlan& plan,
              double* x0, double* x1, double* d_part, int* d_tickets, double* d_vv);

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
  HostPlan hp;
  BatchPlan plan;
  if ((st = make_plan(gp, &hp, &plan))) return st;
  const size_t off = plan.total;
  const int iters = gp->maxIter;
  const TileQueues tiles64 = make_tiles(hp, 64);
  const size_t maxtiles = tiles64.tiles.size();  // the 64-wide tiling has the most tiles
  double* x0 = (double*)workspace("gp_x0", sizeof(double) * off, &st);
  double* x1 = (double*)workspace("gp_x1", sizeof(double) * off, &st);
  // iters x ntiles loop partials and iters x 8 queue tickets, + one slice each
  // for the configuration timing
  double* d_part = (double*)workspace("gp_part", sizeof(double) * maxtiles * (iters + 1), &st);
  int* d_tickets = (int*)workspace("gp_tickets", sizeof(int) * 8 * (iters + 1), &st);
  double* d_vv = (double*)workspace("gp_vv", sizeof(double) * 100, &st);
  if (st) return st;
  // split-precision products (prg_set_precision FP16 / BF16 split): the
  // split kernels, X also carried as hi / lo planes; int8 CRT has no
  // batched kernel, so its parts run with FP64 products
  const int pm = precision_mode();
  plan.split = pm == PRG_PREC_F16SPLIT ? 1 : (pm == PRG_PREC_BF16SPLIT ? 2 : 0);
  BatchConfig force;
  if (const char* e = std::getenv("PRG_BATCH_TILE")) force.panel = std::atoi(e);
  if (plan.split) {
    // two split kernels: 64 x 64 tiles with 64-deep slabs, or tiles of
    // 128-wide panels with 32-deep slabs; timed per batch like the FP64
    // candidates (PRG_BATCH_TILE restricts the panel width)
    char* pl = (char*)workspace("gp_planes", 4 * sizeof(uint16_t) * off, &st);
    if (st) return st;
    for (int k = 0; k < 4; ++k) plan.planes[k] = pl + k * sizeof(uint16_t) * off;
    std::vector<BatchConfig> scand;
    if (const char* e = std::getenv("PRG_BATCH_DEPTH")) force.depth = std::atoi(e);
    for (const BatchConfig& c : kSplitCandidates)
      if ((!force.panel || c.panel == force.panel) && (!force.depth || c.depth == force.depth))
        scand.push_back(c);
    if (scand.empty()) {
      set_error(ERR_ARG, "PRG_BATCH_TILE / PRG_BATCH_DEPTH: no compiled split-precision "
                         "batched SP2 kernel has this configuration");
      return ERR_ARG;
    }
    const auto skey = std::make_pair(size_signature(hp),
                                     -((plan.split * 256 + force.panel) * 8 + force.depth));
    BatchConfig cfg = scand[0];
    auto hit = config_cache().find(skey);
    if (hit != config_cache().end()) {
      cfg = hit->second;
    } else if (scand.size() > 1) {
      if ((st = batched_gather(H, plan, x0, gp->mineval, gp->maxeval)) ||
          (st = batched_split_prepare(plan, x0)))
        return st;
      const TileQueues tiles128 = make_tiles(hp, 128);
      float best = 0.f;
      for (size_t k = 0; k < scand.size(); ++k) {
        const BatchConfig& c = scand[k];
        if ((st = upload_tiles(c.panel == 128 ? tiles128 : tiles64, c.panel, &plan))) return st;
        plan.depth = c.depth;
        plan.kslab = c.kslab;
        float ms = 0.f;
        if ((st = batched_slots(&plan)) ||
            (st = time_step(plan, x0, x1, d_part + (size_t)iters * maxtiles,
                            d_tickets + 8 * iters, &ms)))
          return st;
        if (k == 0 || ms < best) {
          best = ms;
          cfg = c;
        }
      }
    }
    config_cache()[skey] = cfg;
    if ((st = upload_tiles(cfg.panel == 128 ? make_tiles(hp, 128) : tiles64, cfg.panel, &plan)))
      return st;
    plan.depth = cfg.depth;
    plan.kslab = cfg.kslab;
    if ((st = batched_slots(&plan))) return st;
    g_cfg[0] = cfg.panel;
    g_cfg[2] = cfg.depth;
    g_cfg[3] = cfg.kslab;
    return run_batch(H, rho, gp, plan, x0, x1, d_part, d_tickets, d_vv);
  }
  if (const char* e = std::getenv("PRG_BATCH_DEPTH")) force.depth = std::atoi(e);
  if (const char* e = std::getenv("PRG_BATCH_KSLAB")) force.kslab = std::atoi(e);
  std::vector<BatchConfig> cand;
  for (const BatchConfig& c : kCandidates)
    if ((!force.panel || c.panel == force.panel) && (!force.depth || c.depth == force.depth) &&
        (!force.kslab || c.kslab == force.kslab))
      cand.push_back(c);
  if (cand.empty()) {
    set_error(ERR_ARG, "PRG_BATCH_TILE / PRG_BATCH_DEPTH / PRG_BATCH_KSLAB: no compiled batched "
                       "SP2 kernel has this configuration");
    return ERR_ARG;
  }
  const auto key =
      std::make_pair(size_signature(hp), (force.panel * 4 + force.depth) * 64 + force.kslab);
  BatchConfig cfg = cand[0];
  auto hit = config_cache().find(key);
  if (hit != config_cache().end()) {
    cfg = hit->second;
  } else if (cand.size() > 1) {
    if ((st = batched_gather(H, plan, x0, gp->mineval, gp->maxeval))) return st;
    const TileQueues tiles128 = make_tiles(hp, 128);
    float best = 0.f;
    for (size_t k = 0; k < cand.size(); ++k) {
      const BatchConfig& c = cand[k];
      if ((st = upload_tiles(c.panel == 128 ? tiles128 : tiles64, c.panel, &plan))) return st;
      plan.depth = c.depth;
      plan.kslab = c.kslab;
      float ms = 0.f;
      if ((st = batched_slots(&plan)) ||
          (st = time_step(plan, x0, x1, d_part + (size_t)iters * maxtiles, d_tickets + 8 * iters,
                          &ms)))
        return st;
      if (k == 0 || ms < best) {
        best = ms;
        cfg = c;
      }
    }
  }
  config_cache()[key] = cfg;
  if ((st = upload_tiles(cfg.panel == 128 ? make_tiles(hp, 128) : tiles64, cfg.panel, &plan)))
    return st;
  plan.depth = cfg.depth;
  plan.kslab = cfg.kslab;
  if ((st = batched_slots(&plan))) return st;
  g_cfg[0] = cfg.panel;
  g_cfg[2] = cfg.depth;
  g_cfg[3] = cfg.kslab;
  return run_batch(H, rho, gp, plan, x0, x1, d_part, d_tickets, d_vv);
}

// Gather + normalise, the iteration loop, the core-row scatter and the vv
// read-back of one configured batch, with HIP-event phase timing.
int run_batch(const DMat& H, DMat rho, prg_graph_partitioning_t* gp, const BatchPlan& plan,
              double* x0, double* x1, double* d_part, int* d_tickets, double* d_vv) {
  int st = OK;
  const int iters = gp->maxIter;
  PRG_HIP_CHECK(hipMemcpyAsync(d_vv, gp->vv, sizeof(double) * 100, hipMemcpyHostToDevice,
                               stream()));
  hipEvent_t ev[4];
  for (auto& e : ev) PRG_HIP_CHECK(hipEventCreate(&e));
  PRG_HIP_CHECK(hipEventRecord(ev[0], stream()));
  if ((st = batched_gather(H, plan, x0, gp->mineval, gp->maxeval))) return st;
  if (plan.split && (st = batched_split_prepare(plan, x0))) return st;
  PRG_HIP_CHECK(hipEventRecord(ev[1], stream()));
  if ((st = run_loop(plan, x0, x1, gp->pp, iters, d_part, d_tickets, d_vv))) return st;
  PRG_HIP_CHECK(hipEventRecord(ev[2], stream()));
  if ((st = batched_scatter(plan, iters % 2 ? x1 : x0, rho, 2.0))) return st;
  PRG_HIP_CHECK(hipEventRecord(ev[3], stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(gp->vv, d_vv, sizeof(double) * 100, hipMemcpyDeviceToHost,
                               stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  for (int k = 0; k < 3; ++k) PRG_HIP_CHECK(hipEventElapsedTime(&g_ms[k], ev[k], ev[k + 1]));
  for (auto& e : ev) PRG_HIP_CHECK(hipEventDestroy(e));
  return OK;
}

// ----------------------------------------------------------------------------
// Graph partitioner standing in for METIS (prg_metisPartition,
// src/prg_partition_mod.F90:216-311, which the reference links for
// prg_build_densityGP_T0): recursive bisection of breadth-first orderings.
// Each bisection orders the nodes of the current subset by BFS from a
// pseudo-peripheral node (the last node reached by a BFS from the lowest
// index; neighbours in ascending order; components in ascending order of
// their lowest node) and cuts the ordering so that the two sides receive
// node counts proportional to their numbers of parts.  Geometric graphs get
// compact, slab-like parts; the result is deterministic.
namespace {

void bfs_from(int start, const std::vector<std::vector<int>>& adj, const std::vector<char>& in,
              std::vector<char>& mark, std::vector<int>& out) {
  size_t head = out.size();
  out.push_back(start);
  mark[start] = 1;
  while (head < out.size()) {
    const int v = out[head++];
    for (int nb : adj[v])
      if (in[nb] && !mark[nb]) {
        mark[nb] = 1;
        out.push_back(nb);
      }
  }
}

void bisect(std::vector<int> nodes, int k, int first, const std::vector<std::vector<int>>& adj,
            std::vector<char>& in, std::vector<char>& mark, std::vector<int>& part) {
  if (k == 1) {
    for (int v : nodes) part[v] = first;
    return;
  }
  for (int v : nodes) in[v] = 1;
  std::vector<int> order, probe;
  order.reserve(nodes.size());
  for (int v : nodes) {
    if (mark[v]) continue;
    probe.clear();
    bfs_from(v, adj, in, mark, probe);  // the component of v
    for (int u : probe) mark[u] = 0;
    bfs_from(probe.back(), adj, in, mark, order);
  }
  for (int v : nodes) {
    in[v] = 0;
    mark[v] = 0;
  }
  const int kl = k / 2;
  const size_t nl = (size_t)((double)nodes.size() * kl / k + 0.5);
  std::vector<int> left(order.begin(), order.begin() + nl);
  std::vector<int> right(order.begin() + nl, order.end());
  std::sort(left.begin(), left.end());
  std::sort(right.begin(), right.end());
  bisect(std::move(left), kl, first, adj, in, mark, part);
  bisect(std::move(right), k - kl, first + kl, adj, in, mark, part);
}

// Row adjacency (nonzero off-diagonal entries, ascending) of the graph matrix.
int graph_adjacency(bml_matrix_t* g_bml, std::vector<std::vector<int>>* adj) {
  int st;
  if ((st = bml_host_sync(g_bml))) return st;
  const int N = g_bml->N;
  const double* G = g_bml->host;
  adj->assign(N, {});
  // one contiguous pass over the columns: G(i, j) != 0 makes i and j
  // neighbours both ways; each list is then sorted and deduplicated
  for (int j = 0; j < N; ++j) {
    const double* col = G + (size_t)j * N;
    for (int i = 0; i < N; ++i)
      if (i != j && col[i] != 0.0) {
        (*adj)[i].push_back(j);
        (*adj)[j].push_back(i);
      }
  }
  for (auto& l : *adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  return OK;
}

double g_gp_ef = 0.0;

}  // namespace

int bfs_partition(bml_matrix_t* g_bml, int nparts, prg_graph_partitioning_t* gp) {
  std::vector<std::vector<int>> adj;
  int st;
  if ((st = graph_adjacency(g_bml, &adj))) return st;
  const int N = (int)adj.size();
  if (nparts <= 0 || nparts > N) {
    set_error(ERR_ARG, "graph partitioner: nparts must be in [1, N]");
    return ERR_ARG;
  }
  std::vector<int> nodes(N), part(N, -1);
  for (int i = 0; i < N; ++i) nodes[i] = i;
  std::vector<char> in(N, 0), mark(N, 0);
  bisect(std::move(nodes), nparts, 0, adj, in, mark, part);
  if ((st = init_parts(gp, nparts, N))) return st;
  for (int v = 0; v < N; ++v) gp->nnodesInPart[part[v]]++;
  std::vector<int> fill(nparts, 0);
  for (int p = 0; p < nparts; ++p)
    gp->nodeInPart[p] = (int*)std::malloc(sizeof(int) * std::max(1, gp->nnodesInPart[p]));
  for (int v = 0; v < N; ++v) gp->nodeInPart[part[v]][fill[part[v]]++] = v;
  return OK;
}

// prg_build_densityGP_T0 (src/prg_graphsolver_mod.F90:39-191) on one GPU:
// partition the graph (bfs_partition in place of METIS), core+halo lists
// with the double jump, every part's Fermi density at kT = 0.1 from batched
// eigensolves (prg_build_density_T_Fermi, src/prg_densitymatrix_mod.F90:509-582,
// called with bndfil in the threshold slot as the reference does), core rows
// scattered into rho, then the reference's chemical-potential update from the
// trace error and the summed dN/dEf of the parts (:162-177).
int graph_density_t0(const DMat& H, bml_matrix_t* g_bml, DMat rho, double bndfil, double* ef,
                     int nparts, int verbose) {
  int st = ensure_device();
  if (st) return st;
  prg_graph_partitioning_t gp;
  std::memset(&gp, 0, sizeof(gp));
  struct Guard {
    prg_graph_partitioning_t* g;
    ~Guard() { free_parts(g); }
  } guard{&gp};
  if ((st = bfs_partition(g_bml, nparts, &gp))) return st;
  if ((st = halo_indices(g_bml, &gp, 1))) return st;
  HostPlan hp;
  BatchPlan plan;
  if ((st = make_plan(&gp, &hp, &plan))) return st;
  const int np = plan.nparts;
  // eigenvalue offsets: m per part, packed in plan order
  std::vector<size_t> woff(np);
  size_t wtot = 0;
  for (int q = 0; q < np; ++q) {
    woff[q] = wtot;
    wtot += hp.parts[q].m;
  }
  double* a = (double*)workspace("gpt0_a", sizeof(double) * plan.total, &st);
  double* u = (double*)workspace("gpt0_u", sizeof(double) * plan.total, &st);
  double* w = (double*)workspace("gpt0_w", sizeof(double) * wtot, &st);
  size_t* d_woff = (size_t*)workspace("gpt0_woff", sizeof(size_t) * np, &st);
  double* d_dn = (double*)workspace("gpt0_dn", sizeof(double) * np, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(d_woff, woff.data(), sizeof(size_t) * np, hipMemcpyHostToDevice,
                               stream()));
  // padding eigenvalues far above the spectrum: Fermi occupation 0
  double hmax = 0.0;
  {
    double gb[2];
    double* d_b = (double*)workspace("gpt0_bounds", sizeof(double) * 4, &st);
    if (st || (st = gershgorin(H, d_b))) return st;
    PRG_HIP_CHECK(hipMemcpyAsync(gb, d_b, sizeof(gb), hipMemcpyDeviceToHost, stream()));
    PRG_HIP_CHECK(hipStreamSynchronize(stream()));
    if (!std::isfinite(gb[0]) || !std::isfinite(gb[1])) {
      set_error(ERR_NUMERIC, "prg_build_densityGP_T0: non-finite Hamiltonian");
      return ERR_NUMERIC;
    }
    hmax = std::max(std::fabs(gb[0]), std::fabs(gb[1]));
  }
  const double kbt = 0.1;
  const double pad = std::fabs(*ef) + hmax + 1.0e3 * kbt + 1.0;
  if ((st = batched_gather_pad(H, plan, a, pad))) return st;
  // one strided-batched eigensolve per run of equal padded sizes
  for (int q0 = 0; q0 < np;) {
    int q1 = q0;
    while (q1 < np && hp.parts[q1].m == hp.parts[q0].m) ++q1;
    const int m = hp.parts[q0].m;
    if ((st = eig_strided_batched(a + hp.parts[q0].offset, m, q1 - q0, w + woff[q0]))) return st;
    q0 = q1;
  }
  if ((st = batched_fermi_scale(plan, a, w, d_woff, u, *ef, kbt, d_dn))) return st;
  // part densities U V^T
  double* dens = (double*)workspace("gpt0_d", sizeof(double) * plan.total, &st);
  if (st) return st;
  for (int q0 = 0; q0 < np;) {
    int q1 = q0;
    while (q1 < np && hp.parts[q1].m == hp.parts[q0].m) ++q1;
    const size_t o = hp.parts[q0].offset;
    if ((st = gemm_nt_strided_batched(u + o, a + o, dens + o, hp.parts[q0].m, q1 - q0))) return st;
    q0 = q1;
  }
  if ((st = batched_scatter(plan, dens, rho, 1.0))) return st;
  std::vector<double> dn(np);
  PRG_HIP_CHECK(hipMemcpyAsync(dn.data(), d_dn, sizeof(double) * np, hipMemcpyDeviceToHost,
                               stream()));
  ops::Sums sm;
  if ((st = ops::sums_of(rho, &sm))) return st;
  double dnocc = 0.0;
  for (int q = 0; q < np; ++q) dnocc += dn[q];  // parts in packed order
  const double nocc = sm.tr;
  const double target = 2.0 * bndfil * rho.n;
  const double error = std::fabs(nocc - target);
  if (nocc < target) dnocc = -dnocc;
  if (verbose > 1)
    std::printf("prg_build_densityGP_T0: Ef %.10g Nocc %.10g error %.3e gradient %.6g\n", *ef,
                nocc, error, dnocc);
  if (error > 1.0e-6 && std::fabs(dnocc) > 1.0e-6) {
    // the reference steps by nocc (not the error) when the error exceeds 1
    if (error > 1.0)
      *ef = *ef - nocc / dnocc;
    else
      *ef = *ef - error / dnocc;
  }
  return OK;
}

// prg_build_zmatGP (src/prg_graphsolver_mod.F90:202-331) on one GPU: the
// partition (own BFS partitioner in place of METIS), the double-jump core+halo
// lists of the graph (no Hamiltonian, as there), per part the Loewdin factor
// Z_p = U s^-1/2 U^T of the overlap's core+halo block (prg_buildZdiag,
// src/prg_genz_mod.F90:176-276) from batched eigensolves, and the core rows of
// each Z_p written into zmat (bml_submatrix2matrix).  Entries outside the
// core rows' core+halo columns keep their value, as the reference leaves them.
int graph_zmat(const DMat& S, bml_matrix_t* g_bml, DMat Z, int nparts) {
  int st = ensure_device();
  if (st) return st;
  prg_graph_partitioning_t gp;
  std::memset(&gp, 0, sizeof(gp));
  struct Guard {
    prg_graph_partitioning_t* g;
    ~Guard() { free_parts(g); }
  } guard{&gp};
  if ((st = bfs_partition(g_bml, nparts, &gp))) return st;
  if ((st = halo_indices(g_bml, &gp, 1))) return st;
  HostPlan hp;
  BatchPlan plan;
  if ((st = make_plan(&gp, &hp, &plan))) return st;
  const int np = plan.nparts;
  std::vector<size_t> woff(np);
  size_t wtot = 0;
  for (int q = 0; q < np; ++q) {
    woff[q] = wtot;
    wtot += hp.parts[q].m;
  }
  double* a = (double*)workspace("gpt0_a", sizeof(double) * plan.total, &st);
  double* u = (double*)workspace("gpt0_u", sizeof(double) * plan.total, &st);
  double* w = (double*)workspace("gpt0_w", sizeof(double) * wtot, &st);
  size_t* d_woff = (size_t*)workspace("gpt0_woff", sizeof(size_t) * np, &st);
  double* d_wmin = (double*)workspace("gpt0_dn", sizeof(double) * np, &st);
  double* zp = (double*)workspace("gpt0_d", sizeof(double) * plan.total, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(d_woff, woff.data(), sizeof(size_t) * np, hipMemcpyHostToDevice,
                               stream()));
  // padding rows decoupled with eigenvalue 1 (positive, never scattered)
  if ((st = batched_gather_pad(S, plan, a, 1.0))) return st;
  for (int q0 = 0; q0 < np;) {
    int q1 = q0;
    while (q1 < np && hp.parts[q1].m == hp.parts[q0].m) ++q1;
    if ((st = eig_strided_batched(a + hp.parts[q0].offset, hp.parts[q0].m, q1 - q0,
                                  w + woff[q0])))
      return st;
    q0 = q1;
  }
  if ((st = batched_invsqrt_scale(plan, a, w, d_woff, u, d_wmin))) return st;
  for (int q0 = 0; q0 < np;) {
    int q1 = q0;
    while (q1 < np && hp.parts[q1].m == hp.parts[q0].m) ++q1;
    const size_t o = hp.parts[q0].offset;
    if ((st = gemm_nt_strided_batched(u + o, a + o, zp + o, hp.parts[q0].m, q1 - q0))) return st;
    q0 = q1;
  }
  std::vector<double> wmin(np);
  PRG_HIP_CHECK(hipMemcpyAsync(wmin.data(), d_wmin, sizeof(double) * np, hipMemcpyDeviceToHost,
                               stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  for (int q = 0; q < np; ++q) {
    if (!(wmin[q] > 0.0)) {
      // the reference stops: "Overlap matrix is not possitive definite"
      set_error(ERR_NUMERIC, "prg_build_zmatGP: an overlap block is not positive definite");
      return ERR_NUMERIC;
    }
  }
  return batched_scatter(plan, zp, Z, 1.0, false);
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

// src/prg_subgraphloop_mod.F90:36-126 on one GPU: core+halo extraction from
// G and H with the double-jump flag set (as the reference calls it), lock-step SP2 of all
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
  if (halo_indices(g_bml, gp, 1, h_bml)) return;
  DMat H, R;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(rho_bml, &R)) return;
  if (subgraph_sp2_loop(H, R, gp)) return;
  prg_fnormGraph(gp);
}

// The build's partitioner (recursive BFS bisection) in place of METIS; no
// reference counterpart as a C symbol (prg_metisPartition,
// src/prg_partition_mod.F90:216-311, is what the reference calls).
void prg_bfsPartition(prg_graph_partitioning_t* gp, bml_matrix_t* g_bml, int nparts) {
  clear_error();
  if (!gp || !g_bml) {
    set_error(ERR_ARG, "prg_bfsPartition: null argument");
    return;
  }
  bfs_partition(g_bml, nparts, gp);
}

// src/prg_progress_mod.h:803-811 (Fortran src/prg_graphsolver_mod.F90:39-191).
// Ef by value as the reference header declares it; the updated value is
// returned by prg_build_densityGP_T0_ef().
void prg_build_densityGP_T0(bml_matrix_t* ham_bml, bml_matrix_t* g_bml, bml_matrix_t* rho_bml,
                            double threshold, double bndfil, double Ef, int nparts, int verbose) {
  (void)threshold;
  clear_error();
  g_gp_ef = Ef;
  if (check_pair(ham_bml, g_bml, "prg_build_densityGP_T0") ||
      check_pair(ham_bml, rho_bml, "prg_build_densityGP_T0"))
    return;
  DMat H, R;
  if (bml_dev_read(ham_bml, &H) || bml_dev_write(rho_bml, &R)) return;
  double ef = Ef;
  if (graph_density_t0(H, g_bml, R, bndfil, &ef, nparts, verbose)) return;
  g_gp_ef = ef;
}

double prg_build_densityGP_T0_ef(void) { return g_gp_ef; }

// src/prg_progress_mod.h:814-820 (Fortran src/prg_graphsolver_mod.F90:202-331).
// threshold (sparse-algebra threshold of the reference) is ignored on dense
// storage; verbose > 1 prints the part count.
void prg_build_zmatGP(bml_matrix_t* over_bml, bml_matrix_t* g_bml, bml_matrix_t* zmat_bml,
                      double threshold, int nparts, int verbose) {
  (void)threshold;
  clear_error();
  if (check_pair(over_bml, g_bml, "prg_build_zmatGP") ||
      check_pair(over_bml, zmat_bml, "prg_build_zmatGP"))
    return;
  DMat S, Z;
  // zmat's current contents first: the entries the parts do not write keep them
  if (bml_dev_read(over_bml, &S) || bml_dev_read(zmat_bml, &Z) || bml_dev_write(zmat_bml, &Z))
    return;
  if (graph_zmat(S, g_bml, Z, nparts)) return;
  if (verbose > 1) std::printf("prg_build_zmatGP: totalParts = %d\n", nparts);
}

// Device time (ms) of the last prg_subgraphSP2Loop: gather + normalise, the
// purification loop, the core-row scatter.
void prg_gpu_subgraph_timing(double* ms_gather, double* ms_loop, double* ms_scatter) {
  if (ms_gather) *ms_gather = g_ms[0];
  if (ms_loop) *ms_loop = g_ms[1];
  if (ms_scatter) *ms_scatter = g_ms[2];
}

// Kernel configuration of the last prg_subgraphSP2Loop: tile panel width
// (64 or 128), prefetch depth (K slabs), K slab depth, and how its iterations
// were launched (0 eager, 1 captured into a hipGraph, 2 replayed).
void prg_gpu_subgraph_config(int* panel, int* depth, int* kslab, int* graph) {
  if (panel) *panel = g_cfg[0];
  if (depth) *depth = g_cfg[2];
  if (kslab) *kslab = g_cfg[3];
  if (graph) *graph = g_cfg[1];
}

}  // extern "C"
