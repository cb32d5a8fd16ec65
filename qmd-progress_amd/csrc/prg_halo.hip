// Core+halo index lists of a graph partitioning on the GPU
// (bml_matrix2submatrix_index as called at src/prg_subgraphloop_mod.F90:60-63
// and :336-368).
//
// The graph (dense, column-major, in HBM) is first packed into an N x W
// bit matrix, W = ceil(N / 32): bit j of row i is set when G(i, j) != 0 and
// i != j.  One workgroup per part then walks the reference's discovery order
// with a W-word "seen" bitmap in LDS: the core nodes in their given order
// (duplicates dropped), then for every core node in list order its unseen
// neighbours in ascending index order, then -- when the Hamiltonian is given,
// as prg_subgraphSP2Loop passes it -- the core nodes' unseen neighbours in H,
// then (double jump) the unseen graph neighbours of every halo node so far.  Each step is one pass over the node's W words, a
// workgroup prefix sum of the new bits, and an ordered write, so the lists
// are exactly those of the sequential loop.
#include "prg_graph.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace prg {
namespace {

constexpr int NT = 256;
constexpr int kMaxWords = 16000;  // LDS bitmap (64 KB with the statics): N <= 512000

// bits[i * W + w], bit b <-> column 32 w + b of row i.  Block (64 rows, 64
// words): each thread packs 16 words of one row (coalesced column reads),
// the block is transposed through LDS so that rows are written contiguously.
__global__ void graph_bits_kernel(const double* __restrict__ g, int ld, int n, int W,
                                  uint32_t* __restrict__ bits) {
  __shared__ uint32_t t[64][65];
  const int i0 = blockIdx.x * 64, w0 = blockIdx.y * 64;
  const int il = threadIdx.x & 63, wg = threadIdx.x >> 6;
  const int i = i0 + il;
  for (int j = 0; j < 16; ++j) {
    const int wl = wg * 16 + j, w = w0 + wl;
    uint32_t word = 0;
    if (i < n && w < W) {
      for (int b = 0; b < 32; ++b) {
        const int col = 32 * w + b;
        if (col < n && col != i && g[i + (size_t)col * ld] != 0.0) word |= 1u << b;
      }
    }
    t[il][wl] = word;
  }
  __syncthreads();
  for (int r = wg; r < 64; r += 4) {
    const int row = i0 + r, w = w0 + il;
    if (row < n && w < W) bits[(size_t)row * W + w] = t[r][il];
  }
}

// Exclusive prefix sum over the workgroup; returns this thread's offset and
// writes the total to *total (all threads).
__device__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) {
    before += (k < w) ? wsum[k] : 0;
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// One part per workgroup.  out: the part's list (capacity cap); cnt[2p] =
// list length, cnt[2p + 1] = core size after dropping repeated core nodes.
__global__ __launch_bounds__(NT) void halo_lists_kernel(const uint32_t* __restrict__ bits,
                                                        const uint32_t* __restrict__ hbits, int W,
                                                        const int* __restrict__ core,
                                                        const int* __restrict__ core_off,
                                                        int djflag, int cap,
                                                        int* __restrict__ lists,
                                                        int* __restrict__ cnt) {
  extern __shared__ uint32_t seen[];
  __shared__ int s_n;
  const int p = blockIdx.x;
  const int* cp = core + core_off[p];
  const int nc = core_off[p + 1] - core_off[p];
  int* out = lists + (size_t)p * cap;
  // contiguous word range of this thread (ascending node order across threads)
  const int per = (W + NT - 1) / NT;
  const int w0 = min(W, (int)threadIdx.x * per), w1 = min(W, w0 + per);
  for (int w = w0; w < w1; ++w) seen[w] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int k = 0; k < nc; ++k) {
      const int c = cp[k];
      const uint32_t bit = 1u << (c & 31);
      if (!(seen[c >> 5] & bit)) {
        seen[c >> 5] |= bit;
        out[m++] = c;
      }
    }
    s_n = m;
  }
  __syncthreads();
  const int ncore = s_n;
  int n = ncore;
  // expand: for each node of list positions [from, to) in order, append its
  // unseen neighbours (bit matrix m) in ascending order
  auto expand = [&](const uint32_t* m, int from, int to) {
    for (int k = from; k < to; ++k) {
      const uint32_t* row = m + (size_t)out[k] * W;
      uint32_t fresh[8];  // per <= 8 for W <= 2048; larger W loops below
      int c = 0;
      if (per <= 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int w = w0 + j;
          uint32_t f = 0u;
          if (w < w1) {
            f = row[w] & ~seen[w];
            seen[w] |= f;
          }
          fresh[j] = f;
          c += __popc(f);
        }
        int total;
        int pos = n + block_exclusive_scan(c, &total);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          uint32_t f = fresh[j];
          while (f) {
            const int b = __ffs(f) - 1;
            f &= f - 1;
            out[pos++] = 32 * (w0 + j) + b;
          }
        }
        n += total;
      } else {
        for (int w = w0; w < w1; ++w) c += __popc(row[w] & ~seen[w]);
        int total;
        int pos = n + block_exclusive_scan(c, &total);
        for (int w = w0; w < w1; ++w) {
          uint32_t f = row[w] & ~seen[w];
          seen[w] |= f;
          while (f) {
            const int b = __ffs(f) - 1;
            f &= f - 1;
            out[pos++] = 32 * w + b;
          }
        }
        n += total;
      }
      // every thread's writes of this step precede the next step's reads of
      // `out` (the next node index) and of `seen`
      __syncthreads();
    }
  };
  expand(bits, 0, ncore);
  if (hbits) expand(hbits, 0, ncore);
  if (djflag) expand(bits, ncore, n);
  if (threadIdx.x == 0) {
    cnt[2 * p] = n;
    cnt[2 * p + 1] = ncore;
  }
}

// lists[p][0, len_p) -> packed[off_p, off_p + len_p)
__global__ void pack_lists_kernel(const int* __restrict__ lists, int cap,
                                  const int* __restrict__ off, int* __restrict__ packed) {
  const int p = blockIdx.x;
  const int len = off[p + 1] - off[p];
  for (int k = threadIdx.x; k < len; k += NT) packed[off[p] + k] = lists[(size_t)p * cap + k];
}

}  // namespace

int halo_indices_device(const DMat& G, prg_graph_partitioning_t* gp, int djflag, const DMat* H) {
  const int n = G.n;
  const int np = gp->totalParts;
  if (np <= 0 || !gp->nnodesInPart || !gp->nodeInPart) {
    set_error(ERR_ARG, "graph partitioning: no parts");
    return ERR_ARG;
  }
  const int W = (n + 31) / 32;
  if (W > kMaxWords) {
    set_error(ERR_UNSUPPORTED, "core+halo extraction: graphs above 512000 nodes are not supported");
    return ERR_UNSUPPORTED;
  }
  std::vector<int> core, core_off(np + 1, 0);
  for (int p = 0; p < np; ++p) {
    const int nc = gp->nnodesInPart[p];
    if (nc < 0 || (nc > 0 && !gp->nodeInPart[p])) {
      set_error(ERR_ARG, "graph partitioning: bad part");
      return ERR_ARG;
    }
    for (int k = 0; k < nc; ++k) {
      const int c = gp->nodeInPart[p][k];
      if (c < 0 || c >= n) {
        set_error(ERR_ARG, "graph partitioning: node index out of range");
        return ERR_ARG;
      }
      core.push_back(c);
    }
    core_off[p + 1] = (int)core.size();
  }
  int st = OK;
  if (H && H->n != n) {
    set_error(ERR_ARG, "core+halo extraction: graph and Hamiltonian sizes differ");
    return ERR_ARG;
  }
  uint32_t* bits = (uint32_t*)workspace("halo_bits", sizeof(uint32_t) * (size_t)n * W, &st);
  uint32_t* hbits =
      H ? (uint32_t*)workspace("halo_hbits", sizeof(uint32_t) * (size_t)n * W, &st) : nullptr;
  int* d_core = (int*)workspace("halo_core", sizeof(int) * std::max<size_t>(1, core.size()), &st);
  int* d_coff = (int*)workspace("halo_core_off", sizeof(int) * (np + 1), &st);
  int* d_lists = (int*)workspace("halo_lists", sizeof(int) * (size_t)np * n, &st);
  int* d_cnt = (int*)workspace("halo_cnt", sizeof(int) * 2 * np, &st);
  int* d_off = (int*)workspace("halo_off", sizeof(int) * (np + 1), &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(d_core, core.data(), sizeof(int) * core.size(),
                               hipMemcpyHostToDevice, stream()));
  PRG_HIP_CHECK(hipMemcpyAsync(d_coff, core_off.data(), sizeof(int) * (np + 1),
                               hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(graph_bits_kernel, dim3((n + 63) / 64, (W + 63) / 64), dim3(NT), 0, stream(),
                     G.p, G.ld, n, W, bits);
  PRG_HIP_CHECK(hipGetLastError());
  if (H) {
    hipLaunchKernelGGL(graph_bits_kernel, dim3((n + 63) / 64, (W + 63) / 64), dim3(NT), 0,
                       stream(), H->p, H->ld, n, W, hbits);
    PRG_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(halo_lists_kernel, dim3(np), dim3(NT), sizeof(uint32_t) * W, stream(), bits,
                     hbits, W, d_core, d_coff, djflag ? 1 : 0, n, d_lists, d_cnt);
  PRG_HIP_CHECK(hipGetLastError());
  std::vector<int> cnt(2 * (size_t)np);
  PRG_HIP_CHECK(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * 2 * np, hipMemcpyDeviceToHost,
                               stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  std::vector<int> off(np + 1, 0);
  for (int p = 0; p < np; ++p) off[p + 1] = off[p] + cnt[2 * p];
  std::vector<int> packed(std::max(1, off[np]));
  PRG_HIP_CHECK(hipMemcpyAsync(d_off, off.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice,
                               stream()));
  int* d_packed = (int*)workspace("halo_packed", sizeof(int) * packed.size(), &st);
  if (st) return st;
  hipLaunchKernelGGL(pack_lists_kernel, dim3(np), dim3(NT), 0, stream(), d_lists, n, d_off,
                     d_packed);
  PRG_HIP_CHECK(hipGetLastError());
  PRG_HIP_CHECK(hipMemcpyAsync(packed.data(), d_packed, sizeof(int) * off[np],
                               hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  for (int p = 0; p < np; ++p) {
    const int len = cnt[2 * p];
    std::free(gp->core_halo_index[p]);
    gp->core_halo_index[p] = (int*)std::malloc(sizeof(int) * std::max(1, len));
    if (!gp->core_halo_index[p]) {
      set_error(ERR_ARG, "graph partitioning: allocation failed");
      return ERR_ARG;
    }
    std::memcpy(gp->core_halo_index[p], packed.data() + off[p], sizeof(int) * len);
    gp->lsize[p] = len;
    gp->llsize[p] = cnt[2 * p + 1];
  }
  return OK;
}

}  // namespace prg
