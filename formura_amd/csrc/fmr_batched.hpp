// fmr_batched.hpp -- many independent 2D instances of one Formura program,
// of uneven sizes and per-instance scalar parameters, advanced together.
//
// Each instance i has its own periodic grid (nx_i, ny_i) (axis y contiguous),
// its own parameter vector (the program's named scalars, e.g. F, k, dt) and
// two device buffers per state field (ping-pong).  A launch advances every
// instance of a size bucket by T fused time steps: a persistent grid of
// workgroups walks a tile list that covers all instances of the bucket.
//
// Tile = one 64-lane wavefront covering 64*VY consecutive y-points (VY per
// lane) and marching along x over LX rows, with T time levels pipelined in a
// 3-row register window per level, as the 3D march kernel does one dimension
// up.  Neighbours: x +-1 from the window, y +-1 from the lane's own registers
// or, across lanes, by DPP wave shifts.  Ghost zones of T points at both y
// ends and T rows before and after the chunk are recomputed redundantly; reads
// wrap periodically.  The per-site arithmetic is the program's `point`, so an
// FP64 batch computes the same bits as the single-instance library.
#pragma once

#include "fmr_library.hpp"

#include <utility>
#include <vector>

namespace fmr {

template <class real>
struct Instance {
  real* buf[2][MAX_FIELDS];   // [parity][field]
  int nx, ny;
  double prm[MAX_PARAMS];
};

// One tile of the march: instance, first plane, first stored y, and `meta`
// = march chunk (planes) * 4 + lane-width selector (0: one point per lane,
// 1: two, 2: four -- used by the fused kernel; per-bucket launches ignore it).
struct Tile {
  int inst, x0, y0, meta;
};

__host__ __device__ constexpr int vy_of_sel(int sel) { return sel == 0 ? 1 : (sel == 1 ? 2 : 4); }
__host__ __device__ constexpr int sel_of_vy(int vy) { return vy == 1 ? 0 : (vy == 2 ? 1 : 2); }

template <class P, int T, int VY>
struct March2D {
  using real = typename P::real;
  static constexpr int NF = P::NF;
  static constexpr int WY = WAVE * VY;   // y-points per tile, ghosts included

  real w[T][3][VY][NF];
  real nxt[VY][NF];
  int lane;
  int nx, ny;   // the current instance's grid

  template <int L, int PH, int V>
  struct Acc : EvalReuse {
    const March2D& m;
    const double* prm;
    int cx, cy;
    real out[NF];
    template <int F, int D0, int D1, int D2>
    __device__ __forceinline__ real ld() const {
      static_assert(D2 == 0, "2D program");
      static_assert(D0 == 0 || D1 == 0, "star stencil");
      static_assert(D0 >= -1 && D0 <= 1 && D1 >= -1 && D1 <= 1, "radius 1");
      constexpr int NEW = PH, CEN = (PH + 2) % 3, OLD = (PH + 1) % 3;
      if constexpr (D0 == -1) return m.w[L][OLD][V][F];
      else if constexpr (D0 == 1) return m.w[L][NEW][V][F];
      else if constexpr (D1 == -1) {
        if constexpr (V > 0) return m.w[L][CEN][V - 1][F];
        else return lane_prev(m.w[L][CEN][VY - 1][F]);
      } else if constexpr (D1 == 1) {
        if constexpr (V < VY - 1) return m.w[L][CEN][V + 1][F];
        else return lane_next(m.w[L][CEN][0][F]);
      } else return m.w[L][CEN][V][F];
    }
    __device__ __forceinline__ double param(int i) const { return prm[i]; }
    // global coordinates wrapped into the grid: ghost rows and points are
    // periodic images of interior sites and take their index values
    __device__ __forceinline__ int index(int axis) const { return axis == 0 ? wrap(cx, m.nx) : wrap(cy, m.ny); }
    template <int F>
    __device__ __forceinline__ void st(real v) { out[F] = v; }
  };

  __device__ __forceinline__ void load_row(real* const* in, int xrow, int ny, const int (&yoff)[VY]) {
    const size_t base = (size_t)xrow * ny;
#pragma unroll
    for (int v = 0; v < VY; ++v)
#pragma unroll
      for (int f = 0; f < NF; ++f) nxt[v][f] = in[f][base + yoff[v]];
  }

  template <int L, int PH, int V>
  __device__ __forceinline__ void level_point(const double* prm, int q, int ybase, real* const* out, size_t qoff,
                                              const int (&yoff)[VY], bool row_ok, int ylo, int yhi) {
    if constexpr (V < VY) {
      Acc<L, PH, V> a{{}, *this, prm, q, ybase + V, {}};
      P::point(a);
      if constexpr (L + 1 < T) {
#pragma unroll
        for (int f = 0; f < NF; ++f) w[L + 1][PH][V][f] = a.out[f];
      } else {
        const int t = lane * VY + V;   // position within the tile
        if (row_ok && t >= ylo && t < yhi) {
#pragma unroll
          for (int f = 0; f < NF; ++f) __builtin_nontemporal_store(a.out[f], &out[f][qoff + yoff[V]]);
        }
      }
      level_point<L, PH, V + 1>(prm, q, ybase, out, qoff, yoff, row_ok, ylo, yhi);
    }
  }

  template <int L, int PH>
  __device__ __forceinline__ void levels(const double* prm, int xin, int x0, int xend, int ybase, real* const* out,
                                         int ny, const int (&yoff)[VY], int ylo, int yhi) {
    if constexpr (L < T) {
      const int q = xin - (L + 1);
      bool row_ok = false;
      size_t qoff = 0;
      if constexpr (L + 1 == T) {
        row_ok = q >= x0 && q < xend;
        qoff = (size_t)(row_ok ? q : 0) * ny;
      }
      level_point<L, PH, 0>(prm, q, ybase, out, qoff, yoff, row_ok, ylo, yhi);
      levels<L + 1, PH>(prm, xin, x0, xend, ybase, out, ny, yoff, ylo, yhi);
    }
  }

  template <int PH>
  __device__ __forceinline__ void iter(real* const* in, real* const* out, const double* prm, int xin, int x0,
                                       int xend, int nx, int ny, int ybase, const int (&yoff)[VY], int ylo,
                                       int yhi) {
#pragma unroll
    for (int v = 0; v < VY; ++v)
#pragma unroll
      for (int f = 0; f < NF; ++f) w[0][PH][v][f] = nxt[v][f];
    load_row(in, wrap(xin + 1, nx), ny, yoff);
    levels<0, PH>(prm, xin, x0, xend, ybase, out, ny, yoff, ylo, yhi);
  }
};

// March one tile by T steps, buffers [parity] -> [1 - parity], over the
// planes [x0, min(x0 + lx, nx)).
template <class P, int T, int VY>
__device__ __forceinline__ void march2d_tile(const Instance<typename P::real>& I, const Tile& tl, int parity,
                                             int lx) {
  using M = March2D<P, T, VY>;
  using real = typename P::real;
  M m;
  m.lane = threadIdx.x;
  real* const* in = I.buf[parity];
  real* const* out = I.buf[1 - parity];
  const int nx = I.nx, ny = I.ny;
  m.nx = nx;
  m.ny = ny;
  const int x0 = tl.x0;
  const int xend = min(x0 + lx, nx);
  const int ystart = tl.y0 - T;                 // tile position 0
  const int ybase = ystart + m.lane * VY;       // unwrapped y of this lane's first point
  int yoff[VY];
#pragma unroll
  for (int v = 0; v < VY; ++v) yoff[v] = wrap(ybase + v, ny);
  // positions [T, WY - T) are valid; positions past the grid end are
  // periodic duplicates of the first tile and are not stored
  const int ylo = T;
  const int yhi = min(M::WY - T, ny - ystart);
  const double* prm = I.prm;
  m.load_row(in, wrap(x0 - T, nx), ny, yoff);
  const int iters = xend - x0 + 2 * T;
  for (int s = 0; s < iters; s += 3) {
    const int xin = x0 - T + s;
    m.template iter<0>(in, out, prm, xin, x0, xend, nx, ny, ybase, yoff, ylo, yhi);
    m.template iter<1>(in, out, prm, xin + 1, x0, xend, nx, ny, ybase, yoff, ylo, yhi);
    m.template iter<2>(in, out, prm, xin + 2, x0, xend, nx, ny, ybase, yoff, ylo, yhi);
  }
}

// One launch: every tile of `tiles` (one size bucket) advanced by T steps
// with VY points per lane and chunks of lx planes.  Persistent: workgroup b
// handles tiles b, b + G, b + 2G, ...
template <class P, int T, int VY>
__global__ void __launch_bounds__(64) batched_march2d_kernel(const Instance<typename P::real>* insts,
                                                             const Tile* tiles, int ntiles, int parity, int lx) {
  for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    march2d_tile<P, T, VY>(insts[tl.inst], tl, parity, lx);
  }
}

// The fused launch: every tile of every instance, of all size buckets,
// advanced by T steps in one persistent kernel.  Each tile carries its own
// lane width (selector in meta & 3, one of those in the bit mask MASK:
// 1 = one point per lane, 2 = two, 4 = four) and chunk length (meta >> 2),
// the per-bucket choices of the plan.  The selector is uniform over the
// workgroup (one wavefront), so the branch is uniform.
template <class P, int T, int MASK>
__global__ void __launch_bounds__(64) batched_fused_kernel(const Instance<typename P::real>* insts,
                                                           const Tile* tiles, int ntiles, int parity) {
  for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const int sel = tl.meta & 3, lx = tl.meta >> 2;
    if constexpr ((MASK & 1) != 0)
      if (sel == 0) march2d_tile<P, T, 1>(insts[tl.inst], tl, parity, lx);
    if constexpr ((MASK & 2) != 0)
      if (sel == 1) march2d_tile<P, T, 2>(insts[tl.inst], tl, parity, lx);
    if constexpr ((MASK & 4) != 0)
      if (sel == 2) march2d_tile<P, T, 4>(insts[tl.inst], tl, parity, lx);
  }
}

// The plain method for comparison: one site per thread, one instance, one step.
template <class P>
__global__ void __launch_bounds__(256) batched_direct_kernel(const Instance<typename P::real>* insts, int inst,
                                                             int parity) {
  using real = typename P::real;
  const Instance<real>& I = insts[inst];
  Fields<real> in, out;
  for (int f = 0; f < P::NF; ++f) {
    in.f[f] = I.buf[parity][f];
    out.f[f] = I.buf[1 - parity][f];
  }
  Params prm;
  for (int i = 0; i < MAX_PARAMS; ++i) prm.v[i] = I.prm[i];
  const size_t total = (size_t)I.nx * I.ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    DirectAcc<P> a{{}, in, prm, {(int)(i / I.ny), (int)(i % I.ny), 0}, {I.nx, I.ny, 1}, {}};
    P::point(a);
#pragma unroll
    for (int f = 0; f < P::NF; ++f) out.f[f][i] = a.out[f];
  }
}

// ---------------------------------------------------------------------------
// host side
template <class P>
using BatchLauncher = void (*)(const Instance<typename P::real>*, const Tile*, int, int, int, int, hipStream_t);

template <class P, int T, int VY>
void launch_batched(const Instance<typename P::real>* insts, const Tile* tiles, int ntiles, int parity, int lx,
                    int grid, hipStream_t s) {
  hipLaunchKernelGGL((batched_march2d_kernel<P, T, VY>), dim3(grid), dim3(WAVE), 0, s, insts, tiles, ntiles, parity,
                     lx);
}

struct BatchVariant {
  int T, vy;
};

template <class P>
struct BatchVariantEntry {
  BatchVariant v;
  BatchLauncher<P> fn;
};

template <class P>
using FusedLauncher = void (*)(const Instance<typename P::real>*, const Tile*, int, int, int, hipStream_t);

template <class P, int T, int MASK>
void launch_fused(const Instance<typename P::real>* insts, const Tile* tiles, int ntiles, int parity, int grid,
                  hipStream_t s) {
  hipLaunchKernelGGL((batched_fused_kernel<P, T, MASK>), dim3(grid), dim3(WAVE), 0, s, insts, tiles, ntiles, parity);
}

// a fused kernel: depth T, lane widths of the bit mask `mask`
template <class P>
struct FusedEntry {
  int T, mask;
  FusedLauncher<P> fn;
};

template <class P>
class BatchLibrary {
  static_assert(P::NF <= MAX_FIELDS, "more state fields than fmr::MAX_FIELDS");

 public:
  using real = typename P::real;
  struct Bucket {
    std::vector<int> insts;
    int T = 1, vy = 1, lx = 64;
    const BatchVariantEntry<P>* entry = nullptr;
    Tile* d_tiles = nullptr;
    int ntiles = 0;
    int parity = 0;
    hipStream_t stream = nullptr;
    hipEvent_t join = nullptr;
  };

  int n = 0, np = 0;
  std::vector<Instance<real>> host;
  Instance<real>* d_inst = nullptr;
  std::vector<int> bucket_of;
  std::vector<Bucket> buckets;
  const BatchVariantEntry<P>* variants = nullptr;
  int nvariants = 0;
  const FusedEntry<P>* fused = nullptr;
  int nfused = 0;
  // the fused plan (every bucket of one fusion depth): one tile list over
  // all instances, one launch per T-block
  Tile* d_fused = nullptr;
  int fused_ntiles = 0, fused_T = 0, fused_mask = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int grid = 1024;
  // captured graphs, keyed by step count
  std::vector<std::pair<long, hipGraphExec_t>> graphs;
  int last_graph_nodes = 0;   // nodes of the graph captured last
  double* red_part = nullptr;
  double* red_res = nullptr;

  void init(int count, const int* nx, const int* ny, const double* prm, int nprm, const BatchVariantEntry<P>* vs,
            int nv, const FusedEntry<P>* fs, int nf) {
    n = count;
    np = nprm;
    variants = vs;
    nvariants = nv;
    fused = fs;
    nfused = nf;
    FMR_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    FMR_CHECK(hipEventCreate(&ev0));
    FMR_CHECK(hipEventCreate(&ev1));
    host.resize(n);
    for (int i = 0; i < n; ++i) {
      Instance<real>& I = host[i];
      I.nx = nx[i];
      I.ny = ny[i];
      for (int k = 0; k < MAX_PARAMS; ++k) I.prm[k] = k < np ? prm[i * np + k] : 0.0;
      const size_t cnt = (size_t)nx[i] * ny[i];
      for (int b = 0; b < 2; ++b)
        for (int f = 0; f < MAX_FIELDS; ++f) {
          I.buf[b][f] = nullptr;
          if (f < P::NF) {
            FMR_CHECK(hipMalloc(&I.buf[b][f], cnt * sizeof(real)));
            FMR_CHECK(hipMemsetAsync(I.buf[b][f], 0, cnt * sizeof(real), stream));
          }
        }
    }
    FMR_CHECK(hipMalloc(&d_inst, sizeof(Instance<real>) * n));
    FMR_CHECK(hipMemcpyAsync(d_inst, host.data(), sizeof(Instance<real>) * n, hipMemcpyHostToDevice, stream));
    FMR_CHECK(hipMalloc(&red_part, sizeof(double) * 3 * RED_BLOCKS));
    FMR_CHECK(hipMalloc(&red_res, sizeof(double) * 3));
    hipDeviceProp_t prop;
    FMR_CHECK(hipGetDeviceProperties(&prop, 0));
    grid = prop.multiProcessorCount * 4;   // persistent: 4 one-wave workgroups per CU
    bucket_of.assign(n, 0);
    FMR_CHECK(hipStreamSynchronize(stream));
  }

  // the fused tile list of the current plan, when every bucket has one depth
  // T and a fused kernel covers the lane widths in use; the instances are
  // brought to one parity (that of instance 0)
  void plan_fused() {
    if (d_fused) FMR_CHECK(hipFree(d_fused));
    d_fused = nullptr;
    fused_ntiles = fused_T = fused_mask = 0;
    int T = 0, mask = 0;
    for (auto& bk : buckets) {
      if (bk.insts.empty()) continue;
      if (T && bk.T != T) return;
      T = bk.T;
      mask |= 1 << sel_of_vy(bk.vy);
    }
    if (!T || !find_fused(T, mask)) return;
    const int par = buckets[bucket_of[0]].parity;
    std::vector<Tile> tl;
    for (auto& bk : buckets) {
      for (int i : bk.insts) {
        if (bk.parity != par)
          for (int f = 0; f < P::NF; ++f) {
            const size_t cnt = (size_t)host[i].nx * host[i].ny;
            FMR_CHECK(hipMemcpyAsync(host[i].buf[par][f], host[i].buf[bk.parity][f], cnt * sizeof(real),
                                     hipMemcpyDeviceToDevice, stream));
          }
        const int wy = WAVE * bk.vy - 2 * bk.T;
        for (int x0 = 0; x0 < host[i].nx; x0 += bk.lx)
          for (int y0 = 0; y0 < host[i].ny; y0 += wy) tl.push_back(Tile{i, x0, y0, bk.lx * 4 + sel_of_vy(bk.vy)});
      }
      bk.parity = par;
    }
    FMR_CHECK(hipMalloc(&d_fused, sizeof(Tile) * tl.size()));
    FMR_CHECK(hipMemcpyAsync(d_fused, tl.data(), sizeof(Tile) * tl.size(), hipMemcpyHostToDevice, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
    fused_ntiles = (int)tl.size();
    fused_T = T;
    fused_mask = mask;
  }

  // one fused launch per T-block, then the remainder at the deepest fused
  // depth that fits; every bucket's parity flips with each launch
  void enqueue_fused(int steps, hipStream_t s) {
    int left = steps;
    const int g = fused_ntiles < grid ? fused_ntiles : grid;
    while (left > 0) {
      const FusedEntry<P>* e = nullptr;
      for (int t = fused_T <= left ? fused_T : left; t >= 1 && !e; --t) e = find_fused(t, fused_mask);
      if (!e) {
        fprintf(stderr, "formura_amd: no fused batched kernel for the remainder\n");
        abort();
      }
      e->fn(d_inst, d_fused, fused_ntiles, buckets[0].parity, g, s);
      for (auto& bk : buckets) bk.parity ^= 1;
      left -= e->T;
    }
  }

  void advance_parity_fused(int steps) {
    int left = steps;
    while (left > 0) {
      const FusedEntry<P>* e = nullptr;
      for (int t = fused_T <= left ? fused_T : left; t >= 1 && !e; --t) e = find_fused(t, fused_mask);
      for (auto& bk : buckets) bk.parity ^= 1;
      left -= e->T;
    }
  }

  // release every device resource; the library can be initialised again
  void destroy() {
    FMR_CHECK(hipDeviceSynchronize());
    drop_graphs();
    for (auto& bk : buckets) {
      if (bk.d_tiles) FMR_CHECK(hipFree(bk.d_tiles));
      if (bk.stream) FMR_CHECK(hipStreamDestroy(bk.stream));
      if (bk.join) FMR_CHECK(hipEventDestroy(bk.join));
    }
    buckets.clear();
    if (d_fused) FMR_CHECK(hipFree(d_fused));
    d_fused = nullptr;
    fused_ntiles = fused_T = fused_mask = 0;
    for (auto& I : host)
      for (int b = 0; b < 2; ++b)
        for (int f = 0; f < P::NF; ++f)
          if (I.buf[b][f]) FMR_CHECK(hipFree(I.buf[b][f]));
    host.clear();
    if (d_inst) FMR_CHECK(hipFree(d_inst));
    if (red_part) FMR_CHECK(hipFree(red_part));
    if (red_res) FMR_CHECK(hipFree(red_res));
    d_inst = nullptr;
    red_part = red_res = nullptr;
    if (ev0) FMR_CHECK(hipEventDestroy(ev0));
    if (ev1) FMR_CHECK(hipEventDestroy(ev1));
    ev0 = ev1 = nullptr;
    if (stream) FMR_CHECK(hipStreamDestroy(stream));
    stream = nullptr;
    plain_parity = 0;
    n = 0;
  }

  const BatchVariantEntry<P>* find(int T, int vy) const {
    for (int i = 0; i < nvariants; ++i)
      if (variants[i].v.T == T && variants[i].v.vy == vy) return &variants[i];
    return nullptr;
  }

  // the fused kernel of depth T whose lane-width mask covers `mask` with the
  // fewest extra widths (a width the tiles do not use costs registers)
  const FusedEntry<P>* find_fused(int T, int mask) const {
    const FusedEntry<P>* best = nullptr;
    for (int i = 0; i < nfused; ++i)
      if (fused[i].T == T && (fused[i].mask & mask) == mask &&
          (!best || __builtin_popcount(fused[i].mask) < __builtin_popcount(best->mask)))
        best = &fused[i];
    return best;
  }

  void drop_graphs() {
    for (auto& g : graphs) FMR_CHECK(hipGraphExecDestroy(g.second));
    graphs.clear();
  }

  // nb buckets; bucket_of_inst[i] in [0, nb); var[3*b..] = (T, vy, lx)
  int plan(int nb, const int* bucket_of_inst, const int* var) {
    for (int b = 0; b < nb; ++b)
      if (!find(var[3 * b], var[3 * b + 1]) || var[3 * b + 2] < 1) return -1;
    for (int i = 0; i < n; ++i)
      if (bucket_of_inst[i] < 0 || bucket_of_inst[i] >= nb) return -1;
    FMR_CHECK(hipStreamSynchronize(stream));
    // the instances' current parity must survive re-planning
    std::vector<int> par(n, 0);
    for (int i = 0; i < n; ++i) par[i] = buckets.empty() ? 0 : buckets[bucket_of[i]].parity;
    for (auto& bk : buckets) {
      if (bk.d_tiles) FMR_CHECK(hipFree(bk.d_tiles));
      if (bk.stream) FMR_CHECK(hipStreamDestroy(bk.stream));
      if (bk.join) FMR_CHECK(hipEventDestroy(bk.join));
    }
    drop_graphs();
    buckets.assign(nb, Bucket());
    for (int i = 0; i < n; ++i) {
      bucket_of[i] = bucket_of_inst[i];
      buckets[bucket_of[i]].insts.push_back(i);
    }
    for (int b = 0; b < nb; ++b) {
      Bucket& bk = buckets[b];
      bk.T = var[3 * b];
      bk.vy = var[3 * b + 1];
      bk.lx = var[3 * b + 2];
      bk.entry = find(bk.T, bk.vy);
      FMR_CHECK(hipStreamCreateWithFlags(&bk.stream, hipStreamNonBlocking));
      FMR_CHECK(hipEventCreateWithFlags(&bk.join, hipEventDisableTiming));
      std::vector<Tile> tl;
      const int wy = WAVE * bk.vy - 2 * bk.T;
      for (int i : bk.insts) {
        if (par[i] != par[bk.insts[0]]) {
          // bring every instance of the bucket to one parity
          for (int f = 0; f < P::NF; ++f) {
            const size_t cnt = (size_t)host[i].nx * host[i].ny;
            FMR_CHECK(hipMemcpyAsync(host[i].buf[par[bk.insts[0]]][f], host[i].buf[par[i]][f], cnt * sizeof(real),
                                     hipMemcpyDeviceToDevice, stream));
          }
        }
        for (int x0 = 0; x0 < host[i].nx; x0 += bk.lx)
          for (int y0 = 0; y0 < host[i].ny; y0 += wy) tl.push_back(Tile{i, x0, y0, 0});
      }
      bk.parity = bk.insts.empty() ? 0 : par[bk.insts[0]];
      bk.ntiles = (int)tl.size();
      if (bk.ntiles) {
        FMR_CHECK(hipMalloc(&bk.d_tiles, sizeof(Tile) * tl.size()));
        FMR_CHECK(hipMemcpyAsync(bk.d_tiles, tl.data(), sizeof(Tile) * tl.size(), hipMemcpyHostToDevice, stream));
      }
    }
    FMR_CHECK(hipStreamSynchronize(stream));
    plan_fused();
    return 0;
  }

  // launches that advance bucket b by `steps` on stream s (T-deep launches,
  // then the remainder with shallower variants of the same lane width)
  void enqueue_bucket(Bucket& bk, int steps, hipStream_t s) {
    if (!bk.ntiles) return;
    int left = steps;
    const int g = bk.ntiles < grid ? bk.ntiles : grid;
    while (left > 0) {
      int t = bk.T <= left ? bk.T : left;
      const BatchVariantEntry<P>* e = nullptr;
      for (; t >= 1 && !e; --t) e = find(t, bk.vy);
      if (!e) {
        fprintf(stderr, "formura_amd: no batched variant for the remainder\n");
        abort();
      }
      e->fn(d_inst, bk.d_tiles, bk.ntiles, bk.parity, bk.lx, g, s);
      bk.parity ^= 1;
      left -= e->v.T;
    }
  }

  // the parity flips enqueue_bucket(bk, steps) performs, without launching
  void advance_parity(Bucket& bk, int steps) {
    if (!bk.ntiles) return;
    int left = steps;
    while (left > 0) {
      int t = bk.T <= left ? bk.T : left;
      const BatchVariantEntry<P>* e = nullptr;
      for (; t >= 1 && !e; --t) e = find(t, bk.vy);
      bk.parity ^= 1;
      left -= e->v.T;
    }
  }

  // the fused launches when the plan has them; otherwise all buckets on
  // parallel branches: fork from `stream`, join back
  void enqueue_all(int steps) {
    if (fused_ntiles) {
      enqueue_fused(steps, stream);
      return;
    }
    FMR_CHECK(hipEventRecord(ev0, stream));
    for (auto& bk : buckets) {
      FMR_CHECK(hipStreamWaitEvent(bk.stream, ev0, 0));
      enqueue_bucket(bk, steps, bk.stream);
      FMR_CHECK(hipEventRecord(bk.join, bk.stream));
      FMR_CHECK(hipStreamWaitEvent(stream, bk.join, 0));
    }
  }

  // advance every instance by `steps`; with use_graph the launch sequence is
  // captured once per step count into a hipGraph and replayed
  float advance(int steps, int use_graph) {
    if (buckets.empty()) {
      fprintf(stderr, "formura_amd: Formura_Batch_Plan must precede Formura_Batch_Advance\n");
      abort();
    }
    hipEvent_t t0, t1;
    FMR_CHECK(hipEventCreate(&t0));
    FMR_CHECK(hipEventCreate(&t1));
    if (!use_graph) {
      FMR_CHECK(hipEventRecord(t0, stream));
      enqueue_all(steps);
      FMR_CHECK(hipEventRecord(t1, stream));
    } else {
      hipGraphExec_t exec = nullptr;
      // a graph replays fixed kernel arguments, parities included: key it by
      // the step count and the bucket parities it starts from
      long key = steps;
      for (auto& bk : buckets) key = key * 2 + bk.parity;
      for (auto& g : graphs)
        if (g.first == key) exec = g.second;
      if (!exec) {
        hipGraph_t graph;
        FMR_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeGlobal));
        enqueue_all(steps);   // records the launches and advances the parities
        FMR_CHECK(hipStreamEndCapture(stream, &graph));
        size_t nodes = 0;
        FMR_CHECK(hipGraphGetNodes(graph, nullptr, &nodes));
        last_graph_nodes = (int)nodes;
        FMR_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        FMR_CHECK(hipGraphDestroy(graph));
        graphs.push_back({key, exec});
      } else if (fused_ntiles) {
        advance_parity_fused(steps);
      } else {
        for (auto& bk : buckets) advance_parity(bk, steps);
      }
      FMR_CHECK(hipEventRecord(t0, stream));
      FMR_CHECK(hipGraphLaunch(exec, stream));
      FMR_CHECK(hipEventRecord(t1, stream));
    }
    FMR_CHECK(hipEventSynchronize(t1));
    float ms = 0.f;
    FMR_CHECK(hipEventElapsedTime(&ms, t0, t1));
    FMR_CHECK(hipEventDestroy(t0));
    FMR_CHECK(hipEventDestroy(t1));
    return ms;
  }

  // one bucket alone (autotuning): launches on the main stream, no graph;
  // the bucket's parity then differs from the others', so the fused plan is
  // rebuilt by the next Formura_Batch_Plan (the autotuner re-plans anyway)
  float advance_bucket(int b, int steps) {
    if (d_fused) FMR_CHECK(hipFree(d_fused));
    d_fused = nullptr;
    fused_ntiles = 0;
    FMR_CHECK(hipStreamSynchronize(stream));
    FMR_CHECK(hipEventRecord(ev0, stream));
    enqueue_bucket(buckets[b], steps, stream);
    FMR_CHECK(hipEventRecord(ev1, stream));
    FMR_CHECK(hipEventSynchronize(ev1));
    float ms = 0.f;
    FMR_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    return ms;
  }

  // the plain method: one direct launch per instance per step
  float advance_plain(int steps) {
    FMR_CHECK(hipStreamSynchronize(stream));
    FMR_CHECK(hipEventRecord(ev0, stream));
    for (int s = 0; s < steps; ++s) {
      for (int i = 0; i < n; ++i) {
        const size_t cnt = (size_t)host[i].nx * host[i].ny;
        const int blocks = (int)((cnt + 255) / 256 < 65536 ? (cnt + 255) / 256 : 65536);
        hipLaunchKernelGGL((batched_direct_kernel<P>), dim3(blocks), dim3(256), 0, stream, d_inst, i,
                           parity_of(i));
      }
      FMR_CHECK(hipGetLastError());
      // every instance advanced one step: flip every parity
      if (buckets.empty()) plain_parity ^= 1;
      for (auto& bk : buckets) bk.parity ^= 1;
    }
    FMR_CHECK(hipEventRecord(ev1, stream));
    FMR_CHECK(hipEventSynchronize(ev1));
    float ms = 0.f;
    FMR_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    return ms;
  }
  int plain_parity = 0;

  int parity_of(int i) const { return buckets.empty() ? plain_parity : buckets[bucket_of[i]].parity; }

  void set_state(int i, int f, const double* src) {
    const size_t cnt = (size_t)host[i].nx * host[i].ny;
    std::vector<real> tmp(cnt);
    for (size_t k = 0; k < cnt; ++k) tmp[k] = (real)src[k];
    FMR_CHECK(hipMemcpyAsync(host[i].buf[parity_of(i)][f], tmp.data(), cnt * sizeof(real), hipMemcpyHostToDevice,
                             stream));
    FMR_CHECK(hipStreamSynchronize(stream));
  }

  void get_state(int i, int f, double* dst) {
    const size_t cnt = (size_t)host[i].nx * host[i].ny;
    std::vector<real> tmp(cnt);
    FMR_CHECK(hipStreamSynchronize(stream));
    FMR_CHECK(hipMemcpyAsync(tmp.data(), host[i].buf[parity_of(i)][f], cnt * sizeof(real), hipMemcpyDeviceToHost,
                             stream));
    FMR_CHECK(hipStreamSynchronize(stream));
    for (size_t k = 0; k < cnt; ++k) dst[k] = (double)tmp[k];
  }

  void* device_field(int i, int f) { return host[i].buf[parity_of(i)][f]; }

  template <class R>
  void diff(int i, int f, const R* ref, double res[3]) {
    const size_t cnt = (size_t)host[i].nx * host[i].ny;
    hipLaunchKernelGGL((reduce_pass1<real, R, 1>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream,
                       host[i].buf[parity_of(i)][f], ref, cnt, red_part);
    hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(64), 0, stream, red_part, red_res);
    FMR_CHECK(hipGetLastError());
    FMR_CHECK(hipMemcpyAsync(res, red_res, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
  }
};

}  // namespace fmr
