// fmr_library.hpp -- host side of a generated step library: device state,
// kernel variants, the upload / advance / download cycle behind
// `Formura_Forward`, FP64 reductions and the extra device-resident entry
// points declared in include/formura_amd.h.
#pragma once

#include "fmr_march.hpp"
#include "fmr_stencil.hpp"

#include <algorithm>
#include <vector>

namespace fmr {

// ---- FP64 reductions over a device field (deterministic two-pass) ---------
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// block b sums elements i = b*256 + t + k*(1024*256) in a fixed order
template <class A, class B, int MODE>
__global__ void __launch_bounds__(RED_THREADS) reduce_pass1(const A* a, const B* b, size_t n, double* part) {
  __shared__ double s0[RED_THREADS / 64], s1[RED_THREADS / 64], s2[RED_THREADS / 64];
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  for (size_t i = blockIdx.x * (size_t)RED_THREADS + threadIdx.x; i < n; i += (size_t)RED_BLOCKS * RED_THREADS) {
    const double x = (double)a[i];
    if (MODE == 0) {   // sum and sum of squares
      acc0 += x;
      acc1 += x * x;
    } else {           // L2 of difference, L2 of reference, max abs difference
      const double r = (double)b[i];
      const double d = x - r;
      acc0 += d * d;
      acc1 += r * r;
      acc2 = fmax(acc2, fabs(d));
    }
  }
  acc0 = wave_sum(acc0);
  acc1 = wave_sum(acc1);
  acc2 = wave_max(acc2);
  const int wv = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) { s0[wv] = acc0; s1[wv] = acc1; s2[wv] = acc2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int k = 0; k < RED_THREADS / 64; ++k) { t0 += s0[k]; t1 += s1[k]; t2 = fmax(t2, s2[k]); }
    part[3 * blockIdx.x + 0] = t0;
    part[3 * blockIdx.x + 1] = t1;
    part[3 * blockIdx.x + 2] = t2;
  }
}

__global__ void __launch_bounds__(64) reduce_pass2(const double* part, double* res) {
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  for (int k = threadIdx.x; k < RED_BLOCKS; k += 64) {
    t0 += part[3 * k];
    t1 += part[3 * k + 1];
    t2 = fmax(t2, part[3 * k + 2]);
  }
  t0 = wave_sum(t0);
  t1 = wave_sum(t1);
  t2 = wave_max(t2);
  if (threadIdx.x == 0) { res[0] = t0; res[1] = t1; res[2] = t2; }
}

// ---- the same comparison without any allocation ---------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops): the per-block partial sums live in `scratch`, a
// file-scope device array of 3 * RED_BLOCKS doubles (24 KiB) that each GPU
// the library runs on holds until the process ends.  Every block writes its
// own three partials on every call (also a block with no element: zeros), so
// nothing needs clearing between calls; calls that overlap on one GPU race
// on the array.  In host code `scratch` names a host-side stand-in, so the
// device address is looked up once per device (hipGetSymbolAddress) and only
// that address is handed to kernels.
constexpr long long NOALLOC_MAX_ELEMS = 1LL << 36;   // elements per array
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[3 * RED_BLOCKS];

struct ScratchTable {
  double* ptr[NOALLOC_MAX_DEVICES] = {};

  // the current device's copy of `scratch`, looked up on first use; null on error
  double* current() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    if (!ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) {
        fprintf(stderr, "formura_amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
        return nullptr;
      }
      ptr[dev] = (double*)p;
    }
    return ptr[dev];
  }
};

inline ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// res (device memory, 3 doubles) = {sum (a - ref)^2, sum ref^2, max |a - ref|}
// over n elements, enqueued on `s`; the kernels and their order are those of
// StepLibrary::diff, so a whole field gives the same bits.  Returns 0, or -1
// for a bad argument (nothing is launched then).
template <class A, class R>
int diff_noalloc(const A* a, const R* ref, long long n, double* res, hipStream_t s) {
  double* part = scratch_table().current();
  if (!part) return -1;
  hipLaunchKernelGGL((reduce_pass1<A, R, 1>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, a, ref, (size_t)n, part);
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(64), 0, s, (const double*)part, res);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

inline int diff_noalloc_entry(const void* a, int a_elem, const void* ref, int ref_elem, long long n, double* res,
                              hipStream_t s) {
  if (!a || !ref || !res || n < 0) return -1;
  if (n > NOALLOC_MAX_ELEMS) {
    fprintf(stderr, "formura_amd: Formura_Diff_Arrays_NoAlloc: %lld elements exceed the maximum %lld\n", n,
            NOALLOC_MAX_ELEMS);
    return -1;
  }
  if (a_elem == 8 && ref_elem == 8) return diff_noalloc((const double*)a, (const double*)ref, n, res, s);
  if (a_elem == 8 && ref_elem == 4) return diff_noalloc((const double*)a, (const float*)ref, n, res, s);
  if (a_elem == 4 && ref_elem == 8) return diff_noalloc((const float*)a, (const double*)ref, n, res, s);
  if (a_elem == 4 && ref_elem == 4) return diff_noalloc((const float*)a, (const float*)ref, n, res, s);
  return -1;
}

// ---- kernel variants --------------------------------------------------------
// KIND_MARCH: march3d_kernel (fmr_stencil.hpp); KIND_MARCHV1 / KIND_MARCHV2:
// marchv_kernel (fmr_march.hpp) with one / two z-sites per lane;
// KIND_MARCHV1E / KIND_MARCHV2E: the same with march_edge_rows(T) more rows
// on the first and the last wave of a tile (MarchRole)
enum KernelKind {
  KIND_DIRECT = 0, KIND_MARCH = 1, KIND_STAGED = 2, KIND_MARCHV1 = 3, KIND_MARCHV2 = 4, KIND_MARCHV1E = 5,
  KIND_MARCHV2E = 6
};
// Extra rows of the first and the last wave at fusion depth T: (T + 1) / 2,
// which makes their row-levels (T (BY + EY) - T (T + 1) / 2, level l skipped
// on the l rows nearest the tile edge) match the inner waves' T BY.
constexpr int march_edge_rows(int T) { return (T + 1) / 2; }

constexpr bool is_march(int kind) { return kind == KIND_MARCH || (kind >= KIND_MARCHV1 && kind <= KIND_MARCHV2E); }
constexpr bool two_sites(int kind) { return kind == KIND_MARCHV2 || kind == KIND_MARCHV2E; }
// rows of a march tile of W waves of `by` rows
constexpr int march_rows(int kind, int T, int by, int W) {
  return by * W + ((kind == KIND_MARCHV1E || kind == KIND_MARCHV2E) ? 2 * march_edge_rows(T) : 0);
}

// z-sites a wavefront stores per row of a tile: 64 * VZ less the ghost sites
// (march3d_kernel: none at T = 1, T per side otherwise; marchv_kernel: T per
// side rounded up to whole lanes)
constexpr int march_stored_z(int kind, int T) {
  return kind == KIND_MARCH ? (T == 1 ? WAVE : WAVE - 2 * T)
                            : (two_sites(kind) ? 2 * WAVE - 2 * ((T + 1) / 2 * 2) : WAVE - 2 * T);
}

// The z tiling of a march launch over nz sites: stored sites per tile row and
// ghost sites per side.  For two z-sites per lane the stored width is cut
// down to whole 128-byte lines when that takes no more tiles (nz = 512 in
// FP64: 112 instead of 120, five tiles either way): every tile then stores
// whole lines, none shares a line with its neighbour, and the wider ghost
// is lanes the tile computes anyway.  marchv_kernel with one site per lane
// and nz = 64: one tile of whole periodic z-lines (its lane moves rotate), no
// ghost sites.
inline void march_tile_z(int kind, int T, int nz, int elem, int& stored, int& ghost) {
  const int zw = two_sites(kind) ? 2 * WAVE : WAVE;
  stored = march_stored_z(kind, T);
  ghost = (zw - stored) / 2;
  if (kind != KIND_MARCH && !two_sites(kind) && nz == WAVE) {
    stored = zw;
    ghost = 0;
  } else if (two_sites(kind)) {
    const int line = 128 / elem;
    const int s1 = stored / line * line;
    if (s1 > 0 && (nz + s1 - 1) / s1 == (nz + stored - 1) / stored) {
      stored = s1;
      ghost = (zw - s1) / 2;
    }
  }
}

struct Variant {
  int kind;
  int T;       // time levels fused per launch
  int by;      // rows per lane (march kernels)
  int waves;   // wavefronts per tile (march kernels)
  int pf;      // prefetch depth in planes (march kernels)
  int occ;     // minimum waves per SIMD requested from the compiler (march kernels)
};

template <class P>
using MarchLauncher = void (*)(Fields<typename P::real>, Fields<typename P::real>, int, int, int, int, int, int,
                               int, int, Params, hipStream_t);

template <class P, int T, int BY, int W, int PF, int OCC>
void launch_march3d(Fields<typename P::real> in, Fields<typename P::real> out, int nx, int ny, int nz, int groups,
                    int nty, int ntz, int chunk, int /*ghost: T per side*/, Params prm, hipStream_t s) {
  if ((size_t)ny * nz * sizeof(typename P::real) >= ((size_t)1 << 31)) {
    fprintf(stderr, "formura_amd: march kernel needs a plane (NY*NZ sites) smaller than 2 GiB\n");
    abort();
  }
  hipLaunchKernelGGL((march3d_kernel<P, T, BY, W, PF, OCC>), dim3(groups), dim3(WAVE * W), 0, s, in, out, nx, ny, nz,
                     nty, ntz, chunk, prm);
}

template <class P, int T, int BY, int W, int VZ, int PF, int OCC, int EY = 0>
void launch_marchv(Fields<typename P::real> in, Fields<typename P::real> out, int nx, int ny, int nz, int groups,
                   int nty, int ntz, int chunk, int ghost, Params prm, hipStream_t s) {
  if ((size_t)ny * nz * sizeof(typename P::real) >= ((size_t)1 << 31)) {
    fprintf(stderr, "formura_amd: march kernel needs a plane (NY*NZ sites) smaller than 2 GiB\n");
    abort();
  }
  if (nz % VZ != 0) {
    fprintf(stderr, "formura_amd: march kernel with %d z-sites per lane needs NZ divisible by %d\n", VZ, VZ);
    abort();
  }
  if (ghost == 0 ? (VZ != 1 || nz != WAVE)
                 : (ghost < (T + VZ - 1) / VZ * VZ || ghost % VZ != 0 || 2 * ghost >= WAVE * VZ)) {
    fprintf(stderr, "formura_amd: march kernel: bad ghost width %d\n", ghost);
    abort();
  }
  // the x-carry with one z-site per lane (P::XC); with two, the window keeps
  // the plane behind (fewer registers at the peak of a row's evaluation)
  constexpr bool XC = VZ == 1 && P::XC;
  hipLaunchKernelGGL((marchv_kernel<P, T, BY, W, VZ, XC, PF, OCC, EY>), dim3(groups), dim3(WAVE * W), 0, s, in, out,
                     nx, ny, nz, nty, ntz, chunk, ghost, prm);
}

// one manifest node of a multi-stage step (stage_kernel<S>)
template <class P>
struct StageEntry {
  void (*fn)(const Arrays<typename P::real>&, const Outs<typename P::real>&, int, int, int, const Params&, int,
             hipStream_t);
  int nout;                 // manifest nodes evaluated by this kernel
  int to_state[MAX_OUTS];   // 1: writes the new value of state field `index`; 0: stage array `index`
  int index[MAX_OUTS];      // in the array list: state fields first, then stage arrays
};

template <class P, class S, int ROWS>
void launch_stage_rows(const Arrays<typename P::real>& in, const Outs<typename P::real>& out, int n0, int n1, int n2,
                       const Params& prm, hipStream_t s) {
  constexpr int LANE = P::DIM - 1;
  const int n[3] = {n0, n1, n2};
  const int row = LANE == 2 ? 1 : 0, slab = LANE == 2 ? 0 : 2;
  const int gy = (n[row] + ROWS - 1) / ROWS;
  if (gy > 65535 || n[slab] > 65535) {
    fprintf(stderr, "formura_amd: grid %d x %d x %d too large for the stage kernels\n", n0, n1, n2);
    abort();
  }
  const dim3 grid((n[LANE] + 63) / 64, gy, n[slab]);
  for (int a = 0; a < 3; ++a)
    if (n[a] < (a == 0 ? P::RX : a == 1 ? P::RY : P::RZ)) {
      fprintf(stderr, "formura_amd: grid %d x %d x %d is narrower than the stencil reach\n", n0, n1, n2);
      abort();
    }
  const bool wide = (size_t)n0 * n1 * n2 * sizeof(typename P::real) > 0xffffffffull;
  if (wide)
    hipLaunchKernelGGL((stage_kernel<S, LANE, true, ROWS>), grid, dim3(64, ROWS), 0, s, in, out, n0, n1, n2, prm);
  else
    hipLaunchKernelGGL((stage_kernel<S, LANE, false, ROWS>), grid, dim3(64, ROWS), 0, s, in, out, n0, n1, n2, prm);
}

// rows: wavefronts per workgroup, 4 (default, also for 0) or 16
template <class P, class S>
void launch_stage(const Arrays<typename P::real>& in, const Outs<typename P::real>& out, int n0, int n1, int n2,
                  const Params& prm, int rows, hipStream_t s) {
  if (rows == 16)
    launch_stage_rows<P, S, 16>(in, out, n0, n1, n2, prm, s);
  else
    launch_stage_rows<P, S, 4>(in, out, n0, n1, n2, prm, s);
}

template <class P>
struct VariantEntry {
  Variant v;
  MarchLauncher<P> fn;
};

template <class P>
class StepLibrary {
  static_assert(P::NF <= MAX_FIELDS, "more state fields than fmr::MAX_FIELDS");

 public:
  using real = typename P::real;
  int n[3] = {1, 1, 1};
  size_t count = 0;
  Fields<real> buf[2];
  int cur = 0;
  hipStream_t stream = nullptr;
  Params prm{};
  int nparams = 0;
  const VariantEntry<P>* variants = nullptr;
  int nvariants = 0;
  Variant choice{KIND_DIRECT, 1, 0, 1, 1, 1};
  int lx = 64;
  int simds = 1024;   // SIMDs of the device (CUs x 4), read at init
  double* red_part = nullptr;
  double* red_res = nullptr;
  void* staging = nullptr;
  size_t staging_bytes = 0;
  hipEvent_t ev0 = nullptr, evm = nullptr, ev1 = nullptr;
  // multi-stage programs: the manifest-node kernels and their arrays
  const StageEntry<P>* stages = nullptr;
  int nkernels = 0;   // stage kernels per step
  int nstages = 0;    // manifest nodes they evaluate
  int ntemps = 0;
  real* temps[MAX_ARRAYS] = {};
  bool ready = false;

  void init(int n0, int n1, int n2, const VariantEntry<P>* vs, int nv) {
    if (ready) return;
    n[0] = n0; n[1] = n1; n[2] = n2;
    count = (size_t)n0 * n1 * n2;
    variants = vs;
    nvariants = nv;
    FMR_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    {
      int dev = 0, cus = 0;
      FMR_CHECK(hipGetDevice(&dev));
      FMR_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      simds = 4 * (cus > 0 ? cus : 256);
    }
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < MAX_FIELDS; ++f) buf[b].f[f] = nullptr;
    for (int b = 0; b < 2; ++b)
      for (int f = 0; f < P::NF; ++f) {
        FMR_CHECK(hipMalloc(&buf[b].f[f], count * sizeof(real)));
        FMR_CHECK(hipMemsetAsync(buf[b].f[f], 0, count * sizeof(real), stream));
      }
    FMR_CHECK(hipMalloc(&red_part, sizeof(double) * 3 * RED_BLOCKS));
    FMR_CHECK(hipMalloc(&red_res, sizeof(double) * 3));
    FMR_CHECK(hipEventCreate(&ev0));
    FMR_CHECK(hipEventCreate(&evm));
    FMR_CHECK(hipEventCreate(&ev1));
    scratch_table().current();   // the no-allocation entry point's scratch on this device
    // default: the fused variant with T = 2 if compiled in, else the first one
    // (programs without an inlined point function: the stage kernels, set by
    // init_stages)
    choice = Variant{KIND_DIRECT, 1, 0, 1, 1, 1};
    for (int i = 0; i < nvariants; ++i)
      if (choice.kind == KIND_DIRECT || (variants[i].v.T == 2 && choice.T != 2)) choice = variants[i].v;
    FMR_CHECK(hipStreamSynchronize(stream));
    ready = true;
  }

  // the nk stage kernels of a multi-stage step, with nt stage arrays
  void init_stages(const StageEntry<P>* st, int nk, int nt) {
    if (P::NF + nt > MAX_ARRAYS) {
      fprintf(stderr, "formura_amd: %d state fields + %d stage arrays exceed %d\n", P::NF, nt, MAX_ARRAYS);
      abort();
    }
    stages = st;
    nkernels = nk;
    nstages = 0;
    for (int k = 0; k < nk; ++k) nstages += st[k].nout;
    ntemps = nt;
    if (!P::INLINED) choice = Variant{KIND_STAGED, 1, 0, 1, 1, 1};
    for (int t = 0; t < nt; ++t) {
      FMR_CHECK(hipMalloc(&temps[t], count * sizeof(real)));
      FMR_CHECK(hipMemsetAsync(temps[t], 0, count * sizeof(real), stream));
    }
    FMR_CHECK(hipStreamSynchronize(stream));
  }

  // one step as the sequence of manifest-node kernels
  void launch_staged(int steps) {
    for (int s = 0; s < steps; ++s) {
      Arrays<real> a;
      for (int i = 0; i < MAX_ARRAYS; ++i) a.a[i] = nullptr;
      for (int f = 0; f < P::NF; ++f) a.a[f] = buf[cur].f[f];
      for (int t = 0; t < ntemps; ++t) a.a[P::NF + t] = temps[t];
      for (int k = 0; k < nkernels; ++k) {
        const StageEntry<P>& e = stages[k];
        Outs<real> o;
        for (int j = 0; j < MAX_OUTS; ++j) o.o[j] = nullptr;
        for (int j = 0; j < e.nout; ++j)
          o.o[j] = e.to_state[j] ? buf[1 - cur].f[e.index[j]] : temps[e.index[j] - P::NF];
        e.fn(a, o, n[0], n[1], n[2], prm, choice.by, stream);
        FMR_CHECK(hipGetLastError());
      }
      cur = 1 - cur;
    }
  }

  void* host_staging(size_t bytes) {
    if (bytes > staging_bytes) {
      if (staging) FMR_CHECK(hipHostFree(staging));
      FMR_CHECK(hipHostMalloc(&staging, bytes, hipHostMallocDefault));
      staging_bytes = bytes;
    }
    return staging;
  }

  // host array with shape (s0, s1, s2) and interior starting at (o0, o1, o2)
  template <class H>
  void upload(int f, const H* host, const int s[3], const int o[3]) {
    real* st = (real*)host_staging(count * sizeof(real));
    for (int x = 0; x < n[0]; ++x)
      for (int y = 0; y < n[1]; ++y) {
        const H* src = host + ((size_t)(x + o[0]) * s[1] + (y + o[1])) * s[2] + o[2];
        real* dst = st + ((size_t)x * n[1] + y) * n[2];
        for (int z = 0; z < n[2]; ++z) dst[z] = (real)src[z];
      }
    FMR_CHECK(hipMemcpyAsync(buf[cur].f[f], st, count * sizeof(real), hipMemcpyHostToDevice, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
  }

  template <class H>
  void download(int f, H* host, const int s[3], const int o[3]) {
    real* st = (real*)host_staging(count * sizeof(real));
    FMR_CHECK(hipMemcpyAsync(st, buf[cur].f[f], count * sizeof(real), hipMemcpyDeviceToHost, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
    for (int x = 0; x < n[0]; ++x)
      for (int y = 0; y < n[1]; ++y) {
        H* dst = host + ((size_t)(x + o[0]) * s[1] + (y + o[1])) * s[2] + o[2];
        const real* src = st + ((size_t)x * n[1] + y) * n[2];
        for (int z = 0; z < n[2]; ++z) dst[z] = (H)src[z];
      }
  }

  const VariantEntry<P>* find(int kind, int T, int by, int waves, int pf, int occ) const {
    for (int i = 0; i < nvariants; ++i)
      if (variants[i].v.kind == kind && variants[i].v.T == T && variants[i].v.by == by &&
          variants[i].v.waves == waves && variants[i].v.pf == pf && variants[i].v.occ == occ)
        return &variants[i];
    return nullptr;
  }

  void launch_direct(int steps) {
    if constexpr (!P::INLINED) {
      fprintf(stderr, "formura_amd: this program has no inlined point function (too large); use the stage kernels\n");
      abort();
    }
    const int threads = 256;
    const size_t blocks_needed = (count + threads - 1) / threads;
    const int blocks = (int)(blocks_needed < 65536 ? blocks_needed : 65536);
    for (int s = 0; s < steps; ++s) {
      if constexpr (P::INLINED)
        hipLaunchKernelGGL((direct_step_kernel<P>), dim3(blocks), dim3(threads), 0, stream, buf[cur], buf[1 - cur],
                           n[0], n[1], n[2], prm);
      FMR_CHECK(hipGetLastError());
      cur = 1 - cur;
    }
  }

  // work split of a march launch (march_split_of, with the library's lx and
  // the device's SIMD count)
  void march_split(int nty, int ntz, int W, int T, int& groups, int& chunk) const {
    march_split_of((long long)nty * ntz, n[0], W, T, lx, simds, groups, chunk);
  }

  void launch_march(const VariantEntry<P>* e) {
    const int T = e->v.T, rows = march_rows(e->v.kind, T, e->v.by, e->v.waves);
    const int nty = (n[1] + (rows - 2 * T) - 1) / (rows - 2 * T);
    int zw = 0, ghost = 0;
    march_tile_z(e->v.kind, T, n[2], (int)sizeof(real), zw, ghost);
    const int ntz = (n[2] + zw - 1) / zw;
    int groups = 0, chunk = 0;
    march_split(nty, ntz, e->v.waves, T, groups, chunk);
    e->fn(buf[cur], buf[1 - cur], n[0], n[1], n[2], groups, nty, ntz, chunk, ghost, prm, stream);
    FMR_CHECK(hipGetLastError());
    cur = 1 - cur;
  }

  // the fused variant selected by `choice`, or null (direct kernel, or not compiled in)
  const VariantEntry<P>* main_variant() const {
    if (!is_march(choice.kind)) return nullptr;
    return find(choice.kind, choice.T, choice.by, choice.waves, choice.pf, choice.occ);
  }

  // launches of the selected kernel for `steps` steps (the rest is done by advance_rest)
  int advance_main(int steps) {
    if (choice.kind == KIND_DIRECT) {
      launch_direct(steps);
      return steps;
    }
    if (choice.kind == KIND_STAGED) {
      launch_staged(steps);
      return steps;
    }
    const VariantEntry<P>* e = main_variant();
    if (!e) return 0;
    int k = 0;
    for (; (k + 1) * e->v.T <= steps; ++k) launch_march(e);
    return k;
  }

  // `left` < T steps: a variant of the deepest fusion depth that fits (the
  // selected tile shape if it is compiled at that depth, else the first
  // variant of that depth), else the direct kernel
  void advance_rest(int left) {
    while (left > 0) {
      const VariantEntry<P>* e = nullptr;
      int depth = 0;
      for (int i = 0; i < nvariants; ++i)
        if (variants[i].v.T <= left && variants[i].v.T > depth) depth = variants[i].v.T;
      if (depth > 0) e = find(choice.kind, depth, choice.by, choice.waves, choice.pf, choice.occ);
      for (int i = 0; i < nvariants && !e; ++i)
        if (variants[i].v.T == depth) e = &variants[i];
      if (!e) {
        launch_direct(left);
        return;
      }
      launch_march(e);
      left -= e->v.T;
    }
  }

  int steps_of_main(int launches) const {
    const VariantEntry<P>* e = main_variant();
    if (choice.kind == KIND_DIRECT || choice.kind == KIND_STAGED) return launches;
    return e ? launches * e->v.T : 0;
  }

  // ---- hipGraphs of the launch sequences --------------------------------
  // A sequence of launches (the selected kernel's for `steps` steps, or the
  // remainder) is captured once per (kernel choice, work split, step count,
  // starting buffer) and replayed: on a grid small enough for the launches to
  // be short (64^3: a few microseconds each) the host's launch cost would
  // otherwise dominate.  A replay applies the buffer flips the capture made.
  struct GraphKey {
    int kind, T, by, waves, pf, occ, lx, cur, steps, part;
    bool operator==(const GraphKey& o) const { return memcmp(this, &o, sizeof(GraphKey)) == 0; }
  };
  struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec;
    int cur_after, launches;
  };
  std::vector<GraphEntry> graphs;
  bool use_graphs = true;
  bool prepare_only = false;   // capture what is missing, launch nothing

  void drop_graphs() {
    for (auto& g : graphs) FMR_CHECK(hipGraphExecDestroy(g.exec));
    graphs.clear();
  }

  // run `enqueue` (which returns a launch count and flips `cur`) through a
  // captured graph; returns the launch count
  template <class F>
  int graphed(int steps, int part, F enqueue) {
    if (!use_graphs || steps <= 0) return enqueue();
    GraphKey k;
    memset(&k, 0, sizeof k);
    k = GraphKey{choice.kind, choice.T, choice.by, choice.waves, choice.pf, choice.occ, lx, cur, steps, part};
    for (auto& g : graphs)
      if (g.key == k) {
        if (!prepare_only) FMR_CHECK(hipGraphLaunch(g.exec, stream));
        cur = g.cur_after;
        return g.launches;
      }
    hipGraph_t graph;
    FMR_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const int launches = enqueue();
    FMR_CHECK(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t exec;
    FMR_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    FMR_CHECK(hipGraphDestroy(graph));
    graphs.push_back(GraphEntry{k, exec, cur, launches});
    if (!prepare_only) FMR_CHECK(hipGraphLaunch(exec, stream));
    return launches;
  }

  // capture and instantiate the graphs that advance(steps) from the current
  // buffer will replay, without running them or changing the state: a
  // caller can keep the one-time capture out of a timed or latency-bound call
  void prepare(int steps) {
    if (!use_graphs || steps <= 0) return;
    const int cur0 = cur;
    prepare_only = true;
    advance(steps);
    prepare_only = false;
    cur = cur0;
  }

  int advance_main_g(int steps) {
    return graphed(steps, 0, [&] { return advance_main(steps); });
  }
  void advance_rest_g(int left) {
    graphed(left, 1, [&] {
      advance_rest(left);
      return 0;
    });
  }

  // advance the device state by `steps` time steps (asynchronous on `stream`)
  void advance(int steps) { advance_rest_g(steps - steps_of_main(advance_main_g(steps))); }

  // as advance, timed with events on the library stream: returns the total
  // milliseconds; last_main_ms / last_main_launches time the selected kernel's
  // own launches (they run first), separately from the remainder steps
  float last_main_ms = 0.f, last_total_ms = 0.f;
  int last_main_launches = 0;
  float timed_advance(int steps) {
    FMR_CHECK(hipEventRecord(ev0, stream));
    last_main_launches = advance_main_g(steps);
    FMR_CHECK(hipEventRecord(evm, stream));
    advance_rest_g(steps - steps_of_main(last_main_launches));
    FMR_CHECK(hipEventRecord(ev1, stream));
    FMR_CHECK(hipEventSynchronize(ev1));
    FMR_CHECK(hipEventElapsedTime(&last_main_ms, ev0, evm));
    FMR_CHECK(hipEventElapsedTime(&last_total_ms, ev0, ev1));
    return last_total_ms;
  }

  void sync() { FMR_CHECK(hipStreamSynchronize(stream)); }

  // res = {sum, sum of squares, 0}
  void reduce(int f, double res[3]) {
    hipLaunchKernelGGL((reduce_pass1<real, real, 0>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream,
                       buf[cur].f[f], (const real*)nullptr, count, red_part);
    hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(64), 0, stream, red_part, red_res);
    FMR_CHECK(hipGetLastError());
    FMR_CHECK(hipMemcpyAsync(res, red_res, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
  }

  // res = {sum (a-ref)^2, sum ref^2, max |a-ref|} against a device field `ref`
  template <class R>
  void diff(int f, const R* ref, double res[3]) {
    hipLaunchKernelGGL((reduce_pass1<real, R, 1>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, stream, buf[cur].f[f],
                       ref, count, red_part);
    hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(64), 0, stream, red_part, red_res);
    FMR_CHECK(hipGetLastError());
    FMR_CHECK(hipMemcpyAsync(res, red_res, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    FMR_CHECK(hipStreamSynchronize(stream));
  }
};

}  // namespace fmr
