// MI355X (gfx950) hierarchical-matrix matvec engine and its C-ABI.
//
// One matvec y = H x is five kernels on the caller's stream:
//   k_gather   x -> xt[0:n]   (permutation of matvecOriginal, hmat.cc:586-593,
//                              or a copy; converts to the factor precision)
//   k_slab     pass 1: t = W x   (W_k = B_k), each row written through the
//              row map into xt[xt_t : ...]
//   k_reduce   sum of split pass-1 work items (only if any)
//   k_slab     pass 2: y = F x + U t, each y row written once, with the
//              permutation of matvecOriginal (hmat.cc:618-626) fused in
//   k_reduce   sum of split pass-2 work items (only if any)
// Every byte of the factor data is streamed once per matvec with 16-byte
// non-temporal vector loads; x and t are re-read from L2 / MALL.  See
// layout.h for the slab format.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/bigwham_hip.h"
#include "layout.h"
#include "recompress.h"

// Scratch of bw_hip_matvec_noalloc: x gathered in the factor precision and
// the rank vector t (xt), then the partial sums of both passes, in one fixed
// array per GPU (each GPU that loads this library holds its own copy until
// the process ends).  Host code never passes `scratch` itself to a kernel: its
// device address on each GPU comes from hipGetSymbolAddress (noalloc_bind).
static __device__ __align__(256) unsigned char scratch[BW_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("BIGWHAM_VERBOSE");
    return e && e[0] && e[0] != '0';
  }();
  return v;
}

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(BW_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec;
template <>
struct Vec<double> {
  static constexpr int V = 2;
  typedef dbl2 type;
};
template <>
struct Vec<float> {
  static constexpr int V = 4;
  typedef flt4 type;
};

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X
// (workgroup b runs on XCD b % 8).  Each XCD gets a contiguous range of work
// items of about equal bytes (neighbouring row clusters share x entries in
// that XCD's L2, and no XCD is left with the tail); its workgroups take the
// range WAVES WIs at a time (WAVES = waves per workgroup).  Returns the WI of
// this wave, or -1.
struct XcdSplit {
  int32_t s[bwl::kXcds + 1];
};

template <int WAVES = 4>
__device__ __forceinline__ int xcd_work_item(const XcdSplit& xs) {
  const int x = blockIdx.x % bwl::kXcds;
  const int k = blockIdx.x / bwl::kXcds;
  const int w = xs.s[x] + WAVES * k + static_cast<int>(threadIdx.x >> 6);
  return w < xs.s[x + 1] ? w : -1;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ x,
                                                T* __restrict__ xt, int64_t n,
                                                const int64_t* __restrict__ perm, int dof) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t src = i;
  if (perm) {
    const int64_t p = i / dof;
    src = dof * perm[p] + (i - p * dof);
  }
  xt[i] = static_cast<T>(x[src]);
}

template <typename OutT>
__device__ __forceinline__ int64_t out_index(int64_t idx, const int64_t* perm, int dof) {
  if (!perm) return idx;
  // row indices fit in 32 bits (checked when the layout is built)
  const int i = static_cast<int>(idx);
  const int p = i / dof;
  return dof * perm[p] + (i - p * dof);
}

// One wavefront per work item and per workgroup: work items differ in size
// (small pass-1 slabs next to long ones), and a workgroup of several waves
// would hold its slot until its longest wave ends.
// T: factor / accumulation type; OutT: output type (T for pass 1, double
// for pass 2); U: unroll depth (independent 16-byte loads in flight per lane);
// PASS only names the two launches apart in profiles.
template <typename T, typename OutT, int U, int PASS>
__global__ __launch_bounds__(64) void k_slab(const bwl::WorkItem* __restrict__ wis, XcdSplit xs,
                                              const int32_t* __restrict__ gidx,
                                              const T* __restrict__ data,
                                              const T* __restrict__ xin,
                                              OutT* __restrict__ out, T* __restrict__ part,
                                              const int64_t* __restrict__ perm, int dof,
                                              const int32_t* __restrict__ rowmap) {
  typedef typename Vec<T>::type VT;
  constexpr int V = Vec<T>::V;
  const int lane = threadIdx.x & 63;
  const int w = xcd_work_item<1>(xs);
  if (w < 0) return;
  const bwl::WorkItem wi = wis[w];
  const int R = wi.rows;
  const int ng = wi.ngroups;
  const VT* __restrict__ base = reinterpret_cast<const VT*>(data + wi.data_off);
  const int32_t* __restrict__ gi = gidx + wi.gbeg;
  T acc0 = 0, acc1 = 0;
  int nsplit = 1;
  int split = 0;
  if (R > 32) {
    // One column group per iteration for the whole wavefront: group index
    // and input entries are wave-uniform (scalar loads).
    if (lane < R) {
      const VT* p = base + lane;
      int g = 0;
      // Full batches of U groups through two register buffers: the loads of
      // batch b + 1 go out before the FMAs of batch b, so a batch's loads are
      // in flight while the previous one is consumed (one buffer drained to
      // nothing before the next loads went out).  The loop body issues the
      // same loads on every trip and the last one or two batches are peeled,
      // so the waits the compiler places are the same on every path.  Same
      // groups, same accumulator per group and same order: same bits.
      const int nfull = ng / U;
      if (nfull > 0) {
        VT a[U], b[U];
        int za[U], zb[U];
        auto load = [&](int g0, VT (&buf)[U], int (&zz)[U]) {
#pragma unroll
          for (int u = 0; u < U; ++u) zz[u] = gi[g0 + u];
#pragma unroll
          for (int u = 0; u < U; ++u) buf[u] = __builtin_nontemporal_load(p + (g0 + u) * R);
        };
        auto use = [&](const VT (&buf)[U], const int (&zz)[U]) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            // unsigned: no sign extension of the loaded offset is placed
            // (and waited for) next to its load
            const uint32_t z = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(zz[u]));
#pragma unroll
            for (int v = 0; v < V; ++v) {
              if (u & 1)
                acc1 += buf[u][v] * xin[z + v];
              else
                acc0 += buf[u][v] * xin[z + v];
            }
          }
        };
        load(0, a, za);
        int bt = 0;
        for (; bt + 2 < nfull; bt += 2) {
          __builtin_amdgcn_sched_barrier(0);
          load((bt + 1) * U, b, zb);
          __builtin_amdgcn_sched_barrier(0);
          use(a, za);
          __builtin_amdgcn_sched_barrier(0);
          load((bt + 2) * U, a, za);
          __builtin_amdgcn_sched_barrier(0);
          use(b, zb);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (bt + 2 == nfull) {
          load((bt + 1) * U, b, zb);
          __builtin_amdgcn_sched_barrier(0);
          use(a, za);
          __builtin_amdgcn_sched_barrier(0);
          use(b, zb);
        } else {
          use(a, za);
        }
        __builtin_amdgcn_sched_barrier(0);
        g = nfull * U;
      }
      // the last, partial batch in steps of four groups, then one at a time
      for (; g + 4 <= ng; g += 4) {
        VT a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = __builtin_nontemporal_load(p + (g + u) * R);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int z = __builtin_amdgcn_readfirstlane(gi[g + u]);
#pragma unroll
          for (int v = 0; v < V; ++v) acc0 += a[u][v] * xin[z + v];
        }
      }
      for (; g < ng; ++g) {
        const VT a = __builtin_nontemporal_load(p + g * R);
        const int z = __builtin_amdgcn_readfirstlane(gi[g]);
#pragma unroll
        for (int v = 0; v < V; ++v) acc0 += a[v] * xin[z + v];
      }
    }
  } else {
    // Several column groups per iteration: lane = split * R + row.
    nsplit = 64 / R;
    split = lane / R;
    const int row = lane - split * R;
    if (split < nsplit && ng > 0) {
      const VT* p = base + row;
      const int step = U * nsplit;
      // The input offsets of a step are loaded one step ahead, so the input
      // entries are requested together with the factors and one memory
      // round trip covers both (offsets past the end are clamped, unused).
      int zc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) zc[u] = gi[min(split + u * nsplit, ng - 1)];
      int g = split;
      for (; g + (U - 1) * nsplit < ng; g += step) {
        VT a[U];
        T xv[U][V];
        int zn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) a[u] = __builtin_nontemporal_load(p + (g + u * nsplit) * R);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int v = 0; v < V; ++v) xv[u][v] = xin[zc[u] + v];
#pragma unroll
        for (int u = 0; u < U; ++u) zn[u] = gi[min(g + step + u * nsplit, ng - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int v = 0; v < V; ++v) {
            if (u & 1)
              acc1 += a[u][v] * xv[u][v];
            else
              acc0 += a[u][v] * xv[u][v];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) zc[u] = zn[u];
      }
      for (; g < ng; g += nsplit) {
        const VT a = __builtin_nontemporal_load(p + g * R);
        const int z = gi[g];
#pragma unroll
        for (int v = 0; v < V; ++v) acc0 += a[v] * xin[z + v];
      }
    }
  }
  T acc = acc0 + acc1;
  // fold the column splits onto lanes 0..R-1, in split order
  for (int s = 1; s < nsplit; ++s) {
    const T o = __shfl(acc, lane + s * R, 64);
    if (split == 0) acc += o;
  }
  if (wi.part_off < 0) {
    if (lane < R) {
      // pass 1 rows go to their t entries through the row map
      const int64_t orow = rowmap ? rowmap[wi.out_row + lane] : static_cast<int64_t>(wi.out_row) + lane;
      out[out_index<OutT>(orow, perm, dof)] = static_cast<OutT>(acc);
    }
    return;
  }
  // Split slab: the partial sums are added in work-item order by k_reduce.
  // (A last-work-item-reduces variant needs agent-scope release/acquire
  // fences, which write back and invalidate the per-XCD L2 and cost more than
  // the separate pass saves.)
  if (lane < R) part[wi.part_off + lane] = acc;
}

template <typename T, typename OutT>
__global__ __launch_bounds__(64) void k_reduce(const bwl::Reduce* __restrict__ reds, int nred,
                                               const T* __restrict__ part, OutT* out,
                                               const int64_t* __restrict__ perm, int dof,
                                               const int32_t* __restrict__ rowmap) {
  const int r = blockIdx.x;
  if (r >= nred) return;
  const bwl::Reduce rd = reds[r];
  const int t = threadIdx.x;
  if (t >= rd.rows) return;
  // four interleaved partial sums keep four loads in flight; the order of
  // the additions is fixed, so the result is deterministic
  const T* p = part + rd.part_off + t;
  const int R = rd.rows;
  T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int q = 0;
  for (; q + 4 <= rd.nparts; q += 4) {
    s0 += p[(q + 0) * R];
    s1 += p[(q + 1) * R];
    s2 += p[(q + 2) * R];
    s3 += p[(q + 3) * R];
  }
  for (; q < rd.nparts; ++q) s0 += p[q * R];
  const T s = (s0 + s1) + (s2 + s3);
  const int64_t orow = rowmap ? rowmap[rd.out_row + t] : static_cast<int64_t>(rd.out_row) + t;
  out[out_index<OutT>(orow, perm, dof)] = static_cast<OutT>(s);
}

// Sub-matrix selection: the input is scattered into the (zeroed) x part of xt
// at the permuted points of the selected columns, and the output gathered
// from a full permuted y at the selected rows.
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_sel(const double* __restrict__ x,
                                                     T* __restrict__ xt,
                                                     const int64_t* __restrict__ src,
                                                     const int64_t* __restrict__ dst, int64_t n,
                                                     int dof) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * dof) return;
  const int64_t p = i / dof;
  const int64_t c = i - p * dof;
  xt[dof * dst[p] + c] = static_cast<T>(x[dof * src[p] + c]);
}

__global__ __launch_bounds__(256) void k_gather_sel(const double* __restrict__ yfull,
                                                    double* __restrict__ y,
                                                    const int64_t* __restrict__ dst, int64_t n,
                                                    int dof) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * dof) return;
  const int64_t p = i / dof;
  y[i] = yfull[dof * dst[p] + (i - p * dof)];
}

// ---------------------------------------------------------------------------
// Many right-hand sides on FP32 MFMA.
//
// The FP32 slab layout serves as the A operand of v_mfma_f32_32x32x2f32
// without change: lanes 0-31 read rows 0-31 of column group g, lanes 32-63
// rows 0-31 of group g+1, 16 bytes (4 columns) each; element s of a lane's
// vector is A[row][k] of MFMA step s, with k = lane / 32 selecting the group.
// The B operand is the gathered input XB[z][n] (row-major, kRhs columns):
// lane l reads XB[z(g + l/32) + s][l % 32], 128 contiguous bytes per
// half-wavefront.  The 32 x 32 accumulator holds rows x right-hand sides;
// element j of lane l is row 8*(j/4) + 4*(l/32) + j%4, column l % 32.
constexpr int kRhs = 32;
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <typename T>
__global__ __launch_bounds__(256) void k_gather_b(const double* __restrict__ X, int64_t xr,
                                                  int64_t xc,
                                                  int nrhs, T* __restrict__ XB, int64_t n,
                                                  const int64_t* __restrict__ perm, int dof) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kRhs) + threadIdx.x / kRhs;
  const int c = threadIdx.x % kRhs;
  if (i >= n) return;
  int64_t src = i;
  if (perm) {
    const int64_t p = i / dof;
    src = dof * perm[p] + (i - p * dof);
  }
  XB[i * kRhs + c] = c < nrhs ? static_cast<T>(X[src * xr + c * xc]) : T(0);
}

// The same gather for a row-major X (unit column stride, even row stride,
// 16-byte aligned): eight threads per row, four columns each, two 16-byte
// loads and one 16-byte (FP32) or two (FP64) stores per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_b4(const double* __restrict__ X, int64_t xr,
                                                   int nrhs, T* __restrict__ XB, int64_t n,
                                                   const int64_t* __restrict__ perm, int dof) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 8) + threadIdx.x / 8;
  const int c0 = 4 * (threadIdx.x % 8);
  if (i >= n) return;
  int64_t src = i;
  if (perm) {
    const int64_t p = i / dof;
    src = dof * perm[p] + (i - p * dof);
  }
  const double* __restrict__ row = X + src * xr + c0;
  double v[4];
  if (c0 + 4 <= nrhs) {
    const dbl2 a = *reinterpret_cast<const dbl2*>(row);
    const dbl2 b = *reinterpret_cast<const dbl2*>(row + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = c0 + k < nrhs ? row[k] : 0.0;
  }
  T* dst = XB + i * kRhs + c0;
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<flt4*>(dst) = flt4{static_cast<float>(v[0]), static_cast<float>(v[1]),
                                         static_cast<float>(v[2]), static_cast<float>(v[3])};
  } else {
    *reinterpret_cast<dbl2*>(dst) = dbl2{v[0], v[1]};
    *reinterpret_cast<dbl2*>(dst + 2) = dbl2{v[2], v[3]};
  }
}

// KU group pairs per step; PIPE: the loads of step i+1 are issued before the
// MFMAs of step i.  Addresses are clamped into the work item, so every lane
// loads valid data; a group past the end gets B = 0 (its A rows then add
// nothing), and rows past R are never stored.  The mask is applied at the
// MFMA, not at the load, so no VALU op waits on a load of the next step.
// Output of the batched kernels: pass 1 writes t (row through rowmap, kRhs
// values per row), pass 2 writes Y (row through the point permutation, row
// stride yr, column stride yc).
// batched_rows maps N output rows of a work item (indices already clamped
// into range) to the rows they are stored at, with straight-line code per
// case, so that all index loads go out before the first one is waited for:
// pass 1 through rowmap, pass 2 through the point permutation.  Row indices
// and permutation entries are below 2^31 (checked when the layout is built),
// so the low word of each permutation entry is the whole value.
template <int PASS, int N>
__device__ __forceinline__ void batched_rows(const int (&idx)[N], const int32_t* rowmap,
                                             const int64_t* perm, int dof, int (&out)[N]) {
  if (PASS == 1) {
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = rowmap[idx[j]];
  } else if (perm) {
    const int32_t* plo = reinterpret_cast<const int32_t*>(perm);
    int pv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) pv[j] = plo[2 * (idx[j] / dof)];
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = dof * pv[j] + idx[j] % dof;
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] = idx[j];
  }
}

template <int PASS, typename T>
__device__ __forceinline__ void batched_put(T* tout, double* Y, int64_t yr, int64_t yc, int row,
                                            int n, T v) {
  if (PASS == 1)
    tout[static_cast<int64_t>(row) * kRhs + n] = v;
  else
    Y[row * yr + n * yc] = static_cast<double>(v);
}

template <int KU, bool PIPE, bool TWO>
__device__ __forceinline__ void mfma_stream(const flt4* __restrict__ base,
                                            const int32_t* __restrict__ gi,
                                            const float* __restrict__ XB, int ng, int R, int half,
                                            int r0, floatx16& acc0, floatx16& acc1) {
  const int rA0 = min(r0, R - 1), rA1 = min(r0 + 32, R - 1);
  auto load_z = [&](int g0, int (&z)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) z[u] = gi[min(g0 + 2 * u + half, ng - 1)];
  };
  auto load_ab = [&](int g0, const int (&z)[KU], flt4 (&a0)[KU], flt4 (&a1)[KU],
                     float (&b)[KU][4], bool (&ok)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int gg = g0 + 2 * u + half;
      ok[u] = gg < ng;
      const int ga = min(gg, ng - 1);
      a0[u] = __builtin_nontemporal_load(base + ga * R + rA0);
      if (TWO) a1[u] = __builtin_nontemporal_load(base + ga * R + rA1);
#pragma unroll
      for (int s = 0; s < 4; ++s) b[u][s] = XB[static_cast<int64_t>(z[u] + s) * kRhs + r0];
    }
  };
  auto compute = [&](const flt4 (&a0)[KU], const flt4 (&a1)[KU], const float (&b)[KU][4],
                     const bool (&ok)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float bv = ok[u] ? b[u][s] : 0.0f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[u][s], bv, acc0, 0, 0, 0);
        if (TWO) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[u][s], bv, acc1, 0, 0, 0);
      }
  };
  int z[KU];
  load_z(0, z);
  if (!PIPE) {
    for (int g0 = 0; g0 < ng; g0 += 2 * KU) {
      int zn[KU];
      load_z(g0 + 2 * KU, zn);
      flt4 a0[KU], a1[KU];
      float b[KU][4];
      bool ok[KU];
      load_ab(g0, z, a0, a1, b, ok);
      compute(a0, a1, b, ok);
#pragma unroll
      for (int u = 0; u < KU; ++u) z[u] = zn[u];
    }
  } else {
    // ping-pong buffers, two steps per trip, so no register copies sit
    // between the loads of one step and the MFMAs of the other
    flt4 pa0[KU], pa1[KU], qa0[KU], qa1[KU];
    float pb[KU][4], qb[KU][4];
    bool pok[KU], qok[KU];
    load_ab(0, z, pa0, pa1, pb, pok);
    load_z(2 * KU, z);  // offsets of step 1
    for (int g0 = 0; g0 < ng; g0 += 4 * KU) {
      int zn[KU];
      load_z(g0 + 4 * KU, zn);
      load_ab(g0 + 2 * KU, z, qa0, qa1, qb, qok);
      __builtin_amdgcn_sched_barrier(0);
      compute(pa0, pa1, pb, pok);
      if (g0 + 2 * KU >= ng) break;
      load_z(g0 + 6 * KU, z);
      load_ab(g0 + 4 * KU, zn, pa0, pa1, pb, pok);
      __builtin_amdgcn_sched_barrier(0);
      compute(qa0, qa1, qb, qok);
    }
  }
}

// Deep variant: D step buffers in registers used as a ring.  At stage j the
// loads of step s + D - 1 go out before the MFMAs of step s, so D - 1 steps
// of factor and input loads are in flight per wave while its MFMAs run (the
// two-buffer pipeline above keeps one).  Steps past the end load clamped,
// in-bounds addresses and skip their MFMAs (a wave-uniform branch).  zz[k]
// holds the input offsets of the next step to be loaded into buffer k.
template <int KU, int D, bool TWO>
__device__ __forceinline__ void mfma_stream_deep(const flt4* __restrict__ base,
                                                 const int32_t* __restrict__ gi,
                                                 const float* __restrict__ XB, int ng, int R,
                                                 int half, int r0, floatx16& acc0,
                                                 floatx16& acc1) {
  constexpr int G = 2 * KU;  // groups per step
  const int rA0 = min(r0, R - 1), rA1 = min(r0 + 32, R - 1);
  const int nsteps = (ng + G - 1) / G;
  flt4 a0[D][KU], a1[D][KU];
  float b[D][KU][4];
  int zz[D][KU];
  auto gz = [&](int step, int u) { return gi[min(step * G + 2 * u + half, ng - 1)]; };
  auto load = [&](int step, const int (&z)[KU], flt4 (&x0)[KU], flt4 (&x1)[KU],
                  float (&bb)[KU][4]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int ga = min(step * G + 2 * u + half, ng - 1);
      x0[u] = __builtin_nontemporal_load(base + ga * R + rA0);
      if (TWO) x1[u] = __builtin_nontemporal_load(base + ga * R + rA1);
#pragma unroll
      for (int s = 0; s < 4; ++s) bb[u][s] = XB[static_cast<int64_t>(z[u] + s) * kRhs + r0];
    }
  };
  auto compute = [&](int step, const flt4 (&x0)[KU], const flt4 (&x1)[KU],
                     const float (&bb)[KU][4]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const bool ok = step * G + 2 * u + half < ng;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float bv = ok ? bb[u][s] : 0.0f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(x0[u][s], bv, acc0, 0, 0, 0);
        if (TWO) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(x1[u][s], bv, acc1, 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k, u);
#pragma unroll
  for (int k = 0; k < D - 1; ++k) {
    load(k, zz[k], a0[k], a1[k], b[k]);
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k + D, u);
  }
  for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int s = s0 + j;
      const int k = (j + D - 1) % D;  // buffer freed by the previous stage
      load(s + D - 1, zz[k], a0[k], a1[k], b[k]);
#pragma unroll
      for (int u = 0; u < KU; ++u) zz[k][u] = gz(s + 2 * D - 1, u);
      __builtin_amdgcn_sched_barrier(0);
      if (s < nsteps) compute(s, a0[j], a1[j], b[j]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Ring variant with the loads issued in the order they are consumed: at
// stage j the MFMAs of buffer j run (its loads went out one trip, D stages,
// earlier), then buffer j is refilled with the step D ahead and the input
// offsets of the step after that.  The prologue fills the buffers in the same
// order as the loop, so every stage has the same D - 1 younger steps in
// flight on every path (the loop above refills a buffer before the MFMAs and
// its prologue order differs from the loop's, so the wait at the top of the
// loop covered only one step).  Loads and MFMAs are unconditional, so that
// the loop body is one basic block and its waits are the same on every
// path: past the last step the loads read clamped, in-bounds addresses and
// the MFMAs get B = 0 (as groups past the end do), adding nothing.
template <int KU, int D, bool TWO>
__device__ __forceinline__ void mfma_stream_ring(const flt4* __restrict__ base,
                                                 const int32_t* __restrict__ gi,
                                                 const float* __restrict__ XB, int ng, int R,
                                                 int half, int r0, floatx16& acc0,
                                                 floatx16& acc1) {
  constexpr int G = 2 * KU;  // groups per step
  const int rA0 = min(r0, R - 1), rA1 = min(r0 + 32, R - 1);
  const int nsteps = (ng + G - 1) / G;
  flt4 a0[D][KU], a1[D][KU];
  float b[D][KU][4];
  int zz[D][KU];  // input offsets of the step buffer k is refilled with next
  auto gz = [&](int step, int u) { return gi[min(step * G + 2 * u + half, ng - 1)]; };
  auto load = [&](int step, const int (&z)[KU], flt4 (&x0)[KU], flt4 (&x1)[KU],
                  float (&bb)[KU][4]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int ga = min(step * G + 2 * u + half, ng - 1);
      x0[u] = __builtin_nontemporal_load(base + ga * R + rA0);
      if (TWO) x1[u] = __builtin_nontemporal_load(base + ga * R + rA1);
      // 32-bit offsets (xt_size * kRhs < 2^32, checked by the batched calls):
      // no sign extension of the loaded offset sits between its load and here
      const uint32_t zo = static_cast<uint32_t>(z[u]) * kRhs + r0;
#pragma unroll
      for (int s = 0; s < 4; ++s) bb[u][s] = XB[zo + s * kRhs];
    }
  };
  auto compute = [&](int step, const flt4 (&x0)[KU], const flt4 (&x1)[KU],
                     const float (&bb)[KU][4]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const bool ok = step * G + 2 * u + half < ng;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float bv = ok ? bb[u][s] : 0.0f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(x0[u][s], bv, acc0, 0, 0, 0);
        if (TWO) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(x1[u][s], bv, acc1, 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k, u);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    __builtin_amdgcn_sched_barrier(0);
    load(k, zz[k], a0[k], a1[k], b[k]);
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k + D, u);
  }
  for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int s = s0 + j;
      __builtin_amdgcn_sched_barrier(0);
      compute(s, a0[j], a1[j], b[j]);  // past the last step: B = 0, adds nothing
      __builtin_amdgcn_sched_barrier(0);
      load(s + D, zz[j], a0[j], a1[j], b[j]);
#pragma unroll
      for (int u = 0; u < KU; ++u) zz[j][u] = gz(s + 2 * D, u);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

// The load-order ring with the B operands staged through LDS: a column group's
// four input rows (4 x kRhs floats, 512 contiguous bytes) arrive as one
// 16-byte load per lane of its half-wave instead of four 4-byte loads per
// lane, are written to the wave's LDS slot and read back in the MFMA layout
// (row s, column lane % 32).  Steps go: MFMAs of step s (B in registers, read
// from LDS the stage before), then step s + 1's B written to LDS and read
// back, then the refill of step s + D.  One wave per workgroup, so the slot
// is the wave's own; its LDS operations run in order, so a read sees the
// write before it.
template <int KU, int D, bool TWO>
__device__ __forceinline__ void mfma_stream_lds(const flt4* __restrict__ base,
                                                const int32_t* __restrict__ gi,
                                                const float* __restrict__ XB, int ng, int R,
                                                int half, int r0, float* slot, floatx16& acc0,
                                                floatx16& acc1) {
  constexpr int G = 2 * KU;  // groups per step
  const int rA0 = min(r0, R - 1), rA1 = min(r0 + 32, R - 1);
  const int nsteps = (ng + G - 1) / G;
  flt4 a0[D][KU], a1[D][KU];
  flt4 bq[D][KU];  // this lane's 16 bytes of the four input rows of group 2u + half
  int zz[D][KU];    // input offset of group 2u + half for the step buffer k gets next
  float bn[KU][4];  // B of the next step to multiply, read back from LDS
  auto gz = [&](int step, int u) { return gi[min(step * G + 2 * u + half, ng - 1)]; };
  auto load = [&](int step, const int (&z)[KU], flt4 (&x0)[KU], flt4 (&x1)[KU], flt4 (&bb)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int ga = min(step * G + 2 * u + half, ng - 1);
      x0[u] = __builtin_nontemporal_load(base + ga * R + rA0);
      if (TWO) x1[u] = __builtin_nontemporal_load(base + ga * R + rA1);
      // rows z .. z + 3 of XB are 512 contiguous, 128-byte aligned bytes
      const uint32_t zo = static_cast<uint32_t>(z[u]) * kRhs + 4 * r0;
      bb[u] = *reinterpret_cast<const flt4*>(XB + zo);
    }
  };
  auto stage_b = [&](const flt4 (&bb)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u)
      *reinterpret_cast<flt4*>(slot + (2 * u + half) * 4 * kRhs + 4 * r0) = bb[u];
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s) bn[u][s] = slot[(2 * u + half) * 4 * kRhs + s * kRhs + r0];
  };
  auto compute = [&](int step, const flt4 (&x0)[KU], const flt4 (&x1)[KU]) {
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const bool ok = step * G + 2 * u + half < ng;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float bv = ok ? bn[u][s] : 0.0f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(x0[u][s], bv, acc0, 0, 0, 0);
        if (TWO) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(x1[u][s], bv, acc1, 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k, u);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    __builtin_amdgcn_sched_barrier(0);
    load(k, zz[k], a0[k], a1[k], bq[k]);
#pragma unroll
    for (int u = 0; u < KU; ++u) zz[k][u] = gz(k + D, u);
  }
  __builtin_amdgcn_sched_barrier(0);
  stage_b(bq[0]);
  for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int s = s0 + j;
      __builtin_amdgcn_sched_barrier(0);
      compute(s, a0[j], a1[j]);  // past the last step: B = 0, adds nothing
      __builtin_amdgcn_sched_barrier(0);
      stage_b(bq[(j + 1) % D]);
      __builtin_amdgcn_sched_barrier(0);
      load(s + D, zz[j], a0[j], a1[j], bq[j]);
#pragma unroll
      for (int u = 0; u < KU; ++u) zz[j][u] = gz(s + 2 * D, u);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int PASS, int KU, bool PIPE, int D = 0, bool RING = false, bool LDSB = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_slab_mfma(
    const bwl::WorkItem* __restrict__ wis, XcdSplit xs, const int32_t* __restrict__ gidx,
    const float* __restrict__ data, const float* __restrict__ XB,
    float* __restrict__ tout,              // pass 1
    double* __restrict__ Y, int64_t yr, int64_t yc,  // pass 2 (row, column strides)
    int nrhs, float* __restrict__ part, const int64_t* __restrict__ perm, int dof,
    const int32_t* __restrict__ rowmap) {
  const int lane = threadIdx.x & 63;
  const int w = xcd_work_item<1>(xs);
  if (w < 0) return;
  const bwl::WorkItem wi = wis[w];
  const int R = __builtin_amdgcn_readfirstlane(wi.rows);
  const int ng = __builtin_amdgcn_readfirstlane(wi.ngroups);
  const int half = lane >> 5;  // which group of the pair
  const int r0 = lane & 31;
  const flt4* __restrict__ base = reinterpret_cast<const flt4*>(data + wi.data_off);
  const int32_t* __restrict__ gi = gidx + __builtin_amdgcn_readfirstlane(wi.gbeg);
  floatx16 acc0 = {0}, acc1 = {0};
  const bool two = R > 32;
  if (ng > 0) {
    if constexpr (LDSB) {
      __shared__ __align__(16) float slot[2 * KU * 4 * kRhs];
      if (two)
        mfma_stream_lds<KU, D, true>(base, gi, XB, ng, R, half, r0, slot, acc0, acc1);
      else
        mfma_stream_lds<KU, D, false>(base, gi, XB, ng, R, half, r0, slot, acc0, acc1);
    } else if constexpr (RING) {
      if (two)
        mfma_stream_ring<KU, D, true>(base, gi, XB, ng, R, half, r0, acc0, acc1);
      else
        mfma_stream_ring<KU, D, false>(base, gi, XB, ng, R, half, r0, acc0, acc1);
    } else if constexpr (D > 0) {
      if (two)
        mfma_stream_deep<KU, D, true>(base, gi, XB, ng, R, half, r0, acc0, acc1);
      else
        mfma_stream_deep<KU, D, false>(base, gi, XB, ng, R, half, r0, acc0, acc1);
    } else {
      if (two)
        mfma_stream<KU, PIPE, true>(base, gi, XB, ng, R, half, r0, acc0, acc1);
      else
        mfma_stream<KU, PIPE, false>(base, gi, XB, ng, R, half, r0, acc0, acc1);
    }
  }
  const int n = lane & 31;
  auto row_of = [&](int j) { return 8 * (j >> 2) + 4 * half + (j & 3); };
  if (wi.part_off < 0) {
    // 8 output rows at a time: their index loads (clamped rows, all in
    // bounds, none behind a branch) go out together, then the stores, instead
    // of one round trip per stored value
    auto put8 = [&](int r0, int j0, const floatx16& acc) {
      int idx[8], orow[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) idx[j] = wi.out_row + min(r0 + row_of(j0 + j), R - 1);
      batched_rows<PASS>(idx, rowmap, perm, dof, orow);
      if (n >= nrhs) return;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (r0 + row_of(j0 + j) < R) batched_put<PASS>(tout, Y, yr, yc, orow[j], n, acc[j0 + j]);
    };
    put8(0, 0, acc0);
    put8(0, 8, acc0);
    if (two) {
      put8(32, 0, acc1);
      put8(32, 8, acc1);
    }
    return;
  }
  // partial sums of split slabs are added in order by k_reduce_b; they are
  // read once, by it, so they are written (and read) nontemporally: about 1 %
  // of a 32-RHS call on a 150 528-DoF operator, same values
  float* pp = part + static_cast<int64_t>(wi.part_off) * kRhs;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = row_of(j);
    if (row < R) __builtin_nontemporal_store(acc0[j], &pp[row * kRhs + n]);
    if (two && 32 + row < R) __builtin_nontemporal_store(acc1[j], &pp[(32 + row) * kRhs + n]);
  }
}

// The same product on v_mfma_f32_16x16x4f32 with 16-row tiles, so a slab of
// R rows costs ceil(R / 16) tiles instead of 32 or 64 rows (pass-1 slabs of
// 33-48 rows carry a third of their MFMA work as padding on 32-row tiles).
// One k-step covers four column groups: lane l supplies A[i][k] with
// i = l % 16 (row within the tile) and k = l / 16, i.e. element s of row i of
// group 4 step + k is MFMA s's operand, and B[k][n] = XB[z(group) + s][n]
// for n = l % 16 (+ 16 for the second 16 right-hand sides).  The same
// 16-byte slab vectors are read as in the 32 x 32 kernel.  Loads run through
// a ring of D step buffers in load order, as mfma_stream_ring; every tile's
// rows are loaded (clamped, in bounds) and tiles past ceil(R / 16) skip their
// MFMAs.  Accumulator element r of lane l is row 4 (l / 16) + r, column
// l % 16 of its tile.
typedef float floatx4 __attribute__((ext_vector_type(4)));

// The k-steps of k_slab_mfma16 for a slab of NT 16-row tiles (NT a template
// argument, so the loop body is one basic block with the same waits on
// every path).
template <int NT, int D>
__device__ __forceinline__ void mfma16_stream(const flt4* __restrict__ base,
                                              const int32_t* __restrict__ gi,
                                              const float* __restrict__ XB, int ng, int R, int kq,
                                              int i16, floatx4 (&acc)[4][2]) {
  int rowc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) rowc[t] = min(16 * t + i16, R - 1);
  const int nsteps = (ng + 3) >> 2;  // four groups per k-step
  flt4 a[D][NT];
  float b[D][2][4];
  int zz[D];  // input offset of the step buffer q is refilled with next
  auto gz = [&](int step) { return gi[min(4 * step + kq, ng - 1)]; };
  auto load = [&](int step, int z, flt4 (&aa)[NT], float (&bb)[2][4]) {
    const int ga = min(4 * step + kq, ng - 1);
#pragma unroll
    for (int t = 0; t < NT; ++t) aa[t] = __builtin_nontemporal_load(base + ga * R + rowc[t]);
    const uint32_t zo = static_cast<uint32_t>(z) * kRhs + i16;  // < 2^32, checked
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bb[0][s] = XB[zo + s * kRhs];
      bb[1][s] = XB[zo + s * kRhs + 16];
    }
  };
  auto compute = [&](int step, const flt4 (&aa)[NT], const float (&bb)[2][4]) {
    const bool ok = 4 * step + kq < ng;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const float b0 = ok ? bb[0][s] : 0.0f, b1 = ok ? bb[1][s] : 0.0f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[t][s], b0, acc[t][0], 0, 0, 0);
        acc[t][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[t][s], b1, acc[t][1], 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int q = 0; q < D; ++q) zz[q] = gz(q);
#pragma unroll
  for (int q = 0; q < D; ++q) {
    __builtin_amdgcn_sched_barrier(0);
    load(q, zz[q], a[q], b[q]);
    zz[q] = gz(q + D);
  }
  for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const int st = s0 + j;
      __builtin_amdgcn_sched_barrier(0);
      compute(st, a[j], b[j]);  // past the last step: B = 0, adds nothing
      __builtin_amdgcn_sched_barrier(0);
      load(st + D, zz[j], a[j], b[j]);
      zz[j] = gz(st + 2 * D);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int PASS, int D>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(D > 3 ? 3 : 4))) void k_slab_mfma16(
    const bwl::WorkItem* __restrict__ wis, XcdSplit xs, const int32_t* __restrict__ gidx,
    const float* __restrict__ data, const float* __restrict__ XB,
    float* __restrict__ tout,              // pass 1
    double* __restrict__ Y, int64_t yr, int64_t yc,  // pass 2 (row, column strides)
    int nrhs, float* __restrict__ part, const int64_t* __restrict__ perm, int dof,
    const int32_t* __restrict__ rowmap) {
  const int lane = threadIdx.x & 63;
  const int w = xcd_work_item<1>(xs);
  if (w < 0) return;
  const bwl::WorkItem wi = wis[w];
  const int R = wi.rows;
  const int ng = wi.ngroups;
  const int kq = lane >> 4, i16 = lane & 15;
  const flt4* __restrict__ base = reinterpret_cast<const flt4*>(data + wi.data_off);
  const int32_t* __restrict__ gi = gidx + wi.gbeg;
  const int ntile = (R + 15) >> 4;  // wave-uniform
  floatx4 acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  if (ng > 0) {
    switch (ntile) {
      case 1: mfma16_stream<1, D>(base, gi, XB, ng, R, kq, i16, acc); break;
      case 2: mfma16_stream<2, D>(base, gi, XB, ng, R, kq, i16, acc); break;
      case 3: mfma16_stream<3, D>(base, gi, XB, ng, R, kq, i16, acc); break;
      default: mfma16_stream<4, D>(base, gi, XB, ng, R, kq, i16, acc);
    }
  }
  if (wi.part_off >= 0) {
    float* pp = part + static_cast<int64_t>(wi.part_off) * kRhs;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= ntile) break;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * t + 4 * kq + r;
          if (row < R)  // added in order by k_reduce_b
            __builtin_nontemporal_store(acc[t][h][r], &pp[row * kRhs + 16 * h + i16]);
        }
    }
    return;
  }
  // two tiles (8 of this lane's rows) at a time: their output rows first
  // (clamped, in bounds, no branch), so the index loads go out together
#pragma unroll
  for (int t0 = 0; t0 < 4; t0 += 2) {
    if (t0 >= ntile) break;
    int idx[8], orow[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) idx[j] = wi.out_row + min(16 * (t0 + (j >> 2)) + 4 * kq + (j & 3), R - 1);
    batched_rows<PASS>(idx, rowmap, perm, dof, orow);
#pragma unroll
    for (int t = t0; t < t0 + 2; ++t)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * t + 4 * kq + r;
          const int n = 16 * h + i16;
          if (row < R && n < nrhs)
            batched_put<PASS>(tout, Y, yr, yc, orow[4 * (t - t0) + r], n, acc[t][h][r]);
        }
  }
}

template <int PASS>
void launch_mfma(int schedule, int blocks, hipStream_t st, const bwl::WorkItem* wis, XcdSplit xs,
                 const int32_t* gidx, const float* data, const float* XB, float* tout, double* Y,
                 int64_t yr, int64_t yc, int nrhs, float* part, const int64_t* perm, int dof,
                 const int32_t* rowmap) {
  switch (schedule) {
    case 1:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 2, true>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 2:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 4, true>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 3:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 8, false>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 4:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 2, true, 3>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 5:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 2, true, 4>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 6:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 1, true, 6>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 7:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 1, true, 8>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
#define BW_MFMA32_RING(KU, D)                                                              \
  hipLaunchKernelGGL((k_slab_mfma<PASS, KU, true, D, true>), dim3(4 * blocks), dim3(64), 0, st, \
                     wis, xs, gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap)
    case 8: BW_MFMA32_RING(2, 3); break;
    case 9: BW_MFMA32_RING(2, 4); break;
    case 10: BW_MFMA32_RING(1, 6); break;
    case 11: BW_MFMA32_RING(1, 8); break;
#undef BW_MFMA32_RING
#define BW_MFMA32_LDSB(KU, D)                                                                    \
  hipLaunchKernelGGL((k_slab_mfma<PASS, KU, true, D, true, true>), dim3(4 * blocks), dim3(64), 0, \
                     st, wis, xs, gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap)
    case 14: BW_MFMA32_LDSB(2, 3); break;
    case 15: BW_MFMA32_LDSB(2, 4); break;
#undef BW_MFMA32_LDSB
    case 12:
      hipLaunchKernelGGL((k_slab_mfma16<PASS, 3>), dim3(4 * blocks), dim3(64), 0, st, wis, xs, gidx,
                         data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    case 13:
      hipLaunchKernelGGL((k_slab_mfma16<PASS, 4>), dim3(4 * blocks), dim3(64), 0, st, wis, xs, gidx,
                         data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
      break;
    default:
      hipLaunchKernelGGL((k_slab_mfma<PASS, 4, false>), dim3(4 * blocks), dim3(64), 0, st, wis, xs,
                         gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap);
  }
}

template <int PASS, typename T>
__global__ __launch_bounds__(256) void k_reduce_b(const bwl::Reduce* __restrict__ reds, int nred,
                                                  const T* __restrict__ part, T* tout,
                                                  double* Y, int64_t yr, int64_t yc, int nrhs,
                                                  const int64_t* __restrict__ perm, int dof,
                                                  const int32_t* __restrict__ rowmap) {
  // one workgroup per split slab; a thread sums one 16-byte vector of
  // right-hand sides of one row over the parts, in part order, four parts in
  // flight (nontemporal loads: each partial row is read once)
  typedef typename Vec<T>::type VT;
  constexpr int VW = Vec<T>::V;
  const int r = blockIdx.x;
  if (r >= nred) return;
  const bwl::Reduce rd = reds[r];
  const VT* __restrict__ base_p =
      reinterpret_cast<const VT*>(part + static_cast<int64_t>(rd.part_off) * kRhs);
  const int nvec = rd.rows * (kRhs / VW);
  const int64_t stride = static_cast<int64_t>(rd.rows) * (kRhs / VW);
  for (int e = threadIdx.x; e < nvec; e += blockDim.x) {
    VT sum = {};
    int q = 0;
    for (; q + 4 <= rd.nparts; q += 4) {
      const VT a = __builtin_nontemporal_load(&base_p[(q + 0) * stride + e]);
      const VT b = __builtin_nontemporal_load(&base_p[(q + 1) * stride + e]);
      const VT c = __builtin_nontemporal_load(&base_p[(q + 2) * stride + e]);
      const VT d = __builtin_nontemporal_load(&base_p[(q + 3) * stride + e]);
      sum += a;
      sum += b;
      sum += c;
      sum += d;
    }
    for (; q < rd.nparts; ++q) sum += __builtin_nontemporal_load(&base_p[q * stride + e]);
    const int row = e / (kRhs / VW);
    const int n0 = (e % (kRhs / VW)) * VW;
    if (PASS == 1) {
      *reinterpret_cast<VT*>(tout + static_cast<int64_t>(rowmap[rd.out_row + row]) * kRhs + n0) =
          sum;
    } else {
      const int64_t o = out_index<double>(static_cast<int64_t>(rd.out_row) + row, perm, dof);
#pragma unroll
      for (int k = 0; k < VW; ++k)
        if (n0 + k < nrhs) Y[o * yr + (n0 + k) * yc] = static_cast<double>(sum[k]);
    }
  }
}

// ---------------------------------------------------------------------------
// Many right-hand sides in FP64 on MFMA (v_mfma_f64_16x16x4f64), on the FP64
// slab layout data[g][row][v] (V = 2).  One k-step covers two column groups
// (4 columns): lane l supplies A[i][k] with i = l % 16 (row within a 16-row
// tile) and k = l / 16, i.e. column k % 2 of group g0 + k / 2, and
// B[k][n] = XB[z(g0 + k / 2) + k % 2][n] with n = l % 16 (+ 16 for the second
// column tile).  Element r of a lane's accumulator is row l / 16 + 4 r,
// column l % 16 (measured on gfx950 with one-hot operands).  Rows past R read a clamped row and are never stored; groups
// past the end get B = 0.
typedef double dblx4 __attribute__((ext_vector_type(4)));

// D > 0: the k-steps run through a ring of D register buffers (as in
// mfma_stream_deep), D - 1 steps of loads in flight; D = 0: two steps loaded,
// then their MFMAs.  Both issue the same MFMAs in the same order per
// accumulator, so the results agree bit for bit.
template <int PASS, int D, bool RING = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(D > 4 ? 2 : (D > 0 ? 3 : 4)))) void k_slab_mfma64(
    const bwl::WorkItem* __restrict__ wis, XcdSplit xs, const int32_t* __restrict__ gidx,
    const double* __restrict__ data, const double* __restrict__ XB,
    double* __restrict__ tout,             // pass 1
    double* __restrict__ Y, int64_t yr, int64_t yc,  // pass 2 (row, column strides)
    int nrhs, double* __restrict__ part, const int64_t* __restrict__ perm, int dof,
    const int32_t* __restrict__ rowmap) {
  const int lane = threadIdx.x & 63;
  const int w = xcd_work_item<1>(xs);
  if (w < 0) return;
  const bwl::WorkItem wi = wis[w];
  const int R = wi.rows;
  const int ng = wi.ngroups;
  const int k = lane >> 4, i16 = lane & 15;
  const int gsub = k >> 1, v = k & 1;
  const double* __restrict__ base = data + wi.data_off;
  const int32_t* __restrict__ gi = gidx + wi.gbeg;
  const int ntile = (R + 15) >> 4;  // 16-row tiles in use, wave-uniform
  dblx4 acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = dblx4{0.0, 0.0, 0.0, 0.0};
  int rowc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) rowc[t] = min(16 * t + i16, R - 1);
  if constexpr (D > 0) {
    const int nsteps = (ng + 1) >> 1;  // two groups per k-step
    double a[D][4], b[D][2];
    int zz[D];  // input offset of the next step to be loaded into each buffer
    auto gz = [&](int step) { return gi[min(2 * step + gsub, ng - 1)]; };
    auto load = [&](int step, int z, double (&aa)[4], double (&bb)[2]) {
      const int ga = min(2 * step + gsub, ng - 1);
      const double* __restrict__ gp = base + static_cast<int64_t>(ga) * R * 2 + v;
#pragma unroll
      // every tile loads (rows clamped into the slab, in bounds) so that the
      // loads, and hence the waits, are the same on every path; tiles past
      // ntile are never multiplied
      for (int t = 0; t < 4; ++t) aa[t] = gp[rowc[t] * 2];
      const uint32_t zr = static_cast<uint32_t>(z + v) * kRhs + i16;  // < 2^32, checked
      bb[0] = XB[zr];
      bb[1] = XB[zr + 16];
    };
    auto compute = [&](int step, const double (&aa)[4], const double (&bb)[2]) {
      const bool ok = 2 * step + gsub < ng;
      const double b0 = ok ? bb[0] : 0.0, b1 = ok ? bb[1] : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < ntile) {
          acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(aa[t], b0, acc[t][0], 0, 0, 0);
          acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(aa[t], b1, acc[t][1], 0, 0, 0);
        }
    };
#pragma unroll
    for (int q = 0; q < D; ++q) zz[q] = gz(q);
    if constexpr (RING) {
      // buffers consumed in the order they were loaded (as mfma_stream_ring):
      // compute buffer j, then refill it with the step D ahead
#pragma unroll
      for (int q = 0; q < D; ++q) {
        __builtin_amdgcn_sched_barrier(0);
        load(q, zz[q], a[q], b[q]);
        zz[q] = gz(q + D);
      }
      for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
          const int st = s0 + j;
          __builtin_amdgcn_sched_barrier(0);
          compute(st, a[j], b[j]);  // past the last step: B = 0, adds nothing
          __builtin_amdgcn_sched_barrier(0);
          load(st + D, zz[j], a[j], b[j]);
          zz[j] = gz(st + 2 * D);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    } else {
#pragma unroll
    for (int q = 0; q < D - 1; ++q) {
      load(q, zz[q], a[q], b[q]);
      zz[q] = gz(q + D);
    }
    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const int st = s0 + j;
        const int q = (j + D - 1) % D;  // buffer freed by the previous stage
        load(st + D - 1, zz[q], a[q], b[q]);
        zz[q] = gz(st + 2 * D - 1);
        __builtin_amdgcn_sched_barrier(0);
        if (st < nsteps) compute(st, a[j], b[j]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    }
  } else {
  constexpr int S = 2;  // k-steps (two groups each) loaded before their MFMAs
  for (int g0 = 0; g0 < ng; g0 += 2 * S) {
    double a[S][4], b[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int gg = g0 + 2 * s + gsub;
      const bool ok = gg < ng;
      const int ga = min(gg, ng - 1);
      const double* __restrict__ gp = base + static_cast<int64_t>(ga) * R * 2 + v;
#pragma unroll
      for (int t = 0; t < 4; ++t) a[s][t] = t < ntile ? gp[rowc[t] * 2] : 0.0;
      const int64_t zr = static_cast<int64_t>(gi[ga] + v) * kRhs;
      const double b0 = XB[zr + i16], b1 = XB[zr + 16 + i16];
      b[s][0] = ok ? b0 : 0.0;
      b[s][1] = ok ? b1 : 0.0;
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < ntile) {
          acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][t], b[s][0], acc[t][0], 0, 0, 0);
          acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][t], b[s][1], acc[t][1], 0, 0, 0);
        }
  }
  }
  if (wi.part_off >= 0) {
    double* pp = part + static_cast<int64_t>(wi.part_off) * kRhs;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= ntile) break;
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * t + k + 4 * r;
          if (row < R)  // added in order by k_reduce_b
            __builtin_nontemporal_store(acc[t][c][r], &pp[row * kRhs + 16 * c + i16]);
        }
    }
    return;
  }
  // two tiles (8 of this lane's rows) at a time: their output rows first
  // (clamped, in bounds, no branch), so the index loads go out together, then
  // the stores
#pragma unroll
  for (int t0 = 0; t0 < 4; t0 += 2) {
    if (t0 >= ntile) break;
    int idx[8], orow[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) idx[j] = wi.out_row + min(16 * (t0 + (j >> 2)) + k + 4 * (j & 3), R - 1);
    batched_rows<PASS>(idx, rowmap, perm, dof, orow);
#pragma unroll
    for (int t = t0; t < t0 + 2; ++t)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * t + k + 4 * r;
          const int n = 16 * c + i16;
          if (row < R && n < nrhs)
            batched_put<PASS>(tout, Y, yr, yc, orow[4 * (t - t0) + r], n, acc[t][c][r]);
        }
  }
}

template <int PASS>
void launch_mfma64(int schedule, int blocks, hipStream_t st, const bwl::WorkItem* wis, XcdSplit xs,
                   const int32_t* gidx, const double* data, const double* XB, double* tout,
                   double* Y, int64_t yr, int64_t yc, int nrhs, double* part,
                   const int64_t* perm, int dof, const int32_t* rowmap) {
#define BW_MFMA64(D)                                                                        \
  hipLaunchKernelGGL((k_slab_mfma64<PASS, D>), dim3(4 * blocks), dim3(64), 0, st, wis, xs, gidx, \
                     data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap)
#define BW_MFMA64R(D)                                                                         \
  hipLaunchKernelGGL((k_slab_mfma64<PASS, D, true>), dim3(4 * blocks), dim3(64), 0, st, wis, xs, \
                     gidx, data, XB, tout, Y, yr, yc, nrhs, part, perm, dof, rowmap)
  switch (schedule) {
    case 4: BW_MFMA64(3); break;
    case 5: BW_MFMA64(4); break;
    case 6: BW_MFMA64(6); break;
    case 7: BW_MFMA64(8); break;
    case 8: BW_MFMA64R(3); break;
    case 9: BW_MFMA64R(4); break;
    case 10: BW_MFMA64R(6); break;
    case 11: BW_MFMA64R(8); break;
    case 12: BW_MFMA64R(3); break;  // the FP64 kernel has 16-row tiles already
    case 13: BW_MFMA64R(4); break;
    case 14: BW_MFMA64R(3); break;  // FP32-only B staging: the FP64 ring
    case 15: BW_MFMA64R(4); break;
    default: BW_MFMA64(0);
  }
#undef BW_MFMA64
#undef BW_MFMA64R
}

// ---------------------------------------------------------------------------

struct DevPass {
  bwl::WorkItem* wis = nullptr;
  int32_t* gidx = nullptr;
  int32_t* rowmap = nullptr;        // pass 1: t entry of every slab row
  bwl::Reduce* reds = nullptr;      // one per split slab
  void* data = nullptr;
  int nwi = 0, nred = 0;
  XcdSplit xs = {};
  int blocks = 0;  // workgroups of a slab launch
  int64_t data_bytes = 0;
};

// The read-only device layout of one operator (H or H^T) at one precision.
// Everything a matvec writes lives in an OpScratch instead, one per stream.
struct DevOp {
  bool present = false;
  int elem = 8;  // bytes
  DevPass p1, p2;
  int64_t xt_size = 0;
  int64_t xt_t = 0;     // first t entry in xt
  int64_t t_total = 0;  // t entries
  int64_t n_in = 0, n_out = 0;
  const int64_t* perm_in = nullptr;   // device, n_in / dof
  const int64_t* perm_out = nullptr;  // device, n_out / dof
  int64_t stream_bytes = 0;
  int64_t part_elems[2] = {0, 0};
  int unroll[2] = {16, 16};   // loads in flight per lane in k_slab (4, 8 or 16), per pass
  int mfma_schedule = 0;      // k_slab_mfma / k_slab_mfma64 variant (0-15)
};

// What one matvec of an operator writes: xt (x gathered in the factor
// precision, then t = B x), the partial sums of split work items of both
// passes, and the same for the batched products (xt_size x kRhs input and
// rank rows, partial sums x kRhs).  xt and xb are zero where no kernel writes
// (the gap after x and the tail after t, which padded column groups read).
struct OpScratch {
  void* xt = nullptr;
  void* part[2] = {nullptr, nullptr};
  void* xb = nullptr;
  void* partb[2] = {nullptr, nullptr};
};

// The scratch of one stream (or of one captured graph): per operator, plus
// the staging vectors of host-pointer calls and the full permuted output of a
// selection.  Calls on different streams use different StreamCtx, so they
// may run concurrently on one handle.
struct StreamCtx {
  hipStream_t stream = nullptr;
  OpScratch op[2][2];  // [adjoint][fp32]
  double* x_stage = nullptr;
  double* y_stage = nullptr;
  double* y_full = nullptr;
  std::vector<void*> allocs;
};

// Allocation-failure hook for tests of the cleanup paths: with the
// environment variable BWHIP_FAIL_ALLOC_AT=k (k >= 1), the k-th device or
// pinned-host allocation made by one C-ABI call fails as if the memory were
// exhausted.  The count restarts at every call that allocates.
thread_local int64_t g_alloc_seq = 0;
thread_local int64_t g_alloc_fail_at = 0;

void arm_alloc_hook() {
  g_alloc_seq = 0;
  const char* e = std::getenv("BWHIP_FAIL_ALLOC_AT");
  g_alloc_fail_at = e ? std::atoll(e) : 0;
}

bool alloc_hook_fires() { return g_alloc_fail_at > 0 && ++g_alloc_seq == g_alloc_fail_at; }

// hipMalloc of at least 16 bytes, recorded in `allocs`; BW_ENOMEM on failure.
int dev_alloc(void** p, size_t bytes, std::vector<void*>* allocs, const char* what) {
  *p = nullptr;
  if (alloc_hook_fires())
    return fail(BW_ENOMEM, std::string(what) + ": hipMalloc: out of memory (BWHIP_FAIL_ALLOC_AT)");
  const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BW_ENOMEM, std::string(what) + ": hipMalloc: " + hipGetErrorString(e));
  }
  allocs->push_back(*p);
  return BW_OK;
}

int host_pinned_alloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (alloc_hook_fires())
    return fail(BW_ENOMEM, std::string(what) + ": hipHostMalloc: out of memory (BWHIP_FAIL_ALLOC_AT)");
  if (hipHostMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    return fail(BW_ENOMEM, std::string(what) + ": hipHostMalloc failed");
  }
  return BW_OK;
}

// Work-item bytes for the next create; 0 = sized from the matrix (auto_wi_bytes).
std::atomic<int64_t> g_wi_bytes64{0};
std::atomic<int64_t> g_wi_bytes32{0};
std::atomic<int> g_stack_rows{64};
// Relative tolerance of the FP32 copy's low-rank recompression (0: off).
std::atomic<double> g_fp32_rank_tol{0.0};

// About 32k work items: a few rounds of 8 waves on each of the 1024 SIMDs,
// capped where a sweep of the 600k-DoF 3DT0 matrix on MI355X stopped gaining
// (384 KiB for FP64, 192 KiB for FP32) and floored at 8 KiB for small matrices.
int64_t auto_wi_bytes(int64_t entries, int elem_bytes) {
  const int64_t cap = elem_bytes == 8 ? (384 << 10) : (192 << 10);
  return std::min<int64_t>(cap, std::max<int64_t>(8 << 10, entries * elem_bytes / 32768));
}

template <typename T>
int upload_pass(const bwl::Layout& L, const bwl::Program& P, DevPass* dp, hipStream_t st,
                std::vector<void*>* allocs) {
  dp->nwi = static_cast<int>(P.wis.size());
  for (int x = 0; x <= bwl::kXcds; ++x) dp->xs.s[x] = P.xcd[x];
  dp->blocks = P.xcd_blocks;
  dp->nred = static_cast<int>(P.reds.size());
  auto dalloc = [&](void** p, size_t bytes) -> int { return dev_alloc(p, bytes, allocs, "create"); };
  int rc;
  if ((rc = dalloc(reinterpret_cast<void**>(&dp->wis), P.wis.size() * sizeof(bwl::WorkItem))))
    return rc;
  if ((rc = dalloc(reinterpret_cast<void**>(&dp->gidx), P.gidx.size() * sizeof(int32_t)))) return rc;
  if (!P.rowmap.empty()) {
    if ((rc = dalloc(reinterpret_cast<void**>(&dp->rowmap), P.rowmap.size() * sizeof(int32_t))))
      return rc;
    HIPCHK(hipMemcpy(dp->rowmap, P.rowmap.data(), P.rowmap.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
  }
  if ((rc = dalloc(reinterpret_cast<void**>(&dp->reds), P.reds.size() * sizeof(bwl::Reduce))))
    return rc;
  dp->data_bytes = P.data_elems * static_cast<int64_t>(sizeof(T));
  if ((rc = dalloc(&dp->data, static_cast<size_t>(dp->data_bytes)))) return rc;
  if (!P.wis.empty())
    HIPCHK(hipMemcpy(dp->wis, P.wis.data(), P.wis.size() * sizeof(bwl::WorkItem), hipMemcpyHostToDevice));
  if (!P.gidx.empty())
    HIPCHK(hipMemcpy(dp->gidx, P.gidx.data(), P.gidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (!P.reds.empty())
    HIPCHK(hipMemcpy(dp->reds, P.reds.data(), P.reds.size() * sizeof(bwl::Reduce), hipMemcpyHostToDevice));
  // data: fill slabs into pinned staging chunks and stream them up
  const int32_t nslab = static_cast<int32_t>(P.slabs.size());
  if (nslab == 0) return BW_OK;
  int64_t max_slab = 0;
  for (const auto& s : P.slabs) max_slab = std::max<int64_t>(max_slab, s.ngroups * s.rows * L.opt.vec_elems);
  const int64_t chunk_elems = std::max<int64_t>(max_slab, (int64_t(256) << 20) / static_cast<int64_t>(sizeof(T)));
  T* stage[2] = {nullptr, nullptr};
  if ((rc = host_pinned_alloc(reinterpret_cast<void**>(&stage[0]), chunk_elems * sizeof(T),
                              "create: upload staging")))
    return rc;
  if ((rc = host_pinned_alloc(reinterpret_cast<void**>(&stage[1]), chunk_elems * sizeof(T),
                              "create: upload staging"))) {
    (void)hipHostFree(stage[0]);
    return rc;
  }
  hipEvent_t done[2] = {nullptr, nullptr};
  if (hipEventCreate(&done[0]) != hipSuccess || hipEventCreate(&done[1]) != hipSuccess) {
    if (done[0]) (void)hipEventDestroy(done[0]);
    (void)hipHostFree(stage[0]);
    (void)hipHostFree(stage[1]);
    return fail(BW_EHIP, "upload: hipEventCreate");
  }
  const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
  const int nthreads = static_cast<int>(std::min(hc, 32u));
  int buf = 0;
  bool pending[2] = {false, false};
  int32_t s = 0;
  int status = BW_OK;
  while (s < nslab) {
    int32_t e = s;
    int64_t elems = 0;
    while (e < nslab) {
      const int64_t se = P.slabs[e].ngroups * P.slabs[e].rows * L.opt.vec_elems;
      if (elems + se > chunk_elems && e > s) break;
      elems += se;
      ++e;
    }
    if (pending[buf]) {
      // the copy out of this staging buffer must be complete before refilling it
      const hipError_t we = hipEventSynchronize(done[buf]);
      pending[buf] = false;
      if (we != hipSuccess) {
        status = fail(BW_EHIP, std::string("upload: ") + hipGetErrorString(we));
        break;
      }
    }
    T* dst = stage[buf];
    // parallel fill over slabs of this chunk
    std::atomic<int32_t> next{s};
    auto worker = [&]() {
      while (true) {
        const int32_t i = next.fetch_add(8);
        if (i >= e) break;
        const int32_t j = std::min(e, i + 8);
        bwl::fill_slabs<T>(L, P, i, j, dst + (P.slabs[i].data_off - P.slabs[s].data_off));
      }
    };
    std::vector<std::thread> ths;
    for (int t = 1; t < nthreads; ++t) ths.emplace_back(worker);
    worker();
    for (auto& t : ths) t.join();
    hipError_t ce = hipMemcpyAsync(static_cast<T*>(dp->data) + P.slabs[s].data_off, dst,
                                   elems * sizeof(T), hipMemcpyHostToDevice, st);
    if (ce != hipSuccess) {
      status = fail(BW_EHIP, std::string("upload: ") + hipGetErrorString(ce));
      break;
    }
    const hipError_t re = hipEventRecord(done[buf], st);
    if (re != hipSuccess) {
      status = fail(BW_EHIP, std::string("upload: ") + hipGetErrorString(re));
      break;
    }
    pending[buf] = true;
    buf ^= 1;
    if (verbose() && nslab >= 10 && (e * 10 / nslab) != (s * 10 / nslab))
      std::fprintf(stderr, "[bwhip] upload %d / %d slabs\n", e, nslab);
    s = e;
  }
  // every copy has ended (or failed) before the staging buffers are freed
  const hipError_t se = hipStreamSynchronize(st);
  if (se != hipSuccess && status == BW_OK)
    status = fail(BW_EHIP, std::string("upload: ") + hipGetErrorString(se));
  (void)hipEventDestroy(done[0]);
  (void)hipEventDestroy(done[1]);
  (void)hipHostFree(stage[0]);
  (void)hipHostFree(stage[1]);
  return status;
}

}  // namespace

// One direction of a selection operator (0: H, 1: H^T).
struct SelDir {
  int64_t* in_src = nullptr;  // point of each scattered input in the caller's vector
  int64_t* in_dst = nullptr;  // its permuted point (repeated indices: the last one wins)
  int64_t n_pairs = 0;
  int64_t n_in = 0;           // points in the caller's input vector
  int64_t* out_dst = nullptr; // permuted point of each output point
  int64_t n_out = 0;
};

struct SelInfo {
  bool on = false;
  int ordering = BW_ORDER_ORIGINAL;
  SelDir dir[2];
};

struct bw_hip_handle {
  int device = 0;
  int dof = 1;
  int64_t n_row_pts = 0, n_col_pts = 0;
  int64_t nb_entries = 0;
  // algorithmic entries of the factors each precision streams: FR blocks,
  // left factors A, right factors B, and the total rank T ([fp32])
  int64_t entries[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  double fp32_rank_tol = 0.0;
  DevOp op[2][2];  // [adjoint][fp32]
  int64_t* perm0 = nullptr;
  int64_t* perm1 = nullptr;
  std::vector<double> diag_perm, diag_orig;
  std::vector<void*> allocs;
  int64_t device_bytes = 0;
  SelInfo sel;
  // Scratch per stream, allocated on a stream's first call (the null stream's
  // at creation).  stage_elems: doubles of each host-pointer staging vector;
  // yfull_elems: doubles of a selection's full permuted output (0: none).
  std::vector<std::unique_ptr<StreamCtx>> ctxs;
  int64_t stage_elems = 0;
  int64_t yfull_elems = 0;
  // Kernel timing: two alternating sets of events between the kernels of a
  // matvec; a set is read (and accumulated) only when reused two calls later,
  // so timing never stalls the launch pipeline.
  bool timing = false;
  hipEvent_t evs[2][6] = {};
  hipEvent_t* ev = evs[0];
  bool pending[2] = {false, false};
  int cur = 0;
  double sum_ms[6] = {0, 0, 0, 0, 0, 0};
  int64_t n_timed = 0;
  std::mutex mu;  // host-side state (timing ring, lazily allocated batch buffers)
};

namespace {

template <typename T>
int build_op(bw_hip_handle* hd, const bw_flat_hmat* h, bool adjoint, hipStream_t st) {
  DevOp& op = hd->op[adjoint ? 1 : 0][sizeof(T) == 4 ? 1 : 0];
  bwl::OpView view;
  bwl::make_view(h, adjoint, &view);
  bwl::LayoutOptions opt;
  opt.vec_elems = Vec<T>::V;
  opt.elem_bytes = sizeof(T);
  opt.wi_bytes = sizeof(T) == 8 ? g_wi_bytes64.load() : g_wi_bytes32.load();
  opt.stack_rows = g_stack_rows.load();
  if (opt.wi_bytes == 0) {
    int64_t E = 0;
    for (int64_t k = 0; k < h->n_fr; ++k) {
      const int64_t* c = h->fr_pattern + 6 * k;
      E += h->dof * (c[3] - c[1]) * h->dof * (c[4] - c[2]);
    }
    for (int64_t k = 0; k < h->n_lr; ++k) {
      const int64_t* c = h->lr_pattern + 6 * k;
      E += c[5] * h->dof * ((c[3] - c[1]) + (c[4] - c[2]));
    }
    opt.wi_bytes = auto_wi_bytes(E, sizeof(T));
  }
  bwl::Layout L;
  bwl::build_layout(view, opt, &L);
  std::string msg;
  int rc = bwl::check_layout_limits(L, &msg);
  if (rc) return fail(rc, "create: " + msg);
  if ((rc = upload_pass<T>(L, L.pass1, &op.p1, st, &hd->allocs))) return rc;
  if ((rc = upload_pass<T>(L, L.pass2, &op.p2, st, &hd->allocs))) return rc;
  op.elem = sizeof(T);
  // default batched schedule: the fastest on MI355X for the 150k-DoF 3DR0
  // matrix (FP32: two pairs, load-order ring of 3, B staged through LDS;
  // FP64: ring of 4)
  op.mfma_schedule = sizeof(T) == 8 ? 5 : 14;
  op.xt_size = L.xt_size;
  op.xt_t = L.xt_t;
  op.t_total = L.view.t_total;
  op.n_in = static_cast<int64_t>(h->dof) * view.n_col_pts;
  op.n_out = static_cast<int64_t>(h->dof) * view.n_row_pts;
  op.perm_in = adjoint ? hd->perm0 : hd->perm1;
  op.perm_out = adjoint ? hd->perm1 : hd->perm0;
  op.stream_bytes = (L.pass1.data_elems + L.pass2.data_elems) * static_cast<int64_t>(sizeof(T));
  op.part_elems[0] = L.pass1.part_elems;
  op.part_elems[1] = L.pass2.part_elems;
  hd->device_bytes += op.p1.data_bytes + op.p2.data_bytes + L.xt_size * sizeof(T) +
                      (L.pass1.part_elems + L.pass2.part_elems) * sizeof(T) +
                      (L.pass1.wis.size() + L.pass2.wis.size()) * sizeof(bwl::WorkItem) +
                      (L.pass1.gidx.size() + L.pass2.gidx.size()) * sizeof(int32_t);
  op.present = true;
  return BW_OK;
}

// Frees the allocations a failed ensure_* call made after `mark`.
int undo_allocs(StreamCtx* ctx, size_t mark, int rc) {
  for (size_t i = mark; i < ctx->allocs.size(); ++i) (void)hipFree(ctx->allocs[i]);
  ctx->allocs.resize(mark);
  return rc;
}

// Zeroes fresh scratch and waits for it, so that no kernel of any stream can
// see it before the zeros land.  (hipMemset of device memory returns before
// it completes, on the null stream, which does not order non-blocking
// streams such as torch's.)
hipError_t zero_now(void* p, size_t bytes) {
  hipError_t e = hipMemsetAsync(p, 0, bytes, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  return e;
}

// Allocates what `ctx` lacks for operator [adj][fp32]: the single-vector
// scratch, and with `batched` the kRhs-wide one.  xt / xb are zeroed here,
// once (synchronously, before any kernel of the context can use them): the
// kernels only ever write x and t entries, so the gap and tail stay zero.
// On failure nothing of this call stays allocated.
int ensure_op_scratch(bw_hip_handle* hd, StreamCtx* ctx, int adj, int fp32, bool batched) {
  const DevOp& op = hd->op[adj][fp32];
  OpScratch& s = ctx->op[adj][fp32];
  const size_t es = op.elem;
  const size_t mark = ctx->allocs.size();
  int rc;
  if (!s.xt) {
    void* xt = nullptr;
    void* part[2] = {nullptr, nullptr};
    if ((rc = dev_alloc(&xt, op.xt_size * es, &ctx->allocs, "scratch"))) return undo_allocs(ctx, mark, rc);
    for (int p = 0; p < 2; ++p)
      if ((rc = dev_alloc(&part[p], op.part_elems[p] * es, &ctx->allocs, "scratch")))
        return undo_allocs(ctx, mark, rc);
    const hipError_t e = zero_now(xt, op.xt_size * es);
    if (e != hipSuccess)
      return undo_allocs(ctx, mark, fail(BW_EHIP, std::string("scratch: zero fill: ") + hipGetErrorString(e)));
    s.xt = xt;
    s.part[0] = part[0];
    s.part[1] = part[1];
  }
  if (batched && !s.xb) {
    const size_t mark_b = ctx->allocs.size();
    void* xb = nullptr;
    void* partb[2] = {nullptr, nullptr};
    if ((rc = dev_alloc(&xb, static_cast<size_t>(op.xt_size) * kRhs * es, &ctx->allocs, "scratch")))
      return undo_allocs(ctx, mark_b, rc);
    for (int p = 0; p < 2; ++p)
      if ((rc = dev_alloc(&partb[p], static_cast<size_t>(op.part_elems[p]) * kRhs * es, &ctx->allocs,
                          "scratch")))
        return undo_allocs(ctx, mark_b, rc);
    const hipError_t e = zero_now(xb, static_cast<size_t>(op.xt_size) * kRhs * es);
    if (e != hipSuccess)
      return undo_allocs(ctx, mark_b, fail(BW_EHIP, std::string("scratch: zero fill: ") + hipGetErrorString(e)));
    s.xb = xb;
    s.partb[0] = partb[0];
    s.partb[1] = partb[1];
  }
  return BW_OK;
}

// Staging vectors of host-pointer calls and a selection's full output.
int ensure_ctx_vectors(bw_hip_handle* hd, StreamCtx* ctx, bool staging) {
  const size_t mark = ctx->allocs.size();
  int rc;
  if (staging && !ctx->x_stage) {
    void *xs = nullptr, *ys = nullptr;
    if ((rc = dev_alloc(&xs, hd->stage_elems * sizeof(double), &ctx->allocs, "scratch")) ||
        (rc = dev_alloc(&ys, hd->stage_elems * sizeof(double), &ctx->allocs, "scratch")))
      return undo_allocs(ctx, mark, rc);
    ctx->x_stage = static_cast<double*>(xs);
    ctx->y_stage = static_cast<double*>(ys);
  }
  if (hd->yfull_elems > 0 && !ctx->y_full) {
    const size_t mark_y = ctx->allocs.size();
    void* yf = nullptr;
    if ((rc = dev_alloc(&yf, hd->yfull_elems * sizeof(double), &ctx->allocs, "scratch")))
      return undo_allocs(ctx, mark_y, rc);
    ctx->y_full = static_cast<double*>(yf);
  }
  return BW_OK;
}

void free_ctx(StreamCtx* ctx) {
  for (void* p : ctx->allocs) (void)hipFree(p);
  ctx->allocs.clear();
}

// The context of `stream` (created empty on its first use); caller holds hd->mu.
StreamCtx* stream_ctx(bw_hip_handle* hd, hipStream_t stream) {
  for (auto& c : hd->ctxs)
    if (c->stream == stream) return c.get();
  hd->ctxs.emplace_back(new StreamCtx());
  hd->ctxs.back()->stream = stream;
  return hd->ctxs.back().get();
}

void compute_diagonal(bw_hip_handle* hd, const bw_flat_hmat* h) {
  const int64_t d = h->dof;
  const int64_t npt = std::max(h->n_row_pts, h->n_col_pts);
  hd->diag_perm.assign(d * npt, 0.0);
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t* c = h->fr_pattern + 6 * k;
    const int64_t i0 = c[1], j0 = c[2], iend = c[3], jend = c[4];
    const bool lower = (i0 > j0) && (i0 > jend) && (iend > j0) && (iend > jend);
    const bool upper = (i0 < j0) && (i0 < jend) && (iend < j0) && (iend < jend);
    if (lower || upper) continue;
    const int64_t m = d * (iend - i0), n = d * (jend - j0);
    const double* F = h->fr_blocks[k];
    for (int64_t j = 0; j < n; ++j) {
      const int64_t i = j + d * j0 - d * i0;
      if (i >= 0 && i < m) hd->diag_perm[i + d * i0] = F[i + j * m];
    }
  }
  hd->diag_orig.assign(d * npt, 0.0);
  for (int64_t i = 0; i < npt && i < h->n_col_pts; ++i)
    for (int64_t j = 0; j < d; ++j) hd->diag_orig[d * h->perm1[i] + j] = hd->diag_perm[d * i + j];
}

template <typename T, typename OutT, int PASS>
void launch_slab(int unroll, int blocks, hipStream_t st, const bwl::WorkItem* wis, XcdSplit xs,
                 const int32_t* gidx, const T* data, const T* xin, OutT* out, T* part,
                 const int64_t* perm, int dof, const int32_t* rowmap) {
  // `blocks` counts four-wave workgroups (the layout's XCD split); k_slab
  // runs one wave per workgroup
  constexpr int kWavesPerGroup = 4;
  if (unroll == 4)
    hipLaunchKernelGGL((k_slab<T, OutT, 4, PASS>), dim3(kWavesPerGroup * blocks), dim3(64), 0,
                       st, wis, xs, gidx, data, xin, out, part, perm, dof, rowmap);
  else if (unroll == 16)
    hipLaunchKernelGGL((k_slab<T, OutT, 16, PASS>), dim3(kWavesPerGroup * blocks), dim3(64), 0,
                       st, wis, xs, gidx, data, xin, out, part, perm, dof, rowmap);
  else
    hipLaunchKernelGGL((k_slab<T, OutT, 8, PASS>), dim3(kWavesPerGroup * blocks), dim3(64), 0,
                       st, wis, xs, gidx, data, xin, out, part, perm, dof, rowmap);
}

// Zero fill of xt's gap after x and its tail after t, for scratch that other
// operators may have used before (the no-allocation entry point).
template <typename T>
__global__ __launch_bounds__(256) void k_zero_ranges(T* __restrict__ xt, int64_t a0, int64_t a1,
                                                     int64_t b0, int64_t b1) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t na = a1 - a0;
  if (i < na)
    xt[a0 + i] = T(0);
  else if (i - na < b1 - b0)
    xt[b0 + (i - na)] = T(0);
}

// The five (selection: six) kernels of one matvec on stream st, writing only
// the scratch `s` (and y_full for a selection).
template <typename T>
int launch_matvec(bw_hip_handle* hd, DevOp& op, const OpScratch& s, double* y_full,
                  const double* xd, double* yd, int ordering, hipStream_t st,
                  const SelDir* sd = nullptr, bool timed = true) {
  const bool orig = !sd && ordering == BW_ORDER_ORIGINAL;
  const int dof = hd->dof;
  T* xt = static_cast<T*>(s.xt);
  T* part1 = static_cast<T*>(s.part[0]);
  T* part2 = static_cast<T*>(s.part[1]);
  double* const y_user = yd;
  if (sd) yd = y_full;  // the passes write the full permuted y
  const bool timing = timed && hd->timing;
  if (timing) (void)hipEventRecord(hd->ev[0], st);
  if (sd) {
    if (sd->n_pairs > 0) {
      const int blocks = static_cast<int>((sd->n_pairs * dof + 255) / 256);
      hipLaunchKernelGGL(k_scatter_sel<T>, dim3(blocks), dim3(256), 0, st, xd, xt, sd->in_src,
                         sd->in_dst, sd->n_pairs, dof);
    }
  } else {
    const int64_t n = op.n_in;
    const int blocks = static_cast<int>((n + 255) / 256);
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks), dim3(256), 0, st, xd, xt, n,
                       orig ? op.perm_in : nullptr, dof);
  }
  if (timing) (void)hipEventRecord(hd->ev[1], st);
  if (op.p1.nwi > 0) {
    launch_slab<T, T, 1>(op.unroll[0], op.p1.blocks, st, op.p1.wis, op.p1.xs, op.p1.gidx,
                         static_cast<const T*>(op.p1.data), xt, xt, part1, nullptr, dof,
                         op.p1.rowmap);
  }
  if (timing) (void)hipEventRecord(hd->ev[2], st);
  if (op.p1.nred > 0)
    hipLaunchKernelGGL((k_reduce<T, T>), dim3(op.p1.nred), dim3(64), 0, st, op.p1.reds,
                       op.p1.nred, part1, xt, nullptr, dof, op.p1.rowmap);
  if (timing) (void)hipEventRecord(hd->ev[3], st);
  if (op.p2.nwi > 0) {
    launch_slab<T, double, 2>(op.unroll[1], op.p2.blocks, st, op.p2.wis, op.p2.xs, op.p2.gidx,
                              static_cast<const T*>(op.p2.data), xt, yd, part2,
                              orig ? op.perm_out : nullptr, dof, nullptr);
  }
  if (timing) (void)hipEventRecord(hd->ev[4], st);
  if (op.p2.nred > 0)
    hipLaunchKernelGGL((k_reduce<T, double>), dim3(op.p2.nred), dim3(64), 0, st, op.p2.reds,
                       op.p2.nred, part2, yd, orig ? op.perm_out : nullptr, dof, nullptr);
  if (sd && sd->n_out > 0) {
    const int blocks = static_cast<int>((sd->n_out * dof + 255) / 256);
    hipLaunchKernelGGL(k_gather_sel, dim3(blocks), dim3(256), 0, st, yd, y_user, sd->out_dst,
                       sd->n_out, dof);
  }
  if (timing) (void)hipEventRecord(hd->ev[5], st);
  HIPCHK(hipGetLastError());
  return BW_OK;
}

int accumulate_timing(bw_hip_handle* hd, int set) {
  hipEvent_t* e = hd->evs[set];
  HIPCHK(hipEventSynchronize(e[5]));
  float ms;
  for (int i = 0; i < 5; ++i) {
    HIPCHK(hipEventElapsedTime(&ms, e[i], e[i + 1]));
    hd->sum_ms[i] += ms;
  }
  HIPCHK(hipEventElapsedTime(&ms, e[0], e[5]));
  hd->sum_ms[5] += ms;
  hd->n_timed += 1;
  hd->pending[set] = false;
  return BW_OK;
}

int do_matvec(bw_hip_handle* hd, bool adjoint, const double* x, double* y, int ordering,
              int precision, int on_device, void* stream) {
  if (!hd || !x || !y) return fail(BW_EINVAL, "matvec: null argument");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "matvec: ordering must be 0 (permuted) or 1 (original)");
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32)
    return fail(BW_EINVAL, "matvec: precision must be BW_PREC_FP64 or BW_PREC_FP32");
  const int adj = adjoint ? 1 : 0, f32 = precision == BW_PREC_FP32 ? 1 : 0;
  DevOp& op = hd->op[adj][f32];
  if (!op.present)
    return fail(BW_ENOTSUP, adjoint ? "rmatvec: adjoint / precision not built for this handle"
                                    : "matvec: precision not built for this handle");
  const SelDir* sd = nullptr;
  if (hd->sel.on) {
    if (ordering != hd->sel.ordering)
      return fail(BW_ENOTSUP, "matvec: this selection was created for the other ordering");
    sd = &hd->sel.dir[adj];
  }
  const int64_t n_in = sd ? hd->dof * sd->n_in : op.n_in;
  const int64_t n_out = sd ? hd->dof * sd->n_out : op.n_out;
  std::lock_guard<std::mutex> lock(hd->mu);
  arm_alloc_hook();
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(hd->device));
  StreamCtx* ctx = stream_ctx(hd, st);
  int rc;
  if ((rc = ensure_op_scratch(hd, ctx, adj, f32, false)) ||
      (rc = ensure_ctx_vectors(hd, ctx, !on_device)))
    return rc;
  const double* xd = x;
  double* yd = y;
  if (!on_device) {
    HIPCHK(hipMemcpyAsync(ctx->x_stage, x, n_in * sizeof(double), hipMemcpyHostToDevice, st));
    xd = ctx->x_stage;
    yd = ctx->y_stage;
  }
  if (hd->timing) {
    const int set = hd->cur;
    if (hd->pending[set]) {
      int rc0 = accumulate_timing(hd, set);
      if (rc0) return rc0;
    }
    hd->ev = hd->evs[set];
  }
  const OpScratch& s = ctx->op[adj][f32];
  rc = f32 ? launch_matvec<float>(hd, op, s, ctx->y_full, xd, yd, ordering, st, sd)
           : launch_matvec<double>(hd, op, s, ctx->y_full, xd, yd, ordering, st, sd);
  if (rc) return rc;
  if (hd->timing) {
    hd->pending[hd->cur] = true;
    hd->cur ^= 1;
  }
  if (!on_device) {
    HIPCHK(hipMemcpyAsync(y, ctx->y_stage, n_out * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return BW_OK;
}

// Device address of `scratch` on each GPU (index: device ordinal), taken
// once per device by noalloc_bind with that device current; nullptr until
// then, or if hipGetSymbolAddress failed there (noalloc_err holds why).
std::mutex g_noalloc_mu;
std::vector<void*> g_noalloc_ptr;
std::vector<std::string> g_noalloc_err;

// Called by every create on its device (current).  A failure only disables
// bw_hip_matvec_noalloc on that device; the other entry points never use
// the array.
void noalloc_bind(int device) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (static_cast<int>(g_noalloc_ptr.size()) <= device) {
    g_noalloc_ptr.resize(device + 1, nullptr);
    g_noalloc_err.resize(device + 1);
  }
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_noalloc_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_noalloc_ptr[device] = p;
}

int noalloc_matvec(bw_hip_handle* hd, const double* x, double* y, int ordering, int precision,
                   void* stream) {
  if (!hd || !x || !y) return fail(BW_EINVAL, "matvec_noalloc: null argument");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "matvec_noalloc: ordering must be 0 (permuted) or 1 (original)");
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32)
    return fail(BW_EINVAL, "matvec_noalloc: precision must be BW_PREC_FP64 or BW_PREC_FP32");
  if (hd->sel.on) return fail(BW_ENOTSUP, "matvec_noalloc: not available on a selection");
  const int f32 = precision == BW_PREC_FP32 ? 1 : 0;
  DevOp& op = hd->op[0][f32];
  if (!op.present) return fail(BW_ENOTSUP, "matvec_noalloc: precision not built for this handle");
  // xt, then the partial sums of pass 1 and pass 2, each 256-byte aligned
  const auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
  const int64_t off1 = up(op.xt_size * op.elem);
  const int64_t off2 = off1 + up(op.part_elems[0] * op.elem);
  const int64_t need = off2 + up(op.part_elems[1] * op.elem);
  int64_t limit = BW_NOALLOC_SCRATCH_BYTES;
  if (const char* e = std::getenv("BWHIP_NOALLOC_LIMIT")) {
    const long long v = std::atoll(e);
    if (v > 0 && v < limit) limit = v;  // tests of the size check may only lower it
  }
  if (need > limit)
    return fail(BW_ESHAPE, "matvec_noalloc: this operator needs " + std::to_string(need) +
                               " bytes of scratch; the fixed array holds " + std::to_string(limit) +
                               " (use bw_hip_matvec)");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (hd->device < static_cast<int>(g_noalloc_ptr.size())) base = g_noalloc_ptr[hd->device];
    if (!base)
      return fail(BW_EHIP, "matvec_noalloc: no scratch address on this device: " +
                               (hd->device < static_cast<int>(g_noalloc_err.size())
                                    ? g_noalloc_err[hd->device]
                                    : std::string("not bound")));
  }
  std::lock_guard<std::mutex> lock(hd->mu);
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipSetDevice(hd->device));
  unsigned char* b = static_cast<unsigned char*>(base);
  OpScratch s;
  s.xt = b;
  s.part[0] = b + off1;
  s.part[1] = b + off2;
  // Padded column groups read xt past x (the gap) and past t (the tail), so
  // those must hold zeros; an earlier call on this GPU may have left other
  // values there.  x and t are rewritten by every call, and every partial
  // sum is written before it is read.  A fill kernel, not hipMemsetAsync,
  // so the call can be captured into a graph.
  const int64_t na = op.xt_t - op.n_in, nb = op.xt_size - (op.xt_t + op.t_total);
  const int zblocks = static_cast<int>((na + nb + 255) / 256);
  if (zblocks > 0) {
    if (f32)
      hipLaunchKernelGGL(k_zero_ranges<float>, dim3(zblocks), dim3(256), 0, st,
                         static_cast<float*>(s.xt), op.n_in, op.xt_t, op.xt_t + op.t_total,
                         op.xt_size);
    else
      hipLaunchKernelGGL(k_zero_ranges<double>, dim3(zblocks), dim3(256), 0, st,
                         static_cast<double*>(s.xt), op.n_in, op.xt_t, op.xt_t + op.t_total,
                         op.xt_size);
  }
  return f32 ? launch_matvec<float>(hd, op, s, nullptr, x, y, ordering, st, nullptr, false)
             : launch_matvec<double>(hd, op, s, nullptr, x, y, ordering, st, nullptr, false);
}

}  // namespace

extern "C" {

int bw_hip_matvec_noalloc(bw_hip_handle* hd, const double* x, double* y, int ordering,
                          int precision, void* stream) {
  return noalloc_matvec(hd, x, y, ordering, precision, stream);
}

const char* bw_hip_last_error(void) { return g_err.c_str(); }

int bw_hip_set_tuning(int64_t wi64, int64_t wi32) {
  if (wi64 < 0 || wi32 < 0) return fail(BW_EINVAL, "set_tuning: negative size");
  g_wi_bytes64 = wi64;
  g_wi_bytes32 = wi32;
  return BW_OK;
}

int bw_hip_set_fp32_rank_tolerance(double rel_tol) {
  if (!(rel_tol >= 0.0) || !(rel_tol < 1.0))
    return fail(BW_EINVAL, "set_fp32_rank_tolerance: tolerance must be in [0, 1)");
  g_fp32_rank_tol = rel_tol;
  return BW_OK;
}

int bw_hip_entries(bw_hip_handle* hd, int precision, int64_t* e_full, int64_t* e_left,
                   int64_t* e_right, int64_t* total_rank) {
  if (!hd) return fail(BW_EINVAL, "entries: null handle");
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32)
    return fail(BW_EINVAL, "entries: precision must be BW_PREC_FP64 or BW_PREC_FP32");
  const int64_t* e = hd->entries[precision == BW_PREC_FP32 ? 1 : 0];
  if (e_full) *e_full = e[0];
  if (e_left) *e_left = e[1];
  if (e_right) *e_right = e[2];
  if (total_rank) *total_rank = e[3];
  return BW_OK;
}

int bw_hip_set_stacking(int stack_rows) {
  if (stack_rows < 0 || stack_rows > 64) return fail(BW_EINVAL, "set_stacking: rows must be 0-64");
  g_stack_rows = stack_rows;
  return BW_OK;
}

void bw_hip_destroy(bw_hip_handle* hd) {
  if (!hd) return;
  (void)hipSetDevice(hd->device);
  (void)hipDeviceSynchronize();
  for (auto& c : hd->ctxs) free_ctx(c.get());
  for (void* p : hd->allocs) (void)hipFree(p);
  for (auto& set : hd->evs)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  delete hd;
}

}  // extern "C"

namespace bwl {
int compose_flat(const bw_flat_hmat* const* hs, int n, bw_flat_composite* c, std::string* msg);
}

namespace {

// The null stream's single-vector scratch and staging, allocated up front
// (most callers use that stream; other streams get theirs on first use).
int prepare_null_stream(bw_hip_handle* hd) {
  StreamCtx* ctx0 = stream_ctx(hd, nullptr);
  int rc;
  for (int adj = 0; adj < 2; ++adj)
    for (int f32 = 0; f32 < 2; ++f32)
      if (hd->op[adj][f32].present && (rc = ensure_op_scratch(hd, ctx0, adj, f32, false))) return rc;
  return ensure_ctx_vectors(hd, ctx0, true);
}

int create_impl(const bw_flat_hmat* h, long double expected_area, int precision_mask, int flags,
                int device, bw_hip_handle** out, bool prepare_ctx0 = true) {
  if (!out) return fail(BW_EINVAL, "create: null output pointer");
  *out = nullptr;
  std::string msg;
  int rc = bwl::validate_flat(h, &msg, expected_area);
  if (rc) return fail(rc, "create: " + msg);
  if ((precision_mask & (BW_PREC_FP64 | BW_PREC_FP32)) == 0 ||
      (precision_mask & ~(BW_PREC_FP64 | BW_PREC_FP32)))
    return fail(BW_EINVAL, "create: precision_mask must combine BW_PREC_FP64 / BW_PREC_FP32");
  if (flags & ~BW_FLAG_ADJOINT) return fail(BW_EINVAL, "create: unknown flag");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(BW_EHIP, "create: no HIP device available");
  auto* hd = new bw_hip_handle();
  if (device < 0) {
    if (hipGetDevice(&hd->device) != hipSuccess) hd->device = 0;
  } else {
    hd->device = device;
  }
  auto cleanup = [&](int code) {
    bw_hip_destroy(hd);
    return code;
  };
  if (hipSetDevice(hd->device) != hipSuccess) return cleanup(fail(BW_EHIP, "create: hipSetDevice"));
  noalloc_bind(hd->device);  // once per device: the address of bw_hip_matvec_noalloc's array
  hd->dof = h->dof;
  hd->n_row_pts = h->n_row_pts;
  hd->n_col_pts = h->n_col_pts;
  const int64_t d = h->dof;
  auto dmal = [&](void** p, size_t bytes) -> int { return dev_alloc(p, bytes, &hd->allocs, "create"); };
  if ((rc = dmal(reinterpret_cast<void**>(&hd->perm0), h->n_row_pts * sizeof(int64_t)))) return cleanup(rc);
  if ((rc = dmal(reinterpret_cast<void**>(&hd->perm1), h->n_col_pts * sizeof(int64_t)))) return cleanup(rc);
  hd->stage_elems = d * std::max(h->n_row_pts, h->n_col_pts);
  if (hipMemcpy(hd->perm0, h->perm0, h->n_row_pts * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(hd->perm1, h->perm1, h->n_col_pts * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(BW_EHIP, "create: permutation upload"));
  // the FP32 copy may stream recompressed low-rank factors
  bwl::RecompressedFlat rf;
  const bw_flat_hmat* h32 = h;
  hd->fp32_rank_tol = g_fp32_rank_tol.load();
  if ((precision_mask & BW_PREC_FP32) && hd->fp32_rank_tol > 0.0) {
    std::string rmsg;
    if ((rc = bwl::recompress_flat(h, hd->fp32_rank_tol, &rf, &rmsg, expected_area)))
      return cleanup(fail(rc, "create: recompression: " + rmsg));
    h32 = &rf.flat;
  }
  for (int f32 = 0; f32 < 2; ++f32) {
    const bw_flat_hmat* hp = f32 ? h32 : h;
    int64_t* e = hd->entries[f32];
    for (int64_t k = 0; k < hp->n_fr; ++k) {
      const int64_t* c = hp->fr_pattern + 6 * k;
      e[0] += d * (c[3] - c[1]) * d * (c[4] - c[2]);
    }
    for (int64_t k = 0; k < hp->n_lr; ++k) {
      const int64_t* c = hp->lr_pattern + 6 * k;
      e[1] += d * (c[3] - c[1]) * c[5];
      e[2] += c[5] * d * (c[4] - c[2]);
      e[3] += c[5];
    }
  }
  hipStream_t st;
  if (hipStreamCreate(&st) != hipSuccess) return cleanup(fail(BW_EHIP, "create: hipStreamCreate"));
  for (int adj = 0; adj < 2; ++adj) {
    if (adj && !(flags & BW_FLAG_ADJOINT)) continue;
    if (precision_mask & BW_PREC_FP64) {
      rc = build_op<double>(hd, h, adj != 0, st);
      if (rc) {
        (void)hipStreamDestroy(st);
        return cleanup(rc);
      }
    }
    if (precision_mask & BW_PREC_FP32) {
      rc = build_op<float>(hd, h32, adj != 0, st);
      if (rc) {
        (void)hipStreamDestroy(st);
        return cleanup(rc);
      }
    }
  }
  (void)hipStreamDestroy(st);
  // nbOfEntries (hmat.cc:299-312)
  int64_t E = 0;
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t* c = h->fr_pattern + 6 * k;
    E += d * (c[3] - c[1]) * d * (c[4] - c[2]);
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t* c = h->lr_pattern + 6 * k;
    E += d * (c[3] - c[1]) * c[5] + c[5] * d * (c[4] - c[2]);
  }
  hd->nb_entries = E;
  compute_diagonal(hd, h);
  for (auto& set : hd->evs)
    for (auto& e : set)
      if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(BW_EHIP, "create: hipEventCreate"));
  if (prepare_ctx0 && (rc = prepare_null_stream(hd))) return cleanup(rc);
  *out = hd;
  return BW_OK;
}

}  // namespace

// A captured matvec.  It owns its scratch, so replays never race with
// eager calls on the handle (two replays of one graph must be ordered).
struct bw_hip_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  StreamCtx ctx;
};

extern "C" {

int bw_hip_create(const bw_flat_hmat* h, int precision_mask, int flags, int device,
                  bw_hip_handle** out) {
  arm_alloc_hook();
  return create_impl(h, -1.0L, precision_mask, flags, device, out);
}

int bw_hip_create_multi(const bw_flat_hmat* const* hs, int n_mats, int precision_mask, int flags,
                        int device, bw_hip_handle** out, int64_t* row_off, int64_t* col_off) {
  if (!out) return fail(BW_EINVAL, "create_multi: null output pointer");
  *out = nullptr;
  arm_alloc_hook();
  bw_flat_composite* c = nullptr;
  int rc = bw_flat_compose(hs, n_mats, &c);
  if (rc) return fail(rc, std::string("create_multi: ") + bw_flat_compose_last_error());
  const bw_flat_hmat* f = bw_flat_composite_view(c);
  long double area = 0.0L;
  for (int m = 0; m < n_mats; ++m)
    area += static_cast<long double>(hs[m]->n_row_pts) * hs[m]->n_col_pts;
  rc = create_impl(f, area, precision_mask, flags, device, out);
  if (!rc) bw_flat_composite_offsets(c, row_off, col_off);
  bw_flat_composite_free(c);
  return rc;
}

int bw_hip_create_selection(const bw_flat_hmat* h, const int64_t* rows, int64_t n_rows,
                            const int64_t* cols, int64_t n_cols, int ordering,
                            int precision_mask, int flags, int device, bw_hip_handle** out) {
  if (!out) return fail(BW_EINVAL, "create_selection: null output pointer");
  *out = nullptr;
  arm_alloc_hook();
  std::string msg;
  int rc = bwl::validate_flat(h, &msg);
  if (rc) return fail(rc, "create_selection: " + msg);
  if (!rows || !cols || n_rows <= 0 || n_cols <= 0)
    return fail(BW_EINVAL, "create_selection: empty row or column selection");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "create_selection: ordering must be 0 (permuted) or 1 (original)");
  for (int64_t i = 0; i < n_rows; ++i)
    if (rows[i] < 0 || rows[i] >= h->n_row_pts)
      return fail(BW_EINVAL, "create_selection: row index out of bounds");
  for (int64_t i = 0; i < n_cols; ++i)
    if (cols[i] < 0 || cols[i] >= h->n_col_pts)
      return fail(BW_EINVAL, "create_selection: column index out of bounds");
  // selected points in the permuted numbering (HmatSelection::blockSelection,
  // hmat_selection.cc:44-77): the inverse permutation for original indices
  std::vector<int64_t> inv0(h->n_row_pts), inv1(h->n_col_pts);
  for (int64_t i = 0; i < h->n_row_pts; ++i) inv0[h->perm0[i]] = i;
  for (int64_t i = 0; i < h->n_col_pts; ++i) inv1[h->perm1[i]] = i;
  const bool orig = ordering == BW_ORDER_ORIGINAL;
  std::vector<int64_t> prow(n_rows), pcol(n_cols);
  for (int64_t i = 0; i < n_rows; ++i) prow[i] = orig ? inv0[rows[i]] : rows[i];
  for (int64_t i = 0; i < n_cols; ++i) pcol[i] = orig ? inv1[cols[i]] : cols[i];
  // prefix counts of marked points: a block is kept if its row range holds a
  // selected row and its column range a selected column (hmat_selection.cc:82-128)
  std::vector<int64_t> rcnt(h->n_row_pts + 1, 0), ccnt(h->n_col_pts + 1, 0);
  {
    std::vector<char> mr(h->n_row_pts, 0), mc(h->n_col_pts, 0);
    for (int64_t p : prow) mr[p] = 1;
    for (int64_t p : pcol) mc[p] = 1;
    for (int64_t i = 0; i < h->n_row_pts; ++i) rcnt[i + 1] = rcnt[i] + mr[i];
    for (int64_t i = 0; i < h->n_col_pts; ++i) ccnt[i + 1] = ccnt[i] + mc[i];
  }
  auto keep = [&](const int64_t* c) {
    return rcnt[c[3]] > rcnt[c[1]] && ccnt[c[4]] > ccnt[c[2]];
  };
  std::vector<int64_t> fr_pat, lr_pat;
  std::vector<const double*> fr_ptr, lr_a, lr_b;
  long double area = 0.0L;
  for (int64_t k = 0; k < h->n_fr; ++k) {
    const int64_t* c = h->fr_pattern + 6 * k;
    if (!keep(c)) continue;
    fr_pat.insert(fr_pat.end(), c, c + 6);
    fr_ptr.push_back(h->fr_blocks[k]);
    area += static_cast<long double>(c[3] - c[1]) * (c[4] - c[2]);
  }
  for (int64_t k = 0; k < h->n_lr; ++k) {
    const int64_t* c = h->lr_pattern + 6 * k;
    if (!keep(c)) continue;
    lr_pat.insert(lr_pat.end(), c, c + 6);
    lr_a.push_back(h->lr_A[k]);
    lr_b.push_back(h->lr_B[k]);
    area += static_cast<long double>(c[3] - c[1]) * (c[4] - c[2]);
  }
  bw_flat_hmat sub = *h;
  sub.n_fr = static_cast<int64_t>(fr_ptr.size());
  sub.fr_pattern = sub.n_fr ? fr_pat.data() : nullptr;
  sub.fr_blocks = sub.n_fr ? fr_ptr.data() : nullptr;
  sub.n_lr = static_cast<int64_t>(lr_a.size());
  sub.lr_pattern = sub.n_lr ? lr_pat.data() : nullptr;
  sub.lr_A = sub.n_lr ? lr_a.data() : nullptr;
  sub.lr_B = sub.n_lr ? lr_b.data() : nullptr;
  bw_hip_handle* hd = nullptr;
  rc = create_impl(&sub, area, precision_mask, flags, device, &hd, false);
  if (rc) return rc;
  // scatter pairs with repeated targets reduced to the last occurrence, as
  // HmatSelection::matvec overwrites x_full (hmat_selection.cc:197-205)
  auto pairs = [](const std::vector<int64_t>& dst, std::vector<int64_t>* s, std::vector<int64_t>* d) {
    std::vector<int64_t> order(dst.size());
    for (size_t i = 0; i < dst.size(); ++i) order[i] = static_cast<int64_t>(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return dst[a] < dst[b]; });
    for (size_t k = 0; k < order.size(); ++k)
      if (k + 1 == order.size() || dst[order[k + 1]] != dst[order[k]]) {
        s->push_back(order[k]);
        d->push_back(dst[order[k]]);
      }
  };
  auto up = [&](const std::vector<int64_t>& v, int64_t** dp) -> int {
    int rc1 = dev_alloc(reinterpret_cast<void**>(dp), v.size() * sizeof(int64_t), &hd->allocs,
                        "create_selection");
    if (rc1) return rc1;
    if (!v.empty() &&
        hipMemcpy(*dp, v.data(), v.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess)
      return fail(BW_EHIP, "create_selection: upload");
    return BW_OK;
  };
  const std::vector<int64_t>* in_of[2] = {&pcol, &prow};
  const std::vector<int64_t>* out_of[2] = {&prow, &pcol};
  for (int dir = 0; dir < 2; ++dir) {
    SelDir& sd = hd->sel.dir[dir];
    std::vector<int64_t> s_src, s_dst;
    pairs(*in_of[dir], &s_src, &s_dst);
    if ((rc = up(s_src, &sd.in_src)) || (rc = up(s_dst, &sd.in_dst)) ||
        (rc = up(*out_of[dir], &sd.out_dst))) {
      bw_hip_destroy(hd);
      return rc;
    }
    sd.n_pairs = static_cast<int64_t>(s_src.size());
    sd.n_in = static_cast<int64_t>(in_of[dir]->size());
    sd.n_out = static_cast<int64_t>(out_of[dir]->size());
  }
  // every stream's scratch holds the full permuted output of the selected
  // blocks; host-pointer staging must hold the selection vectors, which
  // repeated indices can make longer than the full ones
  const int64_t nfull = h->dof * std::max(h->n_row_pts, h->n_col_pts);
  hd->yfull_elems = nfull;
  hd->stage_elems = std::max(nfull, h->dof * std::max(n_rows, n_cols));
  if ((rc = prepare_null_stream(hd))) {
    bw_hip_destroy(hd);
    return rc;
  }
  hd->sel.on = true;
  hd->sel.ordering = ordering;
  hd->n_row_pts = n_rows;
  hd->n_col_pts = n_cols;
  *out = hd;
  return BW_OK;
}

int bw_hip_matvec(bw_hip_handle* hd, const double* x, double* y, int ordering, int precision,
                  int on_device, void* stream) {
  return do_matvec(hd, false, x, y, ordering, precision, on_device, stream);
}

int bw_hip_rmatvec(bw_hip_handle* hd, const double* x, double* y, int ordering, int precision,
                   int on_device, void* stream) {
  return do_matvec(hd, true, x, y, ordering, precision, on_device, stream);
}

}  // extern "C"

namespace {

// Checks of the batched entry points; on success *op is the operator used.
// X(i, j) = X[i * xr + j * xc] (input index i, right-hand side j); Y likewise
// with (yr, yc).  Column-major is (1, ld), row-major (nrhs, 1).
int check_batched(bw_hip_handle* hd, bool adjoint, int nrhs, const double* X, int64_t xr,
                  int64_t xc, const double* Y, int64_t yr, int64_t yc, int ordering, int precision,
                  bool leading_dims, DevOp** op_out) {
  if (!hd || !X || !Y || nrhs <= 0) return fail(BW_EINVAL, "matvec_batched: bad argument");
  if (xr < 1 || xc < 1 || yr < 1 || yc < 1)
    return fail(BW_ESHAPE, "matvec_batched: strides must be positive");
  if (hd->sel.on) return fail(BW_ENOTSUP, "matvec_batched: not available for a selection");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "matvec_batched: bad ordering");
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32)
    return fail(BW_EINVAL, "matvec_batched: bad precision");
  DevOp& op = hd->op[adjoint ? 1 : 0][precision == BW_PREC_FP32 ? 1 : 0];
  if (!op.present)
    return fail(BW_ENOTSUP, adjoint ? "rmatvec_batched: adjoint / precision not built"
                                    : "matvec_batched: precision not built");
  if (leading_dims && (xc < op.n_in || yc < op.n_out))
    return fail(BW_ESHAPE, "matvec_batched: leading dimension too small");
  // the MFMA kernels index the kRhs-wide input rows with 32-bit offsets
  if (op.xt_size * kRhs > static_cast<int64_t>(UINT32_MAX))
    return fail(BW_ESHAPE, "matvec_batched: input plus rank entries exceed 2^32 / 32");
  // distinct outputs must not share an address: columns apart by at least a
  // column's extent, or rows apart by at least a row's
  if (!(yc >= op.n_out * yr || yr >= static_cast<int64_t>(nrhs) * yc))
    return fail(BW_ESHAPE, "matvec_batched: output strides overlap");
  *op_out = &op;
  return BW_OK;
}

// The kernels of Y = H X (kRhs right-hand sides per round) on stream st,
// writing only the batched scratch `s`.
int launch_batched(bw_hip_handle* hd, DevOp& op, const OpScratch& s, int nrhs, const double* X,
                   int64_t xr, int64_t xc, double* Y, int64_t yr, int64_t yc, int ordering,
                   bool fp64, hipStream_t st) {
  const bool orig = ordering == BW_ORDER_ORIGINAL;
  const int dof = hd->dof;
  for (int k0 = 0; k0 < nrhs; k0 += kRhs) {
    const int nk = std::min(kRhs, nrhs - k0);
    const double* Xk = X + static_cast<int64_t>(k0) * xc;
    double* Yk = Y + static_cast<int64_t>(k0) * yc;
    const int64_t rows_per_block = 256 / kRhs;
    const int gblocks = static_cast<int>((op.n_in + rows_per_block - 1) / rows_per_block);
    // row-major input (unit column stride, even row stride, 16-byte aligned
    // rows): the vector gather, 32 rows per block
    const bool rowmajor = xc == 1 && xr % 2 == 0 && reinterpret_cast<uintptr_t>(Xk) % 16 == 0;
    const int gblocks4 = static_cast<int>((op.n_in + 31) / 32);
    const int64_t* pin = orig ? op.perm_in : nullptr;
    const int64_t* pout = orig ? op.perm_out : nullptr;
    if (fp64) {
      double* xb = static_cast<double*>(s.xb);
      double* pb0 = static_cast<double*>(s.partb[0]);
      double* pb1 = static_cast<double*>(s.partb[1]);
      if (rowmajor)
        hipLaunchKernelGGL(k_gather_b4<double>, dim3(gblocks4), dim3(256), 0, st, Xk, xr, nk, xb,
                           op.n_in, pin, dof);
      else
        hipLaunchKernelGGL(k_gather_b<double>, dim3(gblocks), dim3(256), 0, st, Xk, xr, xc, nk, xb,
                           op.n_in, pin, dof);
      if (op.p1.nwi > 0)
        launch_mfma64<1>(op.mfma_schedule, op.p1.blocks, st, op.p1.wis, op.p1.xs, op.p1.gidx,
                         static_cast<const double*>(op.p1.data), xb, xb, nullptr, 0, 0, kRhs,
                         pb0, nullptr, dof, op.p1.rowmap);
      if (op.p1.nred > 0)
        hipLaunchKernelGGL((k_reduce_b<1, double>), dim3(op.p1.nred), dim3(256), 0, st, op.p1.reds,
                           op.p1.nred, pb0, xb, nullptr, 0, 0, kRhs, nullptr, dof, op.p1.rowmap);
      if (op.p2.nwi > 0)
        launch_mfma64<2>(op.mfma_schedule, op.p2.blocks, st, op.p2.wis, op.p2.xs, op.p2.gidx,
                         static_cast<const double*>(op.p2.data), xb, nullptr, Yk, yr, yc, nk,
                         pb1, pout, dof, nullptr);
      if (op.p2.nred > 0)
        hipLaunchKernelGGL((k_reduce_b<2, double>), dim3(op.p2.nred), dim3(256), 0, st, op.p2.reds,
                           op.p2.nred, pb1, nullptr, Yk, yr, yc, nk, pout, dof, nullptr);
    } else {
      float* xb = static_cast<float*>(s.xb);
      float* pb0 = static_cast<float*>(s.partb[0]);
      float* pb1 = static_cast<float*>(s.partb[1]);
      if (rowmajor)
        hipLaunchKernelGGL(k_gather_b4<float>, dim3(gblocks4), dim3(256), 0, st, Xk, xr, nk, xb,
                           op.n_in, pin, dof);
      else
        hipLaunchKernelGGL(k_gather_b<float>, dim3(gblocks), dim3(256), 0, st, Xk, xr, xc, nk, xb,
                           op.n_in, pin, dof);
      if (op.p1.nwi > 0)
        launch_mfma<1>(op.mfma_schedule, op.p1.blocks, st, op.p1.wis, op.p1.xs, op.p1.gidx,
                       static_cast<const float*>(op.p1.data), xb, xb, nullptr, 0, 0, kRhs, pb0,
                       nullptr, dof, op.p1.rowmap);
      if (op.p1.nred > 0)
        hipLaunchKernelGGL((k_reduce_b<1, float>), dim3(op.p1.nred), dim3(256), 0, st, op.p1.reds,
                           op.p1.nred, pb0, xb, nullptr, 0, 0, kRhs, nullptr, dof, op.p1.rowmap);
      if (op.p2.nwi > 0)
        launch_mfma<2>(op.mfma_schedule, op.p2.blocks, st, op.p2.wis, op.p2.xs, op.p2.gidx,
                       static_cast<const float*>(op.p2.data), xb, nullptr, Yk, yr, yc, nk, pb1, pout,
                       dof, nullptr);
      if (op.p2.nred > 0)
        hipLaunchKernelGGL((k_reduce_b<2, float>), dim3(op.p2.nred), dim3(256), 0, st, op.p2.reds,
                           op.p2.nred, pb1, nullptr, Yk, yr, yc, nk, pout, dof, nullptr);
    }
    HIPCHK(hipGetLastError());
  }
  return BW_OK;
}

int batched_impl(bw_hip_handle* hd, bool adjoint, int nrhs, const double* X, int64_t xr,
                 int64_t xc, double* Y, int64_t yr, int64_t yc, int ordering, int precision,
                 void* stream, bool leading_dims = false) {
  DevOp* op = nullptr;
  int rc = check_batched(hd, adjoint, nrhs, X, xr, xc, Y, yr, yc, ordering, precision,
                         leading_dims, &op);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(hd->mu);
  arm_alloc_hook();
  HIPCHK(hipSetDevice(hd->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int adj = adjoint ? 1 : 0, f32 = precision == BW_PREC_FP32 ? 1 : 0;
  StreamCtx* ctx = stream_ctx(hd, st);
  if ((rc = ensure_op_scratch(hd, ctx, adj, f32, true))) return rc;
  return launch_batched(hd, *op, ctx->op[adj][f32], nrhs, X, xr, xc, Y, yr, yc, ordering, !f32, st);
}

// Captures launch(stream, ctx) on a private stream into a new graph that owns
// ctx (already holding its scratch).  Caller holds hd->mu.
template <typename F>
int capture_graph(bw_hip_handle* hd, std::unique_ptr<bw_hip_graph> g, F launch,
                  bw_hip_graph** out) {
  hipStream_t cs;
  const hipError_t ec = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (ec != hipSuccess) {
    free_ctx(&g->ctx);
    return fail(BW_EHIP, std::string("graph_create: ") + hipGetErrorString(ec));
  }
  const bool timing = hd->timing;
  hd->timing = false;  // no event records inside the graph
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = BW_OK;
  hipGraph_t graph = nullptr;
  if (e == hipSuccess) {
    rc = launch(cs, g->ctx);
    hipError_t e2 = hipStreamEndCapture(cs, &graph);
    if (!rc && e2 != hipSuccess) rc = fail(BW_EHIP, std::string("graph_create: ") + hipGetErrorString(e2));
  } else {
    rc = fail(BW_EHIP, std::string("graph_create: ") + hipGetErrorString(e));
  }
  hd->timing = timing;
  (void)hipStreamDestroy(cs);
  if (!rc) {
    g->graph = graph;
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = fail(BW_EHIP, std::string("graph_create: instantiate: ") + hipGetErrorString(e));
  }
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    free_ctx(&g->ctx);
    return rc;
  }
  *out = g.release();
  return BW_OK;
}

}  // namespace

extern "C" {

int bw_hip_graph_create(bw_hip_handle* hd, const double* x, double* y, int ordering,
                        int precision, int adjoint, bw_hip_graph** out) {
  if (!hd || !x || !y || !out) return fail(BW_EINVAL, "graph_create: null argument");
  *out = nullptr;
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "graph_create: ordering must be 0 (permuted) or 1 (original)");
  if (precision != BW_PREC_FP64 && precision != BW_PREC_FP32)
    return fail(BW_EINVAL, "graph_create: precision must be BW_PREC_FP64 or BW_PREC_FP32");
  const int adj = adjoint ? 1 : 0, f32 = precision == BW_PREC_FP32 ? 1 : 0;
  DevOp& op = hd->op[adj][f32];
  if (!op.present) return fail(BW_ENOTSUP, "graph_create: adjoint / precision not built");
  const SelDir* sd = hd->sel.on ? &hd->sel.dir[adj] : nullptr;
  if (sd && ordering != hd->sel.ordering)
    return fail(BW_ENOTSUP, "graph_create: this selection was created for the other ordering");
  std::lock_guard<std::mutex> lock(hd->mu);
  arm_alloc_hook();
  HIPCHK(hipSetDevice(hd->device));
  std::unique_ptr<bw_hip_graph> g(new bw_hip_graph());
  g->device = hd->device;
  int rc;
  if ((rc = ensure_op_scratch(hd, &g->ctx, adj, f32, false)) ||
      (rc = ensure_ctx_vectors(hd, &g->ctx, false))) {
    free_ctx(&g->ctx);
    return rc;
  }
  return capture_graph(hd, std::move(g), [&](hipStream_t cs, StreamCtx& ctx) {
    const OpScratch& s = ctx.op[adj][f32];
    return f32 ? launch_matvec<float>(hd, op, s, ctx.y_full, x, y, ordering, cs, sd)
               : launch_matvec<double>(hd, op, s, ctx.y_full, x, y, ordering, cs, sd);
  }, out);
}

int bw_hip_graph_create_batched(bw_hip_handle* hd, int adjoint, int nrhs, const double* X,
                                int64_t x_row_stride, int64_t x_col_stride, double* Y,
                                int64_t y_row_stride, int64_t y_col_stride, int ordering,
                                int precision, bw_hip_graph** out) {
  if (!out) return fail(BW_EINVAL, "graph_create_batched: null argument");
  *out = nullptr;
  DevOp* op = nullptr;
  int rc = check_batched(hd, adjoint != 0, nrhs, X, x_row_stride, x_col_stride, Y, y_row_stride,
                         y_col_stride, ordering, precision, false, &op);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(hd->mu);
  arm_alloc_hook();
  HIPCHK(hipSetDevice(hd->device));
  const int adj = adjoint ? 1 : 0, f32 = precision == BW_PREC_FP32 ? 1 : 0;
  std::unique_ptr<bw_hip_graph> g(new bw_hip_graph());
  g->device = hd->device;
  if ((rc = ensure_op_scratch(hd, &g->ctx, adj, f32, true))) {
    free_ctx(&g->ctx);
    return rc;
  }
  return capture_graph(hd, std::move(g), [&](hipStream_t cs, StreamCtx& ctx) {
    return launch_batched(hd, *op, ctx.op[adj][f32], nrhs, X, x_row_stride, x_col_stride, Y,
                          y_row_stride, y_col_stride, ordering, !f32, cs);
  }, out);
}

int bw_hip_graph_launch(bw_hip_graph* g, void* stream) {
  if (!g) return fail(BW_EINVAL, "graph_launch: null graph");
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return BW_OK;
}

void bw_hip_graph_destroy(bw_hip_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  (void)hipDeviceSynchronize();
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  free_ctx(&g->ctx);
  delete g;
}

int bw_hip_stream_release(bw_hip_handle* hd, void* stream) {
  if (!hd) return fail(BW_EINVAL, "stream_release: null handle");
  std::lock_guard<std::mutex> lock(hd->mu);
  HIPCHK(hipSetDevice(hd->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (size_t i = 0; i < hd->ctxs.size(); ++i) {
    if (hd->ctxs[i]->stream != st) continue;
    HIPCHK(hipStreamSynchronize(st));
    free_ctx(hd->ctxs[i].get());
    hd->ctxs.erase(hd->ctxs.begin() + static_cast<std::ptrdiff_t>(i));
    return BW_OK;
  }
  return BW_OK;
}

int bw_hip_matvec_batched(bw_hip_handle* hd, int nrhs, const double* X, int64_t ldx, double* Y,
                          int64_t ldy, int ordering, int precision, void* stream) {
  return batched_impl(hd, false, nrhs, X, 1, ldx, Y, 1, ldy, ordering, precision, stream, true);
}

int bw_hip_rmatvec_batched(bw_hip_handle* hd, int nrhs, const double* X, int64_t ldx, double* Y,
                           int64_t ldy, int ordering, int precision, void* stream) {
  return batched_impl(hd, true, nrhs, X, 1, ldx, Y, 1, ldy, ordering, precision, stream, true);
}

int bw_hip_matvec_batched_strided(bw_hip_handle* hd, int adjoint, int nrhs, const double* X,
                                  int64_t x_row_stride, int64_t x_col_stride, double* Y,
                                  int64_t y_row_stride, int64_t y_col_stride, int ordering,
                                  int precision, void* stream) {
  return batched_impl(hd, adjoint != 0, nrhs, X, x_row_stride, x_col_stride, Y, y_row_stride,
                      y_col_stride, ordering, precision, stream);
}

int bw_hip_diagonal(bw_hip_handle* hd, double* dg, int ordering) {
  if (!hd || !dg) return fail(BW_EINVAL, "diagonal: null argument");
  if (hd->sel.on) return fail(BW_ENOTSUP, "diagonal: not defined for a selection");
  if (ordering != BW_ORDER_PERMUTED && ordering != BW_ORDER_ORIGINAL)
    return fail(BW_EINVAL, "diagonal: bad ordering");
  const auto& v = ordering == BW_ORDER_ORIGINAL ? hd->diag_orig : hd->diag_perm;
  std::memcpy(dg, v.data(), v.size() * sizeof(double));
  return BW_OK;
}

int bw_hip_sizes(bw_hip_handle* hd, int64_t* n_rows, int64_t* n_cols, int64_t* nb_entries,
                 int64_t* device_bytes, int64_t* sb64, int64_t* sb32) {
  if (!hd) return fail(BW_EINVAL, "sizes: null handle");
  if (n_rows) *n_rows = hd->dof * hd->n_row_pts;
  if (n_cols) *n_cols = hd->dof * hd->n_col_pts;
  if (nb_entries) *nb_entries = hd->nb_entries;
  if (device_bytes) *device_bytes = hd->device_bytes;
  if (sb64) *sb64 = hd->op[0][0].present ? hd->op[0][0].stream_bytes : 0;
  if (sb32) *sb32 = hd->op[0][1].present ? hd->op[0][1].stream_bytes : 0;
  return BW_OK;
}

int bw_hip_set_unroll(bw_hip_handle* hd, int unroll_fp64, int unroll_fp32) {
  if (!hd) return fail(BW_EINVAL, "set_unroll: null handle");
  std::lock_guard<std::mutex> lock(hd->mu);
  for (int u : {unroll_fp64, unroll_fp32})
    if (u != 0 && u != 4 && u != 8 && u != 16)
      return fail(BW_EINVAL, "set_unroll: unroll must be 4, 8 or 16 (0 keeps the current one)");
  for (int adj = 0; adj < 2; ++adj)
    for (int pass = 0; pass < 2; ++pass) {
      if (unroll_fp64) hd->op[adj][0].unroll[pass] = unroll_fp64;
      if (unroll_fp32) hd->op[adj][1].unroll[pass] = unroll_fp32;
    }
  return BW_OK;
}

int bw_hip_set_unroll_pass(bw_hip_handle* hd, int precision_mask, int pass, int unroll) {
  if (!hd) return fail(BW_EINVAL, "set_unroll_pass: null handle");
  std::lock_guard<std::mutex> lock(hd->mu);
  if (pass != 1 && pass != 2) return fail(BW_EINVAL, "set_unroll_pass: pass must be 1 or 2");
  if (unroll != 4 && unroll != 8 && unroll != 16)
    return fail(BW_EINVAL, "set_unroll_pass: unroll must be 4, 8 or 16");
  if (precision_mask <= 0 || (precision_mask & ~(BW_PREC_FP64 | BW_PREC_FP32)))
    return fail(BW_EINVAL, "set_unroll_pass: bad precision mask");
  for (int adj = 0; adj < 2; ++adj) {
    if (precision_mask & BW_PREC_FP64) hd->op[adj][0].unroll[pass - 1] = unroll;
    if (precision_mask & BW_PREC_FP32) hd->op[adj][1].unroll[pass - 1] = unroll;
  }
  return BW_OK;
}

int bw_hip_set_mfma_schedule(bw_hip_handle* hd, int precision_mask, int schedule) {
  if (!hd) return fail(BW_EINVAL, "set_mfma_schedule: null handle");
  std::lock_guard<std::mutex> lock(hd->mu);
  if (schedule < 0 || schedule > 15) return fail(BW_EINVAL, "set_mfma_schedule: schedule must be 0-15");
  if (precision_mask <= 0 || (precision_mask & ~(BW_PREC_FP64 | BW_PREC_FP32)))
    return fail(BW_EINVAL, "set_mfma_schedule: bad precision mask");
  for (int adj = 0; adj < 2; ++adj) {
    if (precision_mask & BW_PREC_FP64) hd->op[adj][0].mfma_schedule = schedule;
    if (precision_mask & BW_PREC_FP32) hd->op[adj][1].mfma_schedule = schedule;
  }
  return BW_OK;
}

int bw_hip_set_timing(bw_hip_handle* hd, int enable) {
  if (!hd) return fail(BW_EINVAL, "set_timing: null handle");
  std::lock_guard<std::mutex> lock(hd->mu);
  for (int s = 0; s < 2; ++s)
    if (hd->pending[s]) HIPCHK(hipEventSynchronize(hd->evs[s][5]));
  hd->pending[0] = hd->pending[1] = false;
  hd->cur = 0;
  for (double& v : hd->sum_ms) v = 0.0;
  hd->n_timed = 0;
  hd->timing = enable != 0;
  return BW_OK;
}

int bw_hip_timing_stats(bw_hip_handle* hd, double* avg_ms6, int64_t* count) {
  if (!hd || !avg_ms6) return fail(BW_EINVAL, "timing_stats: null argument");
  std::lock_guard<std::mutex> lock(hd->mu);
  for (int s = 0; s < 2; ++s) {
    const int set = (hd->cur + s) & 1;  // oldest first
    if (hd->pending[set]) {
      int rc = accumulate_timing(hd, set);
      if (rc) return rc;
    }
  }
  for (int i = 0; i < 6; ++i) avg_ms6[i] = hd->n_timed ? hd->sum_ms[i] / hd->n_timed : 0.0;
  if (count) *count = hd->n_timed;
  return BW_OK;
}

}  // extern "C"
