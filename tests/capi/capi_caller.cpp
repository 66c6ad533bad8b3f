// A C++ caller of the engine through its C-ABI only, the way the reference's
// C++ host code would call it from an Hmat adapter or BigWhamIO::MatVec
// (src/io/bigwham_io.cc:919-980): build the H-matrix on the host
// (libbwhost), upload it once (bw_hip_create), release the host copies of the
// low-rank factors as the adapter does, then run the calls the facade makes
// after the build (BigWhamIO::BuildHierarchicalMatrix -> compressionRatio ->
// nbOfEntries, bigwham_io.cc:674; GetFullBlocks, bigwham_io.cc:788-905;
// GetDiagonal) and apply the operator through host pointers, device pointers
// on a stream, in FP32, transposed and batched, and exercise the error path.
// No Python in the loop.
//
//   capi_caller <kernel> <mesh.bin> <out_dir>
//
// mesh.bin: int64 n_nodes, dim, n_elts, n_vertex, then coor (double,
// n_nodes x dim) and conn (int64, n_elts x n_vertex).  The program writes
// raw little-endian doubles to out_dir: x, y64, y32, yt, diag, ybatch
// (n x 3, column-major), fb_val (full-block values), and int64 arrays
// entries (nb_entries, full-block count), fb_row, fb_col.
// tests/test_gpu_capi.py compares them with the oracle.  Exit status 0 on
// success; any failed call prints its message.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bigwham_hip.h"
#include "bigwham_host.h"

namespace {

int fail(const char* what, int rc, const char* msg) {
  std::fprintf(stderr, "capi_caller: %s returned %d: %s\n", what, rc, msg ? msg : "");
  return 1;
}

#define HOST(call)                                                  \
  do {                                                              \
    int rc_ = (call);                                               \
    if (rc_ != 0) return fail(#call, rc_, bwh_last_error());        \
  } while (0)
#define ENG(call)                                                   \
  do {                                                              \
    int rc_ = (call);                                               \
    if (rc_ != BW_OK) return fail(#call, rc_, bw_hip_last_error()); \
  } while (0)
#define RT(call)                                                    \
  do {                                                              \
    hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) return fail(#call, static_cast<int>(e_), hipGetErrorString(e_)); \
  } while (0)

template <typename T>
bool write(const std::string& path, const std::vector<T>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
  return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: capi_caller <kernel> <mesh.bin> <out_dir>\n");
    return 2;
  }
  const std::string kernel = argv[1], out = argv[3];
  FILE* f = std::fopen(argv[2], "rb");
  if (!f) return fail("fopen", -1, argv[2]);
  int64_t hdr[4];
  if (std::fread(hdr, sizeof(int64_t), 4, f) != 4) return fail("fread", -1, "header");
  std::vector<double> coor(hdr[0] * hdr[1]);
  std::vector<int64_t> conn(hdr[2] * hdr[3]);
  if (std::fread(coor.data(), sizeof(double), coor.size(), f) != coor.size() ||
      std::fread(conn.data(), sizeof(int64_t), conn.size(), f) != conn.size())
    return fail("fread", -1, "mesh body");
  std::fclose(f);

  // host construction (BigWhamIO constructor + BuildHierarchicalMatrix)
  const double props[2] = {1.0, 0.25};
  bwh_hmat* h = nullptr;
  HOST(bwh_create(kernel.c_str(), coor.data(), hdr[0], conn.data(), hdr[2], props, 2, 0, 0, 0, &h));
  HOST(bwh_build_hmat(h, 64, 3.0, 1e-6));
  bw_flat_hmat flat;
  HOST(bwh_flat(h, &flat));

  // The adapter hands over the reference's own block storage.  Here the
  // low-rank factors are caller-owned copies, so they can be released after
  // the upload, as the adapter releases LowRank::A / B.
  std::vector<std::vector<double>> lr_a(flat.n_lr), lr_b(flat.n_lr);
  std::vector<const double*> pa(flat.n_lr), pb(flat.n_lr);
  for (int64_t k = 0; k < flat.n_lr; ++k) {
    const int64_t* c = flat.lr_pattern + 6 * k;
    const int64_t r = c[5];
    const int64_t m = flat.dof * (c[3] - c[1]), nc = flat.dof * (c[4] - c[2]);
    lr_a[k].assign(flat.lr_A[k], flat.lr_A[k] + m * r);
    lr_b[k].assign(flat.lr_B[k], flat.lr_B[k] + r * nc);
    pa[k] = lr_a[k].data();
    pb[k] = lr_b[k].data();
  }
  bw_flat_hmat held = flat;
  held.lr_A = pa.data();
  held.lr_B = pb.data();

  // one upload, both precisions and the adjoint layout
  bw_hip_handle* hd = nullptr;
  ENG(bw_hip_create(&held, BW_PREC_FP64 | BW_PREC_FP32, BW_FLAG_ADJOINT, 0, &hd));

  // release the host low-rank payloads (poisoned first, so any later read of
  // them would show in the results), keep the full-rank blocks
  for (int64_t k = 0; k < flat.n_lr; ++k) {
    std::fill(lr_a[k].begin(), lr_a[k].end(), std::nan(""));
    std::fill(lr_b[k].begin(), lr_b[k].end(), std::nan(""));
    std::vector<double>().swap(lr_a[k]);
    std::vector<double>().swap(lr_b[k]);
  }
  held.lr_A = nullptr;
  held.lr_B = nullptr;

  // compressionRatio() -> nbOfEntries(), as BuildHierarchicalMatrix calls it
  int64_t n_rows, n_cols, entries, dev_bytes, s64, s32;
  ENG(bw_hip_sizes(hd, &n_rows, &n_cols, &entries, &dev_bytes, &s64, &s32));
  // GetFullBlocks -> fullBlocksOriginal: full-rank entries from the kept blocks
  int64_t nfb = 0;
  ENG(bw_flat_full_blocks(&held, BW_ORDER_ORIGINAL, nullptr, 0, nullptr, 0, &nfb, nullptr,
                          nullptr, nullptr));
  std::vector<double> fb_val(nfb);
  std::vector<int64_t> fb_row(nfb), fb_col(nfb);
  ENG(bw_flat_full_blocks(&held, BW_ORDER_ORIGINAL, nullptr, 0, nullptr, 0, &nfb, fb_val.data(),
                          fb_row.data(), fb_col.data()));
  const int64_t n = n_rows;
  std::printf("capi_caller: %s n=%lld entries=%lld device bytes=%lld\n", kernel.c_str(),
              static_cast<long long>(n), static_cast<long long>(entries),
              static_cast<long long>(dev_bytes));

  // the diagonal has dof * max(n_row_pts, n_col_pts) entries (n for this square operator)
  std::vector<double> x(n), y64(n), y32(n), yt(n), yd(n), diag(std::max(n_rows, n_cols));
  uint64_t s = 0x9E3779B97F4A7C15ull;  // x in [-1, 1) from a 64-bit LCG
  for (auto& v : x) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    v = static_cast<double>(s >> 11) * 0x1.0p-52 - 1.0;
  }

  // BigWhamIO::MatVec(ArrayView): host vectors, original ordering
  ENG(bw_hip_matvec(hd, x.data(), y64.data(), BW_ORDER_ORIGINAL, BW_PREC_FP64, 0, nullptr));
  ENG(bw_hip_matvec(hd, x.data(), y32.data(), BW_ORDER_ORIGINAL, BW_PREC_FP32, 0, nullptr));
  ENG(bw_hip_rmatvec(hd, x.data(), yt.data(), BW_ORDER_ORIGINAL, BW_PREC_FP64, 0, nullptr));
  ENG(bw_hip_diagonal(hd, diag.data(), BW_ORDER_ORIGINAL));

  // BigWhamIO::MatVec(double* x, double* y): device pointers, asynchronous on a stream
  hipStream_t st;
  RT(hipStreamCreate(&st));
  double *dx = nullptr, *dy = nullptr, *dX = nullptr, *dY = nullptr;
  RT(hipMalloc(&dx, n * sizeof(double)));
  RT(hipMalloc(&dy, n * sizeof(double)));
  RT(hipMemcpyAsync(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  ENG(bw_hip_matvec(hd, dx, dy, BW_ORDER_ORIGINAL, BW_PREC_FP64, 1, st));
  RT(hipMemcpyAsync(yd.data(), dy, n * sizeof(double), hipMemcpyDeviceToHost, st));
  RT(hipStreamSynchronize(st));
  if (std::memcmp(yd.data(), y64.data(), n * sizeof(double)) != 0)
    return fail("device-pointer matvec", -1, "differs from the host-pointer result");

  // three right-hand sides (x, -x, 2x) column-major through the batched entry point
  const int k = 3;
  std::vector<double> X(n * k), Yb(n * k);
  for (int64_t i = 0; i < n; ++i) {
    X[i] = x[i];
    X[n + i] = -x[i];
    X[2 * n + i] = 2.0 * x[i];
  }
  RT(hipMalloc(&dX, n * k * sizeof(double)));
  RT(hipMalloc(&dY, n * k * sizeof(double)));
  RT(hipMemcpyAsync(dX, X.data(), n * k * sizeof(double), hipMemcpyHostToDevice, st));
  ENG(bw_hip_matvec_batched(hd, k, dX, n, dY, n, BW_ORDER_ORIGINAL, BW_PREC_FP64, st));
  RT(hipMemcpyAsync(Yb.data(), dY, n * k * sizeof(double), hipMemcpyDeviceToHost, st));
  RT(hipStreamSynchronize(st));

  // errors come back as codes with a message; the process keeps running
  const int bad = bw_hip_matvec(hd, dx, dy, 7, BW_PREC_FP64, 1, st);
  if (bad != BW_EINVAL || std::strlen(bw_hip_last_error()) == 0)
    return fail("bad ordering", bad, "expected BW_EINVAL with a message");
  const int small = bw_hip_matvec_batched(hd, k, dX, n - 1, dY, n, BW_ORDER_ORIGINAL,
                                          BW_PREC_FP64, st);
  if (small != BW_ESHAPE) return fail("short leading dimension", small, "expected BW_ESHAPE");

  RT(hipFree(dx));
  RT(hipFree(dy));
  RT(hipFree(dX));
  RT(hipFree(dY));
  RT(hipStreamDestroy(st));
  bw_hip_destroy(hd);
  bwh_free(h);

  const std::vector<int64_t> counts = {entries, nfb};
  if (!write(out + "/x.bin", x) || !write(out + "/y64.bin", y64) || !write(out + "/y32.bin", y32) ||
      !write(out + "/yt.bin", yt) || !write(out + "/diag.bin", diag) ||
      !write(out + "/ybatch.bin", Yb) || !write(out + "/fb_val.bin", fb_val) ||
      !write(out + "/fb_row.bin", fb_row) || !write(out + "/fb_col.bin", fb_col) ||
      !write(out + "/entries.bin", counts))
    return fail("write", -1, out.c_str());
  std::printf("capi_caller: ok\n");
  return 0;
}
