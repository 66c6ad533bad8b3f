// A C++ caller of the engine through its C-ABI only, the way the reference's
// C++ host code would call it from an Hmat adapter or BigWhamIO::MatVec
// (src/io/bigwham_io.cc:919-980): build the H-matrix on the host
// (libbwhost), upload it once (bw_hip_create), then apply it through host
// pointers, device pointers on a stream, in FP32, transposed and batched, and
// exercise the error path.  No Python in the loop.
//
//   capi_caller <kernel> <mesh.bin> <out_dir>
//
// mesh.bin: int64 n_nodes, dim, n_elts, n_vertex, then coor (double,
// n_nodes x dim) and conn (int64, n_elts x n_vertex).  The program writes
// raw little-endian doubles to out_dir: x, y64, y32, yt, diag, ybatch
// (n x 3, column-major).  tests/test_gpu_capi.py compares them with the
// oracle.  Exit status 0 on success; any failed call prints its message.
#include <hip/hip_runtime_api.h>

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

bool write(const std::string& path, const std::vector<double>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size();
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

  // one upload, both precisions and the adjoint layout
  bw_hip_handle* hd = nullptr;
  ENG(bw_hip_create(&flat, BW_PREC_FP64 | BW_PREC_FP32, BW_FLAG_ADJOINT, 0, &hd));
  int64_t n_rows, n_cols, entries, dev_bytes, s64, s32;
  ENG(bw_hip_sizes(hd, &n_rows, &n_cols, &entries, &dev_bytes, &s64, &s32));
  const int64_t n = n_rows;
  std::printf("capi_caller: %s n=%lld entries=%lld device bytes=%lld\n", kernel.c_str(),
              static_cast<long long>(n), static_cast<long long>(entries),
              static_cast<long long>(dev_bytes));

  std::vector<double> x(n), y64(n), y32(n), yt(n), yd(n), diag(n);
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

  if (!write(out + "/x.bin", x) || !write(out + "/y64.bin", y64) || !write(out + "/y32.bin", y32) ||
      !write(out + "/yt.bin", yt) || !write(out + "/diag.bin", diag) ||
      !write(out + "/ybatch.bin", Yb))
    return fail("write", -1, out.c_str());
  std::printf("capi_caller: ok\n");
  return 0;
}
