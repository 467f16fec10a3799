// dvr_syev_batch_noalloc: the batched symmetric eigensolver of
// dvr_syev_batch for callers that cannot allocate device memory (code
// under HIP graph capture, tight loops, harnesses that forbid hipMalloc).
//
// Every device buffer of the call -- the staged matrices, the eigenvalues
// and eigenvectors before they are copied out, the inverse-iteration
// workspace -- lives in one file-scope device array, `scratch`, of
// kScratchDoubles doubles (32 MiB), which each GPU the library runs on holds
// until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress, on the first call on that device) and only that
// address is used on the host side.  The call runs on the null stream and
// allocates nothing; calls that overlap on one GPU share the array and race.
//
// The problems are solved by the LDS-resident single-work-group solver
// (k_syev_lds.hip: Householder tridiagonalisation, Sturm multisection,
// inverse iteration, back-transformation), which holds each matrix in LDS:
// every order must be <= 128.
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <cstdio>
#include <mutex>

namespace dvr {

namespace {

constexpr size_t kScratchDoubles = (size_t)32 << 17;   // 4 Mi doubles = 32 MiB
constexpr int kMaxOrder = 128;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once).
double* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    DVR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[dev] = static_cast<double*>(p);
  }
  return g_ptr[dev];
}

// Upper triangle of nb consecutive n x n column-major matrices from their
// lower triangle.
__global__ void __launch_bounds__(256)
k_mirror_lower_batched(double* __restrict__ A, int n, int nb) {
  const long nn = (long)n * n;
  double* M = A + blockIdx.y * nn;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < nn; t += (long)gridDim.x * 256) {
    const int i = (int)(t % n), j = (int)(t / n);
    if (i < j) M[t] = M[j + (long)i * n];
  }
}

// Doubles of scratch a run of nb problems of order n, m vectors each, needs:
// matrices, eigenvalues, vectors, inverse-iteration workspace.
size_t run_doubles(int n, int m, int nb) {
  return (size_t)nb * ((size_t)n * n + n + (size_t)n * m) + syev_lds_scratch_doubles(n, nb);
}

}  // namespace
}  // namespace dvr

using namespace dvr;

extern "C" int dvr_syev_batch_noalloc(int32_t nprob, const int32_t* n, const double* a,
                                      const int32_t* nvec, double* w, double* z) {
  if (nprob <= 0 || !n || !a || !nvec || !w) return DVR_ERR_ARG;
  for (int p = 0; p < nprob; ++p) {
    if (n[p] <= 0 || nvec[p] < 0 || nvec[p] > n[p]) return DVR_ERR_ARG;
    if (n[p] > kMaxOrder) {
      fprintf(stderr, "dvr_syev_batch_noalloc: order %d of problem %d exceeds %d\n", n[p], p,
              kMaxOrder);
      return DVR_ERR_ARG;
    }
  }
  try {
    int dev = 0;
    DVR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return DVR_ERR_ARG;
    // runs of consecutive problems of one order go through one launch each;
    // all of them must fit in the scratch array at once
    size_t need = 0;
    for (int p0 = 0; p0 < nprob;) {
      int p1 = p0, m = 0;
      while (p1 < nprob && n[p1] == n[p0]) m = std::max(m, (int)nvec[p1++]);
      need += run_doubles(n[p0], m, p1 - p0);
      p0 = p1;
    }
    if (need > kScratchDoubles) {
      fprintf(stderr,
              "dvr_syev_batch_noalloc: the problems need %zu doubles of scratch, the maximum is "
              "%zu (%zu MiB)\n",
              need, kScratchDoubles, kScratchDoubles * 8 >> 20);
      return DVR_ERR_ARG;
    }
    double* base = scratch_ptr(dev);
    hipStream_t s = nullptr;   // the null stream: nothing to create
    size_t off = 0;
    long ain = 0, wo = 0, zo = 0;
    for (int p0 = 0; p0 < nprob;) {
      const int np = n[p0];
      int p1 = p0, m = 0;
      while (p1 < nprob && n[p1] == np) m = std::max(m, (int)nvec[p1++]);
      const int nb = p1 - p0;
      const long nn = (long)np * np;
      double* A = base + off;
      double* ev = A + (size_t)nb * nn;
      double* Z = ev + (size_t)nb * np;
      double* ws = Z + (size_t)nb * np * m;
      off += run_doubles(np, m, nb);
      DVR_HIP(hipMemcpyAsync(A, a + ain, sizeof(double) * nb * nn, hipMemcpyHostToDevice, s));
      ain += nb * nn;
      hipLaunchKernelGGL(k_mirror_lower_batched, dim3((unsigned)std::min<long>((nn + 255) / 256, 64), nb),
                         dim3(256), 0, s, A, np, nb);
      DVR_LAUNCH_CHECK();
      launch_syev_lds(s, A, nn, np, np, nb, nullptr, ev, np, Z, (long)np * m, nullptr, m, 0.0, ws);
      for (int p = p0; p < p1; ++p) {
        const int q = p - p0;
        if (nvec[p] > 0) {
          DVR_HIP(hipMemcpyAsync(w + wo, ev + (size_t)q * np, sizeof(double) * nvec[p],
                                 hipMemcpyDeviceToHost, s));
          if (z)
            DVR_HIP(hipMemcpyAsync(z + zo, Z + (size_t)q * np * m,
                                   sizeof(double) * (size_t)np * nvec[p], hipMemcpyDeviceToHost,
                                   s));
        }
        wo += nvec[p];
        zo += (long)np * nvec[p];
      }
      p0 = p1;
    }
    DVR_HIP(hipStreamSynchronize(s));
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_syev_batch_noalloc: %s\n", e.what());
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
