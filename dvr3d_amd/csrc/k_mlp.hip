// Kernels of the learned Davidson preconditioner: a 3-layer perceptron
// (NF = 6 inputs, two tanh layers of width 128, linear output) evaluated per
// (basis function, Ritz pair) row.
//
//   h1 = tanh(W1^T f + b1),  h2 = tanh(W2^T h1 + b2),  w = W3^T h2 + b3
//
// Training (FP32): k_mlp_fwd keeps h1 and h2 (feature-major: element o of row
// r at o*R + r), k_mlp_bwd_rows turns dL/dw into the layer gradients
// ga2 = dL/da2 and ga1 = dL/da1 per row, k_mlp_wgrad reduces the weight
// gradients over rows (deterministic two-pass sums), k_adam updates.
// Inference: k_mlp_infer_bf16 runs the 128 x 128 layer on BF16 MFMA
// (v_mfma_f32_16x16x16bf16_1k) with FP32 accumulation.
//
// Parameter layout (float): W1[f*128 + o], b1[o], W2[k*128 + o], b2[o], W3[o],
// b3.
#include "dvr_internal.h"

#include <hip/hip_bf16.h>

namespace dvr {

namespace {
constexpr int NF = MLP_NF;
constexpr int WD = MLP_WIDTH;
constexpr int OFF_B1 = NF * WD;
constexpr int OFF_W2 = OFF_B1 + WD;
constexpr int OFF_B2 = OFF_W2 + WD * WD;
constexpr int OFF_W3 = OFF_B2 + WD;
constexpr int OFF_B3 = OFF_W3 + WD;
static_assert(OFF_B3 + 1 == MLP_NPARAM, "parameter layout");

__device__ inline float tanh_f(float x) { return tanhf(x); }
}  // namespace

// one thread per row; all parameters staged in LDS (dynamic, MLP_NPARAM floats)
__global__ void __launch_bounds__(256)
k_mlp_fwd(const float* __restrict__ P, long R, const float* __restrict__ F,
          float* __restrict__ W, float* __restrict__ H1, float* __restrict__ H2) {
  extern __shared__ float sp[];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += 256) sp[i] = P[i];
  __syncthreads();
  const long r = blockIdx.x * 256L + threadIdx.x;
  if (r >= R) return;
  float f[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) f[q] = F[r * NF + q];
  float h1[WD];
#pragma unroll
  for (int o = 0; o < WD; ++o) {
    float a = sp[OFF_B1 + o];
#pragma unroll
    for (int q = 0; q < NF; ++q) a += f[q] * sp[q * WD + o];
    h1[o] = tanh_f(a);
    if (H1) H1[(long)o * R + r] = h1[o];
  }
  float z = sp[OFF_B3];
  for (int o = 0; o < WD; ++o) {
    float a = sp[OFF_B2 + o];
#pragma unroll
    for (int k = 0; k < WD; ++k) a += h1[k] * sp[OFF_W2 + k * WD + o];
    const float h = tanh_f(a);
    if (H2) H2[(long)o * R + r] = h;
    z += h * sp[OFF_W3 + o];
  }
  W[r] = z;
}

// per row: ga2[o] = gw W3[o] (1 - h2[o]^2); ga1[k] = (sum_o W2[k,o] ga2[o]) (1 - h1[k]^2)
__global__ void __launch_bounds__(256)
k_mlp_bwd_rows(const float* __restrict__ P, long R, const float* __restrict__ H1,
               const float* __restrict__ H2, const float* __restrict__ GW,
               float* __restrict__ GA1, float* __restrict__ GA2) {
  extern __shared__ float sp[];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += 256) sp[i] = P[i];
  __syncthreads();
  const long r = blockIdx.x * 256L + threadIdx.x;
  if (r >= R) return;
  const float gw = GW[r];
  float ga2[WD];
#pragma unroll
  for (int o = 0; o < WD; ++o) {
    const float h = H2[(long)o * R + r];
    ga2[o] = gw * sp[OFF_W3 + o] * (1.0f - h * h);
    GA2[(long)o * R + r] = ga2[o];
  }
  for (int k = 0; k < WD; ++k) {
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < WD; ++o) s += sp[OFF_W2 + k * WD + o] * ga2[o];
    const float h = H1[(long)k * R + r];
    GA1[(long)k * R + r] = s * (1.0f - h * h);
  }
}

// partial[c][p][q] = sum over rows r of chunk c of X[p*R + r] * G[q*R + r]
// (X: np rows-vectors, G: nq row-vectors; xrow: X element (p, r) is at
// X[r*NF + p] instead, for the feature matrix).  One work-group per (chunk,
// 16 x 16 output block); threads accumulate one (p, q) each.
__global__ void __launch_bounds__(256)
k_mlp_wgrad(const float* __restrict__ X, int np, bool xrow, const float* __restrict__ G, int nq,
            long R, long chunk, float* __restrict__ partial) {
  const int pb = blockIdx.y * 16, qb = blockIdx.z * 16;
  const int tp = threadIdx.x >> 4, tq = threadIdx.x & 15;
  const int p = pb + tp, q = qb + tq;
  const long r0 = blockIdx.x * chunk, r1 = min(R, r0 + chunk);
  __shared__ float xs[16][65];
  __shared__ float gs[16][65];
  float acc = 0.0f;
  for (long rb = r0; rb < r1; rb += 64) {
    for (int i = threadIdx.x; i < 16 * 64; i += 256) {
      const int a = i >> 6, rr = i & 63;
      const long r = rb + rr;
      const int pp = pb + a, qq = qb + a;
      xs[a][rr] = (r < r1 && pp < np) ? (xrow ? X[r * NF + pp] : X[(long)pp * R + r]) : 0.0f;
      gs[a][rr] = (r < r1 && qq < nq) ? G[(long)qq * R + r] : 0.0f;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < 64; ++rr) acc += xs[tp][rr] * gs[tq][rr];
    __syncthreads();
  }
  if (p < np && q < nq) partial[((long)blockIdx.x * np + p) * nq + q] = acc;
}

// out[i] = sum_c partial[c*n + i]
__global__ void __launch_bounds__(256)
k_sum_chunks(const float* __restrict__ partial, int nchunk, int n, float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < nchunk; ++c) s += partial[(long)c * n + i];
  out[i] = s;
}

__global__ void __launch_bounds__(256)
k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ---------------------------------------------------------------------------
// BF16 MFMA inference.  A wavefront evaluates 16 rows; lane l computes the
// layer-1 outputs it feeds to the MFMA: row l%16, inputs k = 16 kb + 4 (l/16)
// + t.  Operand layout of v_mfma_f32_16x16x16bf16_1k: A(i, k) from lane
// i + 16 (k/4), element k%4; B(k, j) from lane j + 16 (k/4); result register
// r of lane l is D(4 (l/16) + r, l%16).
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16(float x) {
  __hip_bfloat16 b = __float2bfloat16(x);
  return *reinterpret_cast<short*>(&b);
}

__global__ void __launch_bounds__(256)
k_mlp_infer_bf16(const float* __restrict__ P, long R, const float* __restrict__ F,
                 float* __restrict__ W) {
  __shared__ float w1[NF * WD + WD];
  __shared__ short w2[WD * WD];        // bf16 bits, [k][o]
  __shared__ float b2w3[2 * WD + 1];
  for (int i = threadIdx.x; i < NF * WD + WD; i += 256) w1[i] = P[i];
  for (int i = threadIdx.x; i < WD * WD; i += 256) w2[i] = to_bf16(P[OFF_W2 + i]);
  for (int i = threadIdx.x; i < 2 * WD + 1; i += 256) b2w3[i] = P[OFF_B2 + i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long row0 = (blockIdx.x * 4L + wave) * 16;
  if (row0 >= R) return;
  const int li = lane & 15, lk = lane >> 4;
  // layer 1 for row row0 + li, inputs 16 kb + 4 lk + t
  const long ra = row0 + li;
  float f[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) f[q] = ra < R ? F[ra * NF + q] : 0.0f;
  bf16x4 a[WD / 16];
#pragma unroll
  for (int kb = 0; kb < WD / 16; ++kb) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = kb * 16 + 4 * lk + t;
      float s = w1[NF * WD + k];
#pragma unroll
      for (int q = 0; q < NF; ++q) s += f[q] * w1[q * WD + k];
      a[kb][t] = to_bf16(tanh_f(s));
    }
  }
  // layer 2 on MFMA, 8 output blocks of 16; layer 3 folded in
  float zpart[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ob = 0; ob < WD / 16; ++ob) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int o = ob * 16 + li;
#pragma unroll
    for (int kb = 0; kb < WD / 16; ++kb) {
      bf16x4 b;
#pragma unroll
      for (int t = 0; t < 4; ++t) b[t] = w2[(kb * 16 + 4 * lk + t) * WD + o];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kb], b, acc, 0, 0, 0);
    }
    const float bo = b2w3[o], w3o = b2w3[WD + o];
#pragma unroll
    for (int r = 0; r < 4; ++r) zpart[r] += tanh_f(acc[r] + bo) * w3o;
  }
  // sum over the 16 lanes that share lane/16 (the output columns)
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float v = zpart[r];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    zpart[r] = v;
  }
  if (li == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long row = row0 + 4 * lk + r;
      if (row < R) W[row] = zpart[r] + b2w3[2 * WD];
    }
  }
}

// ---------------------------------------------------------------------------
// feature rows and the learned diagonal correction
// ---------------------------------------------------------------------------
__device__ inline double clamp_den(double v) {
  return fabs(v) < MLP_DMIN ? copysign(MLP_DMIN, v) : v;
}

// F[(s*N + j)*NF + q] for Ritz pair s (theta[s], shift sigma[s]) and basis j;
// st = [e2 (N), r1 fraction (N), 2D index fraction (N)]
__global__ void __launch_bounds__(256)
k_mlp_features(const double* __restrict__ D, const double* __restrict__ st, int N, int S,
               const double* __restrict__ theta, const double* __restrict__ sigma,
               float* __restrict__ F) {
  const long R = (long)N * S;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < R; t += (long)gridDim.x * 256) {
    const int j = (int)(t % N), s = (int)(t / N);
    const double th = theta[s], sg = sigma[s];
    const double inv = 1.0 / MLP_DELTA;
    float* f = F + t * NF;
    f[0] = (float)asinh((D[j] - th) * inv);
    f[1] = (float)asinh((D[j] - sg) * inv);
    f[2] = (float)asinh((st[j] - th) * inv);
    f[3] = (float)st[N + j];
    f[4] = (float)st[2 * N + j];
    f[5] = (float)asinh((sg - th) * inv);
  }
}

// T[j, s] = w[s*N + j] R[j, s] / (D[j] - theta[s])   (w = nullptr: w = 1)
__global__ void __launch_bounds__(256)
k_mlp_apply(const float* __restrict__ w, const double* __restrict__ D,
            const double* __restrict__ theta, const double* __restrict__ R, int N, int S,
            double* __restrict__ T) {
  const long n = (long)N * S;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int j = (int)(t % N), s = (int)(t / N);
    const double ws = w ? (double)w[(long)s * N + j] : 1.0;
    T[t] = ws * R[t] / clamp_den(D[j] - theta[s]);
  }
}

// gw[s*N + j] = gt[j, s] R[j, s] / (D[j] - theta[s]) * scale
__global__ void __launch_bounds__(256)
k_mlp_gw(const double* __restrict__ gt, const double* __restrict__ D,
         const double* __restrict__ theta, const double* __restrict__ R, int N, int S,
         double scale, float* __restrict__ gw) {
  const long n = (long)N * S;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int j = (int)(t % N), s = (int)(t / N);
    gw[t] = (float)(gt[t] * R[t] / clamp_den(D[j] - theta[s]) * scale);
  }
}

__device__ inline double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// per column s (one work-group): tp = t - x (x.t); loss = 1 - (tp.e)^2 /
// (|tp|^2 |e|^2); G[:, s] = dloss/dt
__global__ void __launch_bounds__(256)
k_mlp_loss(const double* __restrict__ T, const double* __restrict__ X,
           const double* __restrict__ E, int N, double* __restrict__ G,
           double* __restrict__ loss) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  const double* t = T + (long)s * N;
  const double* x = X + (long)s * N;
  const double* e = E + (long)s * N;
  double* g = G + (long)s * N;
  double a = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) a += x[j] * t[j];
  const double xt = block_sum(a, red);
  double c = 0.0, tt = 0.0, ee = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) {
    const double tp = t[j] - x[j] * xt;
    c += tp * e[j];
    tt += tp * tp;
    ee += e[j] * e[j];
  }
  c = block_sum(c, red);
  tt = block_sum(tt, red);
  ee = block_sum(ee, red);
  const double den = fmax(tt * ee, 1e-300);
  const double ga = -2.0 * c / den, gb = 2.0 * c * c / (den * fmax(tt, 1e-300));
  double xg = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) {
    const double tp = t[j] - x[j] * xt;
    const double gv = ga * e[j] + gb * tp;
    g[j] = gv;
    xg += x[j] * gv;
  }
  xg = block_sum(xg, red);
  for (int j = threadIdx.x; j < N; j += 256) g[j] -= x[j] * xg;
  if (threadIdx.x == 0) loss[s] = 1.0 - c * c / den;
}

// out[s] = X[:, s] . Y[:, s]
__global__ void __launch_bounds__(256)
k_coldot(const double* __restrict__ X, const double* __restrict__ Y, int N,
         double* __restrict__ out) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) a += X[(long)s * N + j] * Y[(long)s * N + j];
  a = block_sum(a, red);
  if (threadIdx.x == 0) out[s] = a;
}

// Y[:, s] += alpha * c[s] * X[:, s]
__global__ void __launch_bounds__(256)
k_col_axpy(const double* __restrict__ X, const double* __restrict__ c, double alpha, int N, int S,
           double* __restrict__ Y) {
  const long n = (long)N * S;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int s = (int)(t / N);
    Y[t] += alpha * c[s] * X[t];
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
static int blocks_for(long n) { return (int)((n + 255) / 256); }

static void allow_param_lds(const void* k) {
  DVR_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(MLP_NPARAM * sizeof(float))));
}

void launch_mlp_fwd(hipStream_t s, const float* P, long R, const float* F, float* W, float* H1,
                    float* H2) {
  if (R <= 0) return;
  allow_param_lds((const void*)k_mlp_fwd);
  const double fl = 2.0 * R * (NF * WD + WD * WD + WD);
  DVR_LAUNCH("mlp_fwd", 4.0 * R * (NF + 1 + (H1 ? 2 * WD : 0)), fl, k_mlp_fwd,
             dim3(blocks_for(R)), dim3(256), MLP_NPARAM * sizeof(float), s, P, R, F, W, H1, H2);
}

void launch_mlp_infer_bf16(hipStream_t s, const float* P, long R, const float* F, float* W) {
  if (R <= 0) return;
  const double fl = 2.0 * R * (NF * WD + WD * WD + WD);
  DVR_LAUNCH("mlp_infer_bf16", 4.0 * R * (NF + 1), fl, k_mlp_infer_bf16,
             dim3((int)((R + 63) / 64)), dim3(256), 0, s, P, R, F, W);
}

void launch_mlp_bwd(hipStream_t s, const float* P, long R, const float* F, const float* H1,
                    const float* H2, const float* GW, float* GA1, float* GA2, float* grad,
                    float* partial, long partial_floats) {
  if (R <= 0) return;
  allow_param_lds((const void*)k_mlp_bwd_rows);
  DVR_LAUNCH("mlp_bwd_rows", 4.0 * R * (4 * WD + 1), 4.0 * R * WD * WD, k_mlp_bwd_rows,
             dim3(blocks_for(R)), dim3(256), MLP_NPARAM * sizeof(float), s, P, R, H1, H2, GW, GA1,
             GA2);
  // chunks sized so that the partial sums fit the workspace
  long nchunk = std::max<long>(1, std::min<long>((R + 4095) / 4096, 512));
  while (nchunk > 1 && nchunk * (long)WD * WD > partial_floats) nchunk /= 2;
  const long chunk = (R + nchunk - 1) / nchunk;
  auto wgrad = [&](const float* X, int np, bool xrow, const float* G, int nq, float* out) {
    const dim3 grid((unsigned)nchunk, (np + 15) / 16, (nq + 15) / 16);
    DVR_LAUNCH("mlp_wgrad", 4.0 * R * (np + nq), 2.0 * R * np * nq, k_mlp_wgrad, grid, dim3(256),
               0, s, X, np, xrow, G, nq, R, chunk, partial);
    const int n = np * nq;
    DVR_LAUNCH("mlp_sum", 4.0 * nchunk * n, 0.0, k_sum_chunks, dim3(blocks_for(n)), dim3(256), 0,
               s, partial, (int)nchunk, n, out);
  };
  wgrad(F, NF, true, GA1, WD, grad);                       // W1
  wgrad(H1, WD, false, GA2, WD, grad + OFF_W2);            // W2
  wgrad(H2, WD, false, GW, 1, grad + OFF_W3);              // W3
}

void launch_mlp_bias_grads(hipStream_t s, long R, const float* ones, const float* GA1,
                           const float* GA2, const float* GW, float* grad, float* partial,
                           long partial_floats) {
  long nchunk = std::max<long>(1, std::min<long>((R + 4095) / 4096, 512));
  while (nchunk > 1 && nchunk * (long)WD > partial_floats) nchunk /= 2;
  const long chunk = (R + nchunk - 1) / nchunk;
  auto colsum = [&](const float* G, int nq, float* out) {
    const dim3 grid((unsigned)nchunk, 1, (nq + 15) / 16);
    DVR_LAUNCH("mlp_wgrad", 4.0 * R * (1 + nq), 2.0 * R * nq, k_mlp_wgrad, grid, dim3(256), 0, s,
               ones, 1, false, G, nq, R, chunk, partial);
    DVR_LAUNCH("mlp_sum", 4.0 * nchunk * nq, 0.0, k_sum_chunks, dim3(blocks_for(nq)), dim3(256), 0,
               s, partial, (int)nchunk, nq, out);
  };
  colsum(GA1, WD, grad + OFF_B1);
  colsum(GA2, WD, grad + OFF_B2);
  colsum(GW, 1, grad + OFF_B3);
}

static int grid_cap(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

void launch_mlp_features(hipStream_t s, const double* D, const double* st, int N, int S,
                         const double* theta, const double* sigma, float* F) {
  const long R = (long)N * S;
  DVR_LAUNCH("mlp_features", 8.0 * N + 4.0 * R * NF, 0.0, k_mlp_features, dim3(grid_cap(R)),
             dim3(256), 0, s, D, st, N, S, theta, sigma, F);
}

void launch_mlp_apply(hipStream_t s, const float* w, const double* D, const double* theta,
                      const double* R, int N, int S, double* T) {
  const long n = (long)N * S;
  DVR_LAUNCH("mlp_apply", 20.0 * n, 3.0 * n, k_mlp_apply, dim3(grid_cap(n)), dim3(256), 0, s, w,
             D, theta, R, N, S, T);
}

void launch_mlp_gw(hipStream_t s, const double* gt, const double* D, const double* theta,
                   const double* R, int N, int S, double scale, float* gw) {
  const long n = (long)N * S;
  DVR_LAUNCH("mlp_gw", 20.0 * n, 4.0 * n, k_mlp_gw, dim3(grid_cap(n)), dim3(256), 0, s, gt, D,
             theta, R, N, S, scale, gw);
}

void launch_mlp_loss(hipStream_t s, const double* T, const double* X, const double* E, int N,
                     int S, double* G, double* loss) {
  DVR_LAUNCH("mlp_loss", 40.0 * N * S, 12.0 * N * S, k_mlp_loss, dim3(S), dim3(256), 0, s, T, X, E,
             N, G, loss);
}

void launch_coldot(hipStream_t s, const double* X, const double* Y, int N, int S, double* out) {
  DVR_LAUNCH("coldot", 16.0 * N * S, 2.0 * N * S, k_coldot, dim3(S), dim3(256), 0, s, X, Y, N, out);
}

void launch_col_axpy(hipStream_t s, const double* X, const double* c, double alpha, int N, int S,
                     double* Y) {
  const long n = (long)N * S;
  DVR_LAUNCH("col_axpy", 24.0 * n, 2.0 * n, k_col_axpy, dim3(grid_cap(n)), dim3(256), 0, s, X, c,
             alpha, N, S, Y);
}

void launch_adam(hipStream_t s, float* P, const float* g, float* m, float* v, int n, float lr,
                 int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - std::pow(b1, (float)step));
  const float c2 = 1.0f / (1.0f - std::pow(b2, (float)step));
  DVR_LAUNCH("adam", 20.0 * n, 10.0 * n, k_adam, dim3(blocks_for(n)), dim3(256), 0, s, P, g, m, v,
             n, lr, b1, b2, 1e-8f, c1, c2);
}

}  // namespace dvr
