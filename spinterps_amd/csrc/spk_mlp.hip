// Learned variogram proposal: synthetic Gaussian fields, variogram features
// and a small multilayer perceptron trained and run on the GPU.
//
//  spk_simulate_fields_dev : unconditional Gaussian fields at each step's
//                            active stations, z = mean + L xi with C = L L^T
//                            the step variogram's covariance (nugget on the
//                            diagonal), using the kriging kernels' assembly
//                            and Cholesky
//  spk_vg_features_dev     : 20 equal-width lag bins over (0, mdr * max
//                            distance]: semivariance / variance and the
//                            share of pairs per bin times the bin count
//                            (1 = uniform), 40 features
//  spk_mlp_*               : 40 -> 128 -> 128 -> 128 -> 6 with GELU;
//                            outputs = 3 model logits (Sph, Exp, Gau) and
//                            3 regressions (log sill/var, log range/max
//                            distance, nugget fraction); softmax cross-
//                            entropy + weighted squared error; Adam.
#include "spk_krige.h"

namespace spk {

// ---- synthetic fields ------------------------------------------------------
__global__ void k_sim_diag(double *__restrict__ A, int64_t ld, int64_t stride,
                           const StepSc *__restrict__ sc,
                           const spk_vg *__restrict__ vgs) {
  const int s = blockIdx.y;
  if (sc[s].status) return;
  const int n = sc[s].n;
  const spk_vg &vg = vgs[s];
  double nug = 0.0, tot = 0.0;
  for (int k = 0; k < vg.n; ++k) {
    tot += vg.sill[k];
    if (vg.type[k] == SPK_VG_NUG) nug += vg.sill[k];
  }
  double *As = A + (int64_t)s * stride;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    As[i + (int64_t)i * ld] += nug + 1e-6 * tot;
}

__device__ inline unsigned long long smix(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// standard normal from a counter (Box-Muller on two hashed uniforms)
__device__ inline double normal_at(unsigned long long seed,
                                   unsigned long long k) {
  const unsigned long long a = smix(seed ^ smix(2 * k));
  const unsigned long long b = smix(seed ^ smix(2 * k + 1));
  const double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

// z_i = mean + sum_{k <= i} L_ik xi_k at active station i, NaN elsewhere
__global__ __launch_bounds__(256) void k_sim_apply(
    const double *__restrict__ L, int64_t ld, int64_t stride,
    const StepSc *__restrict__ sc, const int32_t *__restrict__ idx,
    int64_t N, const double *__restrict__ means, unsigned long long seed,
    int64_t t0, double *__restrict__ out) {
  const int s = blockIdx.x;
  double *o = out + (int64_t)s * N;
  for (int64_t i = threadIdx.x; i < N; i += 256) o[i] = NAN;
  __syncthreads();
  if (sc[s].status) return;
  const int n = sc[s].n;
  const double *Ls = L + (int64_t)s * stride;
  const int32_t *my = idx + (int64_t)s * N;
  __shared__ double xi[256];
  const unsigned long long sd = smix(seed ^ smix((unsigned long long)(t0 + s)));
  for (int i0 = 0; i0 < n; i0 += 256) {
    double acc = 0.0;
    const int i = i0 + threadIdx.x;
    for (int k0 = 0; k0 <= i0; k0 += 256) {
      __syncthreads();
      xi[threadIdx.x] =
          (k0 + (int)threadIdx.x < n) ? normal_at(sd, k0 + threadIdx.x)
                                      : 0.0;
      __syncthreads();
      if (i < n) {
        const int kmax = min(256, i - k0 + 1);
        for (int k = 0; k < kmax; ++k)
          acc += Ls[i + (int64_t)(k0 + k) * ld] * xi[k];
      }
    }
    if (i < n) o[my[i]] = means[t0 + s] + acc;
  }
}

// ---- lag-bin features ---------------------------------------------------------
__global__ __launch_bounds__(256) void k_features(
    const int32_t *__restrict__ pi, const int32_t *__restrict__ pj,
    const double *__restrict__ pd, int64_t M, const double *__restrict__ vals,
    int64_t N, double mdr, int nlag, float *__restrict__ feat,
    double *__restrict__ scale) {
  const int s = blockIdx.x;
  const double *z = vals + (int64_t)s * N;
  const int tid = threadIdx.x;
  __shared__ double red[256];
  __shared__ double sums[64];
  __shared__ unsigned long long cnts[64];
  __shared__ double maxd_s, var_s;
  // mean and variance of the step's values
  double a = 0.0, c = 0.0;
  for (int64_t i = tid; i < N; i += 256)
    if (!isnan(z[i])) {
      a += z[i];
      c += 1.0;
    }
  red[tid] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double sa = red[0];
  __syncthreads();
  red[tid] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double sc = red[0];
  __syncthreads();
  const double mean = sa / sc;
  double v = 0.0;
  for (int64_t i = tid; i < N; i += 256)
    if (!isnan(z[i])) v += (z[i] - mean) * (z[i] - mean);
  red[tid] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    var_s = red[0] / sc;
    // largest active pair distance: scan the sorted pairs from the end
    double md = 0.0;
    for (int64_t e = M - 1; e >= 0; --e)
      if (!isnan(z[pi[e]]) && !isnan(z[pj[e]])) {
        md = pd[e];
        break;
      }
    maxd_s = md * mdr;
    scale[2 * s] = maxd_s;
    scale[2 * s + 1] = var_s;
  }
  if (tid < 64) {
    sums[tid] = 0.0;
    cnts[tid] = 0;
  }
  __syncthreads();
  const double md = maxd_s;
  for (int64_t e = tid; md > 0.0 && e < M; e += 256) {
    const double d = pd[e];
    if (d > md) break;  // sorted: nothing beyond
    const double zi = z[pi[e]], zj = z[pj[e]];
    if (isnan(zi) || isnan(zj)) continue;
    int b = (int)(d / md * nlag);
    if (b >= nlag) b = nlag - 1;
    atomicAdd(&sums[b], 0.5 * (zi - zj) * (zi - zj));
    atomicAdd(&cnts[b], 1ull);
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long tot = 0;
    for (int b = 0; b < nlag; ++b) tot += cnts[b];
    float *f = feat + (int64_t)s * 2 * nlag;
    for (int b = 0; b < nlag; ++b) {
      const double g = cnts[b] ? sums[b] / cnts[b] : 0.0;
      f[b] = (float)(var_s > 0 ? g / var_s : 0.0);
      f[nlag + b] = (float)(tot ? (double)cnts[b] * nlag / tot : 0.0);
    }
  }
}

// ---- MLP -----------------------------------------------------------------------
// Layer sizes and the flat parameter layout [W1 b1 W2 b2 W3 b3 W4 b4] with
// W_l row-major [out x in].
constexpr int kIn = 40, kH = 128, kOut = 6, kNReg = 3, kNCls = 3;
__host__ __device__ constexpr int mlp_dim(int l) {
  return l == 0 ? kIn : (l == 4 ? kOut : kH);
}
__host__ __device__ constexpr int64_t w_off(int l) {  // l = 1..4
  int64_t o = 0;
  for (int k = 1; k < l; ++k) o += (int64_t)mlp_dim(k) * mlp_dim(k - 1) + mlp_dim(k);
  return o;
}
__host__ __device__ constexpr int64_t b_off(int l) {
  return w_off(l) + (int64_t)mlp_dim(l) * mlp_dim(l - 1);
}
constexpr int64_t kNParams = w_off(5);
constexpr int kMaxBatch = 1 << 21;  // 32-row tiles in grid.y <= 65535

__device__ inline float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ inline float gelu_d(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// C[r][c] = sum_k A(r, k) B(k, c) over LDS tiles of 32.  A(r, k) =
// A[r * lda + k] (TA = 0) or A[k * lda + r] (TA = 1); B(k, c) likewise.
// beta = 0 overwrites, 1 accumulates.  16 x 16 threads, 2 x 2 outputs each.
template <int TA, int TB>
__global__ __launch_bounds__(256) void k_sgemm(int R, int Cc, int K,
                                               const float *__restrict__ A,
                                               int lda,
                                               const float *__restrict__ B,
                                               int ldb, float *__restrict__ Cm,
                                               int ldc, int beta) {
  __shared__ float As[32][33], Bs[32][33];
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
      const int i = e >> 5, k = e & 31;
      const int r = r0 + i, kk = k0 + k;
      As[i][k] = (r < R && kk < K) ? (TA ? A[(int64_t)kk * lda + r]
                                         : A[(int64_t)r * lda + kk])
                                   : 0.f;
      const int c = c0 + i;
      Bs[i][k] = (c < Cc && kk < K) ? (TB ? B[(int64_t)c * ldb + kk]
                                          : B[(int64_t)kk * ldb + c])
                                    : 0.f;
    }
    __syncthreads();
    for (int k = 0; k < 32; ++k) {
      const float a0 = As[ty * 2][k], a1 = As[ty * 2 + 1][k];
      const float b0 = Bs[tx * 2][k], b1 = Bs[tx * 2 + 1][k];
      acc[0][0] += a0 * b0;
      acc[0][1] += a0 * b1;
      acc[1][0] += a1 * b0;
      acc[1][1] += a1 * b1;
    }
    __syncthreads();
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int r = r0 + ty * 2 + a, c = c0 + tx * 2 + b;
      if (r < R && c < Cc) {
        float *p = &Cm[(int64_t)r * ldc + c];
        *p = beta ? *p + acc[a][b] : acc[a][b];
      }
    }
}

// Z[b][o] += bias[o]; H = gelu(Z) (act) or H = Z
__global__ void k_bias_act(float *__restrict__ Z, float *__restrict__ H,
                           const float *__restrict__ bias, int B, int O,
                           int act) {
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < (int64_t)B * O;
       e += (int64_t)gridDim.x * 256) {
    const float z = Z[e] + bias[e % O];
    Z[e] = z;
    H[e] = act ? gelu(z) : z;
  }
}

// dZ = dH * gelu'(Z)
__global__ void k_act_bwd(const float *__restrict__ Z, float *__restrict__ dH,
                          int64_t n) {
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * 256)
    dH[e] *= gelu_d(Z[e]);
}

// db[o] = sum_b dZ[b][o]
__global__ void k_bias_grad(const float *__restrict__ dZ, int B, int O,
                            float *__restrict__ db) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= O) return;
  float acc = 0.f;
  for (int b = 0; b < B; ++b) acc += dZ[(int64_t)b * O + o];
  db[o] = acc;
}

// Mean loss over the batch and its gradient w.r.t. the outputs:
// L = CE(softmax(y[0:3]), cls) + reg_w * sum_k (y[3+k] - t[k])^2 * m[k]
// where m masks regression targets that are NaN; a class outside 0..2
// drops the cross-entropy term of that sample.
__global__ void k_loss(const float *__restrict__ Y, const int32_t *__restrict__ cls,
                       const float *__restrict__ tgt, int B, float reg_w,
                       float *__restrict__ dY, float *__restrict__ loss) {
  __shared__ float red[256];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) {
    const float *y = Y + (int64_t)b * kOut;
    float *g = dY + (int64_t)b * kOut;
    float mx = fmaxf(fmaxf(y[0], y[1]), y[2]);
    float e0 = expf(y[0] - mx), e1 = expf(y[1] - mx), e2 = expf(y[2] - mx);
    const float se = e0 + e1 + e2;
    const int c = cls[b];
    const float p[3] = {e0 / se, e1 / se, e2 / se};
    const bool has_c = c >= 0 && c < kNCls;  // other labels: no class term
    if (has_c) acc += -(y[c] - mx - logf(se));
    for (int k = 0; k < kNCls; ++k)
      g[k] = has_c ? (p[k] - (k == c ? 1.f : 0.f)) / B : 0.f;
    for (int k = 0; k < kNReg; ++k) {
      const float t = tgt[(int64_t)b * kNReg + k];
      if (isnan(t)) {
        g[kNCls + k] = 0.f;
        continue;
      }
      const float r = y[kNCls + k] - t;
      acc += reg_w * r * r;
      g[kNCls + k] = 2.f * reg_w * r / B;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] / B;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * 256) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void k_init_params(float *__restrict__ p, unsigned long long seed) {
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < kNParams;
       e += (int64_t)gridDim.x * 256) {
    int l = 1;
    while (l < 4 && e >= w_off(l + 1)) ++l;
    const bool is_w = e < b_off(l);
    const double u =
        (double)(smix(seed ^ smix((unsigned long long)e)) >> 11) *
        (1.0 / 9007199254740992.0);
    // uniform(-1/sqrt(in), 1/sqrt(in)) for weights and biases
    const float lim = 1.0f / sqrtf((float)mlp_dim(l - 1));
    p[e] = (float)((2.0 * u - 1.0) * lim) * (is_w ? 1.0f : 0.5f);
  }
}

static int sgemm(hipStream_t st, int ta, int tb, int R, int Cc, int K,
                 const float *A, int lda, const float *B, int ldb, float *Cm,
                 int ldc, int beta) {
  const dim3 g((unsigned)ceil_div(Cc, 32), (unsigned)ceil_div(R, 32));
  if (!ta && tb)
    hipLaunchKernelGGL((k_sgemm<0, 1>), g, dim3(256), 0, st, R, Cc, K, A, lda,
                       B, ldb, Cm, ldc, beta);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_sgemm<1, 0>), g, dim3(256), 0, st, R, Cc, K, A, lda,
                       B, ldb, Cm, ldc, beta);
  else if (!ta && !tb)
    hipLaunchKernelGGL((k_sgemm<0, 0>), g, dim3(256), 0, st, R, Cc, K, A, lda,
                       B, ldb, Cm, ldc, beta);
  else
    hipLaunchKernelGGL((k_sgemm<1, 1>), g, dim3(256), 0, st, R, Cc, K, A, lda,
                       B, ldb, Cm, ldc, beta);
  return check_launch("k_sgemm");
}

// Activations workspace per layer: Z_l and H_l [B x dim(l)], l = 1..4.
static int64_t act_floats(int B) {
  int64_t t = 0;
  for (int l = 1; l <= 4; ++l) t += 2 * (int64_t)B * mlp_dim(l);
  return t;
}

static int mlp_forward(spk_ctx *ctx, const float *P, const float *X, int B,
                       float *acts) {
  hipStream_t st = ctx->stream;
  const float *Hprev = X;
  float *a = acts;
  for (int l = 1; l <= 4; ++l) {
    const int I = mlp_dim(l - 1), O = mlp_dim(l);
    float *Z = a, *H = a + (int64_t)B * O;
    // Z = H_prev W^T
    SPK_TRY(sgemm(st, 0, 1, B, O, I, Hprev, I, P + w_off(l), I, Z, O, 0));
    const int64_t n = (int64_t)B * O;
    hipLaunchKernelGGL(k_bias_act, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 4096)),
                       dim3(256), 0, st, Z, H, P + b_off(l), B, O,
                       l < 4 ? 1 : 0);
    SPK_TRY(check_launch("k_bias_act"));
    Hprev = H;
    a += 2 * n;
  }
  return SPK_OK;
}

}  // namespace spk

using namespace spk;

extern "C" {

int spk_simulate_fields_dev(spk_ctx *ctx, const spk_problem *prob, int64_t T,
                            const double *mask_vals, const spk_vg *vgs,
                            const double *means, uint64_t seed, double *out,
                            int32_t *info) {
  if (!ctx || !prob || !mask_vals || !vgs || !means || !out || !info)
    return set_error(SPK_EARG, "null argument");
  const int64_t N = prob->n_stns;
  if (N < 2) return set_error(SPK_ESHAPE, "need stations");
  if (ctx->pairs_sorted)
    return set_error(SPK_EARG,
                     "field simulation after a station-pair sort on the same "
                     "context is disabled (that order stalled on the GPU with "
                     "no cause found): simulate first, or use another "
                     "context");
  if (T <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  hipStream_t sm = ctx->stream;
  double *sx;
  SPK_TRY(ws_get(ctx, "stn_xy", 2 * N * sizeof(double), (void **)&sx));
  double *sy = sx + N;
  {
    const unsigned gs = (unsigned)std::min<int64_t>(ceil_div(N, 256), 4096);
    hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm,
                       prob->stn_x, sx, N, prob->origin_x);
    hipLaunchKernelGGL(k_recenter, dim3(gs), dim3(256), 0, sm,
                       prob->stn_y, sy, N, prob->origin_y);
  }
  const int64_t ld = ((N + 63) / 64) * 64;
  const int64_t strideM = ld * ld;
  const int64_t ldz = ld;
  const int64_t S = std::max<int64_t>(1, std::min<int64_t>(
      64, (int64_t)((size_t)4 << 30) / (strideM * 8)));
  int32_t *idx;
  double *zact, *A;
  StepSc *sc;
  SPK_TRY(ws_get(ctx, "sim_idx", S * N * 4, (void **)&idx));
  SPK_TRY(ws_get(ctx, "sim_z", S * ldz * 8, (void **)&zact));
  SPK_TRY(ws_get(ctx, "sim_sc", S * sizeof(StepSc), (void **)&sc));
  SPK_TRY(ws_get(ctx, "sim_A", S * strideM * 8, (void **)&A));
  const int tiles = (int)ceil_div(N, 64);
  for (int64_t t0 = 0; t0 < T; t0 += S) {
    const int Sb = (int)std::min<int64_t>(S, T - t0);
    hipLaunchKernelGGL(k_compact, dim3(Sb), dim3(256), 0, sm,
                       mask_vals + t0 * N, N, idx, zact, ldz, sc,
                       (const spk_vg *)nullptr, SPK_M_NNB, -INFINITY, 0.0,
                       (const int32_t *)nullptr);
    hipLaunchKernelGGL(k_assemble, dim3(tiles * tiles, Sb), dim3(256), 0, sm,
                       sx, sy, idx, N, sc, vgs + t0, 1, -INFINITY, A, ld,
                       strideM, (double *)nullptr);
    hipLaunchKernelGGL(k_sim_diag, dim3(8, Sb), dim3(256), 0, sm, A, ld,
                       strideM, sc, vgs + t0);
    hipLaunchKernelGGL(k_potrf, dim3(Sb), dim3(256), 0, sm, A, ld, strideM, sc,
                       vgs + t0, 0);
    hipLaunchKernelGGL(k_sim_apply, dim3(Sb), dim3(256), 0, sm, A, ld, strideM,
                       sc, idx, N, means, (unsigned long long)seed, t0,
                       out + t0 * N);
    hipLaunchKernelGGL(k_status_out, dim3((unsigned)ceil_div(Sb, 256)),
                       dim3(256), 0, sm, sc, Sb, info + t0, (double *)nullptr);
    SPK_TRY(check_launch("simulate"));
  }
  return SPK_OK;
}

int spk_vg_features_dev(spk_ctx *ctx, const int32_t *pi, const int32_t *pj,
                        const double *pd, int64_t M, int64_t N,
                        const double *vals, int64_t T, double mdr, int nlag,
                        float *feat, double *scale) {
  if (!ctx || !pi || !pj || !pd || !vals || !feat || !scale)
    return set_error(SPK_EARG, "null argument");
  if (nlag < 1 || nlag > 64) return set_error(SPK_EARG, "nlag in 1..64");
  if (T <= 0) return SPK_OK;
  SPK_HIP(hipSetDevice(ctx->device));
  for (int64_t t0 = 0; t0 < T; t0 += 65535) {
    const int Tb = (int)std::min<int64_t>(65535, T - t0);
    hipLaunchKernelGGL(k_features, dim3(Tb), dim3(256), 0, ctx->stream, pi, pj,
                       pd, M, vals + t0 * N, N, mdr, nlag,
                       feat + t0 * 2 * nlag, scale + t0 * 2);
    SPK_TRY(check_launch("k_features"));
  }
  return SPK_OK;
}

int64_t spk_mlp_n_params(void) { return kNParams; }

int spk_mlp_init_dev(spk_ctx *ctx, uint64_t seed, float *params) {
  if (!ctx || !params) return set_error(SPK_EARG, "null argument");
  SPK_HIP(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_init_params, dim3(64), dim3(256), 0, ctx->stream,
                     params, (unsigned long long)seed);
  return check_launch("k_init_params");
}

int spk_mlp_forward_dev(spk_ctx *ctx, const float *params, const float *X,
                        int B, float *out) {
  if (!ctx || !params || !X || !out) return set_error(SPK_EARG, "null argument");
  if (B <= 0) return SPK_OK;
  if (B > kMaxBatch) return set_error(SPK_ESHAPE, "batch above %d", kMaxBatch);
  SPK_HIP(hipSetDevice(ctx->device));
  float *acts;
  SPK_TRY(ws_get(ctx, "mlp_acts", act_floats(B) * 4, (void **)&acts));
  SPK_TRY(mlp_forward(ctx, params, X, B, acts));
  // the last H is the output
  int64_t off = 0;
  for (int l = 1; l < 4; ++l) off += 2 * (int64_t)B * mlp_dim(l);
  const float *H4 = acts + off + (int64_t)B * kOut;
  SPK_HIP(hipMemcpyAsync(out, H4, (size_t)B * kOut * 4,
                         hipMemcpyDeviceToDevice, ctx->stream));
  return SPK_OK;
}

int spk_mlp_loss_grad_dev(spk_ctx *ctx, const float *params, const float *X,
                          int B, const int32_t *cls, const float *reg,
                          float reg_w, float *grads, float *loss) {
  if (!ctx || !params || !X || !cls || !reg || !grads || !loss)
    return set_error(SPK_EARG, "null argument");
  if (B <= 0) return set_error(SPK_EARG, "empty batch");
  if (B > kMaxBatch) return set_error(SPK_ESHAPE, "batch above %d", kMaxBatch);
  SPK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  float *acts, *dbuf;
  SPK_TRY(ws_get(ctx, "mlp_acts", act_floats(B) * 4, (void **)&acts));
  SPK_TRY(ws_get(ctx, "mlp_d", 2 * (int64_t)B * kH * 4, (void **)&dbuf));
  SPK_TRY(mlp_forward(ctx, params, X, B, acts));
  // layer activation pointers
  float *Z[5], *H[5];
  {
    float *a = acts;
    for (int l = 1; l <= 4; ++l) {
      Z[l] = a;
      H[l] = a + (int64_t)B * mlp_dim(l);
      a += 2 * (int64_t)B * mlp_dim(l);
    }
  }
  float *dA = dbuf, *dB = dbuf + (int64_t)B * kH;
  // dY = dL/dH4
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, st, H[4], cls, reg, B,
                     reg_w, dA, loss);
  SPK_TRY(check_launch("k_loss"));
  for (int l = 4; l >= 1; --l) {
    const int I = mlp_dim(l - 1), O = mlp_dim(l);
    const float *Hprev = l > 1 ? H[l - 1] : X;
    // dA holds dZ_l (the output layer has no activation)
    if (l < 4) {
      hipLaunchKernelGGL(k_act_bwd, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)B * O, 256), 4096)),
                         dim3(256), 0, st, Z[l], dA, (int64_t)B * O);
      SPK_TRY(check_launch("k_act_bwd"));
    }
    // dW_l = dZ^T H_prev  [O x I]
    SPK_TRY(sgemm(st, 1, 0, O, I, B, dA, O, Hprev, I, grads + w_off(l), I, 0));
    hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)ceil_div(O, 256)),
                       dim3(256), 0, st, dA, B, O, grads + b_off(l));
    SPK_TRY(check_launch("k_bias_grad"));
    if (l > 1) {
      // dH_prev = dZ W  [B x I]
      SPK_TRY(sgemm(st, 0, 0, B, I, O, dA, O, params + w_off(l), I, dB, I, 0));
      float *t = dA;
      dA = dB;
      dB = t;
    }
  }
  return SPK_OK;
}

int spk_mlp_adam_dev(spk_ctx *ctx, float *params, const float *grads,
                     float *m, float *v, int64_t step, float lr, float b1,
                     float b2, float eps) {
  if (!ctx || !params || !grads || !m || !v)
    return set_error(SPK_EARG, "null argument");
  if (step < 1) return set_error(SPK_EARG, "step counts from 1");
  SPK_HIP(hipSetDevice(ctx->device));
  const float c1 = 1.f - powf(b1, (float)step);
  const float c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, ctx->stream, params,
                     grads, m, v, kNParams, lr, b1, b2, eps, c1, c2);
  return check_launch("k_adam");
}

}  // extern "C"
