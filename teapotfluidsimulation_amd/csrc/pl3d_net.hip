// pl3d_net.hip — kernels of the learned pressure preconditioner (network layout in pl3d_net.h) and
// of the flexible CG that applies it.
//
// Convolutions are implicit GEMMs on the 16x16x16 BF16 MFMA with FP32 accumulators:
//   forward      out[v][co] = sum_{tap, ci} in[v + off(tap)][ci] * W[co][tap * cin + ci]
//                M = 16 voxels per wave (consecutive linear indices), N = output channels in
//                16-column tiles, K = 27 cin in 16-wide steps; a lane's 4 K values are 4
//                consecutive channels of one neighbour (one 8-byte load; cin is a multiple of 4)
//   input grad   the same kernel with the flipped, transposed weights (k_wflip)
//   weight grad  dW[co][k] = sum_v dy[v][co] * in[v + off(tap(k))][ci(k)]: M = output channels,
//                N = 16 k columns per wave, K = voxels; each wave sums one chunk of voxels and
//                writes an FP32 partial, and k_adam reduces the partials in a fixed order
//                (deterministic) before the Adam update.
// MFMA lane layout (16x16x16): lane l supplies A[l % 16][4 (l / 16) + e] and
// B[4 (l / 16) + e][l % 16], e = 0..3, and receives D[4 (l / 16) + e][l % 16].
#include "pl3d_net.h"

#include "pl3d_device.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
// FP32 -> BF16, round to nearest even (finite inputs; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__device__ __forceinline__ bool nbr(int i, int j, int k, int tap, int nx, int ny, int nz, int &vn) {
  const int ii = i + tap % 3 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap / 9 - 1;
  if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) return false;
  vn = ii + nx * (jj + ny * kk);
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------------- convolution
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  constexpr int COUTP = (COUT + 15) / 16 * 16;
  constexpr int KP = net_kp(CIN);
  constexpr int NT = COUTP / 16;
  static_assert(CIN % 4 == 0, "channels in groups of 4");
  const int lane = threadIdx.x & 63;
  const int v0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  const int nx = a.nx, ny = a.ny, nz = a.nz, V = nx * ny * nz;
  if (v0 >= V) return;
  const int va = v0 + (lane & 15);
  const bool vin = va < V;
  const int i = va % nx, j = (va / nx) % ny, k = va / (nx * ny);
  const int kq = 4 * (lane >> 4);
  f32x4 acc[NT];
#pragma unroll
  for (int n = 0; n < NT; n++) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kb = 0; kb < KP; kb += 16) {
    const int kk = kb + kq;
    const int tap = kk / CIN, ci = kk % CIN;
    bf16x4 av = {0, 0, 0, 0};
    int vn;
    if (vin && tap < 27 && nbr(i, j, k, tap, nx, ny, nz, vn))
      av = *reinterpret_cast<const bf16x4 *>(a.in + size_t(vn) * CIN + ci);
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const bf16x4 bv =
          *reinterpret_cast<const bf16x4 *>(a.w + size_t(n * 16 + (lane & 15)) * KP + kk);
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[n], 0, 0, 0);
    }
  }
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int co = n * 16 + (lane & 15);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int v = v0 + 4 * (lane >> 4) + e;
      if (v >= V) continue;
      const float x = acc[n][e];
      if constexpr (EPI == EPI_MASK_F32) {
        if (co == 0) a.outf[v] = a.flags[v] == PL3D_FLUID_ ? x : 0.f;
        continue;
      } else {
        if (co >= COUT) continue;
        const size_t o = size_t(v) * COUT + co;
        if constexpr (EPI == EPI_RELU) {
          a.out[o] = f2bf(fmaxf(x, 0.f));
        } else if constexpr (EPI == EPI_RELU_ADD) {
          const float h = fmaxf(x, 0.f);
          const uint16_t hb = f2bf(h);
          a.out[o] = hb;
          a.out2[o] = f2bf(bf2f(hb) + bf2f(a.skip[o]));
        } else if constexpr (EPI == EPI_PLAIN) {
          a.out[o] = f2bf(x);
        } else {  // EPI_GRAD_SPLIT
          a.out[o] = f2bf(bf2f(a.skip[o]) > 0.f ? x : 0.f);
          a.out2[o] = f2bf(x);
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------- weight gradient
template <int CIN, int COUT>
__global__ __launch_bounds__(64) void k_wgrad(WgradArgs a) {
  constexpr int COUTP = (COUT + 15) / 16 * 16;
  constexpr int COUTQ = (COUT + 3) / 4 * 4;  // channels of the gradient tensor dy
  constexpr int KP = net_kp(CIN);
  constexpr int MT = COUTP / 16;
  const int lane = threadIdx.x;
  const int kt = blockIdx.x;     // 16-column tile of k
  const int chunk = blockIdx.y;  // voxel chunk
  const int nx = a.nx, ny = a.ny, nz = a.nz, V = nx * ny * nz;
  const int kk = kt * 16 + (lane & 15);
  const int tap = kk / CIN, ci = kk % CIN;
  const int vq = 4 * (lane >> 4);
  f32x4 acc[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int vb = chunk * a.chunk, ve = min(vb + a.chunk, V);
  for (int v0 = vb; v0 < ve; v0 += 16) {
    bf16x4 bv = {0, 0, 0, 0};
    int vv[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int v = v0 + vq + e;
      vv[e] = v;
      int vn;
      if (v < ve && tap < 27) {
        const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
        if (nbr(i, j, k, tap, nx, ny, nz, vn)) bv[e] = short(a.in[size_t(vn) * CIN + ci]);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co = m * 16 + (lane & 15);
      bf16x4 av = {0, 0, 0, 0};
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (co < COUT && vv[e] < ve) av[e] = short(a.dy[size_t(vv[e]) * COUTQ + co]);
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[m], 0, 0, 0);
    }
  }
  float *P = a.partial + size_t(chunk) * COUTP * KP;
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int e = 0; e < 4; e++) P[size_t(m * 16 + 4 * (lane >> 4) + e) * KP + kk] = acc[m][e];
}

// Partials of `chunks` chunks -> gradient, then Adam on the FP32 master weights (and their BF16
// copy).  Padding entries (co >= cout or k >= 27 cin) keep weight 0.
__global__ __launch_bounds__(256) void k_adam(const float *partial, int chunks, int coutp, int kp,
                                              int cout, int kreal, float *w, float *m, float *v,
                                              uint16_t *wbf, float lr, float b1, float b2,
                                              float eps, float c1, float c2, float *gout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = coutp * kp;
  if (p >= n) return;
  const int co = p / kp, kk = p % kp;
  float g = 0.f;
  for (int c = 0; c < chunks; c++) g += partial[size_t(c) * n + p];
  if (co >= cout || kk >= kreal) g = 0.f;
  if (gout) gout[p] = g;
  if (lr == 0.f) return;  // gradient only
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  float wn = w[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  if (co >= cout || kk >= kreal) wn = 0.f;
  w[p] = wn;
  wbf[p] = f2bf(wn);
}

// Flipped, transposed weights of a layer for its input gradient: the gradient tensor has
// coutq = cout rounded up to 4 channels, so W'[ci][tap' coutq + co] = W[co][(26 - tap') cin + ci].
__global__ __launch_bounds__(256) void k_wflip(const uint16_t *w, int cin, int cout, int kp,
                                               uint16_t *wt, int coutq, int kpt, int cinp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cinp * kpt) return;
  const int ci = p / kpt, q = p % kpt;
  const int tapt = q / coutq, co = q % coutq;
  uint16_t val = 0;
  if (ci < cin && tapt < 27 && co < cout) val = w[size_t(co) * kp + (26 - tapt) * cin + ci];
  wt[p] = val;
}

// --------------------------------------------------------------------------- elementwise kernels
// x0 = [r * inv_s, FLUID, SOLID, 0] from FP32 r (grid order)
__global__ __launch_bounds__(256) void k_pack_input(const float *r, const uint8_t *flags,
                                                    float inv_s, int V, uint16_t *x0) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint8_t f = flags[v];
  x0[4 * v] = f2bf(f == PL3D_FLUID_ ? r[v] * inv_s : 0.f);
  x0[4 * v + 1] = f2bf(f == PL3D_FLUID_ ? 1.f : 0.f);
  x0[4 * v + 2] = f2bf(f == PL3D_SOLID_ ? 1.f : 0.f);
  x0[4 * v + 3] = 0;
}

// coarse voxel of fine (i, j, k) is (i/2, j/2, k/2) on the ceil(n/2) grid
__device__ __forceinline__ int parent_of(int v, int nx, int ny, int cx, int cy) {
  const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
  return (i >> 1) + cx * ((j >> 1) + cy * (k >> 1));
}

// pool: out[p][c] = (sum of the in-grid children) / 8
__global__ __launch_bounds__(256) void k_pool(const uint16_t *in, int C, int nx, int ny, int nz,
                                              uint16_t *out) {
  const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cx * cy * cz * C) return;
  const int p = q / C, c = q % C;
  const int I = p % cx, J = (p / cx) % cy, K = p / (cx * cy);
  float s = 0.f;
  for (int d = 0; d < 8; d++) {
    const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
    if (i < nx && j < ny && k < nz) s += bf2f(in[size_t(i + nx * (j + ny * k)) * C + c]);
  }
  out[q] = f2bf(0.125f * s);
}

// up (nearest): out[v][c] = in[parent(v)][c] on the fine grid (nx, ny, nz)
__global__ __launch_bounds__(256) void k_up(const uint16_t *in, int C, int nx, int ny, int nz,
                                            uint16_t *out) {
  const int cx = (nx + 1) / 2, cy = (ny + 1) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * ny * nz * C) return;
  const int v = q / C, c = q % C;
  out[q] = in[size_t(parent_of(v, nx, ny, cx, cy)) * C + c];
}

// gradient through up + relu (+ skip): g[p] = sum over the children of gf (fine grid nx, ny, nz);
// gpre = g * (act > 0); gskip (optional) = g
__global__ __launch_bounds__(256) void k_upT(const uint16_t *gf, int C, int nx, int ny, int nz,
                                             const uint16_t *act, uint16_t *gpre,
                                             uint16_t *gskip) {
  const int cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cx * cy * cz * C) return;
  const int p = q / C, c = q % C;
  const int I = p % cx, J = (p / cx) % cy, K = p / (cx * cy);
  float s = 0.f;
  for (int d = 0; d < 8; d++) {
    const int i = 2 * I + (d & 1), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d >> 2);
    if (i < nx && j < ny && k < nz) s += bf2f(gf[size_t(i + nx * (j + ny * k)) * C + c]);
  }
  gpre[q] = f2bf(bf2f(act[q]) > 0.f ? s : 0.f);
  if (gskip) gskip[q] = f2bf(s);
}

// gradient through pool, a skip sum and a relu on the fine grid:
// gpre[v] = (gskip[v] + gc[parent(v)] / 8) * (act[v] > 0)
__global__ __launch_bounds__(256) void k_poolT(const uint16_t *gc, const uint16_t *gskip, int C,
                                               int nx, int ny, int nz, const uint16_t *act,
                                               uint16_t *gpre) {
  const int cx = (nx + 1) / 2, cy = (ny + 1) / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * ny * nz * C) return;
  const int v = q / C, c = q % C;
  const float g = bf2f(gskip[q]) + 0.125f * bf2f(gc[size_t(parent_of(v, nx, ny, cx, cy)) * C + c]);
  gpre[q] = f2bf(bf2f(act[q]) > 0.f ? g : 0.f);
}

// loss = sum_FLUID (z - t)^2 * invT2 (per-block partials); gz4 = [2 (z - t) invT2 on FLUID, 0, 0, 0]
__global__ __launch_bounds__(256) void k_loss(const float *z, const float *t, const uint8_t *flags,
                                              float invT2, int V, uint16_t *gz4, double *partial) {
  __shared__ double red[256];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (v < V) {
    const bool f = flags[v] == PL3D_FLUID_;
    const float d = f ? z[v] - t[v] : 0.f;
    l = double(d) * double(d) * invT2;
    gz4[4 * v] = f2bf(2.f * d * invT2);
    gz4[4 * v + 1] = gz4[4 * v + 2] = gz4[4 * v + 3] = 0;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------- flexible CG kernels
// The 7-point operator of the pressure system on the grid (FLUID rows; AIR = 0, SOLID = Neumann),
// solver3d.ts:872-881.  Vectors are zero off the liquid.
__device__ __forceinline__ float apply_a(const float *p, const uint8_t *fl, int v, int nx, int nxy) {
  const int offs[6] = {-1, 1, -nx, nx, -nxy, nxy};
  float s = 0.f;
  int dg = 0;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const uint8_t f = fl[v + offs[q]];
    dg += f != PL3D_SOLID_;
    if (f == PL3D_FLUID_) s += p[v + offs[q]];
  }
  return float(dg) * p[v] - s;
}

__device__ __forceinline__ void block_reduce2(double a, double b, double *pa, double *pb) {
  __shared__ double ra[256], rb[256];
  ra[threadIdx.x] = a;
  rb[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      ra[threadIdx.x] += ra[threadIdx.x + s];
      rb[threadIdx.x] += rb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *pa = ra[0];
    *pb = rb[0];
  }
}

// Ap = A p and partial p.Ap
__global__ __launch_bounds__(256) void k_fcg_spmv(const float *p, const uint8_t *fl, int nx,
                                                  int ny, int V, float *Ap, double *partial,
                                                  const FcgState *st) {
  if (st->done) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (v < V) {
    float a = 0.f;
    if (fl[v] == PL3D_FLUID_) a = apply_a(p, fl, v, nx, nx * ny);
    Ap[v] = a;
    d = double(p[v]) * double(a);
  }
  double dummy;
  block_reduce2(d, 0.0, &partial[blockIdx.x], &dummy);
}

// sum (op 0) or max (op 1) of nb per-block partials, in a fixed order; valid in thread 0
template <class T>
__device__ __forceinline__ double fixed_reduce(const T *part, int nb, int stride, int op) {
  __shared__ double red[256];
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const double x = double(part[size_t(b) * stride]);
    a = op ? fmax(a, x) : a + x;
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      red[threadIdx.x] = op ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                            : red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// alpha = z.r / p.Ap (one block); s.Ap <= 0 ends the solve unconverged (solver3d.ts:916)
__global__ __launch_bounds__(256) void k_fcg_alpha(const double *partial, int nb, FcgState *st) {
  if (st->done) return;
  const double pap = fixed_reduce(partial, nb, 1, 0);
  if (threadIdx.x == 0) {
    st->it += 1;
    if (!(pap > 0.0))
      st->done = 1;
    else
      st->alpha = st->zr / pap;
  }
}

// x += alpha p (FP64), rold = r, r -= alpha Ap; partial max|r|
__global__ __launch_bounds__(256) void k_fcg_update(const float *p, const float *Ap, double *x,
                                                    float *r, float *rold, int V, float *pmax,
                                                    const FcgState *st) {
  if (st->done) return;
  __shared__ float red[256];
  const double alpha = st->alpha;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.f;
  if (v < V) {
    x[v] += alpha * double(p[v]);
    const float ro = r[v];
    rold[v] = ro;
    const float rn = ro - float(alpha) * Ap[v];
    r[v] = rn;
    m = fabsf(rn);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pmax[blockIdx.x] = red[0];
}

// recursive max|r|; a true-residual refresh at the stop test and at least every 25 iterations
__global__ __launch_bounds__(256) void k_fcg_rcheck(const float *pmax, int nb, FcgState *st) {
  if (st->done) return;
  const double m = fixed_reduce(pmax, nb, 1, 1);
  if (threadIdx.x == 0) {
    st->rmax = m;
    st->needTrue = (m <= st->target || st->it - st->lastTrue >= 25) ? 1 : 0;
  }
}

// FP64 true residual r = b - A x on FLUID (0 elsewhere); partial max|r|
__global__ __launch_bounds__(256) void k_fcg_residual(const double *b, const double *x,
                                                      const uint8_t *fl, int nx, int ny, int V,
                                                      float *r, double *pmax,
                                                      const FcgState *st) {
  if (st && (st->done || !st->needTrue)) return;
  __shared__ double red[256];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (v < V) {
    double rv = 0.0;
    if (fl[v] == PL3D_FLUID_) {
      const int nxy = nx * ny;
      const int offs[6] = {-1, 1, -nx, nx, -nxy, nxy};
      double s = 0.0;
      int dg = 0;
      for (int q = 0; q < 6; q++) {
        const uint8_t f = fl[v + offs[q]];
        dg += f != PL3D_SOLID_;
        if (f == PL3D_FLUID_) s += x[v + offs[q]];
      }
      rv = b[v] - (double(dg) * x[v] - s);
    }
    r[v] = float(rv);
    m = fabs(rv);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pmax[blockIdx.x] = red[0];
}

// after a refresh: its max|r| replaces the recursive one and decides convergence
__global__ __launch_bounds__(256) void k_fcg_tcheck(const double *pmax, int nb, FcgState *st) {
  if (st->done || !st->needTrue) return;
  const double m = fixed_reduce(pmax, nb, 1, 1);
  if (threadIdx.x == 0) {
    st->lastTrue = st->it;
    st->rmax = m;
    st->needTrue = 0;
    if (m <= st->target) {
      st->done = 1;
      st->converged = 1;
    }
  }
}

// the network's input [r / max|r|, FLUID, SOLID, 0], the scale read from the state
__global__ __launch_bounds__(256) void k_fcg_pack(const float *r, const uint8_t *flags, int V,
                                                  uint16_t *x0, const FcgState *st) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const float inv_s = float(1.0 / st->rmax);
  const uint8_t f = flags[v];
  x0[4 * v] = f2bf(f == PL3D_FLUID_ ? r[v] * inv_s : 0.f);
  x0[4 * v + 1] = f2bf(f == PL3D_FLUID_ ? 1.f : 0.f);
  x0[4 * v + 2] = f2bf(f == PL3D_SOLID_ ? 1.f : 0.f);
  x0[4 * v + 3] = 0;
}

// z = s * znet (the net's output for r / s; s from the state when sdev is set), or with
// jacobi != 0 the fallback z = r / diag (only when the state asks for it); partials z.r, z.rold,
// z.z and r.r (4 per block)
__global__ __launch_bounds__(256) void k_fcg_dots(const float *znet, float s, const double *sdev,
                                                  const float *r, const float *rold,
                                                  const uint8_t *fl, int nx, int ny, int V,
                                                  int jacobi, float *z, double *part,
                                                  const FcgState *st) {
  if (st->done || (jacobi && !st->fallback)) return;
  if (sdev) s = float(*sdev);
  __shared__ double red[4][256];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  if (v < V) {
    float zz;
    if (jacobi) {
      zz = 0.f;
      if (fl[v] == PL3D_FLUID_) {
        const int nxy = nx * ny;
        const int offs[6] = {-1, 1, -nx, nx, -nxy, nxy};
        int dg = 0;
        for (int o = 0; o < 6; o++) dg += fl[v + offs[o]] != PL3D_SOLID_;
        zz = r[v] / float(dg);
      }
    } else {
      zz = s * znet[v];
    }
    z[v] = zz;
    q[0] = double(zz) * double(r[v]);
    q[1] = double(zz) * double(rold[v]);
    q[2] = double(zz) * double(zz);
    q[3] = double(r[v]) * double(r[v]);
  }
  for (int c = 0; c < 4; c++) red[c][threadIdx.x] = q[c];
  __syncthreads();
  for (int st2 = 128; st2 > 0; st2 >>= 1) {
    if (threadIdx.x < st2)
      for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + st2];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The dot sums of z (one block).  Pass 0 (the network's z): when z is not a descent direction,
// cos(z, r) <= 1e-3, flags the Jacobi fallback for pass 1 and stops; otherwise, and in pass 1,
// beta = (z.r - z.rold) / z.r_old (flexible, Polak-Ribiere; 0 for the first z) and z.r
// becomes the direction's.
__global__ __launch_bounds__(256) void k_fcg_dcheck(const double *part, int nb, int jacobi,
                                                    FcgState *st) {
  if (st->done || (jacobi && !st->fallback)) return;
  double q[4];
  for (int c = 0; c < 4; c++) {
    q[c] = fixed_reduce(part + c, nb, 4, 0);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (!jacobi) {
      if (!(q[0] > 1e-3 * sqrt(q[2] * q[3]))) {
        st->fallback = 1;
        st->fallbacks += 1;
        return;
      }
      st->fallback = 0;
    }
    st->beta = st->init ? 0.0 : (q[0] - q[1]) / st->zr;
    st->zr = q[0];
    st->init = 0;
  }
}

// one damped-Jacobi step on A z = r after the network: zout = s zin + omega (r - A (s zin)) / diag
// (the network captures the smooth part of A^-1 r; the smoothing step the rough part); s from
// the state when sdev is set
__global__ __launch_bounds__(256) void k_fcg_smooth(const float *zin, float s, const double *sdev,
                                                    const float *r, const uint8_t *fl, int nx,
                                                    int ny, int V, float omega, float *zout) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  if (sdev) s = float(*sdev);
  float out = 0.f;
  if (fl[v] == PL3D_FLUID_) {
    const int nxy = nx * ny;
    const int offs[6] = {-1, 1, -nx, nx, -nxy, nxy};
    float sum = 0.f;
    int dg = 0;
    for (int q = 0; q < 6; q++) {
      const uint8_t f = fl[v + offs[q]];
      dg += f != PL3D_SOLID_;
      if (f == PL3D_FLUID_) sum += zin[v + offs[q]];
    }
    const float zv = s * zin[v];
    const float az = float(dg) * zv - s * sum;
    out = zv + omega * (r[v] - az) / float(dg);
  }
  zout[v] = out;
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_fcg_dir(const float *z, int V, float *p,
                                                 const FcgState *st) {
  if (st->done) return;
  const double beta = st->beta;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) p[v] = z[v] + float(beta) * p[v];
}

// ------------------------------------------------------------------------------ instantiations
#define NET_CONV(CI, CO, E) template __global__ void k_conv<CI, CO, E>(ConvArgs);
// forward
NET_CONV(4, 8, EPI_RELU)
NET_CONV(8, 16, EPI_RELU)
NET_CONV(16, 32, EPI_RELU)
NET_CONV(32, 16, EPI_RELU_ADD)
NET_CONV(16, 8, EPI_RELU_ADD)
NET_CONV(8, 1, EPI_MASK_F32)
// input gradients (transposed layers: cin' = cout rounded to 4, cout' = cin)
NET_CONV(4, 8, EPI_GRAD_SPLIT)   // through conv_f: g_d1 -> relu mask of h1, skip part for e1
NET_CONV(8, 16, EPI_PLAIN)       // through conv_e: g_u2
NET_CONV(16, 32, EPI_PLAIN)      // through conv_d: g_u3
NET_CONV(32, 16, EPI_PLAIN)      // through conv_c: g_p2
NET_CONV(16, 8, EPI_PLAIN)       // through conv_b: g_p1
#undef NET_CONV
template __global__ void k_wgrad<4, 8>(WgradArgs);
template __global__ void k_wgrad<8, 16>(WgradArgs);
template __global__ void k_wgrad<16, 32>(WgradArgs);
template __global__ void k_wgrad<32, 16>(WgradArgs);
template __global__ void k_wgrad<16, 8>(WgradArgs);
template __global__ void k_wgrad<8, 1>(WgradArgs);
