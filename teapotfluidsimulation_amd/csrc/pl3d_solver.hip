// pl3d_solver.hip — matrix-free 7-point PCG for the liquid cells of one pour, one workgroup per
// solve (Pcg3D.solve, solver3d.ts:858-937).
//
// Why one workgroup: at 4-5 mm cells a pour has ~6-12 k liquid cells and every CG iteration is a
// chain of dependent reductions; one CU with 1024 threads holds the whole problem in registers
// plus LDS, so an iteration costs barriers instead of grid-wide synchronisation.  The pressure
// and the density-correction solves of a substep are independent and run as two workgroups of
// one launch.
//
// Storage and precision:
//   owner mapping   cell t (list order) is owned by thread t % 1024, register slot t / 1024
//   registers       r, x_lo (FP32 increments of the solution), s (own copy)
//   LDS             s[MMAX] (stencil input), W[MMAX] (A s during the product, then z = M r;
//                   the MIC(0) sweeps run in place in W)
//   global (FP64)   b, x_hi (solution)
//   dot products    FP64 partial sums, fixed-order tree: deterministic
//   refinement      x_hi += x_lo and r = b - A x_hi recomputed in FP64 ("reliable update",
//                   Sleijpen & van der Vorst 1996): after each decade of residual reduction, at
//                   least every RU_PERIOD iterations, and whenever the FP32 residual passes the
//                   stop test; CG restarts from the true residual when the two residuals differ
//                   by more than 1 %.  The solve stops only when the FP64 residual passes the
//                   reference's rule max|r| <= max(tol * max|b|, 1e-14) (solver3d.ts:910, :923-925).
// Preconditioners:
//   MIC(0)  tau = 0.97, safety 0.25 * diag (solver3d.ts:883-900).  The reference applies it in
//           lexicographic order; each cell depends only on its -x/-y/-z neighbours, so processing
//           whole hyperplanes i+j+k = l in increasing l (decreasing for the backward sweep) gives
//           the same factor and the same triangular solves, with one barrier per hyperplane.
//   Jacobi  z = r / diag.
#include "pl3d_device.h"
#include "pl3d_solver.h"

#define NT PL3D_SOLVE_THREADS
#define MMAX PL3D_SOLVE_MMAX
#define CPT (MMAX / NT)
#define NWAVES (NT / 64)
#define RU_PERIOD 25

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

struct Red {
  double sum[2][NWAVES];
  double sum2[2][NWAVES];
  float mx[2][NWAVES];
};

// Block-wide reductions with alternating buffers: one barrier each; every thread adds the 16
// wave partials in the same order, so the result is identical and uniform across the block.
__device__ __forceinline__ double block_sum(Red &R, int &buf, double v) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) R.sum[buf][w] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NWAVES; i++) t += R.sum[buf][i];
  buf ^= 1;
  return t;
}
__device__ __forceinline__ void block_sum_max(Red &R, int &buf, double v, float m, double &vs,
                                              float &ms) {
  v = wave_sum(v);
  m = wave_max(m);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    R.sum[buf][w] = v;
    R.mx[buf][w] = m;
  }
  __syncthreads();
  double t = 0;
  float mm = 0.f;
#pragma unroll
  for (int i = 0; i < NWAVES; i++) {
    t += R.sum[buf][i];
    mm = fmaxf(mm, R.mx[buf][i]);
  }
  buf ^= 1;
  vs = t;
  ms = mm;
}
__device__ __forceinline__ double block_max_d(Red &R, int &buf, double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) R.sum[buf][w] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NWAVES; i++) t = fmax(t, R.sum[buf][i]);
  buf ^= 1;
  return t;
}

__device__ __forceinline__ int nb_lo(uint32_t p) { return int(p & 0xFFFFu); }
__device__ __forceinline__ int nb_hi(uint32_t p) { return int(p >> 16); }

}  // namespace

// One workgroup per solve.  Arguments shared by the solves of one launch in `A`, per-solve
// arguments in `D[blockIdx.x]`.
__global__ __launch_bounds__(NT) void k_pcg(PcgShared A, const PcgDesc *__restrict__ Ds,
                                            const int *__restrict__ activeFlag) {
  if (activeFlag != nullptr && *activeFlag == 0) return;
  const PcgDesc D = Ds[blockIdx.x];
  __shared__ float s_lds[MMAX];
  __shared__ float W[MMAX];
  __shared__ Red R;
  __shared__ int planeStart[PL3D_MAX_PLANES + 1];
  __shared__ int planeFill[PL3D_MAX_PLANES];
  int rb = 0;
  const int tid = threadIdx.x;
  const int m = *A.m;
  double *__restrict__ xhi = D.q;
  const double *__restrict__ b = D.b;
  const uint32_t *__restrict__ nbx = A.nbr;
  const uint32_t *__restrict__ nby = A.nbr + A.stride;
  const uint32_t *__restrict__ nbz = A.nbr + 2 * A.stride;

  if (m > MMAX) {  // beyond capacity: report, leave q = 0
    for (int t = tid; t < m && t < A.stride; t += NT) xhi[t] = 0.0;
    if (tid == 0) {
      *D.iterations = 0;
      *D.converged = 0;
      if (D.failures) atomicAdd(D.failures, 1);
      if (A.overflow) atomicOr(A.overflow, 1);
      if (D.residual) *D.residual = -1.0;
    }
    return;
  }

  // ---- owned cells.  Wave w owns the contiguous chunk [w*ce*64, (w+1)*ce*64) of the list, slot
  // k of lane l is t = (w*ce + k)*64 + l: one base register, slot offsets are immediates.
  const int ce = (m + NT - 1) / NT;
  const int tb = ((tid >> 6) * ce) * 64 + (tid & 63);
#define FOR_OWNED(...)                         \
  _Pragma("unroll") for (int k = 0; k < CPT; k++) { \
    if (k < ce) {                              \
      const int t = tb + k * 64;               \
      if (t < m) { __VA_ARGS__ }               \
    }                                          \
  }
  float r[CPT], xlo[CPT];
  double bmax = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; k++) r[k] = xlo[k] = 0.f;
  FOR_OWNED({
    const double bt = b[t];
    r[k] = float(bt);
    bmax = fmax(bmax, fabs(bt));
    xhi[t] = 0.0;
  })
  const double bnorm = block_max_d(R, rb, bmax);
  if (m == 0 || bnorm < 1e-14) {  // solver3d.ts:882
    if (tid == 0) {
      *D.iterations = 0;
      *D.converged = 1;
      if (D.residual) *D.residual = bnorm;
      if (D.trueResidual) *D.trueResidual = bnorm;
      if (D.bnorm) *D.bnorm = bnorm;
    }
    return;
  }
  const double target = fmax(double(D.tol) * bnorm, 1e-14);
  const bool mic = A.precond == 0;

  // ---- MIC(0): hyperplane ordering and factorisation
  int nplanes = 0;
  __shared__ int planeRange[2];  // first and last non-empty hyperplane
  int *__restrict__ porder = D.porder;      // plane position p -> t
  float *__restrict__ pc2p = D.pc2;         // 1/e per plane position
  uint32_t *__restrict__ nbp = D.nbrPlane;  // 3 x m, plane order
  if (mic) {
    nplanes = A.planes;
    for (int l = tid; l < nplanes; l += NT) planeFill[l] = 0;
    __syncthreads();
    for (int t = tid; t < m; t += NT) atomicAdd(&planeFill[A.plane[t]], 1);
    __syncthreads();
    if (tid == 0) {  // exclusive scan over <= PL3D_MAX_PLANES bins
      int acc = 0, lo = nplanes, hi = -1;
      for (int l = 0; l < nplanes; l++) {
        planeStart[l] = acc;
        if (planeFill[l] > 0) {
          lo = min(lo, l);
          hi = l;
        }
        acc += planeFill[l];
      }
      planeStart[nplanes] = acc;
      planeRange[0] = lo;
      planeRange[1] = hi;
    }
    __syncthreads();
    for (int l = tid; l < nplanes; l += NT) planeFill[l] = 0;
    __syncthreads();
    for (int t = tid; t < m; t += NT) {
      const int l = A.plane[t];
      const int p = planeStart[l] + atomicAdd(&planeFill[l], 1);
      porder[p] = t;
      nbp[p] = nbx[t];
      nbp[p + A.stride] = nby[t];
      nbp[p + 2 * A.stride] = nbz[t];
    }
    __syncthreads();
    // factorisation sweep: W[t] holds 1/e_t (= pc_t^2) once plane(t) is done
    const float tau = 0.97f;
    for (int l = planeRange[0]; l <= planeRange[1]; l++) {
      const int p0 = planeStart[l], p1 = planeStart[l + 1];
      for (int p = p0 + tid; p < p1; p += NT) {
        const int t = porder[p];
        const uint32_t mc = A.mic[t];
        const float d = float(A.diag[t]);
        const int jx = nb_lo(nbp[p]), jy = nb_lo(nbp[p + A.stride]),
                  jz = nb_lo(nbp[p + 2 * A.stride]);
        const float pi2 = jx != int(NO_NBR) ? W[jx] : 0.f;
        const float pj2 = jy != int(NO_NBR) ? W[jy] : 0.f;
        const float pk2 = jz != int(NO_NBR) ? W[jz] : 0.f;
        const float ci = float(mc & 3u), cj = float((mc >> 2) & 3u), ck = float((mc >> 4) & 3u);
        float e = d - pi2 - pj2 - pk2 - tau * (ci * pi2 + cj * pj2 + ck * pk2);
        if (e < 0.25f * d) e = d;
        const float inv = e > 0.f ? 1.0f / e : 0.f;
        W[t] = inv;
        pc2p[p] = inv;
      }
      __syncthreads();
    }
  }

  // W = z = M r.  Owners write W; with MIC(0) the sweeps then run in place.  Ends with a barrier
  // so every owner can read its z_t = W[t].
  auto precondition = [&]() {
    if (!mic) {
      FOR_OWNED({ W[t] = r[k] / float(A.diag[t]); })
      __syncthreads();
      return;
    }
    FOR_OWNED({ W[t] = r[k]; })
    __syncthreads();
    const int lo = planeRange[0], hi = planeRange[1];
    // forward: W_t = (1/e_t) (r_t + sum_{-nb} W_nb).  The next plane's first item is loaded
    // before the barrier so its global-memory latency overlaps the current plane.
    {
      int pt = -1, pj0 = 0, pj1 = 0, pj2 = 0;
      float pinv = 0.f;
      auto fetch = [&](int l) {
        const int p = planeStart[l] + tid;
        pt = -1;
        if (p < planeStart[l + 1]) {
          pt = porder[p];
          pj0 = nb_lo(nbp[p]);
          pj1 = nb_lo(nbp[p + A.stride]);
          pj2 = nb_lo(nbp[p + 2 * A.stride]);
          pinv = pc2p[p];
        }
      };
      if (lo <= hi) fetch(lo);
      for (int l = lo; l <= hi; l++) {
        const int t = pt, j0 = pj0, j1 = pj1, j2 = pj2;
        const float inv = pinv;
        if (l < hi) fetch(l + 1);
        if (t >= 0) {
          float v = W[t];
          if (j0 != int(NO_NBR)) v += W[j0];
          if (j1 != int(NO_NBR)) v += W[j1];
          if (j2 != int(NO_NBR)) v += W[j2];
          W[t] = inv * v;
        }
        for (int p = planeStart[l] + tid + NT; p < planeStart[l + 1]; p += NT) {  // planes > NT
          const int tt = porder[p];
          const int k0 = nb_lo(nbp[p]), k1 = nb_lo(nbp[p + A.stride]),
                    k2 = nb_lo(nbp[p + 2 * A.stride]);
          float v = W[tt];
          if (k0 != int(NO_NBR)) v += W[k0];
          if (k1 != int(NO_NBR)) v += W[k1];
          if (k2 != int(NO_NBR)) v += W[k2];
          W[tt] = pc2p[p] * v;
        }
        __syncthreads();
      }
    }
    // backward: z_t = W_t + (1/e_t) sum_{+nb} z_nb
    {
      int pt = -1, pj0 = 0, pj1 = 0, pj2 = 0;
      float pinv = 0.f;
      auto fetch = [&](int l) {
        const int p = planeStart[l] + tid;
        pt = -1;
        if (p < planeStart[l + 1]) {
          pt = porder[p];
          pj0 = nb_hi(nbp[p]);
          pj1 = nb_hi(nbp[p + A.stride]);
          pj2 = nb_hi(nbp[p + 2 * A.stride]);
          pinv = pc2p[p];
        }
      };
      if (lo <= hi) fetch(hi);
      for (int l = hi; l >= lo; l--) {
        const int t = pt, j0 = pj0, j1 = pj1, j2 = pj2;
        const float inv = pinv;
        if (l > lo) fetch(l - 1);
        if (t >= 0) {
          float v = 0.f;
          if (j0 != int(NO_NBR)) v += W[j0];
          if (j1 != int(NO_NBR)) v += W[j1];
          if (j2 != int(NO_NBR)) v += W[j2];
          W[t] = W[t] + inv * v;
        }
        for (int p = planeStart[l] + tid + NT; p < planeStart[l + 1]; p += NT) {
          const int tt = porder[p];
          const int k0 = nb_hi(nbp[p]), k1 = nb_hi(nbp[p + A.stride]),
                    k2 = nb_hi(nbp[p + 2 * A.stride]);
          float v = 0.f;
          if (k0 != int(NO_NBR)) v += W[k0];
          if (k1 != int(NO_NBR)) v += W[k1];
          if (k2 != int(NO_NBR)) v += W[k2];
          W[tt] = W[tt] + pc2p[p] * v;
        }
        __syncthreads();
      }
    }
  };

  // FP64 true residual r = b - A x_hi (refinement).  Returns max|r_true| over the block and, in
  // gap, max|r_true - r_recursive|.
  auto true_residual = [&](double &gap) -> double {
    double mx = 0.0, gp = 0.0;
    FOR_OWNED({
      xhi[t] += double(xlo[k]);
      xlo[k] = 0.f;
    })
    __syncthreads();
    FOR_OWNED({
      const uint32_t px = nbx[t], py = nby[t], pz = nbz[t];
      double acc = double(A.diag[t]) * xhi[t];
      int j;
      if ((j = nb_hi(px)) != int(NO_NBR)) acc -= xhi[j];
      if ((j = nb_lo(px)) != int(NO_NBR)) acc -= xhi[j];
      if ((j = nb_hi(py)) != int(NO_NBR)) acc -= xhi[j];
      if ((j = nb_lo(py)) != int(NO_NBR)) acc -= xhi[j];
      if ((j = nb_hi(pz)) != int(NO_NBR)) acc -= xhi[j];
      if ((j = nb_lo(pz)) != int(NO_NBR)) acc -= xhi[j];
      const double rt = b[t] - acc;
      gp = fmax(gp, fabs(rt - double(r[k])));
      r[k] = float(rt);
      mx = fmax(mx, fabs(rt));
    })
    gap = block_max_d(R, rb, gp);
    return block_max_d(R, rb, mx);
  };

  // ---- PCG (solver3d.ts:902-936); s lives in LDS only (s_lds), z in W after precondition()
  precondition();
  double sig_part = 0.0;
  FOR_OWNED({
    const float zt = W[t];
    s_lds[t] = zt;
    sig_part += double(zt) * double(r[k]);
  })
  double sigma = block_sum(R, rb, sig_part);
  double res = bnorm;
  double resAtUpdate = bnorm;
  int iters = 0;
  bool converged = false;
  int lastUpdate = 0;
  for (int it = 1; it <= D.maxIter; it++) {
    __syncthreads();  // s complete
    double sas_part = 0.0;
    FOR_OWNED({  // same neighbour order as applyA (solver3d.ts:944-951)
      const uint32_t px = nbx[t], py = nby[t], pz = nbz[t];
      const float st = s_lds[t];
      float acc = float(A.diag[t]) * st;
      int j;
      if ((j = nb_hi(px)) != int(NO_NBR)) acc -= s_lds[j];
      if ((j = nb_lo(px)) != int(NO_NBR)) acc -= s_lds[j];
      if ((j = nb_hi(py)) != int(NO_NBR)) acc -= s_lds[j];
      if ((j = nb_lo(py)) != int(NO_NBR)) acc -= s_lds[j];
      if ((j = nb_hi(pz)) != int(NO_NBR)) acc -= s_lds[j];
      if ((j = nb_lo(pz)) != int(NO_NBR)) acc -= s_lds[j];
      W[t] = acc;  // A s, owner-private until the next precondition
      sas_part += double(st) * double(acc);
    })
    const double sas = block_sum(R, rb, sas_part);
    if (!(sas > 0.0)) break;  // solver3d.ts:916
    const double alpha = sigma / sas;
    const float fa = float(alpha);
    float rmax = 0.f;
    FOR_OWNED({
      xlo[k] += fa * s_lds[t];
      r[k] -= fa * W[t];
      rmax = fmaxf(rmax, fabsf(r[k]));
    })
    __syncthreads();  // every A s read before W is reused
    precondition();
    double sn_part = 0.0;
    FOR_OWNED({ sn_part += double(W[t]) * double(r[k]); })
    double sn;
    float resf;
    block_sum_max(R, rb, sn_part, rmax, sn, resf);
    res = double(resf);
    iters = it;
    bool restart = false;
    // reliable update: at the stop test, after each decade of residual reduction, and at least
    // every RU_PERIOD iterations
    if (res <= target || res < 0.1 * resAtUpdate || it - lastUpdate >= RU_PERIOD) {
      lastUpdate = it;
      double gap;
      res = true_residual(gap);
      resAtUpdate = res;
      if (res <= target) {
        converged = true;
        break;
      }
      if (gap > 0.01 * double(resf)) {  // FP32 drift visible: restart from the true residual
        restart = true;
        precondition();
        double p2 = 0.0;
        FOR_OWNED({ p2 += double(W[t]) * double(r[k]); })
        sn = block_sum(R, rb, p2);
      }
    }
    const float fb = restart ? 0.f : float(sn / sigma);
    FOR_OWNED({ s_lds[t] = W[t] + fb * s_lds[t]; })
    sigma = sn;
  }
  double tres;
  if (!converged) {
    // the loop ran out (or broke on s.As <= 0): the reference reports maxIter (solver3d.ts:936)
    double gap;
    tres = true_residual(gap);  // folds the last increments into x_hi
    iters = D.maxIter;
    converged = tres <= target;
    res = tres;
  } else {
    tres = res;
  }
#undef FOR_OWNED
  if (tid == 0) {
    if (A.cellIterations) {
      atomicAdd(A.cellIterations, (unsigned long long)m * (unsigned long long)iters);
      atomicAdd(A.solveCount, 1ull);
    }
    *D.iterations = iters;
    *D.converged = converged ? 1 : 0;
    if (!converged && D.failures) atomicAdd(D.failures, 1);
    if (D.residual) *D.residual = res;
    if (D.trueResidual) *D.trueResidual = tres;
    if (D.bnorm) *D.bnorm = bnorm;
  }
}
