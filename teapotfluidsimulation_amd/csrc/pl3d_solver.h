// pl3d_solver.h — arguments of the one-workgroup PCG kernel (k_pcg in pl3d_solver.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PL3D_SOLVE_THREADS 1024
#define PL3D_SOLVE_MMAX 16384  // LDS slots per solve; one is the zero slot, so m <= 16383
#define PL3D_MAX_PLANES 1024   // hyperplanes i+j+k of the MIC(0) schedule

// Liquid-cell structure shared by the solves of one launch, all in list order t = rank of the
// FLUID cell in grid index order.
struct PcgShared {
  const int *m;            // device scalar: number of liquid cells
  const uint32_t *nbr;     // 3 x stride words: (-x | +x << 16), (-y | +y << 16), (-z | +z << 16)
  const uint8_t *diag;     // non-solid neighbour count (solver3d.ts:875-876)
  const uint8_t *mic;      // MIC(0) coupling counts of the -x/-y/-z neighbours, 2 bits each
  const uint16_t *plane;   // i + j + k
  int stride;              // capacity of the list arrays (>= m)
  int planes;              // number of hyperplanes (nx + ny + nz - 2)
  int precond;             // 0 = MIC(0), 2 = Jacobi
  int *overflow;           // set when m exceeds PL3D_SOLVE_MMAX
  unsigned long long *cellIterations;  // += m * iterations per solve (nullptr: not counted)
  unsigned long long *solveCount;
  const int *active;  // the substep's run flag (nullptr: always run)
};

// One solve.
struct PcgDesc {
  const double *b;  // rhs, list order
  double *q;        // solution, list order (FP64)
  float tol;
  int maxIter;
  int *iterations;
  int *converged;
  int *failures;         // incremented when not converged (nullptr: not counted)
  double *residual;      // max|r| at exit
  double *trueResidual;  // FP64 max|b - A q| at exit
  double *bnorm;
  // scratch, capacity stride
  int *porder;
  float *pc2;
  uint32_t *nbrPlane;
};

// Grid (solves per member, members): block (x, y) runs solve Ds[y * gridDim.x + x] on the liquid
// cells described by At[y].
__global__ void k_pcg_mic0(const PcgShared *At, const PcgDesc *Ds);
__global__ void k_pcg_jacobi(const PcgShared *At, const PcgDesc *Ds);
