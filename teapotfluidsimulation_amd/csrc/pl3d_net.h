// pl3d_net.h — the learned pressure preconditioner: a three-level encoder-decoder of 3x3x3
// convolutions (8/16/32 channels) over the pressure grid, its BF16-MFMA kernels and the flexible
// CG that applies it (pl3d_net.hip, pl3d_net_api.cpp).  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Network (all convolutions 3x3x3, zero outside the grid, no bias; level l+1 is level l pooled
// 2x2x2 with dims ceil(n/2)):
//   x0 = [r/s, FLUID, SOLID, 0]                         level 1, 4 channels
//   e1 = relu(conv_a(x0))            4 -> 8             level 1
//   e2 = relu(conv_b(pool(e1)))      8 -> 16            level 2
//   e3 = relu(conv_c(pool(e2)))     16 -> 32            level 3
//   h2 = relu(conv_d(up(e3))),  d2 = h2 + e2  32 -> 16  level 2
//   h1 = relu(conv_e(up(d2))),  d1 = h1 + e1  16 -> 8   level 1
//   z  = conv_f(d1) * FLUID          8 -> 1             level 1
// Activations are BF16, channel-last [voxel][C] with C a multiple of 4; products accumulate in FP32.
#define NET_LAYERS 6

// Per-layer constants: input/output channels (padded to 4), output channels padded to the MFMA
// tile (16), the reduction length 27 * cin padded to 16, and the level it runs on (0-based).
struct NetLayer {
  int cin, cout, coutp, kp, level;
  int wOff;  // offset of this layer's [coutp][kp] weights in the parameter arrays
};

__host__ __device__ constexpr int net_kp(int cin) { return (27 * cin + 15) / 16 * 16; }

// Epilogues of the convolution kernel
enum {
  EPI_RELU = 0,      // y = relu(acc) (BF16)
  EPI_RELU_ADD = 1,  // h = relu(acc), y = h + skip: writes h to out and y to out2 (BF16)
  EPI_MASK_F32 = 2,  // z = acc[channel 0] * FLUID (FP32, one value per voxel)
  EPI_PLAIN = 3,     // y = acc (BF16): input gradient of a layer
  EPI_GRAD_SPLIT = 4 // out = acc * (act > 0), out2 = acc (BF16): gradient through a relu + skip
};

struct ConvArgs {
  const uint16_t *in;    // [V][cin] BF16
  const uint16_t *w;     // [coutp][kp] BF16 (kp index = tap * cin + ci, tap = dx + 3 (dy + 3 dz))
  uint16_t *out;         // [V][cout] BF16 (EPI_MASK_F32: unused)
  uint16_t *out2;        // second output of EPI_RELU_ADD / EPI_GRAD_SPLIT
  const uint16_t *skip;  // EPI_RELU_ADD: skip input [V][cout]; EPI_GRAD_SPLIT: activation
  float *outf;           // EPI_MASK_F32 output [V]
  const uint8_t *flags;  // EPI_MASK_F32: cell flags of the level-1 grid
  int nx, ny, nz;
};

struct WgradArgs {
  const uint16_t *in;  // layer input [V][cin]
  const uint16_t *dy;  // output gradient [V][cout rounded up to 4]
  float *partial;      // [chunks][coutp][kp] FP32
  int nx, ny, nz;
  int chunk;           // voxels per chunk (multiple of 16)
};

#define NET_WG_CHUNK 1024
// flexible CG: damped-Jacobi steps applied to the network's output (z += w D^-1 (r - A z))
#define NET_POST_SMOOTH 4
#define NET_SMOOTH_OMEGA 0.6f
// the host reads the flexible CG's state every NET_FCG_POLL iterations (the iterations enqueued
// after convergence find `done` set and change nothing)
#define NET_FCG_POLL 4

// Scalars of the flexible CG, kept on the device so an iteration needs no host round trip; the
// single-block kernels k_fcg_alpha / k_fcg_rcheck / k_fcg_tcheck / k_fcg_dcheck update them
// between the vector kernels, which read alpha, beta and the input scale from here.
struct FcgState {
  double zr;         // z.r of the current search direction's z
  double alpha, beta;
  double rmax;       // max|r|, recursive or after a true-residual refresh (the network's scale)
  double target;     // stop when max|r| <= target
  int it;            // iterations taken
  int lastTrue;      // iteration of the last FP64 true-residual refresh
  int done, converged;
  int needTrue;      // this iteration refreshes the true residual
  int fallback;      // this iteration's z is the Jacobi fallback
  int fallbacks;     // count of those
  int init;          // the first z: beta = 0
};
