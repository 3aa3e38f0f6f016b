// The learned candidate-depth predictor: a 28-32-32-2 perceptron (FP32)
// with hand-written forward, backward and Adam kernels (learn.hip).
//
// Input, per block of the grid (or per particle in radical mode): the
// particle counts of the 3x3x3 blocks around it, divided by the mean count
// per block, and (radical mode) the particle's radius divided by the largest
// radius; the 28th input is 0 otherwise.  Outputs: the worklist depth at
// which the security-radius test stops the cell, /64, and log(1 + nplane
// calls)/8.  The depth used at run time is ceil(64 y0 + margin), the margin
// set so that a chosen fraction of the training cells would fall back.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace vgpu {
namespace learn {

enum { IN = 28, H = 32, OUT = 2 };
enum {
	oW1 = 0,                 // H x IN
	oB1 = oW1 + H * IN,      // H
	oW2 = oB1 + H,           // H x H
	oB2 = oW2 + H * H,       // H
	oW3 = oB2 + H,           // OUT x H
	oB3 = oW3 + OUT * H,     // OUT
	NP = oB3 + OUT
};
enum { ACT = H + H + OUT };  // per-sample activations / deltas

// Device buffers of one training run or one inference pass.
struct MLPBufs {
	float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *part = nullptr;
	float *X = nullptr, *T = nullptr, *A = nullptr, *D = nullptr, *L = nullptr, *loss = nullptr;
	int n = 0, chunks = 0;
};

// Forward + loss + gradient of L = (1/n) sum_s sum_o (y_so - t_so)^2 for n
// samples; params/X/T/grad are device pointers.  Deterministic (fixed
// reduction order).
int mlp_loss_grad_dev(const float *P, const float *X, const float *T, int n, float *A, float *D, float *Lbuf,
		      float *part, int chunks, float *grad, float *loss, hipStream_t s);
// One Adam step.
int mlp_adam_dev(float *P, const float *G, float *M, float *V, float lr, int t, hipStream_t s);
// Depth per sample: d = max(1, ceil(64 y0 + margin)); 64 (no list: the
// block search is predicted) where d >= route (route <= 64), else min(d, 63);
// cost (may be
// null): the predicted log(1 + nplane calls) in steps of 1/8, 0..63.
int mlp_depth_dev(const float *P, const float *X, int n, float margin, int route, unsigned char *depth,
		  unsigned char *cost, hipStream_t s);
// Raw outputs per sample (n x OUT).
int mlp_forward_dev(const float *P, const float *X, int n, float *Y, hipStream_t s);
// Block features (27 neighbour counts / mean; 28th = 0) of the blocks of an
// nx x ny x nz grid.  A periodic container passes its block grid
// (nx x gy x gz, real rows from j0, k0; one sample per block of it, zero for
// image blocks) and wrap = true.
int block_features_dev(int nx, int ny, int nz, const int *bstart, float inv_mean, float *X, hipStream_t s, int gy = 0,
		       int gz = 0, int j0 = 0, int k0 = 0, bool wrap = false);
// Per-cell features for a predictor that takes the radius: the cell's block
// features with input 27 = r / max_r (r read with a stride of 4 doubles).
int cell_features_dev(int nc, const int *cellblk, const double *r, float inv_maxr, const float *Xb, float *Xc,
		      hipStream_t s);
// Per-block labels from per-cell statistics: max worklist exit and max
// nplane calls over the block's cells.
int block_labels_dev(int nc, const int *cellblk, const signed char *st_exit, const short *st_planes, int nblocks,
		     int *lab, hipStream_t s);

}  // namespace learn
}  // namespace vgpu

// The predictor object behind the C ABI.
struct vgpu_predictor {
	int with_radius = 0;
	std::vector<float> X, T;          // training samples (host)
	std::vector<float> params;        // NP floats
	float margin = 0;                 // depth units
	int route = 64;                   // depths from here on: no list (block search predicted)
	bool trained = false;
	long long seen_blocks = 0;
};
