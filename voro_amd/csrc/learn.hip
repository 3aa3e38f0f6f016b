// Kernels and host driver of the learned candidate-depth predictor.
//
// Training is full-batch Adam on the GPU: forward (weights staged in LDS),
// per-sample backward deltas, then each parameter's gradient summed over
// fixed chunks of samples and the chunk sums added in a fixed order, so a
// training run is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/voro_gpu.h"
#include "learn.h"

namespace vgpu {
namespace learn {

namespace {

__device__ __forceinline__ void forward1(const float *W, const float *x, float *h1, float *h2, float *y) {
	for (int i = 0; i < H; i++) {
		float a = W[oB1 + i];
		for (int k = 0; k < IN; k++) a += W[oW1 + i * IN + k] * x[k];
		h1[i] = a > 0.f ? a : 0.f;
	}
	for (int j = 0; j < H; j++) {
		float a = W[oB2 + j];
		for (int i = 0; i < H; i++) a += W[oW2 + j * H + i] * h1[i];
		h2[j] = a > 0.f ? a : 0.f;
	}
	for (int o = 0; o < OUT; o++) {
		float a = W[oB3 + o];
		for (int j = 0; j < H; j++) a += W[oW3 + o * H + j] * h2[j];
		y[o] = a;
	}
}

// Activations per sample: h1 (H), h2 (H), y (OUT).
__global__ void k_fwd(const float *__restrict__ P, const float *__restrict__ X, int n, float *__restrict__ A) {
	__shared__ float W[NP];
	for (int t = threadIdx.x; t < NP; t += blockDim.x) W[t] = P[t];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	float x[IN], h1[H], h2[H], y[OUT];
	for (int k = 0; k < IN; k++) x[k] = X[(size_t)s * IN + k];
	forward1(W, x, h1, h2, y);
	float *a = A + (size_t)s * ACT;
	for (int i = 0; i < H; i++) a[i] = h1[i];
	for (int j = 0; j < H; j++) a[H + j] = h2[j];
	for (int o = 0; o < OUT; o++) a[2 * H + o] = y[o];
}

// Deltas per sample: dz1 (H), dz2 (H), dy (OUT), and the sample's loss.
__global__ void k_bwd(const float *__restrict__ P, const float *__restrict__ A, const float *__restrict__ T, int n,
		      float *__restrict__ D, float *__restrict__ L) {
	__shared__ float W[NP];
	for (int t = threadIdx.x; t < NP; t += blockDim.x) W[t] = P[t];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	const float *a = A + (size_t)s * ACT;
	float dy[OUT], dz2[H];
	float ls = 0.f;
	const float sc = 2.f / (float)n;
	for (int o = 0; o < OUT; o++) {
		const float e = a[2 * H + o] - T[(size_t)s * OUT + o];
		ls += e * e;
		dy[o] = sc * e;
	}
	float *d = D + (size_t)s * ACT;
	for (int j = 0; j < H; j++) {
		float g = 0.f;
		for (int o = 0; o < OUT; o++) g += W[oW3 + o * H + j] * dy[o];
		dz2[j] = a[H + j] > 0.f ? g : 0.f;
		d[H + j] = dz2[j];
	}
	for (int i = 0; i < H; i++) {
		float g = 0.f;
		for (int j = 0; j < H; j++) g += W[oW2 + j * H + i] * dz2[j];
		d[i] = a[i] > 0.f ? g : 0.f;
	}
	for (int o = 0; o < OUT; o++) d[2 * H + o] = dy[o];
	L[s] = ls / (float)n;
}

// Partial gradient of parameter p over sample chunk c (samples c, c+chunks, ...).
__global__ void k_grad(const float *__restrict__ X, const float *__restrict__ A, const float *__restrict__ D, int n,
		       int chunks, float *__restrict__ part) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	const int c = blockIdx.y;
	if (p >= NP) return;
	float acc = 0.f;
	// which product this parameter is
	int kind, r, q;
	if (p < oB1) { kind = 0; r = p / IN; q = p % IN; }
	else if (p < oW2) { kind = 1; r = p - oB1; q = 0; }
	else if (p < oB2) { kind = 2; r = (p - oW2) / H; q = (p - oW2) % H; }
	else if (p < oW3) { kind = 3; r = p - oB2; q = 0; }
	else if (p < oB3) { kind = 4; r = (p - oW3) / H; q = (p - oW3) % H; }
	else { kind = 5; r = p - oB3; q = 0; }
	for (int s = c; s < n; s += chunks) {
		const float *a = A + (size_t)s * ACT, *d = D + (size_t)s * ACT;
		float v;
		switch (kind) {
		case 0: v = d[r] * X[(size_t)s * IN + q]; break;      // dW1 = dz1 x
		case 1: v = d[r]; break;                              // db1
		case 2: v = d[H + r] * a[q]; break;                   // dW2 = dz2 h1
		case 3: v = d[H + r]; break;                          // db2
		case 4: v = d[2 * H + r] * a[H + q]; break;           // dW3 = dy h2
		default: v = d[2 * H + r]; break;                     // db3
		}
		acc += v;
	}
	part[(size_t)c * NP + p] = acc;
}

// out[p] = sum over c of part[c][p], c in increasing order.
__global__ void k_reduce_chunks(const float *__restrict__ part, int chunks, int np, float *__restrict__ out) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= np) return;
	float acc = 0.f;
	for (int c = 0; c < chunks; c++) acc += part[(size_t)c * np + p];
	out[p] = acc;
}

// Sum of n floats by one workgroup, fixed order.
__global__ void k_sum(const float *__restrict__ v, int n, float *__restrict__ out) {
	__shared__ float sh[256];
	float acc = 0.f;
	for (int i = threadIdx.x; i < n; i += 256) acc += v[i];
	sh[threadIdx.x] = acc;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) *out = sh[0];
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M, float *__restrict__ V,
		       float lr, float b1, float b2, float eps, float bc1, float bc2) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP) return;
	const float g = G[p];
	const float m = b1 * M[p] + (1.f - b1) * g;
	const float v = b2 * V[p] + (1.f - b2) * g * g;
	M[p] = m;
	V[p] = v;
	P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

__global__ void k_depth(const float *__restrict__ P, const float *__restrict__ X, int n, float margin, int route,
			unsigned char *__restrict__ depth, unsigned char *__restrict__ cost) {
	__shared__ float W[NP];
	for (int t = threadIdx.x; t < NP; t += blockDim.x) W[t] = P[t];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	float x[IN], h1[H], h2[H], y[OUT];
	for (int k = 0; k < IN; k++) x[k] = X[(size_t)s * IN + k];
	forward1(W, x, h1, h2, y);
	// 64: the cell is predicted to go past the worklist into the block
	// search, which no list covers; it takes the plain path at once
	float g = ceilf(64.f * y[0] + margin);
	if (!(g >= 1.f)) g = 1.f;
	if (route <= 64 && g >= (float)route) g = 64.f;
	else if (g > 63.f) g = 63.f;
	depth[s] = (unsigned char)g;
	if (cost) {
		// predicted log(1 + nplane calls) in 1/8 steps, 0..63
		float c = 64.f * y[1];
		if (!(c >= 0.f)) c = 0.f;
		if (c > 63.f) c = 63.f;
		cost[s] = (unsigned char)c;
	}
}

__global__ void k_forward_out(const float *__restrict__ P, const float *__restrict__ X, int n, float *__restrict__ Y) {
	__shared__ float W[NP];
	for (int t = threadIdx.x; t < NP; t += blockDim.x) W[t] = P[t];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n) return;
	float x[IN], h1[H], h2[H], y[OUT];
	for (int k = 0; k < IN; k++) x[k] = X[(size_t)s * IN + k];
	forward1(W, x, h1, h2, y);
	for (int o = 0; o < OUT; o++) Y[(size_t)s * OUT + o] = y[o];
}

// 27 neighbour-block counts around each block (zero outside the grid).
// Block features: the particle counts of the 27 blocks around a block,
// over the mean count.  Blocks are indexed in a grid nx x gy x gz whose
// real blocks are j0 <= j < j0 + ny, k0 <= k < k0 + nz (a periodic
// container's image rows around them; plain containers have gy = ny,
// j0 = 0); wrap: neighbours wrap around the real grid (periodic containers).
__global__ void k_block_features(int nx, int ny, int nz, int gy, int gz, int j0, int k0, int wrap,
				 const int *__restrict__ bstart, float inv_mean, float *__restrict__ X) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	const int nb = nx * gy * gz;
	if (b >= nb) return;
	const int i = b % nx, j = (b / nx) % gy - j0, k = b / (nx * gy) - k0;
	const bool real = j >= 0 && j < ny && k >= 0 && k < nz;
	int t = 0;
	for (int dk = -1; dk <= 1; dk++)
		for (int dj = -1; dj <= 1; dj++)
			for (int di = -1; di <= 1; di++, t++) {
				int a = i + di, c = j + dj, e = k + dk;
				if (wrap) {
					a = a < 0 ? a + nx : (a >= nx ? a - nx : a);
					c = c < 0 ? c + ny : (c >= ny ? c - ny : c);
					e = e < 0 ? e + nz : (e >= nz ? e - nz : e);
				}
				float v = 0.f;
				if (real && a >= 0 && a < nx && c >= 0 && c < ny && e >= 0 && e < nz) {
					const int q = a + nx * ((c + j0) + gy * (e + k0));
					v = (float)(bstart[q + 1] - bstart[q]) * inv_mean;
				}
				X[(size_t)b * IN + t] = v;
			}
	X[(size_t)b * IN + 27] = 0.f;
}

__global__ void k_cell_features(int nc, const int *__restrict__ cellblk, const double *__restrict__ r, float inv_maxr,
				const float *__restrict__ Xb, float *__restrict__ Xc) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nc) return;
	const float *a = Xb + (size_t)cellblk[c] * IN;
	float *o = Xc + (size_t)c * IN;
	for (int k = 0; k < 27; k++) o[k] = a[k];
	o[27] = (float)r[(size_t)c * 4] * inv_maxr;
}

__global__ void k_block_labels(int nc, const int *__restrict__ cellblk, const signed char *__restrict__ st_exit,
			       const short *__restrict__ st_planes, int *__restrict__ lab) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nc) return;
	const int b = cellblk[c];
	atomicMax(&lab[2 * b], (int)st_exit[c]);
	atomicMax(&lab[2 * b + 1], (int)st_planes[c]);
}

int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

int mlp_loss_grad_dev(const float *P, const float *X, const float *T, int n, float *A, float *D, float *Lbuf,
		      float *part, int chunks, float *grad, float *loss, hipStream_t s) {
	if (n <= 0) return VGPU_INVALID_ARGUMENT;
	hipLaunchKernelGGL(k_fwd, dim3(nblk(n, 256)), dim3(256), 0, s, P, X, n, A);
	hipLaunchKernelGGL(k_bwd, dim3(nblk(n, 256)), dim3(256), 0, s, P, A, T, n, D, Lbuf);
	hipLaunchKernelGGL(k_grad, dim3(nblk(NP, 256), chunks), dim3(256), 0, s, X, A, D, n, chunks, part);
	hipLaunchKernelGGL(k_reduce_chunks, dim3(nblk(NP, 256)), dim3(256), 0, s, part, chunks, (int)NP, grad);
	if (loss) hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, Lbuf, n, loss);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int mlp_adam_dev(float *P, const float *G, float *M, float *V, float lr, int t, hipStream_t s) {
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
	hipLaunchKernelGGL(k_adam, dim3(nblk(NP, 256)), dim3(256), 0, s, P, G, M, V, lr, b1, b2, eps, bc1, bc2);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int mlp_depth_dev(const float *P, const float *X, int n, float margin, int route, unsigned char *depth,
		  unsigned char *cost, hipStream_t s) {
	if (n <= 0) return VGPU_OK;
	hipLaunchKernelGGL(k_depth, dim3(nblk(n, 256)), dim3(256), 0, s, P, X, n, margin, route, depth, cost);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int mlp_forward_dev(const float *P, const float *X, int n, float *Y, hipStream_t s) {
	if (n <= 0) return VGPU_OK;
	hipLaunchKernelGGL(k_forward_out, dim3(nblk(n, 256)), dim3(256), 0, s, P, X, n, Y);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int block_features_dev(int nx, int ny, int nz, const int *bstart, float inv_mean, float *X, hipStream_t s, int gy,
		       int gz, int j0, int k0, bool wrap) {
	if (gy <= 0) { gy = ny; gz = nz; j0 = k0 = 0; }
	const int nb = nx * gy * gz;
	hipLaunchKernelGGL(k_block_features, dim3(nblk(nb, 256)), dim3(256), 0, s, nx, ny, nz, gy, gz, j0, k0,
			   wrap ? 1 : 0, bstart, inv_mean, X);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int cell_features_dev(int nc, const int *cellblk, const double *r, float inv_maxr, const float *Xb, float *Xc,
		      hipStream_t s) {
	if (nc <= 0) return VGPU_OK;
	hipLaunchKernelGGL(k_cell_features, dim3(nblk(nc, 256)), dim3(256), 0, s, nc, cellblk, r, inv_maxr, Xb, Xc);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int block_labels_dev(int nc, const int *cellblk, const signed char *st_exit, const short *st_planes, int nblocks,
		     int *lab, hipStream_t s) {
	if (hipMemsetAsync(lab, 0, (size_t)nblocks * 2 * sizeof(int), s) != hipSuccess) return VGPU_HIP_ERROR;
	if (nc > 0) hipLaunchKernelGGL(k_block_labels, dim3(nblk(nc, 256)), dim3(256), 0, s, nc, cellblk, st_exit, st_planes, lab);
	return hipGetLastError() == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

}  // namespace learn
}  // namespace vgpu

using namespace vgpu::learn;

namespace {

struct DevMem {
	std::vector<void *> ptrs;
	~DevMem() { for (void *p : ptrs) (void)hipFree(p); }
	template<class T> T *get(size_t n) {
		void *p = nullptr;
		if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
		ptrs.push_back(p);
		return (T *)p;
	}
};

void init_params(std::vector<float> &P, unsigned seed) {
	std::mt19937 g(seed);
	P.assign(NP, 0.f);
	auto fill = [&](int off, int rows, int cols) {
		std::normal_distribution<float> d(0.f, std::sqrt(2.f / (float)cols));
		for (int t = 0; t < rows * cols; t++) P[off + t] = d(g);
	};
	fill(oW1, H, IN);
	fill(oW2, H, H);
	fill(oW3, OUT, H);
	for (int o = 0; o < OUT; o++) P[oB3 + o] = 0.5f;
}

}  // namespace

// ---- Loss and gradient without allocating ----
// The scratch of mlp_loss_grad_dev (activations and deltas, n x ACT floats
// each; per-sample losses, n; gradient partial sums, LG_CHUNKS x NP) for up
// to NOALLOC_MAX_SAMPLES samples lives in one file-scope device array: every
// GPU the library runs on holds its own copy (17.6 MB) until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host the
// name is a stand-in, not a device address): each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current.
enum { LG_CHUNKS = 16 };   // the chunk count of vgpu_mlp_loss_grad (same summation order)
enum : size_t {
	NOALLOC_MAX_SAMPLES = VGPU_MLP_NOALLOC_MAX_SAMPLES,
	SCR_A = 0,
	SCR_D = SCR_A + NOALLOC_MAX_SAMPLES * ACT,
	SCR_L = SCR_D + NOALLOC_MAX_SAMPLES * ACT,
	SCR_PART = SCR_L + NOALLOC_MAX_SAMPLES,
	SCR_FLOATS = SCR_PART + (size_t)LG_CHUNKS * NP
};
static __device__ __align__(256) float scratch[SCR_FLOATS];

namespace {

enum { MAX_DEVICES = 64 };
std::mutex g_scratch_mu;
float *g_scratch_ptr[MAX_DEVICES];

// The current device's scratch address (looked up on its first use).
int scratch_of_current_device(float **out) {
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return VGPU_HIP_ERROR;
	if (dev < 0 || dev >= MAX_DEVICES) return VGPU_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		float *ptr = nullptr;
		if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return VGPU_HIP_ERROR;
		g_scratch_ptr[dev] = ptr;
	}
	*out = g_scratch_ptr[dev];
	return VGPU_OK;
}

}  // namespace

extern "C" {

int vgpu_mlp_loss_grad_noalloc(const float *params, const float *X, const float *T, int n, float *loss, float *grad,
			       void *stream) {
	if (!params || !X || !T || !grad || n <= 0) return VGPU_INVALID_ARGUMENT;
	if ((size_t)n > NOALLOC_MAX_SAMPLES) return VGPU_INVALID_ARGUMENT;
	float *scr = nullptr;
	int st = scratch_of_current_device(&scr);
	if (st) return st;
	hipStream_t s = (hipStream_t)stream;
	// every scratch word read is written first in the same call: nothing to clear
	st = mlp_loss_grad_dev(params, X, T, n, scr + SCR_A, scr + SCR_D, scr + SCR_L, scr + SCR_PART, LG_CHUNKS, grad,
			       loss, s);
	if (st) return st;
	return hipStreamSynchronize(s) == hipSuccess ? VGPU_OK : VGPU_HIP_ERROR;
}

int vgpu_predictor_create(int with_radius, vgpu_predictor **out) {
	if (!out) return VGPU_INVALID_ARGUMENT;
	*out = new vgpu_predictor();
	(*out)->with_radius = with_radius ? 1 : 0;
	return VGPU_OK;
}

int vgpu_predictor_destroy(vgpu_predictor *p) {
	delete p;
	return VGPU_OK;
}

int vgpu_predictor_num_samples(vgpu_predictor *p, long long *n) {
	if (!p || !n) return VGPU_INVALID_ARGUMENT;
	*n = (long long)(p->T.size() / OUT);
	return VGPU_OK;
}

int vgpu_predictor_params(vgpu_predictor *p, float *get, const float *set, int n) {
	if (!p || n != NP) return VGPU_INVALID_ARGUMENT;
	if (set) { p->params.assign(set, set + NP); p->trained = true; }
	if (get) {
		if (p->params.size() != (size_t)NP) return VGPU_INVALID_ARGUMENT;
		std::memcpy(get, p->params.data(), sizeof(float) * NP);
	}
	return VGPU_OK;
}

int vgpu_predictor_margin(vgpu_predictor *p, float *get, const float *set) {
	if (!p) return VGPU_INVALID_ARGUMENT;
	if (set) p->margin = *set;
	if (get) *get = p->margin;
	return VGPU_OK;
}

int vgpu_predictor_route(vgpu_predictor *p, int *get, const int *set) {
	if (!p || (set && *set < 1)) return VGPU_INVALID_ARGUMENT;
	if (set) p->route = *set;
	if (get) *get = p->route;
	return VGPU_OK;
}

// Trains on the collected samples (full-batch Adam, `epochs` steps), then
// sets the depth margin so that about `fallback` of the samples would have
// been under-predicted.  Parameters start from a seeded He initialisation
// unless some were set or trained before.
int vgpu_predictor_train(vgpu_predictor *p, int epochs, float lr, unsigned seed, float fallback, float *loss_out) {
	if (!p || epochs < 0 || !(lr > 0) || !(fallback >= 0 && fallback < 1)) return VGPU_INVALID_ARGUMENT;
	const int n = (int)(p->T.size() / OUT);
	if (n <= 0) return VGPU_INVALID_ARGUMENT;
	if (p->params.size() != (size_t)NP) init_params(p->params, seed);
	DevMem mem;
	const int chunks = 64;
	float *P = mem.get<float>(NP), *G = mem.get<float>(NP), *M = mem.get<float>(NP), *V = mem.get<float>(NP);
	float *X = mem.get<float>((size_t)n * IN), *T = mem.get<float>((size_t)n * OUT);
	float *A = mem.get<float>((size_t)n * ACT), *D = mem.get<float>((size_t)n * ACT), *L = mem.get<float>(n);
	float *part = mem.get<float>((size_t)chunks * NP), *loss = mem.get<float>(1), *Y = mem.get<float>((size_t)n * OUT);
	if (!P || !G || !M || !V || !X || !T || !A || !D || !L || !part || !loss || !Y) return VGPU_MEMORY_ERROR;
	hipStream_t s = nullptr;
	if (hipMemcpy(P, p->params.data(), sizeof(float) * NP, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(X, p->X.data(), sizeof(float) * (size_t)n * IN, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(T, p->T.data(), sizeof(float) * (size_t)n * OUT, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(M, 0, sizeof(float) * NP) != hipSuccess || hipMemset(V, 0, sizeof(float) * NP) != hipSuccess)
		return VGPU_HIP_ERROR;
	int st;
	for (int e = 1; e <= epochs; e++) {
		if ((st = mlp_loss_grad_dev(P, X, T, n, A, D, L, part, chunks, G, nullptr, s))) return st;
		if ((st = mlp_adam_dev(P, G, M, V, lr, e, s))) return st;
	}
	if ((st = mlp_loss_grad_dev(P, X, T, n, A, D, L, part, chunks, G, loss, s))) return st;
	if ((st = mlp_forward_dev(P, X, n, Y, s))) return st;
	std::vector<float> y((size_t)n * OUT);
	float lo = 0;
	if (hipMemcpy(p->params.data(), P, sizeof(float) * NP, hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(&lo, loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(y.data(), Y, sizeof(float) * y.size(), hipMemcpyDeviceToHost) != hipSuccess)
		return VGPU_HIP_ERROR;
	// margin: quantile of (label - prediction) in depth units over the
	// samples a list can finish (worklist exit below 64); the others reach
	// the block search whatever the depth
	std::vector<float> r;
	r.reserve(n);
	for (int t = 0; t < n; t++)
		if (p->T[(size_t)t * OUT] < 1.f) r.push_back(64.f * (p->T[(size_t)t * OUT] - y[(size_t)t * OUT]));
	if (r.empty()) p->margin = 0.f;
	else {
		std::sort(r.begin(), r.end());
		const int m = (int)r.size();
		const int q = (int)std::floor((1.0 - fallback) * (m - 1));
		p->margin = std::max(0.f, r[std::min(std::max(q, 0), m - 1)]);
	}
	p->trained = true;
	if (loss_out) *loss_out = lo;
	return VGPU_OK;
}

// Loss and gradient of the network for n samples given in host memory (the
// same kernels as training; used to check them against a CPU computation).
int vgpu_mlp_loss_grad(const float *params, const float *X, const float *T, int n, float *loss, float *grad) {
	if (!params || !X || !T || n <= 0) return VGPU_INVALID_ARGUMENT;
	DevMem mem;
	const int chunks = LG_CHUNKS;
	float *P = mem.get<float>(NP), *G = mem.get<float>(NP), *dX = mem.get<float>((size_t)n * IN),
	      *dT = mem.get<float>((size_t)n * OUT), *A = mem.get<float>((size_t)n * ACT), *D = mem.get<float>((size_t)n * ACT),
	      *L = mem.get<float>(n), *part = mem.get<float>((size_t)chunks * NP), *lo = mem.get<float>(1);
	if (!P || !G || !dX || !dT || !A || !D || !L || !part || !lo) return VGPU_MEMORY_ERROR;
	if (hipMemcpy(P, params, sizeof(float) * NP, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dX, X, sizeof(float) * (size_t)n * IN, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dT, T, sizeof(float) * (size_t)n * OUT, hipMemcpyHostToDevice) != hipSuccess)
		return VGPU_HIP_ERROR;
	int st = mlp_loss_grad_dev(P, dX, dT, n, A, D, L, part, chunks, G, lo, nullptr);
	if (st) return st;
	if (loss && hipMemcpy(loss, lo, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return VGPU_HIP_ERROR;
	if (grad && hipMemcpy(grad, G, sizeof(float) * NP, hipMemcpyDeviceToHost) != hipSuccess) return VGPU_HIP_ERROR;
	return VGPU_OK;
}

// Network outputs for n host samples (n x 2).
int vgpu_mlp_forward(const float *params, const float *X, int n, float *Y) {
	if (!params || !X || !Y || n <= 0) return VGPU_INVALID_ARGUMENT;
	DevMem mem;
	float *P = mem.get<float>(NP), *dX = mem.get<float>((size_t)n * IN), *dY = mem.get<float>((size_t)n * OUT);
	if (!P || !dX || !dY) return VGPU_MEMORY_ERROR;
	if (hipMemcpy(P, params, sizeof(float) * NP, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dX, X, sizeof(float) * (size_t)n * IN, hipMemcpyHostToDevice) != hipSuccess)
		return VGPU_HIP_ERROR;
	int st = mlp_forward_dev(P, dX, n, dY, nullptr);
	if (st) return st;
	if (hipMemcpy(Y, dY, sizeof(float) * (size_t)n * OUT, hipMemcpyDeviceToHost) != hipSuccess) return VGPU_HIP_ERROR;
	return VGPU_OK;
}

int vgpu_mlp_num_params(void) { return NP; }

}  // extern "C"
