// mlp.hip -- the cost model's FP32 multilayer perceptron on the GPU:
// forward, backward and Adam kernels, and the C ABI that drives them
// (declared in include/phase2_hip.h).  No counterpart in the reference.
//
// Network: nin -> nhid -> nhid -> 1, ReLU after both hidden layers.
// Parameters are packed as W1[nhid][nin], b1[nhid], W2[nhid][nhid], b2[nhid],
// W3[nhid], b3.  Loss: 0.5 * mean_b (y_b - t_b)^2.
//
// Every kernel computes one output element per thread and sums in a fixed
// order, so results are deterministic run to run.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "phase2_hip.h"

namespace {

constexpr int kThreads = 256;

// out[b][o] = act(sum_i W[o][i] in[b][i] + bias[o]); z keeps the
// pre-activation when non-null.
__global__ void k_dense_fwd(const float *__restrict__ in,
	const float *__restrict__ W, const float *__restrict__ bias,
	float *__restrict__ z, float *__restrict__ out, uint32_t B, uint32_t I,
	uint32_t O, int relu)
{
	const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
	if (k >= (uint64_t)B * O)
		return;
	const uint32_t b = (uint32_t)(k / O), o = (uint32_t)(k % O);
	float acc = bias[o];
	for (uint32_t i = 0; i < I; i++)
		acc += W[(size_t)o * I + i] * in[(size_t)b * I + i];
	if (z)
		z[k] = acc;
	out[k] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dy[b] = (y[b] - t[b]) / B; loss terms in lt[b].
__global__ void k_loss_grad(const float *__restrict__ y,
	const float *__restrict__ t, float *__restrict__ dy,
	float *__restrict__ lt, uint32_t B)
{
	const uint32_t b = blockIdx.x * kThreads + threadIdx.x;
	if (b >= B)
		return;
	const float d = y[b] - t[b];
	dy[b] = d / (float)B;
	lt[b] = 0.5f * d * d;
}

// din[b][i] = (sum_o W[o][i] dout[b][o]) * (z[b][i] > 0)  (z null: no mask)
__global__ void k_dense_bwd_in(const float *__restrict__ dout,
	const float *__restrict__ W, const float *__restrict__ z,
	float *__restrict__ din, uint32_t B, uint32_t I, uint32_t O)
{
	const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
	if (k >= (uint64_t)B * I)
		return;
	const uint32_t b = (uint32_t)(k / I), i = (uint32_t)(k % I);
	float acc = 0.0f;
	for (uint32_t o = 0; o < O; o++)
		acc += W[(size_t)o * I + i] * dout[(size_t)b * O + o];
	if (z && !(z[k] > 0.0f))
		acc = 0.0f;
	din[k] = acc;
}

// dW[o][i] = sum_b dout[b][o] in[b][i]; db[o] = sum_b dout[b][o]
// (thread index o*I + i for the weights, then O threads for the biases).
__global__ void k_dense_bwd_w(const float *__restrict__ dout,
	const float *__restrict__ in, float *__restrict__ dW,
	float *__restrict__ db, uint32_t B, uint32_t I, uint32_t O)
{
	const uint64_t k = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
	if (k < (uint64_t)O * I) {
		const uint32_t o = (uint32_t)(k / I), i = (uint32_t)(k % I);
		float acc = 0.0f;
		for (uint32_t b = 0; b < B; b++)
			acc += dout[(size_t)b * O + o] * in[(size_t)b * I + i];
		dW[k] = acc;
	} else if (k < (uint64_t)O * I + O) {
		const uint32_t o = (uint32_t)(k - (uint64_t)O * I);
		float acc = 0.0f;
		for (uint32_t b = 0; b < B; b++)
			acc += dout[(size_t)b * O + o];
		db[o] = acc;
	}
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g,
	float *__restrict__ m, float *__restrict__ v, uint32_t n, float lr,
	float b1, float b2, float eps, float c1, float c2)
{
	const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
	if (k >= n)
		return;
	const float gk = g[k];
	const float mk = b1 * m[k] + (1.0f - b1) * gk;
	const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
	m[k] = mk;
	v[k] = vk;
	p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

uint32_t blocks(uint64_t n)
{
	return (uint32_t)((n + kThreads - 1) / kThreads);
}

// splitmix64 for parameter initialisation
uint64_t sm64(uint64_t *x)
{
	uint64_t v = (*x += UINT64_C(0x9e3779b97f4a7c15));
	v = (v ^ (v >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
	v = (v ^ (v >> 27)) * UINT64_C(0x94d049bb133111eb);
	return v ^ (v >> 31);
}

char g_mlp_err[200] = "";

} // namespace

struct ph2hip_mlp {
	uint32_t nin, nhid, maxb;
	size_t np;
	float *p, *g, *m, *v;         // device: params, grads, Adam moments
	float *x, *t, *z1, *a1, *z2, *a2, *y, *dy, *lt, *d2, *d1;
	uint64_t step;
	hipStream_t s;
};

#define MCHK(call)                                                          \
	do {                                                                \
		hipError_t e_ = (call);                                     \
		if (e_ != hipSuccess) {                                     \
			snprintf(g_mlp_err, sizeof g_mlp_err, "%s: %s", #call,  \
				hipGetErrorString(e_));                     \
			return -1;                                          \
		}                                                           \
	} while (0)

namespace {

// parameter block offsets
struct Off {
	size_t W1, b1, W2, b2, W3, b3, n;
};

Off offsets(uint32_t I, uint32_t H)
{
	Off o;
	o.W1 = 0;
	o.b1 = o.W1 + (size_t)H * I;
	o.W2 = o.b1 + H;
	o.b2 = o.W2 + (size_t)H * H;
	o.W3 = o.b2 + H;
	o.b3 = o.W3 + H;
	o.n = o.b3 + 1;
	return o;
}

int forward(ph2hip_mlp *M, uint32_t B)
{
	const Off o = offsets(M->nin, M->nhid);
	const uint32_t I = M->nin, H = M->nhid;
	hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)B * H)),
		dim3(kThreads), 0, M->s, M->x, M->p + o.W1, M->p + o.b1, M->z1,
		M->a1, B, I, H, 1);
	hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)B * H)),
		dim3(kThreads), 0, M->s, M->a1, M->p + o.W2, M->p + o.b2, M->z2,
		M->a2, B, H, H, 1);
	hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(B)), dim3(kThreads), 0, M->s,
		M->a2, M->p + o.W3, M->p + o.b3, (float *)nullptr, M->y, B, H, 1u,
		0);
	MCHK(hipGetLastError());
	return 0;
}

// gradients of the mean loss into M->g; host loss sum in *loss
int backward(ph2hip_mlp *M, uint32_t B, float *loss)
{
	const Off o = offsets(M->nin, M->nhid);
	const uint32_t I = M->nin, H = M->nhid;
	hipLaunchKernelGGL(k_loss_grad, dim3(blocks(B)), dim3(kThreads), 0, M->s,
		M->y, M->t, M->dy, M->lt, B);
	// layer 3: dW3, db3; d2 = W3^T dy masked by z2
	hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)H + 1)),
		dim3(kThreads), 0, M->s, M->dy, M->a2, M->g + o.W3, M->g + o.b3, B,
		H, 1u);
	hipLaunchKernelGGL(k_dense_bwd_in, dim3(blocks((uint64_t)B * H)),
		dim3(kThreads), 0, M->s, M->dy, M->p + o.W3, M->z2, M->d2, B, H,
		1u);
	// layer 2
	hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)H * H + H)),
		dim3(kThreads), 0, M->s, M->d2, M->a1, M->g + o.W2, M->g + o.b2, B,
		H, H);
	hipLaunchKernelGGL(k_dense_bwd_in, dim3(blocks((uint64_t)B * H)),
		dim3(kThreads), 0, M->s, M->d2, M->p + o.W2, M->z1, M->d1, B, H, H);
	// layer 1
	hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)H * I + H)),
		dim3(kThreads), 0, M->s, M->d1, M->x, M->g + o.W1, M->g + o.b1, B,
		I, H);
	MCHK(hipGetLastError());
	if (loss) {
		float *lt = (float *)malloc(sizeof(float) * B);
		if (!lt)
			return -1;
		if (hipMemcpyAsync(lt, M->lt, sizeof(float) * B,
			    hipMemcpyDeviceToHost, M->s) != hipSuccess ||
		    hipStreamSynchronize(M->s) != hipSuccess) {
			free(lt);
			return -1;
		}
		double acc = 0.0;
		for (uint32_t b = 0; b < B; b++)
			acc += lt[b];
		*loss = (float)(acc / B);
		free(lt);
	}
	return 0;
}

int load_batch(ph2hip_mlp *M, const float *x, const float *t, size_t B)
{
	if (B == 0 || B > M->maxb) {
		snprintf(g_mlp_err, sizeof g_mlp_err, "batch %zu outside 1..%u",
			B, M->maxb);
		return -1;
	}
	MCHK(hipMemcpyAsync(M->x, x, sizeof(float) * B * M->nin,
		hipMemcpyHostToDevice, M->s));
	if (t)
		MCHK(hipMemcpyAsync(M->t, t, sizeof(float) * B,
			hipMemcpyHostToDevice, M->s));
	return 0;
}

} // namespace

extern "C" {

const char *ph2hip_mlp_strerror(void)
{
	return g_mlp_err;
}

size_t ph2hip_mlp_nparams(const struct ph2hip_mlp *M)
{
	return M->np;
}

void ph2hip_mlp_destroy(struct ph2hip_mlp *M)
{
	if (!M)
		return;
	(void)hipStreamSynchronize(M->s);
	float *bufs[] = { M->p, M->g, M->m, M->v, M->x, M->t, M->z1, M->a1,
		M->z2, M->a2, M->y, M->dy, M->lt, M->d2, M->d1 };
	for (float *b : bufs)
		(void)hipFree(b);
	(void)hipStreamDestroy(M->s);
	free(M);
}

int ph2hip_mlp_create(struct ph2hip_mlp **out, uint32_t nin, uint32_t nhid,
	uint32_t max_batch, uint64_t seed)
{
	*out = nullptr;
	if (ph2hip_init(-1) < 0)
		return -1;
	ph2hip_mlp *M = (ph2hip_mlp *)calloc(1, sizeof *M);
	if (!M)
		return -1;
	M->nin = nin;
	M->nhid = nhid;
	M->maxb = max_batch;
	const Off o = offsets(nin, nhid);
	M->np = o.n;
	const size_t B = max_batch, H = nhid;
	int rc = 0;
	struct {
		float **p;
		size_t n;
	} alloc[] = { { &M->p, o.n }, { &M->g, o.n }, { &M->m, o.n },
		{ &M->v, o.n }, { &M->x, B * nin }, { &M->t, B },
		{ &M->z1, B * H }, { &M->a1, B * H }, { &M->z2, B * H },
		{ &M->a2, B * H }, { &M->y, B }, { &M->dy, B }, { &M->lt, B },
		{ &M->d2, B * H }, { &M->d1, B * H } };
	if (hipStreamCreateWithFlags(&M->s, hipStreamNonBlocking) != hipSuccess)
		rc = -1;
	for (auto &a : alloc)
		if (!rc && hipMalloc((void **)a.p, sizeof(float) * a.n) !=
			hipSuccess)
			rc = -1;
	// He-uniform weights, zero biases, zero Adam moments
	float *h = (float *)calloc(o.n, sizeof(float));
	if (!h)
		rc = -1;
	if (!rc) {
		uint64_t st = seed ? seed : 1;
		auto fill = [&](size_t off, size_t n, uint32_t fan_in) {
			const double lim = sqrt(6.0 / fan_in);
			for (size_t k = 0; k < n; k++) {
				const double u = (double)(sm64(&st) >> 11) * 0x1.0p-53;
				h[off + k] = (float)((2.0 * u - 1.0) * lim);
			}
		};
		fill(o.W1, (size_t)nhid * nin, nin);
		fill(o.W2, (size_t)nhid * nhid, nhid);
		fill(o.W3, nhid, nhid);
		if (hipMemcpy(M->p, h, sizeof(float) * o.n, hipMemcpyHostToDevice)
			!= hipSuccess ||
		    hipMemset(M->m, 0, sizeof(float) * o.n) != hipSuccess ||
		    hipMemset(M->v, 0, sizeof(float) * o.n) != hipSuccess)
			rc = -1;
	}
	free(h);
	if (rc) {
		snprintf(g_mlp_err, sizeof g_mlp_err, "mlp_create failed");
		ph2hip_mlp_destroy(M);
		return -1;
	}
	*out = M;
	return 0;
}

int ph2hip_mlp_set_params(struct ph2hip_mlp *M, const float *params)
{
	MCHK(hipMemcpyAsync(M->p, params, sizeof(float) * M->np,
		hipMemcpyHostToDevice, M->s));
	MCHK(hipMemsetAsync(M->m, 0, sizeof(float) * M->np, M->s));
	MCHK(hipMemsetAsync(M->v, 0, sizeof(float) * M->np, M->s));
	MCHK(hipStreamSynchronize(M->s));
	M->step = 0;
	return 0;
}

int ph2hip_mlp_get_params(struct ph2hip_mlp *M, float *params)
{
	MCHK(hipMemcpyAsync(params, M->p, sizeof(float) * M->np,
		hipMemcpyDeviceToHost, M->s));
	MCHK(hipStreamSynchronize(M->s));
	return 0;
}

int ph2hip_mlp_forward(struct ph2hip_mlp *M, const float *x, size_t B,
	float *y)
{
	for (size_t b0 = 0; b0 < B; b0 += M->maxb) {
		const size_t nb = B - b0 < M->maxb ? B - b0 : M->maxb;
		if (load_batch(M, x + b0 * M->nin, nullptr, nb) ||
		    forward(M, (uint32_t)nb))
			return -1;
		MCHK(hipMemcpyAsync(y + b0, M->y, sizeof(float) * nb,
			hipMemcpyDeviceToHost, M->s));
	}
	MCHK(hipStreamSynchronize(M->s));
	return 0;
}

int ph2hip_mlp_grad(struct ph2hip_mlp *M, const float *x, const float *t,
	size_t B, float *loss, float *grad)
{
	if (load_batch(M, x, t, B) || forward(M, (uint32_t)B) ||
	    backward(M, (uint32_t)B, loss))
		return -1;
	MCHK(hipMemcpyAsync(grad, M->g, sizeof(float) * M->np,
		hipMemcpyDeviceToHost, M->s));
	MCHK(hipStreamSynchronize(M->s));
	return 0;
}

int ph2hip_mlp_train_step(struct ph2hip_mlp *M, const float *x,
	const float *t, size_t B, float lr, float *loss)
{
	if (load_batch(M, x, t, B) || forward(M, (uint32_t)B) ||
	    backward(M, (uint32_t)B, loss))
		return -1;
	M->step++;
	const float b1 = 0.9f, b2 = 0.999f;
	const float c1 = 1.0f - powf(b1, (float)M->step);
	const float c2 = 1.0f - powf(b2, (float)M->step);
	hipLaunchKernelGGL(k_adam, dim3(blocks(M->np)), dim3(kThreads), 0, M->s,
		M->p, M->g, M->m, M->v, (uint32_t)M->np, lr, b1, b2, 1e-8f, c1,
		c2);
	MCHK(hipGetLastError());
	MCHK(hipStreamSynchronize(M->s));
	return 0;
}

} // extern "C"
