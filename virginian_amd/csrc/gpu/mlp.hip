/*
 * Result-size predictor: a 3-layer perceptron (nin -> H -> H -> 1, ReLU,
 * ReLU, sigmoid; f32) that maps features of (op-program, tablet) to the
 * fraction of the tablet's rows the program selects.  Trained on the GPU
 * by the kernels below (forward, backward, Adam) on per-tablet counts the
 * executor itself reports; the executor uses the prediction only to size
 * single-pass outputs, and an exact two-pass run replaces any run whose
 * result did not fit, so results never depend on it.
 *
 * No counterpart in the reference.  Parameters, flat f32:
 *   W1[nin][H] b1[H] W2[H][H] b2[H] W3[H] b3
 * plus, optionally, a fixed skip: skip * x[0] added to the output's
 * pre-activation (x[0] carries the logit of a classical estimate, which
 * the network then corrects).  Loss: mean over examples of (y - t)^2.
 *
 * The layers are small (H = 32) and the batches a few hundred examples,
 * so each kernel gives one thread one output value; nothing here is worth
 * matrix cores.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <vector>

extern "C" {
#include "virginian.h"
}

/* device scratch for one batch of B examples */
struct mlp_work {
	float *x = nullptr, *t = nullptr, *h1 = nullptr, *h2 = nullptr, *y = nullptr;
	float *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *loss = nullptr;
	unsigned cap = 0;
};

struct vg_mlp {
	unsigned nin = 0, nh = 0, np = 0;
	float skip = 0.0f;           /* output pre-activation += skip * x[0] */
	int loss = 0;                /* 0: mean squared error, 1: mean absolute error, 2: pinball */
	float tau = 0.5f;            /* the pinball loss's quantile */
	float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
	unsigned long long step = 0;
	/* inference: a stream that does not wait for other streams' work,
	 * device scratch and pinned staging kept between calls (grown, never
	 * shrunk), so a forward pass neither allocates nor waits for the
	 * query kernels running beside it */
	hipStream_t s = nullptr;
	mlp_work inf;
	float *hx = nullptr, *hy = nullptr;
	/* the parameters in host memory, refreshed whenever they change on the
	 * device, for vg_mlp_forward_host */
	std::vector<float> hp;
};

/* copy the parameters to the host copy (after training or setting them) */
static int refresh_host_params(vg_mlp *m)
{
	m->hp.resize(m->np);
	return hipMemcpy(m->hp.data(), m->d_p, m->np * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
}

/* offsets of the parameter blocks */
struct mlp_layout {
	size_t w1, b1, w2, b2, w3, b3, n;
};

static mlp_layout layout(unsigned nin, unsigned h)
{
	mlp_layout L;
	L.w1 = 0;
	L.b1 = L.w1 + (size_t)nin * h;
	L.w2 = L.b1 + h;
	L.b2 = L.w2 + (size_t)h * h;
	L.w3 = L.b2 + h;
	L.b3 = L.w3 + h;
	L.n = L.b3 + 1;
	return L;
}

/* out[b][j] = act(bias[j] + skip * xs[b * xs_stride] + sum_i in[b][i] *
 * W[i][j]); act 0 = ReLU, 1 = sigmoid */
__global__ void mlp_layer_fwd(const float *in, unsigned nin, const float *W, const float *bias, float *out,
	unsigned nout, unsigned B, int act, const float *xs, unsigned xs_stride, float skip)
{
	const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= B * nout)
		return;
	const unsigned b = t / nout, j = t % nout;
	float z = bias[j];
	if (skip != 0.0f)
		z += skip * xs[(size_t)b * xs_stride];
	const float *x = in + (size_t)b * nin;
	for (unsigned i = 0; i < nin; i++)
		z += x[i] * W[(size_t)i * nout + j];
	out[t] = act == 0 ? (z > 0.0f ? z : 0.0f) : 1.0f / (1.0f + expf(-z));
}

/* output delta: d(loss)/dz3 = dl / B * y (1 - y), dl = d(loss of one
 * example)/dy: 2 (y - t) for the squared error, sign(y - t) for the
 * absolute error, and for the pinball loss of quantile tau,
 * max(tau (t - y), (1 - tau) (y - t)): 1 - tau above the target, -tau
 * below it (0 on it) */
__global__ void mlp_out_delta(const float *y, const float *target, unsigned B, float *dz, int kind, float tau)
{
	const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= B)
		return;
	const float yy = y[b], r = yy - target[b];
	float dl;
	if (kind == 1)
		dl = r > 0.0f ? 1.0f : r < 0.0f ? -1.0f : 0.0f;
	else if (kind == 2)
		dl = r > 0.0f ? 1.0f - tau : r < 0.0f ? -tau : 0.0f;
	else
		dl = 2.0f * r;
	dz[b] = dl / (float)B * yy * (1.0f - yy);
}

/* hidden delta: dz[b][i] = [h[b][i] > 0] * sum_j dnext[b][j] * Wn[i][j] */
__global__ void mlp_hidden_delta(const float *dnext, const float *Wn, unsigned nnext, const float *h,
	unsigned nh, unsigned B, float *dz)
{
	const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= B * nh)
		return;
	const unsigned b = t / nh, i = t % nh;
	float s = 0.0f;
	for (unsigned j = 0; j < nnext; j++)
		s += dnext[(size_t)b * nnext + j] * Wn[(size_t)i * nnext + j];
	dz[t] = h[t] > 0.0f ? s : 0.0f;
}

/* dW[i][j] = sum_b in[b][i] dz[b][j] for i < nin; row nin: dbias[j] */
__global__ void mlp_wgrad(const float *in, unsigned nin, const float *dz, unsigned nout, unsigned B,
	float *dW, float *dbias)
{
	const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (nin + 1) * nout)
		return;
	const unsigned i = t / nout, j = t % nout;
	float s = 0.0f;
	if (i < nin) {
		for (unsigned b = 0; b < B; b++)
			s += in[(size_t)b * nin + i] * dz[(size_t)b * nout + j];
		dW[t] = s;
	} else {
		for (unsigned b = 0; b < B; b++)
			s += dz[(size_t)b * nout + j];
		dbias[j] = s;
	}
}

/* mean loss into *loss (one workgroup of 256) */
__global__ void mlp_loss(const float *y, const float *target, unsigned B, float *loss, int kind, float tau)
{
	__shared__ float part[256];
	float s = 0.0f;
	for (unsigned b = threadIdx.x; b < B; b += 256) {
		const float d = y[b] - target[b];
		s += kind == 1 ? fabsf(d) : kind == 2 ? (d > 0.0f ? (1.0f - tau) * d : -tau * d) : d * d;
	}
	part[threadIdx.x] = s;
	__syncthreads();
	for (unsigned w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w)
			part[threadIdx.x] += part[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*loss = part[0] / (float)B;
}

/* Adam (Kingma & Ba 2015), bias-corrected */
__global__ void mlp_adam(float *p, const float *g, float *m, float *v, unsigned n, float lr, float b1,
	float b2, float eps, float c1, float c2)
{
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static unsigned blocks(size_t n)
{
	return (unsigned)((n + 255) / 256);
}


static void free_work(mlp_work &w)
{
	float *bufs[] = { w.x, w.t, w.h1, w.h2, w.y, w.d1, w.d2, w.d3, w.loss };
	for (float *p : bufs)
		if (p)
			(void)hipFree(p);
	w = mlp_work();
}

static int alloc_work(mlp_work &w, unsigned B, unsigned nin, unsigned h)
{
	free_work(w);
	const size_t sz[] = { (size_t)B * nin, B, (size_t)B * h, (size_t)B * h, B, (size_t)B * h, (size_t)B * h,
		B, 1 };
	float **bufs[] = { &w.x, &w.t, &w.h1, &w.h2, &w.y, &w.d1, &w.d2, &w.d3, &w.loss };
	for (int k = 0; k < 9; k++)
		if (hipMalloc((void **)bufs[k], (sz[k] ? sz[k] : 1) * sizeof(float)) != hipSuccess) {
			(void)hipGetLastError();
			free_work(w);
			return VIRG_FAIL;
		}
	w.cap = B;
	return VIRG_SUCCESS;
}

/* forward over w.x (B examples) into w.h1, w.h2, w.y */
static void forward(const vg_mlp *m, mlp_work &w, unsigned B, const float *x, hipStream_t st = 0)
{
	const mlp_layout L = layout(m->nin, m->nh);
	const float *P = m->d_p;
	hipLaunchKernelGGL(mlp_layer_fwd, dim3(blocks((size_t)B * m->nh)), dim3(256), 0, st,
		x, m->nin, P + L.w1, P + L.b1, w.h1, m->nh, B, 0, x, m->nin, 0.0f);
	hipLaunchKernelGGL(mlp_layer_fwd, dim3(blocks((size_t)B * m->nh)), dim3(256), 0, st,
		w.h1, m->nh, P + L.w2, P + L.b2, w.h2, m->nh, B, 0, x, m->nin, 0.0f);
	hipLaunchKernelGGL(mlp_layer_fwd, dim3(blocks(B)), dim3(256), 0, st,
		w.h2, m->nh, P + L.w3, P + L.b3, w.y, 1u, B, 1, x, m->nin, m->skip);
}

/* backward after forward: gradient of the mean loss into m->d_g */
static void backward(vg_mlp *m, mlp_work &w, unsigned B, const float *x, const float *t)
{
	const mlp_layout L = layout(m->nin, m->nh);
	const float *P = m->d_p;
	float *G = m->d_g;
	const unsigned h = m->nh;
	hipLaunchKernelGGL(mlp_out_delta, dim3(blocks(B)), dim3(256), 0, 0, w.y, t, B, w.d3, m->loss, m->tau);
	hipLaunchKernelGGL(mlp_wgrad, dim3(blocks((size_t)(h + 1) * 1)), dim3(256), 0, 0,
		w.h2, h, w.d3, 1u, B, G + L.w3, G + L.b3);
	hipLaunchKernelGGL(mlp_hidden_delta, dim3(blocks((size_t)B * h)), dim3(256), 0, 0,
		w.d3, P + L.w3, 1u, w.h2, h, B, w.d2);
	hipLaunchKernelGGL(mlp_wgrad, dim3(blocks((size_t)(h + 1) * h)), dim3(256), 0, 0,
		w.h1, h, w.d2, h, B, G + L.w2, G + L.b2);
	hipLaunchKernelGGL(mlp_hidden_delta, dim3(blocks((size_t)B * h)), dim3(256), 0, 0,
		w.d2, P + L.w2, h, w.h1, h, B, w.d1);
	hipLaunchKernelGGL(mlp_wgrad, dim3(blocks((size_t)(m->nin + 1) * h)), dim3(256), 0, 0,
		x, m->nin, w.d1, h, B, G + L.w1, G + L.b1);
}

/* splitmix64 */
static uint64_t mix(uint64_t &s)
{
	uint64_t z = (s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static float unit(uint64_t &s)
{
	return (float)((mix(s) >> 40) * (1.0 / 16777216.0));
}

static void free_staging(vg_mlp *m)
{
	free_work(m->inf);
	if (m->hx)
		(void)hipHostFree(m->hx);
	if (m->hy)
		(void)hipHostFree(m->hy);
	m->hx = m->hy = nullptr;
}

extern "C" void vg_mlp_free(vg_mlp *m)
{
	if (!m)
		return;
	if (m->s)
		(void)hipStreamSynchronize(m->s);
	free_staging(m);
	if (m->s)
		(void)hipStreamDestroy(m->s);
	float *bufs[] = { m->d_p, m->d_g, m->d_m, m->d_v };
	for (float *p : bufs)
		if (p)
			(void)hipFree(p);
	delete m;
}

/* The loss training and vg_mlp_gradient use: 0 mean squared error (the
 * default), 1 mean absolute error. */
extern "C" int vg_mlp_set_loss(vg_mlp *m, int kind)
{
	if (kind != 0 && kind != 1)
		return VIRG_FAIL;
	m->loss = kind;
	return VIRG_SUCCESS;
}

/* The pinball loss of quantile tau (0 < tau < 1): the network then
 * predicts the tau-quantile of the target given the features. */
extern "C" int vg_mlp_set_quantile(vg_mlp *m, float tau)
{
	if (!(tau > 0.0f && tau < 1.0f))
		return VIRG_FAIL;
	m->loss = 2;
	m->tau = tau;
	return VIRG_SUCCESS;
}

/* A network for `nin` features and `nh` hidden units per layer, weights
 * drawn uniformly in +-sqrt(6 / fan_in) from `seed`, biases 0.  With
 * skip != 0 the output's pre-activation also gets skip * x[0] and W3
 * starts at 0: the network starts as sigmoid(skip * x[0]) and learns a
 * correction to it. */
extern "C" vg_mlp *vg_mlp_create(unsigned nin, unsigned nh, float skip, unsigned long long seed)
{
	if (nin == 0 || nh == 0)
		return nullptr;
	vg_mlp *m = new vg_mlp();
	m->nin = nin;
	m->nh = nh;
	m->skip = skip;
	const mlp_layout L = layout(nin, nh);
	m->np = (unsigned)L.n;
	if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) {
		(void)hipGetLastError();
		m->s = nullptr;
		vg_mlp_free(m);
		return nullptr;
	}
	float **bufs[] = { &m->d_p, &m->d_g, &m->d_m, &m->d_v };
	for (float **b : bufs)
		if (hipMalloc((void **)b, L.n * sizeof(float)) != hipSuccess ||
			hipMemset(*b, 0, L.n * sizeof(float)) != hipSuccess) {
			(void)hipGetLastError();
			vg_mlp_free(m);
			return nullptr;
		}
	std::vector<float> p(L.n, 0.0f);
	uint64_t s = seed;
	auto fill = [&](size_t off, size_t n, unsigned fan_in) {
		const float a = sqrtf(6.0f / (float)fan_in);
		for (size_t i = 0; i < n; i++)
			p[off + i] = (2.0f * unit(s) - 1.0f) * a;
	};
	fill(L.w1, (size_t)nin * nh, nin);
	fill(L.w2, (size_t)nh * nh, nh);
	if (skip == 0.0f)
		fill(L.w3, nh, nh);
	if (hipMemcpy(m->d_p, p.data(), L.n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
		(void)hipGetLastError();
		vg_mlp_free(m);
		return nullptr;
	}
	m->hp = p;
	return m;
}

extern "C" unsigned vg_mlp_num_params(const vg_mlp *m)
{
	return m->np;
}

extern "C" int vg_mlp_get_params(const vg_mlp *m, float *out)
{
	return hipMemcpy(out, m->d_p, m->np * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
		? VIRG_SUCCESS : VIRG_FAIL;
}

extern "C" int vg_mlp_set_params(vg_mlp *m, const float *in)
{
	m->step = 0;
	if (hipMemset(m->d_m, 0, m->np * sizeof(float)) != hipSuccess ||
		hipMemset(m->d_v, 0, m->np * sizeof(float)) != hipSuccess)
		return VIRG_FAIL;
	return hipMemcpy(m->d_p, in, m->np * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
		refresh_host_params(m) ? VIRG_SUCCESS : VIRG_FAIL;
}

/* y[i] = network(x[i]) for n examples (host arrays, x row-major n x nin) */
extern "C" int vg_mlp_forward(vg_mlp *m, const float *x, unsigned n, float *y)
{
	if (n == 0)
		return VIRG_SUCCESS;
	if (n > m->inf.cap) {
		/* grow (rarely: the first call, or a larger batch than before) */
		unsigned cap = 4096;
		while (cap < n)
			cap *= 2;
		(void)hipStreamSynchronize(m->s);
		free_staging(m);
		if (alloc_work(m->inf, cap, m->nin, m->nh) == VIRG_FAIL ||
			hipHostMalloc((void **)&m->hx, (size_t)cap * m->nin * sizeof(float), hipHostMallocDefault) != hipSuccess ||
			hipHostMalloc((void **)&m->hy, (size_t)cap * sizeof(float), hipHostMallocDefault) != hipSuccess) {
			(void)hipGetLastError();
			free_staging(m);
			return VIRG_FAIL;
		}
	}
	mlp_work &w = m->inf;
	memcpy(m->hx, x, (size_t)n * m->nin * sizeof(float));
	int ok = hipMemcpyAsync(w.x, m->hx, (size_t)n * m->nin * sizeof(float), hipMemcpyHostToDevice, m->s) ==
		hipSuccess;
	if (ok) {
		forward(m, w, n, w.x, m->s);
		ok = hipGetLastError() == hipSuccess &&
			hipMemcpyAsync(m->hy, w.y, n * sizeof(float), hipMemcpyDeviceToHost, m->s) == hipSuccess &&
			hipStreamSynchronize(m->s) == hipSuccess;
	}
	if (!ok) {
		(void)hipGetLastError();
		return VIRG_FAIL;
	}
	memcpy(y, m->hy, n * sizeof(float));
	return VIRG_SUCCESS;
}

/* The same network evaluated on the host from the parameters' host copy,
 * in the GPU layers' order of operations (bias, skip, then the inputs in
 * order; f32): for callers that must not wait for the device, such as a
 * serving loop whose batches hold every compute unit. */
extern "C" int vg_mlp_forward_host(const vg_mlp *m, const float *x, unsigned n, float *y)
{
	const mlp_layout L = layout(m->nin, m->nh);
	if (m->hp.size() != L.n)
		return VIRG_FAIL;
	const float *P = m->hp.data();
	const unsigned nin = m->nin, h = m->nh;
	std::vector<float> h1(h), h2(h);
	for (unsigned b = 0; b < n; b++) {
		const float *xb = x + (size_t)b * nin;
		for (unsigned j = 0; j < h; j++) {
			float z = P[L.b1 + j];
			for (unsigned i = 0; i < nin; i++)
				z += xb[i] * P[L.w1 + (size_t)i * h + j];
			h1[j] = z > 0.0f ? z : 0.0f;
		}
		for (unsigned j = 0; j < h; j++) {
			float z = P[L.b2 + j];
			for (unsigned i = 0; i < h; i++)
				z += h1[i] * P[L.w2 + (size_t)i * h + j];
			h2[j] = z > 0.0f ? z : 0.0f;
		}
		float z = P[L.b3];
		if (m->skip != 0.0f)
			z += m->skip * xb[0];
		for (unsigned i = 0; i < h; i++)
			z += h2[i] * P[L.w3 + i];
		y[b] = 1.0f / (1.0f + expf(-z));
	}
	return VIRG_SUCCESS;
}

/* Loss and gradient (parameter order as above) over n examples. */
extern "C" int vg_mlp_gradient(vg_mlp *m, const float *x, const float *t, unsigned n, float *loss, float *grad)
{
	if (n == 0)
		return VIRG_FAIL;
	mlp_work w;
	if (alloc_work(w, n, m->nin, m->nh) == VIRG_FAIL)
		return VIRG_FAIL;
	int ok = hipMemcpy(w.x, x, (size_t)n * m->nin * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
		hipMemcpy(w.t, t, n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
	if (ok) {
		forward(m, w, n, w.x);
		backward(m, w, n, w.x, w.t);
		hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(256), 0, 0, w.y, w.t, n, w.loss, m->loss, m->tau);
		ok = hipGetLastError() == hipSuccess &&
			hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess &&
			hipMemcpy(grad, m->d_g, m->np * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
	}
	free_work(w);
	return ok ? VIRG_SUCCESS : VIRG_FAIL;
}

/* Adam over minibatches of `batch` examples, `epochs` passes over the n
 * examples in an order shuffled once from `seed`; *loss = the mean loss
 * over all n examples after training. */
extern "C" int vg_mlp_train(vg_mlp *m, const float *x, const float *t, unsigned n, unsigned epochs,
	unsigned batch, float lr, unsigned long long seed, float *loss)
{
	if (n == 0 || batch == 0)
		return VIRG_FAIL;
	/* shuffle once on the host */
	std::vector<unsigned> order(n);
	for (unsigned i = 0; i < n; i++)
		order[i] = i;
	uint64_t s = seed;
	for (unsigned i = n - 1; i > 0; i--) {
		const unsigned j = (unsigned)(mix(s) % (i + 1));
		const unsigned tmp = order[i];
		order[i] = order[j];
		order[j] = tmp;
	}
	std::vector<float> xs((size_t)n * m->nin), ts(n);
	for (unsigned i = 0; i < n; i++) {
		memcpy(&xs[(size_t)i * m->nin], x + (size_t)order[i] * m->nin, m->nin * sizeof(float));
		ts[i] = t[order[i]];
	}
	mlp_work w;
	if (alloc_work(w, n, m->nin, m->nh) == VIRG_FAIL)
		return VIRG_FAIL;
	int ok = hipMemcpy(w.x, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
		hipMemcpy(w.t, ts.data(), n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	for (unsigned e = 0; e < epochs && ok; e++)
		for (unsigned o = 0; o < n && ok; o += batch) {
			const unsigned B = n - o < batch ? n - o : batch;
			const float *xb = w.x + (size_t)o * m->nin, *tb = w.t + o;
			forward(m, w, B, xb);
			backward(m, w, B, xb, tb);
			m->step++;
			const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
			hipLaunchKernelGGL(mlp_adam, dim3(blocks(m->np)), dim3(256), 0, 0, m->d_p, m->d_g, m->d_m, m->d_v,
				m->np, lr, b1, b2, eps, c1, c2);
			ok = hipGetLastError() == hipSuccess;
		}
	if (ok) {
		forward(m, w, n, w.x);
		hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(256), 0, 0, w.y, w.t, n, w.loss, m->loss, m->tau);
		ok = hipGetLastError() == hipSuccess &&
			hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
	}
	free_work(w);
	ok = ok && refresh_host_params(m);
	return ok ? VIRG_SUCCESS : VIRG_FAIL;
}
