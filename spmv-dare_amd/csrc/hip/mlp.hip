// FP32 multilayer perceptron for the format selector: forward, backward
// (mean squared error over the masked outputs) and Adam, all as hand-written
// kernels. Layers: n_in -> h1 -> h2 -> n_out, ReLU on the hidden layers,
// linear output. Parameters are one flat float array, per layer W (n_in x
// n_out, row-major) then b (n_out). No counterpart in the reference.
//
// Sizes are tiny (14 -> 64 -> 64 -> 40, batches of a few hundred), so every
// kernel maps one thread to one output element and sums in a fixed order:
// results are deterministic run to run.

#include <cmath>
#include <cstring>
#include <vector>

#include "format.hpp"

namespace spmvhip {

__global__ void mlp_dense_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ b,
		float *__restrict__ Z, float *__restrict__ Y, int B, int nin, int nout, int relu)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= B * nout)
		return;
	int r = t / nout, j = t - r * nout;
	float z = b[j];
	for (int i = 0; i < nin; i++)
		z = fmaf(X[r * nin + i], W[i * nout + j], z);
	Z[t] = z;
	Y[t] = relu ? fmaxf(z, 0.0f) : z;
}

// dY = 2 (Y - T) M / sum(M); loss = sum((Y - T)^2 M) / sum(M). One block.
__global__ void mlp_mse(const float *__restrict__ Y, const float *__restrict__ T, const float *__restrict__ M,
		float *__restrict__ dY, int n, float *__restrict__ loss)
{
	__shared__ float s_l[256], s_m[256];
	float l = 0, cnt = 0;
	for (int t = threadIdx.x; t < n; t += blockDim.x) {
		float mk = M ? M[t] : 1.0f;
		float d = (Y[t] - T[t]) * mk;
		l += d * d;
		cnt += mk;
	}
	s_l[threadIdx.x] = l;
	s_m[threadIdx.x] = cnt;
	__syncthreads();
	for (int o = blockDim.x / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o) {
			s_l[threadIdx.x] += s_l[threadIdx.x + o];
			s_m[threadIdx.x] += s_m[threadIdx.x + o];
		}
		__syncthreads();
	}
	const float norm = s_m[0] > 0 ? s_m[0] : 1.0f;
	for (int t = threadIdx.x; t < n; t += blockDim.x) {
		float mk = M ? M[t] : 1.0f;
		dY[t] = 2.0f * (Y[t] - T[t]) * mk / norm;
	}
	if (threadIdx.x == 0 && loss)
		*loss = s_l[0] / norm;
}

// dW[i][j] = sum_r X[r][i] dZ[r][j]; db[j] = sum_r dZ[r][j]
__global__ void mlp_dense_bwd_w(const float *__restrict__ X, const float *__restrict__ dZ, float *__restrict__ dW,
		float *__restrict__ db, int B, int nin, int nout)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < nin * nout) {
		int i = t / nout, j = t - i * nout;
		float g = 0;
		for (int r = 0; r < B; r++)
			g = fmaf(X[r * nin + i], dZ[r * nout + j], g);
		dW[t] = g;
	} else if (t < nin * nout + nout) {
		int j = t - nin * nout;
		float g = 0;
		for (int r = 0; r < B; r++)
			g += dZ[r * nout + j];
		db[j] = g;
	}
}

// dZprev[r][i] = (Zprev[r][i] > 0) * sum_j dZ[r][j] W[i][j]
__global__ void mlp_dense_bwd_x(const float *__restrict__ dZ, const float *__restrict__ W,
		const float *__restrict__ Zprev, float *__restrict__ dZprev, int B, int nin, int nout)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= B * nin)
		return;
	int r = t / nin, i = t - r * nin;
	float g = 0;
	for (int j = 0; j < nout; j++)
		g = fmaf(dZ[r * nout + j], W[i * nout + j], g);
	dZprev[t] = Zprev[t] > 0.0f ? g : 0.0f;
}

__global__ void mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
		float *__restrict__ m2, int n, float lr, float beta1, float beta2, float eps, float c1, float c2)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n)
		return;
	float gt = g[t];
	float a = beta1 * m1[t] + (1.0f - beta1) * gt;
	float b = beta2 * m2[t] + (1.0f - beta2) * gt * gt;
	m1[t] = a;
	m2[t] = b;
	p[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

} // namespace spmvhip

struct spmvhip_mlp {
	int dims[4];
	long nparams = 0;
	long off_w[3], off_b[3];
	int cap_batch = 0;
	long step = 0;
	spmvhip::DevBuf params, grads, m1, m2;
	spmvhip::DevBuf z[3], a[3], dz[3], dy, loss;
	hipStream_t stream = nullptr;
};

using namespace spmvhip;

namespace {

int ensure_batch(spmvhip_mlp *n, int B)
{
	if (B <= n->cap_batch)
		return SPMVHIP_OK;
	for (int l = 0; l < 3; l++) {
		size_t bytes = (size_t) B * n->dims[l + 1] * sizeof(float);
		SPMVHIP_TRY(n->z[l].alloc(bytes));
		SPMVHIP_TRY(n->a[l].alloc(bytes));
		SPMVHIP_TRY(n->dz[l].alloc(bytes));
	}
	SPMVHIP_TRY(n->dy.alloc((size_t) B * n->dims[3] * sizeof(float)));
	n->cap_batch = B;
	return SPMVHIP_OK;
}

int forward(spmvhip_mlp *n, const float *X, int B)
{
	SPMVHIP_TRY(ensure_batch(n, B));
	const float *in = X;
	const float *P = n->params.as<float>();
	for (int l = 0; l < 3; l++) {
		int nin = n->dims[l], nout = n->dims[l + 1];
		mlp_dense_fwd<<<grid_for((long) B * nout, kBlock), kBlock, 0, n->stream>>>(in, P + n->off_w[l],
				P + n->off_b[l], n->z[l].as<float>(), n->a[l].as<float>(), B, nin, nout, l < 2 ? 1 : 0);
		SPMVHIP_CHECK(hipGetLastError());
		in = n->a[l].as<float>();
	}
	return SPMVHIP_OK;
}

// gradients of the masked MSE into n->grads (the loss value into n->loss)
int backward(spmvhip_mlp *n, const float *X, const float *T, const float *M, int B)
{
	const float *P = n->params.as<float>();
	float *G = n->grads.as<float>();
	mlp_mse<<<1, 256, 0, n->stream>>>(n->a[2].as<float>(), T, M, n->dz[2].as<float>(), B * n->dims[3],
			n->loss.as<float>());
	SPMVHIP_CHECK(hipGetLastError());
	for (int l = 2; l >= 0; l--) {
		int nin = n->dims[l], nout = n->dims[l + 1];
		const float *in = l == 0 ? X : n->a[l - 1].as<float>();
		mlp_dense_bwd_w<<<grid_for((long) nin * nout + nout, kBlock), kBlock, 0, n->stream>>>(in,
				n->dz[l].as<float>(), G + n->off_w[l], G + n->off_b[l], B, nin, nout);
		SPMVHIP_CHECK(hipGetLastError());
		if (l > 0) {
			mlp_dense_bwd_x<<<grid_for((long) B * nin, kBlock), kBlock, 0, n->stream>>>(n->dz[l].as<float>(),
					P + n->off_w[l], n->z[l - 1].as<float>(), n->dz[l - 1].as<float>(), B, nin, nout);
			SPMVHIP_CHECK(hipGetLastError());
		}
	}
	return SPMVHIP_OK;
}

} // namespace

extern "C" {

int spmvhip_mlp_create(int n_in, int h1, int h2, int n_out, unsigned int seed, spmvhip_mlp **out)
{
	if (!out || n_in < 1 || h1 < 1 || h2 < 1 || n_out < 1 || n_in > 4096 || h1 > 4096 || h2 > 4096 || n_out > 4096)
		return set_error(SPMVHIP_EINVAL, "mlp: bad shape");
	*out = nullptr;
	spmvhip_mlp *n = new (std::nothrow) spmvhip_mlp();
	if (!n)
		return set_error(SPMVHIP_ENOMEM, "mlp handle");
	n->dims[0] = n_in;
	n->dims[1] = h1;
	n->dims[2] = h2;
	n->dims[3] = n_out;
	long off = 0;
	for (int l = 0; l < 3; l++) {
		n->off_w[l] = off;
		off += (long) n->dims[l] * n->dims[l + 1];
		n->off_b[l] = off;
		off += n->dims[l + 1];
	}
	n->nparams = off;
	int rc = n->params.alloc(off * sizeof(float));
	if (!rc) rc = n->grads.alloc(off * sizeof(float));
	if (!rc) rc = n->m1.alloc(off * sizeof(float));
	if (!rc) rc = n->m2.alloc(off * sizeof(float));
	if (!rc) rc = n->loss.alloc(sizeof(float));
	hipError_t e = rc ? hipSuccess : hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
	if (!rc && e != hipSuccess)
		rc = set_error(SPMVHIP_EHIP, "mlp stream: %s", hipGetErrorString(e));
	if (rc) {
		delete n;
		return rc;
	}
	// He-uniform initialisation from a fixed LCG (host), zero biases and moments
	std::vector<float> h(off, 0.0f);
	uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
	for (int l = 0; l < 3; l++) {
		float lim = std::sqrt(6.0f / (float) n->dims[l]);
		for (long k = 0; k < (long) n->dims[l] * n->dims[l + 1]; k++) {
			s = s * 6364136223846793005ull + 1442695040888963407ull;
			float u = (float) ((s >> 40) & 0xFFFFFF) / 16777216.0f;
			h[n->off_w[l] + k] = (2.0f * u - 1.0f) * lim;
		}
	}
	std::vector<float> zero(off, 0.0f);
	hipError_t e1 = hipMemcpy(n->params.p, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
	if (e1 == hipSuccess)
		e1 = hipMemcpy(n->m1.p, zero.data(), off * sizeof(float), hipMemcpyHostToDevice);
	if (e1 == hipSuccess)
		e1 = hipMemcpy(n->m2.p, zero.data(), off * sizeof(float), hipMemcpyHostToDevice);
	if (e1 != hipSuccess) {
		spmvhip_mlp_destroy(n);
		return set_error(SPMVHIP_EHIP, "mlp init: %s", hipGetErrorString(e1));
	}
	*out = n;
	return SPMVHIP_OK;
}

long spmvhip_mlp_num_params(const spmvhip_mlp *n)
{
	return n ? n->nparams : 0;
}

int spmvhip_mlp_set_params(spmvhip_mlp *n, const float *host, long count)
{
	if (!n || !host || count != n->nparams)
		return set_error(SPMVHIP_EINVAL, "mlp_set_params: expected %ld floats", n ? n->nparams : 0L);
	SPMVHIP_CHECK(hipMemcpy(n->params.p, host, count * sizeof(float), hipMemcpyHostToDevice));
	n->step = 0;
	std::vector<float> zero(count, 0.0f);
	SPMVHIP_CHECK(hipMemcpy(n->m1.p, zero.data(), count * sizeof(float), hipMemcpyHostToDevice));
	SPMVHIP_CHECK(hipMemcpy(n->m2.p, zero.data(), count * sizeof(float), hipMemcpyHostToDevice));
	return SPMVHIP_OK;
}

int spmvhip_mlp_get_params(const spmvhip_mlp *n, float *host, long count)
{
	if (!n || !host || count != n->nparams)
		return set_error(SPMVHIP_EINVAL, "mlp_get_params: expected %ld floats", n ? n->nparams : 0L);
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	SPMVHIP_CHECK(hipMemcpy(host, n->params.p, count * sizeof(float), hipMemcpyDeviceToHost));
	return SPMVHIP_OK;
}

int spmvhip_mlp_get_adam_state(const spmvhip_mlp *n, float *m1, float *m2, long count, long *step)
{
	if (!n || !m1 || !m2 || !step || count != n->nparams)
		return set_error(SPMVHIP_EINVAL, "mlp_get_adam_state: expected %ld floats per moment", n ? n->nparams : 0L);
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	SPMVHIP_CHECK(hipMemcpy(m1, n->m1.p, count * sizeof(float), hipMemcpyDeviceToHost));
	SPMVHIP_CHECK(hipMemcpy(m2, n->m2.p, count * sizeof(float), hipMemcpyDeviceToHost));
	*step = n->step;
	return SPMVHIP_OK;
}

int spmvhip_mlp_set_adam_state(spmvhip_mlp *n, const float *m1, const float *m2, long count, long step)
{
	if (!n || !m1 || !m2 || count != n->nparams || step < 0)
		return set_error(SPMVHIP_EINVAL, "mlp_set_adam_state: expected %ld floats per moment and step >= 0",
				n ? n->nparams : 0L);
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	SPMVHIP_CHECK(hipMemcpy(n->m1.p, m1, count * sizeof(float), hipMemcpyHostToDevice));
	SPMVHIP_CHECK(hipMemcpy(n->m2.p, m2, count * sizeof(float), hipMemcpyHostToDevice));
	n->step = step;
	return SPMVHIP_OK;
}

// Host in, host out: Y[B][n_out] = mlp(X[B][n_in]).
int spmvhip_mlp_predict(spmvhip_mlp *n, const float *X, int B, float *Y)
{
	if (!n || !X || !Y || B < 1)
		return set_error(SPMVHIP_EINVAL, "mlp_predict: bad arguments");
	DevBuf dx;
	SPMVHIP_TRY(dx.alloc((size_t) B * n->dims[0] * sizeof(float)));
	SPMVHIP_CHECK(hipMemcpyAsync(dx.p, X, (size_t) B * n->dims[0] * sizeof(float), hipMemcpyHostToDevice, n->stream));
	SPMVHIP_TRY(forward(n, dx.as<float>(), B));
	SPMVHIP_CHECK(hipMemcpyAsync(Y, n->a[2].p, (size_t) B * n->dims[3] * sizeof(float), hipMemcpyDeviceToHost,
			n->stream));
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	return SPMVHIP_OK;
}

// Gradients of the masked MSE (M may be NULL) w.r.t. all parameters, to host.
int spmvhip_mlp_gradients(spmvhip_mlp *n, const float *X, const float *T, const float *M, int B, float *grads,
		float *loss)
{
	if (!n || !X || !T || !grads || B < 1)
		return set_error(SPMVHIP_EINVAL, "mlp_gradients: bad arguments");
	size_t nx = (size_t) B * n->dims[0], ny = (size_t) B * n->dims[3];
	DevBuf dx, dt, dm;
	SPMVHIP_TRY(dx.alloc(nx * sizeof(float)));
	SPMVHIP_TRY(dt.alloc(ny * sizeof(float)));
	if (M)
		SPMVHIP_TRY(dm.alloc(ny * sizeof(float)));
	SPMVHIP_CHECK(hipMemcpyAsync(dx.p, X, nx * sizeof(float), hipMemcpyHostToDevice, n->stream));
	SPMVHIP_CHECK(hipMemcpyAsync(dt.p, T, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
	if (M)
		SPMVHIP_CHECK(hipMemcpyAsync(dm.p, M, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
	SPMVHIP_TRY(forward(n, dx.as<float>(), B));
	SPMVHIP_TRY(backward(n, dx.as<float>(), dt.as<float>(), M ? dm.as<float>() : nullptr, B));
	SPMVHIP_CHECK(hipMemcpyAsync(grads, n->grads.p, n->nparams * sizeof(float), hipMemcpyDeviceToHost, n->stream));
	if (loss)
		SPMVHIP_CHECK(hipMemcpyAsync(loss, n->loss.p, sizeof(float), hipMemcpyDeviceToHost, n->stream));
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	return SPMVHIP_OK;
}

// `epochs` passes of Adam over (X, T, M) in mini-batches of B rows, in the
// order given (the caller shuffles). losses[e] = mean batch loss of epoch e.
int spmvhip_mlp_train(spmvhip_mlp *n, const float *X, const float *T, const float *M, long rows, int B, int epochs,
		float lr, float *losses)
{
	if (!n || !X || !T || rows < 1 || B < 1 || epochs < 0)
		return set_error(SPMVHIP_EINVAL, "mlp_train: bad arguments");
	size_t nx = (size_t) rows * n->dims[0], ny = (size_t) rows * n->dims[3];
	DevBuf dx, dt, dm;
	SPMVHIP_TRY(dx.alloc(nx * sizeof(float)));
	SPMVHIP_TRY(dt.alloc(ny * sizeof(float)));
	if (M)
		SPMVHIP_TRY(dm.alloc(ny * sizeof(float)));
	SPMVHIP_CHECK(hipMemcpyAsync(dx.p, X, nx * sizeof(float), hipMemcpyHostToDevice, n->stream));
	SPMVHIP_CHECK(hipMemcpyAsync(dt.p, T, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
	if (M)
		SPMVHIP_CHECK(hipMemcpyAsync(dm.p, M, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	for (int e = 0; e < epochs; e++) {
		double acc = 0;
		long nb = 0;
		for (long r0 = 0; r0 < rows; r0 += B) {
			int bs = (int) (rows - r0 < B ? rows - r0 : B);
			const float *xb = dx.as<float>() + r0 * n->dims[0];
			const float *tb = dt.as<float>() + r0 * n->dims[3];
			const float *mb = M ? dm.as<float>() + r0 * n->dims[3] : nullptr;
			SPMVHIP_TRY(forward(n, xb, bs));
			SPMVHIP_TRY(backward(n, xb, tb, mb, bs));
			n->step++;
			float c1 = 1.0f - std::pow(b1, (float) n->step), c2 = 1.0f - std::pow(b2, (float) n->step);
			mlp_adam<<<grid_for(n->nparams, kBlock), kBlock, 0, n->stream>>>(n->params.as<float>(),
					n->grads.as<float>(), n->m1.as<float>(), n->m2.as<float>(), (int) n->nparams, lr, b1, b2,
					eps, c1, c2);
			SPMVHIP_CHECK(hipGetLastError());
			float l = 0;
			SPMVHIP_CHECK(hipMemcpyAsync(&l, n->loss.p, sizeof(float), hipMemcpyDeviceToHost, n->stream));
			SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
			acc += l;
			nb++;
		}
		if (losses)
			losses[e] = (float) (acc / (nb ? nb : 1));
	}
	return SPMVHIP_OK;
}

// One Adam step with externally supplied gradients (for checking the update).
int spmvhip_mlp_adam_step(spmvhip_mlp *n, const float *grads, float lr)
{
	if (!n || !grads)
		return set_error(SPMVHIP_EINVAL, "mlp_adam_step: bad arguments");
	SPMVHIP_CHECK(hipMemcpy(n->grads.p, grads, n->nparams * sizeof(float), hipMemcpyHostToDevice));
	n->step++;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	float c1 = 1.0f - std::pow(b1, (float) n->step), c2 = 1.0f - std::pow(b2, (float) n->step);
	mlp_adam<<<grid_for(n->nparams, kBlock), kBlock, 0, n->stream>>>(n->params.as<float>(), n->grads.as<float>(),
			n->m1.as<float>(), n->m2.as<float>(), (int) n->nparams, lr, b1, b2, eps, c1, c2);
	SPMVHIP_CHECK(hipGetLastError());
	SPMVHIP_CHECK(hipStreamSynchronize(n->stream));
	return SPMVHIP_OK;
}

void spmvhip_mlp_destroy(spmvhip_mlp *n)
{
	if (!n)
		return;
	if (n->stream)
		(void) hipStreamSynchronize(n->stream);
	hipStream_t s = n->stream;
	delete n;
	if (s)
		(void) hipStreamDestroy(s);
}

} // extern "C"
