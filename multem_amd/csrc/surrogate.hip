// Frozen-phonon surrogate network: 3x3 convolutions on the gfx950 matrix
// cores (bf16 operands, FP32 accumulation), backward pass, Adam.
//
// Every convolution is an implicit GEMM over pixels:
//   forward        out[p][co]  = sum_kk  A[p][kk] W[co][kk],   kk = tap*cin + ci
//   input grad     the same kernel with the flipped, transposed weights
//   weight grad    dW[co][kk]  = sum_p   dy[p][co] A[p][kk]    (split over pixel chunks)
// with A[p][kk] = x[neighbour(p, tap)][ci] under periodic padding.  Each
// wave computes a 16-pixel x (16*NT)-channel block with
// v_mfma_f32_16x16x16_bf16; operands are read straight from global memory
// (the network's whole working set sits in L2), so there is no LDS staging.
// All reductions (weight-gradient chunks, bias gradients, the loss) are
// summed in a fixed order: results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/multem_surrogate.h"

namespace mtg {
void set_last_error(const std::string &msg);
}

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even bfloat16 (finite inputs)
__device__ __forceinline__ short to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// flat pixel index of the neighbour (dy, dx) of pixel p, periodic in y and x
__device__ __forceinline__ int neighbour(int p, int h, int w, int dy, int dx)
{
    const int hw = h * w;
    const int b = p / hw, r = p - b * hw, y = r / w, x = r - (r / w) * w;
    return b * hw + wrap(y + dy, h) * w + wrap(x + dx, w);
}

struct ConvArgs {
    const float *in;      // [npix][cin]
    const short *wt;      // bf16 [cout][9*cin]
    const float *bias;    // [cout] or null
    const float *res;     // added after the (optional) relu: res[p*res_stride + co], or null
    int res_stride;
    const float *mask;    // output zeroed where mask[p*cout + co] <= 0, or null
    float *out;           // [npix][cout]
    int npix, h, w, cin, cout;
    int relu;
};

// NT: 16-channel output tiles per wave (cout <= 16*NT)
template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wave) * 16;
    if (p0 >= a.npix) return;                 // whole waves only: npix is a multiple of 16
    const int m = lane & 15, kq = lane >> 4;
    const int K = 9 * a.cin, ksteps = (K + 15) / 16;
    const int p = p0 + m;
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int s = 0; s < ksteps; ++s) {
        const int kk0 = 16 * s + 4 * kq;
        bf16x4 av = {0, 0, 0, 0};
        if ((a.cin & 3) == 0) {
            if (kk0 < K) {   // the four kk share one tap
                const int tap = kk0 / a.cin, ci = kk0 - tap * a.cin;
                const int q = neighbour(p, a.h, a.w, tap / 3 - 1, tap % 3 - 1);
                const float4 v = *reinterpret_cast<const float4 *>(a.in + (long long)q * a.cin + ci);
                av = bf16x4{to_bf16(v.x), to_bf16(v.y), to_bf16(v.z), to_bf16(v.w)};
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kk = kk0 + i;
                if (kk < K) {
                    const int tap = kk / a.cin, ci = kk - tap * a.cin;
                    const int q = neighbour(p, a.h, a.w, tap / 3 - 1, tap % 3 - 1);
                    av[i] = to_bf16(a.in[(long long)q * a.cin + ci]);
                }
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + m;
            bf16x4 bv = {0, 0, 0, 0};
            if (co < a.cout) {
                const short *wr = a.wt + (long long)co * K;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (kk0 + i < K) bv[i] = wr[kk0 + i];
            }
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nt], 0, 0, 0);
        }
    }
    // lane holds D[4*kq + i][m]: pixel p0 + 4*kq + i, channel nt*16 + m
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = nt * 16 + m;
        if (co >= a.cout) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long pp = p0 + 4 * kq + i;
            float v = acc[nt][i];
            if (a.bias) v += a.bias[co];
            if (a.relu) v = fmaxf(v, 0.f);
            if (a.res) v += a.res[pp * a.res_stride + co];
            if (a.mask && !(a.mask[pp * a.cout + co] > 0.f)) v = 0.f;
            a.out[pp * a.cout + co] = v;
        }
    }
}

struct WgradArgs {
    const float *x;       // layer input [npix][cin]
    const float *dy;      // output gradient [npix][cout]
    float *part;          // [nchunk][cout][9*cin]
    float *part_b;        // [nchunk][cout]
    int npix, h, w, cin, cout, chunk;
};

// One workgroup per chunk of `chunk` pixels; the four waves share the
// (cout/16) x (9 cin/16) output tiles.  Bias partials are summed by the same
// workgroup in a fixed order.
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a)
{
    __shared__ float bsum[256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = blockIdx.x;
    const int pstart = c * a.chunk, pend = min(pstart + a.chunk, a.npix);
    const int K = 9 * a.cin;
    const int MT = (a.cout + 15) / 16, NK = (K + 15) / 16;
    const int kq = lane >> 4, n = lane & 15;
    for (int tile = wave; tile < MT * NK; tile += 4) {
        const int mt = tile / NK, nt = tile - (tile / NK) * NK;
        const int co = mt * 16 + n;          // A row held by this lane
        const int kk = nt * 16 + n;          // B column held by this lane
        const bool kk_ok = kk < K;
        const int tap = kk_ok ? kk / a.cin : 0, ci = kk_ok ? kk - tap * a.cin : 0;
        const int ddy = tap / 3 - 1, ddx = tap % 3 - 1;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int q = pstart; q < pend; q += 16) {
            bf16x4 av = {0, 0, 0, 0}, bv = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int pp = q + 4 * kq + i;
                if (co < a.cout) av[i] = to_bf16(a.dy[(long long)pp * a.cout + co]);
                if (kk_ok) bv[i] = to_bf16(a.x[(long long)neighbour(pp, a.h, a.w, ddy, ddx) * a.cin + ci]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
        // lane holds D[4*kq + i][n]: channel mt*16 + 4*kq + i, kk
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co2 = mt * 16 + 4 * kq + i;
            if (co2 < a.cout && kk_ok) a.part[((long long)c * a.cout + co2) * K + kk] = acc[i];
        }
    }
    // bias partials: thread t sums pixels t/cout, t/cout + G, ... of channel t % cout
    const int G = 256 / a.cout;
    const int t = threadIdx.x, co = t % a.cout, g = t / a.cout;
    float s = 0.f;
    if (g < G)
        for (int pp = pstart + g; pp < pend; pp += G) s += a.dy[(long long)pp * a.cout + co];
    bsum[t] = s;
    __syncthreads();
    if (t < a.cout) {
        float tot = 0.f;
        for (int k = 0; k < G; ++k) tot += bsum[k * a.cout + t];
        a.part_b[(long long)c * a.cout + t] = tot;
    }
}

// grad_w[i] = sum over chunks of part[c][i] (fixed order); same for biases
__global__ void k_chunk_sum(const float *part, int nchunk, int n, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(long long)c * n + i];
    out[i] = s;
}

// bf16 weights of one layer: forward [cout][tap][cin] and flipped for the
// input gradient [cin][tap'][cout] with tap' = 8 - tap
__global__ void k_prep_weights(const float *W, int cout, int cin, short *wt, short *wflip)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int K = 9 * cin;
    if (i >= cout * K) return;
    const int co = i / K, kk = i - co * K, tap = kk / cin, ci = kk - tap * cin;
    const short v = to_bf16(W[i]);
    wt[i] = v;
    wflip[((long long)ci * 9 + (8 - tap)) * cout + co] = v;
}

// squared-error loss partials and dy = 2 (y - t) / n
__global__ __launch_bounds__(256) void k_loss(const float *y, const float *t, int n, float scale, float *dy,
                                              double *part)
{
    __shared__ double red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const float r = y[i] - t[i];
        dy[i] = scale * r;
        s = (double)r * (double)r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_loss_final(const double *part, int nb, double inv_n, double *loss)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    *loss = s * inv_n;
}

__global__ void k_mask_mul(const float *in, const float *mask, float *out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = mask[i] > 0.f ? in[i] : 0.f;
}

// Adam (Kingma & Ba 2015, Algorithm 1); c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
__global__ void k_adam(float *theta, float *m, float *v, const float *g, int n, float lr, float b1, float b2, float c1,
                       float c2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// minibatch gather with augmentation: aug = (image, shift y, shift x, symmetry)
// symmetry bit 0: transpose (square images only), bit 1: flip y, bit 2: flip x
__global__ void k_gather(const float *X, const float *T, const int4 *aug, int h, int w, int batch, float *xb, float *tb)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int hw = h * w;
    if (i >= (long long)batch * hw) return;
    const int j = (int)(i / hw), r = (int)(i % hw), y = r / w, x = r % w;
    const int4 g = aug[j];
    int ys = y, xs = x;
    if (g.w & 1) { const int tmp = ys; ys = xs; xs = tmp; }
    if (g.w & 2) ys = h - 1 - ys;
    if (g.w & 4) xs = w - 1 - xs;
    ys = (ys + g.y) % h;
    xs = (xs + g.z) % w;
    const long long src = (long long)g.x * hw + (long long)ys * w + xs;
    xb[2 * i] = X[2 * src];
    xb[2 * i + 1] = X[2 * src + 1];
    tb[i] = T[src];
}

void hcheck(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void ensure(size_t count)
    {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hcheck(hipMalloc((void **)&p, count * sizeof(T)), "hipMalloc");
        n = count;
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

} // namespace

struct mt_sg_net {
    struct Layer {
        int cin, cout;
        size_t off_w, off_b, off_wt;   // flat parameter offsets; bf16 weight offset
    };
    int n_blocks, ch, device;
    std::vector<Layer> layers;
    size_t n_par = 0, n_wt = 0;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_fwd_ms = 0;
    int adam_t = 0;

    DevBuf<float> theta, am, av, grad;
    DevBuf<short> wt, wflip;
    size_t cap = 0;   // pixel capacity of the activation buffers
    DevBuf<float> xb, tb, y, dy, dh, du, acts;   // acts: h_0..h_nb then u_1..u_nb, each npix*ch
    DevBuf<float> part, part_b;
    DevBuf<double> lpart, losses;
    DevBuf<float> dX, dT;
    DevBuf<int4> aug;

    static constexpr int kChunk = 256;

    mt_sg_net(int nb, int c, int dev) : n_blocks(nb), ch(c), device(dev)
    {
        auto add = [&](int ci, int co) {
            Layer L{ci, co, n_par, n_par + (size_t)co * 9 * ci, n_wt};
            n_par += (size_t)co * 9 * ci + co;
            n_wt += (size_t)co * 9 * ci;
            layers.push_back(L);
        };
        add(2, ch);
        for (int k = 0; k < nb; ++k) {
            add(ch, ch);
            add(ch, ch);
        }
        add(ch, 1);
        hcheck(hipSetDevice(device), "hipSetDevice");
        hcheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        hcheck(hipEventCreate(&e0), "hipEventCreate");
        hcheck(hipEventCreate(&e1), "hipEventCreate");
        theta.ensure(n_par);
        am.ensure(n_par);
        av.ensure(n_par);
        grad.ensure(n_par);
        wt.ensure(n_wt);
        wflip.ensure(n_wt);
        hcheck(hipMemsetAsync(theta.p, 0, n_par * sizeof(float), s), "hipMemsetAsync");
        reset_adam();
        prep_weights();
        hcheck(hipStreamSynchronize(s), "sync");
    }
    ~mt_sg_net()
    {
        if (s) (void)hipStreamSynchronize(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }

    void reset_adam()
    {
        hcheck(hipMemsetAsync(am.p, 0, n_par * sizeof(float), s), "hipMemsetAsync");
        hcheck(hipMemsetAsync(av.p, 0, n_par * sizeof(float), s), "hipMemsetAsync");
        adam_t = 0;
    }

    void prep_weights()
    {
        for (const Layer &L : layers) {
            const int n = L.cout * 9 * L.cin;
            hipLaunchKernelGGL(k_prep_weights, dim3((n + 255) / 256), dim3(256), 0, s, theta.p + L.off_w, L.cout, L.cin,
                               wt.p + L.off_wt, wflip.p + L.off_wt);
        }
        hcheck(hipGetLastError(), "prep weights");
    }

    void ensure(size_t npix)
    {
        if (npix <= cap) return;
        xb.ensure(npix * 2);
        tb.ensure(npix);
        y.ensure(npix);
        dy.ensure(npix);
        dh.ensure(npix * ch);
        du.ensure(npix * ch);
        acts.ensure(npix * ch * (size_t)(2 * n_blocks + 1));
        const size_t nchunk = (npix + kChunk - 1) / kChunk;
        part.ensure(nchunk * (size_t)ch * 9 * ch);
        part_b.ensure(nchunk * (size_t)ch);
        lpart.ensure((npix + 255) / 256);
        cap = npix;
    }

    float *h_act(int k, size_t npix) { return acts.p + (size_t)k * npix * ch; }                 // k = 0..nb
    float *u_act(int k, size_t npix) { return acts.p + (size_t)(n_blocks + 1 + k) * npix * ch; } // k = 0..nb-1

    void conv(const float *in, const Layer &L, bool flipped, const float *bias, const float *res, int res_stride,
              const float *mask, bool relu, float *out, int npix, int h, int w)
    {
        ConvArgs a{};
        a.in = in;
        a.wt = (flipped ? wflip.p : wt.p) + L.off_wt;
        a.bias = bias;
        a.res = res;
        a.res_stride = res_stride;
        a.mask = mask;
        a.out = out;
        a.npix = npix;
        a.h = h;
        a.w = w;
        a.cin = flipped ? L.cout : L.cin;
        a.cout = flipped ? L.cin : L.cout;
        a.relu = relu ? 1 : 0;
        const dim3 grid((npix + 63) / 64), block(256);
        if (a.cout <= 16) hipLaunchKernelGGL(k_conv3x3<1>, grid, block, 0, s, a);
        else if (a.cout <= 32) hipLaunchKernelGGL(k_conv3x3<2>, grid, block, 0, s, a);
        else hipLaunchKernelGGL(k_conv3x3<4>, grid, block, 0, s, a);
        hcheck(hipGetLastError(), "conv3x3");
    }

    void wgrad(const float *x, const float *dyv, const Layer &L, int npix, int h, int w)
    {
        WgradArgs a{};
        a.x = x;
        a.dy = dyv;
        a.part = part.p;
        a.part_b = part_b.p;
        a.npix = npix;
        a.h = h;
        a.w = w;
        a.cin = L.cin;
        a.cout = L.cout;
        a.chunk = kChunk;
        const int nchunk = (npix + kChunk - 1) / kChunk;
        hipLaunchKernelGGL(k_wgrad, dim3(nchunk), dim3(256), 0, s, a);
        const int nw = L.cout * 9 * L.cin;
        hipLaunchKernelGGL(k_chunk_sum, dim3((nw + 255) / 256), dim3(256), 0, s, part.p, nchunk, nw, grad.p + L.off_w);
        hipLaunchKernelGGL(k_chunk_sum, dim3(1), dim3(256), 0, s, part_b.p, nchunk, L.cout, grad.p + L.off_b);
        hcheck(hipGetLastError(), "weight gradient");
    }

    // forward of xb (npix = batch*h*w pixels) into y; activations kept for backward
    void forward(int npix, int h, int w)
    {
        const Layer &L0 = layers[0];
        conv(xb.p, L0, false, theta.p + L0.off_b, nullptr, 0, nullptr, true, h_act(0, npix), npix, h, w);
        for (int k = 0; k < n_blocks; ++k) {
            const Layer &La = layers[1 + 2 * k], &Lb = layers[2 + 2 * k];
            conv(h_act(k, npix), La, false, theta.p + La.off_b, nullptr, 0, nullptr, true, u_act(k, npix), npix, h, w);
            conv(u_act(k, npix), Lb, false, theta.p + Lb.off_b, h_act(k, npix), ch, nullptr, false, h_act(k + 1, npix),
                 npix, h, w);
        }
        const Layer &Lh = layers.back();
        conv(h_act(n_blocks, npix), Lh, false, theta.p + Lh.off_b, xb.p, 2, nullptr, false, y.p, npix, h, w);
    }

    // loss of y against tb into *loss_dev, dy; then all parameter gradients
    void backward(int npix, int h, int w, double *loss_dev)
    {
        const int nb = (npix + 255) / 256;
        hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, y.p, tb.p, npix, (float)(2.0 / npix), dy.p, lpart.p);
        hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, s, lpart.p, nb, 1.0 / npix, loss_dev);
        hcheck(hipGetLastError(), "loss");
        const Layer &Lh = layers.back();
        wgrad(h_act(n_blocks, npix), dy.p, Lh, npix, h, w);
        conv(dy.p, Lh, true, nullptr, nullptr, 0, nullptr, false, dh.p, npix, h, w);
        for (int k = n_blocks - 1; k >= 0; --k) {
            const Layer &La = layers[1 + 2 * k], &Lb = layers[2 + 2 * k];
            wgrad(u_act(k, npix), dh.p, Lb, npix, h, w);
            conv(dh.p, Lb, true, nullptr, nullptr, 0, u_act(k, npix), false, du.p, npix, h, w);
            wgrad(h_act(k, npix), du.p, La, npix, h, w);
            conv(du.p, La, true, nullptr, dh.p, ch, nullptr, false, dh.p, npix, h, w);
        }
        const long long n = (long long)npix * ch;
        hipLaunchKernelGGL(k_mask_mul, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dh.p, h_act(0, npix), du.p, n);
        hcheck(hipGetLastError(), "mask");
        wgrad(xb.p, du.p, layers[0], npix, h, w);
    }

    void adam(float lr)
    {
        ++adam_t;
        const double b1 = 0.9, b2 = 0.999;
        const float c1 = (float)(1.0 / (1.0 - std::pow(b1, adam_t))), c2 = (float)(1.0 / (1.0 - std::pow(b2, adam_t)));
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((n_par + 255) / 256)), dim3(256), 0, s, theta.p, am.p, av.p, grad.p,
                           (int)n_par, lr, (float)b1, (float)b2, c1, c2, 1e-8f);
        hcheck(hipGetLastError(), "adam");
        prep_weights();
    }
};

namespace {

mt_status sg_fail(mt_status c, const std::string &m)
{
    mtg::set_last_error(m);
    return c;
}

template <class F>
mt_status sg_guard(mt_sg_net *net, F &&f)
{
    if (!net) return sg_fail(MT_ERR_INVALID_ARGUMENT, "null network");
    try {
        hcheck(hipSetDevice(net->device), "hipSetDevice");
        mt_status st = f();
        if (st == MT_OK) mtg::set_last_error("");
        return st;
    } catch (const std::bad_alloc &) {
        return sg_fail(MT_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return sg_fail(MT_ERR_DEVICE, e.what());
    }
}

mt_status check_shape(int batch, int h, int w)
{
    if (batch < 1 || h < 3 || w < 3) return sg_fail(MT_ERR_INVALID_ARGUMENT, "batch >= 1 and h, w >= 3 required");
    if (((long long)h * w) % 16) return sg_fail(MT_ERR_INVALID_ARGUMENT, "h*w must be a multiple of 16");
    if ((long long)batch * h * w > (1ll << 26)) return sg_fail(MT_ERR_INVALID_ARGUMENT, "more than 2^26 pixels");
    return MT_OK;
}

} // namespace

extern "C" {

mt_status mt_hip_sg_create(int n_blocks, int channels, int device, mt_sg_net **net)
{
    if (!net) return sg_fail(MT_ERR_INVALID_ARGUMENT, "null output handle");
    *net = nullptr;
    if (n_blocks < 0 || n_blocks > 64) return sg_fail(MT_ERR_INVALID_ARGUMENT, "n_blocks out of range 0..64");
    if (channels < 4 || channels > 64 || channels % 4) return sg_fail(MT_ERR_INVALID_ARGUMENT, "channels: 4..64, multiple of 4");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return sg_fail(MT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return sg_fail(MT_ERR_INVALID_ARGUMENT, "device out of range");
    try {
        *net = new mt_sg_net(n_blocks, channels, device);
    } catch (const std::exception &e) {
        return sg_fail(MT_ERR_DEVICE, e.what());
    }
    mtg::set_last_error("");
    return MT_OK;
}

void mt_hip_sg_destroy(mt_sg_net *net) { delete net; }

size_t mt_hip_sg_param_count(const mt_sg_net *net) { return net ? net->n_par : 0; }

mt_status mt_hip_sg_set_params(mt_sg_net *net, const float *theta, size_t n)
{
    return sg_guard(net, [&] {
        if (!theta || n != net->n_par) return sg_fail(MT_ERR_INVALID_ARGUMENT, "parameter count mismatch");
        hcheck(hipMemcpyAsync(net->theta.p, theta, n * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        net->reset_adam();
        net->prep_weights();
        hcheck(hipStreamSynchronize(net->s), "sync");
        return MT_OK;
    });
}

mt_status mt_hip_sg_get_params(const mt_sg_net *net, float *theta, size_t n)
{
    return sg_guard(const_cast<mt_sg_net *>(net), [&] {
        if (!theta || n != net->n_par) return sg_fail(MT_ERR_INVALID_ARGUMENT, "parameter count mismatch");
        hcheck(hipMemcpyAsync(theta, net->theta.p, n * sizeof(float), hipMemcpyDeviceToHost, net->s), "D2H");
        hcheck(hipStreamSynchronize(net->s), "sync");
        return MT_OK;
    });
}

mt_status mt_hip_sg_forward(mt_sg_net *net, const float *x, int batch, int h, int w, float *y)
{
    return sg_guard(net, [&] {
        if (!x || !y) return sg_fail(MT_ERR_INVALID_ARGUMENT, "null image");
        mt_status st = check_shape(batch, h, w);
        if (st != MT_OK) return st;
        const int npix = batch * h * w;
        net->ensure(npix);
        hcheck(hipMemcpyAsync(net->xb.p, x, (size_t)npix * 2 * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        hcheck(hipEventRecord(net->e0, net->s), "event");
        net->forward(npix, h, w);
        hcheck(hipEventRecord(net->e1, net->s), "event");
        hcheck(hipMemcpyAsync(y, net->y.p, (size_t)npix * sizeof(float), hipMemcpyDeviceToHost, net->s), "D2H");
        hcheck(hipStreamSynchronize(net->s), "sync");
        float ms = 0;
        hcheck(hipEventElapsedTime(&ms, net->e0, net->e1), "event time");
        net->last_fwd_ms = ms;
        return MT_OK;
    });
}

mt_status mt_hip_sg_loss_grad(mt_sg_net *net, const float *x, const float *t, int batch, int h, int w, double *loss,
                              float *grad)
{
    return sg_guard(net, [&] {
        if (!x || !t) return sg_fail(MT_ERR_INVALID_ARGUMENT, "null image");
        mt_status st = check_shape(batch, h, w);
        if (st != MT_OK) return st;
        const int npix = batch * h * w;
        net->ensure(npix);
        net->losses.ensure(1);
        hcheck(hipMemcpyAsync(net->xb.p, x, (size_t)npix * 2 * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        hcheck(hipMemcpyAsync(net->tb.p, t, (size_t)npix * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        net->forward(npix, h, w);
        net->backward(npix, h, w, net->losses.p);
        double l = 0;
        hcheck(hipMemcpyAsync(&l, net->losses.p, sizeof(double), hipMemcpyDeviceToHost, net->s), "D2H");
        if (grad)
            hcheck(hipMemcpyAsync(grad, net->grad.p, net->n_par * sizeof(float), hipMemcpyDeviceToHost, net->s), "D2H");
        hcheck(hipStreamSynchronize(net->s), "sync");
        if (loss) *loss = l;
        return MT_OK;
    });
}

mt_status mt_hip_sg_train(mt_sg_net *net, const float *x, const float *t, int n_images, int h, int w, int batch,
                          int steps, float lr, unsigned long long seed, float *losses)
{
    return sg_guard(net, [&] {
        if (!x || !t) return sg_fail(MT_ERR_INVALID_ARGUMENT, "null data");
        if (n_images < 1 || steps < 1 || !(lr > 0)) return sg_fail(MT_ERR_INVALID_ARGUMENT, "n_images, steps >= 1, lr > 0");
        mt_status st = check_shape(batch, h, w);
        if (st != MT_OK) return st;
        const size_t hw = (size_t)h * w;
        const int npix = batch * h * w;
        net->ensure(npix);
        net->dX.ensure((size_t)n_images * hw * 2);
        net->dT.ensure((size_t)n_images * hw);
        net->losses.ensure(steps);
        hcheck(hipMemcpyAsync(net->dX.p, x, (size_t)n_images * hw * 2 * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        hcheck(hipMemcpyAsync(net->dT.p, t, (size_t)n_images * hw * sizeof(float), hipMemcpyHostToDevice, net->s), "H2D");
        // minibatch draws and augmentations for every step (splitmix64 stream)
        std::vector<int4> aug((size_t)steps * batch);
        uint64_t z = seed;
        auto next = [&z] {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t r = z;
            r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
            r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
            return r ^ (r >> 31);
        };
        for (auto &g : aug) {
            g.x = (int)(next() % (uint64_t)n_images);
            g.y = (int)(next() % (uint64_t)h);
            g.z = (int)(next() % (uint64_t)w);
            g.w = (int)(next() % 8u);
            if (h != w) g.w &= 6;   // transposes only for square images
        }
        net->aug.ensure(aug.size());
        hcheck(hipMemcpyAsync(net->aug.p, aug.data(), aug.size() * sizeof(int4), hipMemcpyHostToDevice, net->s), "H2D");
        for (int k = 0; k < steps; ++k) {
            hipLaunchKernelGGL(k_gather, dim3((npix + 255) / 256), dim3(256), 0, net->s, net->dX.p, net->dT.p,
                               net->aug.p + (size_t)k * batch, h, w, batch, net->xb.p, net->tb.p);
            hcheck(hipGetLastError(), "gather");
            net->forward(npix, h, w);
            net->backward(npix, h, w, net->losses.p + k);
            const float lr_k = (float)(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * k / steps)));
            net->adam(lr_k);
        }
        std::vector<double> l(steps);
        hcheck(hipMemcpyAsync(l.data(), net->losses.p, steps * sizeof(double), hipMemcpyDeviceToHost, net->s), "D2H");
        hcheck(hipStreamSynchronize(net->s), "sync");
        if (losses)
            for (int k = 0; k < steps; ++k) losses[k] = (float)l[k];
        return MT_OK;
    });
}

double mt_hip_sg_last_forward_ms(const mt_sg_net *net) { return net ? net->last_fwd_ms : 0.0; }

} // extern "C"
