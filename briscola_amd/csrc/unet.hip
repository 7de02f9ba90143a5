// The learned preconditioner: a 4-level 3-D convolutional U-Net (channels
// 16/32/64/128, 3x3x3 convolutions, BF16 weights and activations, FP32
// accumulation and FP32 master weights), trained on the device with the
// convolution kernels of unet_kernels.hip (forward, data and weight
// gradients on the matrix cores) and Adam.
//
// Network, on a D x H x W grid divisible by 8, channels-last BF16:
//   in (2 channels: scaled right-hand side, scaled log coefficient; padded
//   to 8) -> E0 conv 8->16 -> pool -> E1 16->32 -> pool -> E2 32->64 ->
//   pool -> E3 64->128 -> up ++ E2 -> D2 192->64 -> up ++ E1 -> D1 96->32 ->
//   up ++ E0 -> D0 48->16 -> 1x1x1 head 16->1 (FP32).
// Every 3x3x3 convolution has a bias and a ReLU; ++ is a channel
// concatenation: each level keeps one buffer [up | skip] that the upsampling
// and the encoder convolution write into, so nothing is copied.
#include "unet_internal.h"
#include "bmg_unet.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace bmg {
namespace unet {

namespace {

#define UCK(expr)                                                             \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            throw UnetError(std::string(#expr) + ": " + hipGetErrorString(e_));\
    } while (0)

template <typename T>
T* dalloc(size_t n)
{
    void* p = nullptr;
    UCK(hipMalloc(&p, std::max<size_t>(n, 1)*sizeof(T)));
    UCK(hipMemset(p, 0, std::max<size_t>(n, 1)*sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// SplitMix64 and Box-Muller for the weight initialisation
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27))*0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11)*(1.0/9007199254740992.0); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0*std::log(u1))*std::cos(6.283185307179586*u2);
    }
};

struct ConvSpec {
    int cin, cout, cinReal, level;
};

// E0, E1, E2, E3, D2, D1, D0
const ConvSpec kConvs[7] = {{8, 16, 2, 0},    {16, 32, 16, 1},
                            {32, 64, 32, 2},  {64, 128, 64, 3},
                            {192, 64, 192, 2}, {96, 32, 96, 1},
                            {48, 16, 48, 0}};

struct Acts {
    Grid3 g[4];
    int M[4];
    bf16* xin = nullptr;       // M0 x 8
    bf16* cat[3] = {nullptr, nullptr, nullptr};  // M0x48, M1x96, M2x192
    bf16* pin[4] = {nullptr, nullptr, nullptr, nullptr};  // pooled inputs
                                                          // of E1..E3 (1..3)
    bf16* y3 = nullptr;        // M3 x 128
    bf16* yd[3] = {nullptr, nullptr, nullptr};  // decoder outputs D0, D1, D2
    float* out = nullptr;      // M0
    // gradients (training only)
    bf16* dcat[3] = {nullptr, nullptr, nullptr};
    bf16* dpin[4] = {nullptr, nullptr, nullptr, nullptr};
    bf16* dy3 = nullptr;
    bf16* dyd[3] = {nullptr, nullptr, nullptr};
    float* gout = nullptr;
    float* partial = nullptr;
    float* scal = nullptr;     // [0] shift, [1] loss
    bool train = false;

    void freeAll()
    {
        dfree(xin);
        dfree(y3);
        dfree(out);
        dfree(dy3);
        dfree(gout);
        dfree(partial);
        dfree(scal);
        for (int i = 0; i < 3; i++)
        {
            dfree(cat[i]); dfree(yd[i]); dfree(dcat[i]); dfree(dyd[i]);
        }
        for (int i = 0; i < 4; i++) { dfree(pin[i]); dfree(dpin[i]); }
    }
};

const int kCatC[3] = {48, 96, 192};     // channels of the level buffers
const int kUpC[3] = {32, 64, 128};      // upsampled part [0, kUpC)
const int kDecC[3] = {16, 32, 64};      // decoder outputs D0, D1, D2
const int kPinC[4] = {0, 16, 32, 64};   // pooled inputs of E1, E2, E3

}  // namespace

struct Net {
    hipStream_t stream = nullptr;
    // parameters: 7 convolutions (weights [cout][kp], bias [cout]) and the
    // head (16 weights + bias), FP32, one contiguous array
    long nparam = 0;
    long woff[7], boff[7], hoff = 0;
    int kp[7], kpt[7];
    float* P = nullptr;
    float* G = nullptr;
    float* Mo = nullptr;
    float* Ve = nullptr;
    bf16* WB[7] = {nullptr};     // BF16 weights [cout][kp]
    bf16* WT[7] = {nullptr};     // flipped BF16 weights [cin][kpt]
    int step = 0;
    Acts acts;

    explicit Net(uint64_t seed)
    {
        UCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        long off = 0;
        for (int i = 0; i < 7; i++)
        {
            kp[i] = round32(27*kConvs[i].cin);
            kpt[i] = round32(27*kConvs[i].cout);
            woff[i] = off;
            off += long(kConvs[i].cout)*kp[i];
            boff[i] = off;
            off += kConvs[i].cout;
        }
        hoff = off;
        off += 17;
        nparam = off;
        std::vector<float> h(nparam, 0.0f);
        Rng r{seed*0x9E3779B97F4A7C15ULL + 12345};
        for (int i = 0; i < 7; i++)
        {
            const double sd = std::sqrt(2.0/(27.0*kConvs[i].cinReal));
            for (int n = 0; n < kConvs[i].cout; n++)
                for (int t = 0; t < 27; t++)
                    for (int c = 0; c < kConvs[i].cinReal; c++)
                        h[woff[i] + long(n)*kp[i] + t*kConvs[i].cin + c] =
                            float(sd*r.normal());
        }
        for (int c = 0; c < 16; c++) h[hoff + c] = float(0.25*r.normal());
        P = dalloc<float>(nparam);
        G = dalloc<float>(nparam);
        Mo = dalloc<float>(nparam);
        Ve = dalloc<float>(nparam);
        UCK(hipMemcpy(P, h.data(), nparam*sizeof(float), hipMemcpyHostToDevice));
        for (int i = 0; i < 7; i++)
        {
            WB[i] = dalloc<bf16>(size_t(kConvs[i].cout)*kp[i]);
            WT[i] = dalloc<bf16>(size_t(kConvs[i].cin)*kpt[i]);
        }
        refreshBF16();
        UCK(hipStreamSynchronize(stream));
    }

    ~Net()
    {
        acts.freeAll();
        dfree(P); dfree(G); dfree(Mo); dfree(Ve);
        for (int i = 0; i < 7; i++) { dfree(WB[i]); dfree(WT[i]); }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void refreshBF16()
    {
        for (int i = 0; i < 7; i++)
            to_bf16(P + woff[i], WB[i], long(kConvs[i].cout)*kp[i], stream);
    }

    void ensureActs(int D, int H, int W, bool train)
    {
        if (D % 8 || H % 8 || W % 8 || D < 8 || H < 8 || W < 8)
            throw UnetError("U-Net grid extents must be multiples of 8");
        Acts& a = acts;
        const bool same = a.xin && a.g[0].D == D && a.g[0].H == H &&
                          a.g[0].W == W;
        if (same && (!train || a.train)) return;
        a.freeAll();
        for (int l = 0; l < 4; l++)
        {
            a.g[l] = Grid3{D >> l, H >> l, W >> l};
            a.M[l] = a.g[l].D*a.g[l].H*a.g[l].W;
        }
        a.xin = dalloc<bf16>(size_t(a.M[0])*8);
        for (int l = 0; l < 3; l++)
        {
            a.cat[l] = dalloc<bf16>(size_t(a.M[l])*kCatC[l]);
            a.yd[l] = dalloc<bf16>(size_t(a.M[l])*kDecC[l]);
        }
        for (int l = 1; l < 4; l++)
            a.pin[l] = dalloc<bf16>(size_t(a.M[l])*kPinC[l]);
        a.y3 = dalloc<bf16>(size_t(a.M[3])*128);
        a.out = dalloc<float>(a.M[0]);
        a.partial = dalloc<float>(size_t((a.M[0] + 255)/256)*17 + 64);
        a.scal = dalloc<float>(4);
        a.train = train;
        if (train)
        {
            for (int l = 0; l < 3; l++)
            {
                a.dcat[l] = dalloc<bf16>(size_t(a.M[l])*kCatC[l]);
                a.dyd[l] = dalloc<bf16>(size_t(a.M[l])*kDecC[l]);
            }
            for (int l = 1; l < 4; l++)
                a.dpin[l] = dalloc<bf16>(size_t(a.M[l])*kPinC[l]);
            a.dy3 = dalloc<bf16>(size_t(a.M[3])*128);
            a.gout = dalloc<float>(a.M[0]);
        }
    }

    ConvArgs conv(int i, const bf16* x, int ldx, bf16* y, int ldy) const
    {
        const ConvSpec& c = kConvs[i];
        ConvArgs a;
        a.g = acts.g[c.level];
        a.x = x;
        a.ldx = ldx;
        a.cin = c.cin;
        a.w = WB[i];
        a.kp = kp[i];
        a.taps = 27;
        a.cout = c.cout;
        a.bias = P + boff[i];
        a.relu = 1;
        a.y = y;
        a.ldy = ldy;
        return a;
    }

    // forward from the packed input acts.xin to acts.out
    void forward(hipStream_t s)
    {
        Acts& a = acts;
        conv_fwd(conv(0, a.xin, 8, a.cat[0] + 32, 48), s);
        pool(a.cat[0] + 32, 48, a.g[0], 16, a.pin[1], 16, s);
        conv_fwd(conv(1, a.pin[1], 16, a.cat[1] + 64, 96), s);
        pool(a.cat[1] + 64, 96, a.g[1], 32, a.pin[2], 32, s);
        conv_fwd(conv(2, a.pin[2], 32, a.cat[2] + 128, 192), s);
        pool(a.cat[2] + 128, 192, a.g[2], 64, a.pin[3], 64, s);
        conv_fwd(conv(3, a.pin[3], 64, a.y3, 128), s);
        up(a.y3, 128, a.g[2], 128, a.cat[2], 192, s);
        conv_fwd(conv(4, a.cat[2], 192, a.yd[2], 64), s);
        up(a.yd[2], 64, a.g[1], 64, a.cat[1], 96, s);
        conv_fwd(conv(5, a.cat[1], 96, a.yd[1], 32), s);
        up(a.yd[1], 32, a.g[0], 32, a.cat[0], 48, s);
        conv_fwd(conv(6, a.cat[0], 48, a.yd[0], 16), s);
        head_fwd(a.yd[0], 16, a.M[0], P + hoff, a.out, s);
    }

    // gradient of one convolution: dz (post-ReLU mask applied) -> weight
    // and bias gradients, and optionally the input gradient dx
    void convBackward(int i, const bf16* x, int ldx, const bf16* dz, int lddz,
                      bf16* dx, int lddx, hipStream_t s)
    {
        const ConvSpec& c = kConvs[i];
        const Grid3 g = acts.g[c.level];
        const int M = g.D*g.H*g.W;
        WgradArgs w;
        w.g = g;
        w.x = x;
        w.ldx = ldx;
        w.cin = c.cin;
        w.dz = dz;
        w.lddz = lddz;
        w.cout = c.cout;
        w.taps = 27;
        w.kp = kp[i];
        w.dw = G + woff[i];
        w.chunk = std::max(256, (M + 255)/256/32*32);
        conv_wgrad(w, s);
        // bias gradient: column sums of dz through a 1x1x1 weight gradient
        // against a constant-one input is avoided; sum dz directly
        biasGrad(dz, lddz, M, c.cout, G + boff[i], s);
        if (dx)
        {
            flip_weights(WB[i], c.cin, c.cout, kp[i], 27, WT[i], kpt[i], s);
            ConvArgs d;
            d.g = g;
            d.x = dz;
            d.ldx = lddz;
            d.cin = c.cout;
            d.w = WT[i];
            d.kp = kpt[i];
            d.taps = 27;
            d.cout = c.cin;
            d.bias = nullptr;
            d.relu = 0;
            d.y = dx;
            d.ldy = lddx;
            conv_fwd(d, s);
        }
    }

    void biasGrad(const bf16* dz, int lddz, int M, int C, float* out,
                  hipStream_t s);

    // backward from acts.gout (dL/d out) to G
    void backward(hipStream_t s)
    {
        Acts& a = acts;
        UCK(hipMemsetAsync(G, 0, nparam*sizeof(float), s));
        // head
        const int nb = head_bwd(a.yd[0], 16, a.M[0], P + hoff, a.gout, a.dyd[0],
                                16, a.partial, s);
        col_sum(a.partial, nb, 17, 1.0f, G + hoff, s);
        // D0
        relu_bwd(a.dyd[0], 16, a.yd[0], 16, a.M[0], 16, s);
        convBackward(6, a.cat[0], 48, a.dyd[0], 16, a.dcat[0], 48, s);
        up_bwd(a.dcat[0], 48, a.g[0], 32, a.dyd[1], 32, s);
        // D1
        relu_bwd(a.dyd[1], 32, a.yd[1], 32, a.M[1], 32, s);
        convBackward(5, a.cat[1], 96, a.dyd[1], 32, a.dcat[1], 96, s);
        up_bwd(a.dcat[1], 96, a.g[1], 64, a.dyd[2], 64, s);
        // D2
        relu_bwd(a.dyd[2], 64, a.yd[2], 64, a.M[2], 64, s);
        convBackward(4, a.cat[2], 192, a.dyd[2], 64, a.dcat[2], 192, s);
        up_bwd(a.dcat[2], 192, a.g[2], 128, a.dy3, 128, s);
        // E3
        relu_bwd(a.dy3, 128, a.y3, 128, a.M[3], 128, s);
        convBackward(3, a.pin[3], 64, a.dy3, 128, a.dpin[3], 64, s);
        // E2: skip gradient + pooled gradient
        pool_bwd(a.dpin[3], 64, a.g[2], 64, a.dcat[2] + 128, 192, s);
        relu_bwd(a.dcat[2] + 128, 192, a.cat[2] + 128, 192, a.M[2], 64, s);
        convBackward(2, a.pin[2], 32, a.dcat[2] + 128, 192, a.dpin[2], 32, s);
        // E1
        pool_bwd(a.dpin[2], 32, a.g[1], 32, a.dcat[1] + 64, 96, s);
        relu_bwd(a.dcat[1] + 64, 96, a.cat[1] + 64, 96, a.M[1], 32, s);
        convBackward(1, a.pin[1], 16, a.dcat[1] + 64, 96, a.dpin[1], 16, s);
        // E0
        pool_bwd(a.dpin[1], 16, a.g[0], 16, a.dcat[0] + 32, 48, s);
        relu_bwd(a.dcat[0] + 32, 48, a.cat[0] + 32, 48, a.M[0], 16, s);
        convBackward(0, a.xin, 8, a.dcat[0] + 32, 48, nullptr, 0, s);
    }
};

namespace {

// per-block partial column sums of a BF16 slice (C <= 256 channels)
__global__ __launch_bounds__(256) void k_bias_partial(const bf16* dz, int ld,
                                                      int M, int C, int rows,
                                                      float* partial)
{
    const int c = threadIdx.x;
    if (c >= C) return;
    const int m0 = blockIdx.x*rows;
    const int m1 = min(M, m0 + rows);
    float s = 0.0f;
    for (int m = m0; m < m1; m++) s += float(dz[size_t(m)*ld + c]);
    partial[size_t(blockIdx.x)*C + c] = s;
}

}  // namespace

void Net::biasGrad(const bf16* dz, int lddz, int M, int C, float* out,
                   hipStream_t s)
{
    const int rows = 1024;
    const int nb = (M + rows - 1)/rows;
    // acts.partial holds at least 17*M/256 floats >= C*M/1024 for C <= 128
    hipLaunchKernelGGL(k_bias_partial, dim3(nb), dim3(256), 0, s, dz, lddz, M,
                       C, rows, acts.partial);
    col_sum(acts.partial, nb, C, 1.0f, out, s);
}

}  // namespace unet
}  // namespace bmg

// ---------------------------------------------------------------------------
// C ABI
using bmg::unet::Net;
using bmg::unet::UnetError;

struct bmg_unet
{
    std::unique_ptr<Net> net;
};

namespace {

thread_local std::string g_unet_err;

template <typename F>
int urun(F&& f)
{
    try
    {
        f();
        return 0;
    }
    catch (const std::exception& e)
    {
        g_unet_err = e.what();
        return -2;
    }
}

}  // namespace

namespace bmg {
namespace unet {

void net_prepare(::bmg_unet* u, int D, int H, int W)
{
    if (!u) throw UnetError("no learned network attached");
    u->net->ensureActs(D, H, W, false);
}

void net_forward(::bmg_unet* u, int D, int H, int W, const float* f,
                 const float* c, float* out, hipStream_t s)
{
    Net& n = *u->net;
    n.ensureActs(D, H, W, false);
    pack_input(f, c, n.acts.M[0], n.acts.xin, s);
    n.forward(s);
    UCK(hipMemcpyAsync(out, n.acts.out, size_t(n.acts.M[0])*sizeof(float),
                       hipMemcpyDeviceToDevice, s));
}

}  // namespace unet
}  // namespace bmg

extern "C" {

const char* bmg_unet_last_error(void) { return g_unet_err.c_str(); }

int bmg_unet_create(bmg_unet** out, unsigned long long seed)
{
    return urun([&] {
        if (!out) throw UnetError("null output pointer");
        auto u = std::make_unique<bmg_unet>();
        u->net = std::make_unique<Net>(seed);
        *out = u.release();
    });
}

int bmg_unet_destroy(bmg_unet* u)
{
    delete u;
    return 0;
}

long bmg_unet_num_params(const bmg_unet* u) { return u ? u->net->nparam : 0; }

int bmg_unet_get_params(const bmg_unet* u, float* host)
{
    return urun([&] {
        UCK(hipStreamSynchronize(u->net->stream));
        UCK(hipMemcpy(host, u->net->P, u->net->nparam*sizeof(float),
                      hipMemcpyDeviceToHost));
    });
}

int bmg_unet_set_params(bmg_unet* u, const float* host)
{
    return urun([&] {
        Net& n = *u->net;
        UCK(hipMemcpy(n.P, host, n.nparam*sizeof(float), hipMemcpyHostToDevice));
        n.refreshBF16();
        UCK(hipStreamSynchronize(n.stream));
    });
}

int bmg_unet_forward(bmg_unet* u, int D, int H, int W, const float* f,
                     const float* c, float* out)
{
    return urun([&] {
        Net& n = *u->net;
        n.ensureActs(D, H, W, false);
        bmg::unet::pack_input(f, c, n.acts.M[0], n.acts.xin, n.stream);
        n.forward(n.stream);
        UCK(hipMemcpyAsync(out, n.acts.out, size_t(n.acts.M[0])*sizeof(float),
                           hipMemcpyDeviceToDevice, n.stream));
        UCK(hipGetLastError());
        UCK(hipStreamSynchronize(n.stream));
    });
}

int bmg_unet_train_step(bmg_unet* u, int D, int H, int W, const float* f,
                        const float* c, const float* target, float lr,
                        float* loss)
{
    return urun([&] {
        Net& n = *u->net;
        n.ensureActs(D, H, W, true);
        hipStream_t s = n.stream;
        bmg::unet::Acts& a = n.acts;
        bmg::unet::pack_input(f, c, a.M[0], a.xin, s);
        n.forward(s);
        // mean-removed squared error (the solution is defined up to a
        // constant for the all-Neumann problems)
        const int nb = bmg::unet::diff_sum(a.out, target, a.M[0], a.partial, s);
        bmg::unet::col_sum(a.partial, nb, 1, 1.0f/float(a.M[0]), a.scal, s);
        const int nb2 = bmg::unet::mse(a.out, target, a.M[0], a.scal, a.gout,
                                       a.partial, s);
        bmg::unet::col_sum(a.partial, nb2, 1, 1.0f/float(a.M[0]), a.scal + 1, s);
        n.backward(s);
        n.step++;
        bmg::unet::adam(n.P, n.Mo, n.Ve, n.G, int(n.nparam), lr, 0.9f, 0.999f,
                        1e-8f, n.step, nullptr, s);
        n.refreshBF16();
        float h = 0.0f;
        UCK(hipMemcpyAsync(&h, a.scal + 1, sizeof(float), hipMemcpyDeviceToHost,
                           s));
        UCK(hipGetLastError());
        UCK(hipStreamSynchronize(s));
        if (loss) *loss = h;
    });
}

int bmg_unet_get_grads(const bmg_unet* u, float* host)
{
    return urun([&] {
        UCK(hipStreamSynchronize(u->net->stream));
        UCK(hipMemcpy(host, u->net->G, u->net->nparam*sizeof(float),
                      hipMemcpyDeviceToHost));
    });
}

int bmg_unet_op_conv(int D, int H, int W, const uint16_t* x, int ldx, int cin,
                     const uint16_t* w, int kp, int taps, int cout,
                     const float* bias, int relu, uint16_t* y, int ldy)
{
    return urun([&] {
        bmg::unet::ConvArgs a;
        a.g = bmg::unet::Grid3{D, H, W};
        a.x = reinterpret_cast<const bmg::unet::bf16*>(x);
        a.ldx = ldx;
        a.cin = cin;
        a.w = reinterpret_cast<const bmg::unet::bf16*>(w);
        a.kp = kp;
        a.taps = taps;
        a.cout = cout;
        a.bias = bias;
        a.relu = relu;
        a.y = reinterpret_cast<bmg::unet::bf16*>(y);
        a.ldy = ldy;
        bmg::unet::conv_fwd(a, nullptr);
        UCK(hipGetLastError());
        UCK(hipDeviceSynchronize());
    });
}

int bmg_unet_op_conv_dgrad(int D, int H, int W, const uint16_t* dz, int lddz,
                           int cout, const uint16_t* w, int kp, int cin,
                           uint16_t* dx, int lddx)
{
    return urun([&] {
        const int kpt = bmg::unet::round32(27*cout);
        bmg::unet::bf16* wt = bmg::unet::dalloc<bmg::unet::bf16>(size_t(cin)*kpt);
        try
        {
            bmg::unet::flip_weights(reinterpret_cast<const bmg::unet::bf16*>(w),
                                    cin, cout, kp, 27, wt, kpt, nullptr);
            bmg::unet::ConvArgs a;
            a.g = bmg::unet::Grid3{D, H, W};
            a.x = reinterpret_cast<const bmg::unet::bf16*>(dz);
            a.ldx = lddz;
            a.cin = cout;
            a.w = wt;
            a.kp = kpt;
            a.taps = 27;
            a.cout = cin;
            a.bias = nullptr;
            a.relu = 0;
            a.y = reinterpret_cast<bmg::unet::bf16*>(dx);
            a.ldy = lddx;
            bmg::unet::conv_fwd(a, nullptr);
            UCK(hipGetLastError());
            UCK(hipDeviceSynchronize());
        }
        catch (...)
        {
            bmg::unet::dfree(wt);
            throw;
        }
        bmg::unet::dfree(wt);
    });
}

int bmg_unet_op_conv_wgrad(int D, int H, int W, const uint16_t* x, int ldx,
                           int cin, const uint16_t* dz, int lddz, int cout,
                           int kp, int chunk, float* dw)
{
    return urun([&] {
        bmg::unet::WgradArgs a;
        a.g = bmg::unet::Grid3{D, H, W};
        a.x = reinterpret_cast<const bmg::unet::bf16*>(x);
        a.ldx = ldx;
        a.cin = cin;
        a.dz = reinterpret_cast<const bmg::unet::bf16*>(dz);
        a.lddz = lddz;
        a.cout = cout;
        a.taps = 27;
        a.kp = kp;
        a.dw = dw;
        a.chunk = chunk;
        bmg::unet::conv_wgrad(a, nullptr);
        UCK(hipGetLastError());
        UCK(hipDeviceSynchronize());
    });
}

}  // extern "C"
