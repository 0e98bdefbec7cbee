// surrogate_capi.hip - the warm-start network behind include/dsmc_surrogate.h: parameters,
// activations and the training set in device memory, forward / backward / Adam as launch
// sequences on one HIP stream, no host synchronisation inside a training run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsmc_surrogate.h"
#include "dsmc_device.h"
#include "surrogate_kernels.h"

using namespace dsmc::net;

namespace {

constexpr int kLayers = 14;
constexpr int kPools = 3;
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B };

struct LayerDef {
    int c1, c2, cout, level;
};

// splitmix64 (Steele, Lea, Flood, OOPSLA'14): a fully specified 64-bit generator for the weight
// initialisation and the shuffles, so both are reproducible from the seed alone.
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()  // Box-Muller
    {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
    }
};

struct Freestream {
    double n, u, t;
};

Freestream freestream(const dsmc_params& p)
{
    const double m = p.molar_mass / dsmc::kAvogadro;
    return {p.n_free, p.mach * std::sqrt(5.0 / 3.0 * dsmc::kBoltzmann * p.t_free / m), p.t_free};
}

NetCase net_case(const dsmc_params& p)
{
    NetCase c{};
    c.box[0] = (float)(p.box_x - 0.5 * p.box_width);
    c.box[1] = (float)(p.box_x + 0.5 * p.box_width);
    c.box[2] = (float)(p.box_y - 0.5 * p.box_height);
    c.box[3] = (float)(p.box_y + 0.5 * p.box_height);
    c.dom[0] = (float)p.x1;
    c.dom[1] = (float)p.x2;
    c.dom[2] = (float)p.y1;
    c.dom[3] = (float)p.y2;
    double an = 1.0, at = 1.0;
    if (p.wall_model == DSMC_WALL_SPECULAR) an = at = 0.0;
    if (p.wall_model == DSMC_WALL_CLL) { an = p.acc_n; at = p.acc_t; }
    const float f[kNumFeatures] = {
        (float)((p.mach - 5.0) / 3.0),
        (float)(p.attack_angle_deg / 30.0),
        (float)(std::log10(p.n_free) - 17.0),
        p.wall_model == DSMC_WALL_DIFFUSE ? 1.f : 0.f,
        p.wall_model == DSMC_WALL_SPECULAR ? 1.f : 0.f,
        p.wall_model == DSMC_WALL_CLL ? 1.f : 0.f,
        (float)an,
        (float)at,
        (float)(p.t_wall / p.t_free - 1.5),
    };
    for (int i = 0; i < kNumFeatures; ++i) c.feat[i] = f[i];
    return c;
}

// The obstacle test of the input raster, in the same float arithmetic as k_encode.
bool raster_in_box(const NetCase& c, int k, int l, int H, int W)
{
    const float fx = ((float)k + 0.5f) / (float)H, fy = ((float)l + 0.5f) / (float)W;
    const float X = c.dom[0] + (c.dom[1] - c.dom[0]) * fx;
    const float Y = c.dom[2] + (c.dom[3] - c.dom[2]) * fy;
    return X > c.box[0] && X < c.box[1] && Y > c.box[2] && Y < c.box[3];
}

int check_config(const dsmc_net_config& c)
{
    if (c.height < 8 || c.width < 8 || c.height % 8 || c.width % 8) return 1;
    if (c.base_channels != 32 && c.base_channels != 64) return 1;
    if (c.max_batch < 1) return 1;
    if ((int64_t)c.max_batch * c.height * c.width > (1ll << 30)) return 1;
    return 0;
}

// Targets and weights of one case (see dsmc_net_normalize).
void normalize_case(const dsmc_net_config& cfg, const dsmc_params& p, const float* f, float* tgt,
                    float* wgt)
{
    const int H = cfg.height, W = cfg.width;
    const size_t HW = (size_t)H * W;
    const Freestream fs = freestream(p);
    const NetCase nc = net_case(p);
    size_t gas = 0;
    for (int k = 0; k < H; ++k)
        for (int l = 0; l < W; ++l) {
            const size_t i = (size_t)k * W + l;
            const double n = f[i], ux = f[HW + i], uy = f[2 * HW + i], t = f[3 * HW + i];
            const bool ok = !raster_in_box(nc, k, l, H, W) && n > 0.0 && t > 0.0
                            && std::isfinite(n) && std::isfinite(ux) && std::isfinite(uy)
                            && std::isfinite(t);
            float* o = tgt + 4 * i;
            if (ok) {
                o[0] = (float)std::log(std::max(n, 1e-3 * fs.n) / fs.n);
                o[1] = (float)(ux / fs.u);
                o[2] = (float)(uy / fs.u);
                o[3] = (float)std::log(std::max(t, 1e-3 * fs.t) / fs.t);
                ++gas;
            } else {
                o[0] = o[1] = o[2] = o[3] = 0.f;
            }
            wgt[i] = ok ? 1.f : 0.f;
        }
    const float w = gas ? (float)(1.0 / (4.0 * (double)gas)) : 0.f;
    for (size_t i = 0; i < HW; ++i) wgt[i] *= w;
}

}  // namespace

struct dsmc_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    dsmc_net_config cfg{};
    int C = 32, H = 0, W = 0, Bmax = 1;
    LayerDef L[kLayers];
    size_t woff[kLayers], boff[kLayers], head_off = 0, nparams = 0;
    std::vector<void*> allocs;
    float* w = nullptr;   // FP32 master parameters
    float* g = nullptr;   // gradient
    float* m = nullptr;   // Adam moments
    float* v = nullptr;
    uint16_t* wb[kLayers] = {};
    uint16_t* wt[kLayers] = {};
    uint16_t* in0 = nullptr;
    uint16_t* act[kLayers] = {};
    uint16_t* pool[kPools] = {};
    uint16_t* dz[kLayers] = {};
    uint16_t* dskip[kPools] = {};
    uint16_t* dpool[kPools] = {};
    uint16_t* dup = nullptr;
    float* out = nullptr;        // [Bmax][HW][4]
    float* gout = nullptr;       // [Bmax][HW][4]
    float* loss_part = nullptr;  // [blocks of the head]
    float* loss_acc = nullptr;   // [1]
    float* part = nullptr;
    size_t part_cap = 0;
    float* stage = nullptr;      // float staging, [Bmax][HW][max(32, C)]
    float* tgt_b = nullptr;      // [Bmax][HW][4]
    float* wgt_b = nullptr;      // [Bmax][HW]
    NetCase* cases_b = nullptr;  // [Bmax]
    // training set
    NetCase* ds_cases = nullptr;
    float* ds_tgt = nullptr;
    float* ds_wgt = nullptr;
    int ds_n = 0;
    int* d_idx = nullptr;
    size_t d_idx_cap = 0;
    float* d_hist = nullptr;
    size_t d_hist_cap = 0;
    int64_t adam_t = 0;
    int last_B = 0;              // batch of the last forward pass
    double last_ms = 0.0, last_flops = 0.0;

    int h(int level) const { return H >> level; }
    int wd(int level) const { return W >> level; }
    size_t pix(int B, int level) const { return (size_t)B * h(level) * wd(level); }
};

namespace {

int fail(dsmc_net* n, int code, const std::string& msg)
{
    if (n) n->err = msg;
    return code;
}

#define NET_TRY(n, expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((n), DSMC_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));    \
    } while (0)

template <typename T>
int nalloc(dsmc_net* n, T** p, size_t count)
{
    void* q = nullptr;
    NET_TRY(n, hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    NET_TRY(n, hipMemsetAsync(q, 0, std::max<size_t>(count, 1) * sizeof(T), n->stream));
    n->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

void nfree(dsmc_net* n, void* p)
{
    if (!p) return;
    for (auto& a : n->allocs)
        if (a == p) {
            (void)hipFree(a);
            a = nullptr;
        }
}

int cin_of(const LayerDef& d) { return d.c1 + d.c2; }

// Split of the pixel dimension of a weight-gradient launch into `splits` ranges of `rows`.
void wgrad_split(const LayerDef& d, size_t P, int& splits, int& rows)
{
    const int cin = cin_of(d);
    const int tco = d.cout % 64 == 0 ? 64 : 32;
    const int tci = cin % 64 == 0 && d.c1 % 64 == 0 ? 64 : 32;
    const size_t tiles = (size_t)9 * (cin / tci) * (d.cout / tco);
    size_t want = (2048 + tiles - 1) / tiles;
    const size_t most = (P + 63) / 64;
    want = std::max<size_t>(1, std::min(want, most));
    rows = (int)(((P + want - 1) / want + 63) / 64 * 64);
    splits = (int)((P + rows - 1) / rows);
}

constexpr int kBiasRows = 1024;
constexpr int kHeadRows = 512;

size_t part_need(const dsmc_net* n)
{
    size_t need = 0;
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = n->L[l];
        const size_t P = n->pix(n->Bmax, d.level);
        int splits, rows;
        wgrad_split(d, P, splits, rows);
        need = std::max(need, (size_t)splits * d.cout * 9 * cin_of(d));
        need = std::max(need, ((P + kBiasRows - 1) / kBiasRows) * d.cout);
    }
    const size_t P0 = n->pix(n->Bmax, 0);
    need = std::max(need, ((P0 + kHeadRows - 1) / kHeadRows) * (4 * n->C + 4));
    return need;
}

void init_weights(dsmc_net* n, uint64_t seed, std::vector<float>& hw)
{
    SplitMix rng{seed};
    hw.assign(n->nparams, 0.f);
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = n->L[l];
        const size_t cnt = (size_t)d.cout * 9 * cin_of(d);
        const double sd = std::sqrt(2.0 / (9.0 * cin_of(d)));
        for (size_t i = 0; i < cnt; ++i) hw[n->woff[l] + i] = (float)(sd * rng.normal());
    }
    const double sd = std::sqrt(1.0 / n->C);
    for (int i = 0; i < 4 * n->C; ++i) hw[n->head_off + i] = (float)(sd * rng.normal());
}

void pack_all(dsmc_net* n)
{
    for (int l = 0; l < kLayers; ++l)
        launch_pack_weights(n->w + n->woff[l], n->L[l].cout, cin_of(n->L[l]), n->wb[l], n->wt[l],
                            n->stream);
}

// Forward pass of batch B (inputs already in in0).
void forward(dsmc_net* n, int B)
{
    const hipStream_t s = n->stream;
    n->last_B = B;
    auto conv = [&](int l, const uint16_t* x1, const uint16_t* x2) {
        const LayerDef& d = n->L[l];
        ConvArgs a{};
        a.x1 = x1;
        a.x2 = x2;
        a.w = n->wb[l];
        a.bias = n->w + n->boff[l];
        a.y1 = n->act[l];
        a.c1 = d.c1;
        a.c2 = d.c2;
        a.cout = d.cout;
        a.csplit = d.cout;
        a.B = B;
        a.H = n->h(d.level);
        a.W = n->wd(d.level);
        launch_conv3x3(a, kConvBiasRelu, s);
    };
    const int C = n->C;
    conv(E0A, n->in0, nullptr);
    conv(E0B, n->act[E0A], nullptr);
    launch_pool2(n->act[E0B], B, n->h(0), n->wd(0), C, n->pool[0], s);
    conv(E1A, n->pool[0], nullptr);
    conv(E1B, n->act[E1A], nullptr);
    launch_pool2(n->act[E1B], B, n->h(1), n->wd(1), 2 * C, n->pool[1], s);
    conv(E2A, n->pool[1], nullptr);
    conv(E2B, n->act[E2A], nullptr);
    launch_pool2(n->act[E2B], B, n->h(2), n->wd(2), 4 * C, n->pool[2], s);
    conv(E3A, n->pool[2], nullptr);
    conv(E3B, n->act[E3A], nullptr);
    conv(D2A, n->act[E2B], n->act[E3B]);
    conv(D2B, n->act[D2A], nullptr);
    conv(D1A, n->act[E1B], n->act[D2B]);
    conv(D1B, n->act[D1A], nullptr);
    conv(D0A, n->act[E0B], n->act[D1B]);
    conv(D0B, n->act[D0A], nullptr);
}

// Forward inputs x1, x2 of layer l.
void layer_inputs(const dsmc_net* n, int l, const uint16_t*& x1, const uint16_t*& x2)
{
    static const int prev[kLayers] = {-1, E0A, -2, E1A, -3, E2A, -4, E3A, -5, D2A, -6, D1A, -7, D0A};
    x2 = nullptr;
    switch (l) {
    case E0A: x1 = n->in0; break;
    case E1A: x1 = n->pool[0]; break;
    case E2A: x1 = n->pool[1]; break;
    case E3A: x1 = n->pool[2]; break;
    case D2A: x1 = n->act[E2B]; x2 = n->act[E3B]; break;
    case D1A: x1 = n->act[E1B]; x2 = n->act[D2B]; break;
    case D0A: x1 = n->act[E0B]; x2 = n->act[D1B]; break;
    default: x1 = n->act[prev[l]]; break;
    }
}

// The head on the last activation; with targets, its loss terms and gradients. Returns the
// head's block count (the number of loss partials).
int head(dsmc_net* n, int B, const int* idx, const float* tgt, const float* wgt, float* out)
{
    HeadArgs a{};
    a.h = n->act[D0B];
    a.wh = n->w + n->head_off;
    a.bh = n->w + n->head_off + 4 * n->C;
    a.C = n->C;
    a.P = (int)n->pix(B, 0);
    a.HW = n->H * n->W;
    a.idx = idx;
    a.tgt = tgt;
    a.wgt = wgt;
    a.scale = 1.f / (float)B;
    a.out = out;
    a.g = n->gout;
    a.dh = n->dz[D0B];
    a.loss_part = n->loss_part;
    return launch_head(a, n->stream);
}

// Backward pass after head(): every parameter's gradient into n->g.
void backward(dsmc_net* n, int B)
{
    const hipStream_t s = n->stream;
    const int C = n->C;
    {
        const int P0 = (int)n->pix(B, 0);
        const int splits = (P0 + kHeadRows - 1) / kHeadRows;
        launch_head_wgrad(n->act[D0B], n->gout, C, P0, kHeadRows, splits, n->part, s);
        launch_reduce(n->part, splits, 4 * C + 4, n->g + n->head_off, s);
    }
    auto dgrad = [&](int l, ConvMode mode, const uint16_t* mask, uint16_t* y1, uint16_t* y2,
                     int csplit) {
        const LayerDef& d = n->L[l];
        ConvArgs a{};
        a.x1 = n->dz[l];
        a.c1 = d.cout;
        a.c2 = 0;
        a.w = n->wt[l];
        a.mask = mask;
        a.y1 = y1;
        a.y2 = y2;
        a.cout = cin_of(d);
        a.csplit = csplit;
        a.B = B;
        a.H = n->h(d.level);
        a.W = n->wd(d.level);
        launch_conv3x3(a, mode, s);
    };
    for (int l = kLayers - 1; l >= 0; --l) {
        const LayerDef& d = n->L[l];
        const int Hl = n->h(d.level), Wl = n->wd(d.level);
        const size_t P = n->pix(B, d.level);
        WgradArgs wa{};
        wa.dz = n->dz[l];
        layer_inputs(n, l, wa.x1, wa.x2);
        wa.cout = d.cout;
        wa.c1 = d.c1;
        wa.c2 = d.c2;
        wa.B = B;
        wa.H = Hl;
        wa.W = Wl;
        int splits;
        wgrad_split(d, P, splits, wa.rows);
        wa.part = n->part;
        launch_wgrad(wa, splits, s);
        launch_reduce(n->part, splits, (size_t)d.cout * 9 * cin_of(d), n->g + n->woff[l], s);
        const int bsplits = (int)((P + kBiasRows - 1) / kBiasRows);
        launch_bias_grad(n->dz[l], d.cout, (int)P, kBiasRows, bsplits, n->part, s);
        launch_reduce(n->part, bsplits, d.cout, n->g + n->boff[l], s);
        switch (l) {
        case D0B: case D1B: case D2B: case E3B: case E2B: case E1B: case E0B:
            dgrad(l, kConvMask, n->act[l - 1], n->dz[l - 1], nullptr, cin_of(d));
            break;
        case D0A:
            dgrad(l, kConvPlain, nullptr, n->dskip[0], n->dup, C);
            launch_up_back(n->dup, n->act[D1B], B, Hl, Wl, 2 * C, n->dz[D1B], s);
            break;
        case D1A:
            dgrad(l, kConvPlain, nullptr, n->dskip[1], n->dup, 2 * C);
            launch_up_back(n->dup, n->act[D2B], B, Hl, Wl, 4 * C, n->dz[D2B], s);
            break;
        case D2A:
            dgrad(l, kConvPlain, nullptr, n->dskip[2], n->dup, 4 * C);
            launch_up_back(n->dup, n->act[E3B], B, Hl, Wl, 8 * C, n->dz[E3B], s);
            break;
        case E3A:
            dgrad(l, kConvPlain, nullptr, n->dpool[2], nullptr, cin_of(d));
            launch_pool_back_add(n->dskip[2], n->dpool[2], n->act[E2B], B, n->h(2), n->wd(2),
                                 4 * C, n->dz[E2B], s);
            break;
        case E2A:
            dgrad(l, kConvPlain, nullptr, n->dpool[1], nullptr, cin_of(d));
            launch_pool_back_add(n->dskip[1], n->dpool[1], n->act[E1B], B, n->h(1), n->wd(1),
                                 2 * C, n->dz[E1B], s);
            break;
        case E1A:
            dgrad(l, kConvPlain, nullptr, n->dpool[0], nullptr, cin_of(d));
            launch_pool_back_add(n->dskip[0], n->dpool[0], n->act[E0B], B, n->h(0), n->wd(0), C,
                                 n->dz[E0B], s);
            break;
        default:
            break;  // E0A: no gradient with respect to the inputs
        }
    }
}

double conv_flops_per_step(const dsmc_net* n, int B)
{
    double f = 0.0;
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = n->L[l];
        const double mac = (double)n->pix(B, d.level) * d.cout * 9.0 * cin_of(d);
        f += 2.0 * mac * (l == E0A ? 2.0 : 3.0);
    }
    return f;
}

int check_net(const dsmc_net* n) { return n ? 0 : DSMC_E_ARGUMENT; }

int upload_inputs(dsmc_net* n, const float* inputs, int B)
{
    const size_t cnt = n->pix(B, 0) * DSMC_NET_IN_CHANNELS;
    NET_TRY(n, hipMemcpyAsync(n->stage, inputs, cnt * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    launch_to_bf16(n->stage, cnt, n->in0, n->stream);
    return 0;
}

int sync(dsmc_net* n)
{
    NET_TRY(n, hipStreamSynchronize(n->stream));
    NET_TRY(n, hipGetLastError());
    return 0;
}

int set_device(dsmc_net* n)
{
    NET_TRY(n, hipSetDevice(n->device));
    return 0;
}

}  // namespace

extern "C" {

void dsmc_net_config_default(dsmc_net_config* cfg)
{
    if (!cfg) return;
    cfg->height = 256;
    cfg->width = 256;
    cfg->base_channels = 32;
    cfg->max_batch = 8;
    cfg->seed = 1;
}

int dsmc_net_create(const dsmc_net_config* cfg, int device, dsmc_net** out)
{
    if (!out) return DSMC_E_ARGUMENT;
    *out = nullptr;
    if (!cfg || check_config(*cfg)) return DSMC_E_ARGUMENT;
    dsmc_net* n = new dsmc_net();
    *out = n;
    n->cfg = *cfg;
    n->device = device;
    n->C = cfg->base_channels;
    n->H = cfg->height;
    n->W = cfg->width;
    n->Bmax = cfg->max_batch;
    const int C = n->C;
    const LayerDef defs[kLayers] = {
        {DSMC_NET_IN_CHANNELS, 0, C, 0}, {C, 0, C, 0},
        {C, 0, 2 * C, 1},                {2 * C, 0, 2 * C, 1},
        {2 * C, 0, 4 * C, 2},            {4 * C, 0, 4 * C, 2},
        {4 * C, 0, 8 * C, 3},            {8 * C, 0, 8 * C, 3},
        {4 * C, 8 * C, 4 * C, 2},        {4 * C, 0, 4 * C, 2},
        {2 * C, 4 * C, 2 * C, 1},        {2 * C, 0, 2 * C, 1},
        {C, 2 * C, C, 0},                {C, 0, C, 0},
    };
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
        n->L[l] = defs[l];
        n->woff[l] = off;
        off += (size_t)defs[l].cout * 9 * cin_of(defs[l]);
        n->boff[l] = off;
        off += defs[l].cout;
    }
    n->head_off = off;
    off += 4 * C + 4;
    n->nparams = off;
    if (int rc = set_device(n)) return rc;
    NET_TRY(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    int rc = 0;
    const int B = n->Bmax;
    if ((rc = nalloc(n, &n->w, n->nparams)) || (rc = nalloc(n, &n->g, n->nparams))
        || (rc = nalloc(n, &n->m, n->nparams)) || (rc = nalloc(n, &n->v, n->nparams)))
        return rc;
    for (int l = 0; l < kLayers; ++l) {
        const size_t cnt = (size_t)n->L[l].cout * 9 * cin_of(n->L[l]);
        if ((rc = nalloc(n, &n->wb[l], cnt)) || (rc = nalloc(n, &n->wt[l], cnt))) return rc;
        if ((rc = nalloc(n, &n->act[l], n->pix(B, n->L[l].level) * n->L[l].cout))) return rc;
        if ((rc = nalloc(n, &n->dz[l], n->pix(B, n->L[l].level) * n->L[l].cout))) return rc;
    }
    if ((rc = nalloc(n, &n->in0, n->pix(B, 0) * DSMC_NET_IN_CHANNELS))) return rc;
    for (int k = 0; k < kPools; ++k) {
        const size_t ck = (size_t)C << k;
        if ((rc = nalloc(n, &n->pool[k], n->pix(B, k + 1) * ck))
            || (rc = nalloc(n, &n->dpool[k], n->pix(B, k + 1) * ck))
            || (rc = nalloc(n, &n->dskip[k], n->pix(B, k) * ck)))
            return rc;
    }
    // the decoder's upsampled-gradient buffer: largest of 2C at level 0, 4C at 1, 8C at 2
    size_t dup_n = 0;
    for (int k = 0; k < kPools; ++k) dup_n = std::max(dup_n, n->pix(B, k) * ((size_t)C << (k + 1)));
    const size_t P0 = n->pix(B, 0);
    if ((rc = nalloc(n, &n->dup, dup_n)) || (rc = nalloc(n, &n->out, P0 * 4))
        || (rc = nalloc(n, &n->gout, P0 * 4)) || (rc = nalloc(n, &n->loss_part, (P0 + 255) / 256))
        || (rc = nalloc(n, &n->loss_acc, 1))
        || (rc = nalloc(n, &n->stage, P0 * std::max(DSMC_NET_IN_CHANNELS, C)))
        || (rc = nalloc(n, &n->tgt_b, P0 * 4)) || (rc = nalloc(n, &n->wgt_b, P0))
        || (rc = nalloc(n, &n->cases_b, B)))
        return rc;
    n->part_cap = part_need(n);
    if ((rc = nalloc(n, &n->part, n->part_cap))) return rc;
    std::vector<float> hw;
    init_weights(n, cfg->seed, hw);
    return dsmc_net_set_weights(n, hw.data(), hw.size());
}

void dsmc_net_destroy(dsmc_net* n)
{
    if (!n) return;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void* p : n->allocs)
        if (p) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

const char* dsmc_net_last_error(const dsmc_net* n) { return n ? n->err.c_str() : "no network"; }

int64_t dsmc_net_num_params(const dsmc_net* n) { return n ? (int64_t)n->nparams : -1; }

int dsmc_net_get_weights(dsmc_net* n, float* out, size_t cnt)
{
    if (int rc = check_net(n)) return rc;
    if (!out || cnt < n->nparams) return fail(n, DSMC_E_ARGUMENT, "weight array too small");
    if (int rc = set_device(n)) return rc;
    NET_TRY(n, hipMemcpyAsync(out, n->w, n->nparams * sizeof(float), hipMemcpyDeviceToHost,
                              n->stream));
    return sync(n);
}

int dsmc_net_set_weights(dsmc_net* n, const float* in, size_t cnt)
{
    if (int rc = check_net(n)) return rc;
    if (!in || cnt != n->nparams) return fail(n, DSMC_E_ARGUMENT, "weight count mismatch");
    if (int rc = set_device(n)) return rc;
    NET_TRY(n, hipMemcpyAsync(n->w, in, n->nparams * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    NET_TRY(n, hipMemsetAsync(n->m, 0, n->nparams * sizeof(float), n->stream));
    NET_TRY(n, hipMemsetAsync(n->v, 0, n->nparams * sizeof(float), n->stream));
    n->adam_t = 0;
    pack_all(n);
    return sync(n);
}

int dsmc_net_encode(dsmc_net* n, const dsmc_params* cases, int ncases, float* inputs)
{
    if (int rc = check_net(n)) return rc;
    if (!cases || !inputs || ncases < 0) return fail(n, DSMC_E_ARGUMENT, "bad arguments");
    if (int rc = set_device(n)) return rc;
    const size_t per = (size_t)n->H * n->W * DSMC_NET_IN_CHANNELS;
    for (int b0 = 0; b0 < ncases; b0 += n->Bmax) {
        const int B = std::min(n->Bmax, ncases - b0);
        std::vector<NetCase> hc(B);
        for (int i = 0; i < B; ++i) hc[i] = net_case(cases[b0 + i]);
        NET_TRY(n, hipMemcpyAsync(n->cases_b, hc.data(), B * sizeof(NetCase),
                                  hipMemcpyHostToDevice, n->stream));
        launch_encode(n->cases_b, nullptr, B, n->H, n->W, n->in0, n->stream);
        launch_from_bf16(n->in0, per * B, n->stage, n->stream);
        NET_TRY(n, hipMemcpyAsync(inputs + per * b0, n->stage, per * B * sizeof(float),
                                  hipMemcpyDeviceToHost, n->stream));
        if (int rc = sync(n)) return rc;
    }
    return 0;
}

int dsmc_net_normalize(const dsmc_net_config* cfg, const dsmc_params* cases, int n,
                       const float* fields, float* targets, float* weights)
{
    if (!cfg || check_config(*cfg) || !cases || !fields || !targets || !weights || n < 0)
        return DSMC_E_ARGUMENT;
    const size_t HW = (size_t)cfg->height * cfg->width;
    for (int i = 0; i < n; ++i)
        normalize_case(*cfg, cases[i], fields + 4 * HW * i, targets + 4 * HW * i,
                       weights + HW * i);
    return 0;
}

int dsmc_net_forward_raw(dsmc_net* n, const float* inputs, int B, float* outputs)
{
    if (int rc = check_net(n)) return rc;
    if (!inputs || !outputs || B < 1 || B > n->Bmax) return fail(n, DSMC_E_ARGUMENT, "bad batch");
    if (int rc = set_device(n)) return rc;
    if (int rc = upload_inputs(n, inputs, B)) return rc;
    forward(n, B);
    head(n, B, nullptr, nullptr, nullptr, n->out);
    NET_TRY(n, hipMemcpyAsync(outputs, n->out, n->pix(B, 0) * 4 * sizeof(float),
                              hipMemcpyDeviceToHost, n->stream));
    return sync(n);
}

int dsmc_net_get_activation(dsmc_net* n, int layer, float* out, size_t cap, size_t* count)
{
    if (int rc = check_net(n)) return rc;
    if (layer < 0 || layer >= 2 * kLayers + kPools) return fail(n, DSMC_E_ARGUMENT, "bad layer");
    if (!n->last_B) return fail(n, DSMC_E_ARGUMENT, "no forward pass yet");
    if (int rc = set_device(n)) return rc;
    const uint16_t* src;
    size_t cnt;
    if (layer < kLayers) {
        src = n->act[layer];
        cnt = n->pix(n->last_B, n->L[layer].level) * n->L[layer].cout;
    } else if (layer >= kLayers + kPools) {
        const int l = layer - kLayers - kPools;
        src = n->dz[l];
        cnt = n->pix(n->last_B, n->L[l].level) * n->L[l].cout;
    } else {
        const int k = layer - kLayers;
        src = n->pool[k];
        cnt = n->pix(n->last_B, k + 1) * ((size_t)n->C << k);
    }
    if (count) *count = cnt;
    if (!out) return 0;
    cnt = std::min(cnt, cap);
    // the staging buffer holds P0 * max(32, C) floats, at least any one activation
    launch_from_bf16(src, cnt, n->stage, n->stream);
    NET_TRY(n, hipMemcpyAsync(out, n->stage, cnt * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    return sync(n);
}

int dsmc_net_loss_grad_raw(dsmc_net* n, const float* inputs, const float* targets,
                           const float* weights, int B, float* loss, float* grad, size_t cnt)
{
    if (int rc = check_net(n)) return rc;
    if (!inputs || !targets || !weights || B < 1 || B > n->Bmax)
        return fail(n, DSMC_E_ARGUMENT, "bad batch");
    if (grad && cnt < n->nparams) return fail(n, DSMC_E_ARGUMENT, "gradient array too small");
    if (int rc = set_device(n)) return rc;
    if (int rc = upload_inputs(n, inputs, B)) return rc;
    const size_t P0 = n->pix(B, 0);
    NET_TRY(n, hipMemcpyAsync(n->tgt_b, targets, P0 * 4 * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    NET_TRY(n, hipMemcpyAsync(n->wgt_b, weights, P0 * sizeof(float), hipMemcpyHostToDevice,
                              n->stream));
    forward(n, B);
    const int blocks = head(n, B, nullptr, n->tgt_b, n->wgt_b, nullptr);
    launch_reduce(n->loss_part, blocks, 1, n->loss_acc, n->stream);
    backward(n, B);
    if (loss)
        NET_TRY(n, hipMemcpyAsync(loss, n->loss_acc, sizeof(float), hipMemcpyDeviceToHost,
                                  n->stream));
    if (grad)
        NET_TRY(n, hipMemcpyAsync(grad, n->g, n->nparams * sizeof(float), hipMemcpyDeviceToHost,
                                  n->stream));
    return sync(n);
}

int dsmc_net_set_dataset(dsmc_net* n, const dsmc_params* cases, const float* fields, int cnt)
{
    if (int rc = check_net(n)) return rc;
    if (!cases || !fields || cnt < 1) return fail(n, DSMC_E_ARGUMENT, "empty training set");
    if (int rc = set_device(n)) return rc;
    if (int rc = sync(n)) return rc;
    nfree(n, n->ds_cases);
    nfree(n, n->ds_tgt);
    nfree(n, n->ds_wgt);
    n->ds_cases = nullptr;
    n->ds_tgt = n->ds_wgt = nullptr;
    n->ds_n = 0;
    const size_t HW = (size_t)n->H * n->W;
    int rc = 0;
    if ((rc = nalloc(n, &n->ds_cases, cnt)) || (rc = nalloc(n, &n->ds_tgt, HW * 4 * cnt))
        || (rc = nalloc(n, &n->ds_wgt, HW * cnt)))
        return rc;
    std::vector<NetCase> hc(cnt);
    for (int i = 0; i < cnt; ++i) hc[i] = net_case(cases[i]);
    NET_TRY(n, hipMemcpyAsync(n->ds_cases, hc.data(), cnt * sizeof(NetCase), hipMemcpyHostToDevice,
                              n->stream));
    std::vector<float> t(HW * 4), w(HW);
    for (int i = 0; i < cnt; ++i) {
        normalize_case(n->cfg, cases[i], fields + 4 * HW * i, t.data(), w.data());
        NET_TRY(n, hipMemcpyAsync(n->ds_tgt + HW * 4 * i, t.data(), t.size() * sizeof(float),
                                  hipMemcpyHostToDevice, n->stream));
        NET_TRY(n, hipMemcpyAsync(n->ds_wgt + HW * i, w.data(), w.size() * sizeof(float),
                                  hipMemcpyHostToDevice, n->stream));
        if (int r2 = sync(n)) return r2;  // t and w are reused
    }
    n->ds_n = cnt;
    return 0;
}

int dsmc_net_train(dsmc_net* n, int epochs, int batch, float lr, uint64_t shuffle_seed,
                   float* loss_history, size_t history_cap)
{
    if (int rc = check_net(n)) return rc;
    if (!n->ds_n) return fail(n, DSMC_E_ARGUMENT, "no training set");
    if (epochs < 1 || batch < 1 || batch > n->Bmax || batch > n->ds_n || !(lr > 0.f))
        return fail(n, DSMC_E_ARGUMENT, "bad training arguments");
    if (int rc = set_device(n)) return rc;
    const int per_epoch = n->ds_n / batch;
    const size_t steps = (size_t)epochs * per_epoch;
    // every step's sample indices: one shuffle (Fisher-Yates) per epoch
    std::vector<int> idx(steps * batch);
    std::vector<int> perm(n->ds_n);
    SplitMix rng{shuffle_seed};
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < n->ds_n; ++i) perm[i] = i;
        for (int i = n->ds_n - 1; i > 0; --i) {
            const int j = (int)(rng.next() % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        std::memcpy(&idx[(size_t)e * per_epoch * batch], perm.data(),
                    sizeof(int) * (size_t)per_epoch * batch);
    }
    int rc = 0;
    if (idx.size() > n->d_idx_cap) {
        nfree(n, n->d_idx);
        if ((rc = nalloc(n, &n->d_idx, idx.size()))) return rc;
        n->d_idx_cap = idx.size();
    }
    if (steps > n->d_hist_cap) {
        nfree(n, n->d_hist);
        if ((rc = nalloc(n, &n->d_hist, steps))) return rc;
        n->d_hist_cap = steps;
    }
    NET_TRY(n, hipMemcpyAsync(n->d_idx, idx.data(), idx.size() * sizeof(int),
                              hipMemcpyHostToDevice, n->stream));
    if ((rc = sync(n))) return rc;
    hipEvent_t e0, e1;
    NET_TRY(n, hipEventCreate(&e0));
    NET_TRY(n, hipEventCreate(&e1));
    NET_TRY(n, hipEventRecord(e0, n->stream));
    const float b1 = 0.9f, b2 = 0.999f, lr_min = lr / 20.f;
    for (size_t t = 0; t < steps; ++t) {
        const int* bidx = n->d_idx + t * batch;
        launch_encode(n->ds_cases, bidx, batch, n->H, n->W, n->in0, n->stream);
        forward(n, batch);
        const int blocks = head(n, batch, bidx, n->ds_tgt, n->ds_wgt, nullptr);
        launch_reduce(n->loss_part, blocks, 1, n->d_hist + t, n->stream);
        backward(n, batch);
        ++n->adam_t;
        const double cosf = 0.5 * (1.0 + std::cos(3.14159265358979323846 * (double)t / (double)steps));
        const float lrt = lr_min + (lr - lr_min) * (float)cosf;
        const float bc1 = 1.f - (float)std::pow(0.9, (double)n->adam_t);
        const float bc2 = 1.f - (float)std::pow(0.999, (double)n->adam_t);
        launch_adam(n->w, n->g, n->m, n->v, n->nparams, lrt, b1, b2, 1e-8f, bc1, bc2, n->stream);
        pack_all(n);
    }
    NET_TRY(n, hipEventRecord(e1, n->stream));
    if ((rc = sync(n))) return rc;
    float ms = 0.f;
    NET_TRY(n, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    n->last_ms = ms;
    n->last_flops = conv_flops_per_step(n, batch) * (double)steps;
    if (loss_history) {
        const size_t k = std::min(history_cap, steps);
        NET_TRY(n, hipMemcpy(loss_history, n->d_hist, k * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

int dsmc_net_evaluate(dsmc_net* n, const dsmc_params* cases, const float* fields, int cnt,
                      float* loss)
{
    if (int rc = check_net(n)) return rc;
    if (!cases || !fields || cnt < 1 || !loss) return fail(n, DSMC_E_ARGUMENT, "bad arguments");
    if (int rc = set_device(n)) return rc;
    const size_t HW = (size_t)n->H * n->W;
    std::vector<float> t(HW * 4 * n->Bmax), w(HW * n->Bmax);
    std::vector<NetCase> hc(n->Bmax);
    double total = 0.0;
    for (int b0 = 0; b0 < cnt; b0 += n->Bmax) {
        const int B = std::min(n->Bmax, cnt - b0);
        for (int i = 0; i < B; ++i) {
            hc[i] = net_case(cases[b0 + i]);
            normalize_case(n->cfg, cases[b0 + i], fields + 4 * HW * (b0 + i), t.data() + 4 * HW * i,
                           w.data() + HW * i);
        }
        NET_TRY(n, hipMemcpyAsync(n->cases_b, hc.data(), B * sizeof(NetCase),
                                  hipMemcpyHostToDevice, n->stream));
        NET_TRY(n, hipMemcpyAsync(n->tgt_b, t.data(), HW * 4 * B * sizeof(float),
                                  hipMemcpyHostToDevice, n->stream));
        NET_TRY(n, hipMemcpyAsync(n->wgt_b, w.data(), HW * B * sizeof(float),
                                  hipMemcpyHostToDevice, n->stream));
        launch_encode(n->cases_b, nullptr, B, n->H, n->W, n->in0, n->stream);
        forward(n, B);
        const int blocks = head(n, B, nullptr, n->tgt_b, n->wgt_b, nullptr);
        launch_reduce(n->loss_part, blocks, 1, n->loss_acc, n->stream);
        float l = 0.f;
        NET_TRY(n, hipMemcpyAsync(&l, n->loss_acc, sizeof(float), hipMemcpyDeviceToHost, n->stream));
        if (int rc = sync(n)) return rc;
        total += (double)l * B;  // the head averages over the batch
    }
    *loss = (float)(total / cnt);
    return 0;
}

int dsmc_net_predict(dsmc_net* n, const dsmc_params* cases, int cnt, float* fields)
{
    if (int rc = check_net(n)) return rc;
    if (!cases || !fields || cnt < 0) return fail(n, DSMC_E_ARGUMENT, "bad arguments");
    if (int rc = set_device(n)) return rc;
    const int H = n->H, W = n->W;
    const size_t HW = (size_t)H * W;
    std::vector<float> o(HW * 4 * n->Bmax);
    std::vector<NetCase> hc(n->Bmax);
    for (int b0 = 0; b0 < cnt; b0 += n->Bmax) {
        const int B = std::min(n->Bmax, cnt - b0);
        for (int i = 0; i < B; ++i) hc[i] = net_case(cases[b0 + i]);
        NET_TRY(n, hipMemcpyAsync(n->cases_b, hc.data(), B * sizeof(NetCase),
                                  hipMemcpyHostToDevice, n->stream));
        launch_encode(n->cases_b, nullptr, B, H, W, n->in0, n->stream);
        forward(n, B);
        head(n, B, nullptr, nullptr, nullptr, n->out);
        NET_TRY(n, hipMemcpyAsync(o.data(), n->out, HW * 4 * B * sizeof(float),
                                  hipMemcpyDeviceToHost, n->stream));
        if (int rc = sync(n)) return rc;
        for (int i = 0; i < B; ++i) {
            const Freestream fs = freestream(cases[b0 + i]);
            float* dst = fields + 4 * HW * (b0 + i);
            for (int k = 0; k < H; ++k)
                for (int l = 0; l < W; ++l) {
                    const size_t p = (size_t)k * W + l;
                    const float* op = o.data() + 4 * (HW * i + p);
                    const bool in = raster_in_box(hc[i], k, l, H, W);
                    dst[p] = in ? 0.f : (float)(fs.n * std::exp((double)op[0]));
                    dst[HW + p] = (float)(fs.u * op[1]);
                    dst[2 * HW + p] = (float)(fs.u * op[2]);
                    dst[3 * HW + p] = (float)(fs.t * std::exp((double)op[3]));
                }
        }
    }
    return 0;
}

int dsmc_net_last_train_stats(dsmc_net* n, double* ms, double* flops)
{
    if (int rc = check_net(n)) return rc;
    if (ms) *ms = n->last_ms;
    if (flops) *flops = n->last_flops;
    return 0;
}

}  // extern "C"
