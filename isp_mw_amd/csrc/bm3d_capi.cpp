// C ABI of the MI355X BM3D denoiser: parameters, context, workspace and the
// orchestration of one or two BM3D steps on one GPU.
//
// Reference behaviour mirrored here:
//   parameter presets and thMSE_Default   include/BM3D.h:15-168
//   OPP-norm scaling of sigma and thMSE   include/BM3D.h:237-259
//   copy-through on degenerate parameters source/BM3D.cpp:82-87, 143-151, 284-290
//   step orchestration                    source/BM3D.cpp:464-527
#include "../../include/isp_bm3d.h"
#include "bm3d_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ispbm3d;

struct isp_bm3d_ctx {
    isp_bm3d_params params;
    int device = 0;
    int precision = ISP_BM3D_PREC_FP32;
    float *d_tables = nullptr;  // C8 (64) then the padded C_N matrices
    int cn_off[kMaxGroup + 1] = {0};
    double *d_tables64 = nullptr;  // C8 (64) then the unpadded C_N matrices, FP64
    int cnd_off[kMaxGroup + 1] = {0};
    uint16_t *d_mfma[2] = {nullptr, nullptr};  // MFMA operand tables, [0] FP16, [1] BF16
    void *ws = nullptr;
    size_t ws_bytes = 0;
    char *pin = nullptr;  // page-locked staging for host-memory calls
    size_t pin_bytes = 0;
    std::string err;
    // groups of the last run of each step (device, inside ws)
    int *last_cnt[2] = {nullptr, nullptr};
    int *last_pos[2] = {nullptr, nullptr};
    size_t last_refs[2] = {0, 0};
    int last_G[2] = {0, 0};
    // per-stage timing
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    struct Pending {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double stage_ms[ISP_BM3D_NUM_STAGES] = {0};
    long stage_calls[ISP_BM3D_NUM_STAGES] = {0};
};

namespace {

thread_local std::string g_create_err = "no error";

enum Stage {
    ST_OPP = 0,
    ST_FILL,
    ST_MATCH_BASIC,
    ST_FILTER_BASIC,
    ST_MATCH_FINAL,
    ST_FILTER_FINAL,
    ST_FINALIZE,
    ST_COPY
};
const char *const kStageNames[ISP_BM3D_NUM_STAGES] = {
    "opp", "fill", "match_basic", "filter_basic", "match_final", "filter_final", "finalize", "copy"};

hipEvent_t take_event(isp_bm3d_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets the launches of its scope with events when timing is enabled.
struct StageTimer {
    isp_bm3d_ctx *ctx;
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr;
    StageTimer(isp_bm3d_ctx *c, int st, hipStream_t stream) : ctx(c), stage(st), s(stream)
    {
        if (ctx->timing && (a = take_event(ctx))) (void)hipEventRecord(a, s);
    }
    ~StageTimer()
    {
        if (!a) return;
        hipEvent_t b = take_event(ctx);
        if (!b) {
            ctx->event_pool.push_back(a);
            return;
        }
        (void)hipEventRecord(b, s);
        ctx->pending.push_back({stage, a, b});
    }
};

int fail(isp_bm3d_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else g_create_err = buf;
    return code;
}

#define HIP_CHECK(ctx, expr)                                                                   \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), ISP_BM3D_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                   \
    } while (0)

bool known_profile(const char *p)
{
    return !strcmp(p, "fast") || !strcmp(p, "lc") || !strcmp(p, "np") || !strcmp(p, "high") ||
           !strcmp(p, "vn");
}

// OPP norms (include/BM3D.h:241-248 with include/Specification.h:391-407).
void opp_norms(double norm[3])
{
    const double Yr = (double)(1.0L / 3.0L), Yg = Yr, Yb = Yr;
    const double Ur = (double)(1.0L / 2.0L), Ug = 0.0, Ub = (double)(-1.0L / 2.0L);
    const double Vr = (double)(1.0L / 4.0L), Vg = (double)(-1.0L / 2.0L), Vb = (double)(1.0L / 4.0L);
    norm[0] = std::sqrt(Yr * Yr + Yg * Yg + Yb * Yb);
    norm[1] = std::sqrt(Ur * Ur + Ug * Ug + Ub * Ub);
    norm[2] = std::sqrt(Vr * Vr + Vg * Vg + Vb * Vb);
}

int next_pow2(int n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Parameters of one step after the constructor's scaling.
struct Derived {
    bool degenerate;  // copy src through (Kernel's early return)
    int range;
    bool bypass;
    double thSSE, distMul;
    float thr[3], sig2[3];
    double thr64[3];
    float band[3];
    int active[3];
    int ngx, ngy;
};

int derive(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, int nplanes, int W, int H, Derived &d)
{
    double norm[3];
    opp_norms(norm);
    const double thMSE = p.thMSE * norm[0];
    bool nosigma = true;
    for (int c = 0; c < nplanes; ++c) nosigma = nosigma && !(p.sigma[c] > 0);
    d.degenerate = nosigma || p.group_size == 0 || p.block_size <= 0 || p.bm_range <= 0 ||
                   p.bm_range < p.bm_step || thMSE <= 0;
    if (d.degenerate) return ISP_BM3D_OK;
    if (p.block_size != kBlock)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "BlockSize %d: the GPU kernels implement BlockSize 8",
                    p.block_size);
    if (p.group_size < 0 || p.group_size > kMaxGroup)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "GroupSize %d outside 1..%d", p.group_size, kMaxGroup);
    if (p.block_step <= 0 || p.bm_step <= 0)
        return fail(ctx, ISP_BM3D_E_ARG, "BlockStep and BMstep must be positive");
    if (W < kBlock || H < kBlock)
        return fail(ctx, ISP_BM3D_E_SIZE, "image %dx%d is smaller than one %dx%d block", W, H, kBlock, kBlock);
    if (W > 65535 || H > 65535)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "image %dx%d: width and height must be < 65536", W, H);
    d.range = p.bm_range / p.bm_step * p.bm_step;
    d.bypass = p.group_size == 1 || thMSE <= 0;
    // include/Block.h:582-584 with range 1 (Plane_FL default Floor 0 / Ceil 1)
    const double MSE2SSE = (double)(kBlock * kBlock) * 1.0 * 1.0 / (double)(255 * 255);
    d.distMul = 1.0 / MSE2SSE;
    d.thSSE = thMSE * MSE2SSE;
    for (int c = 0; c < 3; ++c) {
        d.active[c] = c < nplanes && p.sigma[c] > 0;
        const double sigma = p.sigma[c] / 255.0 * norm[c];
        d.thr[c] = (float)(sigma * p.lambda);
        d.sig2[c] = (float)(sigma * sigma);
        d.thr64[c] = sigma * p.lambda;
        // FP32 coefficients of [0,1] data carry absolute errors of a few 1e-6;
        // decisions closer than this to the threshold are re-made in FP64.
        d.band[c] = (float)(3e-5 + 1e-4 * d.thr64[c]);
    }
    d.ngx = (W - kBlock) / p.block_step + 1 + ((W - kBlock) % p.block_step ? 1 : 0);
    d.ngy = (H - kBlock) / p.block_step + 1 + ((H - kBlock) % p.block_step ? 1 : 0);
    return ISP_BM3D_OK;
}

// Workspace carving: a bump allocator over one device allocation.
struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base((char *)b) {}
    template <typename T>
    T *take(size_t n)
    {
        off = (off + 255) & ~size_t(255);
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += n * sizeof(T);
        return p;
    }
};

int ensure_ws(isp_bm3d_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->ws_bytes) return ISP_BM3D_OK;
    if (ctx->ws) {
        HIP_CHECK(ctx, hipDeviceSynchronize());
        HIP_CHECK(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
        for (int k = 0; k < 2; ++k) {
            ctx->last_cnt[k] = ctx->last_pos[k] = nullptr;
            ctx->last_refs[k] = 0;
        }
    }
    bytes = (bytes + (1u << 20)) & ~size_t((1u << 20) - 1);
    HIP_CHECK(ctx, hipMalloc(&ctx->ws, bytes));
    ctx->ws_bytes = bytes;
    return ISP_BM3D_OK;
}

// Host-memory calls stage each plane through page-locked memory: the CPU packs
// rows at the device pitch, one DMA per plane runs on the caller's stream, and
// the call synchronises that stream before it returns (so no DMA from an
// earlier call is in flight when the next one writes the staging area).
struct HostStage {
    isp_bm3d_ctx *ctx;
    hipStream_t s;
    int P, W, H;
    size_t plane_bytes() const { return (size_t)P * H * 4; }
    char *slot(int k) const { return ctx->pin + (size_t)k * plane_bytes(); }

    static int reserve(isp_bm3d_ctx *ctx, size_t bytes)
    {
        if (bytes <= ctx->pin_bytes) return ISP_BM3D_OK;
        if (ctx->pin) {
            HIP_CHECK(ctx, hipHostFree(ctx->pin));
            ctx->pin = nullptr;
            ctx->pin_bytes = 0;
        }
        HIP_CHECK(ctx, hipHostMalloc((void **)&ctx->pin, bytes, hipHostMallocDefault));
        ctx->pin_bytes = bytes;
        return ISP_BM3D_OK;
    }
    // host plane (row stride `stride` elements of 4 bytes) -> device plane of pitch P
    int upload(int k, void *dev, const void *host, size_t stride) const
    {
        char *p = slot(k);
        for (int y = 0; y < H; ++y)
            std::memcpy(p + (size_t)y * P * 4, (const char *)host + (size_t)y * stride * 4, (size_t)W * 4);
        HIP_CHECK(ctx, hipMemcpyAsync(dev, p, plane_bytes(), hipMemcpyHostToDevice, s));
        return ISP_BM3D_OK;
    }
    int download_start(int k, const void *dev) const
    {
        HIP_CHECK(ctx, hipMemcpyAsync(slot(k), dev, plane_bytes(), hipMemcpyDeviceToHost, s));
        return ISP_BM3D_OK;
    }
    // after hipStreamSynchronize(s)
    void download_finish(int k, void *host, size_t stride) const
    {
        const char *p = slot(k);
        for (int y = 0; y < H; ++y)
            std::memcpy((char *)host + (size_t)y * stride * 4, p + (size_t)y * P * 4, (size_t)W * 4);
    }
};

void host_copy_plane(void *dst, const void *src, size_t stride, int W, int H)
{
    if (dst == src) return;
    for (int y = 0; y < H; ++y)
        std::memcpy((char *)dst + (size_t)y * stride * 4, (const char *)src + (size_t)y * stride * 4, (size_t)W * 4);
}

// Per-call device buffers.
struct Work {
    float *plane[3];   // src YUV (RGB path) or uploaded src (host plane path)
    float *refp[3];    // uploaded/converted guide planes
    float *basic[3];   // step-1 estimate
    unsigned long long *num[3], *den[3];  // fixed-point accumulators
    float *out[3];     // host-mode output staging
    double *y64;       // FP64 matching guide (RGB path)
    double *ref64;     // FP64 matching guide of a separate ref frame
    int *cnt[2], *pos[2];  // per step
    double *keys;
};

size_t layout(Work *w, void *base, int P, int H, size_t refs, bool keys)
{
    Carver c(base);
    const size_t np = (size_t)P * H;
    for (int k = 0; k < 3; ++k) w->plane[k] = c.take<float>(np);
    for (int k = 0; k < 3; ++k) w->refp[k] = c.take<float>(np);
    for (int k = 0; k < 3; ++k) w->basic[k] = c.take<float>(np);
    for (int k = 0; k < 3; ++k) w->num[k] = c.take<unsigned long long>(np);
    for (int k = 0; k < 3; ++k) w->den[k] = c.take<unsigned long long>(np);
    for (int k = 0; k < 3; ++k) w->out[k] = c.take<float>(np);
    w->y64 = c.take<double>(np);
    w->ref64 = c.take<double>(np);
    for (int k = 0; k < 2; ++k) {
        w->cnt[k] = c.take<int>(refs);
        w->pos[k] = c.take<int>(refs * kMaxGroup);
    }
    w->keys = keys ? c.take<double>(refs * kMaxGroup) : nullptr;
    return c.off + 256;
}

int prepare(isp_bm3d_ctx *ctx, Work &w, int P, int H, int W, bool keys)
{
    // grid size bound over both steps
    size_t refs = 1;
    const isp_bm3d_step_params *ps[2] = {&ctx->params.basic, &ctx->params.final_};
    for (auto *p : ps) {
        const int s = p->block_step > 0 ? p->block_step : 1;
        const size_t gx = (size_t)(W / s + 2), gy = (size_t)(H / s + 2);
        refs = std::max(refs, gx * gy);
    }
    const size_t bytes = layout(&w, nullptr, P, H, refs, keys);
    int rc = ensure_ws(ctx, bytes);
    if (rc) return rc;
    layout(&w, ctx->ws, P, H, refs, keys);
    return ISP_BM3D_OK;
}

struct StepIO {
    int nplanes;
    const float *src[3];
    const float *guide[3];     // Wiener guides (final step)
    const float *mguide32;     // matching guide, FP32 ...
    const double *mguide64;    // ... or FP64
    int pitch;                 // of src, guide, mguide32 and the accumulators
    int pitch64;
    // finalize: output format (see FinalizeArgs)
    int to_rgb, out_int, clip;
    double gain_y, gain_c, offset, q_gain, lo, hi;
    void *out[3];
    size_t out_pitch;
};

int run_match(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const Derived &d, Work &w,
              const float *g32, const double *g64, int gpitch, int W, int H, double *keys,
              hipStream_t s, int stage)
{
    MatchArgs a{};
    a.W = W;
    a.H = H;
    a.ngx = d.ngx;
    a.ngy = d.ngy;
    a.bstep = p.block_step;
    a.range = d.range;
    a.step = p.bm_step;
    a.G = p.group_size;
    a.bypass = d.bypass;
    a.thSSE = d.thSSE;
    a.distMul = d.distMul;
    // the plain FP32 path reproduces the reference's tie order bit for bit;
    // the mixed-precision forms make no parity claim and use fused FMA
    a.exact = ctx->precision == ISP_BM3D_PREC_FP32;
    a.tileW = (kTileRX - 1) * p.block_step + 2 * d.range + kBlock + (kStrip - 1) * p.bm_step;
    a.tileH = (kTileRY - 1) * p.block_step + 2 * d.range + kBlock;
    a.tilePitch = a.tileW | 1;
    const int side = 2 * (d.range / p.bm_step) + 1;
    a.maxCand = side * side;
    const int k = stage == ST_MATCH_FINAL ? 1 : 0;
    a.cnt_out = w.cnt[k];
    a.pos_out = w.pos[k];
    ctx->last_cnt[k] = w.cnt[k];
    ctx->last_pos[k] = w.pos[k];
    ctx->last_refs[k] = (size_t)d.ngx * d.ngy;
    ctx->last_G[k] = p.group_size;
    a.key_out = keys;
    if (match_lds_bytes(a) > 160 * 1024)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED,
                    "BMrange %d / BMstep %d / BlockStep %d need %zu bytes of LDS for matching (max 163840)",
                    p.bm_range, p.bm_step, p.block_step, match_lds_bytes(a));
    StageTimer tm(ctx, stage, s);
    if (g64) HIP_CHECK(ctx, launch_match_f64(g64, gpitch, a, s));
    else HIP_CHECK(ctx, launch_match_f32(g32, gpitch, a, s));
    return ISP_BM3D_OK;
}

// One BM3D_Base::Kernel (source/BM3D.cpp:80-218) followed by finalize.
int run_step(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, bool wiener, const StepIO &io, Work &w,
             int W, int H, hipStream_t s)
{
    Derived d{};
    int rc = derive(ctx, p, io.nplanes, W, H, d);
    if (rc) return rc;
    FinalizeArgs fa{};
    fa.W = W;
    fa.H = H;
    fa.pitch = io.pitch;
    fa.nplanes = io.nplanes;
    fa.to_rgb = io.to_rgb;
    fa.out_int = io.out_int;
    fa.clip = io.clip;
    fa.gain_y = io.gain_y;
    fa.gain_c = io.gain_c;
    fa.offset = io.offset;
    fa.q_gain = io.q_gain;
    fa.lo = io.lo;
    fa.hi = io.hi;
    fa.out_pitch = io.out_pitch;
    for (int c = 0; c < 3; ++c) {
        fa.src[c] = io.src[c];
        fa.out[c] = io.out[c];
        fa.num[c] = w.num[c];
        fa.den[c] = w.den[c];
        fa.active[c] = 0;
    }
    if (!d.degenerate) {
        rc = run_match(ctx, p, d, w, io.mguide32, io.mguide64, io.mguide64 ? io.pitch64 : io.pitch, W, H,
                       nullptr, s, wiener ? ST_MATCH_FINAL : ST_MATCH_BASIC);
        if (rc) return rc;
        FilterArgs f{};
        f.W = W;
        f.H = H;
        f.pitch = io.pitch;
        f.ngx = d.ngx;
        f.ngy = d.ngy;
        f.bstep = p.block_step;
        f.range = d.range;
        f.G = p.group_size;
        f.cnt = w.cnt[wiener ? 1 : 0];
        f.pos = w.pos[wiener ? 1 : 0];
        for (int c = 0; c < 3; ++c) {
            f.src[c] = io.src[c];
            f.guide[c] = io.guide[c];
            f.num[c] = w.num[c];
            f.den[c] = w.den[c];
            f.active[c] = d.active[c];
            f.thr[c] = d.thr[c];
            f.sig2[c] = d.sig2[c];
            fa.active[c] = d.active[c];
            if (d.active[c]) {
                StageTimer tm(ctx, ST_FILL, s);
                // zero bits of the 64-bit accumulators
                HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(w.num[c]), (size_t)io.pitch * H * 2, 0.f, s));
                HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(w.den[c]), (size_t)io.pitch * H * 2, 0.f, s));
            }
        }
        const bool mfma = ctx->precision != ISP_BM3D_PREC_FP32;
        const int trx = mfma ? kMTileRX : kTileRX, try_ = mfma ? kMTileRY : kTileRY;
        f.FW = (trx - 1) * p.block_step + 2 * d.range + kBlock;
        f.FH = (try_ - 1) * p.block_step + 2 * d.range + kBlock;
        f.c8 = ctx->d_tables;
        f.cn = ctx->d_tables;
        for (int n = 0; n <= kMaxGroup; ++n) {
            f.cn_off[n] = ctx->cn_off[n];
            f.cnd_off[n] = ctx->cnd_off[n];
        }
        f.c8d = ctx->d_tables64;
        f.cnd = ctx->d_tables64;
        for (int c = 0; c < 3; ++c) {
            f.thr64[c] = d.thr64[c];
            f.band[c] = d.band[c];
        }
        if (mfma) {
            const uint16_t *t = ctx->d_mfma[ctx->precision == ISP_BM3D_PREC_FP16 ? 0 : 1];
            f.kb = t;
            f.ka = t + 8 * 64 * 8;
            f.cna = t + 16 * 64 * 8;
            f.cnb = f.cna + (size_t)kMaxGroup * 2 * 64 * 8;
        }
        const size_t lds = mfma ? filter_mfma_lds_bytes(f, wiener) : filter_lds_bytes(f, wiener);
        if (lds > 160 * 1024)
            return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "parameters need %zu bytes of LDS for filtering", lds);
        StageTimer tm(ctx, wiener ? ST_FILTER_FINAL : ST_FILTER_BASIC, s);
        if (mfma) HIP_CHECK(ctx, launch_filter_mfma(f, wiener, ctx->precision, s));
        else HIP_CHECK(ctx, launch_filter(f, wiener, s));
    }
    StageTimer tm(ctx, ST_FINALIZE, s);
    HIP_CHECK(ctx, launch_finalize(fa, s));
    return ISP_BM3D_OK;
}

int check_common(isp_bm3d_ctx *ctx, int W, int H, size_t stride, int mode, int mem)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (mode < ISP_BM3D_BASIC || mode > ISP_BM3D_BOTH) return fail(ctx, ISP_BM3D_E_ARG, "bad mode %d", mode);
    if (mem != ISP_BM3D_MEM_HOST && mem != ISP_BM3D_MEM_DEVICE)
        return fail(ctx, ISP_BM3D_E_ARG, "bad mem %d", mem);
    if (W <= 0 || H <= 0) return fail(ctx, ISP_BM3D_E_SIZE, "empty image %dx%d", W, H);
    if (stride < (size_t)W) return fail(ctx, ISP_BM3D_E_ARG, "stride %zu < width %d", stride, W);
    if (stride > (size_t)1 << 30) return fail(ctx, ISP_BM3D_E_ARG, "stride %zu too large", stride);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return ISP_BM3D_OK;
}

int internal_pitch(int W, size_t stride, int mem)
{
    return mem == ISP_BM3D_MEM_DEVICE ? (int)stride : ((W + 63) / 64) * 64;
}

void set_default_step(isp_bm3d_step_params *p, const char *profile, int final_step)
{
    std::memset(p, 0, sizeof(*p));
    std::strncpy(p->profile, profile, sizeof(p->profile) - 1);
    p->sigma[0] = p->sigma[1] = p->sigma[2] = 10.0;
    p->block_size = 8;
    p->bm_range = 16;
    p->bm_step = 1;
    const std::string pr = profile;
    if (pr == "fast" || pr == "lc") p->bm_range = 9;
    if (!final_step) {
        p->block_step = 4;
        p->group_size = 16;
        p->lambda = 2.7;
        if (pr == "fast") p->block_step = 8;
        else if (pr == "lc") p->block_step = 6;
        else if (pr == "high") p->block_step = 3;
        else if (pr == "vn") {
            p->block_step = 4;
            p->group_size = 32;
            p->lambda = 2.8;
        }
    } else {
        p->block_step = 3;
        p->group_size = 32;
        // BM3D_Final_Para leaves lambda unset; the final step never reads it.
        p->lambda = 2.7;
        if (pr == "fast") {
            p->block_step = 7;
            p->group_size = 16;
        } else if (pr == "lc") {
            p->block_step = 5;
            p->group_size = 16;
        } else if (pr == "high") {
            p->block_step = 2;
        } else if (pr == "vn") {
            p->block_size = 11;
            p->block_step = 6;
        }
    }
}

uint16_t to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

uint16_t to_f16(float f)
{
    _Float16 hv = (_Float16)f;
    uint16_t u;
    std::memcpy(&u, &hv, 2);
    return u;
}

// MFMA operand tables for filter_mfma_kernel (layouts in bm3d_internal.h).
// h = lane / 32, c = lane % 32; a 16-long k chunk m of an accumulator-fed
// operand is ordered 8*(e/4) + 4h + e%4 to match the accumulator rows.
std::vector<float> mfma_tables_f32(size_t *n_kb, size_t *n_ka, size_t *n_cn)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    auto c8 = [&](int k, int n) {
        return (double)((k == 0 ? std::sqrt(1.0L / 8) : std::sqrt(2.0L / 8)) * std::cos(PI * (2 * n + 1) * k / 16.0L));
    };
    auto k64 = [&](int p, int q) { return c8(q >> 3, p >> 3) * c8(q & 7, p & 7); };
    auto cn = [&](int N, int k, int z) -> double {
        if (k >= N || z >= N) return 0.0;
        return (double)((k == 0 ? std::sqrt(1.0L / N) : std::sqrt(2.0L / N)) *
                        std::cos(PI * (2 * z + 1) * k / (2.0L * N)));
    };
    auto perm = [](int m, int h, int e) { return 16 * m + 8 * (e >> 2) + 4 * h + (e & 3); };
    std::vector<float> t;
    for (int s = 0; s < 4; ++s)
        for (int q = 0; q < 2; ++q)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) t.push_back((float)k64(16 * s + 8 * (l >> 5) + e, 32 * q + (l & 31)));
    *n_kb = t.size();
    for (int r = 0; r < 2; ++r)
        for (int s = 0; s < 4; ++s)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) t.push_back((float)k64(32 * r + (l & 31), perm(s, l >> 5, e)));
    *n_ka = t.size() - *n_kb;
    for (int N = 1; N <= kMaxGroup; ++N)
        for (int m = 0; m < 2; ++m)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) t.push_back((float)cn(N, l & 31, perm(m, l >> 5, e)));
    for (int N = 1; N <= kMaxGroup; ++N)
        for (int m = 0; m < 2; ++m)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) t.push_back((float)cn(N, perm(m, l >> 5, e), l & 31));
    *n_cn = (t.size() - *n_kb - *n_ka) / 2;
    return t;
}

}  // namespace

extern "C" {

int isp_bm3d_thmse_default(isp_bm3d_step_params *p, int final_step)
{
    if (!p) return fail(nullptr, ISP_BM3D_E_ARG, "null params");
    const bool vn = !std::strcmp(p->profile, "vn");
    if (!final_step) p->thMSE = vn ? 1000 + p->sigma[0] * 150 : 400 + p->sigma[0] * 80;
    else p->thMSE = vn ? 400 + p->sigma[0] * 40 : 200 + p->sigma[0] * 10;
    return ISP_BM3D_OK;
}

int isp_bm3d_step_params_default(isp_bm3d_step_params *p, const char *profile, int final_step)
{
    if (!p) return fail(nullptr, ISP_BM3D_E_ARG, "null params");
    if (!profile) profile = "fast";
    if (!known_profile(profile))
        return fail(nullptr, ISP_BM3D_E_PROFILE, "unrecognized profile \"%s\", should be \"fast\", \"lc\", "
                    "\"np\", \"vn\" or \"high\"", profile);
    set_default_step(p, profile, final_step);
    return isp_bm3d_thmse_default(p, final_step);
}

int isp_bm3d_params_default(isp_bm3d_params *params, const char *profile)
{
    if (!params) return fail(nullptr, ISP_BM3D_E_ARG, "null params");
    int rc = isp_bm3d_step_params_default(&params->basic, profile, 0);
    if (rc) return rc;
    return isp_bm3d_step_params_default(&params->final_, profile, 1);
}

const char *isp_bm3d_last_error(const isp_bm3d_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int isp_bm3d_create(isp_bm3d_ctx **out, const isp_bm3d_params *params, int device)
{
    if (!out || !params) return fail(nullptr, ISP_BM3D_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(nullptr, ISP_BM3D_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(nullptr, ISP_BM3D_E_ARG, "device %d of %d", device, ndev);
    isp_bm3d_ctx *ctx = new isp_bm3d_ctx();
    ctx->params = *params;
    ctx->device = device;
    ctx->err = "no error";
    // Orthonormal DCT-II matrices: C8 [k][n], then C_N for N = 1..32, each
    // padded to NP x NP (NP = next power of two) with zeros.
    std::vector<float> tab(64);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < 8; ++k)
        for (int n = 0; n < 8; ++n)
            tab[k * 8 + n] = (float)((k == 0 ? std::sqrt(1.0L / 8) : std::sqrt(2.0L / 8)) *
                                     std::cos(PI * (2 * n + 1) * k / 16.0L));
    ctx->cn_off[0] = 0;
    for (int N = 1; N <= kMaxGroup; ++N) {
        const int NP = next_pow2(N);
        ctx->cn_off[N] = (int)tab.size();
        const size_t base = tab.size();
        tab.resize(base + (size_t)NP * NP, 0.f);
        for (int k = 0; k < N; ++k)
            for (int z = 0; z < N; ++z)
                tab[base + (size_t)k * NP + z] =
                    (float)((k == 0 ? std::sqrt(1.0L / N) : std::sqrt(2.0L / N)) *
                            std::cos(PI * (2 * z + 1) * k / (2.0L * N)));
    }
    std::vector<double> tab64(64);
    for (int k = 0; k < 8; ++k)
        for (int n = 0; n < 8; ++n)
            tab64[k * 8 + n] = (double)((k == 0 ? std::sqrt(1.0L / 8) : std::sqrt(2.0L / 8)) *
                                        std::cos(PI * (2 * n + 1) * k / 16.0L));
    for (int N = 1; N <= kMaxGroup; ++N) {
        ctx->cnd_off[N] = (int)tab64.size();
        for (int k = 0; k < N; ++k)
            for (int z = 0; z < N; ++z)
                tab64.push_back((double)((k == 0 ? std::sqrt(1.0L / N) : std::sqrt(2.0L / N)) *
                                         std::cos(PI * (2 * z + 1) * k / (2.0L * N))));
    }
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_tables, tab.size() * sizeof(float))) != hipSuccess ||
        (e = hipMemcpy(ctx->d_tables, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMalloc(&ctx->d_tables64, tab64.size() * sizeof(double))) != hipSuccess ||
        (e = hipMemcpy(ctx->d_tables64, tab64.data(), tab64.size() * sizeof(double), hipMemcpyHostToDevice)) !=
            hipSuccess) {
        int rc = fail(nullptr, ISP_BM3D_E_HIP, "create: %s", hipGetErrorString(e));
        if (ctx->d_tables) (void)hipFree(ctx->d_tables);
        if (ctx->d_tables64) (void)hipFree(ctx->d_tables64);
        delete ctx;
        return rc;
    }
    size_t n_kb, n_ka, n_cn;
    std::vector<float> mt = mfma_tables_f32(&n_kb, &n_ka, &n_cn);
    for (int k = 0; k < 2; ++k) {
        std::vector<uint16_t> bits(mt.size());
        for (size_t i = 0; i < mt.size(); ++i) bits[i] = k == 0 ? to_f16(mt[i]) : to_bf16(mt[i]);
        if ((e = hipMalloc(&ctx->d_mfma[k], bits.size() * 2)) != hipSuccess ||
            (e = hipMemcpy(ctx->d_mfma[k], bits.data(), bits.size() * 2, hipMemcpyHostToDevice)) != hipSuccess)
            break;
    }
    if (e != hipSuccess) {
        int rc = fail(nullptr, ISP_BM3D_E_HIP, "create: %s", hipGetErrorString(e));
        (void)hipFree(ctx->d_tables);
        (void)hipFree(ctx->d_tables64);
        for (int k = 0; k < 2; ++k)
            if (ctx->d_mfma[k]) (void)hipFree(ctx->d_mfma[k]);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return ISP_BM3D_OK;
}

int isp_bm3d_set_timing(isp_bm3d_ctx *ctx, int enable)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    ctx->timing = enable != 0;
    return ISP_BM3D_OK;
}

int isp_bm3d_get_timing(isp_bm3d_ctx *ctx, const char **names, double *total_ms, long *launches, int max)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        HIP_CHECK(ctx, hipEventSynchronize(p.b));
        HIP_CHECK(ctx, hipEventElapsedTime(&ms, p.a, p.b));
        ctx->stage_ms[p.stage] += ms;
        ctx->stage_calls[p.stage] += 1;
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
    for (int i = 0; i < max && i < ISP_BM3D_NUM_STAGES; ++i) {
        if (names) names[i] = kStageNames[i];
        if (total_ms) total_ms[i] = ctx->stage_ms[i];
        if (launches) launches[i] = ctx->stage_calls[i];
    }
    return ISP_BM3D_NUM_STAGES;
}

int isp_bm3d_reset_timing(isp_bm3d_ctx *ctx)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    const char *n[ISP_BM3D_NUM_STAGES];
    int rc = isp_bm3d_get_timing(ctx, n, nullptr, nullptr, 0);
    if (rc < 0) return rc;
    for (int i = 0; i < ISP_BM3D_NUM_STAGES; ++i) {
        ctx->stage_ms[i] = 0;
        ctx->stage_calls[i] = 0;
    }
    return ISP_BM3D_OK;
}

int isp_bm3d_destroy(isp_bm3d_ctx *ctx)
{
    if (!ctx) return ISP_BM3D_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->d_tables) (void)hipFree(ctx->d_tables);
    if (ctx->d_tables64) (void)hipFree(ctx->d_tables64);
    for (int k = 0; k < 2; ++k)
        if (ctx->d_mfma[k]) (void)hipFree(ctx->d_mfma[k]);
    delete ctx;
    return ISP_BM3D_OK;
}

int isp_bm3d_set_precision(isp_bm3d_ctx *ctx, int precision)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (precision < ISP_BM3D_PREC_FP32 || precision > ISP_BM3D_PREC_BF16)
        return fail(ctx, ISP_BM3D_E_ARG, "bad precision %d", precision);
    ctx->precision = precision;
    return ISP_BM3D_OK;
}

int isp_bm3d_last_groups(isp_bm3d_ctx *ctx, int final_step, int *counts, int *pos, size_t capacity)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    const int k = final_step ? 1 : 0;
    if (!ctx->last_cnt[k]) return fail(ctx, ISP_BM3D_E_ARG, "no groups recorded for this step yet");
    const size_t refs = ctx->last_refs[k];
    if (capacity < refs) return fail(ctx, ISP_BM3D_E_ARG, "capacity %zu < %zu grid points", capacity, refs);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    HIP_CHECK(ctx, hipDeviceSynchronize());
    if (counts) HIP_CHECK(ctx, hipMemcpy(counts, ctx->last_cnt[k], refs * sizeof(int), hipMemcpyDeviceToHost));
    if (pos)
        HIP_CHECK(ctx, hipMemcpy(pos, ctx->last_pos[k], refs * ctx->last_G[k] * sizeof(int), hipMemcpyDeviceToHost));
    return (int)refs;
}

int isp_bm3d_grid_size(const isp_bm3d_ctx *ctx, int W, int H, int final_step, int *ngx, int *ngy)
{
    if (!ctx || !ngx || !ngy) return ISP_BM3D_E_ARG;
    const isp_bm3d_step_params &p = final_step ? ctx->params.final_ : ctx->params.basic;
    if (p.block_step <= 0 || p.block_size <= 0 || W < p.block_size || H < p.block_size) return ISP_BM3D_E_ARG;
    const int lx = W - p.block_size, ly = H - p.block_size;
    *ngx = lx / p.block_step + 1 + (lx % p.block_step ? 1 : 0);
    *ngy = ly / p.block_step + 1 + (ly % p.block_step ? 1 : 0);
    return ISP_BM3D_OK;
}

}  // extern "C"

namespace {

// Sample format of the caller's planes: float (Plane_FL: luma [0,1]) or
// int32 with the quantisation of a Plane (include/Image_Type.hpp:85-117).
struct Fmt {
    bool is_int;
    int floor_, neutral, ceil_;
    bool chroma() const { return floor_ < neutral; }                                      // isChroma
    bool pc_chroma() const { return floor_ < neutral && (floor_ + ceil_) % 2 == 1; }       // isPCChroma
};

void set_plain_out(StepIO &io, void *out, size_t pitch)
{
    io.to_rgb = 0;
    io.out_int = 0;
    io.clip = 0;
    io.out[0] = out;
    io.out_pitch = pitch;
}

// RangeConvert float -> int with clip (include/Conversion.hpp:241-256)
void set_int_plane_out(StepIO &io, const Fmt &f, void *out, size_t pitch)
{
    io.to_rgb = 0;
    io.out_int = 1;
    io.q_gain = (double)(f.ceil_ - f.floor_) / 1.0;
    io.offset = -0.0 * io.q_gain + (double)(f.chroma() ? f.neutral : f.floor_) + (f.pc_chroma() ? 0.499999 : 0.5);
    io.clip = 1;
    io.lo = f.floor_;
    io.hi = f.ceil_;
    io.out[0] = out;
    io.out_pitch = pitch;
}

int copy_plane(isp_bm3d_ctx *ctx, void *dst, size_t dpitch, const void *src, size_t spitch, int W, int H,
               hipStream_t s)
{
    StageTimer tm(ctx, ST_COPY, s);
    HIP_CHECK(ctx, launch_copy2d((float *)dst, dpitch, (const float *)src, spitch, W, H, s));
    return ISP_BM3D_OK;
}

// One plane: Plane_FL (source/BM3D.cpp:80-136, 239-245, 464-472) or, for int
// samples, Plane (RangeConvert in and out, :248-279, 475-498).
int plane_common(isp_bm3d_ctx *ctx, void *dst, const void *src, const void *ref, int W, int H, size_t stride,
                 const Fmt &fmt, int mode, int mem, void *stream)
{
    int rc = check_common(ctx, W, H, stride, mode, mem);
    if (rc) return rc;
    if (!dst || !src) return fail(ctx, ISP_BM3D_E_ARG, "null dst/src");
    if (fmt.is_int && !(fmt.floor_ < fmt.ceil_ && fmt.floor_ <= fmt.neutral && fmt.neutral <= fmt.ceil_))
        return fail(ctx, ISP_BM3D_E_ARG, "bad quantisation floor %d neutral %d ceil %d", fmt.floor_, fmt.neutral,
                    fmt.ceil_);
    hipStream_t s = (hipStream_t)stream;
    const bool host = mem == ISP_BM3D_MEM_HOST;
    const int P = internal_pitch(W, stride, mem);
    Work w{};
    if ((rc = prepare(ctx, w, P, H, W, false))) return rc;
    const bool sep_ref = ref && ref != src;
    // BM3D_Base::process_Plane's early return for the single steps on int planes
    if (fmt.is_int && mode != ISP_BM3D_BOTH) {
        const isp_bm3d_step_params &p = mode == ISP_BM3D_BASIC ? ctx->params.basic : ctx->params.final_;
        double norm[3];
        opp_norms(norm);
        if (!(p.sigma[0] > 0) || p.group_size == 1 || p.block_size <= 0 || p.bm_range <= 0 ||
            p.bm_range < p.bm_step || p.thMSE * norm[0] <= 0) {
            if (host) {
                host_copy_plane(dst, src, stride, W, H);
                return ISP_BM3D_OK;
            }
            return dst == src ? ISP_BM3D_OK : copy_plane(ctx, dst, stride, src, stride, W, H, s);
        }
    }
    // source and guide as float planes of pitch P
    const void *rsrc = src, *rref = sep_ref ? ref : src;
    size_t rpitch = stride;
    const HostStage hs{ctx, s, P, W, H};
    if (host) {
        if ((rc = HostStage::reserve(ctx, 2 * hs.plane_bytes()))) return rc;
        float *stage_s = fmt.is_int ? w.out[1] : w.plane[0];
        float *stage_r = fmt.is_int ? w.out[2] : w.refp[0];
        if ((rc = hs.upload(0, stage_s, src, stride))) return rc;
        rsrc = stage_s;
        rref = stage_s;
        if (sep_ref) {
            if ((rc = hs.upload(1, stage_r, ref, stride))) return rc;
            rref = stage_r;
        }
        rpitch = (size_t)P;
    }
    const float *fsrc = (const float *)rsrc, *fref = (const float *)rref;
    if (fmt.is_int) {
        // RangeConvert int -> float, no clip except for PC-range chroma
        ToFloatArgs t{};
        t.W = W;
        t.H = H;
        t.spitch = rpitch;
        t.dpitch = P;
        t.gain = 1.0 / (double)(fmt.ceil_ - fmt.floor_);
        t.offset = -(double)(fmt.chroma() ? fmt.neutral : fmt.floor_) * t.gain + 0.0;
        t.clip = fmt.pc_chroma();
        t.lo = fmt.chroma() ? -0.5 : 0.0;
        t.hi = fmt.chroma() ? 0.5 : 1.0;
        {
            // the matching guide also in FP64, as the reference holds it
            StageTimer tm(ctx, ST_OPP, s);
            t.src = (const int *)rsrc;
            t.dst = w.plane[0];
            t.dst64 = sep_ref ? nullptr : w.y64;
            HIP_CHECK(ctx, launch_to_float(t, s));
            fsrc = w.plane[0];
            fref = fsrc;
            if (sep_ref) {
                t.src = (const int *)rref;
                t.dst = w.refp[0];
                t.dst64 = w.y64;
                HIP_CHECK(ctx, launch_to_float(t, s));
                fref = w.refp[0];
            }
        }
    }
    void *out = host ? (void *)w.out[0] : dst;
    const size_t out_pitch = host ? (size_t)P : stride;
    StepIO io{};
    io.nplanes = 1;
    io.pitch = P;
    io.src[0] = fsrc;
    auto final_out = [&](StepIO &x) {
        if (fmt.is_int) set_int_plane_out(x, fmt, out, out_pitch);
        else set_plain_out(x, out, out_pitch);
    };
    io.pitch64 = P;
    const double *g64 = fmt.is_int ? w.y64 : nullptr;
    if (mode == ISP_BM3D_BASIC || mode == ISP_BM3D_FINAL) {
        const bool fin = mode == ISP_BM3D_FINAL;
        io.guide[0] = fref;
        io.mguide32 = fref;
        io.mguide64 = g64;
        final_out(io);
        rc = run_step(ctx, fin ? ctx->params.final_ : ctx->params.basic, fin, io, w, W, H, s);
    } else {
        io.mguide32 = fref;
        io.mguide64 = g64;
        io.guide[0] = fref;
        set_plain_out(io, w.basic[0], (size_t)P);
        rc = run_step(ctx, ctx->params.basic, false, io, w, W, H, s);
        if (!rc) {
            io.mguide32 = w.basic[0];
            io.mguide64 = nullptr;
            io.guide[0] = w.basic[0];
            final_out(io);
            rc = run_step(ctx, ctx->params.final_, true, io, w, W, H, s);
        }
    }
    if (rc) return rc;
    if (host) {
        if ((rc = hs.download_start(0, out))) return rc;
        HIP_CHECK(ctx, hipStreamSynchronize(s));
        hs.download_finish(0, dst, stride);
    }
    return ISP_BM3D_OK;
}

// BM3D_Base::process_Frame's wrapper test for a single step (source/BM3D.cpp:284-290)
bool frame_degenerate(const isp_bm3d_ctx *ctx, int mode)
{
    const isp_bm3d_step_params &p = mode == ISP_BM3D_BASIC ? ctx->params.basic : ctx->params.final_;
    double norm[3];
    opp_norms(norm);
    return (!(p.sigma[0] > 0) && !(p.sigma[1] > 0) && !(p.sigma[2] > 0)) || p.group_size == 1 ||
           p.block_size <= 0 || p.bm_range <= 0 || p.bm_range < p.bm_step || p.thMSE * norm[0] <= 0;
}

// RGB frame: process_Frame (source/BM3D.cpp:282-313, 501-527) on float RGB in
// [0,1] or on int32 RGB with Floor / Ceil.
int rgb_common(isp_bm3d_ctx *ctx, void *const *dst, const void *const *src, const void *const *ref, int W, int H,
               size_t stride, const Fmt &fmt, int mode, int mem, void *stream)
{
    int rc = check_common(ctx, W, H, stride, mode, mem);
    if (rc) return rc;
    if (!dst[0] || !dst[1] || !dst[2] || !src[0] || !src[1] || !src[2])
        return fail(ctx, ISP_BM3D_E_ARG, "null dst/src plane");
    if (fmt.is_int && !(fmt.floor_ < fmt.ceil_)) return fail(ctx, ISP_BM3D_E_ARG, "bad quantisation");
    const bool has_ref = ref[0] || ref[1] || ref[2];
    if (has_ref && !(ref[0] && ref[1] && ref[2])) return fail(ctx, ISP_BM3D_E_ARG, "ref needs all three planes");
    const bool sep_ref = has_ref && (ref[0] != src[0] || ref[1] != src[1] || ref[2] != src[2]);
    hipStream_t s = (hipStream_t)stream;
    const bool host = mem == ISP_BM3D_MEM_HOST;
    const int P = internal_pitch(W, stride, mem);
    Work w{};
    if ((rc = prepare(ctx, w, P, H, W, false))) return rc;

    const void *rgb[3] = {src[0], src[1], src[2]};
    const void *rrgb[3] = {ref[0], ref[1], ref[2]};
    size_t in_pitch = stride;
    const HostStage hs{ctx, s, P, W, H};
    if (host) {
        // BM3D_Base::process_Frame's own early return needs no device work
        if (mode != ISP_BM3D_BOTH && frame_degenerate(ctx, mode)) {
            for (int c = 0; c < 3; ++c) host_copy_plane(dst[c], src[c], stride, W, H);
            return ISP_BM3D_OK;
        }
        // stage RGB in the out/basic planes, which are free until later
        if ((rc = HostStage::reserve(ctx, (sep_ref ? 6 : 3) * hs.plane_bytes()))) return rc;
        for (int c = 0; c < 3; ++c) {
            if ((rc = hs.upload(c, w.out[c], rgb[c], stride))) return rc;
            rgb[c] = w.out[c];
            if (sep_ref) {
                if ((rc = hs.upload(3 + c, w.basic[c], rrgb[c], stride))) return rc;
                rrgb[c] = w.basic[c];
            }
        }
        in_pitch = (size_t)P;
    }
    // BM3D_Base::process_Frame's own early return (source/BM3D.cpp:284-290)
    if (mode != ISP_BM3D_BOTH && frame_degenerate(ctx, mode)) {
        for (int c = 0; c < 3; ++c)
            if (dst[c] != rgb[c] && (rc = copy_plane(ctx, dst[c], stride, rgb[c], in_pitch, W, H, s))) return rc;
        return ISP_BM3D_OK;
    }
    // RGB -> OPP (float planes + FP64 luma guide)
    OppArgs o{};
    o.W = W;
    o.H = H;
    o.r = rgb[0];
    o.g = rgb[1];
    o.b = rgb[2];
    o.in_int = fmt.is_int;
    {
        const double sRange = fmt.is_int ? (double)(fmt.ceil_ - fmt.floor_) : 1.0;
        const double sFloor = fmt.is_int ? (double)fmt.floor_ : 0.0;
        o.gainY = 1.0 / (sRange * 3.0);
        o.offsetY = -sFloor * 3.0 * o.gainY + 0.0;
        o.gainU = 1.0 / (sRange * 2.0);
        o.gainV = 1.0 / (sRange * 4.0);
        o.offsetC = 0.0;
    }
    o.in_pitch = in_pitch;
    o.y = w.plane[0];
    o.u = w.plane[1];
    o.v = w.plane[2];
    o.pitch = P;
    o.y64 = w.y64;
    o.pitch64 = P;
    {
        StageTimer tm_opp(ctx, ST_OPP, s);
        HIP_CHECK(ctx, launch_opp(o, s));
        if (sep_ref) {
            OppArgs q = o;
            q.r = rrgb[0];
            q.g = rrgb[1];
            q.b = rrgb[2];
            q.y64 = w.ref64;
            if (mode == ISP_BM3D_FINAL) {  // BM3D_Final::RGB2YUV converts the whole guide
                q.y = w.refp[0];
                q.u = w.refp[1];
                q.v = w.refp[2];
            } else {                       // BM3D_Basic::RGB2YUV: ConvertToY only
                q.y = q.u = q.v = nullptr;
            }
            HIP_CHECK(ctx, launch_opp(q, s));
        }
    }
    void *out[3] = {dst[0], dst[1], dst[2]};
    size_t out_pitch = stride;
    if (host) {
        for (int c = 0; c < 3; ++c) out[c] = w.out[c];
        out_pitch = (size_t)P;
    }
    auto rgb_out = [&](StepIO &x) {
        // MatrixConvert_YUV2RGB OPP with clip (include/Conversion.hpp:660-762)
        x.to_rgb = 1;
        x.out_int = fmt.is_int;
        x.clip = 1;
        const double dRange = fmt.is_int ? (double)(fmt.ceil_ - fmt.floor_) : 1.0;
        x.gain_y = dRange / 1.0;
        x.gain_c = dRange / 1.0;
        x.offset = fmt.is_int ? (double)fmt.floor_ + 0.5 : 0.0;
        x.lo = fmt.is_int ? fmt.floor_ : 0.0;
        x.hi = fmt.is_int ? fmt.ceil_ : 1.0;
        for (int c = 0; c < 3; ++c) x.out[c] = out[c];
        x.out_pitch = out_pitch;
    };
    StepIO io{};
    io.nplanes = 3;
    io.pitch = P;
    io.pitch64 = P;
    for (int c = 0; c < 3; ++c) io.src[c] = w.plane[c];
    if (mode == ISP_BM3D_BASIC) {
        io.mguide64 = sep_ref ? w.ref64 : w.y64;
        for (int c = 0; c < 3; ++c) io.guide[c] = w.plane[c];
        rgb_out(io);
        rc = run_step(ctx, ctx->params.basic, false, io, w, W, H, s);
    } else if (mode == ISP_BM3D_FINAL) {
        io.mguide64 = sep_ref ? w.ref64 : w.y64;
        for (int c = 0; c < 3; ++c) io.guide[c] = sep_ref ? w.refp[c] : w.plane[c];
        rgb_out(io);
        rc = run_step(ctx, ctx->params.final_, true, io, w, W, H, s);
    } else {
        io.mguide64 = sep_ref ? w.ref64 : w.y64;
        for (int c = 0; c < 3; ++c) {
            io.guide[c] = w.plane[c];
            io.out[c] = w.basic[c];
        }
        io.to_rgb = 0;
        io.out_int = 0;
        io.clip = 0;
        io.out_pitch = (size_t)P;
        rc = run_step(ctx, ctx->params.basic, false, io, w, W, H, s);
        if (!rc) {
            io.mguide64 = nullptr;
            io.mguide32 = w.basic[0];
            for (int c = 0; c < 3; ++c) io.guide[c] = w.basic[c];
            rgb_out(io);
            rc = run_step(ctx, ctx->params.final_, true, io, w, W, H, s);
        }
    }
    if (rc) return rc;
    if (host) {
        for (int c = 0; c < 3; ++c)
            if ((rc = hs.download_start(c, out[c]))) return rc;
        HIP_CHECK(ctx, hipStreamSynchronize(s));
        for (int c = 0; c < 3; ++c) hs.download_finish(c, dst[c], stride);
    }
    return ISP_BM3D_OK;
}

}  // namespace

extern "C" {

int isp_bm3d_plane_f32(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref, int W, int H,
                       size_t stride, int mode, int mem, void *stream)
{
    const Fmt f{false, 0, 0, 1};
    return plane_common(ctx, dst, src, ref, W, H, stride, f, mode, mem, stream);
}

int isp_bm3d_plane_i32(isp_bm3d_ctx *ctx, int *dst, const int *src, const int *ref, int W, int H, size_t stride,
                       int floor_, int neutral, int ceil_, int mode, int mem, void *stream)
{
    const Fmt f{true, floor_, neutral, ceil_};
    return plane_common(ctx, dst, src, ref, W, H, stride, f, mode, mem, stream);
}

int isp_bm3d_rgb_f32(isp_bm3d_ctx *ctx, float *dst_r, float *dst_g, float *dst_b, const float *src_r,
                     const float *src_g, const float *src_b, const float *ref_r, const float *ref_g,
                     const float *ref_b, int W, int H, size_t stride, int mode, int mem, void *stream)
{
    void *d[3] = {dst_r, dst_g, dst_b};
    const void *sv[3] = {src_r, src_g, src_b}, *r[3] = {ref_r, ref_g, ref_b};
    const Fmt f{false, 0, 0, 1};
    return rgb_common(ctx, d, sv, r, W, H, stride, f, mode, mem, stream);
}

int isp_bm3d_rgb_i32(isp_bm3d_ctx *ctx, int *dst_r, int *dst_g, int *dst_b, const int *src_r, const int *src_g,
                     const int *src_b, const int *ref_r, const int *ref_g, const int *ref_b, int W, int H,
                     size_t stride, int floor_, int ceil_, int mode, int mem, void *stream)
{
    void *d[3] = {dst_r, dst_g, dst_b};
    const void *sv[3] = {src_r, src_g, src_b}, *r[3] = {ref_r, ref_g, ref_b};
    const Fmt f{true, floor_, floor_, ceil_};
    return rgb_common(ctx, d, sv, r, W, H, stride, f, mode, mem, stream);
}

int isp_bm3d_match(isp_bm3d_ctx *ctx, const float *guide, int W, int H, size_t stride, int final_step,
                   int mem, int *counts, int *pos, double *keys)
{
    int rc = check_common(ctx, W, H, stride, ISP_BM3D_BASIC, mem);
    if (rc) return rc;
    if (!guide || !counts || !pos) return fail(ctx, ISP_BM3D_E_ARG, "null argument");
    const isp_bm3d_step_params &p = final_step ? ctx->params.final_ : ctx->params.basic;
    Derived d{};
    if ((rc = derive(ctx, p, 1, W, H, d))) return rc;
    if (d.degenerate) return fail(ctx, ISP_BM3D_E_ARG, "parameters are degenerate (no matching runs)");
    const int P = internal_pitch(W, stride, mem);
    Work w{};
    if ((rc = prepare(ctx, w, P, H, W, true))) return rc;
    hipStream_t s = nullptr;
    const float *g = guide;
    if (mem == ISP_BM3D_MEM_HOST) {
        const HostStage hs{ctx, s, P, W, H};
        if ((rc = HostStage::reserve(ctx, hs.plane_bytes()))) return rc;
        if ((rc = hs.upload(0, w.plane[0], guide, stride))) return rc;
        g = w.plane[0];
    }
    rc = run_match(ctx, p, d, w, g, nullptr, P, W, H, w.keys, s, final_step ? ST_MATCH_FINAL : ST_MATCH_BASIC);
    if (rc) return rc;
    const size_t refs = (size_t)d.ngx * d.ngy, G = (size_t)p.group_size;
    const int k = final_step ? 1 : 0;
    HIP_CHECK(ctx, hipMemcpyAsync(counts, w.cnt[k], refs * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(ctx, hipMemcpyAsync(pos, w.pos[k], refs * G * sizeof(int), hipMemcpyDeviceToHost, s));
    if (keys) HIP_CHECK(ctx, hipMemcpyAsync(keys, w.keys, refs * G * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(ctx, hipStreamSynchronize(s));
    return ISP_BM3D_OK;
}

}  // extern "C"
