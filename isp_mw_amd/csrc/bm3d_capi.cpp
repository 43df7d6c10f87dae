// C ABI of the MI355X BM3D denoiser: parameters, context, workspace and the
// orchestration of one or two BM3D steps on one GPU.
//
// Reference behaviour mirrored here:
//   parameter presets and thMSE_Default   include/BM3D.h:15-168
//   OPP-norm scaling of sigma and thMSE   include/BM3D.h:237-259
//   copy-through on degenerate parameters source/BM3D.cpp:82-87, 143-151, 284-290
//   step orchestration                    source/BM3D.cpp:464-527
#include "../../include/isp_bm3d.h"
#include "../../include/isp_nlm.h"
#include "bm3d_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace ispbm3d;

struct isp_bm3d_ctx {
    isp_bm3d_params params;
    int device = 0;
    int precision = ISP_BM3D_PREC_FP32;
    bool plain_valu = false;  // the plain form's filter on the vector ALUs (ISP_BM3D_FP32_VALU)
    // FP32-distance matching by column walks where eligible; the environment
    // variable ISP_BM3D_MATCH_WALK=0 at context creation keeps every search
    // on the list kernel (the two give identical groups)
    bool match_walk = true;
    float *d_tables = nullptr;  // C8 (64) then the padded C_N matrices
    int cn_off[kMaxGroup + 1] = {0};
    double *d_tables64 = nullptr;  // C8 (64) then the unpadded C_N matrices, FP64
    int cnd_off[kMaxGroup + 1] = {0};
    // MFMA operand tables: [0] FP16, [1] BF16, [2] the low halves of [0]
    // (fp16(v - fp16(v)), the FP32_MFMA form's second operand halves)
    uint16_t *d_mfma[3] = {nullptr, nullptr, nullptr};
    // generic kernels (BlockSize != 8 or GroupSize > 32), per step [basic, final]:
    // the B x B orthonormal DCT-II, then C_N for N = 1..G, FP64
    double *d_gen[2] = {nullptr, nullptr};
    // variant modes (isp_bm3d_set_variant): every step on the generic kernels
    // with these tables per step: B x B DCT-II, C_N for N = 1..G (the Haar
    // matrix for a power-of-two N in the Haar mode), then the B x B window
    int variant = 0;
    double kaiser_beta = 0.0;
    double *d_var[2] = {nullptr, nullptr};
    size_t var_kaiser_off[2] = {0, 0};
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // this device's copy of the allocation-free entry point's scratch array
    // (scratch.hip), looked up at creation; null if the lookup failed
    void *scratch = nullptr;
    char *pin = nullptr;  // page-locked staging for host-memory calls
    size_t pin_bytes = 0;
    // tile shapes (reference blocks per workgroup) of the single-image launches
    struct Tiles {
        int match_x = ispbm3d::kTileRX, match_y = ispbm3d::kTileRY;
        // FP32-distance matching (the mixed forms) walks tile columns: tall tiles
        int walk_x = ispbm3d::kWalkTileX, walk_y = ispbm3d::kWalkTileY;
        int filter_x = ispbm3d::kTileRX, filter_y = ispbm3d::kTileRY;
        int mfma_x = ispbm3d::kMTileRX, mfma_y = ispbm3d::kMTileRY;
    } tiles;
    std::map<int, isp_bm3d_tuning> tune;  // size class -> tuned tiles (and group size)
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
    ST_COPY,
    ST_NLM_MATCH,
    ST_NLM_AVERAGE
};
const char *const kStageNames[ISP_BM3D_NUM_STAGES] = {
    "opp", "fill", "match_basic", "filter_basic", "match_final", "filter_final", "finalize", "copy",
    "nlm_match", "nlm_average"};

hipEvent_t take_event(isp_bm3d_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    // timestamps only: the caller's stream synchronisation makes results
    // visible, so the events skip the system-scope fence (cheaper records)
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
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
    double thr64[3], sig2_64[3];
    float band[3];
    float wmax;  // cap of the Wiener group weight (aggregation headroom)
    int active[3];
    int ngx, ngy;
    int B;         // BlockSize
    bool generic;  // BlockSize != 8 or GroupSize > 32 (or a variant mode): the generic kernels
    const double *gen_tab;  // their transform tables (GenFilterArgs::cb, then cn)
    int haar;               // variant: Haar along the group axis
    const double *kaiser;   // variant: B x B aggregation window, or null
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
    const int B = p.block_size;
    d.B = B;
    // BlockSize 8 with GroupSize <= 32 runs the tuned kernels; anything else
    // within the generic kernels' bounds runs those
    d.generic = B != kBlock || p.group_size > kMaxGroup || ctx->variant != 0;
    if (B > kGenMaxBlock)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "BlockSize %d outside 1..%d", B, kGenMaxBlock);
    if (p.group_size < 0 || p.group_size > kGenMaxGroup)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "GroupSize %d outside 1..%d", p.group_size, kGenMaxGroup);
    if (d.generic && p.group_size * B * B > kGenMaxElems)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "GroupSize %d x BlockSize %d^2 above %d samples per group",
                    p.group_size, B, kGenMaxElems);
    d.gen_tab = nullptr;
    d.haar = 0;
    d.kaiser = nullptr;
    for (int k = 0; d.generic && k < 2; ++k) {
        // the tables built for a step of this BlockSize hold C_N up to its GroupSize
        const isp_bm3d_step_params &q = k ? ctx->params.final_ : ctx->params.basic;
        if (q.block_size != B || p.group_size > q.group_size) continue;
        if (ctx->variant) {
            if (!ctx->d_var[k]) continue;
            d.gen_tab = ctx->d_var[k];
            d.haar = (ctx->variant & ISP_BM3D_VARIANT_HAAR) != 0;
            d.kaiser = (ctx->variant & ISP_BM3D_VARIANT_KAISER) ? ctx->d_var[k] + ctx->var_kaiser_off[k] : nullptr;
        } else if (ctx->d_gen[k]) {
            d.gen_tab = ctx->d_gen[k];
        }
    }
    if (d.generic && !d.gen_tab)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "no transform tables for BlockSize %d / GroupSize %d", B,
                    p.group_size);
    if (p.block_step <= 0 || p.bm_step <= 0)
        return fail(ctx, ISP_BM3D_E_ARG, "BlockStep and BMstep must be positive");
    if (W < B || H < B)
        return fail(ctx, ISP_BM3D_E_SIZE, "image %dx%d is smaller than one %dx%d block", W, H, B, B);
    // positions are packed (y << 16) | x in an int32 and decoded with an
    // arithmetic shift: y must stay below 2^15
    if (W > 65535 || H > 32767)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "image %dx%d: width must be < 65536 and height < 32768", W, H);
    d.range = p.bm_range / p.bm_step * p.bm_step;
    d.wmax = wiener_weight_cap(B, p.block_step, d.range, p.bm_step, p.group_size, 0);
    if (!(d.wmax > 0))
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED,
                    "BMrange %d / BlockStep %d / GroupSize %d: a pixel's aggregation sums could overflow",
                    p.bm_range, p.block_step, p.group_size);
    d.bypass = p.group_size == 1 || thMSE <= 0;
    // include/Block.h:582-584 with range 1 (Plane_FL default Floor 0 / Ceil 1)
    const double MSE2SSE = (double)(B * B) * 1.0 * 1.0 / (double)(255 * 255);
    d.distMul = 1.0 / MSE2SSE;
    d.thSSE = thMSE * MSE2SSE;
    for (int c = 0; c < 3; ++c) {
        d.active[c] = c < nplanes && p.sigma[c] > 0;
        const double sigma = p.sigma[c] / 255.0 * norm[c];
        d.thr[c] = (float)(sigma * p.lambda);
        d.sig2[c] = (float)(sigma * sigma);
        d.sig2_64[c] = sigma * sigma;
        d.thr64[c] = sigma * p.lambda;
        // FP32 coefficients of [0,1] data carry absolute errors of a few 1e-6;
        // the plain form re-makes decisions closer than this to the threshold
        // in FP64; the FP32-throughout form does not (a negative band)
        d.band[c] = ctx->precision == ISP_BM3D_PREC_FP32 ? (float)(3e-5 + 1e-4 * d.thr64[c]) : -1.f;
    }
    d.ngx = (W - B) / p.block_step + 1 + ((W - B) % p.block_step ? 1 : 0);
    d.ngy = (H - B) / p.block_step + 1 + ((H - B) % p.block_step ? 1 : 0);
    return ISP_BM3D_OK;
}

// Largest GroupSize of the context's two steps (the group lists' row length
// in the workspace), at least kMaxGroup.
int max_group(const isp_bm3d_ctx *ctx)
{
    int G = kMaxGroup;
    for (const isp_bm3d_step_params *p : {&ctx->params.basic, &ctx->params.final_})
        G = std::max(G, std::min(p->group_size, kGenMaxGroup));
    return G;
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

size_t layout(Work *w, void *base, int P, int H, size_t refs, int G, bool keys)
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
        w->pos[k] = c.take<int>(refs * (size_t)G);
    }
    w->keys = keys ? c.take<double>(refs * (size_t)G) : nullptr;
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
    const int G = max_group(ctx);
    const size_t bytes = layout(&w, nullptr, P, H, refs, G, keys);
    int rc = ensure_ws(ctx, bytes);
    if (rc) return rc;
    layout(&w, ctx->ws, P, H, refs, G, keys);
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

int size_class(int W, int H)
{
    const unsigned long long n = (unsigned long long)W * (unsigned long long)H;
    int k = 0;
    while (k < 63 && (1ull << k) < n) ++k;
    return k;
}

// The context's engines: the FP16/BF16 matrix-core filter (engine 1 or 2),
// the split FP16 matrix-core filter with FP32-accurate products (engine 3:
// the FP32_MFMA form, and the plain form, which adds the FP64 re-decision
// near the hard threshold; ISP_BM3D_FP32_VALU=1 puts the plain form back on
// the vector filter) or the FP32 vector filter (engine 0: the FP32_FAST
// form); matching with FP64 distances in the reference's order (the plain
// form) or FP32 distances (every other).
inline int filter_engine(const isp_bm3d_ctx *ctx)
{
    switch (ctx->precision) {
    case ISP_BM3D_PREC_FP16: return 1;
    case ISP_BM3D_PREC_BF16: return 2;
    case ISP_BM3D_PREC_FP32_MFMA: return 3;
    case ISP_BM3D_PREC_FP32: return ctx->plain_valu ? 0 : 3;
    default: return 0;
    }
}
// Multi-frame groups: the same engines on the stack kernels.
inline int stack_engine(const isp_bm3d_ctx *ctx) { return filter_engine(ctx); }
inline bool exact_matching(const isp_bm3d_ctx *ctx) { return ctx->precision == ISP_BM3D_PREC_FP32; }

// Tile shapes of an image: the tuning entry of its size class, else the defaults.
isp_bm3d_ctx::Tiles tiles_for(const isp_bm3d_ctx *ctx, int W, int H)
{
    isp_bm3d_ctx::Tiles t = ctx->tiles;
    if (!exact_matching(ctx)) {
        t.match_x = t.walk_x;
        t.match_y = t.walk_y;
    }
    auto it = ctx->tune.find(size_class(W, H));
    if (it != ctx->tune.end()) {
        t.match_x = it->second.match_tile[0];
        t.match_y = it->second.match_tile[1];
        if (filter_engine(ctx)) {
            t.mfma_x = it->second.filter_tile[0];
            t.mfma_y = it->second.filter_tile[1];
        } else {
            t.filter_x = it->second.filter_tile[0];
            t.filter_y = it->second.filter_tile[1];
        }
    }
    if (filter_engine(ctx) == 3) {
        // the FP32_MFMA form stages both halves of its footprint and of the
        // K64 fragments: where a step's LDS would exceed a CU's, the tile
        // shrinks (rows first) until it fits (tiles never change results)
        auto fits = [&](int tx, int ty) {
            for (const isp_bm3d_step_params *q : {&ctx->params.basic, &ctx->params.final_}) {
                FilterArgs probe{};
                probe.FW = (tx - 1) * q->block_step + 2 * q->bm_range + kBlock;
                probe.FH = (ty - 1) * q->block_step + 2 * q->bm_range + kBlock;
                probe.kbl = ctx->d_mfma[2];
                if (filter_mfma_lds_bytes(probe, true) > 160 * 1024) return false;
            }
            return true;
        };
        while (t.mfma_x * t.mfma_y > 1 && !fits(t.mfma_x, t.mfma_y)) {
            if (t.mfma_y >= t.mfma_x) t.mfma_y = (t.mfma_y + 1) / 2;
            else t.mfma_x = (t.mfma_x + 1) / 2;
        }
    }
    return t;
}

// Arguments of one matching launch on one image (guide: FP32 or FP64 plane).
// Row pitch (elements) of matching's staged tile.  Lane s of a read (strip s
// of the nsx strips of STRIP offsets per lattice row of the window) reads
// consecutive elements of the tile from element (s / nsx) * st * pitch +
// (s % nsx) * STRIP * st + const.  Bank model, fitted to a probe of these
// reads at every pitch from 41 to 167 (8 waves per CU; correlation 0.999 for
// the column walk's lanes at BMrange 12, 0.996 for NL-means' strips of 6 at
// BMstep 3, 0.97 for the FP64 list kernel): 4-byte elements are read by
// groups of 32 lanes over 32 banks, 8-byte elements by groups of 16 lanes,
// and a group costs its largest number of lanes on one bank.  The pitch is
// the one of tileW .. tileW + 31 with the smallest model cost over the
// window's strips (tileW | 1 unless another is strictly cheaper); the model's
// choices are the measured fastest in all three probed cases (column walk:
// 2.55 clocks per CU per element read at pitch 57 or 89, 4.41 at 47, 10.1 at
// 59).  STRIP follows the kernel the step takes (the column walk: strips of 6
// where one round of 64 covers the window, else 5; the list kernel: kStrip),
// nsx a block whose window the image border does not clip.  The odd pitch
// tileW | 1 is kept where the chosen one would lower the workgroups a CU holds.
int match_pitch(const MatchArgs &a, const Derived &d, const isp_bm3d_step_params &p)
{
    const int st = p.bm_step, NM = 2 * d.range / st + 1, nm = 2 * (d.range / st) + 1;
    int strip = kStrip;
    if (a.walk) {
        const bool multi = nm * ((nm + 4) / 5) > 128, one = nm * ((nm + 5) / 6) <= 64;
        strip = !multi && one ? 6 : 5;
    }
    const int nsx = (NM + strip - 1) / strip, nstrips = std::min(NM * nsx, 256);
    const int words = !a.walk && a.exact ? 2 : 1, G = 32 / words;  // FP64 tiles in the exact list kernel
    auto cost = [&](int P) {
        int t = 0;
        for (int g0 = 0; g0 < nstrips; g0 += G) {
            int cnt[32] = {}, mx = 0;
            for (int s = g0; s < std::min(g0 + G, nstrips); ++s) {
                const long o = ((long)(s / nsx) * st * P + (long)(s % nsx) * strip * st) * words;
                for (int k = 0; k < words; ++k) mx = std::max(mx, ++cnt[(o + k) % 32]);
            }
            t += mx;
        }
        return t;
    };
    MatchArgs c = a;
    c.tilePitch = a.tileW | 1;
    int bc = cost(c.tilePitch);
    const size_t l0 = match_lds_bytes(c);
    for (int P = a.tileW; P < a.tileW + 32; ++P) {
        const int cp = cost(P);
        if (cp < bc) {
            bc = cp;
            c.tilePitch = P;
        }
    }
    const size_t l1 = match_lds_bytes(c);
    constexpr size_t cu_lds = 160 * 1024;
    return std::min<size_t>(4, cu_lds / l1) >= std::min<size_t>(4, cu_lds / l0) ? c.tilePitch : (a.tileW | 1);
}

MatchArgs build_match(const isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const Derived &d, int W, int H,
                      const void *guide, int gpitch, int trx, int try_, int *cnt, int *pos, double *keys)
{
    MatchArgs a{};
    a.guide = guide;
    a.gpitch = gpitch;
    a.trx = trx;
    a.try_ = try_;
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
    a.exact = exact_matching(ctx);
    const int side = 2 * (d.range / p.bm_step) + 1;
    a.maxCand = std::min(side * side, kMatchListCap);
    a.cnt_out = cnt;
    a.pos_out = pos;
    a.key_out = keys;
    a.walk = ctx->match_walk && match_walk_eligible(a) ? 1 : 0;
    // the staged tile: the window of every block plus the widest lane strip
    a.tileW = (trx - 1) * p.block_step + 2 * d.range + kBlock + ((a.walk ? kStripMax : kStrip) - 1) * p.bm_step;
    a.tileH = (try_ - 1) * p.block_step + 2 * d.range + kBlock;
    a.tilePitch = match_pitch(a, d, p);
    return a;
}

int check_match_lds(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const MatchArgs &a)
{
    if (match_lds_bytes(a) > 160 * 1024)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED,
                    "BMrange %d / BMstep %d / BlockStep %d need %zu bytes of LDS for matching (max 163840)",
                    p.bm_range, p.bm_step, p.block_step, match_lds_bytes(a));
    return ISP_BM3D_OK;
}

// Arguments of one filtering launch on one image; the engine (FP32 VALU or
// the matrix cores) follows the context's precision.
FilterArgs build_filter(const isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const Derived &d, int W, int H,
                        int pitch, const float *const src[3], const float *const guide[3],
                        unsigned long long *const num[3], unsigned long long *const den[3], const int *cnt,
                        const int *pos, int trx, int try_)
{
    FilterArgs f{};
    f.trx = trx;
    f.try_ = try_;
    f.W = W;
    f.H = H;
    f.pitch = pitch;
    f.ngx = d.ngx;
    f.ngy = d.ngy;
    f.bstep = p.block_step;
    f.range = d.range;
    f.G = p.group_size;
    f.cnt = cnt;
    f.pos = pos;
    for (int c = 0; c < 3; ++c) {
        f.src[c] = src[c];
        f.guide[c] = guide[c];
        f.num[c] = num[c];
        f.den[c] = den[c];
        f.active[c] = d.active[c];
        f.thr[c] = d.thr[c];
        f.sig2[c] = d.sig2[c];
        f.thr64[c] = d.thr64[c];
        f.band[c] = d.band[c];
    }
    f.wmax = d.wmax;
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
    if (const int eng = filter_engine(ctx)) {
        // the FP32_MFMA form: FP16 high halves plus the low-half tables
        const uint16_t *t = ctx->d_mfma[eng == 3 ? 0 : eng - 1];
        f.kb = t;
        f.ka = t + 8 * 64 * 8;
        f.cna = t + 16 * 64 * 8;
        f.cnb = f.cna + (size_t)kMaxGroup * 2 * 64 * 8;
        f.cnap = f.cnb + (size_t)kMaxGroup * 2 * 64 * 8;
        f.cnbp = f.cnap + (size_t)kPairTables * 2 * 64 * 8;
        if (eng == 3) {
            const uint16_t *l = ctx->d_mfma[2];
            f.kbl = l;
            f.kal = l + 8 * 64 * 8;
            f.cnal = l + 16 * 64 * 8;
            f.cnbl = f.cnal + (size_t)kMaxGroup * 2 * 64 * 8;
            f.cnapl = f.cnbl + (size_t)kMaxGroup * 2 * 64 * 8;
            f.cnbpl = f.cnapl + (size_t)kPairTables * 2 * 64 * 8;
        }
    }
    return f;
}

size_t filter_lds(const isp_bm3d_ctx *ctx, const FilterArgs &f, bool wiener)
{
    return filter_engine(ctx) ? filter_mfma_lds_bytes(f, wiener) : filter_lds_bytes(f, wiener);
}

hipError_t launch_filter_any(const isp_bm3d_ctx *ctx, const FilterArgs &f, bool wiener, hipStream_t s,
                             const BatchLaunch *b = nullptr)
{
    if (filter_engine(ctx)) return launch_filter_mfma(f, wiener, filter_engine(ctx), s, b);
    return launch_filter(f, wiener, s, b);
}

// The output side of one step: num/den, then the plane / RGB conversion.
FinalizeArgs build_finalize(int W, int H, const StepIO &io, const Derived &d, unsigned long long *const num[3],
                            unsigned long long *const den[3])
{
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
        fa.num[c] = num[c];
        fa.den[c] = den[c];
        fa.active[c] = d.degenerate ? 0 : d.active[c];
    }
    // the tiled filter kernels add den per block origin, the generic ones per pixel
    fa.den_box = d.generic ? 0 : kBlock;
    return fa;
}

// Matching arguments of the generic kernel: any BlockSize, FP64 distances in
// the reference's order whatever the context's precision; the tile shrinks
// until its LDS fits a CU.
int build_match_generic(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const Derived &d, int W, int H,
                        const void *guide, int gpitch, int *cnt, int *pos, double *keys, MatchArgs &a)
{
    int trx = kTileRX, try_ = kTileRY;
    for (;;) {
        a = build_match(ctx, p, d, W, H, guide, gpitch, trx, try_, cnt, pos, keys);
        a.B = d.B;
        a.exact = 1;
        a.walk = 0;
        a.tileW = (trx - 1) * p.block_step + 2 * d.range + d.B + (kStrip - 1) * p.bm_step;
        a.tileH = (try_ - 1) * p.block_step + 2 * d.range + d.B;
        a.tilePitch = a.tileW | 1;
        if (match_generic_lds_bytes(a) <= 160 * 1024) return ISP_BM3D_OK;
        if (trx == 1 && try_ == 1)
            return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "BMrange %d / BlockSize %d need %zu bytes of LDS for matching",
                        p.bm_range, d.B, match_generic_lds_bytes(a));
        if (trx >= try_ && trx > 1) trx = (trx + 1) / 2;
        else try_ = (try_ + 1) / 2;
    }
}

int run_match(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, const Derived &d, Work &w,
              const float *g32, const double *g64, int gpitch, int W, int H, double *keys,
              hipStream_t s, int stage)
{
    const int k = stage == ST_MATCH_FINAL ? 1 : 0;
    if (d.generic) {
        MatchArgs a{};
        int rc = build_match_generic(ctx, p, d, W, H, g64 ? (const void *)g64 : (const void *)g32, gpitch,
                                     w.cnt[k], w.pos[k], keys, a);
        if (rc) return rc;
        ctx->last_cnt[k] = w.cnt[k];
        ctx->last_pos[k] = w.pos[k];
        ctx->last_refs[k] = (size_t)d.ngx * d.ngy;
        ctx->last_G[k] = p.group_size;
        StageTimer tm(ctx, stage, s);
        HIP_CHECK(ctx, launch_match_generic(a, g64 != nullptr, s));
        return ISP_BM3D_OK;
    }
    const isp_bm3d_ctx::Tiles t = tiles_for(ctx, W, H);
    const MatchArgs a = build_match(ctx, p, d, W, H, g64 ? (const void *)g64 : (const void *)g32, gpitch,
                                    t.match_x, t.match_y, w.cnt[k], w.pos[k], keys);
    ctx->last_cnt[k] = w.cnt[k];
    ctx->last_pos[k] = w.pos[k];
    ctx->last_refs[k] = (size_t)d.ngx * d.ngy;
    ctx->last_G[k] = p.group_size;
    int rc = check_match_lds(ctx, p, a);
    if (rc) return rc;
    StageTimer tm(ctx, stage, s);
    HIP_CHECK(ctx, launch_match(a, g64 != nullptr, s));
    return ISP_BM3D_OK;
}

// One BM3D_Base::Kernel (source/BM3D.cpp:80-218) followed by finalize.
int run_step(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, bool wiener, const StepIO &io, Work &w,
             int W, int H, hipStream_t s)
{
    Derived d{};
    int rc = derive(ctx, p, io.nplanes, W, H, d);
    if (rc) return rc;
    const FinalizeArgs fa = build_finalize(W, H, io, d, w.num, w.den);
    if (!d.degenerate) {
        rc = run_match(ctx, p, d, w, io.mguide32, io.mguide64, io.mguide64 ? io.pitch64 : io.pitch, W, H,
                       nullptr, s, wiener ? ST_MATCH_FINAL : ST_MATCH_BASIC);
        if (rc) return rc;
        {
            StageTimer tm(ctx, ST_FILL, s);
            // zero bits of the 64-bit accumulators: num[0..2] and den[0..2] are
            // one contiguous span of the workspace (layout), cleared by one launch
            char *lo = reinterpret_cast<char *>(w.num[0]);
            char *hi = reinterpret_cast<char *>(w.den[2] + (size_t)io.pitch * H);
            HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(lo), (size_t)(hi - lo) / 4, 0.f, s));
        }
        const int k = wiener ? 1 : 0;
        if (d.generic) {
            GenFilterArgs g{};
            g.W = W;
            g.H = H;
            g.pitch = io.pitch;
            g.ngx = d.ngx;
            g.ngy = d.ngy;
            g.B = d.B;
            g.G = p.group_size;
            g.cnt = w.cnt[k];
            g.pos = w.pos[k];
            for (int c = 0; c < 3; ++c) {
                g.src[c] = io.src[c];
                g.guide[c] = io.guide[c];
                g.num[c] = w.num[c];
                g.den[c] = w.den[c];
                g.active[c] = d.active[c];
                g.thr[c] = d.thr64[c];
                g.sig2[c] = d.sig2_64[c];
            }
            g.wmax = d.wmax;
            g.cb = d.gen_tab;
            g.cn = d.gen_tab + (size_t)d.B * d.B;
            g.haar = d.haar;
            g.kaiser = d.kaiser;
            StageTimer tm(ctx, wiener ? ST_FILTER_FINAL : ST_FILTER_BASIC, s);
            HIP_CHECK(ctx, launch_filter_generic(g, wiener, s));
        } else {
            const bool mfma = filter_engine(ctx) != 0;
            const isp_bm3d_ctx::Tiles t = tiles_for(ctx, W, H);
            const FilterArgs f = build_filter(ctx, p, d, W, H, io.pitch, io.src, io.guide, w.num, w.den,
                                              w.cnt[k], w.pos[k], mfma ? t.mfma_x : t.filter_x,
                                              mfma ? t.mfma_y : t.filter_y);
            const size_t lds = filter_lds(ctx, f, wiener);
            if (lds > 160 * 1024)
                return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "parameters need %zu bytes of LDS for filtering", lds);
            StageTimer tm(ctx, wiener ? ST_FILTER_FINAL : ST_FILTER_BASIC, s);
            HIP_CHECK(ctx, launch_filter_any(ctx, f, wiener, s));
        }
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
    // two groups of N1, N2 <= 16 blocks in one operand: block-diagonal
    // diag(C_N1, C_N2), group 1 on rows / columns 0..15, group 2 on 16..31
    auto cpair = [&](int N1, int N2, int k, int z) -> double {
        if (k < 16 && z < 16) return cn(N1, k, z);
        if (k >= 16 && z >= 16 && N2 > 0) return cn(N2, k - 16, z - 16);
        return 0.0;
    };
    for (int N1 = 1; N1 <= 16; ++N1)
        for (int N2 = 0; N2 <= 16; ++N2)
            for (int m = 0; m < 2; ++m)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) t.push_back((float)cpair(N1, N2, l & 31, perm(m, l >> 5, e)));
    for (int N1 = 1; N1 <= 16; ++N1)
        for (int N2 = 0; N2 <= 16; ++N2)
            for (int m = 0; m < 2; ++m)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) t.push_back((float)cpair(N1, N2, perm(m, l >> 5, e), l & 31));
    return t;
}


// ---------------------------------------------------------------------------
// Batches.  Every image gets its own workspace; per step, one launch of each
// stage covers all images through a workgroup -> (image, tile) map and an
// array of per-image arguments in device memory.  The launches of a run are
// captured once into a hipGraph.
struct BatchImg {
    int W, H, P;            // P: row pitch of the caller's planes and of the workspace planes
    isp_bm3d_params prm;
    Derived d[2];
    isp_bm3d_ctx::Tiles tiles;
    float *yuv[3];          // RGB: OPP planes
    float *refyuv[3];       // RGB, final step alone: OPP planes of the guide frame
    double *y64;            // RGB: FP64 luma guide (of ref when given)
    float *basic[3];
    unsigned long long *num[3], *den[3];
    int *cnt[2], *pos[2];
};

struct BatchStepPlan {
    bool any = false;          // some image runs matching/filtering in this step
    MatchArgs hm{};            // launch header: largest LDS
    FilterArgs hf{};
    FinalizeArgs hfin{};
    std::vector<MatchArgs> m;
    std::vector<FilterArgs> f;
    std::vector<FinalizeArgs> fin;
    std::vector<uint32_t> mt, ft;
    MatchArgs *dm = nullptr;
    FilterArgs *df = nullptr;
    FinalizeArgs *dfin = nullptr;
    uint32_t *dmt = nullptr, *dft = nullptr;
};

}  // namespace

struct isp_bm3d_batch {
    isp_bm3d_ctx *ctx = nullptr;
    int count = 0, nplanes = 1, mode = ISP_BM3D_BOTH, flags = 0;
    bool guide_y64 = false;          // RGB: step-1 guide is FP64
    std::vector<isp_bm3d_image> img;
    std::vector<BatchImg> bi;
    BatchStepPlan st[2];
    OppArgs ho{};
    std::vector<OppArgs> o;          // per image: RGB -> OPP
    std::vector<OppArgs> oref;       // per image: ConvertToY of its guide frame (when it has one)
    OppArgs *dop = nullptr, *dopref = nullptr;
    std::vector<uint32_t> rows;      // (image << 20) | y over all images
    std::vector<uint32_t> rows_ref;  // the same over the images with a guide frame
    uint32_t *drows = nullptr, *drows_ref = nullptr;
    int maxW = 0;
    void *mem = nullptr;
    size_t mem_bytes = 0;
    unsigned long long *acc = nullptr;
    size_t acc_words = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int launches = 0;
    unsigned max_wg = 0;
    // tile overrides of the autotuner (x <= 0: none)
    int ov_match[2] = {0, 0}, ov_filter[2] = {0, 0}, ov_group = 0;
};

namespace {

// The steps a batch runs, as kinds (0 basic, 1 final), in order.
int batch_kinds(const isp_bm3d_batch *b, int kinds[2])
{
    if (b->mode == ISP_BM3D_BOTH) {
        kinds[0] = 0;
        kinds[1] = 1;
        return 2;
    }
    kinds[0] = b->mode == ISP_BM3D_FINAL ? 1 : 0;
    return 1;
}

void batch_free(isp_bm3d_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->mem) (void)hipFree(b->mem);
    b->exec = nullptr;
    b->graph = nullptr;
    b->cap = nullptr;
    b->mem = nullptr;
}

// Per-image parameters: the context's, with the image's sigma and thMSE_Default.
int batch_params(isp_bm3d_batch *b, int k)
{
    isp_bm3d_ctx *ctx = b->ctx;
    const isp_bm3d_image &im = b->img[k];
    BatchImg &x = b->bi[k];
    x.prm = ctx->params;
    if (im.sigma > 0) {
        for (int c = 0; c < 3; ++c) x.prm.basic.sigma[c] = x.prm.final_.sigma[c] = im.sigma;
        int rc = isp_bm3d_thmse_default(&x.prm.basic, 0);
        if (!rc) rc = isp_bm3d_thmse_default(&x.prm.final_, 1);
        if (rc) return fail(ctx, ISP_BM3D_E_ARG, "image %d: thMSE_Default failed", k);
    }
    x.tiles = tiles_for(ctx, x.W, x.H);
    if (b->ov_match[0] > 0) {
        x.tiles.match_x = b->ov_match[0];
        x.tiles.match_y = b->ov_match[1];
    }
    if (b->ov_filter[0] > 0) {
        x.tiles.filter_x = x.tiles.mfma_x = b->ov_filter[0];
        x.tiles.filter_y = x.tiles.mfma_y = b->ov_filter[1];
    }
    int g = 0;
    if (b->ov_group > 0) g = b->ov_group;
    else if (b->flags & ISP_BM3D_BATCH_AUTO_GROUP) {
        auto it = ctx->tune.find(size_class(x.W, x.H));
        if (it != ctx->tune.end()) g = it->second.final_group_size;
    }
    if (g > 0) x.prm.final_.group_size = g;
    for (int s = 0; s < 2; ++s) {
        int rc = derive(ctx, s ? x.prm.final_ : x.prm.basic, b->nplanes, x.W, x.H, x.d[s]);
        if (rc) return rc;
        if (x.d[s].generic)
            return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "batches need BlockSize 8 and GroupSize <= %d", kMaxGroup);
    }
    return ISP_BM3D_OK;
}

// Builds the per-image launch arguments of every step (pointers included).
int batch_args(isp_bm3d_batch *b)
{
    isp_bm3d_ctx *ctx = b->ctx;
    const bool mfma = filter_engine(ctx) != 0;
    int kinds[2];
    const int nst = batch_kinds(b, kinds);
    const bool final_only = b->mode == ISP_BM3D_FINAL;
    b->o.clear();
    b->oref.clear();
    for (int s = 0; s < 2; ++s) {
        BatchStepPlan &sp = b->st[s];
        sp.m.clear();
        sp.f.clear();
        sp.fin.clear();
    }
    for (int k = 0; k < b->count; ++k) {
        const isp_bm3d_image &im = b->img[k];
        BatchImg &x = b->bi[k];
        const bool rgb = b->nplanes == 3;
        const bool sep = im.ref[0] && (im.ref[0] != im.src[0]);
        const float *src[3] = {nullptr, nullptr, nullptr};
        // guide planes of a final step run alone: BM3D_Final::RGB2YUV converts
        // the whole guide frame (source/BM3D.cpp:392-408); without one, src
        const float *fguide[3] = {nullptr, nullptr, nullptr};
        if (rgb) {
            OppArgs q{};
            q.W = x.W;
            q.H = x.H;
            q.r = im.src[0];
            q.g = im.src[1];
            q.b = im.src[2];
            q.in_int = 0;
            q.gainY = 1.0 / 3.0;
            q.offsetY = 0.0;
            q.gainU = 1.0 / 2.0;
            q.gainV = 1.0 / 4.0;
            q.offsetC = 0.0;
            q.in_pitch = (size_t)x.P;
            q.y = x.yuv[0];
            q.u = x.yuv[1];
            q.v = x.yuv[2];
            q.pitch = x.P;
            q.y64 = sep ? nullptr : x.y64;
            q.pitch64 = x.P;
            b->o.push_back(q);
            // the guide frame: ConvertToY for step 1 (BM3D_Basic::RGB2YUV,
            // source/BM3D.cpp:320-336), the whole OPP for a final step alone
            OppArgs r = q;
            r.r = sep ? im.ref[0] : im.src[0];
            r.g = sep ? im.ref[1] : im.src[1];
            r.b = sep ? im.ref[2] : im.src[2];
            r.y = r.u = r.v = nullptr;
            if (final_only && sep) {
                r.y = x.refyuv[0];
                r.u = x.refyuv[1];
                r.v = x.refyuv[2];
            }
            r.y64 = sep ? x.y64 : nullptr;
            b->oref.push_back(r);
            for (int c = 0; c < 3; ++c) {
                src[c] = x.yuv[c];
                fguide[c] = sep && final_only ? x.refyuv[c] : x.yuv[c];
            }
        } else {
            src[0] = im.src[0];
            fguide[0] = sep ? im.ref[0] : im.src[0];
        }
        for (int j = 0; j < nst; ++j) {
            const int kind = kinds[j];
            BatchStepPlan &sp = b->st[kind];
            const bool fin = kind == 1;
            const isp_bm3d_step_params &p = fin ? x.prm.final_ : x.prm.basic;
            const Derived &d = x.d[kind];
            // step input / guides / output (source/BM3D.cpp:501-527, 464-472, 392-408)
            const float *guide[3] = {src[0], src[1], src[2]};
            const void *mg;
            if (fin && !final_only) {
                for (int c = 0; c < 3; ++c) guide[c] = x.basic[c];
                mg = x.basic[0];
            } else if (fin) {
                for (int c = 0; c < 3; ++c) guide[c] = fguide[c];
                mg = rgb ? (const void *)x.y64 : (const void *)fguide[0];
            } else if (rgb) {
                mg = x.y64;
            } else {
                mg = sep ? im.ref[0] : im.src[0];
            }
            StepIO io{};
            io.nplanes = b->nplanes;
            io.pitch = x.P;
            for (int c = 0; c < 3; ++c) io.src[c] = src[c];
            const bool last = j == nst - 1;
            io.out_pitch = (size_t)x.P;
            if (last) {
                for (int c = 0; c < 3; ++c) io.out[c] = im.dst[c];
                if (rgb) {
                    // MatrixConvert_YUV2RGB OPP with clip (include/Conversion.hpp:660-762)
                    io.to_rgb = 1;
                    io.clip = 1;
                    io.gain_y = io.gain_c = 1.0;
                    io.offset = 0.0;
                    io.lo = 0.0;
                    io.hi = 1.0;
                }
            } else {
                for (int c = 0; c < 3; ++c) io.out[c] = x.basic[c];
            }
            sp.fin.push_back(build_finalize(x.W, x.H, io, d, x.num, x.den));
            if (d.degenerate) {
                sp.m.push_back(MatchArgs{});
                sp.f.push_back(FilterArgs{});
                continue;
            }
            const MatchArgs ma = build_match(ctx, p, d, x.W, x.H, mg, x.P, x.tiles.match_x, x.tiles.match_y,
                                             x.cnt[kind], x.pos[kind], nullptr);
            int rc = check_match_lds(ctx, p, ma);
            if (rc) return rc;
            sp.m.push_back(ma);
            const FilterArgs fa =
                build_filter(ctx, p, d, x.W, x.H, x.P, src, guide, x.num, x.den, x.cnt[kind], x.pos[kind],
                             mfma ? x.tiles.mfma_x : x.tiles.filter_x, mfma ? x.tiles.mfma_y : x.tiles.filter_y);
            if (filter_lds(ctx, fa, fin) > 160 * 1024)
                return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "image %d needs %zu bytes of LDS for filtering", k,
                            filter_lds(ctx, fa, fin));
            sp.f.push_back(fa);
        }
    }
    // one matching kernel per batched launch: the column walk only when every
    // image of the step can take it
    for (int kind = 0; kind < 2; ++kind) {
        BatchStepPlan &sp = b->st[kind];
        bool all = true;
        for (size_t k = 0; k < sp.m.size(); ++k)
            if (!b->bi[k].d[kind].degenerate && !sp.m[k].walk) all = false;
        for (size_t k = 0; k < sp.m.size(); ++k)
            if (!all && sp.m[k].walk) {
                const BatchImg &x = b->bi[k];
                sp.m[k].walk = 0;
                sp.m[k].tilePitch = match_pitch(sp.m[k], x.d[kind], kind == 1 ? x.prm.final_ : x.prm.basic);
            }
    }
    return ISP_BM3D_OK;
}

// Tile maps, launch headers and the device copies of the argument arrays.
int batch_upload(isp_bm3d_batch *b)
{
    isp_bm3d_ctx *ctx = b->ctx;
    int kinds[2];
    const int nst = batch_kinds(b, kinds);
    for (int j = 0; j < nst; ++j) {
        const int s = kinds[j];
        BatchStepPlan &sp = b->st[s];
        sp.mt.clear();
        sp.ft.clear();
        sp.any = false;
        size_t lm = 0, lf = 0;
        for (int k = 0; k < b->count; ++k) {
            if (b->bi[k].d[s].degenerate) continue;
            const MatchArgs &m = sp.m[k];
            const FilterArgs &f = sp.f[k];
            const int mx = (m.ngx + m.trx - 1) / m.trx, my = (m.ngy + m.try_ - 1) / m.try_;
            const int fx = (f.ngx + f.trx - 1) / f.trx, fy = (f.ngy + f.try_ - 1) / f.try_;
            if (mx > kMaxBatchTiles || my > kMaxBatchTiles || fx > kMaxBatchTiles || fy > kMaxBatchTiles)
                return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "image %d: more than %d tiles along an axis", k,
                            kMaxBatchTiles);
            for (int ty = 0; ty < my; ++ty)
                for (int tx = 0; tx < mx; ++tx)
                    sp.mt.push_back(((uint32_t)k << kBatchImageShift) | ((uint32_t)ty << 10) | (uint32_t)tx);
            for (int ty = 0; ty < fy; ++ty)
                for (int tx = 0; tx < fx; ++tx)
                    sp.ft.push_back(((uint32_t)k << kBatchImageShift) | ((uint32_t)ty << 10) | (uint32_t)tx);
            const size_t a = match_lds_bytes(m), c = filter_lds(ctx, f, s == 1);
            if (!sp.any || a > lm) {
                sp.hm = m;
                lm = a;
            }
            if (!sp.any || c > lf) {
                sp.hf = f;
                lf = c;
            }
            sp.any = true;
        }
        sp.hfin = sp.fin[0];
        b->max_wg = std::max<unsigned>(b->max_wg, (unsigned)std::max(sp.mt.size(), sp.ft.size()));
    }
    // device copies
    auto put = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) HIP_CHECK(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return ISP_BM3D_OK;
    };
    int rc;
    for (int j = 0; j < nst; ++j) {
        BatchStepPlan &sp = b->st[kinds[j]];
        if ((rc = put(sp.dm, sp.m.data(), sp.m.size() * sizeof(MatchArgs)))) return rc;
        if ((rc = put(sp.df, sp.f.data(), sp.f.size() * sizeof(FilterArgs)))) return rc;
        if ((rc = put(sp.dfin, sp.fin.data(), sp.fin.size() * sizeof(FinalizeArgs)))) return rc;
        if ((rc = put(sp.dmt, sp.mt.data(), sp.mt.size() * 4))) return rc;
        if ((rc = put(sp.dft, sp.ft.data(), sp.ft.size() * 4))) return rc;
    }
    if (b->nplanes == 3) {
        if ((rc = put(b->dop, b->o.data(), b->o.size() * sizeof(OppArgs)))) return rc;
        if ((rc = put(b->dopref, b->oref.data(), b->oref.size() * sizeof(OppArgs)))) return rc;
        if ((rc = put(b->drows_ref, b->rows_ref.data(), b->rows_ref.size() * 4))) return rc;
        b->ho = b->o[0];
    }
    return put(b->drows, b->rows.data(), b->rows.size() * 4);
}

// All launches of one run, in stream order on s.
int batch_enqueue(isp_bm3d_batch *b, hipStream_t s)
{
    isp_bm3d_ctx *ctx = b->ctx;
    int n = 0;
    const unsigned nrows = (unsigned)b->rows.size();
    if (b->nplanes == 3) {
        // OPP of every image, then ConvertToY of the guides that differ from src
        BatchLaunch bl{b->drows, b->dop, nrows, b->maxW};
        HIP_CHECK(ctx, launch_opp(b->ho, s, &bl));
        ++n;
        if (!b->rows_ref.empty()) {
            BatchLaunch bg{b->drows_ref, b->dopref, (unsigned)b->rows_ref.size(), b->maxW};
            HIP_CHECK(ctx, launch_opp(b->ho, s, &bg));
            ++n;
        }
    }
    int kinds[2];
    const int nst = batch_kinds(b, kinds);
    for (int j = 0; j < nst; ++j) {
        BatchStepPlan &sp = b->st[kinds[j]];
        const bool fin = kinds[j] == 1;
        // FP64 guides: the RGB luma of step 1, and of the guide frame of a final step alone
        const bool g64 = b->nplanes == 3 && (!fin || b->mode == ISP_BM3D_FINAL);
        if (sp.any) {
            HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(b->acc), b->acc_words * 2, 0.f, s));
            BatchLaunch bm{sp.dmt, sp.dm, (unsigned)sp.mt.size(), 0};
            HIP_CHECK(ctx, launch_match(sp.hm, g64, s, &bm));
            BatchLaunch bf{sp.dft, sp.df, (unsigned)sp.ft.size(), 0};
            HIP_CHECK(ctx, launch_filter_any(ctx, sp.hf, fin, s, &bf));
            n += 3;
        }
        BatchLaunch bz{b->drows, sp.dfin, nrows, b->maxW};
        HIP_CHECK(ctx, launch_finalize(sp.hfin, s, &bz));
        ++n;
    }
    b->launches = n;
    return ISP_BM3D_OK;
}

// Device memory of a batch: accumulators of all images first (one zero-fill
// span), then per-image planes and groups, then the argument arrays and maps.
size_t batch_layout(isp_bm3d_batch *b, void *base)
{
    Carver c(base);
    int kinds[2];
    const int nst = batch_kinds(b, kinds);
    bool used_kind[2] = {false, false};
    for (int j = 0; j < nst; ++j) used_kind[kinds[j]] = true;
    size_t acc0 = 0;
    for (int k = 0; k < b->count; ++k) {
        BatchImg &x = b->bi[k];
        const size_t np = (size_t)x.P * x.H;
        for (int pl = 0; pl < 3; ++pl) {
            x.num[pl] = pl < b->nplanes ? c.take<unsigned long long>(np) : nullptr;
            x.den[pl] = pl < b->nplanes ? c.take<unsigned long long>(np) : nullptr;
            if (k == 0 && pl == 0) acc0 = (size_t)((char *)x.num[0] - (char *)base);
        }
    }
    b->acc = reinterpret_cast<unsigned long long *>(base ? (char *)base + acc0 : nullptr);
    b->acc_words = (c.off - acc0 + 7) / 8;
    size_t mt[2] = {0, 0}, ft[2] = {0, 0};
    for (int k = 0; k < b->count; ++k) {
        BatchImg &x = b->bi[k];
        const size_t np = (size_t)x.P * x.H;
        const bool sep = b->img[k].ref[0] && b->img[k].ref[0] != b->img[k].src[0];
        for (int pl = 0; pl < 3; ++pl) {
            x.yuv[pl] = b->nplanes == 3 ? c.take<float>(np) : nullptr;
            x.refyuv[pl] = (b->nplanes == 3 && b->mode == ISP_BM3D_FINAL && sep) ? c.take<float>(np) : nullptr;
            x.basic[pl] = (pl < b->nplanes && b->mode == ISP_BM3D_BOTH) ? c.take<float>(np) : nullptr;
        }
        x.y64 = b->nplanes == 3 ? c.take<double>(np) : nullptr;
        for (int s = 0; s < 2; ++s) {
            const Derived &d = x.d[s];
            const bool used = used_kind[s] && !d.degenerate;
            const size_t refs = used ? (size_t)d.ngx * d.ngy : 0;
            const int G = s ? x.prm.final_.group_size : x.prm.basic.group_size;
            x.cnt[s] = used ? c.take<int>(refs) : nullptr;
            x.pos[s] = used ? c.take<int>(refs * (size_t)std::max(G, 1)) : nullptr;
            if (used) {
                const bool mf = filter_engine(b->ctx) != 0;
                const int tfx = mf ? x.tiles.mfma_x : x.tiles.filter_x, tfy = mf ? x.tiles.mfma_y : x.tiles.filter_y;
                mt[s] += (size_t)((d.ngx + x.tiles.match_x - 1) / x.tiles.match_x) *
                         ((d.ngy + x.tiles.match_y - 1) / x.tiles.match_y);
                ft[s] += (size_t)((d.ngx + tfx - 1) / tfx) * ((d.ngy + tfy - 1) / tfy);
            }
        }
    }
    for (int j = 0; j < nst; ++j) {
        const int s = kinds[j];
        BatchStepPlan &sp = b->st[s];
        sp.dm = c.take<MatchArgs>(b->count);
        sp.df = c.take<FilterArgs>(b->count);
        sp.dfin = c.take<FinalizeArgs>(b->count);
        sp.dmt = c.take<uint32_t>(std::max<size_t>(mt[s], 1));
        sp.dft = c.take<uint32_t>(std::max<size_t>(ft[s], 1));
    }
    b->dop = b->nplanes == 3 ? c.take<OppArgs>(b->count) : nullptr;
    b->dopref = b->nplanes == 3 ? c.take<OppArgs>(b->count) : nullptr;
    b->drows = c.take<uint32_t>(b->rows.size());
    b->drows_ref = c.take<uint32_t>(std::max<size_t>(b->rows_ref.size(), 1));
    return c.off + 256;
}

int batch_build(isp_bm3d_batch *b, const isp_bm3d_image *images)
{
    isp_bm3d_ctx *ctx = b->ctx;
    batch_free(b);
    b->img.assign(images, images + b->count);
    b->bi.assign(b->count, BatchImg{});
    b->rows.clear();
    b->rows_ref.clear();
    b->maxW = 0;
    b->max_wg = 0;
    for (int k = 0; k < b->count; ++k) {
        const isp_bm3d_image &im = b->img[k];
        for (int c = 0; c < b->nplanes; ++c)
            if (!im.dst[c] || !im.src[c]) return fail(ctx, ISP_BM3D_E_ARG, "image %d: null plane %d", k, c);
        if (b->nplanes == 3 && im.ref[0] && !(im.ref[1] && im.ref[2]))
            return fail(ctx, ISP_BM3D_E_ARG, "image %d: ref needs all three planes", k);
        if (im.width <= 0 || im.height <= 0) return fail(ctx, ISP_BM3D_E_SIZE, "image %d is empty", k);
        if (im.stride < (size_t)im.width || im.stride > ((size_t)1 << 30))
            return fail(ctx, ISP_BM3D_E_ARG, "image %d: bad stride %zu", k, im.stride);
        BatchImg &x = b->bi[k];
        x.W = im.width;
        x.H = im.height;
        x.P = (int)im.stride;
        int rc = batch_params(b, k);
        if (rc) return rc;
        b->maxW = std::max(b->maxW, x.W);
        const bool sep = b->nplanes == 3 && im.ref[0] && im.ref[0] != im.src[0];
        for (int y = 0; y < x.H; ++y) {
            b->rows.push_back(((uint32_t)k << kBatchImageShift) | (uint32_t)y);
            if (sep) b->rows_ref.push_back(((uint32_t)k << kBatchImageShift) | (uint32_t)y);
        }
    }
    if (b->rows.size() > (size_t)kMaxBatchRows)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "a batch holds at most %d rows in all", kMaxBatchRows);
    const size_t bytes = batch_layout(b, nullptr);
    HIP_CHECK(ctx, hipMalloc(&b->mem, bytes));
    b->mem_bytes = bytes;
    batch_layout(b, b->mem);
    int rc = batch_args(b);
    if (!rc) rc = batch_upload(b);
    if (rc) return rc;
    HIP_CHECK(ctx, hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    HIP_CHECK(ctx, hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
    rc = batch_enqueue(b, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIP_CHECK(ctx, e);
    b->graph = g;
    HIP_CHECK(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return ISP_BM3D_OK;
}

int batch_make(isp_bm3d_ctx *ctx, isp_bm3d_batch **out, const isp_bm3d_image *images, int count, int nplanes,
               int mode, int flags, const int *ov_match, const int *ov_filter, int ov_group)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (!out || !images) return fail(ctx, ISP_BM3D_E_ARG, "null argument");
    *out = nullptr;
    if (count < 1 || count > kMaxBatchImages)
        return fail(ctx, ISP_BM3D_E_ARG, "batch of %d images (1..%d)", count, kMaxBatchImages);
    if (nplanes != 1 && nplanes != 3) return fail(ctx, ISP_BM3D_E_ARG, "nplanes %d (1 or 3)", nplanes);
    if (mode < ISP_BM3D_BASIC || mode > ISP_BM3D_BOTH) return fail(ctx, ISP_BM3D_E_ARG, "bad mode %d", mode);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    isp_bm3d_batch *b = new isp_bm3d_batch;
    b->ctx = ctx;
    b->count = count;
    b->nplanes = nplanes;
    b->mode = mode;
    b->flags = flags;
    if (ov_match) {
        b->ov_match[0] = ov_match[0];
        b->ov_match[1] = ov_match[1];
    }
    if (ov_filter) {
        b->ov_filter[0] = ov_filter[0];
        b->ov_filter[1] = ov_filter[1];
    }
    b->ov_group = ov_group;
    const int rc = batch_build(b, images);
    if (rc) {
        batch_free(b);
        delete b;
        return rc;
    }
    *out = b;
    return ISP_BM3D_OK;
}

// Seeded synthetic grey image of the measurement workloads: a gradient,
// 20 flat rectangles and 4 gratings, clipped to [0,1].
struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u1 = 1.0 - uni(), u2 = uni();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

void synth_clean(std::vector<float> &img, int W, int H, uint64_t seed)
{
    SplitMix r{seed};
    img.assign((size_t)W * H, 0.f);
    const double gx = r.uni() * 0.6 - 0.3, gy = r.uni() * 0.6 - 0.3;
    std::vector<double> v((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) v[(size_t)y * W + x] = 0.25 + gx * ((double)x / W - 0.5) + gy * ((double)y / H - 0.5);
    for (int k = 0; k < 20; ++k) {
        const int x0 = (int)(r.uni() * W), y0 = (int)(r.uni() * H);
        const int w = (int)(W / 32.0 + r.uni() * (W / 4.0 - W / 32.0)), h = (int)(H / 32.0 + r.uni() * (H / 4.0 - H / 32.0));
        const double val = r.uni();
        for (int y = y0; y < std::min(H, y0 + h); ++y)
            for (int x = x0; x < std::min(W, x0 + w); ++x) v[(size_t)y * W + x] = val;
    }
    for (int k = 0; k < 4; ++k) {
        const double a = 0.05 + 0.15 * r.uni(), T = 4.0 + 28.0 * r.uni(), th = 3.141592653589793 * r.uni();
        const double ph = 6.283185307179586 * r.uni(), cx = std::cos(th), sy = std::sin(th);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) v[(size_t)y * W + x] += a * std::sin(6.283185307179586 * (x * cx + y * sy) / T + ph);
    }
    for (size_t i = 0; i < v.size(); ++i) img[i] = (float)std::min(1.0, std::max(0.0, v[i]));
}

double psnr01(const float *a, const float *b, size_t n)
{
    double se = 0;
    for (size_t i = 0; i < n; ++i) {
        const double d = (double)a[i] - (double)b[i];
        se += d * d;
    }
    return se > 0 ? 10.0 * std::log10((double)n / se) : 200.0;
}

// Milliseconds per run of a batch (graph replay), best of `reps` after a warm-up.
int batch_time(isp_bm3d_ctx *ctx, isp_bm3d_batch *b, hipStream_t s, int reps, double *ms)
{
    hipEvent_t e0, e1;
    HIP_CHECK(ctx, hipEventCreate(&e0));
    HIP_CHECK(ctx, hipEventCreate(&e1));
    double best = 1e30;
    hipError_t err = hipGraphLaunch(b->exec, s);
    for (int r = 0; r < reps && err == hipSuccess; ++r) {
        (void)hipEventRecord(e0, s);
        err = hipGraphLaunch(b->exec, s);
        (void)hipEventRecord(e1, s);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float t = 0;
        if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
        best = std::min(best, (double)t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_CHECK(ctx, err);
    *ms = best;
    return ISP_BM3D_OK;
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
    if (const char *mw = std::getenv("ISP_BM3D_MATCH_WALK")) ctx->match_walk = std::strcmp(mw, "0") != 0;
    if (const char *pv = std::getenv("ISP_BM3D_FP32_VALU")) ctx->plain_valu = std::strcmp(pv, "0") != 0;
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
    for (int k = 0; k < 3; ++k) {
        std::vector<uint16_t> bits(mt.size());
        for (size_t i = 0; i < mt.size(); ++i)
            bits[i] = k == 0 ? to_f16(mt[i]) : k == 1 ? to_bf16(mt[i]) : to_f16(mt[i] - (float)(_Float16)mt[i]);
        if ((e = hipMalloc(&ctx->d_mfma[k], bits.size() * 2)) != hipSuccess ||
            (e = hipMemcpy(ctx->d_mfma[k], bits.data(), bits.size() * 2, hipMemcpyHostToDevice)) != hipSuccess)
            break;
    }
    if (e != hipSuccess) {
        int rc = fail(nullptr, ISP_BM3D_E_HIP, "create: %s", hipGetErrorString(e));
        (void)hipFree(ctx->d_tables);
        (void)hipFree(ctx->d_tables64);
        for (int k = 0; k < 3; ++k)
            if (ctx->d_mfma[k]) (void)hipFree(ctx->d_mfma[k]);
        delete ctx;
        return rc;
    }
    // transform tables of the generic kernels, for a step whose BlockSize is
    // not 8 or whose GroupSize is above 32 (the reference plans a transform
    // for every group length of every step, source/BM3D.cpp:9-24)
    for (int k = 0; k < 2; ++k) {
        const isp_bm3d_step_params &q = k ? ctx->params.final_ : ctx->params.basic;
        const int B = q.block_size, G = q.group_size;
        if (!(B != kBlock || G > kMaxGroup) || B < 1 || B > kGenMaxBlock || G < 1 || G > kGenMaxGroup) continue;
        std::vector<double> gt;
        auto dct = [&](int n) {
            for (int r = 0; r < n; ++r)
                for (int c = 0; c < n; ++c)
                    gt.push_back((double)((r == 0 ? std::sqrt(1.0L / n) : std::sqrt(2.0L / n)) *
                                          std::cos(PI * (2 * c + 1) * r / (2.0L * n))));
        };
        dct(B);
        for (int N = 1; N <= G; ++N) dct(N);
        if ((e = hipMalloc(&ctx->d_gen[k], gt.size() * sizeof(double))) != hipSuccess ||
            (e = hipMemcpy(ctx->d_gen[k], gt.data(), gt.size() * sizeof(double), hipMemcpyHostToDevice)) !=
                hipSuccess) {
            int rc = fail(nullptr, ISP_BM3D_E_HIP, "create: %s", hipGetErrorString(e));
            isp_bm3d_destroy(ctx);
            return rc;
        }
    }
    // the scratch array's address on this device (the device is current):
    // the allocation-free entry point then needs no runtime lookup
    if (scratch_address(device, &ctx->scratch) != hipSuccess) ctx->scratch = nullptr;
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
    if (ctx->d_mfma[2]) (void)hipFree(ctx->d_mfma[2]);
    for (int k = 0; k < 2; ++k) {
        if (ctx->d_mfma[k]) (void)hipFree(ctx->d_mfma[k]);
        if (ctx->d_gen[k]) (void)hipFree(ctx->d_gen[k]);
        if (ctx->d_var[k]) (void)hipFree(ctx->d_var[k]);
    }
    delete ctx;
    return ISP_BM3D_OK;
}

int isp_bm3d_set_variant(isp_bm3d_ctx *ctx, int flags, double kaiser_beta)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (flags & ~(ISP_BM3D_VARIANT_HAAR | ISP_BM3D_VARIANT_KAISER))
        return fail(ctx, ISP_BM3D_E_ARG, "unknown variant flags %d", flags);
    if ((flags & ISP_BM3D_VARIANT_KAISER) && !(kaiser_beta >= 0 && kaiser_beta <= 100))
        return fail(ctx, ISP_BM3D_E_ARG, "Kaiser beta %g outside 0..100", kaiser_beta);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    HIP_CHECK(ctx, hipDeviceSynchronize());
    for (int k = 0; k < 2; ++k) {
        if (ctx->d_var[k]) HIP_CHECK(ctx, hipFree(ctx->d_var[k]));
        ctx->d_var[k] = nullptr;
    }
    ctx->variant = 0;
    ctx->kaiser_beta = 0.0;
    if (!flags) return ISP_BM3D_OK;
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < 2; ++k) {
        const isp_bm3d_step_params &q = k ? ctx->params.final_ : ctx->params.basic;
        const int B = q.block_size, G = q.group_size;
        if (B < 1 || B > kGenMaxBlock || G < 1 || G > kGenMaxGroup) continue;  // degenerate or refused later
        std::vector<double> t;
        auto dct = [&](int n) {
            for (int r = 0; r < n; ++r)
                for (int c = 0; c < n; ++c)
                    t.push_back((double)((r == 0 ? std::sqrt(1.0L / n) : std::sqrt(2.0L / n)) *
                                         std::cos(PI * (2 * c + 1) * r / (2.0L * n))));
        };
        // orthonormal Haar matrix: H_1 = [1], H_2m = [H_m (x) [1 1]; I_m (x) [1 -1]] / sqrt 2
        auto haar = [&](int n) {
            std::vector<long double> h(1, 1.0L);
            for (int m = 1; m < n; m *= 2) {
                std::vector<long double> g(4 * m * m, 0.0L);
                const long double r = 1.0L / std::sqrt(2.0L);
                for (int i = 0; i < m; ++i)
                    for (int c = 0; c < m; ++c) g[i * 2 * m + 2 * c] = g[i * 2 * m + 2 * c + 1] = h[i * m + c] * r;
                for (int i = 0; i < m; ++i) {
                    g[(m + i) * 2 * m + 2 * i] = r;
                    g[(m + i) * 2 * m + 2 * i + 1] = -r;
                }
                h.swap(g);
            }
            for (long double v : h) t.push_back((double)v);
        };
        dct(B);
        for (int N = 1; N <= G; ++N) {
            if ((flags & ISP_BM3D_VARIANT_HAAR) && (N & (N - 1)) == 0) haar(N);
            else dct(N);
        }
        ctx->var_kaiser_off[k] = t.size();
        // Kaiser window: I0(beta sqrt(1 - s^2)) / I0(beta), s = (2i - (B - 1)) / (B - 1)
        auto i0 = [](long double x) {
            long double sum = 1.0L, term = 1.0L;
            for (int m = 1; m < 200; ++m) {
                term *= (x / (2.0L * m)) * (x / (2.0L * m));
                sum += term;
                if (term < 1e-21L * sum) break;
            }
            return sum;
        };
        std::vector<long double> kw(B, 1.0L);
        if (flags & ISP_BM3D_VARIANT_KAISER)
            for (int i = 0; i < B && B > 1; ++i) {
                const long double sv = (2.0L * i - (B - 1)) / (long double)(B - 1);
                kw[i] = i0(kaiser_beta * std::sqrt(std::max(0.0L, 1.0L - sv * sv))) / i0(kaiser_beta);
            }
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) t.push_back((double)(kw[y] * kw[x]));
        HIP_CHECK(ctx, hipMalloc(&ctx->d_var[k], t.size() * sizeof(double)));
        HIP_CHECK(ctx, hipMemcpy(ctx->d_var[k], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    ctx->variant = flags;
    ctx->kaiser_beta = kaiser_beta;
    return ISP_BM3D_OK;
}

int isp_bm3d_set_precision(isp_bm3d_ctx *ctx, int precision)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (precision < ISP_BM3D_PREC_FP32 || precision > ISP_BM3D_PREC_FP32_MFMA)
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

int isp_bm3d_aggregation_limits(const isp_bm3d_step_params *p, int radius, double *max_contributions,
                                double *weight_cap)
{
    if (!p || !max_contributions || !weight_cap || radius < 0 || p->block_step <= 0 || p->bm_step <= 0 ||
        p->bm_range < 0 || p->group_size < 1)
        return ISP_BM3D_E_ARG;
    const int range = p->bm_range / p->bm_step * p->bm_step;
    *max_contributions = aggregation_bound(kBlock, p->block_step, range, p->bm_step, p->group_size, radius);
    *weight_cap = wiener_weight_cap(kBlock, p->block_step, range, p->bm_step, p->group_size, radius);
    return *weight_cap > 0 ? ISP_BM3D_OK : ISP_BM3D_E_UNSUPPORTED;
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

// Work of one float device plane carved from the scratch array: the step-1
// estimate, one plane of num / den (adjacent, cleared by one fill), and one
// group list shared by the two steps (step 2 matches after step 1's filter
// has read its groups, in stream order).  Returns the bytes needed.
size_t noalloc_layout(Work *w, void *base, int P, int H, size_t refs, int G)
{
    Carver c(base);
    const size_t np = (size_t)P * H;
    w->basic[0] = c.take<float>(np);
    w->num[0] = c.take<unsigned long long>(np);
    w->den[0] = c.take<unsigned long long>(np);
    for (int k = 1; k < 3; ++k) {
        // inactive planes: never read or written with one plane; the fill
        // of run_step clears [num[0], den[2] + P * H) = num[0] and den[0]
        w->basic[k] = w->basic[0];
        w->num[k] = w->num[0];
        w->den[k] = w->den[0];
    }
    w->cnt[0] = w->cnt[1] = c.take<int>(refs);
    w->pos[0] = w->pos[1] = c.take<int>(refs * (size_t)G);
    return c.off;
}

int plane_noalloc(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref, int W, int H,
                  size_t stride, int mode, void *stream)
{
    int rc = check_common(ctx, W, H, stride, mode, ISP_BM3D_MEM_DEVICE);
    if (rc) return rc;
    if (!dst || !src) return fail(ctx, ISP_BM3D_E_ARG, "null dst/src");
    if (!ctx->scratch)
        return fail(ctx, ISP_BM3D_E_HIP, "the scratch array's address was not found on device %d", ctx->device);
    const int P = (int)stride;
    size_t refs = 1;
    for (const isp_bm3d_step_params *p : {&ctx->params.basic, &ctx->params.final_}) {
        if (p->block_step <= 0) continue;
        const int B = p->block_size > 0 ? p->block_size : kBlock;
        const size_t ngx = (size_t)std::max(W - B, 0) / p->block_step + 2;
        const size_t ngy = (size_t)std::max(H - B, 0) / p->block_step + 2;
        refs = std::max(refs, ngx * ngy);
    }
    const int G = max_group(ctx);
    Work w{};
    const size_t need = noalloc_layout(&w, nullptr, P, H, refs, G);
    if (need > kScratchBytes)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED,
                    "a %dx%d plane (stride %zu) needs %zu bytes of scratch; the allocation-free entry point "
                    "has %zu",
                    W, H, stride, need, kScratchBytes);
    noalloc_layout(&w, ctx->scratch, P, H, refs, G);
    // stage timers would create events: off for this call
    struct TimingOff {
        isp_bm3d_ctx *c;
        bool was;
        ~TimingOff() { c->timing = was; }
    } off{ctx, ctx->timing};
    ctx->timing = false;
    hipStream_t s = (hipStream_t)stream;
    const float *fref = ref ? ref : src;
    StepIO io{};
    io.nplanes = 1;
    io.pitch = P;
    io.pitch64 = P;
    io.src[0] = src;
    if (mode == ISP_BM3D_BASIC || mode == ISP_BM3D_FINAL) {
        const bool fin = mode == ISP_BM3D_FINAL;
        io.guide[0] = fref;
        io.mguide32 = fref;
        set_plain_out(io, dst, stride);
        return run_step(ctx, fin ? ctx->params.final_ : ctx->params.basic, fin, io, w, W, H, s);
    }
    io.mguide32 = fref;
    io.guide[0] = fref;
    set_plain_out(io, w.basic[0], (size_t)P);
    if ((rc = run_step(ctx, ctx->params.basic, false, io, w, W, H, s))) return rc;
    io.mguide32 = w.basic[0];
    io.guide[0] = w.basic[0];
    set_plain_out(io, dst, stride);
    return run_step(ctx, ctx->params.final_, true, io, w, W, H, s);
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

int isp_bm3d_plan(const isp_bm3d_params *params, int W, int H, int nplanes, isp_bm3d_plan_info *info)
{
    if (!params || !info || (nplanes != 1 && nplanes != 3))
        return fail(nullptr, ISP_BM3D_E_ARG, "plan: null argument or nplanes %d", nplanes);
    if (W <= 0 || H <= 0) return fail(nullptr, ISP_BM3D_E_SIZE, "plan: empty image %dx%d", W, H);
    std::memset(info, 0, sizeof(*info));
    // a context holding only the parameters: derive() reads them and whether
    // generic tables would exist (their contents are never read here)
    isp_bm3d_ctx c;
    c.params = *params;
    static const double no_table[1] = {0.0};
    for (int k = 0; k < 2; ++k) {
        const isp_bm3d_step_params &q = k ? c.params.final_ : c.params.basic;
        if ((q.block_size != kBlock || q.group_size > kMaxGroup) && q.block_size >= 1 &&
            q.block_size <= kGenMaxBlock && q.group_size >= 1 && q.group_size <= kGenMaxGroup)
            c.d_gen[k] = const_cast<double *>(no_table);
    }
    for (int k = 0; k < 2; ++k) {
        Derived d{};
        const int rc = derive(&c, k ? c.params.final_ : c.params.basic, nplanes, W, H, d);
        c.d_gen[k] = nullptr;
        if (rc) {
            g_create_err = c.err;
            for (int j = 0; j < 2; ++j) c.d_gen[j] = nullptr;
            return rc;
        }
        info->degenerate[k] = d.degenerate;
        info->generic[k] = !d.degenerate && d.generic;
        info->ngx[k] = d.degenerate ? 0 : d.ngx;
        info->ngy[k] = d.degenerate ? 0 : d.ngy;
        info->weight_cap[k] = d.degenerate ? 0.0 : (double)d.wmax;
    }
    for (int j = 0; j < 2; ++j) c.d_gen[j] = nullptr;
    size_t refs = 1;
    for (const isp_bm3d_step_params *p : {&c.params.basic, &c.params.final_}) {
        const int s = p->block_step > 0 ? p->block_step : 1;
        refs = std::max(refs, (size_t)(W / s + 2) * (size_t)(H / s + 2));
    }
    Work w{};
    info->workspace_bytes = layout(&w, nullptr, ((W + 63) / 64) * 64, H, refs, max_group(&c), false);
    size_t nrefs = 1;
    for (const isp_bm3d_step_params *p : {&c.params.basic, &c.params.final_}) {
        if (p->block_step <= 0) continue;
        const int B = p->block_size > 0 ? p->block_size : kBlock;
        nrefs = std::max(nrefs, ((size_t)std::max(W - B, 0) / p->block_step + 2) *
                                    ((size_t)std::max(H - B, 0) / p->block_step + 2));
    }
    info->noalloc_bytes = noalloc_layout(&w, nullptr, W, H, nrefs, max_group(&c));
    info->noalloc_capacity = kScratchBytes;
    return ISP_BM3D_OK;
}

int isp_bm3d_plane_f32(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref, int W, int H,
                       size_t stride, int mode, int mem, void *stream)
{
    const Fmt f{false, 0, 0, 1};
    return plane_common(ctx, dst, src, ref, W, H, stride, f, mode, mem, stream);
}

int isp_bm3d_plane_f32_noalloc(isp_bm3d_ctx *ctx, float *dst, const float *src, const float *ref, int W, int H,
                               size_t stride, int mode, void *stream)
{
    return plane_noalloc(ctx, dst, src, ref, W, H, stride, mode, stream);
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


int isp_bm3d_rgb_to_opp(float *y, float *u, float *v, const float *r, const float *g, const float *b, int W, int H,
                        size_t stride, void *stream)
{
    if (!y || !u || !v || !r || !g || !b || W <= 0 || H <= 0 || stride < (size_t)W) return ISP_BM3D_E_ARG;
    OppArgs o{};
    o.W = W;
    o.H = H;
    o.r = r;
    o.g = g;
    o.b = b;
    o.in_int = 0;
    o.gainY = 1.0 / 3.0;
    o.gainU = 1.0 / 2.0;
    o.gainV = 1.0 / 4.0;
    o.in_pitch = stride;
    o.y = y;
    o.u = u;
    o.v = v;
    o.pitch = (int)stride;
    return launch_opp(o, (hipStream_t)stream) == hipSuccess ? ISP_BM3D_OK : ISP_BM3D_E_HIP;
}

int isp_bm3d_opp_to_rgb(float *r, float *g, float *b, const float *y, const float *u, const float *v, int W, int H,
                        size_t stride, int clip, void *stream)
{
    if (!y || !u || !v || !r || !g || !b || W <= 0 || H <= 0 || stride < (size_t)W) return ISP_BM3D_E_ARG;
    return launch_opp_inverse(r, g, b, y, u, v, W, H, stride, clip, (hipStream_t)stream) == hipSuccess
               ? ISP_BM3D_OK
               : ISP_BM3D_E_HIP;
}

int isp_bm3d_batch_create(isp_bm3d_ctx *ctx, isp_bm3d_batch **batch, const isp_bm3d_image *images, int count,
                          int nplanes, int mode, int flags)
{
    return batch_make(ctx, batch, images, count, nplanes, mode, flags, nullptr, nullptr, 0);
}

int isp_bm3d_batch_run(isp_bm3d_batch *b, void *stream, int use_graph)
{
    if (!b) return fail(nullptr, ISP_BM3D_E_ARG, "null batch");
    isp_bm3d_ctx *ctx = b->ctx;
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (use_graph) {
        HIP_CHECK(ctx, hipGraphLaunch(b->exec, (hipStream_t)stream));
        return ISP_BM3D_OK;
    }
    return batch_enqueue(b, (hipStream_t)stream);
}

int isp_bm3d_batch_update(isp_bm3d_batch *b, const isp_bm3d_image *images)
{
    if (!b || !images) return fail(b ? b->ctx : nullptr, ISP_BM3D_E_ARG, "null argument");
    isp_bm3d_ctx *ctx = b->ctx;
    for (int k = 0; k < b->count; ++k)
        if (images[k].width != b->img[k].width || images[k].height != b->img[k].height)
            return fail(ctx, ISP_BM3D_E_ARG, "image %d changed size; create a new batch", k);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    HIP_CHECK(ctx, hipDeviceSynchronize());
    return batch_build(b, images);
}

int isp_bm3d_batch_info(const isp_bm3d_batch *b, int *launches, size_t *device_bytes, int *max_workgroups)
{
    if (!b) return fail(nullptr, ISP_BM3D_E_ARG, "null batch");
    if (launches) *launches = b->launches;
    if (device_bytes) *device_bytes = b->mem_bytes;
    if (max_workgroups) *max_workgroups = (int)b->max_wg;
    return ISP_BM3D_OK;
}

int isp_bm3d_batch_destroy(isp_bm3d_batch *b)
{
    if (!b) return ISP_BM3D_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    delete b;
    return ISP_BM3D_OK;
}

int isp_bm3d_size_class(int width, int height)
{
    if (width <= 0 || height <= 0) return ISP_BM3D_E_ARG;
    return size_class(width, height);
}

int isp_bm3d_set_tuning(isp_bm3d_ctx *ctx, int cls, const isp_bm3d_tuning *t)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (cls < 0 || cls > 62) return fail(ctx, ISP_BM3D_E_ARG, "size class %d", cls);
    if (!t) {
        ctx->tune.erase(cls);
        return ISP_BM3D_OK;
    }
    for (int k = 0; k < 2; ++k)
        if (t->match_tile[k] < 1 || t->filter_tile[k] < 1 || t->match_tile[0] * t->match_tile[1] > kMaxTileRefs ||
            t->filter_tile[0] * t->filter_tile[1] > kMaxTileRefs)
            return fail(ctx, ISP_BM3D_E_ARG, "tile shapes must be positive with at most %d blocks", kMaxTileRefs);
    if (t->final_group_size < 0 || t->final_group_size > kMaxGroup)
        return fail(ctx, ISP_BM3D_E_ARG, "final_group_size %d", t->final_group_size);
    ctx->tune[cls] = *t;
    return ISP_BM3D_OK;
}

int isp_bm3d_get_tuning(const isp_bm3d_ctx *ctx, int cls, isp_bm3d_tuning *t)
{
    if (!ctx || !t) return ISP_BM3D_E_ARG;
    auto it = ctx->tune.find(cls);
    if (it == ctx->tune.end()) return 0;
    *t = it->second;
    return 1;
}

int isp_bm3d_autotune(isp_bm3d_ctx *ctx, int W, int H, double sigma, int flags, isp_bm3d_tuning *out)
{
    if (!ctx) return fail(nullptr, ISP_BM3D_E_ARG, "null context");
    if (W < kBlock || H < kBlock) return fail(ctx, ISP_BM3D_E_SIZE, "image %dx%d", W, H);
    if ((size_t)W * H > ((size_t)1 << 26)) return fail(ctx, ISP_BM3D_E_ARG, "tuning image too large");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int cls = size_class(W, H);
    ctx->tune.erase(cls);
    // the synthetic workload: K noisy copies of one image (RGB: three planes),
    // enough tiles to fill the GPU
    const int np_ = (flags & ISP_BM3D_TUNE_RGB) ? 3 : 1;
    std::vector<float> clean, noisy, plane;
    for (int c = 0; c < np_; ++c) {
        synth_clean(plane, W, H, 0x5EED0000ull + (uint64_t)cls * 8 + (uint64_t)c);
        clean.insert(clean.end(), plane.begin(), plane.end());
    }
    noisy = clean;
    SplitMix rn{0xA5A5ull + (uint64_t)W * 131 + (uint64_t)H};
    const double sg = (sigma > 0 ? sigma : ctx->params.basic.sigma[0]) / 255.0;
    for (auto &v : noisy) v = (float)(v + sg * rn.normal());
    Derived d{};
    int rc = derive(ctx, ctx->params.basic, 1, W, H, d);
    if (rc) return rc;
    if (d.generic)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "autotuning needs BlockSize 8 and GroupSize <= %d", kMaxGroup);
    const long tiles = (long)((d.ngx + 7) / 8) * ((d.ngy + 7) / 8);
    const int K = (int)std::max(1L, std::min(32L, (4096 + tiles - 1) / std::max(1L, tiles)));
    const size_t np = (size_t)W * H;
    float *dmem = nullptr;
    HIP_CHECK(ctx, hipMalloc(&dmem, np * 4 * np_ * (K + 1)));
    struct Guard {
        float *p;
        ~Guard() { (void)hipFree(p); }
    } guard{dmem};
    HIP_CHECK(ctx, hipMemcpy(dmem, noisy.data(), np * 4 * np_, hipMemcpyHostToDevice));
    std::vector<isp_bm3d_image> imgs(K);
    for (int k = 0; k < K; ++k) {
        isp_bm3d_image &im = imgs[k];
        std::memset(&im, 0, sizeof(im));
        for (int c = 0; c < np_; ++c) {
            im.dst[c] = dmem + np * ((size_t)np_ * (k + 1) + c);
            im.src[c] = dmem + np * c;
        }
        im.width = W;
        im.height = H;
        im.stride = (size_t)W;
        im.sigma = sigma;
    }
    hipStream_t s = nullptr;
    HIP_CHECK(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct SGuard {
        hipStream_t s;
        ~SGuard() { (void)hipStreamDestroy(s); }
    } sguard{s};
    auto measure = [&](const int *mt, const int *ft, int g, double *ms, double *psnr) -> int {
        isp_bm3d_batch *b = nullptr;
        int r = batch_make(ctx, &b, imgs.data(), K, np_, ISP_BM3D_BOTH, 0, mt, ft, g);
        if (r) return r;
        r = batch_time(ctx, b, s, 3, ms);
        if (!r && psnr) {
            std::vector<float> outp(np * np_);
            hipError_t e = hipMemcpy(outp.data(), imgs[0].dst[0], np * 4 * np_, hipMemcpyDeviceToHost);
            if (e != hipSuccess) r = fail(ctx, ISP_BM3D_E_HIP, "copy: %s", hipGetErrorString(e));
            else *psnr = psnr01(outp.data(), clean.data(), np * np_);
        }
        (void)isp_bm3d_batch_destroy(b);
        return r;
    };
    const bool mf = filter_engine(ctx) != 0, walk = !exact_matching(ctx);
    const int mcands_v[][2] = {{8, 4}, {8, 8}, {16, 4}, {4, 8}, {16, 8}};
    const int mcands_m[][2] = {{4, 16}, {8, 16}, {4, 8}, {8, 8}, {8, 4}};
    const int fcands_m[][2] = {{8, 8}, {16, 8}, {8, 16}, {16, 16}};
    const int fcands_v[][2] = {{8, 4}, {8, 8}, {16, 4}, {4, 8}};
    int best_m[2] = {walk ? ctx->tiles.walk_x : ctx->tiles.match_x, walk ? ctx->tiles.walk_y : ctx->tiles.match_y};
    int best_f[2] = {mf ? ctx->tiles.mfma_x : ctx->tiles.filter_x, mf ? ctx->tiles.mfma_y : ctx->tiles.filter_y};
    double best = 1e30, psnr_base = 0;
    for (const auto &c : walk ? mcands_m : mcands_v) {
        double ms = 0;
        rc = measure(c, best_f, 0, &ms, nullptr);
        if (rc == ISP_BM3D_E_UNSUPPORTED) continue;
        if (rc) return rc;
        if (ms < best) {
            best = ms;
            best_m[0] = c[0];
            best_m[1] = c[1];
        }
    }
    const int nf = 4;
    for (int i = 0; i < nf; ++i) {
        const int *c = mf ? fcands_m[i] : fcands_v[i];
        double ms = 0;
        rc = measure(best_m, c, 0, &ms, nullptr);
        if (rc == ISP_BM3D_E_UNSUPPORTED) continue;
        if (rc) return rc;
        if (ms < best) {
            best = ms;
            best_f[0] = c[0];
            best_f[1] = c[1];
        }
    }
    rc = measure(best_m, best_f, 0, &best, &psnr_base);
    if (rc) return rc;
    int group = 0;
    double psnr = psnr_base;
    if ((flags & ISP_BM3D_BATCH_AUTO_GROUP) && ctx->params.final_.group_size > 16) {
        double ms16 = 0, p16 = 0;
        rc = measure(best_m, best_f, 16, &ms16, &p16);
        if (rc) return rc;
        if (ms16 < best && psnr_base - p16 <= 0.05) {
            group = 16;
            best = ms16;
            psnr = p16;
        }
    }
    isp_bm3d_tuning t{};
    t.match_tile[0] = best_m[0];
    t.match_tile[1] = best_m[1];
    t.filter_tile[0] = best_f[0];
    t.filter_tile[1] = best_f[1];
    t.final_group_size = group;
    t.ms_per_image = best / K;
    t.psnr_db = psnr;
    ctx->tune[cls] = t;
    if (out) *out = t;
    return ISP_BM3D_OK;
}
}  // extern "C"

// ---------------------------------------------------------------------------
// NL-means (include/NLMeans.h, source/NLMeans.cpp) on the matching kernel.

namespace {

// NLMeans::process_* early return (source/NLMeans.cpp:152-158, 234-240, 263-269)
bool nlm_degenerate(const isp_nlm_params &p)
{
    return p.strength <= 0 || p.group_size == 1 || p.block_size <= 0 || p.bm_range <= 0 ||
           p.bm_range < p.bm_step || p.thMSE <= 0;
}

struct NlmWork {
    double *g64;                  // FP64 matching guide (int forms)
    double *s64;                  // FP64 source plane (int plane form)
    float *stage[6];              // host uploads: src planes, then ref planes (4-byte samples)
    unsigned long long *num[3];   // fixed-point sums
    float *out[3];                // host-mode output staging (4-byte samples)
    int *cnt, *pos;
    double *keys;
};

size_t nlm_layout(NlmWork *w, void *base, int P, int H, size_t refs, int G)
{
    Carver c(base);
    const size_t np = (size_t)P * H;
    w->g64 = c.take<double>(np);
    w->s64 = c.take<double>(np);
    for (int k = 0; k < 6; ++k) w->stage[k] = c.take<float>(np);
    for (int k = 0; k < 3; ++k) w->num[k] = c.take<unsigned long long>(np);
    for (int k = 0; k < 3; ++k) w->out[k] = c.take<float>(np);
    w->cnt = c.take<int>(refs);
    w->pos = c.take<int>(refs * G);
    w->keys = c.take<double>(refs * G);
    return c.off + 256;
}

// The three forms: nplanes 1 (float plane, or int plane when fmt.is_int) or
// 3 (int RGB frame).
int nlm_common(isp_bm3d_ctx *ctx, const isp_nlm_params *pp, int nplanes, void *const *dst,
               const void *const *src, const void *const *ref, int W, int H, size_t stride, const Fmt &fmt,
               int mem, void *stream)
{
    int rc = check_common(ctx, W, H, stride, ISP_BM3D_BASIC, mem);
    if (rc) return rc;
    if (!pp) return fail(ctx, ISP_BM3D_E_ARG, "null NL-means parameters");
    const isp_nlm_params p = *pp;
    for (int c = 0; c < nplanes; ++c)
        if (!dst[c] || !src[c]) return fail(ctx, ISP_BM3D_E_ARG, "null dst/src plane");
    if (fmt.is_int && !(fmt.floor_ < fmt.ceil_ && fmt.floor_ <= fmt.neutral && fmt.neutral <= fmt.ceil_))
        return fail(ctx, ISP_BM3D_E_ARG, "bad quantisation floor %d neutral %d ceil %d", fmt.floor_, fmt.neutral,
                    fmt.ceil_);
    bool has_ref = false, all_ref = true, sep_ref = false;
    for (int c = 0; c < nplanes; ++c) {
        has_ref = has_ref || ref[c];
        all_ref = all_ref && ref[c];
        sep_ref = sep_ref || (ref[c] && ref[c] != src[c]);
    }
    if (has_ref && !all_ref) return fail(ctx, ISP_BM3D_E_ARG, "ref needs every plane");
    hipStream_t s = (hipStream_t)stream;
    const bool host = mem == ISP_BM3D_MEM_HOST;
    if (nlm_degenerate(p)) {
        for (int c = 0; c < nplanes; ++c) {
            if (host) host_copy_plane(dst[c], src[c], stride, W, H);
            else if (dst[c] != src[c] && (rc = copy_plane(ctx, dst[c], stride, src[c], stride, W, H, s)))
                return rc;
        }
        return ISP_BM3D_OK;
    }
    if (p.block_size != kBlock)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "BlockSize %d: the GPU kernels implement BlockSize 8",
                    p.block_size);
    if (p.group_size < 2 || p.group_size > kMaxGroup)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "GroupSize %d outside 2..%d", p.group_size, kMaxGroup);
    if (p.block_step <= 0 || p.bm_step <= 0)
        return fail(ctx, ISP_BM3D_E_ARG, "BlockStep and BMstep must be positive");
    if (W < kBlock || H < kBlock)
        return fail(ctx, ISP_BM3D_E_SIZE, "image %dx%d is smaller than one %dx%d block", W, H, kBlock, kBlock);
    // positions are packed (y << 16) | x in an int32: y must stay below 2^15
    if (W > 65535 || H > 32767)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "image %dx%d: width must be < 65536 and height < 32768", W, H);

    // matching: BlockMatchingMulti with the unscaled thMSE (source/NLMeans.cpp:203, 332)
    isp_bm3d_step_params mp{};
    mp.block_size = kBlock;
    mp.block_step = p.block_step;
    mp.group_size = p.group_size;
    mp.bm_range = p.bm_range;
    mp.bm_step = p.bm_step;
    mp.thMSE = p.thMSE;
    Derived d{};
    d.range = p.bm_range / p.bm_step * p.bm_step;
    d.bypass = false;
    const double MSE2SSE = (double)(kBlock * kBlock) * 1.0 * 1.0 / (double)(255 * 255);
    d.distMul = 1.0 / MSE2SSE;
    d.thSSE = p.thMSE * MSE2SSE;
    d.ngx = (W - kBlock) / p.block_step + 1 + ((W - kBlock) % p.block_step ? 1 : 0);
    d.ngy = (H - kBlock) / p.block_step + 1 + ((H - kBlock) % p.block_step ? 1 : 0);
    const size_t refs = (size_t)d.ngx * d.ngy;

    const int P = internal_pitch(W, stride, mem);
    NlmWork w{};
    if ((rc = ensure_ws(ctx, nlm_layout(&w, nullptr, P, H, refs, p.group_size)))) return rc;
    nlm_layout(&w, ctx->ws, P, H, refs, p.group_size);

    const void *in[3] = {src[0], nplanes > 1 ? src[1] : nullptr, nplanes > 1 ? src[2] : nullptr};
    const void *rin[3] = {ref[0], nplanes > 1 ? ref[1] : nullptr, nplanes > 1 ? ref[2] : nullptr};
    size_t ipitch = stride;
    const HostStage hs{ctx, s, P, W, H};
    if (host) {
        if ((rc = HostStage::reserve(ctx, (size_t)(sep_ref ? 2 : 1) * nplanes * hs.plane_bytes()))) return rc;
        for (int c = 0; c < nplanes; ++c) {
            if ((rc = hs.upload(c, w.stage[c], in[c], stride))) return rc;
            in[c] = w.stage[c];
            if (sep_ref) {
                if ((rc = hs.upload(nplanes + c, w.stage[3 + c], rin[c], stride))) return rc;
                rin[c] = w.stage[3 + c];
            }
        }
        ipitch = (size_t)P;
    }
    if (!sep_ref)
        for (int c = 0; c < nplanes; ++c) rin[c] = in[c];

    NlmArgs a{};
    a.W = W;
    a.H = H;
    a.ngx = d.ngx;
    a.ngy = d.ngy;
    a.bstep = p.block_step;
    a.G = p.group_size;
    a.cnt = w.cnt;
    a.pos = w.pos;
    a.keys = w.keys;
    a.nplanes = nplanes;
    a.correction = p.correction != 0;
    a.exp_mul = -1 / (p.strength * p.strength);
    a.pitch = P;
    for (int c = 0; c < nplanes; ++c) a.num[c] = w.num[c];
    NlmFinalizeArgs fa{};
    fa.W = W;
    fa.H = H;
    fa.pitch = P;
    fa.bstep = p.block_step;
    fa.nplanes = nplanes;
    const void *mguide = nullptr;
    bool g64 = false;
    if (nplanes == 3) {
        // R, G, B averaged as int samples; matching on ConvertToY(ref) with clip
        OppArgs o{};
        o.W = W;
        o.H = H;
        o.r = rin[0];
        o.g = rin[1];
        o.b = rin[2];
        o.in_int = 1;
        const double sRange = (double)(fmt.ceil_ - fmt.floor_);
        o.gainY = 1.0 / (sRange * 3.0);
        o.offsetY = -(double)fmt.floor_ * 3.0 * o.gainY + 0.0;
        o.in_pitch = ipitch;
        o.pitch = P;
        o.y64 = w.g64;
        o.pitch64 = P;
        o.clipY = 1;
        {
            StageTimer tm(ctx, ST_OPP, s);
            HIP_CHECK(ctx, launch_opp(o, s));
        }
        mguide = w.g64;
        g64 = true;
        a.src_type = 2;
        for (int c = 0; c < 3; ++c) a.src[c] = in[c];
        a.spitch = ipitch;
        const double sigma = p.sigma * (double)(fmt.ceil_ - fmt.floor_) / 255;
        a.varN = sigma * sigma;
        // |samples| < 2^(e+5): 64 blocks per pixel stay below 2^63
        const int M = std::max({std::abs(fmt.floor_), std::abs(fmt.ceil_), 1});
        int e = 0;
        while (e < 31 && ((int64_t)1 << e) < (int64_t)M) ++e;
        a.fix = std::ldexp(1.0, 52 - e);
        fa.out_int = 1;
        fa.gain = 1.0;
        fa.offset = 0.5;
        fa.clip = 0;
    } else if (fmt.is_int) {
        // Plane_FL src_data(src): RangeConvert with clip (source/NLMeans.cpp:242-243,
        // include/Conversion.hpp:277-296)
        ToFloatArgs t{};
        t.W = W;
        t.H = H;
        t.spitch = ipitch;
        t.dpitch = P;
        t.gain = 1.0 / (double)(fmt.ceil_ - fmt.floor_);
        t.offset = -(double)(fmt.chroma() ? fmt.neutral : fmt.floor_) * t.gain + 0.0;
        t.clip = 1;
        t.lo = fmt.chroma() ? -0.5 : 0.0;
        t.hi = fmt.chroma() ? 0.5 : 1.0;
        {
            StageTimer tm(ctx, ST_OPP, s);
            t.src = (const int *)in[0];
            t.dst = w.out[2];  // FP32 copy unused
            t.dst64 = w.s64;
            HIP_CHECK(ctx, launch_to_float(t, s));
            mguide = w.s64;
            if (sep_ref) {
                t.src = (const int *)rin[0];
                t.dst64 = w.g64;
                HIP_CHECK(ctx, launch_to_float(t, s));
                mguide = w.g64;
            }
        }
        g64 = true;
        a.src_type = 1;
        a.src[0] = w.s64;
        a.spitch = (size_t)P;
        const double sigma = p.sigma * 1.0 / 255;
        a.varN = sigma * sigma;
        a.fix = std::ldexp(1.0, 52);
        // RangeConvert back, +0.5 (0.499999 for PC-range chroma); values are
        // convex combinations of clipped samples, so the clip changes nothing
        fa.out_int = 1;
        fa.gain = (double)(fmt.ceil_ - fmt.floor_) / 1.0;
        fa.offset = -0.0 * fa.gain + (double)(fmt.chroma() ? fmt.neutral : fmt.floor_) +
                    (fmt.pc_chroma() ? 0.499999 : 0.5);
        fa.clip = 1;
        fa.lo = fmt.floor_;
        fa.hi = fmt.ceil_;
    } else {
        mguide = rin[0];
        g64 = false;
        a.src_type = 0;
        a.src[0] = in[0];
        a.spitch = ipitch;
        const double sigma = p.sigma * 1.0 / 255;
        a.varN = sigma * sigma;
        a.fix = std::ldexp(1.0, 36);
        fa.out_int = 0;
    }
    fa.inv_fix = 1.0 / a.fix;
    const int gpitch = g64 ? P : (int)ipitch;

    const isp_bm3d_ctx::Tiles tl = tiles_for(ctx, W, H);
    const MatchArgs ma = build_match(ctx, mp, d, W, H, mguide, gpitch, tl.match_x, tl.match_y, w.cnt, w.pos,
                                     w.keys);
    if ((rc = check_match_lds(ctx, mp, ma))) return rc;
    {
        StageTimer tm(ctx, ST_NLM_MATCH, s);
        HIP_CHECK(ctx, launch_match(ma, g64, s));
    }
    {
        StageTimer tm(ctx, ST_FILL, s);
        for (int c = 0; c < nplanes; ++c) HIP_CHECK(ctx, launch_fill((float *)w.num[c], (size_t)P * H * 2, 0.0f, s));
    }
    // tile of reference blocks per workgroup: 8 x 8, smaller where the
    // footprint's accumulators would not fit in 64 KiB of LDS
    int t = 8;
    for (;; --t) {
        a.trx = a.try_ = t;
        a.FW = (t - 1) * p.block_step + kBlock;
        a.FH = a.FW;
        if (t == 1 || nlm_lds_bytes(a) <= 65536) break;
    }
    {
        StageTimer tm(ctx, ST_NLM_AVERAGE, s);
        HIP_CHECK(ctx, launch_nlm(a, s));
    }
    void *out[3] = {dst[0], nplanes > 1 ? dst[1] : nullptr, nplanes > 1 ? dst[2] : nullptr};
    fa.out_pitch = stride;
    if (host) {
        for (int c = 0; c < nplanes; ++c) out[c] = w.out[c];
        fa.out_pitch = (size_t)P;
    }
    for (int c = 0; c < nplanes; ++c) {
        fa.num[c] = w.num[c];
        fa.out[c] = out[c];
    }
    {
        StageTimer tm(ctx, ST_FINALIZE, s);
        HIP_CHECK(ctx, launch_nlm_finalize(fa, s));
    }
    if (host) {
        for (int c = 0; c < nplanes; ++c)
            if ((rc = hs.download_start(c, out[c]))) return rc;
        HIP_CHECK(ctx, hipStreamSynchronize(s));
        for (int c = 0; c < nplanes; ++c) hs.download_finish(c, dst[c], stride);
    }
    return ISP_BM3D_OK;
}

}  // namespace

extern "C" {

int isp_nlm_params_default(isp_nlm_params *p, int correction, double sigma)
{
    if (!p) return ISP_BM3D_E_ARG;
    // NLMeans_Para (include/NLMeans.h:11-22)
    p->correction = correction != 0;
    p->sigma = sigma;
    p->strength = correction ? sigma * 5 : sigma * 1.5;
    p->block_size = 8;
    p->block_step = 5;
    p->group_size = 16;
    p->bm_range = 24;
    p->bm_step = 3;
    p->thMSE = correction ? sigma * 50 : sigma * 25;
    return ISP_BM3D_OK;
}

int isp_nlm_plane_f32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, float *dst, const float *src, const float *ref,
                      int W, int H, size_t stride, int mem, void *stream)
{
    void *d[3] = {dst, nullptr, nullptr};
    const void *sv[3] = {src, nullptr, nullptr}, *r[3] = {ref, nullptr, nullptr};
    const Fmt f{false, 0, 0, 1};
    return nlm_common(ctx, p, 1, d, sv, r, W, H, stride, f, mem, stream);
}

int isp_nlm_plane_i32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, int *dst, const int *src, const int *ref, int W,
                      int H, size_t stride, int floor_, int neutral, int ceil_, int mem, void *stream)
{
    void *d[3] = {dst, nullptr, nullptr};
    const void *sv[3] = {src, nullptr, nullptr}, *r[3] = {ref, nullptr, nullptr};
    const Fmt f{true, floor_, neutral, ceil_};
    return nlm_common(ctx, p, 1, d, sv, r, W, H, stride, f, mem, stream);
}

int isp_nlm_rgb_i32(isp_bm3d_ctx *ctx, const isp_nlm_params *p, int *dst_r, int *dst_g, int *dst_b,
                    const int *src_r, const int *src_g, const int *src_b, const int *ref_r, const int *ref_g,
                    const int *ref_b, int W, int H, size_t stride, int floor_, int ceil_, int mem, void *stream)
{
    void *d[3] = {dst_r, dst_g, dst_b};
    const void *sv[3] = {src_r, src_g, src_b}, *r[3] = {ref_r, ref_g, ref_b};
    const Fmt f{true, floor_, floor_, ceil_};
    return nlm_common(ctx, p, 3, d, sv, r, W, H, stride, f, mem, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multi-frame grouping: BM3D over a stack of frames with Pos3 groups
// (BlockGroup, include/Block.h:843-846, 927-936, 1141-1208).
namespace {

struct StackWork {
    float *src[3];                       // plane c of frame t at src[c] + t * fs (YUV for RGB)
    float *refp[3];                      // a separate ref stack's planes
    float *basic[3];                     // step-1 estimate
    unsigned long long *num[3], *den[3];
    float *out[3];                       // host-mode staging (4-byte samples)
    double *y64, *r64;                   // FP64 luma guides of RGB src / ref frames
    int *lcnt[kMaxStackLists], *lpos[kMaxStackLists];
    double *lkey[kMaxStackLists];
    int *cnt, *pos, *slot;               // merged groups of every reference frame (slot: member frame)
};

size_t stack_layout(StackWork *w, void *base, size_t fs, int T, size_t refs, int G, int nlists)
{
    Carver c(base);
    const size_t tf = fs * (size_t)T;
    for (int k = 0; k < 3; ++k) w->src[k] = c.take<float>(tf);
    for (int k = 0; k < 3; ++k) w->refp[k] = c.take<float>(tf);
    for (int k = 0; k < 3; ++k) w->basic[k] = c.take<float>(tf);
    for (int k = 0; k < 3; ++k) w->num[k] = c.take<unsigned long long>(tf);
    for (int k = 0; k < 3; ++k) w->den[k] = c.take<unsigned long long>(tf);
    for (int k = 0; k < 3; ++k) w->out[k] = c.take<float>(tf);
    w->y64 = c.take<double>(tf);
    w->r64 = c.take<double>(tf);
    for (int l = 0; l < nlists; ++l) {
        w->lcnt[l] = c.take<int>(refs);
        w->lpos[l] = c.take<int>(refs * G);
        w->lkey[l] = c.take<double>(refs * G);
    }
    w->cnt = c.take<int>(refs * T);
    w->pos = c.take<int>(refs * G * T);
    w->slot = c.take<int>(refs * G * T);
    return c.off + 256;
}

// Frames of the window of t, own frame first, then t-1, t+1, t-2, ...
int stack_order(int t, int T, int radius, int *order)
{
    int n = 0;
    order[n++] = t;
    for (int d = 1; d <= radius; ++d) {
        if (t - d >= 0) order[n++] = t - d;
        if (t + d < T) order[n++] = t + d;
    }
    return n;
}

// Planes of one step over the stack.
struct StackIO {
    int nplanes, T, radius, P;
    size_t fs;
    const float *src[3];                 // frame 0 of each plane
    const float *guide[3];               // Wiener guides (final step)
    const float *mg32;                   // matching guide frames, FP32 ...
    const double *mg64;                  // ... or FP64
    // finalize of frame t: out[t * nplanes + c], out_pitch
    void *const *out;
    size_t out_pitch;
    int to_rgb;
};

// One BM3D step (BM3D_Base::Kernel with multi-frame groups): the groups of
// every reference frame, one filter launch over the whole stack, then
// num/den of every frame.
int stack_step(isp_bm3d_ctx *ctx, const isp_bm3d_step_params &p, bool wiener, const StackIO &io, StackWork &w,
               int W, int H, hipStream_t s)
{
    Derived d{};
    int rc = derive(ctx, p, io.nplanes, W, H, d);
    if (rc) return rc;
    if (d.generic)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "multi-frame groups need BlockSize 8 and GroupSize <= %d",
                    kMaxGroup);
    const size_t fs = io.fs;
    const int T = io.T, P = io.P;
    const bool mfma = stack_engine(ctx) != 0;
    if (!d.degenerate) {
        const isp_bm3d_ctx::Tiles tl = tiles_for(ctx, W, H);
        const size_t refs = (size_t)d.ngx * d.ngy;
        for (int t = 0; t < T; ++t) {
            int order[kMaxStackLists];
            const int nl = d.bypass ? 1 : stack_order(t, T, io.radius, order);
            if (d.bypass) order[0] = t;
            StackMergeArgs mg{};
            mg.refs = (int)refs;
            mg.G = p.group_size;
            mg.nlists = nl;
            for (int l = 0; l < nl; ++l) {
                const int z = order[l];
                const void *gt = io.mg64 ? (const void *)(io.mg64 + fs * t) : (const void *)(io.mg32 + fs * t);
                const void *gz = io.mg64 ? (const void *)(io.mg64 + fs * z) : (const void *)(io.mg32 + fs * z);
                MatchArgs a = build_match(ctx, p, d, W, H, gt, P, tl.match_x, tl.match_y, w.lcnt[l], w.lpos[l],
                                          w.lkey[l]);
                if (z != t) {
                    a.search = gz;
                    a.walk = 0;
                    a.tilePitch = match_pitch(a, d, p);
                    a.refW = (tl.match_x - 1) * p.block_step + kBlock;
                    a.refH = (tl.match_y - 1) * p.block_step + kBlock;
                    a.refPitch = a.refW | 1;
                }
                if ((rc = check_match_lds(ctx, p, a))) return rc;
                StageTimer tm(ctx, wiener ? ST_MATCH_FINAL : ST_MATCH_BASIC, s);
                HIP_CHECK(ctx, launch_match(a, io.mg64 != nullptr, s));
                mg.cnt[l] = w.lcnt[l];
                mg.pos[l] = w.lpos[l];
                mg.key[l] = w.lkey[l];
                mg.slot[l] = z;
            }
            mg.cnt_out = w.cnt + refs * t;
            mg.pos_out = w.pos + refs * p.group_size * t;
            mg.slot_out = w.slot + refs * p.group_size * t;
            StageTimer tm(ctx, wiener ? ST_MATCH_FINAL : ST_MATCH_BASIC, s);
            HIP_CHECK(ctx, launch_stack_merge(mg, s));
        }
        for (int c = 0; c < 3; ++c)
            if (d.active[c]) {
                StageTimer tm(ctx, ST_FILL, s);
                HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(w.num[c]), fs * T * 2, 0.f, s));
                HIP_CHECK(ctx, launch_fill(reinterpret_cast<float *>(w.den[c]), fs * T * 2, 0.f, s));
            }
        // one launch over the stack: each workgroup walks the frames for its tile
        int trx = mfma ? tl.mfma_x : tl.filter_x, try_ = mfma ? tl.mfma_y : tl.filter_y;
        unsigned long long *np_[3], *dp[3];
        for (int c = 0; c < 3; ++c) {
            np_[c] = w.num[c];
            dp[c] = w.den[c];
        }
        FilterArgs f{};
        for (;;) {
            f = build_filter(ctx, p, d, W, H, P, io.src, io.guide, np_, dp, w.cnt, w.pos, trx, try_);
            f.slot = w.slot;
            f.frames = T;
            f.radius = io.radius;
            // groups of 2 * radius + 1 reference frames reach each frame
            f.wmax = wiener_weight_cap(kBlock, p.block_step, d.range, p.bm_step, p.group_size, io.radius);
            f.nslots = std::min(2 * io.radius + 1, T);
            f.fstride = fs;
            if (filter_stack_lds_bytes(f, wiener, stack_engine(ctx)) <= 160 * 1024) break;
            if (trx == 1 && try_ == 1)
                return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "radius %d needs more LDS than a CU has", io.radius);
            if (trx >= try_ && trx > 1) trx = (trx + 1) / 2;
            else try_ = (try_ + 1) / 2;
        }
        if (!(f.wmax > 0))
            return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "radius %d: a pixel's aggregation sums could overflow",
                        io.radius);
        StageTimer tm(ctx, wiener ? ST_FILTER_FINAL : ST_FILTER_BASIC, s);
        HIP_CHECK(ctx, launch_filter_stack(f, wiener, stack_engine(ctx), s));
    }
    for (int t = 0; t < T; ++t) {
        FinalizeArgs fa{};
        fa.W = W;
        fa.H = H;
        fa.pitch = P;
        fa.nplanes = io.nplanes;
        fa.to_rgb = io.to_rgb;
        fa.out_int = 0;
        fa.clip = io.to_rgb;
        fa.gain_y = fa.gain_c = 1.0;
        fa.offset = 0.0;
        fa.lo = 0.0;
        fa.hi = 1.0;
        fa.out_pitch = io.out_pitch;
        fa.den_box = kBlock;  // the stack filters add den per block origin
        for (int c = 0; c < 3; ++c) {
            fa.src[c] = io.src[c] ? io.src[c] + fs * t : nullptr;
            fa.num[c] = w.num[c] + fs * t;
            fa.den[c] = w.den[c] + fs * t;
            fa.active[c] = d.degenerate ? 0 : d.active[c];
            fa.out[c] = (io.to_rgb || c < io.nplanes) ? io.out[t * (io.to_rgb ? 3 : io.nplanes) + c] : nullptr;
        }
        StageTimer tm(ctx, ST_FINALIZE, s);
        HIP_CHECK(ctx, launch_finalize(fa, s));
    }
    return ISP_BM3D_OK;
}

int stack_common(isp_bm3d_ctx *ctx, float *const *dst, const float *const *src, const float *const *ref, int T,
                 int nplanes, int W, int H, size_t stride, int radius, int mode, int mem, void *stream)
{
    int rc = check_common(ctx, W, H, stride, mode, mem);
    if (rc) return rc;
    if (T < 1 || T > 4096) return fail(ctx, ISP_BM3D_E_ARG, "frame count %d", T);
    if (nplanes != 1 && nplanes != 3) return fail(ctx, ISP_BM3D_E_ARG, "nplanes %d (1 or 3)", nplanes);
    if (radius < 0 || radius > kMaxStackRadius)
        return fail(ctx, ISP_BM3D_E_UNSUPPORTED, "temporal radius %d outside 0..%d", radius, kMaxStackRadius);
    if (!dst || !src) return fail(ctx, ISP_BM3D_E_ARG, "null dst/src");
    const int n = T * nplanes;
    for (int k = 0; k < n; ++k)
        if (!dst[k] || !src[k] || (ref && !ref[k])) return fail(ctx, ISP_BM3D_E_ARG, "null plane %d", k);
    hipStream_t s = (hipStream_t)stream;
    const bool host = mem == ISP_BM3D_MEM_HOST;
    const bool has_ref = ref != nullptr;
    // the single steps' early return (as isp_bm3d_plane_f32 / rgb_f32)
    if (mode != ISP_BM3D_BOTH) {
        const bool degen = nplanes == 3 ? frame_degenerate(ctx, mode) : [&] {
            const isp_bm3d_step_params &p = mode == ISP_BM3D_BASIC ? ctx->params.basic : ctx->params.final_;
            Derived d{};
            return derive(ctx, p, 1, W, H, d) == ISP_BM3D_OK && d.degenerate;
        }();
        if (degen) {
            for (int k = 0; k < n; ++k) {
                if (host) host_copy_plane(dst[k], src[k], stride, W, H);
                else if (dst[k] != src[k] && (rc = copy_plane(ctx, dst[k], stride, src[k], stride, W, H, s)))
                    return rc;
            }
            return ISP_BM3D_OK;
        }
    }
    const int P = ((W + 63) / 64) * 64;
    const size_t fs = (size_t)P * H;
    size_t refs = 1;
    int G = 1;
    for (const isp_bm3d_step_params *p : {&ctx->params.basic, &ctx->params.final_}) {
        const int st = p->block_step > 0 ? p->block_step : 1;
        refs = std::max(refs, (size_t)(W / st + 2) * (size_t)(H / st + 2));
        G = std::max(G, std::min(std::max(p->group_size, 1), kMaxGroup));
    }
    const int nlists = 2 * radius + 1;
    StackWork w{};
    if ((rc = ensure_ws(ctx, stack_layout(&w, nullptr, fs, T, refs, G, nlists)))) return rc;
    stack_layout(&w, ctx->ws, fs, T, refs, G, nlists);
    for (int k = 0; k < 2; ++k) {
        ctx->last_cnt[k] = nullptr;
        ctx->last_refs[k] = 0;
    }

    // frames into the workspace: grey planes as they are, RGB as OPP planes
    // plus the FP64 luma guide; a separate ref likewise
    const HostStage hs{ctx, s, P, W, H};
    const int nin = has_ref ? 2 * n : n;
    if (host && (rc = HostStage::reserve(ctx, (size_t)nin * hs.plane_bytes()))) return rc;
    for (int pass = 0; pass < (has_ref ? 2 : 1); ++pass) {
        const float *const *in = pass == 0 ? src : ref;
        float *const *dstp = pass == 0 ? w.src : w.refp;
        double *g64 = pass == 0 ? w.y64 : w.r64;
        for (int t = 0; t < T; ++t) {
            const void *pl[3];
            size_t ip = stride;
            for (int c = 0; c < nplanes; ++c) {
                pl[c] = in[t * nplanes + c];
                if (host) {
                    // RGB lands in the output staging planes, grey directly in place
                    float *land = nplanes == 3 ? w.out[c] + fs * t : dstp[c] + fs * t;
                    if ((rc = hs.upload(pass * n + t * nplanes + c, land, pl[c], stride))) return rc;
                    pl[c] = land;
                }
            }
            if (host) ip = (size_t)P;
            if (nplanes == 3) {
                OppArgs o{};
                o.W = W;
                o.H = H;
                o.r = pl[0];
                o.g = pl[1];
                o.b = pl[2];
                o.gainY = 1.0 / (1.0 * 3.0);
                o.offsetY = -0.0 * 3.0 * o.gainY + 0.0;
                o.gainU = 1.0 / (1.0 * 2.0);
                o.gainV = 1.0 / (1.0 * 4.0);
                o.offsetC = 0.0;
                o.in_pitch = ip;
                o.y = dstp[0] + fs * t;
                o.u = dstp[1] + fs * t;
                o.v = dstp[2] + fs * t;
                o.pitch = P;
                o.y64 = g64 + fs * t;
                o.pitch64 = P;
                StageTimer tm(ctx, ST_OPP, s);
                HIP_CHECK(ctx, launch_opp(o, s));
            } else if (!host) {
                if ((rc = copy_plane(ctx, dstp[0] + fs * t, P, pl[0], stride, W, H, s))) return rc;
            }
        }
    }

    // outputs: the caller's planes (device) or the staging planes (host)
    std::vector<void *> outs((size_t)n);
    for (int t = 0; t < T; ++t)
        for (int c = 0; c < nplanes; ++c)
            outs[(size_t)t * nplanes + c] = host ? (void *)(w.out[c] + fs * t) : (void *)dst[t * nplanes + c];
    std::vector<void *> basics((size_t)T * 3);
    for (int t = 0; t < T; ++t)
        for (int c = 0; c < 3; ++c) basics[(size_t)t * 3 + c] = w.basic[c] + fs * t;
    StackIO io{};
    io.nplanes = nplanes;
    io.T = T;
    io.radius = radius;
    io.P = P;
    io.fs = fs;
    for (int c = 0; c < 3; ++c) io.src[c] = c < nplanes ? w.src[c] : nullptr;
    const bool rgb = nplanes == 3;
    auto final_out = [&](StackIO &x) {
        x.out = outs.data();
        x.out_pitch = host ? (size_t)P : stride;
        x.to_rgb = rgb;
    };
    if (mode == ISP_BM3D_BASIC || mode == ISP_BM3D_FINAL) {
        const bool fin = mode == ISP_BM3D_FINAL;
        // step-1 matching on ref's luma; a final step's ref is its Wiener guide
        for (int c = 0; c < 3; ++c) io.guide[c] = c < nplanes ? (has_ref && fin ? w.refp[c] : w.src[c]) : nullptr;
        if (rgb) io.mg64 = has_ref ? w.r64 : w.y64;
        else io.mg32 = has_ref ? w.refp[0] : w.src[0];
        final_out(io);
        rc = stack_step(ctx, fin ? ctx->params.final_ : ctx->params.basic, fin, io, w, W, H, s);
    } else {
        for (int c = 0; c < 3; ++c) io.guide[c] = c < nplanes ? w.src[c] : nullptr;
        if (rgb) io.mg64 = has_ref ? w.r64 : w.y64;
        else io.mg32 = has_ref ? w.refp[0] : w.src[0];
        io.out = basics.data();
        io.out_pitch = (size_t)P;
        io.to_rgb = 0;
        // step 1 writes plane c of frame t to basics[t * 3 + c]: index with 3 planes
        StackIO io1 = io;
        std::vector<void *> b1((size_t)n);
        for (int t = 0; t < T; ++t)
            for (int c = 0; c < nplanes; ++c) b1[(size_t)t * nplanes + c] = basics[(size_t)t * 3 + c];
        io1.out = b1.data();
        rc = stack_step(ctx, ctx->params.basic, false, io1, w, W, H, s);
        if (!rc) {
            StackIO io2 = io;
            for (int c = 0; c < 3; ++c) io2.guide[c] = c < nplanes ? w.basic[c] : nullptr;
            io2.mg64 = nullptr;
            io2.mg32 = w.basic[0];
            final_out(io2);
            rc = stack_step(ctx, ctx->params.final_, true, io2, w, W, H, s);
        }
    }
    if (rc) return rc;
    if (host) {
        for (int k = 0; k < n; ++k) {
            const int t = k / nplanes, c = k % nplanes;
            HIP_CHECK(ctx, hipMemcpyAsync(hs.slot(k), w.out[c] + fs * t, hs.plane_bytes(), hipMemcpyDeviceToHost, s));
        }
        HIP_CHECK(ctx, hipStreamSynchronize(s));
        for (int k = 0; k < n; ++k) hs.download_finish(k, dst[k], stride);
    }
    return ISP_BM3D_OK;
}

}  // namespace

extern "C" int isp_bm3d_stack_f32(isp_bm3d_ctx *ctx, float *const *dst, const float *const *src,
                                  const float *const *ref, int frames, int nplanes, int width, int height,
                                  size_t stride, int radius, int mode, int mem, void *stream)
{
    return stack_common(ctx, dst, src, ref, frames, nplanes, width, height, stride, radius, mode, mem, stream);
}
