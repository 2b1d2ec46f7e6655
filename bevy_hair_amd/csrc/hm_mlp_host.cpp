// Host side of the learned shading term: weights, sample collection,
// training loop and the MLP shading entry point (see hm_mlp.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "hm_ctx.h"
#include "hm_mlp.h"

using hm::MLP_IN;
using hm::MLP_NBIAS;
using hm::MLP_NOUT;
using hm::MLP_NW;

struct hm_mlp_state {
    float* w = nullptr;      // float32 master weights
    float* b = nullptr;
    __bf16* packed = nullptr;
    float* m = nullptr;      // Adam moments (weights then biases)
    float* v = nullptr;
    float* gw = nullptr;
    float* gb = nullptr;
    float* loss = nullptr;
    uint32_t batch_cap = 0;
    __bf16* act0 = nullptr;
    __bf16* act[3] = {nullptr, nullptr, nullptr};
    float* dz[4] = {nullptr, nullptr, nullptr, nullptr};
    float* sx = nullptr;     // collected samples
    float* st = nullptr;
    uint32_t* scount = nullptr;
    uint32_t scap = 0;
    uint32_t* flags = nullptr;
    uint32_t flags_cap = 0;
    float lo[MLP_IN], hi[MLP_IN];
    bool ready = false;      // weights set
    bool range = false;      // trained feature range known
    bool borrowed = false;   // a batch lane's: w, b and packed are the parent's
    int step = 0;
};

namespace {

template <class T> hipError_t alloc(T** p, size_t bytes) {
    *p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(p), bytes ? bytes : 16);
}
template <class T> void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

#define MLP_HIP(ctx, call)                                             \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return hm_hip_fail((ctx), e_, #call);    \
    } while (0)

hm_status state(hm_ctx* ctx, hm_mlp_state** out) {
    if (!ctx) return HM_ERR_INVALID;
    if (!ctx->mlp) {
        (void)hipSetDevice(ctx->device);
        ctx->alloc_gen++;
        hm_mlp_state* s = new hm_mlp_state();
        hipError_t e = alloc(&s->w, MLP_NW * 4);
        if (e == hipSuccess) e = alloc(&s->b, MLP_NBIAS * 4);
        if (e == hipSuccess) e = alloc(&s->packed, hm::MLP_TILES * 64 * 8 * 2);
        if (e == hipSuccess) e = alloc(&s->m, (MLP_NW + MLP_NBIAS) * 4);
        if (e == hipSuccess) e = alloc(&s->v, (MLP_NW + MLP_NBIAS) * 4);
        if (e == hipSuccess) e = alloc(&s->gw, MLP_NW * 4);
        if (e == hipSuccess) e = alloc(&s->gb, MLP_NBIAS * 4);
        if (e == hipSuccess) e = alloc(&s->loss, 16);
        if (e == hipSuccess) e = alloc(&s->scount, 16);
        if (e == hipSuccess) e = hipMemset(s->gw, 0, MLP_NW * 4);
        if (e == hipSuccess) e = hipMemset(s->gb, 0, MLP_NBIAS * 4);
        if (e == hipSuccess) e = hipMemset(s->scount, 0, 16);
        ctx->mlp = s;
        if (e != hipSuccess) {
            hm_mlp_release(ctx);
            return hm_hip_fail(ctx, e, "mlp state");
        }
    }
    *out = ctx->mlp;
    return HM_OK;
}

hm_status ensure_batch(hm_ctx* ctx, hm_mlp_state* s, uint32_t batch) {
    if (batch <= s->batch_cap) return HM_OK;
    release(s->act0);
    for (int i = 0; i < 3; ++i) release(s->act[i]);
    for (int i = 0; i < 4; ++i) release(s->dz[i]);
    s->batch_cap = 0;
    MLP_HIP(ctx, alloc(&s->act0, (size_t)batch * 32 * 2));
    for (int i = 0; i < 3; ++i) MLP_HIP(ctx, alloc(&s->act[i], (size_t)batch * 64 * 2));
    for (int i = 0; i < 3; ++i) MLP_HIP(ctx, alloc(&s->dz[i], (size_t)batch * 64 * 4));
    MLP_HIP(ctx, alloc(&s->dz[3], (size_t)batch * 16 * 4));
    s->batch_cap = batch;
    return HM_OK;
}

hm::TrainArgs train_args(hm_mlp_state* s, const float* x, const float* t, uint32_t n, const float scale[MLP_NOUT]) {
    hm::TrainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = x;
    a.t = t;
    a.n = n;
    a.perm_a = 1;
    a.perm_b = 0;
    a.packed = s->packed;
    a.w = s->w;
    a.bias = s->b;
    a.gw = s->gw;
    a.gb = s->gb;
    a.act0 = s->act0;
    for (int i = 0; i < 3; ++i) a.act[i] = s->act[i];
    for (int i = 0; i < 4; ++i) a.dz[i] = s->dz[i];
    for (int i = 0; i < MLP_NOUT; ++i) a.scale[i] = scale ? scale[i] : 1.0f;
    a.loss = s->loss;
    return a;
}

hm_status upload_weights(hm_ctx* ctx, hm_mlp_state* s, const float* w, const float* b) {
    MLP_HIP(ctx, hipMemcpy(s->w, w, MLP_NW * 4, hipMemcpyHostToDevice));
    MLP_HIP(ctx, hipMemcpy(s->b, b, MLP_NBIAS * 4, hipMemcpyHostToDevice));
    MLP_HIP(ctx, hipMemset(s->m, 0, (MLP_NW + MLP_NBIAS) * 4));
    MLP_HIP(ctx, hipMemset(s->v, 0, (MLP_NW + MLP_NBIAS) * 4));
    MLP_HIP(ctx, hm::launch_mlp_pack(ctx->stream, s->w, s->packed));
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->step = 0;
    s->ready = true;
    return HM_OK;
}

}  // namespace

void hm_mlp_release(hm_ctx* ctx) {
    if (!ctx || !ctx->mlp) return;
    hm_mlp_state* s = ctx->mlp;
    if (s->borrowed) {
        s->w = nullptr;
        s->b = nullptr;
        s->packed = nullptr;
    }
    release(s->w); release(s->b); release(s->packed); release(s->m); release(s->v);
    release(s->gw); release(s->gb); release(s->loss); release(s->act0);
    for (int i = 0; i < 3; ++i) release(s->act[i]);
    for (int i = 0; i < 4; ++i) release(s->dz[i]);
    release(s->sx); release(s->st); release(s->scount); release(s->flags);
    delete s;
    ctx->mlp = nullptr;
}

hm_status hm_mlp_borrow(hm_ctx* lane, const hm_ctx* parent) {
    if (!parent->mlp) return HM_OK;
    const hm_mlp_state* p = parent->mlp;
    hm_mlp_state* s = new hm_mlp_state();
    s->borrowed = true;
    s->w = p->w;
    s->b = p->b;
    s->packed = p->packed;
    std::memcpy(s->lo, p->lo, sizeof(s->lo));
    std::memcpy(s->hi, p->hi, sizeof(s->hi));
    s->ready = p->ready;
    s->range = p->range;
    lane->mlp = s;
    return HM_OK;
}

extern "C" {

hm_status hm_mlp_init(hm_ctx* ctx, uint32_t seed) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    // He-uniform weights, zero biases; zero padding rows/columns stay zero.
    std::mt19937_64 rng(seed);
    std::vector<float> w(MLP_NW, 0.0f), b(MLP_NBIAS, 0.0f);
    const int outs[4] = {64, 64, 64, 16}, ins[4] = {32, 64, 64, 64};
    for (int L = 0; L < 4; ++L) {
        int fan_in = L == 0 ? hm::MLP_NFEAT : ins[L];
        std::uniform_real_distribution<float> U(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
        for (int o = 0; o < outs[L]; ++o)
            for (int i = 0; i < ins[L]; ++i) {
                bool pad = (L == 0 && i >= hm::MLP_NFEAT) || (L == 3 && o >= MLP_NOUT);
                float v = U(rng);
                w[hm::mlp_w_offset(L) + o * ins[L] + i] = pad ? 0.0f : v;
            }
    }
    return upload_weights(ctx, s, w.data(), b.data());
}

hm_status hm_mlp_set_weights(hm_ctx* ctx, const float* w, const float* b) {
    if (!w || !b) return hm_fail(ctx, HM_ERR_INVALID, "null weights");
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    return upload_weights(ctx, s, w, b);
}

hm_status hm_mlp_get_weights(hm_ctx* ctx, float* w, float* b) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!s->ready) return hm_fail(ctx, HM_ERR_NOT_READY, "network weights not set");
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (w) MLP_HIP(ctx, hipMemcpy(w, s->w, MLP_NW * 4, hipMemcpyDeviceToHost));
    if (b) MLP_HIP(ctx, hipMemcpy(b, s->b, MLP_NBIAS * 4, hipMemcpyDeviceToHost));
    return HM_OK;
}

hm_status hm_mlp_forward(hm_ctx* ctx, const float* x, uint32_t n, float* y) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!s->ready) return hm_fail(ctx, HM_ERR_NOT_READY, "network weights not set");
    if (!x || !y) return hm_fail(ctx, HM_ERR_INVALID, "null array");
    float *dx = nullptr, *dy = nullptr;
    hipError_t e = alloc(&dx, (size_t)n * 32 * 4);
    if (e == hipSuccess) e = alloc(&dy, (size_t)n * 16 * 4);
    if (e == hipSuccess) e = hipMemcpy(dx, x, (size_t)n * 32 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hm::launch_mlp_forward(ctx->stream, dx, n, s->packed, s->b, dy);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * 16 * 4, hipMemcpyDeviceToHost);
    release(dx);
    release(dy);
    if (e != hipSuccess) return hm_hip_fail(ctx, e, "hm_mlp_forward");
    return HM_OK;
}

hm_status hm_mlp_gradients(hm_ctx* ctx, const float* x, const float* t, uint32_t n, const float scale[MLP_NOUT],
                           float* loss, float* dw, float* db) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!s->ready) return hm_fail(ctx, HM_ERR_NOT_READY, "network weights not set");
    if (!x || !t || !scale || n == 0) return hm_fail(ctx, HM_ERR_INVALID, "bad gradient arguments");
    if ((st = ensure_batch(ctx, s, n)) != HM_OK) return st;
    float *dx = nullptr, *dt = nullptr;
    MLP_HIP(ctx, alloc(&dx, (size_t)n * 32 * 4));
    hipError_t e = alloc(&dt, (size_t)n * MLP_NOUT * 4);
    if (e == hipSuccess) e = hipMemcpy(dx, x, (size_t)n * 32 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, t, (size_t)n * MLP_NOUT * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->gw, 0, MLP_NW * 4);
    if (e == hipSuccess) e = hipMemset(s->gb, 0, MLP_NBIAS * 4);
    if (e == hipSuccess) e = hipMemset(s->loss, 0, 4);
    hm::TrainArgs a = train_args(s, dx, dt, n, scale);
    a.first = 0;
    a.batch = n;
    a.inv_count = 1.0f / ((float)n * MLP_NOUT);
    if (e == hipSuccess) e = hm::launch_mlp_grads(ctx->stream, a);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && dw) e = hipMemcpy(dw, s->gw, MLP_NW * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && db) e = hipMemcpy(db, s->gb, MLP_NBIAS * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, s->loss, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemset(s->gw, 0, MLP_NW * 4);
    if (e == hipSuccess) e = hipMemset(s->gb, 0, MLP_NBIAS * 4);
    release(dx);
    release(dt);
    if (e != hipSuccess) return hm_hip_fail(ctx, e, "hm_mlp_gradients");
    return HM_OK;
}

hm_status hm_mlp_reserve_samples(hm_ctx* ctx, uint32_t capacity) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (capacity != s->scap) {
        release(s->sx);
        release(s->st);
        s->scap = 0;
        MLP_HIP(ctx, alloc(&s->sx, (size_t)capacity * 32 * 4));
        MLP_HIP(ctx, alloc(&s->st, (size_t)capacity * MLP_NOUT * 4));
        s->scap = capacity;
    }
    MLP_HIP(ctx, hipMemset(s->scount, 0, 4));
    return HM_OK;
}

static void shade_args(hm_ctx* ctx, const hm_params* p, float exposure, hm::ShadeArgs& a) {
    std::memcpy(&a.view, p, sizeof(hm_params));
    a.center = ctx->center;
    a.beta = ctx->beta;
    a.dom = ctx->dom;
    hm_shade_lights(ctx, a);
    a.exposure = exposure;
    a.shaded = ctx->shaded;
    a.shaded_depth = ctx->shaded_depth;
}

hm_status hm_mlp_collect(hm_ctx* ctx, const hm_params* cam, float keep, uint32_t seed, uint32_t* total) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!ctx->gb_slab) return hm_fail(ctx, HM_ERR_NOT_READY, "targets not allocated");
    if (!s->sx) return hm_fail(ctx, HM_ERR_NOT_READY, "no sample buffer (hm_mlp_reserve_samples)");
    if (!cam || (uint32_t)cam->screen[0] != ctx->width || (uint32_t)cam->screen[1] != ctx->height)
        return hm_fail(ctx, HM_ERR_INVALID, "camera params size differs from the targets");
    hm::ShadeArgs a;
    shade_args(ctx, cam, 1.0f, a);
    hm::SampleArgs sa = {s->sx, s->st, s->scount, s->scap, seed, keep};
    MLP_HIP(ctx, hm::launch_mlp_samples(ctx->stream, a, sa, ctx->width, ctx->height));
    uint32_t n = 0;
    MLP_HIP(ctx, hipMemcpyAsync(&n, s->scount, 4, hipMemcpyDeviceToHost, ctx->stream));
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (total) *total = std::min(n, s->scap);
    return HM_OK;
}

hm_status hm_mlp_read_samples(hm_ctx* ctx, float* x, float* t, uint32_t n) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (n > s->scap) return hm_fail(ctx, HM_ERR_INVALID, "more samples than held");
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (x) MLP_HIP(ctx, hipMemcpy(x, s->sx, (size_t)n * 32 * 4, hipMemcpyDeviceToHost));
    if (t) MLP_HIP(ctx, hipMemcpy(t, s->st, (size_t)n * MLP_NOUT * 4, hipMemcpyDeviceToHost));
    return HM_OK;
}

hm_status hm_mlp_set_samples(hm_ctx* ctx, const float* x, const float* t, uint32_t n) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!x || !t) return hm_fail(ctx, HM_ERR_INVALID, "null samples");
    if (n > s->scap) return hm_fail(ctx, HM_ERR_CAPACITY, "more samples than the reserved capacity");
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MLP_HIP(ctx, hipMemcpy(s->sx, x, (size_t)n * 32 * 4, hipMemcpyHostToDevice));
    MLP_HIP(ctx, hipMemcpy(s->st, t, (size_t)n * MLP_NOUT * 4, hipMemcpyHostToDevice));
    MLP_HIP(ctx, hipMemcpy(s->scount, &n, 4, hipMemcpyHostToDevice));
    return HM_OK;
}

hm_status hm_mlp_set_range(hm_ctx* ctx, const float lo[32], const float hi[32]) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!lo || !hi) {
        s->range = false;
        return HM_OK;
    }
    std::memcpy(s->lo, lo, sizeof(s->lo));
    std::memcpy(s->hi, hi, sizeof(s->hi));
    s->range = true;
    return HM_OK;
}

hm_status hm_mlp_train(hm_ctx* ctx, uint32_t epochs, uint32_t batch, float lr, const float scale[MLP_NOUT], uint32_t seed,
                       float* final_loss) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!s->ready) return hm_fail(ctx, HM_ERR_NOT_READY, "network weights not set (hm_mlp_init)");
    if (!scale || batch == 0 || batch > (1u << 22)) return hm_fail(ctx, HM_ERR_INVALID, "bad training arguments");
    uint32_t n = 0;
    MLP_HIP(ctx, hipMemcpy(&n, s->scount, 4, hipMemcpyDeviceToHost));
    n = std::min(n, s->scap);
    if (n == 0) return hm_fail(ctx, HM_ERR_NOT_READY, "no samples collected");
    if ((st = ensure_batch(ctx, s, std::min(batch, n))) != HM_OK) return st;
    std::mt19937_64 rng(seed);
    float last = 0.0f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        uint64_t pa = 1 + rng() % n;
        while (std::gcd<uint64_t, uint64_t>(pa, n) != 1) pa = 1 + rng() % n;
        uint64_t pb = rng() % n;
        MLP_HIP(ctx, hipMemsetAsync(s->loss, 0, 4, ctx->stream));
        uint32_t steps = 0;
        for (uint32_t first = 0; first < n; first += batch) {
            hm::TrainArgs a = train_args(s, s->sx, s->st, n, scale);
            a.perm_a = pa;
            a.perm_b = pb;
            a.first = first;
            a.batch = std::min(batch, n - first);
            a.inv_count = 1.0f / ((float)a.batch * MLP_NOUT);
            s->step += 1;
            MLP_HIP(ctx, hm::launch_mlp_train_step(ctx->stream, a, s->w, s->b, s->m, s->v, lr, s->step));
            ++steps;
        }
        float sum = 0.0f;
        MLP_HIP(ctx, hipMemcpyAsync(&sum, s->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
        MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        last = sum / (float)steps;
    }
    if (final_loss) *final_loss = last;
    return HM_OK;
}

hm_status hm_shade_mlp(hm_ctx* ctx, const hm_params* p, float exposure, float tolerance, uint32_t* fallback_tiles) {
    return hm_shade_mlp_impl(ctx, p, exposure, tolerance, fallback_tiles, hm_lists());
}

}  // extern "C"

hm_status hm_shade_mlp_impl(hm_ctx* ctx, const hm_params* p, float exposure, float tolerance, uint32_t* fallback,
                            const hm_lists& lists) {
    hm_mlp_state* s;
    hm_status st = state(ctx, &s);
    if (st != HM_OK) return st;
    if (!s->ready) return hm_fail(ctx, HM_ERR_NOT_READY, "network weights not set");
    if (!ctx->gb_slab) return hm_fail(ctx, HM_ERR_NOT_READY, "targets not allocated");
    if (!p || (uint32_t)p->screen[0] != ctx->width || (uint32_t)p->screen[1] != ctx->height ||
        (uint32_t)p->dom_info[0] != ctx->dom_size)
        return hm_fail(ctx, HM_ERR_INVALID, "shade params differ from the targets");
    hm::ShadeArgs a;
    shade_args(ctx, p, exposure, a);
    struct T {
        hm_ctx* c;
        size_t base;
        bool on;
        T(hm_ctx* ctx_) : c(ctx_) {
            on = c->slots && c->frame < c->slots;
            base = ((size_t)c->frame * HM_K_COUNT + HM_K_SHADE) * 2;
            if (on) (void)hipEventRecord(c->ev[base], c->stream);
        }
        ~T() {
            if (!on) return;
            (void)hipEventRecord(c->ev[base + 1], c->stream);
            c->ran[base / 2] = 1;
        }
    } timer(ctx);
    hm::PixelList L;
    if ((st = hm_shade_list(ctx, lists, &L)) != HM_OK) return st;
    // At most LIST_SUBS * cap listed pixels: at most cap groups of 64.
    const uint32_t groups_cap = L.cap;
    if (groups_cap > s->flags_cap) {
        ctx->alloc_gen++;
        release(s->flags);
        s->flags_cap = 0;
        MLP_HIP(ctx, alloc(&s->flags, (size_t)groups_cap * 4));
        s->flags_cap = groups_cap;
    }
    hm::MlpShadeArgs m;
    std::memset(&m, 0, sizeof(m));
    m.packed = s->packed;
    m.bias = s->b;
    m.group_flags = s->flags;
    m.groups_cap = groups_cap;
    for (int k = 0; k < MLP_IN; ++k) {
        m.lo[k] = s->range ? s->lo[k] : -3.4e38f;
        m.hi[k] = s->range ? s->hi[k] : 3.4e38f;
    }
    m.tol = tolerance;
    m.floor = 1e-4f;
    if (fallback)  // flags of groups past this frame's list would be stale: clear them for the count
        MLP_HIP(ctx, hm::launch_fill32(ctx->stream, s->flags, groups_cap, 0u, ctx->cu_count));
    MLP_HIP(ctx, hm::launch_shade_mlp(ctx->stream, a, m, L, ctx->cu_count * 8));
    if (fallback) {
        std::vector<uint32_t> f(groups_cap);
        MLP_HIP(ctx, hipMemcpyAsync(f.data(), s->flags, f.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        uint32_t c = 0;
        for (uint32_t v : f) c += (v & 3u) != 0u;
        *fallback = c;
    }
    return HM_OK;
}
