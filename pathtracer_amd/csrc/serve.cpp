// serve.cpp — render requests served on one context (pt_server_*).
//
// The reference renders one scene interactively and accumulates until the user
// stops (src/pathtracer.cpp:4005-4086). A server instead takes a stream of
// requests (scene JSON, resolution, sample budget, relative-error target) and
// traces each until its estimated relative RMSE reaches the target:
//   - every admitted request keeps its scene, BVH and per-pixel moment buffers
//     resident in HBM; admission stops at the server's HBM budget and request
//     count;
//   - a step traces one wave holding the paths of every active request that
//     fits (continuous batching): each request gets a quantum of min_spp
//     samples per pixel first, then its estimated remaining samples, in
//     policy order while the wave has room, so requests that arrive while
//     others are being traced join the next wave;
//   - the policy orders the requests: shortest estimated remaining GPU time
//     first (PT_SERVE_SRPT), or only the oldest request, one to completion at a
//     time (PT_SERVE_FIFO, the one-job-at-a-time baseline);
//   - after each quantum the relative RMSE is estimated on the GPU from the
//     per-pixel second moments (k_ext_rel_error): sqrt(sum var/n) / sqrt(sum mean^2)
//     over pixels and XYZ channels;
//   - a request that asks for denoising also keeps the sums of its odd-index
//     samples and the first-hit features of a 4-sample camera-ray pass; its
//     denoised image's error is estimated from the two half-sample images,
//     each denoised and compared with the other (ptk::dn_request_estimate). It
//     finishes denoised when that estimate reaches the target before the
//     plain one does; at its sample budget it returns whichever image has the
//     lower estimate (the plain mean is the fallback).
// The extended tracer's random numbers are keyed on (scene seed, pixel, sample,
// bounce, dimension), so a request's image does not depend on how its samples
// were split into quanta or interleaved with other requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pathtracer.h"
#include "tracer.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int id = 0;
    std::string json;
    int W = 0, H = 0, P = 0;
    int spp_max = 0, path_length = 0;
    float target = 0.01f;
    pt_scene* scene = nullptr;
    pt_push_constants pc{};
    DevAlloc mem;  // sum (3P floats), sumsq (3P floats), nsamp (P ints); denoised requests also:
                   // sum_odd (3P), n_odd (P), aov (12P), aov_n (P), den (3P)
    float* sum = nullptr;
    float* sumsq = nullptr;
    int* nsamp = nullptr;
    bool denoise = false;  // asked for, and the network takes the resolution
    float* sum_odd = nullptr;
    int* n_odd = nullptr;
    float* aov = nullptr;
    int* aov_n = nullptr;
    float* den = nullptr;
    double rel_dn = INFINITY;  // dual-buffer estimate of the denoised image
    int dn_spp = 0;            // samples at that estimate
    bool denoised = false;     // the result is the denoised image
    double mean2 = 0.0;        // sum of mean^2 over pixels and XYZ (per-sample units)
    int64_t bytes = 0;  // scene + moment buffers, counted against the server budget
    int spp = 0;
    double rel = INFINITY;
    int state = PT_REQ_WAITING;
    double t_arrive = 0.0, t_admit = -1.0, t_done = -1.0, gpu_s = 0.0;
    uint64_t segments = 0;
};

}  // namespace

struct pt_server {
    pt_ctx* ctx = nullptr;
    pt_server_config cfg{};
    Clock::time_point t0;
    int next_id = 0;
    std::map<int, std::unique_ptr<Request>> reqs;
    std::deque<int> waiting;  // arrival order
    std::vector<int> active;
    int64_t bytes_in_use = 0;
    double sec_per_path = 0.0;  // running estimate of GPU seconds per traced path
    double* d_rel = nullptr;    // per request of a wave: 64 lines x 32 doubles
    pt_denoiser* dn = nullptr;  // for requests that ask for denoising
    std::string err;
};

namespace {

constexpr int kRelLines = 64, kRelStride = 32;

double now_s(const pt_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

int fail(pt_server* s, int st, const std::string& msg) {
    s->err = msg;
    return st;
}

// Bytes a request will hold once admitted: the moment buffers plus an estimate
// of its scene tables and BVH (a few hundred bytes per primitive).
int64_t estimate_bytes(const Request& r) {
    int64_t prims = 0;
    for (size_t i = r.json.find("\"radius\""); i != std::string::npos; i = r.json.find("\"radius\"", i + 1)) ++prims;
    for (size_t i = r.json.find("\"edge1\""); i != std::string::npos; i = r.json.find("\"edge1\"", i + 1)) ++prims;
    return (int64_t)r.P * (r.denoise ? 112 : 28) + prims * 320 + (1 << 20);
}

// Paths per batched wave: the context's wave capacity, else 2^22.
int64_t server_wave_paths(const pt_ctx* ctx) { return ctx->cfg.wave_paths > 0 ? ctx->cfg.wave_paths : (1 << 22); }

void release_request(pt_server* s, Request& r) {
    if (r.scene) pt_scene_release(r.scene);
    r.scene = nullptr;
    ptk::dev_free(s->ctx, r.mem);
    s->bytes_in_use -= r.bytes;
    r.bytes = 0;
}

int admit(pt_server* s, Request& r, double now) {
    pt_ctx* ctx = s->ctx;
    int st = pt_scene_from_json(ctx, r.json.data(), r.json.size(), &r.scene);
    if (st != PT_OK) return fail(s, st, std::string("request scene: ") + pt_last_error(ctx));
    st = pt_scene_camera(r.scene, 1, &r.pc);
    if (st != PT_OK) return fail(s, st, std::string("request camera: ") + pt_last_error(ctx));
    r.pc.resolution[0] = r.W;
    r.pc.resolution[1] = r.H;
    r.pc.pathLength = r.path_length;
    const size_t P = (size_t)r.P;
    st = ptk::dev_alloc(ctx, r.mem, P * (r.denoise ? 112 : 28) + 8 * 256);
    if (st != PT_OK) return fail(s, st, ctx->err);
    char* p = static_cast<char*>(r.mem.p);
    auto carve = [&](size_t n) {
        char* q = p;
        p += (n + 255) / 256 * 256;
        return q;
    };
    r.sum = reinterpret_cast<float*>(carve(3 * P * 4));
    r.sumsq = reinterpret_cast<float*>(carve(3 * P * 4));
    r.nsamp = reinterpret_cast<int*>(carve(P * 4));
    if (r.denoise) {
        r.sum_odd = reinterpret_cast<float*>(carve(3 * P * 4));
        r.n_odd = reinterpret_cast<int*>(carve(P * 4));
        r.aov = reinterpret_cast<float*>(carve(12 * P * 4));
        r.aov_n = reinterpret_cast<int*>(carve(P * 4));
        r.den = reinterpret_cast<float*>(carve(3 * P * 4));
    }
    hipError_t e = hipMemsetAsync(r.mem.p, 0, r.mem.bytes, ctx->stream);
    if (e != hipSuccess) return fail(s, ptk::hip_fail(ctx, e, "request buffers"), ctx->err);
    // build the extended scene (BVH upload) now, so its bytes are known
    st = ptk::ext_scene_build(ctx, r.scene);
    if (st != PT_OK) return fail(s, st, ctx->err);
    if (r.denoise) {
        // first-hit features: 4 camera rays per pixel, traced one segment deep
        pt_push_constants fc = r.pc;
        fc.pathLength = 1;
        fc.samplesPerFrame = 4;
        fc.frame = 4;
        fc.currentSamples = 4;
        st = ptk::ext_dispatch_features(ctx, r.scene, &fc, r.aov, r.aov_n);
        if (st != PT_OK) return fail(s, st, ctx->err);
    }
    int64_t sb = 0;
    (void)pt_scene_device_bytes(r.scene, &sb);
    r.bytes = (int64_t)r.mem.bytes + sb;
    s->bytes_in_use += r.bytes;
    r.state = PT_REQ_ACTIVE;
    r.t_admit = now;
    return PT_OK;
}

// Samples still needed for the target, from rel ~ 1/sqrt(spp), and at least
// up to the minimum before a convergence decision.
int remaining_spp(const Request& r, int min_spp) {
    const int left = r.spp_max - r.spp;
    double rel = r.rel;
    if (r.denoise && r.dn_spp > 0) rel = std::min(rel, r.rel_dn * std::sqrt((double)r.dn_spp / r.spp));
    if (r.spp < 2 || !std::isfinite(rel)) return left;
    double need = (double)r.spp * ((rel / r.target) * (rel / r.target) - 1.0);
    need = std::max(need, (double)(min_spp - r.spp));
    return (int)std::min<double>(left, std::max(1.0, std::ceil(need)));
}

double remaining_cost(const pt_server* s, const Request& r) {
    return (double)remaining_spp(r, s->cfg.min_spp) * r.P * (s->sec_per_path > 0.0 ? s->sec_per_path : 1e-9);
}

ptk::DnRequest dn_view(const Request& r) {
    ptk::DnRequest q{};
    q.sum = r.sum;
    q.sumsq = r.sumsq;
    q.nsamp = r.nsamp;
    q.sum_odd = r.sum_odd;
    q.n_odd = r.n_odd;
    q.aov = r.aov;
    q.aov_n = r.aov_n;
    q.W = r.W;
    q.H = r.H;
    q.exposure = r.pc.apertureSize * r.pc.apertureSize * (float)r.pc.ISO;
    return q;
}

// After a quantum of a request that asks for denoising and has not met its
// target plainly: estimate the denoised image's error when it may have met
// the target (its last estimate, extrapolated as 1/sqrt(samples), within 20 %)
// or at the sample budget; finish denoised when the estimate meets the target,
// or at the budget when it beats the plain estimate. The denoiser's time is
// the request's GPU time.
int denoise_step(pt_server* s, Request& r, bool last) {
    const bool due = r.dn_spp == 0 || last ||
                     r.rel_dn * std::sqrt((double)r.dn_spp / r.spp) <= 1.2 * r.target;
    if (!due || r.spp < 2 || !(r.mean2 > 0.0)) return PT_OK;
    const auto t = Clock::now();
    const ptk::DnRequest q = dn_view(r);
    // radiance of order one for the network: the RMS of the mean image
    const float scale = (float)(q.exposure * std::sqrt(r.mean2 / (3.0 * r.P)));
    double rel = INFINITY;
    int st = ptk::dn_request_estimate(s->dn, q, scale, r.spp, s->d_rel, &rel);
    if (st != PT_OK) return fail(s, st, std::string("denoiser: ") + pt_dn_last_error(s->dn));
    r.rel_dn = rel;
    r.dn_spp = r.spp;
    if (rel <= r.target || (last && rel < r.rel)) {
        st = ptk::dn_request_denoise(s->dn, q, scale, r.den);
        if (st != PT_OK) return fail(s, st, std::string("denoiser: ") + pt_dn_last_error(s->dn));
        hipError_t e = hipStreamSynchronize(s->ctx->stream);
        if (e != hipSuccess) return fail(s, ptk::hip_fail(s->ctx, e, "denoiser"), s->ctx->err);
        r.denoised = true;
    }
    r.gpu_s += std::chrono::duration<double>(Clock::now() - t).count();
    return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_server_create(pt_ctx* ctx, const pt_server_config* cfg, pt_server** out) {
    if (!ctx || !out) return PT_ERR_INVALID_ARG;
    *out = nullptr;
    pt_server* s = new (std::nothrow) pt_server();
    if (!s) return PT_ERR_INVALID_ARG;
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget <= 0) s->cfg.hbm_budget = (int64_t)64 << 30;
    if (s->cfg.max_active <= 0) s->cfg.max_active = 16;
    if (s->cfg.min_spp <= 0) s->cfg.min_spp = 8;
    if (s->cfg.policy != PT_SERVE_FIFO && s->cfg.policy != PT_SERVE_HRRN) s->cfg.policy = PT_SERVE_SRPT;
    s->t0 = Clock::now();
    (void)hipSetDevice(ctx->device);
    if (hipMalloc(&s->d_rel, (size_t)ptk::kMaxBatch * kRelLines * kRelStride * sizeof(double)) != hipSuccess) {
        delete s;
        return ptk::hip_fail(ctx, hipGetLastError(), "pt_server_create");
    }
    *out = s;
    return PT_OK;
}

void pt_server_destroy(pt_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto& kv : s->reqs) release_request(s, *kv.second);
    if (s->d_rel) (void)hipFree(s->d_rel);
    delete s;
}

const char* pt_server_last_error(const pt_server* s) { return s ? s->err.c_str() : ""; }

double pt_server_now(const pt_server* s) { return s ? now_s(s) : 0.0; }

pt_status pt_server_submit(pt_server* s, const pt_request_desc* d, double t_arrive, int32_t* id) {
    if (!s || !d || !d->json || !id) return PT_ERR_INVALID_ARG;
    if (d->width <= 0 || d->height <= 0 || (int64_t)d->width * d->height > (int64_t)1 << 26 || d->spp_max <= 0 ||
        d->path_length <= 0 || !(d->target_rel > 0.0f))
        return fail(s, PT_ERR_INVALID_ARG, "pt_server_submit: bad request");
    if (d->denoise && !s->dn)
        return fail(s, PT_ERR_STATE, "pt_server_submit: denoising asked for but no denoiser attached");
    // a batched wave traces whole requests: one sample of every pixel must fit it
    if ((int64_t)d->width * d->height > server_wave_paths(s->ctx))
        return fail(s, PT_ERR_CAPACITY, "pt_server_submit: request has more pixels than a batched wave holds paths");
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->json.assign(d->json, d->len);
    r->W = d->width;
    r->H = d->height;
    r->P = d->width * d->height;
    r->spp_max = d->spp_max;
    r->path_length = d->path_length;
    r->target = d->target_rel;
    // a resolution the network does not take is served without denoising
    r->denoise = d->denoise != 0 && ptk::dn_fits(s->dn, r->W, r->H);
    r->t_arrive = t_arrive;
    *id = r->id;
    s->waiting.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    return PT_OK;
}

pt_status pt_server_step(pt_server* s, pt_request_status* stepped, int32_t cap, int32_t* n_stepped) {
    if (!s || (cap > 0 && !stepped)) return PT_ERR_INVALID_ARG;
    pt_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    if (n_stepped) *n_stepped = 0;
    // admission in arrival order, within the request count and HBM budget
    while (!s->waiting.empty() && (int)s->active.size() < s->cfg.max_active) {
        Request& r = *s->reqs[s->waiting.front()];
        if (!s->active.empty() && s->bytes_in_use + estimate_bytes(r) > s->cfg.hbm_budget) break;
        const auto ta = Clock::now();
        int st = admit(s, r, now_s(s));
        r.gpu_s += std::chrono::duration<double>(Clock::now() - ta).count();
        s->waiting.pop_front();
        if (st != PT_OK) {
            r.state = PT_REQ_FAILED;
            r.t_done = now_s(s);
            release_request(s, r);
            if (cap > 0 && n_stepped) {
                (void)pt_server_status(s, r.id, &stepped[0]);
                *n_stepped = 1;
            }
            return st;
        }
        s->active.push_back(r.id);
    }
    if (s->active.empty()) return PT_SERVE_IDLE;

    // the wave: requests in policy order, each with its quantum (min_spp samples
    // first, then the estimated remaining ones), while paths fit the wave;
    // FIFO traces only the oldest request
    std::vector<int> order = s->active;
    if (s->cfg.policy == PT_SERVE_SRPT) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return remaining_cost(s, *s->reqs[a]) < remaining_cost(s, *s->reqs[b]);
        });
    } else if (s->cfg.policy == PT_SERVE_HRRN) {
        // response ratio (age + remaining) / remaining, largest first (Brinch
        // Hansen's highest-response-ratio-next); equal ratios keep arrival order
        const double now = now_s(s);
        std::vector<std::pair<double, int>> key;
        for (int id : order) {
            const Request& r = *s->reqs[id];
            const double rem = std::max(remaining_cost(s, r), 1e-9);
            key.emplace_back(-(std::max(now - r.t_arrive, 0.0) + rem) / rem, id);
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
        for (size_t i = 0; i < key.size(); ++i) order[i] = key[i].second;
    }
    const int64_t wave = server_wave_paths(ctx);
    int64_t left = wave;
    std::vector<ptk::ExtBatchItem> items;
    std::vector<Request*> in;
    for (int id : order) {
        Request& r = *s->reqs[id];
        int want = r.spp == 0 ? s->cfg.min_spp : (int)std::ceil(1.05 * remaining_spp(r, s->cfg.min_spp));
        want = std::max(1, std::min(want, r.spp_max - r.spp));
        const int q = (int)std::min<int64_t>(want, left / std::max(1, r.P));
        if (q < 1) continue;
        ptk::ExtBatchItem it;
        it.scene = r.scene;
        it.pc = r.pc;
        it.pc.samplesPerFrame = q;
        it.pc.frame = r.spp + q;
        it.pc.currentSamples = r.spp + q;
        it.sum = r.sum;
        it.sumsq = r.sumsq;
        it.nsamp = r.nsamp;
        it.sum_odd = r.sum_odd;
        it.n_odd = r.n_odd;
        items.push_back(it);
        in.push_back(&r);
        left -= (int64_t)q * r.P;
        if (s->cfg.policy == PT_SERVE_FIFO || (int)items.size() == ptk::kMaxBatch) break;
    }
    if (items.empty()) return fail(s, PT_ERR_CAPACITY, "pt_server_step: no request fits the wave");

    pt_counters c0{}, c1{};
    (void)pt_counters_get(ctx, &c0);
    const auto t_start = Clock::now();
    // Once a wave may have added samples to the moment buffers, an error leaves
    // those requests' sums ahead of their sample counts; retracing the same
    // sample indices would count them twice, so such requests fail instead.
    auto fail_wave = [&](int code) -> pt_status {
        for (Request* r : in) {
            r->state = PT_REQ_FAILED;
            r->t_done = now_s(s);
            auto it = std::find(s->active.begin(), s->active.end(), r->id);
            if (it != s->active.end()) s->active.erase(it);
        }
        return fail(s, code, ctx->err);
    };
    int st = ptk::ext_dispatch_batch(ctx, items.data(), (int)items.size());
    if (st != PT_OK) return fail_wave(st);
    const size_t nline = (size_t)kRelLines * kRelStride;
    hipError_t e = hipMemsetAsync(s->d_rel, 0, items.size() * nline * sizeof(double), ctx->stream);
    if (e != hipSuccess) return fail_wave(ptk::hip_fail(ctx, e, "pt_server_step"));
    for (size_t i = 0; i < items.size(); ++i) {
        st = ptk::ext_rel_error(ctx, in[i]->P, in[i]->sum, in[i]->sumsq, in[i]->nsamp, s->d_rel + i * nline);
        if (st != PT_OK) return fail_wave(st);
    }
    std::vector<double> lines(items.size() * nline);
    e = hipMemcpyAsync(lines.data(), s->d_rel, lines.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail_wave(ptk::hip_fail(ctx, e, "pt_server_step"));
    const double dt = std::chrono::duration<double>(Clock::now() - t_start).count();
    (void)pt_counters_get(ctx, &c1);
    const double total_paths = (double)(wave - left);
    s->sec_per_path = s->sec_per_path > 0.0 ? 0.8 * s->sec_per_path + 0.2 * dt / total_paths : dt / total_paths;
    // every traced request takes its samples even when a denoiser run fails;
    // the failing request ends PT_REQ_FAILED and the first error is returned
    pt_status first_err = PT_OK;
    std::string first_msg;
    for (size_t i = 0; i < items.size(); ++i) {
        Request& r = *in[i];
        double v = 0.0, m = 0.0;
        for (int l = 0; l < kRelLines; ++l) {
            v += lines[i * nline + (size_t)l * kRelStride];
            m += lines[i * nline + (size_t)l * kRelStride + 1];
        }
        const int q = items[i].pc.samplesPerFrame;
        const double share = (double)q * r.P / total_paths;
        r.spp += q;
        r.rel = m > 0.0 ? std::sqrt(v / m) : INFINITY;
        r.mean2 = m;
        r.gpu_s += dt * share;
        r.segments += (uint64_t)((double)(c1.segments - c0.segments) * share);
        bool done = (r.spp >= s->cfg.min_spp && r.rel <= r.target) || r.spp >= r.spp_max;
        bool failed = false;
        if (r.denoise && r.spp >= s->cfg.min_spp && !(r.rel <= r.target)) {
            st = denoise_step(s, r, r.spp >= r.spp_max);
            if (st != PT_OK) {
                if (first_err == PT_OK) {
                    first_err = st;
                    first_msg = s->err;
                }
                r.state = PT_REQ_FAILED;
                r.t_done = now_s(s);
                s->active.erase(std::find(s->active.begin(), s->active.end(), r.id));
                failed = true;
            }
            done = done || r.denoised;
        }
        if (done && !failed) {
            r.state = PT_REQ_DONE;
            r.t_done = now_s(s);
            s->active.erase(std::find(s->active.begin(), s->active.end(), r.id));
        }
        if (n_stepped && (int)i < cap) {
            (void)pt_server_status(s, r.id, &stepped[i]);
            *n_stepped = (int32_t)i + 1;
        }
    }
    if (first_err != PT_OK) {
        s->err = first_msg;
        return first_err;
    }
    return PT_OK;
}

pt_status pt_server_status(pt_server* s, int32_t id, pt_request_status* out) {
    if (!s || !out) return PT_ERR_INVALID_ARG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(s, PT_ERR_INVALID_ARG, "pt_server_status: unknown request");
    const Request& r = *it->second;
    std::memset(out, 0, sizeof *out);
    out->id = r.id;
    out->state = r.state;
    out->spp = r.spp;
    const double rel = r.denoised ? r.rel_dn : r.rel;
    out->converged = r.spp >= s->cfg.min_spp && rel <= r.target;
    out->rel_error = std::isfinite(rel) ? rel : -1.0;
    out->denoised = r.denoised;
    out->rel_plain = std::isfinite(r.rel) ? r.rel : -1.0;
    out->t_arrive = r.t_arrive;
    out->t_admit = r.t_admit;
    out->t_done = r.t_done;
    out->gpu_seconds = r.gpu_s;
    out->segments = r.segments;
    out->width = r.W;
    out->height = r.H;
    return PT_OK;
}

pt_status pt_server_read(pt_server* s, int32_t id, float* host_rgba, size_t n_floats) {
    if (!s || !host_rgba) return PT_ERR_INVALID_ARG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(s, PT_ERR_INVALID_ARG, "pt_server_read: unknown request");
    Request& r = *it->second;
    if (!r.mem.p) return fail(s, PT_ERR_STATE, "pt_server_read: request not admitted or released");
    if (n_floats < (size_t)r.P * 4) return fail(s, PT_ERR_CAPACITY, "pt_server_read: buffer too small");
    const size_t P = (size_t)r.P;
    pt_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    if (r.denoised) {
        std::vector<float> den(3 * P);
        hipError_t e = hipMemcpyAsync(den.data(), r.den, 3 * P * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return fail(s, ptk::hip_fail(ctx, e, "pt_server_read"), ctx->err);
        for (size_t i = 0; i < P; ++i) {
            for (int c = 0; c < 3; ++c) host_rgba[4 * i + c] = den[3 * i + c];
            host_rgba[4 * i + 3] = 1.0f;
        }
        return PT_OK;
    }
    std::vector<float> sum(3 * P);
    std::vector<int> n(P);
    hipError_t e = hipMemcpyAsync(sum.data(), r.sum, 3 * P * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(n.data(), r.nsamp, P * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(s, ptk::hip_fail(ctx, e, "pt_server_read"), ctx->err);
    // the texel units of k_ext_finish: mean per-sample XYZ x apertureSize^2 x ISO
    const float ex = r.pc.apertureSize * r.pc.apertureSize * (float)r.pc.ISO;
    for (size_t i = 0; i < P; ++i) {
        const float inv = n[i] > 0 ? ex / (float)n[i] : 0.0f;
        host_rgba[4 * i + 0] = sum[i] * inv;
        host_rgba[4 * i + 1] = sum[P + i] * inv;
        host_rgba[4 * i + 2] = sum[2 * P + i] * inv;
        host_rgba[4 * i + 3] = 1.0f;
    }
    return PT_OK;
}

pt_status pt_server_release(pt_server* s, int32_t id) {
    if (!s) return PT_ERR_INVALID_ARG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(s, PT_ERR_INVALID_ARG, "pt_server_release: unknown request");
    Request& r = *it->second;
    if (r.state == PT_REQ_ACTIVE) return fail(s, PT_ERR_STATE, "pt_server_release: request still active");
    if (r.state == PT_REQ_WAITING) s->waiting.erase(std::find(s->waiting.begin(), s->waiting.end(), id));
    (void)hipStreamSynchronize(s->ctx->stream);
    release_request(s, r);
    s->reqs.erase(it);
    return PT_OK;
}

pt_status pt_server_set_denoiser(pt_server* s, pt_denoiser* dn) {
    if (!s) return PT_ERR_INVALID_ARG;
    if (dn && ptk::dn_ctx(dn) != s->ctx) return fail(s, PT_ERR_INVALID_ARG, "pt_server_set_denoiser: another context");
    for (const auto& kv : s->reqs)
        if (kv.second->denoise && (kv.second->state == PT_REQ_ACTIVE || kv.second->state == PT_REQ_WAITING))
            return fail(s, PT_ERR_STATE, "pt_server_set_denoiser: requests in flight use the denoiser");
    s->dn = dn;
    return PT_OK;
}

pt_status pt_server_counts(const pt_server* s, int32_t* waiting, int32_t* active, int64_t* bytes_in_use) {
    if (!s) return PT_ERR_INVALID_ARG;
    if (waiting) *waiting = (int32_t)s->waiting.size();
    if (active) *active = (int32_t)s->active.size();
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
    return PT_OK;
}

}  // extern "C"
