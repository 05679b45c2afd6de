// fsg_serve.hip -- continuous-batching request loop over resident worlds.
//
// One worker thread owns every world of the server.  Each step it
//   1. admits queued requests in arrival order while the context's HBM state
//      budget allows (evicting idle resident worlds, least recently used
//      first); a request whose world is busy waits for the earlier one;
//   2. writes the cells of newly created worlds with the device generator
//      (fsg_world_generate) and uploads their bodies;
//   3. advances all admitted requests together: one batched tick launch per
//      tick (16-tick HIP graphs when no world has bodies), with the rigid
//      bodies of every world stamped before and unstamped after each tick in
//      one claim-round batch, the way game::tick brackets world::tick with the
//      stamp and unstamp loops (game.cpp:1760-1815, 1838-1840, 1896-1983);
//   4. completes requests whose tick count is reached.
// A world's evolution depends only on its seed, its ticks and its bodies, so
// its result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fsg_internal.hpp"

using namespace fsg_impl;

extern "C" uint64_t fsg_now_ns(void) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

namespace {

// Host rigid body standing in for a Box2D body: the server steps a simple
// kinematic model (gravity, wall bounces) because Box2D is not on the path;
// only the pose crosses into the tick, as b2Body::GetPosition/GetAngle do.
struct HostBody {
    double x, y, vx, vy, ang, w;
    int mw, mh;
    uint32_t off;
};

struct Resident {
    uint64_t id = 0;
    fsg_world *w = nullptr;
    bool busy = false;
    uint64_t last_used = 0;
    uint64_t checksum = 0;
    bool checksum_known = false;
    std::vector<HostBody> hb;
    std::vector<fsg_body> poses;
    std::vector<fsg_cell> px;  // body pixels as uploaded (kept while the copy may be in flight)
};

struct Req {
    uint64_t id = 0;
    fsg_serve_request r{};
    uint64_t t_submit = 0, t_admit = 0;
    int done_ticks = 0;
    bool cancelled = false;
    Resident *res = nullptr;
};


constexpr double kDt = 1.0 / 30.3;  // b2world->Step interval (world.cpp:2276 at maxTps 30)
constexpr double kGravity = 20.0;   // b2World gravity (0, 20), world.cpp:147-148

uint64_t world_bytes(int W, int H) {
    fsg_config cfg;
    fsg_default_config(&cfg);  // served worlds take the default config
    return fsg_impl::world_state_bytes(W, H, cfg.n_materials);
}

// Deterministic bodies of a served world: boxes 4-64 px and convex polygons
// of 3-8 vertices (rasterised), stone or wood, random poses in the upper
// third, small random velocities.
void make_bodies(uint64_t seed, int n, int W, int H, std::vector<HostBody> &hb, std::vector<fsg_cell> &px) {
    uint64_t st = fsg_splitmix64(seed ^ 0xB0D1E5ull);
    auto u01 = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        return (double)(fsg_splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
    };
    hb.clear();
    px.clear();
    for (int i = 0; i < n; i++) {
        HostBody b;
        const bool box = u01() < 0.5;
        b.mw = 4 + (int)(u01() * 61);
        b.mh = box ? 4 + (int)(u01() * 61) : b.mw;
        b.off = (uint32_t)px.size();
        const uint16_t mat = u01() < 0.25 ? 7 : 1;  // wood or stone
        std::vector<double> vx, vy;
        if (!box) {
            const int k = 3 + (int)(u01() * 6);
            for (int v = 0; v < k; v++) {
                const double a = 2.0 * M_PI * (v + 0.8 * u01()) / k;
                vx.push_back(0.5 * b.mw * (1.0 + std::cos(a)));
                vy.push_back(0.5 * b.mh * (1.0 + std::sin(a)));
            }
        }
        for (int ty = 0; ty < b.mh; ty++)
            for (int tx = 0; tx < b.mw; tx++) {
                bool inside = true;
                if (!box) {  // convex: the centre is on the same side of every edge
                    const double cx = tx + 0.5, cy = ty + 0.5;
                    const size_t k = vx.size();
                    for (size_t e = 0; e < k && inside; e++) {
                        const size_t f = (e + 1) % k;
                        inside = (vx[f] - vx[e]) * (cy - vy[e]) - (vy[f] - vy[e]) * (cx - vx[e]) >= 0;
                    }
                }
                fsg_cell c;
                c.mat = inside ? mat : 0;
                c.temp = 320;
                c.rgba = inside ? (mat == 7 ? 0xFF2D5280u : 0xFF707070u) : 0u;
                px.push_back(c);
            }
        b.x = u01() * W;
        b.y = u01() * H / 3.0;
        b.ang = u01() * 2.0 * M_PI;
        b.vx = (u01() - 0.5) * 10.0;
        b.vy = (u01() - 0.5) * 4.0;
        b.w = (u01() - 0.5) * 2.0;
        hb.push_back(b);
    }
}

void poses_of(const std::vector<HostBody> &hb, std::vector<fsg_body> &out) {
    out.resize(hb.size());
    for (size_t i = 0; i < hb.size(); i++) {
        const HostBody &b = hb[i];
        fsg_body &p = out[i];
        p.pos_x_q16 = (int32_t)std::llround(b.x * 65536.0);
        p.pos_y_q16 = (int32_t)std::llround(b.y * 65536.0);
        p.cos_q16 = (int32_t)std::llround(std::cos(b.ang) * 65536.0);
        p.sin_q16 = (int32_t)std::llround(std::sin(b.ang) * 65536.0);
        p.mat_w = b.mw;
        p.mat_h = b.mh;
        p.pixel_offset = b.off;
        p.enabled = 1;
    }
}

void step_bodies(std::vector<HostBody> &hb, int W, int H) {
    for (HostBody &b : hb) {
        b.vy += kGravity * kDt;
        b.x += b.vx * kDt;
        b.y += b.vy * kDt;
        b.ang += b.w * kDt;
        if (b.y > H - 8) {
            b.y = H - 8;
            b.vy = -0.5 * std::fabs(b.vy);
        }
        if (b.x < 0) {
            b.x = 0;
            b.vx = std::fabs(b.vx);
        }
        if (b.x > W - 1) {
            b.x = W - 1;
            b.vx = -std::fabs(b.vx);
        }
    }
}

}  // namespace

struct fsg_server {
    fsg_ctx *ctx = nullptr;
    fsg_serve_config cfg{};
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Req> queue;                 // submitted, not admitted   (mu)
    std::vector<fsg_serve_result> done;    // completed, not polled     (mu)
    std::vector<uint64_t> releases;        // world ids to drop         (mu)
    std::vector<uint64_t> corruptions;     // test hook requests        (mu)
    uint64_t next_id = 1;
    bool stop = false;
    fsg_serve_stats st{};
    // worker-owned
    std::unordered_map<uint64_t, std::unique_ptr<Resident>> res;
    std::vector<Req> active;
    // per-step staging in pinned host memory: the stamp tables of the step
    // (tick t: stamp at [2tn, (2t+1)n), unstamp at [(2t+1)n, (2t+2)n)) and
    // every world's body poses of every tick (one upload); reused after the
    // step's final synchronisation
    fsg::StampWorld *d_table = nullptr, *h_table = nullptr;
    int table_cap = 0;
    fsg_body *d_pose = nullptr, *h_pose = nullptr;
    size_t pose_cap = 0;
    // census results of the requests a step completes (18 words each, pinned)
    unsigned long long *h_sum = nullptr;
    size_t sum_cap = 0;
    std::vector<size_t> pose_off;
    std::thread th;
};

namespace {

// sum: the world's census words from stats_enqueue, or null to compute them.
void finish(fsg_server *S, Req &q, int status, std::vector<fsg_serve_result> &out,
            const unsigned long long *sum = nullptr) {
    fsg_serve_result r{};
    r.request_id = q.id;
    r.world_id = q.r.world_id;
    r.status = status;
    r.ticks_done = q.done_ticks;
    r.t_submit_ns = q.t_submit;
    r.t_admit_ns = q.t_admit;
    r.t_done_ns = fsg_now_ns();
    if (q.res && q.res->w) {
        int W = 0, H = 0;
        uint64_t t = 0;
        fsg_world_info(q.res->w, &W, &H, &t);
        r.tick = t;
        r.cells = (uint64_t)W * H;
        if (status == FSG_OK) {
            fsg_stats s;
            if (sum) {
                r.checksum = sum[16];
                q.res->checksum = sum[16];
                q.res->checksum_known = true;
            } else if (fsg_world_stats(q.res->w, &s) == FSG_OK) {
                r.checksum = s.checksum;
                q.res->checksum = s.checksum;
                q.res->checksum_known = true;
            } else {
                r.status = FSG_ERR_DEVICE;
            }
        }
        q.res->busy = false;
        q.res->last_used = r.t_done_ns;
    }
    (void)S;
    out.push_back(r);
}

void drop_world(fsg_server *S, uint64_t id) {
    auto it = S->res.find(id);
    if (it == S->res.end() || it->second->busy) return;
    fsg_world_destroy(it->second->w);
    S->res.erase(it);
}

// Frees idle resident worlds (least recently used first) until `need` bytes
// fit the budget.  True if they fit.
bool make_room(fsg_server *S, uint64_t need) {
    fsg_ctx *c = S->ctx;
    if (!c->budget) return true;
    for (;;) {
        uint64_t used = 0;
        fsg_ctx_usage(c, &used);
        if (used + need <= c->budget) return true;
        Resident *lru = nullptr;
        for (auto &kv : S->res)
            if (!kv.second->busy && (!lru || kv.second->last_used < lru->last_used)) lru = kv.second.get();
        if (!lru) return false;
        drop_world(S, lru->id);
        S->st.evictions++;
    }
}

// Admission (called with S->mu held).
void admit(fsg_server *S, std::vector<fsg_serve_result> &out) {
    const uint64_t now = fsg_now_ns();
    const size_t cap = (size_t)(S->cfg.max_active > 0 ? S->cfg.max_active : 1024);
    std::vector<uint64_t> blocked;  // worlds with an earlier queued request still waiting
    for (auto it = S->queue.begin(); it != S->queue.end();) {
        Req &q = *it;
        if (q.cancelled || (q.r.deadline_ns && now > q.r.deadline_ns)) {
            finish(S, q, q.cancelled ? FSG_ERR_CANCELLED : FSG_ERR_TIMEOUT, out);
            it = S->queue.erase(it);
            continue;
        }
        const uint64_t wid = q.r.world_id;
        if (std::find(blocked.begin(), blocked.end(), wid) != blocked.end() ||
            S->active.size() >= cap) {
            blocked.push_back(wid);
            ++it;
            continue;
        }
        auto rit = S->res.find(wid);
        if (rit != S->res.end()) {
            Resident *r = rit->second.get();
            if (r->busy) {
                blocked.push_back(wid);
                ++it;
                continue;
            }
            q.res = r;
            r->busy = true;
            if (S->cfg.verify_resident && r->checksum_known) {
                fsg_stats s;
                if (fsg_world_stats(r->w, &s) != FSG_OK || s.checksum != r->checksum) {
                    // the resident state no longer matches what the last request left
                    r->busy = false;
                    q.res = nullptr;
                    finish(S, q, FSG_ERR_CHECKSUM, out);
                    drop_world(S, wid);
                    it = S->queue.erase(it);
                    continue;
                }
            }
            q.t_admit = now;
            S->active.push_back(q);
            it = S->queue.erase(it);
            continue;
        }
        // a new world
        const fsg_serve_request &rq = q.r;
        if (rq.width <= 0 || rq.height <= 0) {
            finish(S, q, FSG_ERR_INVALID, out);
            it = S->queue.erase(it);
            continue;
        }
        const uint64_t need = world_bytes(rq.width, rq.height);
        if (!make_room(S, need)) {
            if (S->active.empty()) {  // can never fit
                finish(S, q, FSG_ERR_BUDGET, out);
                it = S->queue.erase(it);
                continue;
            }
            break;  // wait for running requests to free memory; keep arrival order
        }
        fsg_config cfg;
        fsg_default_config(&cfg);
        fsg_world *w = nullptr;
        const int rc = fsg_world_create(S->ctx, rq.width, rq.height, rq.seed, &cfg, &w);
        if (rc != FSG_OK) {
            if (rc == FSG_ERR_BUDGET && !S->active.empty()) break;
            finish(S, q, rc, out);
            it = S->queue.erase(it);
            continue;
        }
        auto r = std::make_unique<Resident>();
        r->id = wid;
        r->w = w;
        r->busy = true;
        Resident *rp = r.get();
        q.res = rp;
        q.t_admit = now;
        S->res[wid] = std::move(r);
        // the world's cells come from the device generator; bodies from the
        // host kinematic model
        int grc = fsg_world_generate(w, rq.kind, rq.seed);
        if (grc == FSG_OK && rq.n_bodies > 0) {
            make_bodies(rq.seed, std::min(rq.n_bodies, 4096), rq.width, rq.height, rp->hb, rp->px);
            poses_of(rp->hb, rp->poses);
            grc = set_bodies(w, rp->poses.data(), (int)rp->poses.size(), rp->px.data(), rp->px.size());
        }
        if (grc != FSG_OK) {
            finish(S, q, grc, out);
            rp->busy = false;
            drop_world(S, wid);
            it = S->queue.erase(it);
            continue;
        }
        S->active.push_back(q);
        it = S->queue.erase(it);
    }
}

// Claim rounds of the worlds with bodies for tick t of the step: stamp or
// unstamp table slice 2t + unstamp, bodies read from tick t's poses.  The
// table is built when the rounds are enqueued, after the ticks before it, so
// it names the world's buffers of that tick.
int stamp_batch(fsg_server *S, const std::vector<Resident *> &bw, bool unstamp, int t) {
    fsg_ctx *c = S->ctx;
    const int n = (int)bw.size();
    const size_t slice = (size_t)(2 * t + (unstamp ? 1 : 0)) * n;
    fsg::StampWorld *h = S->h_table + slice;
    uint32_t total = 0, max_npix = 0;
    for (int i = 0; i < n; i++) {
        h[i] = stamp_world_of(bw[i]->w, total);
        h[i].bodies = S->d_pose + S->pose_off[(size_t)t * n + i];
        bw[i]->w->prev_valid = false;  // claim slots stay in the spare buffer (nothing renders served worlds)
        total += (uint32_t)bw[i]->w->npix;
        max_npix = std::max(max_npix, (uint32_t)bw[i]->w->npix);
    }
    fsg::StampWorld *d = S->d_table + slice;
    HIPCHK(hipMemcpyAsync(d, h, (size_t)n * sizeof(fsg::StampWorld), hipMemcpyHostToDevice, c->stream));
    return stamp_rounds(c, d, n, total, max_npix, bw[0]->w->d.mat, unstamp, false);
}

// Grows the pinned host / device staging of a step.
int ensure_staging(fsg_server *S, int nworlds, size_t nbodies) {
    fsg_ctx *c = S->ctx;
    if (nworlds > S->table_cap) {
        const int cap = std::max(nworlds, 2 * S->table_cap);
        dev_free(S->d_table);
        if (S->h_table) (void)hipHostFree(S->h_table);
        S->d_table = nullptr;
        S->h_table = nullptr;
        S->table_cap = 0;
        void *p;
        if (dev_alloc(c, &p, 2 * (size_t)cap * sizeof(fsg::StampWorld))) return FSG_ERR_NOMEM;
        S->d_table = (fsg::StampWorld *)p;
        if (hipHostMalloc(&p, 2 * (size_t)cap * sizeof(fsg::StampWorld)) != hipSuccess) return FSG_ERR_NOMEM;
        S->h_table = (fsg::StampWorld *)p;
        S->table_cap = cap;
    }
    if (nbodies > S->pose_cap) {
        const size_t cap = std::max(nbodies, 2 * S->pose_cap);
        dev_free(S->d_pose);
        if (S->h_pose) (void)hipHostFree(S->h_pose);
        S->d_pose = nullptr;
        S->h_pose = nullptr;
        S->pose_cap = 0;
        void *p;
        if (dev_alloc(c, &p, cap * sizeof(fsg_body))) return FSG_ERR_NOMEM;
        S->d_pose = (fsg_body *)p;
        if (hipHostMalloc(&p, cap * sizeof(fsg_body)) != hipSuccess) return FSG_ERR_NOMEM;
        S->h_pose = (fsg_body *)p;
        S->pose_cap = cap;
    }
    return FSG_OK;
}

// Enqueues T ticks of every active request.  Without bodies: one batched
// launch sequence of T ticks.  With bodies, each tick is stamp, batched tick,
// unstamp; the host body model (which reads nothing back from the device)
// gives the poses of all T ticks up front, uploaded once, so a step with
// bodies also needs only one wait.  The caller synchronises.
int run_step(fsg_server *S, int T) {
    std::vector<fsg_world *> ws;
    std::vector<Resident *> bw;  // worlds with bodies
    for (Req &q : S->active) {
        ws.push_back(q.res->w);
        if (!q.res->hb.empty()) bw.push_back(q.res);
    }
    fsg_ctx *c = S->ctx;
    std::lock_guard<std::mutex> lk(c->batch_mu);  // the context's batch staging
    bool graph_ok = false;
    if (bw.empty()) {
        if (int rc = batch_prepare(c, ws.data(), (int)ws.size(), &graph_ok)) return rc;
        return batch_run(c, ws.data(), (int)ws.size(), T, graph_ok);
    }
    size_t nb = 0;  // bodies per tick
    for (Resident *r : bw) nb += r->hb.size();
    if (int rc = ensure_staging(S, (int)bw.size() * T, nb * T)) return rc;
    S->pose_off.clear();
    size_t off = 0;
    for (int t = 0; t < T; t++) {  // poses of tick t at pose_off[t * n + i]
        for (Resident *r : bw) {
            poses_of(r->hb, r->poses);
            std::memcpy(S->h_pose + off, r->poses.data(), r->poses.size() * sizeof(fsg_body));
            S->pose_off.push_back(off);
            off += r->poses.size();
            step_bodies(r->hb, r->w->d.W, r->w->d.H);
        }
    }
    HIPCHK(hipMemcpyAsync(S->d_pose, S->h_pose, off * sizeof(fsg_body), hipMemcpyHostToDevice, c->stream));
    if (int rc = batch_prepare(c, ws.data(), (int)ws.size(), &graph_ok)) return rc;
    for (int t = 0; t < T; t++) {
        if (int rc = stamp_batch(S, bw, false, t)) return rc;
        if (int rc = batch_run(c, ws.data(), (int)ws.size(), 1, graph_ok)) return rc;
        if (int rc = stamp_batch(S, bw, true, t)) return rc;
    }
    return FSG_OK;
}

// One loop iteration: enqueue a step of the active requests, admit new ones
// while it runs (their worlds are created and generated on the same stream,
// behind the step; admission never touches a world of the running step, which
// is busy), then wait for the step and account for the requests that ran in
// it -- the first n_run entries of `active`, since admission only appends.
void worker(fsg_server *S) {
    for (;;) {
        std::vector<fsg_serve_result> out;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop) break;
            for (uint64_t id : S->releases) drop_world(S, id);
            S->releases.clear();
            for (uint64_t id : S->corruptions) {
                auto it = S->res.find(id);
                if (it != S->res.end() && !it->second->busy) {
                    fsg_cell_edit e{};
                    e.cell.mat = 2;
                    e.cell.temp = 12345;
                    fsg_world_set_cells(it->second->w, &e, 1);
                }
            }
            S->corruptions.clear();
            if (S->active.empty()) {
                admit(S, out);
                if (S->active.empty() && out.empty()) {
                    S->cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
        }
        if (!S->active.empty()) {
            // cancelled or late active requests leave before the step
            const uint64_t now = fsg_now_ns();
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto it = S->active.begin(); it != S->active.end();) {
                if (it->cancelled || (it->r.deadline_ns && now > it->r.deadline_ns)) {
                    finish(S, *it, it->cancelled ? FSG_ERR_CANCELLED : FSG_ERR_TIMEOUT, out);
                    it = S->active.erase(it);
                } else {
                    ++it;
                }
            }
        }
        const size_t n_run = S->active.size();
        int T = 0, rc = FSG_OK;
        if (n_run) {
            T = S->cfg.block_ticks > 0 ? S->cfg.block_ticks : 16;
            for (Req &q : S->active) T = std::min(T, q.r.ticks - q.done_ticks);
            rc = run_step(S, T);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            admit(S, out);  // overlaps the step
        }
        if (n_run) {
            // the census of every request this step completes, behind the step,
            // read back with the step's one wait (the pinned staging of this
            // step is rewritten by the next one, so the wait is always made)
            size_t ndone = 0;
            if (rc == FSG_OK) {
                for (size_t k = 0; k < n_run; k++)
                    if (S->active[k].done_ticks + T >= S->active[k].r.ticks) ndone++;
                if (ndone > S->sum_cap) {
                    if (S->h_sum) (void)hipHostFree(S->h_sum);
                    S->h_sum = nullptr;
                    S->sum_cap = 0;
                    void *p = nullptr;
                    if (hipHostMalloc(&p, std::max<size_t>(ndone, 64) * 18 * 8) != hipSuccess) rc = FSG_ERR_NOMEM;
                    else {
                        S->h_sum = static_cast<unsigned long long *>(p);
                        S->sum_cap = std::max<size_t>(ndone, 64);
                    }
                }
                size_t j = 0;
                for (size_t k = 0; k < n_run && rc == FSG_OK; k++)
                    if (S->active[k].done_ticks + T >= S->active[k].r.ticks)
                        rc = stats_enqueue(S->active[k].res->w, S->h_sum + 18 * j++);
            }
            if (hipStreamSynchronize(S->ctx->stream) != hipSuccess && rc == FSG_OK) rc = FSG_ERR_DEVICE;
            uint64_t cu = 0;
            std::lock_guard<std::mutex> lk(S->mu);
            size_t i = 0, j = 0;
            for (size_t k = 0; k < n_run; k++) {
                Req &q = S->active[i];
                if (rc != FSG_OK) {
                    finish(S, q, rc, out);
                    S->active.erase(S->active.begin() + (long)i);
                    continue;
                }
                q.done_ticks += T;
                cu += (uint64_t)T * q.res->w->d.W * q.res->w->d.H;
                if (q.done_ticks >= q.r.ticks) {
                    finish(S, q, FSG_OK, out, S->h_sum + 18 * j++);
                    S->active.erase(S->active.begin() + (long)i);
                } else {
                    ++i;
                }
            }
            S->st.steps++;
            S->st.cell_updates += cu;
        }
        if (!out.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto &r : out) {
                if (r.status == FSG_OK) S->st.completed++;
                else S->st.failed++;
                S->done.push_back(r);
            }
            S->cv_done.notify_all();
        }
    }
    // shutdown: fail what is left
    std::vector<fsg_serve_result> out;
    std::lock_guard<std::mutex> lk(S->mu);
    for (Req &q : S->active) finish(S, q, FSG_ERR_CANCELLED, out);
    for (Req &q : S->queue) finish(S, q, FSG_ERR_CANCELLED, out);
    S->active.clear();
    S->queue.clear();
    for (auto &r : out) S->done.push_back(r);
}

}  // namespace

extern "C" int fsg_serve_create(fsg_ctx *ctx, const fsg_serve_config *cfg, fsg_server **out) {
    if (!ctx || !out) return FSG_ERR_INVALID;
    *out = nullptr;
    fsg_server *S = new (std::nothrow) fsg_server();
    if (!S) return FSG_ERR_NOMEM;
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    if (S->cfg.block_ticks <= 0 || S->cfg.block_ticks > 16) S->cfg.block_ticks = 16;
    S->th = std::thread(worker, S);
    *out = S;
    return FSG_OK;
}

extern "C" int fsg_serve_destroy(fsg_server *S) {
    if (!S) return FSG_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    S->th.join();
    for (auto &kv : S->res) fsg_world_destroy(kv.second->w);
    S->res.clear();
    dev_free(S->d_table);
    dev_free(S->d_pose);
    if (S->h_table) (void)hipHostFree(S->h_table);
    if (S->h_pose) (void)hipHostFree(S->h_pose);
    if (S->h_sum) (void)hipHostFree(S->h_sum);
    delete S;
    return FSG_OK;
}

extern "C" int fsg_serve_submit(fsg_server *S, const fsg_serve_request *req, uint64_t *id) {
    if (!S || !req || req->ticks < 1) return FSG_ERR_INVALID;
    Req q;
    q.r = *req;
    q.t_submit = fsg_now_ns();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return FSG_ERR_STATE;
        q.id = S->next_id++;
        S->queue.push_back(q);
        S->st.submitted++;
    }
    S->cv.notify_all();
    if (id) *id = q.id;
    return FSG_OK;
}

extern "C" int fsg_serve_cancel(fsg_server *S, uint64_t id) {
    if (!S) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    for (Req &q : S->queue)
        if (q.id == id) {
            q.cancelled = true;
            S->cv.notify_all();
            return FSG_OK;
        }
    for (Req &q : S->active)  // read by the worker between steps under mu
        if (q.id == id) {
            q.cancelled = true;
            return FSG_OK;
        }
    return FSG_ERR_NOTFOUND;
}

extern "C" int fsg_serve_poll(fsg_server *S, fsg_serve_result *out, int cap, int *n) {
    if (!S || !n || cap < 0 || (cap && !out)) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    const int k = std::min<int>(cap, (int)S->done.size());
    for (int i = 0; i < k; i++) out[i] = S->done[i];
    S->done.erase(S->done.begin(), S->done.begin() + k);
    *n = k;
    return FSG_OK;
}

extern "C" int fsg_serve_drain(fsg_server *S, uint64_t timeout_ns) {
    if (!S) return FSG_ERR_INVALID;
    const uint64_t end = fsg_now_ns() + timeout_ns;
    std::unique_lock<std::mutex> lk(S->mu);
    while (S->st.completed + S->st.failed < S->st.submitted) {
        if (fsg_now_ns() > end) return FSG_ERR_TIMEOUT;
        S->cv_done.wait_for(lk, std::chrono::milliseconds(5));
    }
    return FSG_OK;
}

extern "C" int fsg_serve_release(fsg_server *S, uint64_t world_id) {
    if (!S) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->releases.push_back(world_id);
    S->cv.notify_all();
    return FSG_OK;
}

extern "C" int fsg_serve_corrupt(fsg_server *S, uint64_t world_id) {
    if (!S) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->corruptions.push_back(world_id);
    S->cv.notify_all();
    return FSG_OK;
}

extern "C" int fsg_serve_stats_get(fsg_server *S, fsg_serve_stats *out) {
    if (!S || !out) return FSG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->queued = S->queue.size();
    out->active = S->active.size();
    out->resident_worlds = S->res.size();
    uint64_t used = 0;
    fsg_ctx_usage(S->ctx, &used);
    out->resident_bytes = used;
    return FSG_OK;
}
