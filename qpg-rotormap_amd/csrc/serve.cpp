// serve.cpp -- continuous-batching serving loop on one GPU.
//
// Requests (packed fragment words on the host) wait in a FIFO.  A dispatcher
// thread packs the FIFO head into the next free launch slot row by row: a
// request larger than the slot's row capacity is spread over several
// launches, and a launch carries the tails and heads of several requests.
// Each slot owns a stream, its device buffers (words, encoded queries,
// split partials, outputs) and pinned output staging, all allocated once at
// create time, so the loop itself never allocates device memory.  A slot's
// launch is: H2D of its rows, encoder, fused GEMM + top-k, merge / rescoring,
// D2H of (vals, locs), an event.  A completer thread waits on the slots'
// events in launch order, scatters the rows back to their requests and
// frees the slot.  With two or more slots, the host work and the copies of
// one launch overlap the kernels of another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "rm_internal.h"

namespace {

using clk = std::chrono::steady_clock;

enum { Q_QUEUED = 0, Q_ADMITTED = 1, Q_DONE = 2, Q_CANCELLED = 3, Q_FAILED = 4 };

struct Req {
    int64_t id = 0, n = 0;
    std::vector<uint32_t> own;       // copied words (rm_server_submit)
    const uint32_t *words = nullptr; // own.data() or the caller's buffer
    int64_t next_row = 0;            // first row not yet placed in a launch
    int64_t rows_done = 0;
    int pending = 0;                 // pieces of this request in launch slots
    int state = Q_QUEUED;
    double t_sub = 0, t_adm = -1, t_fin = -1;
    std::vector<float> vals;         // n x ktop row-major
    std::vector<int32_t> locs;
    int err = 0;
    std::string msg;
};

struct Piece {
    Req *r;
    int64_t lo, hi, row;  // rows [lo, hi) of r sit at launch rows [row, row + hi - lo)
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    uint32_t *d_words = nullptr;
    void *d_q = nullptr;
    float *d_q32 = nullptr;
    float *d_pv = nullptr;
    int32_t *d_pc = nullptr;
    size_t pbytes = 0;
    float *d_vals = nullptr;
    int32_t *d_locs = nullptr;
    float *h_vals = nullptr;  // pinned
    int32_t *h_locs = nullptr;
    std::vector<Piece> pieces;
    int64_t rows = 0;
};

}  // namespace

struct rm_server {
    rm_ctx *ctx = nullptr;
    const rm_index *idx = nullptr;
    int k = 0, s = 0, m = 0, c = 0, ktop = 0, kprime = 0, mode = 0;
    int64_t max_rows = 0, W = 0, rdim = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req *> queue;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::deque<Slot *> inflight;
    std::vector<Slot *> free_slots;
    int64_t next_id = 0;
    int policy = RM_SCHED_FIFO;
    // low-load mode (rm_server_set_lowload): while at most lowload_backlog
    // rows wait unlaunched, one launch is in flight at a time and it holds
    // whole requests only, at most lowload_rows rows
    int64_t lowload_rows = 0, lowload_backlog = 0;
    bool stop = false;
    bool disp_done = false;  // the dispatcher has exited: no launch will join inflight
    int64_t launches = 0, rows_done = 0, max_inflight = 0;
    clk::time_point t0 = clk::now();
    std::thread disp, comp;
    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

namespace {

using namespace rm;

// device buffers of one slot; part bytes cover the largest split plan any
// batch of <= max_rows rows gets from the plans of the server's creation
// (launch_slot lowers the split count of a plan set later to what fits)
int64_t slot_bytes(rm_server *sv, size_t *pbytes_out) {
    const int KP = sv->mode == RM_MODE_FP8_RESCORE ? sv->kprime : sv->ktop;
    size_t pb = 0;
    constexpr int step = 64;  // the smallest tile: the split rule changes at most this often
    for (int64_t R = 1; R <= cdiv(sv->max_rows, step); ++R) {
        const int64_t M = std::min<int64_t>(R * step, sv->max_rows);
        for (int64_t Mi : {M, std::max<int64_t>(1, M - step + 1)}) {
            const int S = plan_for(sv->ctx, sv->idx, Mi, sv->mode, KP).S;
            pb = std::max(pb, part_bytes(sv->idx, Mi, KP, S));
        }
    }
    if (pbytes_out) *pbytes_out = pb;
    const int64_t M = sv->max_rows;
    int64_t b = M * sv->W * 4;
    b += M * sv->rdim * (sv->mode == RM_MODE_FP16 ? 2 : 1);
    if (sv->mode == RM_MODE_FP8_RESCORE) b += M * sv->rdim * 4;
    b += 2 * (int64_t)pb;
    b += M * (int64_t)sv->ktop * 8;
    return b;
}

void free_slot(Slot *sl) {
    if (sl->done) hipEventDestroy(sl->done);
    if (sl->st) hipStreamDestroy(sl->st);
    for (void *p : {(void *)sl->d_words, sl->d_q, (void *)sl->d_q32, (void *)sl->d_pv,
                    (void *)sl->d_pc, (void *)sl->d_vals, (void *)sl->d_locs})
        if (p) hipFree(p);
    if (sl->h_vals) hipHostFree(sl->h_vals);
    if (sl->h_locs) hipHostFree(sl->h_locs);
}

int alloc_slot(rm_server *sv, Slot *sl, size_t pbytes) {
    const int64_t M = sv->max_rows;
    RM_HIP(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
    RM_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    RM_HIP(hipMalloc(&sl->d_words, (size_t)M * sv->W * 4));
    RM_HIP(hipMalloc(&sl->d_q, (size_t)M * sv->rdim * (sv->mode == RM_MODE_FP16 ? 2 : 1)));
    if (sv->mode == RM_MODE_FP8_RESCORE) RM_HIP(hipMalloc(&sl->d_q32, (size_t)M * sv->rdim * 4));
    sl->pbytes = std::max<size_t>(pbytes, 256);
    RM_HIP(hipMalloc(&sl->d_pv, sl->pbytes));
    RM_HIP(hipMalloc(&sl->d_pc, sl->pbytes));
    RM_HIP(hipMalloc(&sl->d_vals, (size_t)M * sv->ktop * 4));
    RM_HIP(hipMalloc(&sl->d_locs, (size_t)M * sv->ktop * 4));
    RM_HIP(hipHostMalloc((void **)&sl->h_vals, (size_t)M * sv->ktop * 4, hipHostMallocDefault));
    RM_HIP(hipHostMalloc((void **)&sl->h_locs, (size_t)M * sv->ktop * 4, hipHostMallocDefault));
    return RM_OK;
}

// enqueue one slot's launch on its stream (the slot's pieces are set)
int launch_slot(rm_server *sv, Slot *sl) {
    RM_HIP(hipSetDevice(sv->ctx->device));
    const int64_t rows = sl->rows, W = sv->W;
    for (const Piece &p : sl->pieces)
        RM_HIP(hipMemcpyAsync(sl->d_words + p.row * W, p.r->words + p.lo * W,
                              (size_t)(p.hi - p.lo) * W * 4, hipMemcpyHostToDevice, sl->st));
    int rc;
    {
        std::lock_guard<std::recursive_mutex> lk(sv->ctx->mu);
        const int KP = sv->mode == RM_MODE_FP8_RESCORE ? sv->kprime : sv->ktop;
        const int elem = sv->mode == RM_MODE_FP16 ? 1 : 0;
        if ((rc = encode_into(sv->ctx, sl->d_words, rows, sv->k, sv->s, sv->m, sv->c, sv->mode,
                              sv->rdim, nullptr, sl->d_q, sl->d_q32, sl->st)))
            return rc;
        // the context's plan may have changed since the slot was sized (a later
        // rm_autotune / rm_plan_set): keep at most the splits whose partials
        // fit the slot -- results do not depend on the split count
        Plan pl = plan_for(sv->ctx, sv->idx, rows, sv->mode, KP);
        pl.S = max_splits_within(sv->idx, rows, KP, pl.S, sl->pbytes);
        if ((rc = topk_core_bufs(sv->ctx, sv->idx, sl->d_q, sv->rdim * (elem ? 2 : 1), elem,
                                 sl->d_q32, sv->rdim, rows, sv->ktop, sv->kprime, sv->mode,
                                 sl->d_vals, sl->d_locs, rows, pl, sl->d_pv, sl->d_pc, sl->pbytes,
                                 sl->st)))
            return rc;
    }
    RM_HIP(hipMemcpyAsync(sl->h_vals, sl->d_vals, (size_t)rows * sv->ktop * 4,
                          hipMemcpyDeviceToHost, sl->st));
    RM_HIP(hipMemcpyAsync(sl->h_locs, sl->d_locs, (size_t)rows * sv->ktop * 4,
                          hipMemcpyDeviceToHost, sl->st));
    RM_HIP(hipEventRecord(sl->done, sl->st));
    return RM_OK;
}

void fail_pieces(rm_server *sv, Slot *sl, int code, const std::string &msg) {
    const double t = sv->now();
    for (const Piece &p : sl->pieces) {
        Req *r = p.r;
        --r->pending;
        if (r->state == Q_FAILED) continue;
        r->state = Q_FAILED;
        r->err = code;
        r->msg = msg;
        r->t_fin = t;
        // rows of this request still waiting in the FIFO are dropped with it
        auto it = std::find(sv->queue.begin(), sv->queue.end(), r);
        if (it != sv->queue.end()) sv->queue.erase(it);
    }
}

// low-load mode: a small backlog of unlaunched rows (caller holds sv->mu)
bool low_load(const rm_server *sv) {
    if (sv->lowload_rows <= 0) return false;
    int64_t backlog = 0;
    for (const Req *r : sv->queue) backlog += r->n - r->next_row;
    return backlog <= sv->lowload_backlog;
}

void dispatcher(rm_server *sv) {
    hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        // at low load a launch waits for the GPU to be idle, so a request
        // that arrives meanwhile is not queued behind a second launch
        sv->cv.wait(lk, [&] {
            return sv->stop || (!sv->queue.empty() && !sv->free_slots.empty() &&
                                (sv->inflight.empty() || !low_load(sv)));
        });
        if (sv->stop) break;
        const bool low = low_load(sv);
        Slot *sl = sv->free_slots.back();
        sv->free_slots.pop_back();
        sl->pieces.clear();
        int64_t rows = 0;
        const double t = sv->now();
        bool emptied = false;
        const int64_t cap = low ? std::min(sv->max_rows, sv->lowload_rows) : sv->max_rows;
        auto take_rows = [&](Req *r, int64_t want) {
            const int64_t take = std::min({r->n - r->next_row, cap - rows, want});
            if (take <= 0) return;
            sl->pieces.push_back({r, r->next_row, r->next_row + take, rows});
            ++r->pending;
            if (r->state == Q_QUEUED) {
                r->state = Q_ADMITTED;
                r->t_adm = t;
            }
            r->next_row += take;
            rows += take;
            emptied |= r->next_row == r->n;
        };
        // the queue stays in arrival order; SRPT visits it by remaining rows,
        // after the oldest request has taken up to a quarter of the launch
        // (so a large request keeps progressing while small ones keep coming)
        std::vector<Req *> order(sv->queue.begin(), sv->queue.end());
        auto srpt_sort = [&] {
            std::stable_sort(order.begin(), order.end(), [](const Req *x, const Req *y) {
                return x->n - x->next_row < y->n - y->next_row;
            });
        };
        if (low) {
            // whole requests in the policy's order while they fit; a first
            // request larger than the cap gets a slice of it.  A small
            // request then rides a short launch of its own instead of one
            // padded with a large request's rows.
            if (sv->policy == RM_SCHED_SRPT) srpt_sort();
            for (Req *r : order) {
                const int64_t left = r->n - r->next_row;
                if (rows > 0 && left > cap - rows) break;
                take_rows(r, cap);
            }
        } else {
            if (sv->policy == RM_SCHED_SRPT) {
                take_rows(order.front(), std::max<int64_t>(1, cap / 4));
                srpt_sort();
            }
            for (Req *r : order) {
                if (rows == cap) break;
                take_rows(r, cap);
            }
        }
        if (emptied)
            sv->queue.erase(std::remove_if(sv->queue.begin(), sv->queue.end(),
                                           [](const Req *r) { return r->next_row == r->n; }),
                            sv->queue.end());
        sl->rows = rows;
        lk.unlock();
        const int rc = launch_slot(sv, sl);
        const std::string msg = rc ? std::string(rm_last_error()) : std::string();
        lk.lock();
        if (rc) {
            // nothing of this slot reached its event: fail its requests and
            // hand the slot back once its stream has drained
            lk.unlock();
            hipStreamSynchronize(sl->st);
            lk.lock();
            fail_pieces(sv, sl, rc, msg);
            sv->free_slots.push_back(sl);
            sv->cv.notify_all();
            continue;
        }
        sv->inflight.push_back(sl);
        ++sv->launches;
        sv->max_inflight = std::max<int64_t>(sv->max_inflight, (int64_t)sv->inflight.size());
        sv->cv.notify_all();
    }
}

void completer(rm_server *sv) {
    hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        // the dispatcher may still be enqueueing a launch after stop is set:
        // exit only once it is gone and nothing is in flight
        sv->cv.wait(lk, [&] { return !sv->inflight.empty() || sv->disp_done; });
        if (sv->inflight.empty()) break;
        Slot *sl = sv->inflight.front();
        lk.unlock();
        const hipError_t e = hipEventSynchronize(sl->done);
        if (e == hipSuccess) {
            // scatter (launch row, rank) -> (request row, rank) outside the
            // lock: the pieces' requests stay alive while pending > 0
            const int kt = sv->ktop;
            const int64_t rows = sl->rows;
            for (const Piece &p : sl->pieces) {
                Req *r = p.r;
                for (int64_t i = 0; i < p.hi - p.lo; ++i)
                    for (int j = 0; j < kt; ++j) {
                        r->vals[(p.lo + i) * kt + j] = sl->h_vals[p.row + i + (int64_t)j * rows];
                        r->locs[(p.lo + i) * kt + j] = sl->h_locs[p.row + i + (int64_t)j * rows];
                    }
            }
        }
        lk.lock();
        sv->inflight.pop_front();
        if (e != hipSuccess) {
            fail_pieces(sv, sl, RM_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
        } else {
            const double t = sv->now();
            for (const Piece &p : sl->pieces) {
                Req *r = p.r;
                --r->pending;
                if (r->state == Q_FAILED) continue;
                r->rows_done += p.hi - p.lo;
                sv->rows_done += p.hi - p.lo;
                if (r->rows_done == r->n) {
                    r->state = Q_DONE;
                    r->t_fin = t;
                }
            }
        }
        sl->pieces.clear();
        sv->free_slots.push_back(sl);
        sv->cv.notify_all();
    }
}

int submit_common(rm_server *srv, const uint32_t *words, int64_t n, bool copy, int64_t *req_id) {
    RM_CHECK(srv && req_id, RM_ERR_ARG, "NULL argument");
    RM_CHECK(n >= 0, RM_ERR_ARG, "n < 0");
    RM_CHECK(n == 0 || words, RM_ERR_ARG, "words is NULL");
    auto r = std::make_unique<Req>();
    r->n = n;
    if (copy) {
        r->own.assign(words, words + n * srv->W);
        r->words = r->own.data();
    } else {
        r->words = words;
    }
    r->vals.assign((size_t)n * srv->ktop, 0.0f);
    r->locs.assign((size_t)n * srv->ktop, 0);
    std::lock_guard<std::mutex> lk(srv->mu);
    RM_CHECK(!srv->stop, RM_ERR_STATE, "server is shutting down");
    r->id = srv->next_id++;
    r->t_sub = srv->now();
    if (n == 0) {  // nothing to compute: done on arrival
        r->state = Q_DONE;
        r->t_adm = r->t_fin = r->t_sub;
    } else {
        srv->queue.push_back(r.get());
    }
    *req_id = r->id;
    srv->reqs[r->id] = std::move(r);
    srv->cv.notify_all();
    return RM_OK;
}

}  // namespace

extern "C" {

int rm_server_create(rm_ctx *ctx, const rm_index *idx, int k, int s, int m, int c, int ktop,
                     int kprime, int mode, int64_t max_rows_per_launch, int n_streams,
                     int64_t hbm_budget_bytes, rm_server **out) {
    RM_CHECK(ctx && idx && out, RM_ERR_ARG, "NULL argument");
    *out = nullptr;
    int rc;
    if ((rc = check_encoder(k, s, m, c, 0))) return rc;
    if ((rc = check_query(idx, ktop, kprime, mode))) return rc;
    const int64_t rdim = 2LL * m * (1LL << (2 * c));
    RM_CHECK(rdim == idx->rdim, RM_ERR_SHAPE, "encoder dim 2*m*4^c = %lld != index rdim %lld",
             (long long)rdim, (long long)idx->rdim);
    RM_CHECK(max_rows_per_launch >= 1 && max_rows_per_launch <= (1LL << 22), RM_ERR_ARG,
             "max_rows_per_launch must be in 1..4194304");
    RM_CHECK(n_streams >= 1 && n_streams <= 8, RM_ERR_ARG, "n_streams must be in 1..8");
    RM_HIP(hipSetDevice(ctx->device));
    auto sv = std::make_unique<rm_server>();
    sv->ctx = ctx;
    sv->idx = idx;
    sv->k = k, sv->s = s, sv->m = m, sv->c = c;
    sv->ktop = ktop, sv->kprime = kprime, sv->mode = mode;
    sv->max_rows = max_rows_per_launch;
    sv->W = (k + 15) / 16;
    sv->rdim = rdim;
    // admission: every slot's buffers must fit the free HBM (less 96 MiB of
    // headroom, as the resident-index rule keeps) and the caller's budget
    size_t pb = 0;
    const int64_t per = slot_bytes(sv.get(), &pb);
    const int64_t need = per * n_streams;
    int64_t avail = rm_device_free_bytes(ctx) - (96LL << 20);
    if (hbm_budget_bytes > 0) avail = std::min(avail, hbm_budget_bytes);
    RM_CHECK(need <= avail, RM_ERR_NOMEM,
             "server needs %lld bytes (%d slots x %lld rows); %lld bytes available",
             (long long)need, n_streams, (long long)max_rows_per_launch, (long long)avail);
    const int32_t *tab;
    if ((rc = table_get(ctx, k, s, m, &tab))) return rc;
    for (int i = 0; i < n_streams; ++i) {
        sv->slots.push_back(std::make_unique<Slot>());
        if ((rc = alloc_slot(sv.get(), sv->slots.back().get(), pb))) {
            for (auto &sl : sv->slots) free_slot(sl.get());
            return rc;
        }
        sv->free_slots.push_back(sv->slots.back().get());
    }
    sv->disp = std::thread(dispatcher, sv.get());
    sv->comp = std::thread(completer, sv.get());
    *out = sv.release();
    return RM_OK;
}

int rm_server_submit(rm_server *srv, const uint32_t *host_words, int64_t n, int64_t *req_id) {
    return submit_common(srv, host_words, n, true, req_id);
}

int rm_server_submit_ref(rm_server *srv, const uint32_t *host_words, int64_t n, int64_t *req_id) {
    return submit_common(srv, host_words, n, false, req_id);
}

int rm_server_wait(rm_server *srv, int64_t req_id, int timeout_ms, float *host_vals,
                   int32_t *host_locs, double *t_submit, double *t_admit, double *t_finish) {
    RM_CHECK(srv, RM_ERR_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    RM_CHECK(it != srv->reqs.end(), RM_ERR_ARG, "unknown request id %lld", (long long)req_id);
    Req *r = it->second.get();
    // a failed request returns once none of its rows is in flight any more:
    // its words (submit_ref) are no longer read by a queued copy
    auto ready = [&] { return r->state >= Q_DONE && !(r->state == Q_FAILED && r->pending > 0); };
    if (timeout_ms < 0) {
        srv->cv.wait(lk, ready);
    } else if (!srv->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return fail(RM_ERR_STATE, "request %lld not finished within %d ms", (long long)req_id,
                    timeout_ms);
    }
    if (t_submit) *t_submit = r->t_sub;
    if (t_admit) *t_admit = r->t_adm;
    if (t_finish) *t_finish = r->t_fin;
    const int state = r->state;
    const int err = r->err;
    const std::string msg = r->msg;
    if (state == Q_DONE) {
        if (host_vals) std::memcpy(host_vals, r->vals.data(), r->vals.size() * 4);
        if (host_locs) std::memcpy(host_locs, r->locs.data(), r->locs.size() * 4);
    }
    srv->reqs.erase(req_id);  // (another waiter of the same id may have erased it)
    if (state == Q_CANCELLED) return fail(RM_ERR_STATE, "request %lld was cancelled", (long long)req_id);
    if (state == Q_FAILED) return fail(err ? err : RM_ERR_HIP, "request %lld failed: %s", (long long)req_id, msg.c_str());
    clear_error();
    return RM_OK;
}

int rm_server_cancel(rm_server *srv, int64_t req_id) {
    RM_CHECK(srv, RM_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(req_id);
    RM_CHECK(it != srv->reqs.end(), RM_ERR_ARG, "unknown request id %lld", (long long)req_id);
    Req *r = it->second.get();
    RM_CHECK(r->state == Q_QUEUED, RM_ERR_STATE, "request %lld is already admitted",
             (long long)req_id);
    auto q = std::find(srv->queue.begin(), srv->queue.end(), r);
    if (q != srv->queue.end()) srv->queue.erase(q);
    r->state = Q_CANCELLED;
    r->t_fin = srv->now();
    srv->cv.notify_all();
    return RM_OK;
}

int rm_server_set_policy(rm_server *srv, int policy) {
    RM_CHECK(srv, RM_ERR_ARG, "NULL argument");
    RM_CHECK(policy == RM_SCHED_FIFO || policy == RM_SCHED_SRPT, RM_ERR_ARG,
             "policy must be RM_SCHED_FIFO or RM_SCHED_SRPT");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->policy = policy;
    return RM_OK;
}

int rm_server_set_lowload(rm_server *srv, int64_t launch_rows, int64_t backlog_rows) {
    RM_CHECK(srv, RM_ERR_ARG, "NULL argument");
    RM_CHECK(launch_rows >= 0 && backlog_rows >= 0, RM_ERR_ARG,
             "launch_rows and backlog_rows must be >= 0");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->lowload_rows = launch_rows;
    srv->lowload_backlog = backlog_rows;
    return RM_OK;
}

int rm_server_stats(rm_server *srv, int64_t *launches, int64_t *rows_done, int64_t *max_inflight) {
    RM_CHECK(srv, RM_ERR_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (launches) *launches = srv->launches;
    if (rows_done) *rows_done = srv->rows_done;
    if (max_inflight) *max_inflight = srv->max_inflight;
    return RM_OK;
}

double rm_server_now(rm_server *srv) { return srv ? srv->now() : -1.0; }

int rm_server_destroy(rm_server *srv) {
    if (!srv) return RM_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        // requests still in the FIFO end as cancelled (never admitted) or
        // failed (admitted, rows left)
        for (Req *r : srv->queue) {
            if (r->state == Q_QUEUED) {
                r->state = Q_CANCELLED;
            } else {
                r->state = Q_FAILED;
                r->err = RM_ERR_STATE;
                r->msg = "server destroyed before the request's last rows were launched";
            }
        }
        srv->queue.clear();
        srv->cv.notify_all();
    }
    // the dispatcher finishes the launch it may be enqueueing, then exits;
    // the completer drains every launch in flight before it exits
    if (srv->disp.joinable()) srv->disp.join();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->disp_done = true;
        srv->cv.notify_all();
    }
    if (srv->comp.joinable()) srv->comp.join();
    hipSetDevice(srv->ctx->device);
    for (auto &sl : srv->slots) free_slot(sl.get());
    delete srv;
    return RM_OK;
}

}  // extern "C"
