// Native serving engine (mdz_serve_*): the admission and launch loop for
// mixed-codec decode requests, in C++ behind the C ABI.  No counterpart in
// the reference, whose drivers decode one file per process.
//
// A request names a codec (RLE v1, RLE v2, LZSS: the lane-transposed
// containers; Deflate: zlib streams of one chunk each) and its container in
// host memory (pinned memory makes the uploads asynchronous DMA).  Each
// mdz_serve_poll call advances every request one step:
//   * admission in submission order while the HBM budget allows (compressed
//     input + output + metadata; a request larger than the budget runs
//     alone);
//   * a lane-codec request is uploaded and decoded on one of the engine's
//     streams (rotating), its chunk statuses copied back behind the kernel,
//     and an event marks it done;
//   * a Deflate request is uploaded, and once the upload's event has passed
//     its streams are published to the live inflate queue: one persistent
//     inflate kernel that runs for the engine's life (reopened when its idle
//     workers time out); it is done when its statuses are no longer pending;
//   * done requests are reported once (tag, first failing status, device
//     output) and keep their device buffers until mdz_serve_release.
// Device buffers come from size-binned free lists and are freed only by
// mdz_serve_close, so no hipFree runs while the live kernel does.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "mdz_internal.h"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Size-binned device (or pinned host) buffers, reused and never freed
// before the engine closes.
struct BinPool {
    bool pinned = false;
    std::map<uint64_t, std::vector<void*>> free;
    std::vector<void*> all;
    static uint64_t bin(uint64_t n) {
        uint64_t b = 256;
        while (b < n) b <<= 1;
        return b;
    }
    int get(uint64_t n, void** p, uint64_t* b_out) {
        const uint64_t b = bin(n);
        auto& v = free[b];
        if (!v.empty()) {
            *p = v.back();
            v.pop_back();
        } else {
            const hipError_t e = pinned ? hipHostMalloc(p, b, hipHostMallocDefault) : hipMalloc(p, b);
            if (e != hipSuccess) return mdz::set_error(MDZ_ERR_NOMEM, "serve: allocation of %llu bytes",
                                                       (unsigned long long)b);
            all.push_back(*p);
        }
        *b_out = b;
        return MDZ_OK;
    }
    void put(void* p, uint64_t b) { free[b].push_back(p); }
    void release_all() {
        for (void* p : all) (void)(pinned ? hipHostFree(p) : hipFree(p));
        all.clear();
        free.clear();
    }
};

enum State { kWaiting, kUploading, kLaunched, kPublished, kDone };

struct Req {
    mdz_serve_request_t r;
    uint64_t hbm = 0;
    State st = kWaiting;
    hipEvent_t ev = nullptr;
    std::vector<std::pair<void*, uint64_t>> dbufs;      // device buffers and their bins
    void* d_out = nullptr;
    void* h_status = nullptr;                           // pinned copy of the chunk statuses
    uint64_t h_status_bin = 0;
    std::vector<mdz_inflate_in_t> desc;                 // Deflate: one descriptor per stream
    mdz_inflate_live_t* q = nullptr;
    uint32_t first = 0;
    int32_t result = MDZ_ST_OK;
    double t_active = 0;                                // admitted, or published
};

}  // namespace

struct mdz_server {
    uint64_t budget = 0, in_use = 0;
    std::vector<hipStream_t> streams;
    uint32_t next = 0;
    BinPool dev, host;
    std::deque<Req*> waiting;
    std::vector<Req*> active;
    std::unordered_map<uint64_t, Req*> done;           // reported, not released
    std::unordered_set<uint64_t> tags;                  // every submitted, unreleased tag
    std::vector<hipEvent_t> spare_events;
    mdz_inflate_live_t* live = nullptr;
    uint32_t live_cap = 1u << 16, live_used = 0;
    uint32_t live_blocks = 0, idle_ms = 0;
    double stall_s = 0.05;
};

namespace {

int event_get(mdz_server* s, hipEvent_t* e) {
    if (!s->spare_events.empty()) {
        *e = s->spare_events.back();
        s->spare_events.pop_back();
        return MDZ_OK;
    }
    return mdz::record_hip(hipEventCreateWithFlags(e, hipEventDisableTiming), "serve event");
}

uint64_t hbm_bytes(const mdz_serve_request_t& r) {
    uint64_t m = 0;
    if (r.codec == MDZ_SERVE_DEFLATE) m = 32 * r.n_streams;
    else m = 8 * (r.n_chunks + 1) + 256 * r.n_chunks + (r.col_map ? 32 * r.n_chunks : 0) + 4 * r.n_chunks;
    return r.payload_bytes + r.out_bytes + m;
}

int upload(mdz_server* s, Req* q, const void* src, uint64_t n, hipStream_t st, void** d) {
    uint64_t b = 0;
    int rc = s->dev.get(n + 16, d, &b);
    if (rc) return rc;
    q->dbufs.push_back({*d, b});
    if (n && hipMemcpyAsync(*d, src, n, hipMemcpyHostToDevice, st) != hipSuccess)
        return mdz::set_error(MDZ_ERR_HIP, "serve: upload of %llu bytes", (unsigned long long)n);
    return MDZ_OK;
}

// Admits request q: uploads it and, for a lane codec, enqueues its decode.
int start(mdz_server* s, Req* q) {
    const mdz_serve_request_t& r = q->r;
    hipStream_t st = s->streams[s->next];
    s->next = (s->next + 1) % (uint32_t)s->streams.size();
    uint64_t ob = 0;
    int rc = s->dev.get(r.out_bytes ? r.out_bytes : 1, &q->d_out, &ob);
    if (rc) return rc;
    q->dbufs.push_back({q->d_out, ob});
    void* d_pay = nullptr;
    if ((rc = upload(s, q, r.payload, r.payload_bytes, st, &d_pay))) return rc;
    if ((rc = event_get(s, &q->ev))) return rc;
    q->t_active = now_s();
    if (r.codec == MDZ_SERVE_DEFLATE) {
        const uint64_t k = r.n_streams;
        q->desc.resize(k);
        std::vector<uint64_t> order(k);
        for (uint64_t i = 0; i < k; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint64_t a, uint64_t b) { return r.src_len[a] > r.src_len[b]; });   // longest first
        for (uint64_t j = 0; j < k; ++j) {
            const uint64_t i = order[j], d0 = i * r.chunk_bytes;
            mdz_inflate_in_t& d = q->desc[j];
            d.src = static_cast<const uint8_t*>(d_pay) + r.src_off[i];
            d.src_bytes = r.src_len[i];
            d.dst = static_cast<uint8_t*>(q->d_out) + d0;
            d.dst_bytes = d0 < r.out_bytes ? std::min<uint64_t>(r.chunk_bytes, r.out_bytes - d0) : 0;
        }
        if (hipEventRecord(q->ev, st) != hipSuccess) return mdz::set_error(MDZ_ERR_HIP, "serve: event record");
        q->st = kUploading;
        return MDZ_OK;
    }
    const uint64_t n = r.n_chunks;
    void *d_off = nullptr, *d_col = nullptr, *d_map = nullptr, *d_st = nullptr;
    if ((rc = upload(s, q, r.blk_off, 8 * (n + 1), st, &d_off)) ||
        (rc = upload(s, q, r.col_len, 8 * 32 * n, st, &d_col)) ||
        (r.col_map && (rc = upload(s, q, r.col_map, 32 * n, st, &d_map))))
        return rc;
    uint64_t sb = 0;
    if ((rc = s->dev.get(4 * n, &d_st, &sb))) return rc;
    q->dbufs.push_back({d_st, sb});
    if ((rc = s->host.get(4 * n, &q->h_status, &q->h_status_bin))) return rc;
    mdz_lanes_t L = {d_pay, static_cast<const uint64_t*>(d_off), static_cast<const uint64_t*>(d_col),
                     static_cast<const uint8_t*>(d_map), n, r.chunk_bytes, r.elem_bytes, r.lane_line,
                     r.payload_bytes};
    if (r.codec == MDZ_SERVE_RLEV1)
        rc = mdz::rlev1_decompress(&L, q->d_out, r.out_bytes, static_cast<int32_t*>(d_st), nullptr, st);
    else if (r.codec == MDZ_SERVE_RLEV2)
        rc = mdz::rlev2_decompress(&L, q->d_out, r.out_bytes, static_cast<int32_t*>(d_st), nullptr, st);
    else
        rc = mdz::lzss_decompress(&L, q->d_out, r.out_bytes, static_cast<int32_t*>(d_st), st);
    if (rc) return rc;
    if (n && hipMemcpyAsync(q->h_status, d_st, 4 * n, hipMemcpyDeviceToHost, st) != hipSuccess)
        return mdz::set_error(MDZ_ERR_HIP, "serve: status copy");
    if (hipEventRecord(q->ev, st) != hipSuccess) return mdz::set_error(MDZ_ERR_HIP, "serve: event record");
    q->st = kLaunched;
    return MDZ_OK;
}

int publish(mdz_server* s, Req* q);

// Ends the live queue; its requests with items it never took (its workers
// gave up idle) are published again to a new one.
int retire_live(mdz_server* s, bool republish) {
    mdz_inflate_live_t* old = s->live;
    s->live = nullptr;
    s->live_used = 0;
    std::vector<mdz_inflate_status_t> fin(std::max<uint32_t>(s->live_cap, 1));
    int rc = mdz_inflate_live_close(old, fin.data());
    if (rc) return rc;
    for (Req* q : s->active) {
        if (q->st != kPublished || q->q != old) continue;
        bool pending = false;
        int32_t res = MDZ_ST_OK;
        for (uint32_t j = 0; j < q->desc.size(); ++j) {
            const int32_t v = (int32_t)fin[q->first + j].status;
            if (v == MDZ_ST_LIVE_PENDING) pending = true;
            else if (v != MDZ_ST_OK && res == MDZ_ST_OK) res = v;
        }
        if (!pending) {
            q->result = res;
            q->st = kDone;
        } else if (republish) {
            if ((rc = publish(s, q))) return rc;
        }
    }
    return MDZ_OK;
}

int publish(mdz_server* s, Req* q) {
    const uint32_t k = (uint32_t)q->desc.size();
    if (k > s->live_cap) return mdz::set_error(MDZ_ERR_UNSUPPORTED, "serve: %u streams in one request", k);
    int rc;
    if (s->live && (mdz_inflate_live_expired(s->live) == 1 || s->live_cap - s->live_used < k))
        if ((rc = retire_live(s, true))) return rc;
    if (!s->live) {
        if ((rc = mdz_inflate_live_open(s->live_cap, MDZ_HDR_ZLIB, s->live_blocks, s->idle_ms, &s->live)))
            return rc;
        s->live_used = 0;
    }
    uint32_t first = 0;
    if ((rc = mdz_inflate_live_submit(s->live, q->desc.data(), k, &first))) return rc;
    s->live_used += k;
    q->q = s->live;
    q->first = first;
    q->st = kPublished;
    q->t_active = now_s();
    return MDZ_OK;
}

}  // namespace

extern "C" int mdz_serve_open(uint64_t hbm_budget, uint32_t n_streams, uint32_t live_blocks, uint32_t idle_ms,
                              mdz_server_t** out) {
    if (!out || n_streams == 0 || n_streams > 16)
        return mdz::set_error(MDZ_ERR_INVALID, "serve: %u streams (1..16)", n_streams);
    mdz_server* s = new mdz_server();
    s->budget = hbm_budget;
    s->live_blocks = live_blocks;
    s->idle_ms = idle_ms;
    s->host.pinned = true;
    for (uint32_t i = 0; i < n_streams; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return mdz::set_error(MDZ_ERR_HIP, "serve: stream create");
        }
        s->streams.push_back(st);
    }
    *out = s;
    return MDZ_OK;
}

extern "C" int mdz_serve_submit(mdz_server_t* s, const mdz_serve_request_t* r) {
    if (!s || !r) return mdz::set_error(MDZ_ERR_INVALID, "serve submit: null argument");
    const bool lanes = r->codec == MDZ_SERVE_RLEV1 || r->codec == MDZ_SERVE_RLEV2 || r->codec == MDZ_SERVE_LZSS;
    if (!lanes && r->codec != MDZ_SERVE_DEFLATE)
        return mdz::set_error(MDZ_ERR_INVALID, "serve submit: codec %d", r->codec);
    if ((r->payload_bytes && !r->payload) || r->chunk_bytes == 0 ||
        (lanes && (!r->blk_off || !r->col_len || (r->codec == MDZ_SERVE_LZSS && !r->col_map))) ||
        (!lanes && r->n_streams && (!r->src_off || !r->src_len)))
        return mdz::set_error(MDZ_ERR_INVALID, "serve submit: request %llu incomplete",
                              (unsigned long long)r->tag);
    if (s->tags.count(r->tag)) return mdz::set_error(MDZ_ERR_INVALID, "serve submit: tag in use");
    s->tags.insert(r->tag);
    Req* q = new Req();
    q->r = *r;
    q->hbm = hbm_bytes(*r);
    s->waiting.push_back(q);
    return MDZ_OK;
}

extern "C" int mdz_serve_poll(mdz_server_t* s, mdz_serve_done_t* done, uint32_t cap, uint32_t* n_done) {
    if (!s || !n_done || (!done && cap)) return mdz::set_error(MDZ_ERR_INVALID, "serve poll: null argument");
    *n_done = 0;
    int rc;
    // admission, in submission order
    while (!s->waiting.empty()) {
        Req* q = s->waiting.front();
        const bool fits = s->in_use + q->hbm <= s->budget;
        const bool alone = s->active.empty() && s->done.empty();
        if (!fits && !alone) break;
        s->waiting.pop_front();
        s->in_use += q->hbm;
        s->active.push_back(q);
        if ((rc = start(s, q))) return rc;
    }
    // progress
    const mdz_inflate_status_t* lst = s->live ? mdz_inflate_live_statuses(s->live) : nullptr;
    bool live_busy = false;
    for (Req* q : s->active) {
        if (q->st == kUploading) {
            const hipError_t e = hipEventQuery(q->ev);
            if (e == hipSuccess) {
                if ((rc = publish(s, q))) return rc;
                lst = mdz_inflate_live_statuses(s->live);
            } else if (e != hipErrorNotReady) {
                return mdz::record_hip(e, "serve: upload event");
            }
        } else if (q->st == kLaunched) {
            const hipError_t e = hipEventQuery(q->ev);
            if (e == hipSuccess) {
                const int32_t* h = static_cast<const int32_t*>(q->h_status);
                for (uint64_t c = 0; c < q->r.n_chunks; ++c)
                    if (h[c] != MDZ_ST_OK) { q->result = h[c]; break; }
                q->st = kDone;
            } else if (e != hipErrorNotReady) {
                return mdz::record_hip(e, "serve: decode event");
            }
        }
        if (q->st == kPublished && q->q == s->live && lst) {
            bool pending = false;
            int32_t res = MDZ_ST_OK;
            for (uint32_t j = 0; j < q->desc.size(); ++j) {
                const int32_t v = (int32_t)__atomic_load_n(&lst[q->first + j].status, __ATOMIC_ACQUIRE);
                if (v == MDZ_ST_LIVE_PENDING) { pending = true; break; }
                if (v != MDZ_ST_OK && res == MDZ_ST_OK) res = v;
            }
            if (!pending) {
                q->result = res;
                q->st = kDone;
            } else {
                live_busy = true;
            }
        }
    }
    if (s->live) {
        if (live_busy && mdz_inflate_live_expired(s->live) == 1) {
            if ((rc = retire_live(s, true))) return rc;        // items its idle workers never took
        } else if (!live_busy) {
            // work that waits while the live kernel idles may sit behind it on
            // a shared hardware queue: end the idle kernel
            const double t = now_s();
            for (Req* q : s->active)
                if ((q->st == kUploading || q->st == kLaunched) && t - q->t_active > s->stall_s) {
                    if ((rc = retire_live(s, false))) return rc;
                    break;
                }
        }
    }
    // report
    std::vector<Req*> still;
    for (Req* q : s->active) {
        if (q->st == kDone && *n_done < cap) {
            mdz_serve_done_t& d = done[(*n_done)++];
            d.tag = q->r.tag;
            d.status = q->result;
            d.reserved = 0;
            d.d_out = q->d_out;
            d.out_bytes = q->r.out_bytes;
            s->done[q->r.tag] = q;
        } else {
            still.push_back(q);
        }
    }
    s->active.swap(still);
    return MDZ_OK;
}

extern "C" int mdz_serve_release(mdz_server_t* s, uint64_t tag) {
    if (!s) return mdz::set_error(MDZ_ERR_INVALID, "serve release: null server");
    auto it = s->done.find(tag);
    if (it == s->done.end()) return mdz::set_error(MDZ_ERR_INVALID, "serve release: tag not done");
    Req* q = it->second;
    s->done.erase(it);
    s->tags.erase(tag);
    for (auto& b : q->dbufs) s->dev.put(b.first, b.second);
    if (q->h_status) s->host.put(q->h_status, q->h_status_bin);
    if (q->ev) s->spare_events.push_back(q->ev);
    s->in_use -= q->hbm;
    delete q;
    return MDZ_OK;
}

extern "C" int mdz_serve_pending(const mdz_server_t* s, uint64_t* waiting, uint64_t* active) {
    if (!s) return mdz::set_error(MDZ_ERR_INVALID, "serve pending: null server");
    if (waiting) *waiting = s->waiting.size();
    if (active) *active = s->active.size();
    return MDZ_OK;
}

extern "C" int mdz_serve_close(mdz_server_t* s) {
    if (!s) return MDZ_OK;
    int rc = MDZ_OK;
    if (s->live) rc = retire_live(s, false);
    for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
    auto drop = [&](Req* q) {
        if (q->ev) (void)hipEventDestroy(q->ev);
        delete q;
    };
    for (Req* q : s->waiting) delete q;
    for (Req* q : s->active) drop(q);
    for (auto& kv : s->done) drop(kv.second);
    for (hipEvent_t e : s->spare_events) (void)hipEventDestroy(e);
    s->dev.release_all();
    s->host.release_all();
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    delete s;
    return rc;
}
