// pdmd_serve.cpp -- continuous-batching request loop over one engine context.
//
// Trajectory requests (a PRIME20 system and a number of calendar entries to
// commit) queue on the host.  Each step first prepares the host half of the
// start-up (topology, H-bond restore, tables) of the requests at the head of
// the queue while the launch in flight still runs, then waits for it,
// retires the requests it finished (final state downloaded, device memory
// freed), admits queued requests that fit the device-memory budget (first
// fit in arrival order, so a large request does not block smaller ones
// behind it; a prepared request needs only its uploads and start-up kernel
// here), and launches the next quantum for every resident request in one
// persistent launch.  A request's trajectory does not depend on which
// others share its launches or where the quanta fall: each system has its
// own wavefront, calendar and random streams.
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pdmd.h"
#include "pdmd_host.h"
#include "pdmd_internal.h"

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { kQueued = 0, kResident = 1, kDone = 2, kFailed = 3 };

// queued requests whose host half of the start-up is kept ready at most
// (their host copies of the system's tables and records)
constexpr int kPrepAhead = 16;

struct PrepFree {
    void operator()(pdmd::PreparedSystem* p) const { pdmd::prepared_free(p); }
};
using PrepPtr = std::unique_ptr<pdmd::PreparedSystem, PrepFree>;

struct Req {
    int64_t id = 0;
    // owned copies of the description's arrays
    std::vector<int32_t> chain_len, species, restype, hp, bptnr;
    std::vector<double> sv;
    pdmd_system_desc desc{};
    PrepPtr prep;  // host half of the start-up, ready for admission
    int64_t ncoll = 0, done = 0, bytes = 0, coll0 = 0;
    int32_t sys = -1;
    int state = kQueued, status = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> sv_out;
    pdmd_stats_t stats{};
};

}  // namespace

struct pdmd_server {
    pdmd_ctx* ctx = nullptr;
    int64_t budget = 0, used = 0, quantum = 0;
    double slice_ms = 0;  // device wall-clock slice per launch (0: event quanta only)
    int32_t max_resident = 0;
    double t0 = 0;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::deque<Req*> queue;
    std::vector<Req*> resident;
    std::vector<int64_t> inc;  // increments of the launch in flight, per resident request
    bool inflight = false;
    int64_t committed = 0;     // calendar entries committed by finished launches
    int nprep = 0;             // queued requests holding a prepared start-up
};

namespace {

int serve_fail(const char* m) { return pdmd::report_error(m); }

// retire the finished launch: advance counters, finish requests at target
int retire(pdmd_server* s, int32_t* ndone) {
    int rc = pdmd_wait(s->ctx);
    if (rc) return rc;
    s->inflight = false;
    const double t = now_s() - s->t0;
    std::vector<Req*> keep;
    for (size_t k = 0; k < s->resident.size(); ++k) {
        Req* r = s->resident[k];
        pdmd_stats_t st;
        rc = pdmd_stats(s->ctx, r->sys, &st);
        if (rc) return rc;
        const int64_t now_done = st.coll - r->coll0;
        s->committed += now_done - r->done;
        r->done = now_done;
        if (st.status != 0 || r->done >= r->ncoll) {
            r->stats = st;
            r->status = st.status;
            r->sv_out.resize(r->sv.size());
            rc = pdmd_download(s->ctx, r->sys, r->sv_out.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            if (rc) return rc;
            pdmd_system_remove(s->ctx, r->sys);
            r->sys = -1;
            s->used -= r->bytes;
            r->state = st.status ? kFailed : kDone;
            r->t_done = t;
            if (ndone) ++*ndone;
        } else
            keep.push_back(r);
    }
    s->resident.swap(keep);
    s->inc.clear();
    return 0;
}

// host halves of the start-up for the head of the queue (no device call)
void prepare_ahead(pdmd_server* s) {
    for (Req* r : s->queue) {
        if (s->nprep >= kPrepAhead) break;
        if (r->prep) continue;
        r->prep.reset(pdmd::prepared_new(s->ctx, &r->desc));
        if (r->prep) ++s->nprep;
    }
}

int admit(pdmd_server* s) {
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Req* r = *it;
        if ((int32_t)s->resident.size() >= s->max_resident) break;
        if (s->used + r->bytes > s->budget) {
            ++it;
            continue;
        }
        int32_t sid = -1;
        int rc;
        if (r->prep) {
            rc = pdmd::prepared_add(s->ctx, r->prep.get(), &sid);
            r->prep.reset();
            --s->nprep;
        } else
            rc = pdmd_system_add(s->ctx, &r->desc, &sid);
        if (rc) {
            r->state = kFailed;
            r->status = rc;
            r->t_done = now_s() - s->t0;
            it = s->queue.erase(it);
            continue;
        }
        r->sys = sid;
        pdmd_stats_t st;
        if (pdmd_stats(s->ctx, sid, &st) == 0) r->coll0 = st.coll;
        r->state = kResident;
        r->t_admit = now_s() - s->t0;
        s->used += r->bytes;
        s->resident.push_back(r);
        it = s->queue.erase(it);
    }
    return 0;
}

int launch(pdmd_server* s) {
    if (s->resident.empty()) return 0;
    std::vector<int32_t> ids;
    std::vector<int64_t> nev;
    s->inc.clear();
    for (Req* r : s->resident) {
        const int64_t left = r->ncoll - r->done;
        const int64_t q = left < s->quantum ? left : s->quantum;
        ids.push_back(r->sys);
        nev.push_back(q);
        s->inc.push_back(q);
    }
    int rc = s->slice_ms > 0 ? pdmd_advance_budget_async(s->ctx, (int32_t)ids.size(), ids.data(), nev.data(), s->slice_ms)
                             : pdmd_advance_async(s->ctx, (int32_t)ids.size(), ids.data(), nev.data());
    if (rc) return rc;
    s->inflight = true;
    return 0;
}

}  // namespace

extern "C" {

int pdmd_server_create(pdmd_ctx* ctx, int64_t mem_budget_bytes, int64_t quantum, double slice_ms, int32_t max_resident,
                       pdmd_server** out) {
    if (!ctx || !out || mem_budget_bytes <= 0 || quantum <= 0 || slice_ms < 0)
        return serve_fail("pdmd_server_create: bad arguments");
    auto s = std::make_unique<pdmd_server>();
    s->ctx = ctx;
    s->budget = mem_budget_bytes;
    s->quantum = quantum;
    s->slice_ms = slice_ms;
    s->max_resident = max_resident > 0 ? (max_resident < 4096 ? max_resident : 4096) : 4096;
    s->t0 = now_s();
    // each step's launch is a replay of the captured event-loop launch
    // while the resident set keeps its shape
    if (int rc = pdmd_set_graph_launch(ctx, 1)) return rc;
    *out = s.release();
    return 0;
}

int pdmd_server_destroy(pdmd_server* s) {
    if (!s) return -1;
    if (s->inflight) pdmd_wait(s->ctx);
    for (Req* r : s->resident)
        if (r->sys >= 0) pdmd_system_remove(s->ctx, r->sys);
    pdmd_set_graph_launch(s->ctx, 0);
    delete s;
    return 0;
}

int pdmd_server_submit(pdmd_server* s, const pdmd_system_desc* d, int64_t ncoll, int64_t* req_id) {
    if (!s || !d || !req_id || ncoll <= 0 || d->nchains <= 0 || !d->chain_len || !d->restype || !d->hp || !d->sv)
        return serve_fail("pdmd_server_submit: bad arguments");
    auto r = std::make_unique<Req>();
    const int nc = d->nchains;
    r->chain_len.assign(d->chain_len, d->chain_len + nc);
    int nres = 0;
    for (int c = 0; c < nc; ++c) nres += d->chain_len[c];
    const int N = 4 * nres;
    if (d->species) r->species.assign(d->species, d->species + nc);
    r->restype.assign(d->restype, d->restype + nres);
    r->hp.assign(d->hp, d->hp + nres);
    r->sv.assign(d->sv, d->sv + 6 * (size_t)N);
    if (d->bptnr) r->bptnr.assign(d->bptnr, d->bptnr + N);
    r->desc = *d;
    r->desc.chain_len = r->chain_len.data();
    r->desc.species = d->species ? r->species.data() : nullptr;
    r->desc.restype = r->restype.data();
    r->desc.hp = r->hp.data();
    r->desc.sv = r->sv.data();
    r->desc.bptnr = d->bptnr ? r->bptnr.data() : nullptr;
    // the host half of the start-up sizes the request; it is kept for
    // admission while few queued requests hold one
    PrepPtr prep(pdmd::prepared_new(s->ctx, &r->desc));
    if (!prep) return -1;
    r->bytes = pdmd::prepared_bytes(prep.get());
    if (r->bytes > s->budget) return serve_fail("pdmd_server_submit: request exceeds the memory budget");
    if (s->nprep < kPrepAhead) {
        r->prep = std::move(prep);
        ++s->nprep;
    }
    r->ncoll = ncoll;
    r->id = s->next_id++;
    r->t_submit = now_s() - s->t0;
    *req_id = r->id;
    s->queue.push_back(r.get());
    s->reqs[r->id] = std::move(r);
    return 0;
}

int pdmd_server_step(pdmd_server* s, int32_t* ncompleted) {
    if (!s) return -1;
    if (ncompleted) *ncompleted = 0;
    prepare_ahead(s);  // overlaps the launch in flight
    if (s->inflight) {
        int rc = retire(s, ncompleted);
        if (rc) return rc;
    }
    int rc = admit(s);
    if (rc) return rc;
    return launch(s);
}

int pdmd_server_pending(pdmd_server* s, int32_t* queued, int32_t* resident, int64_t* committed) {
    if (!s) return -1;
    if (queued) *queued = (int32_t)s->queue.size();
    if (resident) *resident = (int32_t)s->resident.size();
    if (committed) *committed = s->committed;
    return 0;
}

int pdmd_server_result(pdmd_server* s, int64_t req_id, pdmd_request_info_t* info, double* sv) {
    if (!s) return -1;
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return serve_fail("pdmd_server_result: unknown request");
    Req* r = it->second.get();
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->state = r->state;
        info->status = r->status;
        info->ncoll = r->ncoll;
        info->done = r->done;
        info->nbeads = (int32_t)(r->sv.size() / 6);
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
        info->stats = r->stats;
    }
    if (sv && (r->state == kDone || r->state == kFailed) && !r->sv_out.empty())
        std::memcpy(sv, r->sv_out.data(), sizeof(double) * r->sv_out.size());
    return 0;
}

int pdmd_server_release(pdmd_server* s, int64_t req_id) {
    if (!s) return -1;
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return -1;
    if (it->second->state == kQueued || it->second->state == kResident) return -1;
    s->reqs.erase(it);
    return 0;
}

}  // extern "C"
