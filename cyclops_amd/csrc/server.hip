// server.hip -- the fit-serving loop behind ccd_server_* (include/cyclops_amd.h).
//
// Continuous admission: `slots` worker threads each own a HIP stream (one engine
// per request). The HBM budget is reserved once as a device region; a request is
// admitted, first-fit in arrival order, as soon as a slot is free, a block of the
// region holds its footprint, and (resident mode) its persistent launch fits what
// the running requests leave of the device's co-resident workgroups -- so
// concurrent persistent kernels can always all be resident and never wait on each
// other. Admitted fits run concurrently on
// the device. A request's kernel shape depends only on its own data, so its result
// is the same alone and in any mix.
//
// The reference has no serving loop; its concurrency is thread-level clones of
// the engine (AbstractCrossValidationDriver.cpp:44-129, Thread.h:75-130), which
// this generalises to independent requests.
#include "engine.h"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

enum ReqState { QUEUED = 0, RUNNING = 1, DONE = 2 };

struct Request {
    int64_t id = 0;
    // owned copies of the caller's description and arguments
    std::vector<int32_t> format, rows, pid;
    std::vector<int64_t> colPtr, densePtr;
    std::vector<double> values, dense, y, time, weights, beta;
    std::vector<uint8_t> exclude, fixed;
    ccd_data_desc desc{};
    ccd_fit_args args{};
    int32_t J = 0;
    // admission
    double footprint = 0.0;
    double capFraction = 0.0;
    size_t blockOff = 0, blockSize = 0;
    size_t heldOff = 0, heldSize = 0;   // the region block this request frees when done (a packed group's first member holds the group's block)
    int model = 0, prec = 0;
    bool packable = false;   // resident fit of one workgroup: may share a launch with others of its model and precision
    // outcome
    ReqState state = QUEUED;
    int rc = CCD_OK;
    std::string error;
    ccd_fit_result result{};
    ccd_request_stats stats{};
};

template <class T>
void copyIn(std::vector<T>& dst, const T* src, size_t n) {
    if (src) dst.assign(src, src + n);
    else dst.clear();
}

// First-fit allocator of blocks inside one device region reserved at server start.
struct Region {
    char* base = nullptr;
    size_t cap = 0;
    std::map<size_t, size_t> holes;   // offset -> size, disjoint, coalesced

    bool alloc(size_t n, size_t& off) {
        for (auto it = holes.begin(); it != holes.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            holes.erase(it);
            if (rest) holes[off + n] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        auto it = holes.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != holes.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            holes.erase(nx);
        }
        if (it != holes.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                holes.erase(it);
            }
        }
    }
};

// Device bytes a request's engine may carve from its block: the admission estimate
// with a margin (an engine that needs more allocates the rest with hipMalloc).
size_t blockBytes(double footprint) { return (((size_t)(footprint * 1.25) + (size_t(1) << 20)) + 255) & ~size_t(255); }

}  // namespace

struct ccd_server {
    ccd_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex m;
    std::condition_variable work;   // a request may be admissible / stop
    std::condition_variable done;   // a request finished
    std::deque<std::shared_ptr<Request>> queue;
    std::map<int64_t, std::shared_ptr<Request>> all;
    double usedCap = 0.0;
    Region region;   // all requests' device memory
    int64_t nextId = 1;
    bool stopping = false;
    std::vector<std::thread> workers;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // A request's persistent launch takes at most a quarter of the device's CUs in
    // workgroups, so that no single large fit (a Cox fit asks for one workgroup per
    // 1,024 rows) holds the whole co-resident capacity while others queue behind it.
    int maxWorkgroups() const {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg.device) != hipSuccess) return 0;
        return std::max(1, cus / 4);
    }

    // First queued request that fits the remaining budget (caller holds m). Later
    // requests may overtake one that does not fit yet. Once the oldest has waited 2 s
    // it holds a reservation: a later request starts only if the co-resident capacity
    // and an HBM block for the oldest would still be left, so overtakers never take
    // what the oldest is waiting for (a large request cannot starve, and small ones
    // keep flowing past it).
    std::shared_ptr<Request> pick() {
        if (queue.empty()) return nullptr;
        const auto& head = queue.front();
        const bool headWaited = now() - head->stats.submit_time > 2.0;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            const auto& r = *it;
            const bool reserve = headWaited && it != queue.begin();
            if (usedCap + r->capFraction + (reserve ? head->capFraction : 0.0) > 1.0 + 1e-9) continue;
            size_t off = 0;
            if (!region.alloc(r->blockSize, off)) continue;
            if (reserve) {
                size_t hoff = 0;
                if (!region.alloc(head->blockSize, hoff)) {
                    region.release(off, r->blockSize);
                    continue;
                }
                region.release(hoff, head->blockSize);
            }
            r->blockOff = off;
            auto q = r;
            queue.erase(it);
            return q;
        }
        return nullptr;
    }

    // The admitted request r plus up to cfg.pack - 1 later queued requests of its model
    // and precision whose resident fit is one workgroup (caller holds m). The group
    // takes one region block, the sum of its members' blocks: its engines carve their
    // memory from it one after another, so a packed launch calls no hipMalloc either.
    std::vector<std::shared_ptr<Request>> pickGroup() {
        std::vector<std::shared_ptr<Request>> g;
        auto r = pick();
        if (!r) return g;
        g.push_back(r);
        if (cfg.pack <= 1 || cfg.mode != CCD_MODE_RESIDENT || !r->packable) return g;
        size_t total = r->blockSize;
        double cap = usedCap + r->capFraction;
        // a waiting oldest request keeps its reserved share (see pick) unless it joins
        // the group; it is the first one the walk below meets
        const std::shared_ptr<Request> front = queue.empty() ? nullptr : queue.front();
        bool reserve = front && now() - front->stats.submit_time > 2.0;
        std::vector<std::deque<std::shared_ptr<Request>>::iterator> take;
        for (auto it = queue.begin(); it != queue.end() && (int)(g.size() + take.size()) < cfg.pack; ++it) {
            const auto& q = *it;
            if (!q->packable || q->model != r->model || q->prec != r->prec) continue;
            const double held = (reserve && q != front) ? front->capFraction : 0.0;
            if (cap + q->capFraction + held > 1.0 + 1e-9) continue;
            cap += q->capFraction;
            total += q->blockSize;
            take.push_back(it);
            if (q == front) reserve = false;
        }
        if (take.empty()) return g;
        // one block for the group, in place of r's own (and, for a waiting oldest
        // request left out of it, only if its block would still fit beside it)
        size_t off = 0;
        region.release(r->blockOff, r->blockSize);
        bool ok = region.alloc(total, off);
        if (ok && reserve) {
            size_t hoff = 0;
            if (region.alloc(front->blockSize, hoff)) region.release(hoff, front->blockSize);
            else {
                region.release(off, total);
                ok = false;
            }
        }
        if (!ok) {
            region.alloc(r->blockSize, off);   // the block just released fits again
            r->blockOff = off;
            return g;
        }
        r->blockOff = off;
        for (auto it : take) g.push_back(*it);
        for (auto k = take.rbegin(); k != take.rend(); ++k) queue.erase(*k);
        r->heldOff = off;
        r->heldSize = total;
        for (size_t i = 1; i < g.size(); ++i) g[i]->heldSize = 0;
        return g;
    }

    // A request's engine carves its device memory from the request's block of the
    // server's region, so starting or finishing a request calls no hipMalloc / hipFree
    // (which synchronise the whole device and would wait for the other slots' kernels).
    void run(Request& r, ccd::Arena& arena, bool useArena) {
        CCD_HIP(hipSetDevice(cfg.device));
        if (useArena) {
            arena.base = region.base + r.blockOff;
            arena.cap = r.blockSize;
            arena.used = 0;
        }
        ccd_handle h = nullptr;
        int rc = ccd_create(&r.desc, cfg.device, &h);
        if (rc == CCD_OK) {
            // the engine holds its own copy in HBM: release the host copy now
            std::vector<int32_t>().swap(r.rows);
            std::vector<double>().swap(r.values);
            std::vector<double>().swap(r.dense);
            rc = ccd_fit(h, &r.args, cfg.mode, r.beta.data(), &r.result);
        }
        if (rc != CCD_OK) r.error = ccd_last_error();
        if (h) ccd_destroy(h);
        arena.base = nullptr;
        arena.cap = arena.used = 0;
        r.rc = rc;
    }

    void worker() {
        ccd::Arena arena;
        bool ready = true;
        try {
            CCD_HIP(hipSetDevice(cfg.device));
            ccd::setThreadMaxWorkgroups(maxWorkgroups());
            CCD_HIP(hipStreamCreateWithFlags(&arena.stream, hipStreamNonBlocking));
            CCD_HIP(hipHostMalloc(reinterpret_cast<void**>(&arena.host), 64 * sizeof(double), hipHostMallocDefault));
            ccd::setThreadArena(&arena);
        } catch (const std::exception&) {
            ready = false;   // no arena: this slot allocates per request
            ccd::setThreadArena(nullptr);
        }
        loop(arena, ready);
        ccd::setThreadArena(nullptr);
        if (arena.host) (void)hipHostFree(arena.host);
        if (arena.stream) (void)hipStreamDestroy(arena.stream);
    }

    // Requests packed into one launch: one engine each, carved one after another from
    // the group's block, then ccd_fit_multi over all of them (each fit keeps the team
    // shape of its own resident launch, so its result equals that launch's).
    void runGroup(std::vector<std::shared_ptr<Request>>& g, ccd::Arena& arena, bool useArena) {
        CCD_HIP(hipSetDevice(cfg.device));
        if (useArena) {
            arena.base = region.base + g[0]->blockOff;
            arena.cap = g[0]->heldSize;
            arena.used = 0;
        }
        const size_t n = g.size();
        std::vector<ccd_handle> h(n, nullptr);
        std::vector<ccd_fit_args> args(n);
        std::vector<double*> betas(n);
        std::vector<ccd_fit_result> res(n);
        int rc = CCD_OK;
        for (size_t i = 0; i < n && rc == CCD_OK; ++i) {
            Request& r = *g[i];
            rc = ccd_create(&r.desc, cfg.device, &h[i]);
            if (rc == CCD_OK) {
                std::vector<int32_t>().swap(r.rows);
                std::vector<double>().swap(r.values);
                std::vector<double>().swap(r.dense);
            }
            args[i] = r.args;
            betas[i] = r.beta.data();
        }
        if (rc == CCD_OK) rc = ccd_fit_multi(h.data(), (int32_t)n, args.data(), betas.data(), res.data());
        const std::string err = rc == CCD_OK ? std::string() : std::string(ccd_last_error());
        for (size_t i = 0; i < n; ++i) {
            if (h[i]) ccd_destroy(h[i]);
            g[i]->rc = rc;
            g[i]->error = err;
            if (rc == CCD_OK) g[i]->result = res[i];
        }
        arena.base = nullptr;
        arena.cap = arena.used = 0;
    }

    void loop(ccd::Arena& arena, bool useArena) {
        for (;;) {
            std::vector<std::shared_ptr<Request>> g;
            {
                std::unique_lock<std::mutex> lk(m);
                work.wait(lk, [&] { return !(g = pickGroup()).empty() || (stopping && queue.empty()); });
                if (g.empty()) return;
                const double t = now();
                for (auto& r : g) {
                    usedCap += r->capFraction;
                    r->state = RUNNING;
                    r->stats.start_time = t;
                }
            }
            try {
                if (g.size() == 1) run(*g[0], arena, useArena);
                else runGroup(g, arena, useArena);
            } catch (const std::exception& ex) {
                for (auto& r : g) {
                    r->rc = CCD_ERR_HIP;
                    r->error = ex.what();
                }
            }
            {
                std::lock_guard<std::mutex> lk(m);
                if (g.size() == 1) region.release(g[0]->blockOff, g[0]->blockSize);
                else region.release(g[0]->heldOff, g[0]->heldSize);
                const double t = now();
                for (auto& r : g) {
                    usedCap -= r->capFraction;
                    r->stats.end_time = t;
                    r->state = DONE;
                    // keep only what poll returns
                    std::vector<int32_t>().swap(r->format);
                    std::vector<int64_t>().swap(r->colPtr);
                    std::vector<int64_t>().swap(r->densePtr);
                    std::vector<double>().swap(r->y);
                    std::vector<double>().swap(r->time);
                    std::vector<int32_t>().swap(r->pid);
                    std::vector<double>().swap(r->weights);
                    std::vector<uint8_t>().swap(r->exclude);
                    std::vector<uint8_t>().swap(r->fixed);
                    r->args.exclude = r->args.fixed = nullptr;
                    r->args.weights = nullptr;
                }
            }
            done.notify_all();
            work.notify_all();
        }
    }
};

namespace {
// Errors go to the same per-thread message as every other entry point (ccd_last_error).
template <class F>
int sguard(F&& f) {
    try {
        return f();
    } catch (const ccd::HipError& ex) {
        ccd::setLastError(ex.what());
        return CCD_ERR_HIP;
    } catch (const std::invalid_argument& ex) {
        ccd::setLastError(ex.what());
        return CCD_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        ccd::setLastError("out of host memory");
        return CCD_ERR_NOMEM;
    } catch (const std::exception& ex) {
        ccd::setLastError(ex.what());
        return CCD_ERR_STATE;
    }
}
}  // namespace

extern "C" {

int ccd_server_start(const ccd_server_config* cfg, ccd_server_handle* out) {
    return sguard([&] {
        if (!cfg || !out) throw std::invalid_argument("null argument");
        int n = 0;
        CCD_HIP(hipGetDeviceCount(&n));
        if (cfg->device < 0 || cfg->device >= n) throw std::invalid_argument("no such device");
        if (cfg->slots < 1 || cfg->slots > 64) throw std::invalid_argument("slots must be in [1, 64]");
        if (!(cfg->hbm_budget_bytes > 0)) throw std::invalid_argument("hbm_budget_bytes must be > 0");
        if (cfg->mode != CCD_MODE_PLAIN && cfg->mode != CCD_MODE_RESIDENT) throw std::invalid_argument("bad mode");
        if (cfg->pack < 0 || cfg->pack > 64) throw std::invalid_argument("pack must be in [0, 64]");
        auto* s = new ccd_server();
        s->cfg = *cfg;
        s->t0 = std::chrono::steady_clock::now();
        // the admission budget, reserved once (at most 90% of what the device has free)
        CCD_HIP(hipSetDevice(cfg->device));
        size_t freeB = 0, totalB = 0;
        CCD_HIP(hipMemGetInfo(&freeB, &totalB));
        size_t want = std::min((size_t)cfg->hbm_budget_bytes, (size_t)(0.9 * (double)freeB));
        want &= ~size_t(255);
        if (want < (size_t(1) << 20)) {
            delete s;
            throw std::invalid_argument("hbm_budget_bytes leaves no room for a request");
        }
        if (hipMalloc(reinterpret_cast<void**>(&s->region.base), want) != hipSuccess) {
            delete s;
            throw ccd::HipError("could not reserve the serving region of " + std::to_string(want) + " bytes");
        }
        s->region.cap = want;
        s->region.holes[0] = want;
        s->cfg.hbm_budget_bytes = (double)want;
        for (int i = 0; i < cfg->slots; ++i) s->workers.emplace_back([s] { s->worker(); });
        *out = s;
        return CCD_OK;
    });
}

int ccd_server_submit(ccd_server_handle s, const ccd_data_desc* d, const ccd_fit_args* a, const double* beta0,
                      int64_t* request_id) {
    return sguard([&] {
        if (!s || !d || !a || !request_id) throw std::invalid_argument("null argument");
        if (d->K <= 0 || d->J <= 0 || !d->format || !d->col_ptr || !d->y)
            throw std::invalid_argument("incomplete data description");
        auto r = std::make_shared<Request>();
        const int64_t K = d->K;
        const int32_t J = d->J;
        r->J = J;
        copyIn(r->format, d->format, J);
        copyIn(r->colPtr, d->col_ptr, J + 1);
        const int64_t nnz = d->col_ptr[J];
        if (nnz < 0) throw std::invalid_argument("col_ptr[J] < 0");
        copyIn(r->rows, d->rows, nnz);
        copyIn(r->values, d->values, nnz);
        if (d->dense_ptr) {
            copyIn(r->densePtr, d->dense_ptr, J + 1);
            copyIn(r->dense, d->dense_values, d->dense_ptr[J]);
        }
        copyIn(r->y, d->y, K);
        copyIn(r->time, d->time, K);
        copyIn(r->pid, d->pid, K);
        r->desc = *d;
        r->desc.format = r->format.data();
        r->desc.col_ptr = r->colPtr.data();
        r->desc.rows = r->rows.empty() ? nullptr : r->rows.data();
        r->desc.values = r->values.empty() ? nullptr : r->values.data();
        r->desc.dense_ptr = r->densePtr.empty() ? nullptr : r->densePtr.data();
        r->desc.dense_values = r->dense.empty() ? nullptr : r->dense.data();
        r->desc.y = r->y.data();
        r->desc.time = r->time.empty() ? nullptr : r->time.data();
        r->desc.pid = r->pid.empty() ? nullptr : r->pid.data();
        r->args = *a;
        copyIn(r->exclude, a->exclude, J);
        copyIn(r->fixed, a->fixed, J);
        copyIn(r->weights, a->weights, K);
        r->args.exclude = r->exclude.empty() ? nullptr : r->exclude.data();
        r->args.fixed = r->fixed.empty() ? nullptr : r->fixed.data();
        r->args.weights = r->weights.empty() ? nullptr : r->weights.data();
        r->beta.assign(J, 0.0);
        if (beta0) std::copy(beta0, beta0 + J, r->beta.begin());
        // admission footprint: the engine plus the resident workspace
        double bytes = 0.0;
        if (ccd_estimate_footprint(&r->desc, &bytes) != CCD_OK) throw std::invalid_argument(ccd_last_error());
        r->model = d->model_type;
        r->prec = ccd::precisionOf(d->precision);
        if (s->cfg.mode == CCD_MODE_RESIDENT) {
            const ccd::ResidentNeeds n = ccd::residentNeeds(r->desc, s->cfg.device, s->maxWorkgroups());
            bytes += n.workspaceBytes;
            r->capFraction = n.capacityFraction;
            r->packable = n.workgroups == 1;
            // starting values on more than 32 SPARSE/INDICATOR columns: the fit forms xb
            // from a row-major copy of those columns (row pointers, column ids, values)
            int live = 0;
            int64_t idxNnz = 0;
            for (int j = 0; j < J; ++j) {
                const int f = r->format[j];
                if (f != CCD_FORMAT_SPARSE && f != CCD_FORMAT_INDICATOR) continue;
                idxNnz += r->colPtr[j + 1] - r->colPtr[j];
                if ((j == 0 && d->has_offset) || r->beta[j] != 0.0) ++live;
            }
            if (live > 32)
                bytes += (double)(K + 1) * 8.0 + (double)idxNnz * (4.0 + (ccd::precisionOf(d->precision) == CCD_FP32 ? 4.0 : 8.0));
        }
        r->footprint = bytes;
        r->blockSize = blockBytes(bytes);
        if ((double)r->blockSize > s->cfg.hbm_budget_bytes)
            throw std::invalid_argument("request needs " + std::to_string(r->blockSize) +
                                        " bytes of HBM, over the budget of " + std::to_string(s->cfg.hbm_budget_bytes));
        {
            std::lock_guard<std::mutex> lk(s->m);
            if (s->stopping) throw std::runtime_error("server is stopping");
            r->id = s->nextId++;
            r->stats.submit_time = s->now();
            r->stats.footprint_bytes = bytes;
            s->queue.push_back(r);
            s->all[r->id] = r;
            *request_id = r->id;
        }
        s->work.notify_all();
        return CCD_OK;
    });
}

int ccd_server_poll(ccd_server_handle s, int64_t id, int wait, double* beta_out, ccd_fit_result* result,
                    ccd_request_stats* stats) {
    return sguard([&] {
        if (!s) throw std::invalid_argument("null server handle");
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->all.find(id);
        if (it == s->all.end()) throw std::invalid_argument("unknown request id");
        auto r = it->second;
        if (r->state != DONE) {
            if (!wait) {
                ccd::setLastError("request still running");
                return CCD_ERR_STATE;
            }
            s->done.wait(lk, [&] { return r->state == DONE; });
        }
        if (stats) *stats = r->stats;
        if (result) *result = r->result;
        if (beta_out) std::copy(r->beta.begin(), r->beta.end(), beta_out);
        if (r->rc != CCD_OK) ccd::setLastError(r->error);
        return r->rc;
    });
}

int ccd_server_release(ccd_server_handle s, int64_t id) {
    return sguard([&] {
        if (!s) throw std::invalid_argument("null server handle");
        std::lock_guard<std::mutex> lk(s->m);
        auto it = s->all.find(id);
        if (it == s->all.end()) throw std::invalid_argument("unknown request id");
        if (it->second->state != DONE) {
            ccd::setLastError("request still running");
            return CCD_ERR_STATE;
        }
        s->all.erase(it);
        return CCD_OK;
    });
}

int ccd_server_stop(ccd_server_handle s) {
    return sguard([&] {
        if (!s) throw std::invalid_argument("null server handle");
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->stopping = true;   // queued requests still run; new ones are refused
        }
        s->work.notify_all();
        for (auto& t : s->workers) t.join();
        if (s->region.base) (void)hipFree(s->region.base);
        delete s;
        return CCD_OK;
    });
}

}  // extern "C"
