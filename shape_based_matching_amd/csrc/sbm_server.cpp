// Continuous-batching request loop over one detector's resident template bank.
//
// Requests (images of any admissible size) queue up as they arrive.  A dispatcher thread owns a fixed
// set of slots -- pipelines, each with its own stream, level buffers and captured graphs -- and admits
// the oldest queued request whenever a slot is free and the bank plus every slot's device memory, with
// this slot's for that image size, fits the budget (an idle slot's buffers are released first when that
// makes room).  Queued requests of
// the same size, channels, mask presence, threshold and selection join it in the same launch, up to
// max_per_launch and as far as the budget allows: every kernel of the launch covers all of them.
// Launches on different slots run concurrently, so work from several requests overlaps on the device;
// each request returns exactly what Detector::match returns for it alone.  A finished launch wakes the
// dispatcher through a host callback on its stream.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "sbm_server.h"

namespace sbm {

Server::Server(Detector& det, int slots, size_t budget, int max_per_launch)
    : det_(det), budget_(budget), max_per_launch_(std::max(1, std::min(max_per_launch, kMaxBatch))) {
    det_.ensure_device();
    det_.ensure_bank();
    bank_version_ = det_.bank().version;
    device_ = det_.device();
    // the budget covers everything the server's detector holds on the device: the resident bank and the
    // slots; the detector's own pipelines stay idle while it serves, so their buffers are given back
    det_.release_match_buffers();
    fixed_ = det_.device_bytes();
    const int n = std::max(1, std::min(slots, 16));
    for (int i = 0; i < n; ++i) {
        slots_.emplace_back(new Pipeline(device_, det_.num_cu()));
        slots_.back()->tuning = det_.tuning;
    }
    running_.resize(n);
    thread_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    if (thread_.joinable()) thread_.join();
    // a finished launch's host callback may still be queued behind its completion event: let every slot's
    // stream drain while this server (its mutex and condition variable) still exists
    for (auto& p : slots_) (void)hipStreamSynchronize(p->stream());
}

long long Server::submit(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                         float threshold, const std::vector<std::string>& class_ids) {
    det_.check_channels(channels);
    if (stride < cols * channels) throw Error(kEInval, "stride smaller than a row");
    det_.check_sizes(rows, cols);  // a malformed request is refused here; queued work is unaffected
    if (det_.bank().version != bank_version_) throw Error(kEInval, "the detector's bank changed under the server");
    auto r = std::make_shared<Request>();
    r->rows = rows;
    r->cols = cols;
    r->channels = channels;
    r->threshold = threshold;
    r->img.resize((size_t)rows * cols * channels);
    for (int y = 0; y < rows; ++y)
        memcpy(&r->img[(size_t)y * cols * channels], img + (size_t)y * stride, (size_t)cols * channels);
    if (mask) r->mask.assign(mask, mask + (size_t)rows * cols);
    det_.selection(class_ids, true, r->sel);
    r->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) throw Error(kEInval, "server is shutting down");
        r->id = next_id_++;
        queue_.push_back(r);
        ++submitted_;
    }
    cv_work_.notify_all();
    return r->id;
}

int Server::wait(long long id, int timeout_ms, std::vector<MatchRec>* out, double* latency_ms, std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto ready = [&] { return done_.count(id) != 0; };
    if (timeout_ms < 0) cv_done_.wait(lk, ready);
    else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return kETimeout;
    auto r = done_[id];
    done_.erase(id);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->code != kOk) {
        if (err) *err = r->message;
        return r->code;
    }
    if (out) *out = std::move(r->result);
    return kOk;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServerStats s = stats_;
    s.submitted = submitted_;
    s.budget = (long long)budget_;
    return s;
}

size_t Server::slots_bytes() const {
    size_t s = 0;
    for (const auto& p : slots_) s += p->device_bytes();
    return s;
}

// Picks a free slot for r and makes room for it within the budget.  Returns the slot; -2 when every
// slot is busy; -1 when r must wait for memory (never_fits: it cannot fit even with nothing running).
int Server::admit_slot(const Request& r, int n, bool any_running, bool* never_fits, bool may_release) {
    *never_fits = false;
    const Bank& bank = det_.bank();
    const bool m = !r.mask.empty();
    auto after = [&](int s) { return slots_[s]->bytes_after(bank, r.rows, r.cols, r.channels, m, n); };
    int best = -1;
    size_t best_grow = 0;
    for (size_t s = 0; s < slots_.size(); ++s) {
        if (!running_[s].empty()) continue;
        const size_t grow = after((int)s) - slots_[s]->device_bytes();
        if (best < 0 || grow < best_grow) {
            best = (int)s;
            best_grow = grow;
        }
    }
    if (best < 0) return -2;
    auto fits = [&] { return fixed_ + slots_bytes() - slots_[best]->device_bytes() + after(best) <= budget_; };
    if (fits()) return best;
    if (!may_release) return -1;
    // give back the memory of other idle slots, largest first, then the chosen slot's own
    for (;;) {
        int victim = -1;
        size_t vb = 0;
        for (size_t s = 0; s < slots_.size(); ++s)
            if (running_[s].empty() && (int)s != best && slots_[s]->device_bytes() > vb) {
                victim = (int)s;
                vb = slots_[s]->device_bytes();
            }
        if (victim < 0) break;
        slots_[victim]->release();
        ++stats_.releases;
        if (fits()) return best;
    }
    if (slots_[best]->device_bytes() > 0) {
        slots_[best]->release();
        ++stats_.releases;
        if (fits()) return best;
    }
    if (!any_running) *never_fits = true;  // nothing else holds memory and it still does not fit
    return -1;
}

void Server::finish(int s) {
    auto group = std::move(running_[s]);
    running_[s].clear();
    int code = kOk;
    std::string msg;
    std::vector<std::vector<MatchRec>> res;
    try {
        res = slots_[s]->graph_finish_batch(det_.bank());
    } catch (const Error& e) {
        code = e.code;
        msg = e.what();
    }
    const auto now = std::chrono::steady_clock::now();
    for (size_t k = 0; k < group.size(); ++k) {
        auto& r = group[k];
        r->code = code;
        r->message = msg;
        if (code == kOk) r->result = std::move(res[k]);
        r->t_done = now;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& r : group) {
            done_[r->id] = r;
            if (r->code == kOk) ++stats_.completed;
            else ++stats_.failed;
        }
    }
    cv_done_.notify_all();
}

void Server::on_done(void* server) {
    Server* self = static_cast<Server*>(server);
    {
        std::lock_guard<std::mutex> lk(self->mu_);
        self->wake_ = true;
    }
    self->cv_work_.notify_all();
}

void Server::fail(const std::shared_ptr<Request>& r, int code, const std::string& msg) {
    r->code = code;
    r->message = msg;
    r->t_done = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        done_[r->id] = r;
        ++stats_.failed;
    }
    cv_done_.notify_all();
}

void Server::loop() {
    (void)hipSetDevice(device_);
    auto same_launch = [](const Request& a, const Request& b) {
        return a.rows == b.rows && a.cols == b.cols && a.channels == b.channels && a.mask.empty() == b.mask.empty() &&
               a.threshold == b.threshold && a.sel == b.sel;
    };
    for (;;) {
        bool progressed = false;
        bool any_running = false;
        for (size_t s = 0; s < slots_.size(); ++s) {
            if (running_[s].empty()) continue;
            bool ready = false;
            try {
                ready = slots_[s]->graph_ready();
            } catch (const Error& e) {
                ready = true;  // graph_finish reports the error for these requests
            }
            if (ready) {
                finish((int)s);
                progressed = true;
            } else {
                any_running = true;
            }
        }
        // admissions, oldest first; queued requests that can share its launch join the oldest
        for (;;) {
            std::vector<std::shared_ptr<Request>> group;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (queue_.empty()) break;
                group.push_back(queue_.front());
                for (size_t i = 1; i < queue_.size() && (int)group.size() < max_per_launch_; ++i)
                    if (same_launch(*queue_[i], *group[0])) group.push_back(queue_[i]);
            }
            const auto& r = group[0];
            bool never = false;
            int s = -1, n = (int)group.size();
            int code = kOk;
            std::string msg;
            try {
                // as many as the budget allows, in powers of two (few captured graphs per size); idle slots
                // give their memory back only to make room for one request
                while (n & (n - 1)) n &= n - 1;
                for (; n >= 1; n /= 2) {
                    s = admit_slot(*r, n, any_running, &never, n == 1);
                    if (s != -1) break;
                }
                if (never) {
                    code = kEBudget;
                    msg = "request needs more device memory than the server's budget";
                }
            } catch (const Error& e) {
                code = e.code;
                msg = e.what();
            }
            if (code != kOk) {
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    queue_.pop_front();
                }
                fail(r, code, msg);
                progressed = true;
                continue;
            }
            if (s == -2) break;
            if (s < 0) {  // a slot is free but the budget is not: deferred until memory frees up
                if (!r->counted_deferral) {
                    r->counted_deferral = true;
                    std::lock_guard<std::mutex> lk(mu_);
                    ++stats_.deferrals;
                }
                break;
            }
            group.resize(n);
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (const auto& g : group)
                    for (auto it = queue_.begin(); it != queue_.end(); ++it)
                        if (*it == g) {
                            queue_.erase(it);
                            break;
                        }
                stats_.admitted += n;
                ++stats_.launches;
            }
            std::vector<ImageArgs> ims(n);
            for (int k = 0; k < n; ++k) {
                const Request& q = *group[k];
                ims[k].img = q.img.data();
                ims[k].rows = q.rows;
                ims[k].cols = q.cols;
                ims[k].stride = q.cols * q.channels;
                ims[k].channels = q.channels;
                ims[k].mask = q.mask.empty() ? nullptr : q.mask.data();
                ims[k].device = false;
            }
            try {
                slots_[s]->tuning = det_.tuning;
                slots_[s]->graph_launch_batch(det_.bank(), ims.data(), n, r->threshold, r->sel);
                hip_check(hipLaunchHostFunc(slots_[s]->stream(), &Server::on_done, this), "host callback");
                running_[s] = group;
                any_running = true;
                const auto now = std::chrono::steady_clock::now();
                for (auto& g : group) g->t_admit = now;
                const size_t b = fixed_ + slots_bytes();
                std::lock_guard<std::mutex> lk(mu_);
                if ((long long)b > stats_.peak_bytes) stats_.peak_bytes = (long long)b;
            } catch (const Error& e) {
                if (slots_[s]->graph_pending()) {  // the launch went out; its results are never read
                    try {
                        (void)slots_[s]->graph_finish_batch(det_.bank());
                    } catch (...) {
                    }
                }
                for (auto& g : group) fail(g, e.code, e.what());
            }
            progressed = true;
        }
        if (progressed) continue;
        std::unique_lock<std::mutex> lk(mu_);
        if (stop_ && queue_.empty() && !any_running) break;
        // a finished launch (host callback), a submission or shutdown wakes the loop; the timeout only
        // bounds a missed wake-up
        const long long seen = submitted_;
        cv_work_.wait_for(lk, std::chrono::milliseconds(any_running ? 2 : 5),
                          [&] { return wake_ || stop_ || submitted_ != seen; });
        wake_ = false;
    }
}

}  // namespace sbm
