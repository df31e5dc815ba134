// Continuous-batching request loop (see sbm_server.cpp).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "sbm_detector.h"

namespace sbm {

constexpr int kETimeout = -11;

struct ServerStats {
    long long submitted = 0, admitted = 0, completed = 0, failed = 0, deferrals = 0, releases = 0;
    long long peak_bytes = 0, budget = 0;  // device bytes of the bank and the slots, at most the budget
    long long launches = 0;  // slot launches; admitted / launches is the mean number of requests per launch
};

class Server {
public:
    // slots: pipelines (streams) that run requests concurrently; budget: device bytes the slots may hold;
    // max_per_launch: queued requests of one size, threshold and selection that one slot launch may take
    Server(Detector& det, int slots, size_t budget, int max_per_launch = 8);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    long long submit(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                     float threshold, const std::vector<std::string>& class_ids);
    // kOk with the matches, the request's error code, or kETimeout (timeout_ms < 0 waits forever)
    int wait(long long id, int timeout_ms, std::vector<MatchRec>* out, double* latency_ms, std::string* err);
    ServerStats stats();

private:
    struct Request {
        long long id = 0;
        std::vector<uint8_t> img, mask;
        int rows = 0, cols = 0, channels = 1;
        float threshold = 0.f;
        std::vector<int> sel;
        std::chrono::steady_clock::time_point t_submit, t_admit, t_done;
        int code = 0;
        std::string message;
        std::vector<MatchRec> result;
        bool counted_deferral = false;
    };
    void loop();
    // a free slot for a launch of n requests like r within the budget (-2: none free, -1: no room);
    // may_release: idle slots' buffers may be released to make room
    int admit_slot(const Request& r, int n, bool any_running, bool* never_fits, bool may_release);
    void finish(int s);
    static void on_done(void* server);  // a slot's stream finished its launch: wake the dispatcher
    void fail(const std::shared_ptr<Request>& r, int code, const std::string& msg);
    size_t slots_bytes() const;

    Detector& det_;
    int device_ = 0;
    size_t budget_;
    size_t fixed_ = 0;  // the detector's device bytes outside the slots (the bank), counted in the budget
    uint64_t bank_version_ = 0;
    int max_per_launch_ = 8;
    std::vector<std::unique_ptr<Pipeline>> slots_;
    std::vector<std::vector<std::shared_ptr<Request>>> running_;  // the requests of each slot's launch
    std::deque<std::shared_ptr<Request>> queue_;
    std::unordered_map<long long, std::shared_ptr<Request>> done_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::thread thread_;
    bool stop_ = false;
    bool wake_ = false;  // set by a finished launch's host callback
    long long next_id_ = 1;
    long long submitted_ = 0;
    ServerStats stats_;
};

}  // namespace sbm
