// server.hpp -- served simulation jobs with continuous batching.
//
// Jobs (a validated system plus the number of MCS it asks for) queue in
// arrival order.  Each scheduling round admits queued jobs while their device
// bytes fit the HBM budget, runs one chunk of MCS on every admitted job in one
// batched launch sequence (their work overlaps on the device), and retires the
// jobs that are done: the copy of their positions into pinned memory is queued
// behind their last chunk and their device memory is freed in stream order,
// which lets the next round admit more.  No round waits for the device unless
// nothing else can proceed: a retiring job is reported done by the first round
// that finds its copy arrived.  A job's state stays resident on the device
// between chunks.  A job's result is bit-identical to
// running it alone (its trajectory depends only on its seed and age).
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"

namespace pscbfm {

class Server {
public:
    enum State { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3, kCancelled = 4 };

    Server(int gpu, uint64_t hbm_budget, uint32_t chunk);

    uint64_t submit(HostSystem &&sys, uint64_t seed, int64_t age, uint32_t n_mcs);
    // One round; returns the number of jobs still queued or running.
    uint32_t step();
    void cancel(uint64_t job);
    void release(uint64_t job);

    State state(uint64_t job) const;
    uint32_t mcsDone(uint64_t job) const;
    const std::string &error(uint64_t job) const;
    // Final positions of a done job (3 * n int32, by id), from pinned memory
    // where its retirement left them; the job holds them until released.
    void result(uint64_t job, int32_t *xyz, uint32_t n) const;
    // seconds since the server was created: submission, admission, completion
    void times(uint64_t job, double &submitted, double &admitted, double &finished) const;

    uint32_t queued() const { return uint32_t(mQueue.size()); }
    // running jobs, including those whose result is still on its way
    uint32_t running() const { return uint32_t(mRunning.size() + mRetiring.size()); }
    uint64_t budget() const { return mBudget; }
    uint64_t bytesInUse() const { return mInUse; }
    uint64_t attemptsDone() const { return mAttempts; }
    // Host memory held for jobs: queued systems, results, the engine's members.
    size_t hostBytes() const;
    uint32_t jobsHeld() const { return uint32_t(mJobs.size()); }
    BatchEngine &engine() { return mEngine; }

private:
    struct Job {
        HostSystem sys;      // until admission (a job of 0 MCS: its result)
        uint32_t n = 0;      // monomers
        uint64_t seed = 0;
        int64_t age = 0;
        uint32_t n_mcs = 0;
        uint32_t done = 0;
        uint64_t bytes = 0;
        State state = kQueued;
        uint32_t member = 0;
        bool admitted = false;   // the engine still holds its member
        BatchEngine::Download download{};   // while retiring
        std::string error;
        double t_submit = 0, t_admit = -1, t_finish = -1;
    };

    double now() const;
    // Reports the retiring jobs whose positions have arrived; with `block`,
    // first waits for the oldest one.
    void collect(bool block);
    uint32_t pending() const { return uint32_t(mQueue.size() + mRunning.size() + mRetiring.size()); }
    Job &job(uint64_t id);
    const Job &job(uint64_t id) const;

    BatchEngine mEngine;
    uint64_t mBudget;
    uint32_t mChunk;
    uint64_t mInUse = 0;
    uint64_t mNext = 1;
    uint64_t mAttempts = 0;
    std::map<uint64_t, std::unique_ptr<Job>> mJobs;
    std::deque<uint64_t> mQueue;
    std::vector<uint64_t> mRunning;
    std::vector<uint64_t> mRetiring;   // done on the device, positions on their way
    std::chrono::steady_clock::time_point mT0;
};

}  // namespace pscbfm
