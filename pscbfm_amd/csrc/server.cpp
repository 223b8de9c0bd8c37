// server.cpp -- continuous batching of simulation jobs on one device.
#include "server.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "updater.hpp"

namespace pscbfm {

Server::Server(int gpu, uint64_t hbm_budget, uint32_t chunk)
    : mEngine(gpu), mBudget(hbm_budget), mChunk(chunk), mT0(std::chrono::steady_clock::now())
{
    if (chunk == 0)
        throw std::invalid_argument("server: chunk must be positive");
    if (mBudget == 0) {
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        mBudget = uint64_t(double(free_b) * 0.8);
    }
}

double Server::now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - mT0).count();
}

Server::Job &Server::job(uint64_t id)
{
    auto it = mJobs.find(id);
    if (it == mJobs.end())
        throw std::invalid_argument("server: unknown job id");
    return *it->second;
}

const Server::Job &Server::job(uint64_t id) const
{
    auto it = mJobs.find(id);
    if (it == mJobs.end())
        throw std::invalid_argument("server: unknown job id");
    return *it->second;
}

uint64_t Server::submit(HostSystem &&sys, uint64_t seed, int64_t age, uint32_t n_mcs)
{
    auto j = std::make_unique<Job>();
    j->n = sys.n;
    j->bytes = BatchEngine::bytesFor(sys);
    j->sys = std::move(sys);
    j->seed = seed;
    j->age = age;
    j->n_mcs = n_mcs;
    j->t_submit = now();
    const uint64_t id = mNext++;
    if (j->bytes > mBudget) {
        std::ostringstream msg;
        msg << "job needs " << j->bytes << " device bytes, more than the whole budget of " << mBudget;
        j->state = kFailed;
        j->error = msg.str();
        j->t_finish = j->t_submit;
        j->sys = HostSystem();
    } else if (n_mcs == 0) {
        j->state = kDone;
        j->t_admit = j->t_finish = j->t_submit;
    } else {
        mQueue.push_back(id);
    }
    mJobs[id] = std::move(j);
    return id;
}

void Server::collect(bool block)
{
    std::vector<uint64_t> left;
    for (size_t k = 0; k < mRetiring.size(); ++k) {
        Job &j = job(mRetiring[k]);
        if (block && k == 0)
            mEngine.wait(j.download);
        else if (!mEngine.finished(j.download)) {
            left.push_back(mRetiring[k]);
            continue;
        }
        // the positions are in pinned memory, unsorted and unwrapped: the
        // member's host arrays are no longer needed
        mEngine.forget(j.member);
        j.admitted = false;
        j.state = kDone;
        j.t_finish = now();
    }
    mRetiring.swap(left);
}

uint32_t Server::step()
{
    collect(false);
    // admission in arrival order while the budget allows (no overtaking, so a
    // large job is not starved by small ones behind it)
    while (!mQueue.empty()) {
        Job &j = job(mQueue.front());
        if (mInUse + j.bytes > mBudget)
            break;
        try {
            // a re-sort falls inside the job only if one of its MCS after the
            // first has an age that is a multiple of the sort period
            const int64_t P = j.sys.steps_between_sortings;
            const bool may_sort = j.sys.spatial_order && P > 0 && (j.age / P + 1) * P <= j.age + int64_t(j.n_mcs) - 1;
            j.member = mEngine.add(std::move(j.sys), j.seed, j.age, false, may_sort);   // the engine owns it now
            j.sys = HostSystem();
            j.admitted = true;
            j.state = kRunning;
            j.t_admit = now();
            mInUse += j.bytes;
            mRunning.push_back(mQueue.front());
        } catch (const std::exception &e) {
            j.state = kFailed;
            j.error = e.what();
            j.t_finish = now();
        }
        mQueue.pop_front();
    }
    if (mRunning.empty()) {
        // nothing to run (so nothing to admit): wait for the results instead
        // of returning to a caller that would spin
        if (!mRetiring.empty())
            collect(true);
        return pending();
    }

    // jobs with less than a chunk left run their remainder first
    std::vector<uint64_t> full;
    std::map<uint32_t, std::vector<uint32_t>> partial;
    std::vector<uint32_t> members;
    for (uint64_t id : mRunning) {
        Job &j = job(id);
        const uint32_t left = j.n_mcs - j.done;
        if (left >= mChunk)
            members.push_back(j.member);
        else
            partial[left].push_back(j.member);
    }
    for (auto &kv : partial)
        mEngine.runMembers(kv.second, kv.first, kv.first);
    if (!members.empty())
        mEngine.runMembers(members, mChunk, mChunk);

    std::vector<uint64_t> still;
    for (uint64_t id : mRunning) {
        Job &j = job(id);
        const uint32_t ran = std::min(mChunk, j.n_mcs - j.done);
        j.done += ran;
        mAttempts += uint64_t(ran) * j.n;
        if (j.done >= j.n_mcs) {
            j.download = mEngine.download(j.member, true);
            j.age = mEngine.age(j.member);
            mEngine.remove(j.member);   // freed after the download, in stream order
            mInUse -= j.bytes;
            mRetiring.push_back(id);
        } else {
            still.push_back(id);
        }
    }
    mRunning.swap(still);
    collect(false);
    return pending();
}

void Server::cancel(uint64_t id)
{
    Job &j = job(id);
    if (j.state == kQueued) {
        mQueue.erase(std::find(mQueue.begin(), mQueue.end(), id));
    } else if (j.state == kRunning) {
        if (std::find(mRetiring.begin(), mRetiring.end(), id) != mRetiring.end())
            return;   // finished on the device; its result is on its way
        mEngine.forget(j.member);   // device memory freed in stream order, host arrays now
        j.admitted = false;
        mInUse -= j.bytes;
        mRunning.erase(std::find(mRunning.begin(), mRunning.end(), id));
    } else {
        return;
    }
    j.state = kCancelled;
    j.t_finish = now();
}

void Server::release(uint64_t id)
{
    Job &j = job(id);
    if (j.state == kQueued || j.state == kRunning)
        throw std::invalid_argument("server: release of a job that has not finished");
    if (j.admitted)
        mEngine.forget(j.member);
    mEngine.give(j.download);
    mJobs.erase(id);
}

Server::State Server::state(uint64_t id) const { return job(id).state; }
uint32_t Server::mcsDone(uint64_t id) const { return job(id).done; }
const std::string &Server::error(uint64_t id) const { return job(id).error; }

size_t Server::hostBytes() const
{
    size_t b = mEngine.hostBytes();
    for (const auto &kv : mJobs) {
        const HostSystem &s = kv.second->sys;
        b += sizeof(int32_t) * (s.xyz.capacity() + s.attribute.capacity()) + s.nnbr.capacity() +
             sizeof(uint32_t) * (s.nbr.capacity() + s.colour.capacity() + s.to_sorted.capacity() +
                                 s.ppos_sorted.capacity() + s.ids_sorted.capacity() + s.nbr_sorted.capacity() +
                                 s.lattice.capacity()) +
             sizeof(int32_t) * s.image_sorted.capacity() + mEngine.pinnedBytes(kv.second->download);
    }
    return b;
}

void Server::result(uint64_t id, int32_t *xyz, uint32_t n) const
{
    const Job &j = job(id);
    if (j.state != kDone)
        throw std::invalid_argument("server: the job has no result (not done)");
    if (n != j.n)
        throw std::invalid_argument("server: the result has as many positions as the job has monomers");
    if (j.download.held)
        mEngine.copyOut(j.download, xyz);          // unsorted and unwrapped on the device
    else
        std::memcpy(xyz, j.sys.xyz.data(), sizeof(int32_t) * 3 * size_t(n));   // a job of 0 MCS
}

void Server::times(uint64_t id, double &submitted, double &admitted, double &finished) const
{
    const Job &j = job(id);
    submitted = j.t_submit;
    admitted = j.t_admit;
    finished = j.t_finish;
}

}  // namespace pscbfm
