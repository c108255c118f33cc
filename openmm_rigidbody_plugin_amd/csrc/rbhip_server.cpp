// Job server over the rbhip C ABI (include/rbhip.h, rbhip_server_*).
//
// Jobs are independent rigid-body systems with a step count.  They wait in
// arrival order and are admitted while their device state fits the HBM
// budget; admitted jobs stay resident and are stepped together as one batch
// (one launch per phase over all their tiles).  Each round advances every
// resident job by the same number of steps, at most stepsPerRound and never
// past the end of the job that finishes first, so a job leaves at exactly
// its step count; finished jobs hand their final state to host memory and
// free their device state, and the next waiting jobs are admitted in the
// following round (continuous batching).
//
// A resident job is stepped with the same kernels and tiles as when it is
// stepped alone with the same workgroup size, so its result does not depend
// on which other jobs share its rounds.
//
// Threads: run_round (the caller's thread) only hands waiting jobs that fit
// the budget to the admission threads, hands finished jobs to a collection
// thread, updates the batch members and queues the next round; it waits for
// the device only to keep at most kRoundsInFlight rounds queued.  The
// admission threads build new jobs' device state, several at a time (each
// set-up waits for its copies and reductions on the device); the collection
// thread downloads a finished job's state and frees it.  A context belongs
// to one thread at a time: created by an admission thread, stepped by
// run_round, released by the collection thread.
//
// Streams: each of the two batches runs on a stream of its own, the
// resident jobs share one (the batch orders it after each advance), each
// admission thread has one and finished jobs rotate over kRetireStreams more,
// so neither setting up a new job nor collecting a finished one waits behind
// the queued steps when the streams map to different hardware queues.  HIP
// maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by
// default); streams that share one run in order, so a set-up can wait behind
// queued rounds, which is why several set-ups run at a time.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "rbhip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    long long id = 0;
    int numAtoms = 0;
    // the caller's arrays, read once at admission (velocities and the tether
    // may be null)
    const double *masses = nullptr, *positions = nullptr, *velocities = nullptr;
    const double *tetherK = nullptr, *tetherR0 = nullptr;
    const int* bodyIndices = nullptr;
    double stepSize = 0.0;
    int rotationMode = 0;
    long long steps = 0, done = 0;
    long long bodies = 0;
    size_t bytes = 0;                 // admission estimate
    rbhip_context* ctx = nullptr;
    Clock::time_point submitted, admitted, finished;
    std::vector<double> outPos, outVel;
    double ke[2] = {0.0, 0.0};
};

struct ServerError : std::runtime_error {
    int code;
    ServerError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_err;

void need(bool ok, const std::string& msg, int code = RBHIP_E_ARG) {
    if (!ok) throw ServerError(code, msg);
}

void call(int status) {
    if (status != RBHIP_OK) throw ServerError(status, rbhip_last_error());
}

template <class F> int guard(F&& f) {
    try {
        f();
        return RBHIP_OK;
    } catch (const ServerError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return RBHIP_E_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return RBHIP_E_ARG;
    }
}

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

constexpr int kAdmitPerRound = 8;   // jobs handed to the set-up thread per round, at most
constexpr int kRoundsInFlight = 2;  // rounds queued on the device ahead of the host, at most
constexpr int kAdmitThreads = 4;    // jobs set up in parallel (each waits on the device in turns)

struct rbhip_server {
    int device = 0;
    size_t budget = 0;
    int stepsPerRound = 0;
    int precision = RBHIP_PRECISION_MIXED;
    int block = 256;
    int exactKind = -1;                 // rotation kind of the jobs: 1 exact, 0 NO-SQUISH
    int admissionOrder = RBHIP_ADMIT_FIFO;
    long long nextId = 1;
    long long bodySteps = 0;
    long long rounds = 0;
    long long shortRounds = 0;          // rounds cut below stepsPerRound by a finishing job
    // seconds: admit and collect (worker threads), waiting for a queued
    // round, rebuilding a member list, queueing a round (run_round thread),
    // and the rounds' own duration on the device (events)
    double phase[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

    // Shared with the worker threads (under mtx): jobs chosen for admission,
    // jobs whose contexts are ready, finished jobs to collect, completed jobs,
    // the device bytes held, and the first worker error.  kAdmitThreads
    // threads set up new jobs, one more collects finished ones, so an
    // admission never waits behind a download.
    std::mutex mtx;
    std::condition_variable wake;       // the worker threads wait for work
    std::condition_variable ready;      // run_round waits for an admission
    std::deque<std::unique_ptr<Job>> queue;          // waiting (run_round thread only)
    std::deque<std::unique_ptr<Job>> toAdmit;
    std::vector<std::unique_ptr<Job>> admitted;
    std::deque<std::unique_ptr<Job>> toCollect;
    std::map<long long, std::unique_ptr<Job>> completed;
    size_t inUse = 0;
    int busy[2] = {0, 0};               // jobs being admitted, being collected
    bool stop = false;
    int workerStatus = RBHIP_OK;
    std::string workerError;
    std::thread workers[kAdmitThreads + 1];   // admission threads, then the collection thread

    // run_round thread only
    std::vector<std::unique_ptr<Job>> resident;     // batch members
    // two batches: one runs the queued round while the other takes a new
    // member list; members[b] holds the ids of the jobs batch b was last
    // given (ids, not context addresses: a finished job's context may be
    // freed and a new job's context created at the same address)
    rbhip_batch* batches[2] = {nullptr, nullptr};
    std::vector<long long> members[2];
    int cur = 0;
    hipStream_t batchStreams[2] = {nullptr, nullptr};
    hipStream_t memberStream = nullptr;
    hipStream_t admitStreams[kAdmitThreads] = {};
    // finished jobs move to retire stream (round index mod kRetireStreams):
    // a download then waits only for the rounds queued on that stream, not
    // for the jobs that finished after it
    static constexpr int kRetireStreams = 4;
    hipStream_t retireStreams[kRetireStreams] = {nullptr, nullptr, nullptr, nullptr};
    long long queuedRounds = 0;
    struct RoundEvents { hipEvent_t start, end; };
    std::deque<RoundEvents> inFlight;   // the queued rounds, oldest first
    std::vector<RoundEvents> spareEvents;
    double deviceSeconds = 0.0;         // sum of the retired rounds' durations

    std::vector<hipStream_t> allStreams() const {
        std::vector<hipStream_t> v = {batchStreams[0], batchStreams[1], memberStream};
        v.insert(v.end(), admitStreams, admitStreams + kAdmitThreads);
        v.insert(v.end(), retireStreams, retireStreams + kRetireStreams);
        return v;
    }

    ~rbhip_server() {
        {
            std::lock_guard<std::mutex> lock(mtx);
            stop = true;
        }
        wake.notify_all();
        for (std::thread& t : workers)
            if (t.joinable()) t.join();
        for (hipStream_t st : allStreams())
            if (st) (void)hipStreamSynchronize(st);
        for (rbhip_batch* b : batches)
            if (b) (void)rbhip_batch_destroy(b);
        for (auto& j : resident)
            if (j->ctx) (void)rbhip_destroy(j->ctx);
        for (auto& j : admitted)
            if (j->ctx) (void)rbhip_destroy(j->ctx);
        for (auto& j : toCollect)
            if (j->ctx) (void)rbhip_destroy(j->ctx);
        for (RoundEvents e : inFlight) {
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.end);
        }
        for (RoundEvents e : spareEvents) {
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.end);
        }
        for (hipStream_t st : allStreams())
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
            }
    }

    // ---- worker threads ----
    // creates a job's context on an admission stream and moves it to the
    // resident jobs' stream
    void admit(Job& j, hipStream_t admitStream) {
        rbhip_context* c = nullptr;
        call(rbhip_create(j.numAtoms, j.masses, j.bodyIndices, j.stepSize, j.rotationMode, precision, device, &c));
        try {
            call(rbhip_set_stream(c, admitStream));
            call(rbhip_set_launch(c, block, 0));
            if (j.tetherK) call(rbhip_set_tether(c, j.tetherK, j.tetherR0));
            call(rbhip_set_positions(c, j.positions));
            if (j.velocities) call(rbhip_set_velocities(c, j.velocities));
            call(rbhip_set_stream(c, memberStream));
        } catch (...) {
            (void)rbhip_destroy(c);
            throw;
        }
        j.ctx = c;
        j.admitted = Clock::now();
        // the context holds its own copies now
        j.masses = j.positions = j.velocities = j.tetherK = j.tetherR0 = nullptr;
        j.bodyIndices = nullptr;
    }

    // downloads a finished job's final state (on the retire stream, after the
    // round it finished in) and frees its device state
    void collect(Job& j) {
        j.outPos.resize(3 * (size_t)j.numAtoms);
        j.outVel.resize(3 * (size_t)j.numAtoms);
        call(rbhip_get_positions(j.ctx, j.outPos.data()));
        call(rbhip_get_velocities(j.ctx, j.outVel.data()));
        call(rbhip_kinetic_energies(j.ctx, j.ke));
        call(rbhip_destroy(j.ctx));
        j.ctx = nullptr;
        j.finished = Clock::now();
    }

    // an admission thread (slot < kAdmitThreads) or the collection thread
    void work(int slot) {
        const bool collecting = slot == kAdmitThreads;
        std::unique_lock<std::mutex> lock(mtx);
        if (hipSetDevice(device) != hipSuccess) {
            workerStatus = RBHIP_E_HIP;
            workerError = "worker thread: hipSetDevice failed";
            ready.notify_all();
            return;
        }
        std::deque<std::unique_ptr<Job>>& in = collecting ? toCollect : toAdmit;
        for (;;) {
            wake.wait(lock, [&] { return stop || !in.empty(); });
            if (stop) return;
            std::unique_ptr<Job> j = std::move(in.front());
            in.pop_front();
            busy[collecting]++;
            lock.unlock();
            Clock::time_point t0 = Clock::now();
            int status = RBHIP_OK;
            std::string error;
            try {
                if (collecting) collect(*j);
                else admit(*j, admitStreams[slot]);
            } catch (const ServerError& e) {
                status = e.code;
                error = e.what();
            } catch (const std::exception& e) {
                status = RBHIP_E_ARG;
                error = e.what();
            }
            double dt = seconds(t0, Clock::now());
            lock.lock();
            busy[collecting]--;
            phase[collecting ? 1 : 0] += dt;
            if (status != RBHIP_OK) {
                if (workerStatus == RBHIP_OK) {
                    workerStatus = status;
                    workerError = (collecting ? "collecting job " : "admitting job ") + std::to_string(j->id) +
                                  ": " + error;
                }
                if (j->ctx) (void)rbhip_destroy(j->ctx);
                inUse -= std::min(inUse, j->bytes);
            } else if (collecting) {
                inUse -= std::min(inUse, j->bytes);
                long long id = j->id;
                completed[id] = std::move(j);
            } else {
                admitted.push_back(std::move(j));
            }
            ready.notify_all();
        }
    }

    // ---- run_round thread ----
    void checkWorker() {
        if (workerStatus != RBHIP_OK) throw ServerError(workerStatus, workerError);
    }

    // waits for the oldest queued round to finish
    void retireOldestRound() {
        RoundEvents e = inFlight.front();
        inFlight.pop_front();
        float ms = 0.0f;
        if (hipEventSynchronize(e.end) != hipSuccess || hipEventElapsedTime(&ms, e.start, e.end) != hipSuccess)
            throw ServerError(RBHIP_E_HIP, "round event failed");
        deviceSeconds += 1e-3 * ms;
        spareEvents.push_back(e);
    }

    // One round.
    //  1. Jobs whose contexts the set-up thread has finished join the
    //     resident jobs; waiting jobs that fit the budget go to the set-up
    //     thread (at most kAdmitPerRound per round).  With nothing resident,
    //     the round waits for the first admission.
    //  2. Jobs that reached their step count move to the retire stream, which
    //     continues after the round they finished in, and go to the set-up
    //     thread for collection.
    //  3. With kRoundsInFlight rounds queued, the oldest is waited for.  The
    //     next round is queued on the same batch if the member list is
    //     unchanged, else on the other batch (idle: its tables are rebuilt
    //     while the queued round runs); it starts when that round ends.
    int round() {
        std::vector<std::unique_ptr<Job>> done, next;
        for (auto& j : resident) (j->done >= j->steps ? done : next).push_back(std::move(j));
        resident.clear();
        for (auto& j : done) call(rbhip_set_stream(j->ctx, retireStreams[queuedRounds % kRetireStreams]));
        int finished = 0;
        {
            std::unique_lock<std::mutex> lock(mtx);
            checkWorker();
            for (auto& j : done) toCollect.push_back(std::move(j));
            size_t before = completed.size();
            int handed = 0;
            while (!queue.empty() && inUse + queue.front()->bytes <= budget && handed < kAdmitPerRound) {
                inUse += queue.front()->bytes;
                toAdmit.push_back(std::move(queue.front()));
                queue.pop_front();
                handed++;
            }
            if (handed || !toCollect.empty()) wake.notify_all();
            if (next.empty() && admitted.empty() && (!toAdmit.empty() || busy[0] > 0))
                ready.wait(lock, [&] { return !admitted.empty() || workerStatus != RBHIP_OK ||
                                              (toAdmit.empty() && busy[0] == 0); });
            checkWorker();
            for (auto& j : admitted) next.push_back(std::move(j));
            admitted.clear();
            finished = (int)(completed.size() - before);
        }
        Clock::time_point t0 = Clock::now();
        std::vector<rbhip_context*> m;
        std::vector<long long> ids;
        for (auto& j : next) {
            m.push_back(j->ctx);
            ids.push_back(j->id);
        }
        if (!m.empty() && (int)inFlight.size() >= kRoundsInFlight) retireOldestRound();
        Clock::time_point t1 = Clock::now();
        if (!m.empty() && ids != members[cur]) {
            cur = 1 - cur;
            if (ids != members[cur]) {
                call(rbhip_batch_set_members(batches[cur], (int)m.size(), m.data()));
                members[cur] = ids;
            }
        }
        Clock::time_point t2 = Clock::now();
        long long n = stepsPerRound;
        for (auto& j : next) n = std::min(n, j->steps - j->done);
        if (!next.empty() && n > 0) {
            RoundEvents e{nullptr, nullptr};
            if (!spareEvents.empty()) {
                e = spareEvents.back();
                spareEvents.pop_back();
            } else if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.end) != hipSuccess) {
                throw ServerError(RBHIP_E_HIP, "event creation failed");
            }
            // rounds run one after another (consecutive rounds share members);
            // waiting here makes the start event mark this round's own start
            if (!inFlight.empty() && hipStreamWaitEvent(batchStreams[cur], inFlight.back().end, 0) != hipSuccess)
                throw ServerError(RBHIP_E_HIP, "stream wait failed");
            if (hipEventRecord(e.start, batchStreams[cur]) != hipSuccess)
                throw ServerError(RBHIP_E_HIP, "event record failed");
            call(rbhip_batch_execute(batches[cur], (int)n));
            if (hipEventRecord(e.end, batchStreams[cur]) != hipSuccess) throw ServerError(RBHIP_E_HIP, "event record failed");
            inFlight.push_back(e);
            queuedRounds++;
            if (n < stepsPerRound) shortRounds++;
            for (auto& j : next) {
                j->done += n;
                bodySteps += j->bodies * n;
            }
        }
        resident.swap(next);
        // a batch whose member list held finished jobs forgets it (their
        // handles are gone)
        for (int b = 0; b < 2; b++)
            if (b != cur || resident.empty()) members[b].clear();
        Clock::time_point t3 = Clock::now();
        std::lock_guard<std::mutex> lock(mtx);
        rounds = queuedRounds;
        phase[2] += seconds(t0, t1);
        phase[3] += seconds(t1, t2);
        phase[4] += seconds(t2, t3);
        phase[5] = deviceSeconds;
        return finished;
    }
};

extern "C" {

const char* rbhip_server_last_error(void) { return g_err.c_str(); }

int rbhip_server_create(int device, size_t hbmBudget, int stepsPerRound, int precision, int blockSize,
                        rbhip_server** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        *out = nullptr;
        need(hbmBudget > 0, "HBM budget must be positive");
        need(stepsPerRound > 0, "steps per round must be positive");
        need(precision == RBHIP_PRECISION_MIXED || precision == RBHIP_PRECISION_DOUBLE, "unknown precision");
        need(blockSize >= 64 && blockSize <= 512 && blockSize % 64 == 0, "block size must be 64..512, a multiple of 64");
        std::unique_ptr<rbhip_server> s(new rbhip_server());
        s->device = device;
        s->budget = hbmBudget;
        s->stepsPerRound = stepsPerRound;
        s->precision = precision;
        s->block = blockSize;
        std::vector<hipStream_t*> streams = {&s->batchStreams[0], &s->batchStreams[1], &s->memberStream};
        for (hipStream_t& st : s->admitStreams) streams.push_back(&st);
        for (hipStream_t& st : s->retireStreams) streams.push_back(&st);
        for (hipStream_t* st : streams)
            if (hipStreamCreateWithFlags(st, hipStreamNonBlocking) != hipSuccess)
                throw ServerError(RBHIP_E_HIP, "stream creation failed");
        for (int b = 0; b < 2; b++) {
            call(rbhip_batch_create(device, &s->batches[b]));
            call(rbhip_batch_set_stream(s->batches[b], s->batchStreams[b]));
            call(rbhip_batch_set_launch(s->batches[b], blockSize, 0));
        }
        rbhip_server* p = s.get();
        for (int w = 0; w <= kAdmitThreads; w++) s->workers[w] = std::thread([p, w] { p->work(w); });
        *out = s.release();
    });
}

int rbhip_server_destroy(rbhip_server* s) {
    return guard([&] { delete s; });
}

int rbhip_server_set_admission_order(rbhip_server* s, int order) {
    return guard([&] {
        need(s != nullptr, "null server");
        need(order == RBHIP_ADMIT_FIFO || order == RBHIP_ADMIT_LONGEST_FIRST, "unknown admission order");
        s->admissionOrder = order;
        // job ids grow with submission: FIFO is ascending id
        std::stable_sort(s->queue.begin(), s->queue.end(),
                         [&](const std::unique_ptr<Job>& a, const std::unique_ptr<Job>& b) {
                             return order == RBHIP_ADMIT_FIFO ? a->id < b->id : a->steps > b->steps;
                         });
    });
}

int rbhip_server_submit(rbhip_server* s, int numAtoms, const double* masses, const int* bodyIndices,
                        const double* positions, const double* velocities, const double* tetherK,
                        const double* tetherR0, double stepSize, int rotationMode, long long steps,
                        long long* jobId) {
    return guard([&] {
        need(s && numAtoms > 0 && masses && bodyIndices && positions && jobId, "bad arguments");
        need(steps > 0, "a job needs a positive step count");
        need(rotationMode >= 0, "Rotation mode cannot be negative");
        need(!tetherK || tetherR0, "tether anchors missing");
        int kind = rotationMode == 0 ? 1 : 0;
        need(s->exactKind < 0 || s->exactKind == kind,
             "jobs of one server share the rotation kind (exact or NO-SQUISH)", RBHIP_E_UNSUPPORTED);
        need(!(kind == 1 && s->block > 256), "exact-rotation jobs need a block size <= 256", RBHIP_E_UNSUPPORTED);
        size_t bytes = 0;
        int numBodies = 0;
        call(rbhip_estimate_bytes(numAtoms, bodyIndices, s->precision, &bytes, &numBodies));
        need(bytes <= s->budget, "job needs " + std::to_string(bytes) + " device bytes, more than the whole budget",
             RBHIP_E_NOMEM);
        std::unique_ptr<Job> j(new Job());
        j->id = s->nextId++;
        j->numAtoms = numAtoms;
        j->masses = masses;
        j->bodyIndices = bodyIndices;
        j->positions = positions;
        j->velocities = velocities;
        j->tetherK = tetherK;
        j->tetherR0 = tetherK ? tetherR0 : nullptr;
        j->bodies = numBodies;
        j->stepSize = stepSize;
        j->rotationMode = rotationMode;
        j->steps = steps;
        j->bytes = bytes;
        j->submitted = Clock::now();
        s->exactKind = kind;
        *jobId = j->id;
        auto at = s->queue.end();
        if (s->admissionOrder == RBHIP_ADMIT_LONGEST_FIRST)
            at = std::find_if(s->queue.begin(), s->queue.end(),
                              [&](const std::unique_ptr<Job>& w) { return w->steps < j->steps; });
        s->queue.insert(at, std::move(j));
    });
}

int rbhip_server_run_round(rbhip_server* s, int* finished) {
    return guard([&] {
        need(s != nullptr, "null server");
        int f = s->round();
        if (finished) *finished = f;
    });
}

int rbhip_server_status(rbhip_server* s, long long* counts3, size_t* bytesInUse, long long* bodySteps) {
    return guard([&] {
        need(s != nullptr, "null server");
        std::lock_guard<std::mutex> lock(s->mtx);
        if (counts3) {
            // jobs in set-up or collection hold device memory: resident
            counts3[0] = (long long)s->queue.size();
            counts3[1] = (long long)(s->resident.size() + s->toAdmit.size() + s->admitted.size() +
                                     s->toCollect.size() + s->busy[0] + s->busy[1]);
            counts3[2] = (long long)s->completed.size();
        }
        if (bytesInUse) *bytesInUse = s->inUse;
        if (bodySteps) *bodySteps = s->bodySteps;
    });
}

int rbhip_server_profile(rbhip_server* s, long long* rounds2, double* seconds6) {
    return guard([&] {
        need(s != nullptr, "null server");
        std::lock_guard<std::mutex> lock(s->mtx);
        if (rounds2) {
            rounds2[0] = s->rounds;
            rounds2[1] = s->shortRounds;
        }
        if (seconds6)
            for (int i = 0; i < 6; i++) seconds6[i] = s->phase[i];
    });
}

int rbhip_server_completed(rbhip_server* s, long long* ids, int max, int* count) {
    return guard([&] {
        need(s && count && (max == 0 || ids), "bad arguments");
        std::lock_guard<std::mutex> lock(s->mtx);
        int n = 0;
        for (auto& kv : s->completed) {
            if (n >= max) break;
            ids[n++] = kv.first;
        }
        *count = n;
    });
}

int rbhip_server_take(rbhip_server* s, long long jobId, double* positions, double* velocities, double* ke2,
                      double* times3) {
    return guard([&] {
        need(s != nullptr, "null server");
        std::unique_ptr<Job> owned;
        {
            std::lock_guard<std::mutex> lock(s->mtx);
            auto it = s->completed.find(jobId);
            need(it != s->completed.end(), "no completed job with id " + std::to_string(jobId), RBHIP_E_STATE);
            owned = std::move(it->second);
            s->completed.erase(it);
        }
        Job& j = *owned;
        if (positions) std::copy(j.outPos.begin(), j.outPos.end(), positions);
        if (velocities) std::copy(j.outVel.begin(), j.outVel.end(), velocities);
        if (ke2) {
            ke2[0] = j.ke[0];
            ke2[1] = j.ke[1];
        }
        if (times3) {
            times3[0] = seconds(j.submitted, j.admitted);    // waiting for admission
            times3[1] = seconds(j.admitted, j.finished);     // resident
            times3[2] = seconds(j.submitted, j.finished);    // completion latency
        }
    });
}

}  // extern "C"
