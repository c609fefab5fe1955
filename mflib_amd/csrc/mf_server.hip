// Request server of the C ABI (mf_server_* in include/mflib_gpu.h).
//
// HBM layout:
//   library   per add call: FP64 raw and prepared templates [n][Lt], and
//             one spectrum per supported FFT size [n][logM][nfft/2+1]
//   arena     one allocation of the budget, carved per request into
//             [inputs copied from host][rho rows (template, channel)]
//   slots     per in-flight batch: descriptor tables and dead-trace partials
//
// Threads: callers submit and wait; a scheduler thread forms batches and
// launches them; a completion thread waits on batch events in launch order
// and wakes the waiters.  Every request owns its lanes (one per channel),
// rows and outputs, so its numbers do not depend on what it is batched
// with.
#include "mflib_gpu.h"
#include "mf_host.hpp"
#include "mf_kernels.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

using namespace mfh;

namespace {

using Clock = std::chrono::steady_clock;

double msBetween(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

enum class ReqState { Pending, Launched, Done, Failed };  // Failed includes cancelled

struct LibTemplate {
    int Lt = 0;
    const double *tmpl = nullptr;                      // prepared FP64 template
    const void *spec[mfk::kMaxLogMFloat + 1] = {};     // per logM, or null
    float hmax[mfk::kMaxLogMFloat + 1] = {};
};

struct LibChunk {
    DevBuf raw, tmpl, spec, hmax;
};

struct Request {
    long long id = 0;
    int nC = 0, nT = 0, nb = 0;
    bool onDevice = false;
    std::vector<int> len;
    std::vector<const void *> src;
    std::vector<int> tmplIds;
    std::vector<LibTemplate> tpl;  // library entries, copied at submission
    std::vector<int> order;        // template indices sorted by length
    std::vector<std::vector<Bucket>> buckets;  // per channel: length buckets (lanes)
    std::vector<int> nOut;
    size_t ldOut = 0;
    size_t inBytes = 0, bytes = 0, off = 0;
    double corrSamples = 0;
    ReqState state = ReqState::Pending;
    bool allocated = false;  // holds [off, off + bytes) of the arena
    int status = MF_OK;
    Clock::time_point tSubmit, tLaunch, tDone;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    PinnedBuf pin;
    DevBuf desc, part;
    std::vector<Request *> reqs;
    bool busy = false;
    int launchStatus = MF_OK;
    std::string launchError;
};

}  // namespace

struct mf_server {
    int precision = MF_PRECISION_FLOAT;
    size_t tsize = 4;
    int device = 0;
    int maxBatch = 1, maxInflight = 1;
    std::vector<LibTemplate> lib;
    std::vector<std::unique_ptr<LibChunk>> chunks;
    std::map<int, TwTables> tw;
    DevBuf arena;
    std::map<size_t, size_t> freeList;  // offset -> bytes, coalesced
    std::map<long long, std::unique_ptr<Request>> reqs;
    std::deque<Request *> pending;
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<int> inflight;
    std::mutex mu;
    std::condition_variable cvSched, cvComplete, cvDone;
    std::thread scheduler, completer;
    bool stop = false;
    long long memEpoch = 0;
    long long nextId = 1;
    long long nBatches = 0, nDone = 0;
    double doneSamples = 0;
    hipStream_t copyStream = nullptr;
    // window statistics (mf_server_window_stats): the device's running
    // counts of windows recomputed exactly (all, in full), and the windows
    // and solo-row windows of the launched batches
    DevBuf dFlagged;
    std::atomic<long long> statWindows{0}, statSolo{0};

    ~mf_server();
};

namespace {

constexpr size_t kAlign = 256;

bool arenaAlloc(mf_server *s, size_t bytes, size_t *off)
{
    for (auto it = s->freeList.begin(); it != s->freeList.end(); ++it) {
        if (it->second >= bytes) {
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t restOff = it->first + bytes;
            s->freeList.erase(it);
            if (rest > 0) { s->freeList[restOff] = rest; }
            return true;
        }
    }
    return false;
}

void arenaFree(mf_server *s, size_t off, size_t bytes)
{
    auto it = s->freeList.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != s->freeList.end() && it->first + it->second == next->first) {
        it->second += next->second;
        s->freeList.erase(next);
    }
    if (it != s->freeList.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            s->freeList.erase(it);
        }
    }
}

Request *findRequest(mf_server *s, long long id)
{
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) { fail(MF_ERR_INVALID_ARGUMENT, "unknown request " + std::to_string(id)); }
    return it->second.get();
}

// Builds the descriptors of one batch in the slot's pinned buffer, copies
// them and any host inputs to the device, and launches the dead-trace pass
// and one correlation launch per FFT size.  Runs without the server lock:
// the batch's requests belong to the launch path until they complete.
void launchBatch(mf_server *s, Slot &slot)
{
    std::vector<mfk::Seg> segs;
    std::vector<mfk::Row> rows;
    long long windows = 0;
    // one launch per (FFT size, kernel variant); each lane's variant depends
    // on the lane alone
    std::map<std::pair<int, int>, std::vector<int>> segsByPlan;
    std::map<std::pair<int, int>, std::vector<long long>> blocksByPlan;
    long long maxN = 1;
    char *arena = static_cast<char *>(s->arena.p);
    for (Request *r : slot.reqs) {
        char *base = arena + r->off;
        char *in = base;
        char *out = base + r->inBytes;
        for (int c = 0; c < r->nC; ++c) {
            const void *x = r->src[c];
            if (!r->onDevice) {
                MF_HIP(hipMemcpyAsync(in, r->src[c], static_cast<size_t>(r->len[c]) * s->tsize,
                                      hipMemcpyHostToDevice, slot.stream));
                x = in;
                in += roundUp(static_cast<size_t>(r->len[c]) * s->tsize, kAlign);
            }
            // one lane per length bucket; rows in template-length order (the
            // kernels walk them so)
            for (const Bucket &bk : r->buckets[c]) {
                const int lg = bk.logM;
                const int L = blockLength(lg, bk.nb);
                mfk::Seg sg{};
                sg.flagged = s->dFlagged.as<unsigned long long>();
                sg.x = x;
                sg.n = r->len[c];
                sg.L = L;
                sg.nOut = r->nOut[c];
                sg.rowBegin = static_cast<int>(rows.size());
                sg.dead = 1;
                float hmax = 0;
                for (int k = bk.begin; k < bk.end; ++k) {
                    const int i = r->order[k];
                    const LibTemplate &t = r->tpl[i];
                    mfk::Row row;
                    row.H = t.spec[lg];
                    row.tmpl = t.tmpl;
                    row.out = out + (static_cast<size_t>(i) * r->nC + c) * r->ldOut * s->tsize;
                    row.nOut = r->nOut[c];
                    row.Lt = t.Lt;
                    row.x = x;
                    rows.push_back(row);
                    windows += row.nOut;
                    hmax = std::max(hmax, t.hmax[lg]);
                }
                sg.rowEnd = static_cast<int>(rows.size());
                sg.hmax = hmax;
                const std::pair<int, int> plan{lg, laneVariant(s->precision, lg, 0, bk.mix)};
                segsByPlan[plan].push_back(static_cast<int>(segs.size()));
                blocksByPlan[plan].push_back((r->nOut[c] + L - 1) / L);
                maxN = std::max<long long>(maxN, r->len[c]);
                segs.push_back(sg);
            }
        }
    }
    struct Launch {
        int logM, first, count, variant;
    };
    std::vector<mfk::Item> items;
    std::vector<Launch> launches;
    for (auto &kv : segsByPlan) {
        Launch la{kv.first.first, static_cast<int>(items.size()), 0, kv.first.second};
        appendXcdItems(kv.second, blocksByPlan[kv.first], items);
        la.count = static_cast<int>(items.size()) - la.first;
        launches.push_back(la);
    }
    long long solo = 0;
    for (const auto &la : launches) {
        if (la.variant != 2) { continue; }
        solo += soloWindows(items.data(), la.first, la.count, segs.data(), rows.data(),
                            mfk::corr2_shared_scale(static_cast<int>(s->tsize), 1 << la.logM));
    }
    s->statWindows += windows;
    s->statSolo += solo;
    const size_t segBytes = roundUp(segs.size() * sizeof(mfk::Seg), kAlign);
    const size_t rowBytes = roundUp(rows.size() * sizeof(mfk::Row), kAlign);
    const size_t itemBytes = roundUp(std::max<size_t>(items.size(), 1) * sizeof(mfk::Item), kAlign);
    const size_t total = segBytes + rowBytes + itemBytes;
    slot.pin.reserve(total);
    if (slot.desc.bytes < total) { slot.desc.alloc(total); }
    char *pin = static_cast<char *>(slot.pin.p);
    std::memcpy(pin, segs.data(), segs.size() * sizeof(mfk::Seg));
    std::memcpy(pin + segBytes, rows.data(), rows.size() * sizeof(mfk::Row));
    std::memcpy(pin + segBytes + rowBytes, items.data(), items.size() * sizeof(mfk::Item));
    MF_HIP(hipMemcpyAsync(slot.desc.p, pin, total, hipMemcpyHostToDevice, slot.stream));
    auto *dSegs = reinterpret_cast<mfk::Seg *>(slot.desc.p);
    auto *dRows = reinterpret_cast<mfk::Row *>(static_cast<char *>(slot.desc.p) + segBytes);
    auto *dItems = reinterpret_cast<mfk::Item *>(static_cast<char *>(slot.desc.p) + segBytes + rowBytes);
    const size_t chunks = static_cast<size_t>((maxN + mfk::kDeadChunk - 1) / mfk::kDeadChunk);
    const size_t partBytes = 2 * sizeof(double) * segs.size() * chunks;
    if (slot.part.bytes < partBytes) { slot.part.alloc(partBytes); }
    MF_HIP(mfk::launch_seg_dead(s->precision, dSegs, static_cast<int>(segs.size()), maxN,
                                slot.part.as<double>(), slot.stream));
    for (const auto &la : launches) {
        const TwTables &t = s->tw.at(la.logM);
        const bool f = s->precision == MF_PRECISION_FLOAT;
        MF_HIP(mfk::launch_corr(s->precision, la.logM, la.variant, dItems + la.first,
                                la.count, dSegs, dRows, f ? t.lo.p : t.lo64.p, f ? t.hi.p : t.hi64.p,
                                mfk::FixQueue{}, slot.stream));
    }
    MF_HIP(hipEventRecord(slot.done, slot.stream));
}

void schedulerLoop(mf_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    long long seenEpoch = -1;
    while (true) {
        s->cvSched.wait(lk, [&] {
            if (s->stop) { return true; }
            if (s->pending.empty()) { return false; }
            bool free = false;
            for (const auto &sl : s->slots) { free = free || !sl->busy; }
            return free && seenEpoch != s->memEpoch;
        });
        if (s->stop) { break; }
        int si = 0;
        while (s->slots[si]->busy) { ++si; }
        Slot &slot = *s->slots[si];
        slot.reqs.clear();
        while (!s->pending.empty() && static_cast<int>(slot.reqs.size()) < s->maxBatch) {
            Request *r = s->pending.front();
            if (!arenaAlloc(s, r->bytes, &r->off)) { break; }  // arrival order: no overtaking
            r->allocated = true;
            s->pending.pop_front();
            r->state = ReqState::Launched;
            r->tLaunch = Clock::now();
            slot.reqs.push_back(r);
        }
        if (slot.reqs.empty()) {
            seenEpoch = s->memEpoch;  // head does not fit: wait for a release
            continue;
        }
        seenEpoch = -1;
        slot.busy = true;
        ++s->nBatches;
        lk.unlock();
        slot.launchStatus = MF_OK;
        try {
            launchBatch(s, slot);
        } catch (const MfError &e) {
            slot.launchStatus = e.code;
            slot.launchError = e.msg;
            (void)hipEventRecord(slot.done, slot.stream);
        }
        lk.lock();
        s->inflight.push_back(si);
        s->cvComplete.notify_one();
    }
}

void completerLoop(mf_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cvComplete.wait(lk, [&] { return !s->inflight.empty() || s->stop; });
        if (s->inflight.empty()) { break; }  // stop requested and drained
        const int si = s->inflight.front();
        Slot &slot = *s->slots[si];
        lk.unlock();
        const hipError_t e = hipEventSynchronize(slot.done);
        const auto now = Clock::now();
        lk.lock();
        s->inflight.pop_front();
        const int st = slot.launchStatus != MF_OK ? slot.launchStatus : (e == hipSuccess ? MF_OK : MF_ERR_DEVICE);
        for (Request *r : slot.reqs) {
            r->tDone = now;
            r->status = st;
            r->state = st == MF_OK ? ReqState::Done : ReqState::Failed;
            if (st == MF_OK) {
                ++s->nDone;
                s->doneSamples += r->corrSamples;
            }
        }
        slot.reqs.clear();
        slot.busy = false;
        s->cvDone.notify_all();
        s->cvSched.notify_one();
    }
}

void stopThreads(mf_server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Request *r : s->pending) {
            r->state = ReqState::Failed;
            r->status = MF_ERR_BAD_STATE;
        }
        s->pending.clear();
    }
    s->cvSched.notify_all();
    if (s->scheduler.joinable()) { s->scheduler.join(); }
    s->cvComplete.notify_all();
    if (s->completer.joinable()) { s->completer.join(); }
    s->cvDone.notify_all();
}

// Library spectra: every supported FFT size with at least one output per
// block, so any plan a request gets finds its spectra resident.
void addTemplates(mf_server *s, int n, const int *lengths, const double *const *samples, int *firstId)
{
    if (n < 1 || lengths == nullptr || samples == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "no templates"); }
    const int hiLog = maxLogM(s->precision);
    size_t tmplTotal = 0, specTotal = 0;
    std::vector<size_t> tOff(n);
    std::vector<std::vector<size_t>> sOff(n, std::vector<size_t>(hiLog + 1, SIZE_MAX));
    for (int i = 0; i < n; ++i) {
        if (lengths[i] < 1 || samples[i] == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "Waveform not yet set on template"); }
        tOff[i] = tmplTotal;
        tmplTotal += roundUp(lengths[i], 32);
        for (int lg = mfk::kMinLogM; lg <= hiLog; ++lg) {
            if ((2LL << lg) - lengths[i] + 1 < 2) { continue; }
            sOff[i][lg] = specTotal;
            specTotal += roundUp((1u << lg) + 1, 32);
        }
    }
    auto chunk = std::make_unique<LibChunk>();
    chunk->raw.alloc(tmplTotal * sizeof(double));
    chunk->tmpl.alloc(tmplTotal * sizeof(double));
    chunk->spec.alloc(std::max<size_t>(specTotal, 1) * 2 * s->tsize);
    chunk->hmax.alloc(static_cast<size_t>(n) * (hiLog + 1) * sizeof(float));
    std::vector<double> raw(tmplTotal, 0.0);
    for (int i = 0; i < n; ++i) { std::copy(samples[i], samples[i] + lengths[i], raw.begin() + tOff[i]); }
    MF_HIP(hipMemcpy(chunk->raw.p, raw.data(), raw.size() * sizeof(double), hipMemcpyHostToDevice));
    for (int lg = mfk::kMinLogM; lg <= hiLog; ++lg) {
        std::vector<mfk::TmplJob> jobs;
        for (int i = 0; i < n; ++i) {
            if (sOff[i][lg] == SIZE_MAX) { continue; }
            mfk::TmplJob j;
            j.t = chunk->raw.as<double>() + tOff[i];
            j.Lt = lengths[i];
            j.tOut = chunk->tmpl.as<double>() + tOff[i];
            j.H = static_cast<char *>(chunk->spec.p) + sOff[i][lg] * 2 * s->tsize;
            j.hmax = chunk->hmax.as<float>() + static_cast<size_t>(i) * (hiLog + 1) + lg;
            jobs.push_back(j);
        }
        if (jobs.empty()) { continue; }
        DevBuf dj;
        dj.alloc(jobs.size() * sizeof(mfk::TmplJob));
        MF_HIP(hipMemcpy(dj.p, jobs.data(), dj.bytes, hipMemcpyHostToDevice));
        const TwTables &t = s->tw.at(lg);
        MF_HIP(mfk::launch_tmpl(s->precision, lg, dj.as<mfk::TmplJob>(), static_cast<int>(jobs.size()), t.lo64.p,
                                t.hi64.p, s->copyStream));
        MF_HIP(hipStreamSynchronize(s->copyStream));
    }
    std::vector<float> hm(static_cast<size_t>(n) * (hiLog + 1));
    MF_HIP(hipMemcpy(hm.data(), chunk->hmax.p, hm.size() * sizeof(float), hipMemcpyDeviceToHost));
    std::lock_guard<std::mutex> lk(s->mu);
    *firstId = static_cast<int>(s->lib.size());
    for (int i = 0; i < n; ++i) {
        LibTemplate t;
        t.Lt = lengths[i];
        t.tmpl = chunk->tmpl.as<double>() + tOff[i];
        for (int lg = mfk::kMinLogM; lg <= hiLog; ++lg) {
            if (sOff[i][lg] == SIZE_MAX) { continue; }
            t.spec[lg] = static_cast<char *>(chunk->spec.p) + sOff[i][lg] * 2 * s->tsize;
            t.hmax[lg] = hm[static_cast<size_t>(i) * (hiLog + 1) + lg];
        }
        s->lib.push_back(t);
    }
    s->chunks.push_back(std::move(chunk));
}

std::unique_ptr<Request> planRequest(mf_server *s, int nC, const int *lengths, const void *const *channels,
                                     int onDevice, int nT, const int *ids)
{
    if (nC < 1 || lengths == nullptr || channels == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "no channels"); }
    if (nT < 1 || ids == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "no templates"); }
    auto r = std::make_unique<Request>();
    r->nC = nC;
    r->nT = nT;
    r->onDevice = onDevice != 0;
    r->len.assign(lengths, lengths + nC);
    r->src.assign(channels, channels + nC);
    r->tmplIds.assign(ids, ids + nT);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int id : r->tmplIds) {
            if (id < 0 || id >= static_cast<int>(s->lib.size())) {
                fail(MF_ERR_INVALID_ARGUMENT, "unknown template id " + std::to_string(id));
            }
            r->tpl.push_back(s->lib[id]);
            r->nb = std::max(r->nb, s->lib[id].Lt);
        }
    }
    r->order.resize(nT);
    for (int i = 0; i < nT; ++i) { r->order[i] = i; }
    std::stable_sort(r->order.begin(), r->order.end(), [&](int a, int b) { return r->tpl[a].Lt < r->tpl[b].Lt; });
    std::vector<int> lts;
    for (int i : r->order) { lts.push_back(r->tpl[i].Lt); }
    int maxOut = 0;
    r->nOut.resize(nC);
    r->buckets.resize(nC);
    for (int c = 0; c < nC; ++c) {
        if (channels[c] == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "channel " + std::to_string(c) + " is NULL"); }
        const long long no = static_cast<long long>(lengths[c]) - r->nb + 1;
        if (no < 1) {
            fail(MF_ERR_INVALID_ARGUMENT, "channel " + std::to_string(c) + " has " + std::to_string(lengths[c]) +
                                              " samples, fewer than the template length " + std::to_string(r->nb));
        }
        r->nOut[c] = static_cast<int>(no);
        r->buckets[c] = (c > 0 && lengths[c] == lengths[c - 1])
                            ? r->buckets[c - 1]
                            : planBuckets(s->precision, 0, lts, std::vector<long long>(lts.size(), no));
        maxOut = std::max(maxOut, r->nOut[c]);
        r->corrSamples += static_cast<double>(no) * nT;
        if (!r->onDevice) { r->inBytes += roundUp(static_cast<size_t>(lengths[c]) * s->tsize, kAlign); }
    }
    r->ldOut = roundUp(static_cast<size_t>(maxOut), 64);
    r->bytes = r->inBytes + roundUp(r->ldOut * nT * nC * s->tsize, kAlign);
    if (r->bytes > s->arena.bytes) {
        fail(MF_ERR_OUT_OF_MEMORY, "request needs " + std::to_string(r->bytes) + " bytes, more than the budget of " +
                                       std::to_string(s->arena.bytes));
    }
    return r;
}

Request *doneRequest(mf_server *s, long long id)
{
    Request *r = findRequest(s, id);
    if (r->state == ReqState::Failed) { fail(r->status, "request " + std::to_string(id) + " failed"); }
    if (r->state != ReqState::Done) { fail(MF_ERR_BAD_STATE, "request " + std::to_string(id) + " has not completed"); }
    return r;
}

void rowOf(mf_server *s, Request *r, int i, int c, const void **p, int *len)
{
    if (i < 0 || i >= r->nT) { fail(MF_ERR_INVALID_ARGUMENT, "template index " + std::to_string(i) + " out of range"); }
    if (c < 0 || c >= r->nC) { fail(MF_ERR_INVALID_ARGUMENT, "channel " + std::to_string(c) + " out of range"); }
    *p = static_cast<const char *>(s->arena.p) + r->off + r->inBytes +
         (static_cast<size_t>(i) * r->nC + c) * r->ldOut * s->tsize;
    *len = r->nOut[c];
}

}  // namespace

mf_server::~mf_server()
{
    stopThreads(this);
    for (auto &sl : slots) {
        if (sl->done) { (void)hipEventDestroy(sl->done); }
        if (sl->stream) { (void)hipStreamDestroy(sl->stream); }
    }
    if (copyStream) { (void)hipStreamDestroy(copyStream); }
}

extern "C" {

int mf_server_create(mf_server **out, int precision, double budgetBytes, int maxBatch, int maxInflight)
{
    return guard([&] {
        if (out == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server pointer is NULL"); }
        *out = nullptr;
        if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(precision));
        }
        if (maxBatch < 1 || maxBatch > 4096 || maxInflight < 1 || maxInflight > 8) {
            fail(MF_ERR_INVALID_ARGUMENT, "maxBatch must be in [1,4096] and maxInflight in [1,8]");
        }
        if (!(budgetBytes >= (1 << 20))) { fail(MF_ERR_INVALID_ARGUMENT, "budget must be at least 1 MiB"); }
        requireDevice();
        auto s = std::make_unique<mf_server>();
        MF_HIP(hipGetDevice(&s->device));
        s->precision = precision;
        s->tsize = precision == MF_PRECISION_FLOAT ? 4 : 8;
        s->maxBatch = maxBatch;
        s->maxInflight = maxInflight;
        size_t freeB = 0, totalB = 0;
        MF_HIP(hipMemGetInfo(&freeB, &totalB));
        const size_t margin = size_t(2) << 30;
        const size_t want = roundUp(static_cast<size_t>(budgetBytes), kAlign);
        if (want + margin > freeB) {
            fail(MF_ERR_OUT_OF_MEMORY, "budget of " + std::to_string(want) + " bytes exceeds free device memory (" +
                                           std::to_string(freeB) + " bytes, 2 GiB kept free)");
        }
        s->arena.alloc(want);
        s->freeList[0] = want;
        s->dFlagged.alloc(2 * sizeof(unsigned long long));
        MF_HIP(hipMemset(s->dFlagged.p, 0, 2 * sizeof(unsigned long long)));
        // twiddle tables of every FFT size, read-only from here on
        for (int lg = mfk::kMinLogM; lg <= maxLogM(precision); ++lg) {
            makeTwiddles(s->tw[lg], lg, precision == MF_PRECISION_FLOAT);
        }
        MF_HIP(hipStreamCreateWithFlags(&s->copyStream, hipStreamNonBlocking));
        for (int i = 0; i < maxInflight; ++i) {
            s->slots.push_back(std::make_unique<Slot>());
            Slot &sl = *s->slots.back();
            MF_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
            MF_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        }
        mf_server *p = s.get();
        p->scheduler = std::thread(schedulerLoop, p);
        p->completer = std::thread(completerLoop, p);
        *out = s.release();
    });
}

int mf_server_destroy(mf_server *s)
{
    return guard([&] { delete s; });
}

int mf_server_add_templates(mf_server *s, int n, const int *lengths, const double *const *samples, int *firstId)
{
    return guard([&] {
        if (s == nullptr || firstId == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server or id pointer is NULL"); }
        addTemplates(s, n, lengths, samples, firstId);
    });
}

int mf_server_submit(mf_server *s, int nChannels, const int *lengths, const void *const *channels, int onDevice,
                     int nTemplates, const int *templateIds, long long *requestId)
{
    return guard([&] {
        if (s == nullptr || requestId == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server or id pointer is NULL"); }
        auto r = planRequest(s, nChannels, lengths, channels, onDevice, nTemplates, templateIds);
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) { fail(MF_ERR_BAD_STATE, "server is stopping"); }
        r->id = s->nextId++;
        r->tSubmit = Clock::now();
        *requestId = r->id;
        s->pending.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        s->cvSched.notify_one();
    });
}

int mf_server_wait(mf_server *s, long long id, int *status)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        std::unique_lock<std::mutex> lk(s->mu);
        Request *r = findRequest(s, id);
        s->cvDone.wait(lk, [&] { return r->state == ReqState::Done || r->state == ReqState::Failed; });
        if (status) { *status = r->status; }
    });
}

int mf_server_poll(mf_server *s, long long id, int *done)
{
    return guard([&] {
        if (s == nullptr || done == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server or flag pointer is NULL"); }
        std::lock_guard<std::mutex> lk(s->mu);
        const Request *r = findRequest(s, id);
        *done = (r->state == ReqState::Done || r->state == ReqState::Failed) ? 1 : 0;
    });
}

int mf_server_get_rho(mf_server *s, long long id, int i, int c, void *out, int cap, int *len)
{
    return guard([&] {
        if (s == nullptr || out == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server or output is NULL"); }
        const void *p = nullptr;
        int n = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            rowOf(s, doneRequest(s, id), i, c, &p, &n);
        }
        if (len) { *len = n; }
        if (cap < n) { fail(MF_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(cap) + " < " + std::to_string(n)); }
        MF_HIP(hipMemcpyAsync(out, p, static_cast<size_t>(n) * s->tsize, hipMemcpyDeviceToHost, s->copyStream));
        MF_HIP(hipStreamSynchronize(s->copyStream));
    });
}

int mf_server_get_rho_device(mf_server *s, long long id, int i, int c, const void **dptr, int *len)
{
    return guard([&] {
        if (s == nullptr || dptr == nullptr || len == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "NULL argument"); }
        std::lock_guard<std::mutex> lk(s->mu);
        rowOf(s, doneRequest(s, id), i, c, dptr, len);
    });
}

int mf_server_latency(mf_server *s, long long id, double *queuedMs, double *totalMs)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        std::lock_guard<std::mutex> lk(s->mu);
        const Request *r = doneRequest(s, id);
        if (queuedMs) { *queuedMs = msBetween(r->tSubmit, r->tLaunch); }
        if (totalMs) { *totalMs = msBetween(r->tSubmit, r->tDone); }
    });
}

int mf_server_release(mf_server *s, long long id)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        std::lock_guard<std::mutex> lk(s->mu);
        Request *r = findRequest(s, id);
        if (r->state == ReqState::Pending || r->state == ReqState::Launched) {
            fail(MF_ERR_BAD_STATE, "request " + std::to_string(id) + " is still running");
        }
        if (r->allocated) { arenaFree(s, r->off, r->bytes); }
        s->reqs.erase(id);
        ++s->memEpoch;
        s->cvSched.notify_one();
    });
}

int mf_server_cancel(mf_server *s, long long id, int *cancelled)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        std::lock_guard<std::mutex> lk(s->mu);
        Request *r = findRequest(s, id);
        const bool pending = r->state == ReqState::Pending;
        if (pending) {
            s->pending.erase(std::find(s->pending.begin(), s->pending.end(), r));
            r->state = ReqState::Failed;
            r->status = MF_ERR_CANCELLED;
            r->tDone = Clock::now();
            ++s->memEpoch;  // the queue head may have changed: admission retries
            s->cvDone.notify_all();
            s->cvSched.notify_one();
        }
        if (cancelled) { *cancelled = pending ? 1 : 0; }
    });
}

int mf_server_window_stats(mf_server *s, long long *windows, long long *flagged, long long *full, long long *solo)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        unsigned long long nf[2] = {0, 0};
        MF_HIP(hipMemcpy(nf, s->dFlagged.p, sizeof(nf), hipMemcpyDeviceToHost));
        if (windows) { *windows = s->statWindows.load(); }
        if (flagged) { *flagged = static_cast<long long>(nf[0]); }
        if (full) { *full = static_cast<long long>(nf[1]); }
        if (solo) { *solo = s->statSolo.load(); }
    });
}

int mf_debug_fail_allocations(int n, int *remaining)
{
    return guard([&] {
        if (n < 0) { fail(MF_ERR_INVALID_ARGUMENT, "n must be non-negative"); }
        const int before = injectedAllocFailures().exchange(n);
        if (remaining) { *remaining = before; }
    });
}

int mf_server_stats(mf_server *s, long long *batches, long long *requests, double *corrSamples)
{
    return guard([&] {
        if (s == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "server is NULL"); }
        std::lock_guard<std::mutex> lk(s->mu);
        if (batches) { *batches = s->nBatches; }
        if (requests) { *requests = s->nDone; }
        if (corrSamples) { *corrSamples = s->doneSamples; }
    });
}

}  // extern "C"
