// Host side of the C ABI declared in include/mflib_gpu.h.
//
// A handle owns: the template rows (FP64 host copies), the channel layout,
// one HIP stream, and after mf_initialize the device state:
//   channels   [C][n_c]            samples in the handle precision
//   templates  [rows][Lt]          FP64 demeaned unit-norm templates
//   spectra    [rows][nfft/2+1]    conj(T^)/(2 nfft), handle precision
//   rho        rows back to back   each row padded to a 64-sample multiple
// plus the small descriptor tables the kernels read (lanes, rows, work
// items) and the two-level twiddle tables of every FFT size in use.
//
// Work layout: a lane is (channel, template-length bucket, FFT size); its
// blocks are the work items.  Items are ordered so that all blocks of a
// lane land on one XCD
// (workgroup i runs on XCD i mod 8), which keeps that channel's template
// spectra resident in that XCD's L2.
#include "mflib_gpu.h"
#include "mf_host.hpp"
#include "mf_kernels.hpp"
#include "stream_peaks.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

using namespace mfh;

// blocks of the exact-recomputation launches after the correlation
// launches (grid-stride over the queued windows: 8 blocks per CU)
constexpr int kFixBlocks = 2048;

struct mf_handle {
    int precision = MF_PRECISION_FLOAT;
    int mode = MF_MODE_MULTI_CHANNEL;
    int device = 0;  // the HIP device current at mf_create; all buffers live there
    size_t tsize = 4;
    int signalSize = 0;
    int absval = 1;
    int userNfft = 0;
    int userVariant = 0;  // 0: best available correlation kernel
    bool useGraph = false;            // replay apply's launches as one HIP graph
    // streaming: every channel keeps its last `carry` samples in front of the
    // newest chunk, so consecutive chunks give continuous rows and stacks
    bool streaming = false;
    long long carry = 0;
    // streaming PeakFinder (mf_stream_peaks): one state per stacking group
    long long applies = 0;  // applies since initialize
    struct StreamPeakState {
        bool active = false;
        long long origin = 0, lastApply = 0;  // apply count of the stream's first and latest push
        double threshold = 0;
        int minDist = 0;
        std::unique_ptr<mfstream::StreamPeaks<float>> f;
        std::unique_ptr<mfstream::StreamPeaks<double>> d;
    };
    std::vector<StreamPeakState> streamPeaks;
    DevBuf dStreamTab, dStreamIdx, dStreamEdge;  // mf_stream_peaks_all: segment table, edge gather
    DevBuf dPushScratch;  // mf_push_chunks_device: every channel's carry while its buffer is rewritten
    DevBuf dCarry;
    DevBuf dDeadPart;  // streaming: per-chunk min/max of the dead-trace pass on the device
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    // templates
    std::vector<std::vector<double>> tmpl;
    bool haveShifts = false;
    std::vector<int> shifts;
    std::vector<double> weights;
    // batched layout
    bool haveLayout = false;
    int nChannels = 0;
    int nGroupsIn = 0;
    std::vector<int> chLenIn, rowChIn, rowGroupIn;
    // derived layout
    struct Channel {
        DevBuf buf;
        long long n = 0;
        bool dead = true;
        bool set = false;
    };
    struct RowP {
        int ch = 0, group = 0, Lt = 0;
        int nOut = 0;
        int lane = 0;
        int logM = 0;
        size_t specOff = 0;  // elements of complex T
        size_t rhoOff = 0;   // elements of T
        size_t tmplOff = 0;  // doubles
        float hmax = 0;
    };
    struct Lane {
        int ch = 0, logM = 0, L = 0, nOut = 0;
        int td = 0, KT = 0;  // time-domain lane: kernel kind (3/4) and padded taps
        RowMix mix;
        std::vector<int> rows;  // FFT lanes: by decreasing nOut, then Lt; time-domain lanes: by Lt
    };
    struct Group {
        std::vector<int> rows;
        int nb = 0;
        int nOut = 0;
    };
    std::vector<Channel> ch;
    std::vector<RowP> rows;
    std::vector<Lane> lanes;
    std::vector<Group> groups;
    size_t ldOut = 0;
    bool initialized = false;
    bool applied = false;
    bool tablesDirty = true;
    hipStream_t stream = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr, evC = nullptr, evD = nullptr;
    double corrMs = 0, stackMs = 0;
    // device state
    DevBuf dRho, dSpec, dTmpl, dTmplRaw, dHmax;
    DevBuf dSegs, dRowTab, dItems;
    // window statistics (mf_get_window_stats): the device's running counts
    // of windows recomputed exactly (all, in full), and per apply the
    // windows computed and the windows of solo rows (both fixed by the plan)
    DevBuf dFlagged;
    long long winPerApply = 0, soloPerApply = 0, statApplies = 0;
    DevBuf dFrag, dTdRowTab, dTdItems;  // time-domain lanes
    DevBuf dTdFix;                      //   their queue of flagged windows (count, then entries)
    mfk::FixQueue tdFix{};
    DevBuf dFftFix;                     // queue of flagged windows of the FP32 wave-kernel rows
    mfk::FixQueue fftFix{};
    DevBuf dStackRows, dStackOut;
    DevBuf dReduceRows, dRedDet, dRedId, dRedPk, dRedPkId;  // Detector reduce step
    DevBuf dProdJobs, dProdOut;                             // per-detection products
    PinnedBuf hProd;                                        // their host staging (jobs, then results)
    // the Detector's per-template peak lists (rows taken in chunks)
    DevBuf dDetGroupTab, dDetCounts, dDetCandIdx, dDetCandVal, dDetPackIdx, dDetPackVal;
    PinnedBuf hDetPack;
    // stacking groups (built on first stack/detect)
    struct StackInfo {
        int first = 0, count = 0, smax = 0, minShift = 0;
        long long len = 0;
    };
    bool stacksReady = false;
    std::vector<StackInfo> stackInfo;
    size_t ldStack = 0;
    size_t candCap = 0;
    DevBuf dStacks, dCounts, dCandIdx, dCandVal, dGroupTab;
    DevBuf dPackIdx, dPackVal;  // candidates of all groups, packed for one readback
    std::vector<mfk::StackGroup> hGroupTab;  // host copy of dGroupTab
    PinnedBuf hPack;            // host side of that readback
    std::map<int, TwTables> tw;
    struct Launch {
        int logM, first, count, variant;
        int KT = 0;  // time-domain launches (variant 3/4): padded taps; items in dTdItems
    };
    std::vector<Launch> launches;
    std::vector<mfk::Seg> segHost;
    std::vector<int> rowTabIndex;  // row table position of each row

    void dropGraph()
    {
        if (graphExec) { (void)hipGraphExecDestroy(graphExec); }
        if (graph) { (void)hipGraphDestroy(graph); }
        graphExec = nullptr;
        graph = nullptr;
    }

    ~mf_handle()
    {
        dropGraph();
        if (evA) { (void)hipEventDestroy(evA); }
        if (evB) { (void)hipEventDestroy(evB); }
        if (evC) { (void)hipEventDestroy(evC); }
        if (evD) { (void)hipEventDestroy(evD); }
        if (stream) { (void)hipStreamSynchronize(stream); (void)hipStreamDestroy(stream); }
    }
};

namespace {

void requireHandle(const mf_handle *h)
{
    if (h == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "handle is NULL"); }
}

void requireInit(const mf_handle *h)
{
    requireHandle(h);
    if (!h->initialized) { fail(MF_ERR_BAD_STATE, "Class is not initialized"); }
}

void dropDevice(mf_handle *h)
{
    if (h->stream) { (void)hipStreamSynchronize(h->stream); }
    h->dropGraph();
    h->ch.clear();
    h->rows.clear();
    h->lanes.clear();
    h->groups.clear();
    h->dRho.release();
    h->dSpec.release();
    h->dTmpl.release();
    h->dTmplRaw.release();
    h->dHmax.release();
    h->dSegs.release();
    h->dRowTab.release();
    h->dItems.release();
    h->dFrag.release();
    h->dTdRowTab.release();
    h->dTdFix.release();
    h->tdFix = mfk::FixQueue{};
    h->dFftFix.release();
    h->fftFix = mfk::FixQueue{};
    h->dTdItems.release();
    h->dStackRows.release();
    h->stacksReady = false;
    h->stackInfo.clear();
    h->dStacks.release();
    h->dCounts.release();
    h->dCandIdx.release();
    h->dCandVal.release();
    h->dPackIdx.release();
    h->dPackVal.release();
    h->dGroupTab.release();
    h->dStackOut.release();
    h->dCarry.release();
    h->dPushScratch.release();
    h->dDeadPart.release();
    h->carry = 0;
    h->applies = 0;
    h->streamPeaks.clear();
    h->tw.clear();
    h->launches.clear();
    h->segHost.clear();
    h->rowTabIndex.clear();
    h->initialized = false;
    h->applied = false;
}

void deriveLayout(mf_handle *h, std::vector<int> &rowCh, std::vector<int> &rowGroup,
                  std::vector<long long> &chLen, int &nGroups)
{
    const int nT = static_cast<int>(h->tmpl.size());
    if (h->mode == MF_MODE_BATCHED) {
        if (!h->haveLayout) { fail(MF_ERR_INVALID_ARGUMENT, "batch layout not set"); }
        if (static_cast<int>(h->rowChIn.size()) != nT) {
            fail(MF_ERR_INVALID_ARGUMENT, "batch layout does not match the number of templates");
        }
        rowCh = h->rowChIn;
        rowGroup = h->rowGroupIn;
        chLen.assign(h->chLenIn.begin(), h->chLenIn.end());
        nGroups = h->nGroupsIn;
        return;
    }
    if (h->signalSize < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Parameters are invalid: signal size not set"); }
    rowGroup.assign(nT, 0);
    nGroups = 1;
    if (h->mode == MF_MODE_MULTI_CHANNEL) {
        rowCh.resize(nT);
        std::iota(rowCh.begin(), rowCh.end(), 0);
        chLen.assign(nT, h->signalSize);
    } else {
        rowCh.assign(nT, 0);
        chLen.assign(1, h->signalSize);
    }
}

// Windows of the launches whose row computes its window scale on its own
// wave (corr2_kernel's solo rows); the block-wide kernel computes every
// scale for the whole block: no solo rows.
long long soloWindows(const mf_handle *h, const std::vector<mfk::Item> &items, const std::vector<mfk::Row> &rowTab)
{
    long long solo = 0;
    for (const auto &la : h->launches) {
        if (la.KT > 0 || la.variant != 2) { continue; }
        solo += mfh::soloWindows(items.data(), la.first, la.count, h->segHost.data(), rowTab.data(),
                                 mfk::corr2_shared_scale(static_cast<int>(h->tsize), 1 << la.logM));
    }
    return solo;
}

void initialize(mf_handle *h)
{
    dropDevice(h);
    const int nT = static_cast<int>(h->tmpl.size());
    if (nT < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Parameters are invalid: no templates"); }
    std::vector<int> rowCh, rowGroup;
    std::vector<long long> chLen;
    int nGroups = 0;
    deriveLayout(h, rowCh, rowGroup, chLen, nGroups);
    const int nC = static_cast<int>(chLen.size());
    h->carry = 0;
    if (h->streaming) {
        // nb - 1 samples for the windows ending in the chunk, plus the
        // largest moveout for the stacks of the chunk's lags
        int nbMax = 0, sMax = 0;
        for (int r = 0; r < nT; ++r) { nbMax = std::max(nbMax, static_cast<int>(h->tmpl[r].size())); }
        if (h->haveShifts) { sMax = *std::max_element(h->shifts.begin(), h->shifts.end()); }
        h->carry = static_cast<long long>(nbMax - 1) + std::max(sMax, 0);
        for (auto &n : chLen) {
            n += h->carry;
            // the device dead-trace pass covers 65535 chunks of kDeadChunk samples
            if ((n + mfk::kDeadChunk - 1) / mfk::kDeadChunk > 65535) {
                fail(MF_ERR_UNSUPPORTED, "streaming chunk of " + std::to_string(n) + " samples is too long");
            }
        }
        if (h->carry > 0) { h->dCarry.alloc(static_cast<size_t>(h->carry) * h->tsize); }
    }

    // groups: nb = max template length, nOut = min channel length - nb + 1
    h->groups.assign(nGroups, {});
    h->rows.assign(nT, {});
    for (int r = 0; r < nT; ++r) {
        auto &rp = h->rows[r];
        rp.ch = rowCh[r];
        rp.group = rowGroup[r];
        rp.Lt = static_cast<int>(h->tmpl[r].size());
        h->groups[rp.group].rows.push_back(r);
        h->groups[rp.group].nb = std::max(h->groups[rp.group].nb, rp.Lt);
    }
    long long maxOut = 0;
    for (int g = 0; g < nGroups; ++g) {
        auto &gr = h->groups[g];
        if (gr.rows.empty()) { fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(g) + " has no templates"); }
        long long nmin = -1;
        for (int r : gr.rows) { nmin = nmin < 0 ? chLen[rowCh[r]] : std::min(nmin, chLen[rowCh[r]]); }
        // computeOptimalFFTAndBlockLength(nb, n + nb - 1) rejects n <= 1
        if (nmin + gr.nb - 1 <= gr.nb) {
            fail(MF_ERR_INVALID_ARGUMENT, "Signal length = " + std::to_string(nmin + gr.nb - 1) +
                                              " must exceed filter length = " + std::to_string(gr.nb));
        }
        if (nmin < gr.nb) {
            fail(MF_ERR_INVALID_ARGUMENT, "signal length " + std::to_string(nmin) +
                                              " is shorter than the template length " +
                                              std::to_string(gr.nb));
        }
        gr.nOut = static_cast<int>(nmin - gr.nb + 1);
        maxOut = std::max<long long>(maxOut, gr.nOut);
        for (int r : gr.rows) { h->rows[r].nOut = gr.nOut; }
    }
    // lanes: one per (channel, template-length bucket) carrying rows
    std::vector<std::vector<int>> chRows(nC);
    for (int r = 0; r < nT; ++r) { chRows[rowCh[r]].push_back(r); }
    for (int c = 0; c < nC; ++c) {
        if (chRows[c].empty()) { continue; }
        std::vector<int> sorted = chRows[c];
        std::stable_sort(sorted.begin(), sorted.end(),
                         [&](int a, int b) { return h->rows[a].Lt < h->rows[b].Lt; });
        std::vector<int> lts;
        std::vector<long long> outs;
        for (int r : sorted) {
            lts.push_back(h->rows[r].Lt);
            outs.push_back(h->rows[r].nOut);
        }
        // time-domain path (variant 3/4, FP32): the rows of at most kTdMaxTaps
        // taps form one lane; longer rows keep FFT lanes
        size_t nTd = 0;
        if ((h->userVariant == mfk::kTdFp16 || h->userVariant == mfk::kTdBf16) &&
            h->precision == MF_PRECISION_FLOAT) {
            while (nTd < lts.size() && mfk::td_supported(lts[nTd])) { ++nTd; }
        }
        if (nTd > 0) {
            mf_handle::Lane ln;
            ln.ch = c;
            ln.td = h->userVariant;
            ln.KT = mfk::td_taps(lts[nTd - 1]);
            ln.logM = mfk::kMinLogM;  // template spectra for mf_template_spectrum only
            ln.L = mfk::kTdOutputsPerItem;
            ln.mix = rowMix(std::vector<int>(lts.begin(), lts.begin() + nTd));
            ln.rows.assign(sorted.begin(), sorted.begin() + nTd);
            for (int r : ln.rows) {
                ln.nOut = std::max(ln.nOut, h->rows[r].nOut);
                h->rows[r].lane = static_cast<int>(h->lanes.size());
                h->rows[r].logM = ln.logM;
            }
            h->lanes.push_back(std::move(ln));
            sorted.erase(sorted.begin(), sorted.begin() + nTd);
            lts.erase(lts.begin(), lts.begin() + nTd);
            outs.erase(outs.begin(), outs.begin() + nTd);
            if (sorted.empty()) { continue; }
        }
        for (const Bucket &bk : planBuckets(h->precision, h->userNfft == 0 ? 0 : h->userNfft, lts, outs)) {
            mf_handle::Lane ln;
            ln.ch = c;
            ln.mix = bk.mix;
            ln.logM = bk.logM;
            ln.L = blockLength(ln.logM, bk.nb);
            ln.rows.assign(sorted.begin() + bk.begin, sorted.begin() + bk.end);
            // decreasing nOut (then increasing length): the rows with
            // outputs in a block are a prefix, which the block's item names
            std::stable_sort(ln.rows.begin(), ln.rows.end(), [&](int a, int b) {
                if (h->rows[a].nOut != h->rows[b].nOut) { return h->rows[a].nOut > h->rows[b].nOut; }
                return h->rows[a].Lt < h->rows[b].Lt;
            });
            for (int r : ln.rows) {
                ln.nOut = std::max(ln.nOut, h->rows[r].nOut);
                h->rows[r].lane = static_cast<int>(h->lanes.size());
                h->rows[r].logM = ln.logM;
            }
            h->lanes.push_back(std::move(ln));
        }
    }

    // device buffers
    h->ldOut = roundUp(static_cast<size_t>(maxOut), 64);
    size_t rhoTotal = 0;
    for (int r = 0; r < nT; ++r) {
        h->rows[r].rhoOff = rhoTotal;
        rhoTotal += roundUp(static_cast<size_t>(std::max(h->rows[r].nOut, 1)), 64);
    }
    h->dRho.alloc(rhoTotal * h->tsize);
    size_t specTotal = 0, tmplTotal = 0;
    for (int r = 0; r < nT; ++r) {
        auto &rp = h->rows[r];
        rp.specOff = specTotal;
        specTotal += roundUp((1u << rp.logM) + 1, 32);
        rp.tmplOff = tmplTotal;
        tmplTotal += roundUp(rp.Lt, 32);
    }
    h->dSpec.alloc(specTotal * 2 * h->tsize);
    h->dTmpl.alloc(tmplTotal * sizeof(double));
    h->dTmplRaw.alloc(tmplTotal * sizeof(double));
    h->dHmax.alloc(nT * sizeof(float));
    h->ch.clear();
    h->ch.resize(nC);
    for (int c = 0; c < nC; ++c) {
        h->ch[c].n = chLen[c];
        h->ch[c].buf.alloc(static_cast<size_t>(chLen[c]) * h->tsize);
        MF_HIP(hipMemsetAsync(h->ch[c].buf.p, 0, h->ch[c].buf.bytes, h->stream));
        h->ch[c].dead = true;
        h->ch[c].set = false;
    }
    // twiddle tables
    for (const auto &ln : h->lanes) {
        if (h->tw.count(ln.logM)) { continue; }
        makeTwiddles(h->tw[ln.logM], ln.logM, h->precision == MF_PRECISION_FLOAT);
    }
    // templates: upload raw FP64, prepare + transform on the GPU
    {
        std::vector<double> raw(tmplTotal, 0.0);
        for (int r = 0; r < nT; ++r) {
            std::copy(h->tmpl[r].begin(), h->tmpl[r].end(), raw.begin() + h->rows[r].tmplOff);
        }
        MF_HIP(hipMemcpy(h->dTmplRaw.p, raw.data(), raw.size() * sizeof(double), hipMemcpyHostToDevice));
        std::map<int, std::vector<mfk::TmplJob>> jobs;
        for (int r = 0; r < nT; ++r) {
            const auto &rp = h->rows[r];
            mfk::TmplJob j;
            j.t = h->dTmplRaw.as<double>() + rp.tmplOff;
            j.Lt = rp.Lt;
            j.tOut = h->dTmpl.as<double>() + rp.tmplOff;
            j.H = static_cast<char *>(h->dSpec.p) + rp.specOff * 2 * h->tsize;
            j.hmax = h->dHmax.as<float>() + r;
            jobs[rp.logM].push_back(j);
        }
        for (auto &kv : jobs) {
            DevBuf dj;
            dj.alloc(kv.second.size() * sizeof(mfk::TmplJob));
            MF_HIP(hipMemcpy(dj.p, kv.second.data(), dj.bytes, hipMemcpyHostToDevice));
            const TwTables &t = h->tw.at(kv.first);
            MF_HIP(mfk::launch_tmpl(h->precision, kv.first, dj.as<mfk::TmplJob>(),
                                    static_cast<int>(kv.second.size()), t.lo64.p, t.hi64.p, h->stream));
            MF_HIP(hipStreamSynchronize(h->stream));
        }
        std::vector<float> hm(nT);
        MF_HIP(hipMemcpy(hm.data(), h->dHmax.p, nT * sizeof(float), hipMemcpyDeviceToHost));
        for (int r = 0; r < nT; ++r) { h->rows[r].hmax = hm[r]; }
    }
    // row table (lane order), work items (XCD-aware order)
    std::vector<mfk::Row> rowTab;
    h->rowTabIndex.assign(nT, 0);
    h->segHost.clear();
    h->dFlagged.alloc(2 * sizeof(unsigned long long));
    MF_HIP(hipMemset(h->dFlagged.p, 0, 2 * sizeof(unsigned long long)));
    h->statApplies = 0;
    h->winPerApply = 0;
    for (int r = 0; r < nT; ++r) { h->winPerApply += h->rows[r].nOut; }
    for (size_t li = 0; li < h->lanes.size(); ++li) {
        const auto &ln = h->lanes[li];
        mfk::Seg s{};
        s.flagged = h->dFlagged.as<unsigned long long>();
        s.n = h->ch[ln.ch].n;
        s.L = ln.L;
        s.nOut = ln.nOut;
        s.rowBegin = static_cast<int>(rowTab.size());
        float hmax = 0;
        if (ln.td) {
            // time-domain lane: its rows go to the TD row table (built below)
            s.rowEnd = s.rowBegin;
            s.hmax = 0;
            s.x = h->ch[ln.ch].buf.p;
            s.dead = 1;
            h->segHost.push_back(s);
            continue;
        }
        for (int r : ln.rows) {
            const auto &rp = h->rows[r];
            mfk::Row row;
            row.H = static_cast<char *>(h->dSpec.p) + rp.specOff * 2 * h->tsize;
            row.tmpl = h->dTmpl.as<double>() + rp.tmplOff;
            row.out = static_cast<char *>(h->dRho.p) + rp.rhoOff * h->tsize;
            row.nOut = rp.nOut;
            row.Lt = rp.Lt;
            row.x = h->ch[ln.ch].buf.p;
            h->rowTabIndex[r] = static_cast<int>(rowTab.size());
            rowTab.push_back(row);
            hmax = std::max(hmax, rp.hmax);
        }
        s.rowEnd = static_cast<int>(rowTab.size());
        s.hmax = hmax;
        s.x = h->ch[ln.ch].buf.p;
        s.dead = 1;
        h->segHost.push_back(s);
    }
    h->dRowTab.alloc(std::max<size_t>(rowTab.size(), 1) * sizeof(mfk::Row));
    if (!rowTab.empty()) {
        MF_HIP(hipMemcpy(h->dRowTab.p, rowTab.data(), rowTab.size() * sizeof(mfk::Row), hipMemcpyHostToDevice));
    }
    // time-domain lanes: split taps, row table, work items
    std::vector<mfk::TdRow> tdRows;
    std::vector<mfk::TdItem> tdItems;
    long long tdWindows = 0;
    std::map<std::pair<int, int>, std::pair<int, int>> tdLaunch;  // (kind, KT) -> (first, count)
    {
        size_t fragBytes = 0;
        for (const auto &ln : h->lanes) {
            if (ln.td) { fragBytes += ln.rows.size() * mfk::td_frag_bytes(ln.td, ln.KT); }
        }
        if (fragBytes > 0) {
            h->dFrag.alloc(fragBytes);
            std::map<std::pair<int, int>, std::vector<mfk::TdPrepJob>> jobs;
            std::map<std::pair<int, int>, std::vector<mfk::TdItem>> itemsByKind;
            size_t off = 0;
            for (size_t li = 0; li < h->lanes.size(); ++li) {
                const auto &ln = h->lanes[li];
                if (!ln.td) { continue; }
                const int rowBegin = static_cast<int>(tdRows.size());
                for (int r : ln.rows) {
                    const auto &rp = h->rows[r];
                    mfk::TdRow tr{};
                    tr.frag = static_cast<char *>(h->dFrag.p) + off;
                    tr.tmpl = h->dTmpl.as<double>() + rp.tmplOff;
                    tr.out = static_cast<char *>(h->dRho.p) + rp.rhoOff * h->tsize;
                    tr.nOut = rp.nOut;
                    tr.Lt = rp.Lt;
                    const double eq = mfk::td_eps(ln.td, ln.KT) / mfk::kTdRhoTol;
                    tr.sqrtLt = static_cast<float>(std::sqrt(static_cast<double>(rp.Lt)));
                    tr.kc = static_cast<float>(2.0e-9 * rp.Lt * rp.Lt);
                    tr.kn = static_cast<float>(eq * eq * rp.Lt);
                    tr.x = static_cast<const float *>(h->ch[ln.ch].buf.p);
                    tdRows.push_back(tr);
                    tdWindows += rp.nOut;
                    jobs[{ln.td, ln.KT}].push_back({tr.tmpl, rp.Lt, 0, const_cast<void *>(tr.frag)});
                    off += mfk::td_frag_bytes(ln.td, ln.KT);
                }
                const int nRows = static_cast<int>(ln.rows.size());
                for (long long s0 = 0; s0 < ln.nOut; s0 += mfk::kTdOutputsPerItem) {
                    for (int g = 0; g < nRows; g += 128) {
                        mfk::TdItem it{};
                        it.seg = static_cast<int>(li);
                        it.rowBegin = rowBegin + g;
                        it.seg0 = s0;
                        it.nRows = std::min(128, nRows - g);
                        itemsByKind[{ln.td, ln.KT}].push_back(it);
                    }
                }
            }
            for (auto &kv : jobs) {
                DevBuf dj;
                dj.alloc(kv.second.size() * sizeof(mfk::TdPrepJob));
                MF_HIP(hipMemcpy(dj.p, kv.second.data(), dj.bytes, hipMemcpyHostToDevice));
                MF_HIP(mfk::launch_td_prep(kv.first.first, dj.as<mfk::TdPrepJob>(), static_cast<int>(kv.second.size()),
                                           kv.first.second, h->stream));
                MF_HIP(hipStreamSynchronize(h->stream));
            }
            for (auto &kv : itemsByKind) {
                tdLaunch[kv.first] = {static_cast<int>(tdItems.size()), static_cast<int>(kv.second.size())};
                tdItems.insert(tdItems.end(), kv.second.begin(), kv.second.end());
            }
        }
    }
    h->dTdRowTab.alloc(std::max<size_t>(tdRows.size(), 1) * sizeof(mfk::TdRow));
    if (!tdRows.empty()) {
        // room for 1 in 32 windows flagged (more are recomputed in place)
        const long long cap = std::min<long long>(std::max<long long>(tdWindows / 32, 1 << 16), 1 << 25);
        h->dTdFix.alloc(256 + static_cast<size_t>(cap) * 8);
        h->tdFix.count = h->dTdFix.as<unsigned long long>();
        h->tdFix.entries = reinterpret_cast<int *>(static_cast<char *>(h->dTdFix.p) + 256);
        h->tdFix.cap = static_cast<unsigned long long>(cap);
    }
    h->dTdItems.alloc(std::max<size_t>(tdItems.size(), 1) * sizeof(mfk::TdItem));
    if (!tdRows.empty()) {
        MF_HIP(hipMemcpy(h->dTdRowTab.p, tdRows.data(), tdRows.size() * sizeof(mfk::TdRow), hipMemcpyHostToDevice));
        MF_HIP(hipMemcpy(h->dTdItems.p, tdItems.data(), tdItems.size() * sizeof(mfk::TdItem), hipMemcpyHostToDevice));
    }
    h->dSegs.alloc(h->segHost.size() * sizeof(mfk::Seg));
    std::vector<mfk::Item> items;
    // one launch per (FFT size, kernel variant)
    std::map<std::pair<int, int>, std::vector<int>> lanesByPlan;
    for (size_t li = 0; li < h->lanes.size(); ++li) {
        const auto &ln = h->lanes[li];
        if (ln.td) { continue; }
        lanesByPlan[{ln.logM, laneVariant(h->precision, ln.logM, h->userVariant, ln.mix)}].push_back(
            static_cast<int>(li));
    }
    for (auto &kv : lanesByPlan) {
        std::vector<long long> blocks;
        for (int li : kv.second) {
            const auto &ln = h->lanes[li];
            blocks.push_back((ln.nOut + ln.L - 1) / ln.L);
        }
        mf_handle::Launch la;
        la.logM = kv.first.first;
        la.variant = kv.first.second;
        la.first = static_cast<int>(items.size());
        const std::vector<int> &laneIds = kv.second;
        appendXcdItems(laneIds, blocks, items, [&](int i, long long k) {
            const auto &ln = h->lanes[laneIds[i]];
            const long long start = k * ln.L;
            int n = 0;
            while (n < static_cast<int>(ln.rows.size()) && h->rows[ln.rows[n]].nOut > start) { ++n; }
            return n;
        });
        la.count = static_cast<int>(items.size()) - la.first;
        h->launches.push_back(la);
    }
    h->soloPerApply = soloWindows(h, items, rowTab);
    // FP32 wave-kernel rows: flagged outputs are written as NaN and their
    // row blocks recorded (at most two records per row block: a block's two
    // halves at M = 2048) for one exact recomputation launch
    if (h->precision == MF_PRECISION_FLOAT) {
        long long rowBlocks = 0;
        for (const auto &la : h->launches) {
            if (la.KT > 0 || la.variant != 2) { continue; }
            for (int k = la.first; k < la.first + la.count; ++k) {
                const mfk::Item &it = items[k];
                const mfk::Seg &sg = h->segHost[it.seg];
                const int end = it.nRows < 0 ? sg.rowEnd : std::min(sg.rowEnd, sg.rowBegin + it.nRows);
                rowBlocks += end - sg.rowBegin;
            }
        }
        if (rowBlocks > 0) {
            h->dFftFix.alloc(256 + static_cast<size_t>(2 * rowBlocks) * 16);
            h->fftFix.overCount = h->dFftFix.as<unsigned long long>();
            h->fftFix.over = reinterpret_cast<int *>(static_cast<char *>(h->dFftFix.p) + 256);
        }
    }
    for (const auto &kv : tdLaunch) {
        mf_handle::Launch la;
        la.logM = 0;
        la.variant = kv.first.first;
        la.KT = kv.first.second;
        la.first = kv.second.first;
        la.count = kv.second.second;
        h->launches.push_back(la);
    }
    h->dItems.alloc(std::max<size_t>(items.size(), 1) * sizeof(mfk::Item));
    if (!items.empty()) {
        MF_HIP(hipMemcpy(h->dItems.p, items.data(), items.size() * sizeof(mfk::Item), hipMemcpyHostToDevice));
    }
    h->dStackOut.alloc(std::max<size_t>(h->ldOut, 64) * h->tsize);
    MF_HIP(hipStreamSynchronize(h->stream));
    h->tablesDirty = true;
    h->initialized = true;
    h->applied = false;
}

template <typename T>
bool isDead(const T *x, long long n)
{
    auto mm = std::minmax_element(x, x + n);
    return std::abs(*mm.first) == std::abs(*mm.second);
}

void checkChannel(const mf_handle *h, int c)
{
    const int nC = static_cast<int>(h->ch.size());
    if (c < 0 || c >= nC) {
        fail(MF_ERR_INVALID_ARGUMENT, "it = " + std::to_string(c) + " must be in the range [0," +
                                          std::to_string(nC - 1) + "]");
    }
}

void checkRow(const mf_handle *h, int r)
{
    const int nT = static_cast<int>(h->rows.size());
    if (r < 0 || r >= nT) {
        fail(MF_ERR_INVALID_ARGUMENT, "it = " + std::to_string(r) + " must be in the range [0," +
                                          std::to_string(nT - 1) + "]");
    }
}

void checkSignalLength(const mf_handle *h, int c, int n)
{
    if (n != h->ch[c].n) {
        fail(MF_ERR_INVALID_ARGUMENT, "Number of samples = " + std::to_string(n) + " must equal " +
                                          std::to_string(h->ch[c].n));
    }
}

void apply(mf_handle *h)
{
    requireInit(h);
    h->applied = false;
    if (h->tablesDirty) {
        for (size_t li = 0; li < h->lanes.size(); ++li) {
            h->segHost[li].dead = h->ch[h->lanes[li].ch].dead ? 1 : 0;
        }
        MF_HIP(hipMemcpyAsync(h->dSegs.p, h->segHost.data(), h->dSegs.bytes, hipMemcpyHostToDevice,
                              h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        h->tablesDirty = false;
    }
    auto launchAll = [&]() {
        bool td = false;
        if (h->fftFix.over != nullptr) {
            MF_HIP(mfk::launch_fill_int(reinterpret_cast<int *>(h->fftFix.overCount), 2, 0, h->stream));
        }
        for (const auto &la : h->launches) {
            if (la.KT > 0) {
                if (!td) {
                    // the queue of flagged windows starts empty (a fill
                    // kernel: no memset node in a captured graph)
                    MF_HIP(mfk::launch_fill_int(reinterpret_cast<int *>(h->tdFix.count), 2, 0, h->stream));
                    td = true;
                }
                MF_HIP(mfk::launch_td(la.variant, la.KT, h->dTdItems.as<mfk::TdItem>() + la.first, la.count,
                                      h->dSegs.as<mfk::Seg>(), h->dTdRowTab.as<mfk::TdRow>(), h->tdFix, h->stream));
                continue;
            }
            const TwTables &t = h->tw.at(la.logM);
            const void *lo = h->precision == MF_PRECISION_FLOAT ? t.lo.p : t.lo64.p;
            const void *hi = h->precision == MF_PRECISION_FLOAT ? t.hi.p : t.hi64.p;
            MF_HIP(mfk::launch_corr(h->precision, la.logM, la.variant, h->dItems.as<mfk::Item>() + la.first,
                                    la.count, h->dSegs.as<mfk::Seg>(), h->dRowTab.as<mfk::Row>(), lo, hi,
                                    h->fftFix, h->stream));
        }
        if (td) {
            MF_HIP(mfk::launch_fix(h->dTdRowTab.as<mfk::TdRow>(), h->tdFix, kFixBlocks, h->stream));
        }
        MF_HIP(mfk::launch_fix(h->dRowTab.as<mfk::Row>(), h->fftFix, kFixBlocks, h->stream));
    };
    if (h->useGraph && h->graphExec == nullptr && !h->launches.empty()) {
        // capture the correlation launches once; their descriptor tables and
        // buffers stay in place until the next initialize.  The graph holds
        // kernel nodes only (no memset nodes).
        MF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        try {
            launchAll();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(h->stream, &g);
            if (g) { (void)hipGraphDestroy(g); }
            throw;
        }
        MF_HIP(hipStreamEndCapture(h->stream, &h->graph));
        MF_HIP(hipGraphInstantiate(&h->graphExec, h->graph, nullptr, nullptr, 0));
    }
    if (h->streaming && !h->segHost.empty()) {
        // streaming pushes leave the dead-trace rule to the device: every
        // segment's flag from its channel's carry + chunk, before the launches
        long long maxN = 0;
        for (const auto &c : h->ch) { maxN = std::max(maxN, c.n); }
        const size_t chunks = static_cast<size_t>((maxN + mfk::kDeadChunk - 1) / mfk::kDeadChunk);
        const size_t partBytes = 2 * sizeof(double) * h->segHost.size() * chunks;
        if (h->dDeadPart.bytes < partBytes) { h->dDeadPart.alloc(partBytes); }
        MF_HIP(mfk::launch_seg_dead(h->precision, h->dSegs.as<mfk::Seg>(), static_cast<int>(h->segHost.size()), maxN,
                                    h->dDeadPart.as<double>(), h->stream));
    }
    MF_HIP(hipEventRecord(h->evA, h->stream));
    if (h->useGraph && h->graphExec != nullptr) {
        MF_HIP(hipGraphLaunch(h->graphExec, h->stream));
    } else {
        launchAll();
    }
    MF_HIP(hipEventRecord(h->evB, h->stream));
    h->applied = true;
    ++h->applies;
    ++h->statApplies;
}

// Stack weights exactly as shiftAndStack forms them
// (src/matchedFilter.cpp:1330-1343): xnorm = 1/sum(w) accumulated in T,
// weight = xnorm * w in T.
template <typename T>
void stackWeights(const mf_handle *h, const std::vector<int> &rows, std::vector<mfk::StackRow> &out,
                  int &smax)
{
    T xnorm = 0;
    smax = h->shifts[rows[0]];
    for (int r : rows) {
        smax = std::max(smax, h->shifts[r]);
        xnorm = xnorm + static_cast<T>(h->weights[r]);
    }
    if (xnorm > 0) { xnorm = 1 / xnorm; }
    for (int r : rows) {
        mfk::StackRow sr{};
        sr.rho = static_cast<const char *>(h->dRho.p) + h->rows[r].rhoOff * h->tsize;
        sr.shift = h->shifts[r];
        sr.weight = static_cast<double>(static_cast<T>(xnorm * static_cast<T>(h->weights[r])));
        out.push_back(sr);
    }
}

// Stack tables and per-group output/candidate buffers (built on first use
// after initialize; they depend only on the template rows).
void prepareStacks(mf_handle *h)
{
    if (h->stacksReady) { return; }
    const int nG = static_cast<int>(h->groups.size());
    std::vector<mfk::StackRow> all;
    h->stackInfo.assign(nG, {});
    size_t ldStack = 64;
    for (int g = 0; g < nG; ++g) {
        auto &gi = h->stackInfo[g];
        gi.first = static_cast<int>(all.size());
        int smax = 0;
        if (h->precision == MF_PRECISION_FLOAT) {
            stackWeights<float>(h, h->groups[g].rows, all, smax);
        } else {
            stackWeights<double>(h, h->groups[g].rows, all, smax);
        }
        gi.count = static_cast<int>(all.size()) - gi.first;
        gi.smax = smax;
        gi.minShift = smax;
        for (int r : h->groups[g].rows) { gi.minShift = std::min(gi.minShift, h->shifts[r]); }
        gi.len = static_cast<long long>(h->groups[g].nOut) - smax;
        ldStack = std::max(ldStack, roundUp(static_cast<size_t>(std::max<long long>(gi.len, 1)), 64));
    }
    h->ldStack = ldStack;
    h->dStackRows.alloc(std::max<size_t>(all.size(), 1) * sizeof(mfk::StackRow));
    MF_HIP(hipMemcpy(h->dStackRows.p, all.data(), all.size() * sizeof(mfk::StackRow), hipMemcpyHostToDevice));
    h->dStacks.alloc(ldStack * nG * h->tsize);
    h->dCounts.alloc(nG * sizeof(int));
    const size_t candCap = ldStack / 2 + 2;
    h->candCap = candCap;
    h->dCandIdx.alloc(candCap * nG * sizeof(int));
    h->dCandVal.alloc(candCap * nG * h->tsize);
    h->dPackIdx.alloc((candCap + 1) * nG * sizeof(int));
    h->dPackVal.alloc(candCap * nG * h->tsize);
    std::vector<mfk::StackGroup> tab(nG);
    for (int g = 0; g < nG; ++g) {
        const auto &gi = h->stackInfo[g];
        tab[g].rows = h->dStackRows.as<mfk::StackRow>() + gi.first;
        tab[g].nRows = gi.count;
        tab[g].candCap = static_cast<int>(candCap);
        tab[g].len = std::max<long long>(gi.len, 0);
        tab[g].out = static_cast<char *>(h->dStacks.p) + static_cast<size_t>(g) * ldStack * h->tsize;
        tab[g].count = h->dCounts.as<int>() + g;
        tab[g].candIdx = h->dCandIdx.as<int>() + static_cast<size_t>(g) * candCap;
        tab[g].candVal = static_cast<char *>(h->dCandVal.p) + static_cast<size_t>(g) * candCap * h->tsize;
    }
    h->dGroupTab.alloc(nG * sizeof(mfk::StackGroup));
    MF_HIP(hipMemcpy(h->dGroupTab.p, tab.data(), tab.size() * sizeof(mfk::StackGroup), hipMemcpyHostToDevice));
    h->hGroupTab = tab;
    h->stacksReady = true;
}

void checkStackable(mf_handle *h, int group)
{
    requireInit(h);
    if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtered signals not yet computed"); }
    if (!h->haveShifts) { fail(MF_ERR_BAD_STATE, "Templates missing info - can't shift"); }
    const int nG = static_cast<int>(h->groups.size());
    if (group < -1 || group >= nG) { fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " out of range"); }
    prepareStacks(h);
    for (int g = (group < 0 ? 0 : group); g < (group < 0 ? nG : group + 1); ++g) {
        const auto &gi = h->stackInfo[g];
        if (gi.minShift < 0) { fail(MF_ERR_INVALID_ARGUMENT, "negative moveout shifts are not supported"); }
        if (gi.len < 0) {
            fail(MF_ERR_INVALID_ARGUMENT, "maximum shift " + std::to_string(gi.smax) +
                                              " exceeds the filtered signal length " +
                                              std::to_string(h->groups[g].nOut));
        }
    }
}

// Stack of one group (group >= 0) or of every group (-1) into the
// handle's stack buffers; asynchronous on the handle's stream.
void stackGroups(mf_handle *h, int group)
{
    const int nG = static_cast<int>(h->groups.size());
    const int g0 = group < 0 ? 0 : group;
    const int ng = group < 0 ? nG : 1;
    long long maxLen = 0;
    for (int g = g0; g < g0 + ng; ++g) { maxLen = std::max(maxLen, h->stackInfo[g].len); }
    MF_HIP(hipEventRecord(h->evC, h->stream));
    MF_HIP(mfk::launch_stack_groups(h->precision, h->dGroupTab.as<mfk::StackGroup>() + g0, ng, maxLen,
                                    h->absval, h->stream));
    MF_HIP(hipEventRecord(h->evD, h->stream));
}

const void *stackPtr(const mf_handle *h, int g)
{
    return static_cast<const char *>(h->dStacks.p) + static_cast<size_t>(g) * h->ldStack * h->tsize;
}

void stack(mf_handle *h, int group, void *dOut, int cap, int *len)
{
    checkStackable(h, group);
    if (group < 0) { fail(MF_ERR_INVALID_ARGUMENT, "group must be non-negative"); }
    const long long n = h->stackInfo[group].len;
    if (len) { *len = static_cast<int>(n); }
    if (cap < n) {
        fail(MF_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(cap) + " < " + std::to_string(n) + " samples");
    }
    stackGroups(h, group);
    if (dOut != nullptr && n > 0) {
        MF_HIP(hipMemcpyAsync(dOut, stackPtr(h, group), static_cast<size_t>(n) * h->tsize,
                              hipMemcpyDeviceToDevice, h->stream));
    }
    MF_HIP(hipStreamSynchronize(h->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, h->evC, h->evD) == hipSuccess) { h->stackMs = ms; }
}

// Peak finding: candidates on the GPU, pruning on the host.
std::mutex g_peakMutex;

template <typename T>
void pruneCandidates(std::vector<std::pair<T, int>> &cand, int minDist, int *idx, int cap, int *nPeaks)
{
    std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a.second < b.second; });
    std::vector<int> keep;
    if (minDist > 0 && !cand.empty()) {
        // greedy in order of decreasing value (ties: ascending index): a
        // candidate is kept when no kept index lies within minDist of it
        // (src/peakFinder.cpp:536-574).  The kept indices live in an ordered
        // set, so each test is one lower_bound instead of a scan over every
        // kept index.
        std::stable_sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        std::set<long long> kept;
        for (const auto &p : cand) {
            const long long i = p.second;
            const auto it = kept.lower_bound(i - minDist);
            if (it == kept.end() || *it > i + minDist) { kept.insert(i); }
        }
        keep.assign(kept.begin(), kept.end());
    } else {
        for (const auto &p : cand) { keep.push_back(p.second); }
    }
    *nPeaks = static_cast<int>(keep.size());
    if (static_cast<int>(keep.size()) > cap) {
        fail(MF_ERR_INVALID_ARGUMENT, "maxx = " + std::to_string(cap) + " must be at least = " +
                                          std::to_string(keep.size()));
    }
    std::copy(keep.begin(), keep.end(), idx);
}

// Device workspace of the standalone peak finder, grown on demand and kept
// for the life of the process (guarded by g_peakMutex).
struct PeakWork {
    hipStream_t stream = nullptr;
    DevBuf count, idx, val, staged;
    PinnedBuf host;  // count and the first candidates, read back together
};
PeakWork &peakWork()
{
    static PeakWork *w = new PeakWork();  // never destroyed: outlives static teardown order
    return *w;
}

void growTo(DevBuf &b, size_t bytes)
{
    if (b.bytes < bytes) { b.alloc(bytes); }
}

// Device workspace of mf_amplitude_factors, kept for the life of the
// process (guarded by g_amplitudeMutex).
struct AmplitudeWork {
    hipStream_t stream = nullptr;
    DevBuf tmpl, windows, out;
};
std::mutex g_amplitudeMutex;
AmplitudeWork &amplitudeWork()
{
    static AmplitudeWork *w = new AmplitudeWork();  // never destroyed: outlives static teardown order
    return *w;
}

// Threshold + local-maximum candidates of a device signal (value, index;
// unordered), read back to the host.
template <typename T>
void rawCandidates(const void *dx, long long n, double threshold, std::vector<std::pair<T, int>> &cand,
                   hipStream_t stream, T *edges = nullptr)
{
    PeakWork &w = peakWork();
    const int candCap = static_cast<int>(std::min<long long>(n / 2 + 2, 0x7fffffffLL));
    growTo(w.count, sizeof(int));
    growTo(w.idx, static_cast<size_t>(candCap) * sizeof(int));
    growTo(w.val, static_cast<size_t>(candCap) * sizeof(T));
    MF_HIP(hipMemsetAsync(w.count.p, 0, sizeof(int), stream));
    MF_HIP(mfk::launch_peaks(sizeof(T) == 4 ? 0 : 1, dx, n, threshold, w.count.as<int>(), w.idx.as<int>(), w.val.p,
                             candCap, stream));
    // one readback of the count and the first kPrefix candidates (usually
    // all of them); the rest, if any, in a second
    constexpr int kPrefix = 8192;
    const int pre = std::min(kPrefix, candCap);
    const size_t valOff = (1 + static_cast<size_t>(pre)) * sizeof(int);
    w.host.reserve(valOff + static_cast<size_t>(pre) * sizeof(T));
    int *hCount = static_cast<int *>(w.host.p);
    int *hIdx = hCount + 1;
    T *hVal = reinterpret_cast<T *>(static_cast<char *>(w.host.p) + valOff);
    MF_HIP(hipMemcpyAsync(hCount, w.count.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    MF_HIP(hipMemcpyAsync(hIdx, w.idx.p, static_cast<size_t>(pre) * sizeof(int), hipMemcpyDeviceToHost, stream));
    MF_HIP(hipMemcpyAsync(hVal, w.val.p, static_cast<size_t>(pre) * sizeof(T), hipMemcpyDeviceToHost, stream));
    if (edges != nullptr) {  // the signal's first two and last two samples, in the same wait
        const long long at[4] = {0, std::min(1LL, n - 1), std::max(n - 2, 0LL), n - 1};
        for (int i = 0; i < 4; ++i) {
            MF_HIP(hipMemcpyAsync(&edges[i], static_cast<const T *>(dx) + at[i], sizeof(T), hipMemcpyDeviceToHost,
                                  stream));
        }
    }
    MF_HIP(hipStreamSynchronize(stream));
    const int count = *hCount;
    if (count > candCap) { fail(MF_ERR_DEVICE, "peak candidate buffer overflow"); }
    cand.resize(count);
    if (count <= pre) {
        for (int i = 0; i < count; ++i) { cand[i] = {hVal[i], hIdx[i]}; }
        return;
    }
    std::vector<int> ci(count);
    std::vector<T> cv(count);
    MF_HIP(hipMemcpyAsync(ci.data(), w.idx.p, count * sizeof(int), hipMemcpyDeviceToHost, stream));
    MF_HIP(hipMemcpyAsync(cv.data(), w.val.p, count * sizeof(T), hipMemcpyDeviceToHost, stream));
    MF_HIP(hipStreamSynchronize(stream));
    for (int i = 0; i < count; ++i) { cand[i] = {cv[i], ci[i]}; }
}

template <typename T>
void peaksDevice(const void *dx, long long n, double threshold, int minDist, int *idx, int cap, int *nPeaks,
                 hipStream_t stream)
{
    std::vector<std::pair<T, int>> cand;
    rawCandidates<T>(dx, n, threshold, cand, stream);
    pruneCandidates<T>(cand, minDist, idx, cap, nPeaks);
}

// Stack every group and find the candidates of every stack (one launch),
// then prune per group on the host.
template <typename T>
void detectT(mf_handle *h, double threshold, int minDist, int *idx, int cap, int *offsets)
{
    const int nG = static_cast<int>(h->groups.size());
    long long maxLen = 0;
    for (const auto &gi : h->stackInfo) { maxLen = std::max(maxLen, gi.len); }
    MF_HIP(hipMemsetAsync(h->dCounts.p, 0, nG * sizeof(int), h->stream));
    // stacks and their candidates in one pass (the stacks are not read back)
    MF_HIP(hipEventRecord(h->evC, h->stream));
    MF_HIP(mfk::launch_stack_peak_groups(h->precision, h->dGroupTab.as<mfk::StackGroup>(), nG, maxLen, h->absval,
                                         threshold, h->stream));
    MF_HIP(hipEventRecord(h->evD, h->stream));
    MF_HIP(mfk::launch_pack_candidates(h->precision, h->dGroupTab.as<mfk::StackGroup>(), nG, h->dPackIdx.as<int>(),
                                       h->dPackVal.p, h->stream));
    // one readback of the counts and the first kPrefix candidates (usually
    // all of them); a second only when there are more
    constexpr size_t kPrefix = 8192;
    const size_t pre = std::min(kPrefix, h->candCap * static_cast<size_t>(nG));
    const size_t hdrBytes = (static_cast<size_t>(nG) + pre) * sizeof(int);
    h->hPack.reserve(hdrBytes + pre * sizeof(T));
    int *hIdx = static_cast<int *>(h->hPack.p);
    T *hVal = reinterpret_cast<T *>(static_cast<char *>(h->hPack.p) + hdrBytes);
    MF_HIP(hipMemcpyAsync(hIdx, h->dPackIdx.p, hdrBytes, hipMemcpyDeviceToHost, h->stream));
    MF_HIP(hipMemcpyAsync(hVal, h->dPackVal.p, pre * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    MF_HIP(hipStreamSynchronize(h->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, h->evC, h->evD) == hipSuccess) { h->stackMs = ms; }
    std::vector<int> counts(nG);
    std::vector<size_t> first(nG + 1, 0);
    for (int g = 0; g < nG; ++g) {
        counts[g] = hIdx[g];
        if (counts[g] > static_cast<int>(h->candCap)) { fail(MF_ERR_DEVICE, "peak candidate buffer overflow"); }
        first[g + 1] = first[g] + static_cast<size_t>(counts[g]);
    }
    const size_t total = first[nG];
    std::vector<int> ci(total);
    std::vector<T> cv(total);
    std::copy(hIdx + nG, hIdx + nG + std::min(total, pre), ci.begin());
    std::copy(hVal, hVal + std::min(total, pre), cv.begin());
    if (total > pre) {
        MF_HIP(hipMemcpyAsync(ci.data() + pre, h->dPackIdx.as<int>() + nG + pre, (total - pre) * sizeof(int),
                              hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipMemcpyAsync(cv.data() + pre, static_cast<const T *>(h->dPackVal.p) + pre, (total - pre) * sizeof(T),
                              hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    }
    int off = 0;
    for (int g = 0; g < nG; ++g) {
        offsets[g] = off;
        const int cnt = h->stackInfo[g].len <= 0 ? 0 : counts[g];
        std::vector<std::pair<T, int>> cand(cnt);
        for (int i = 0; i < cnt; ++i) { cand[i] = {cv[first[g] + i], ci[first[g] + i]}; }
        int np = 0;
        pruneCandidates<T>(cand, minDist, idx + off, cap - off, &np);
        off += np;
    }
    offsets[nG] = off;
}

void peaks(const void *x, bool onDevice, int n, int precision, double threshold, int minDist, int *idx, int cap,
           int *nPeaks)
{
    if (x == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "x is NULL"); }
    if (n < 1) { fail(MF_ERR_INVALID_ARGUMENT, "No points in x"); }
    if (nPeaks == nullptr || (idx == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "peaks is NULL"); }
    if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
        fail(MF_ERR_INVALID_ARGUMENT, "unknown precision");
    }
    if (threshold < 0 || threshold > 1) {
        fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
    }
    const size_t ts = precision == MF_PRECISION_FLOAT ? 4 : 8;
    std::lock_guard<std::mutex> lock(g_peakMutex);
    PeakWork &w = peakWork();
    if (w.stream == nullptr) { MF_HIP(hipStreamCreate(&w.stream)); }
    hipStream_t stream = w.stream;
    const void *dx = x;
    if (!onDevice) {
        growTo(w.staged, static_cast<size_t>(n) * ts);
        MF_HIP(hipMemcpyAsync(w.staged.p, x, static_cast<size_t>(n) * ts, hipMemcpyHostToDevice, stream));
        dx = w.staged.p;
    }
    if (precision == MF_PRECISION_FLOAT) {
        peaksDevice<float>(dx, n, static_cast<float>(threshold), minDist, idx, cap, nPeaks, stream);
    } else {
        peaksDevice<double>(dx, n, threshold, minDist, idx, cap, nPeaks, stream);
    }
}

// Scratch of mf_peaks_device_noalloc: the candidate counter and candidate
// list of one call, in a file-scope device array that every GPU the library
// runs on holds until the process ends (24 MiB).  Host code never passes
// `scratch` itself to a kernel: its name is a host-side stand-in, so the
// device address comes from hipGetSymbolAddress, once per device.
constexpr long long kNoAllocMaxSamples = 1LL << 22;                     // 4,194,304 samples
constexpr int kNoAllocCandCap = static_cast<int>(kNoAllocMaxSamples / 2 + 2);  // worst case n/2 + 2
constexpr size_t kScratchIdxOff = 256;                                  // [0, 256): the counter
constexpr size_t kScratchValOff = kScratchIdxOff + ((kNoAllocCandCap * sizeof(int) + 255) / 256) * 256;
constexpr size_t kScratchBytes = kScratchValOff + static_cast<size_t>(kNoAllocCandCap) * sizeof(double);
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_noAllocMutex;  // serializes every call (they share one array per GPU)
void *g_scratchPtr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
void *scratchOnCurrentDevice()
{
    int dev = -1;
    MF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) { fail(MF_ERR_UNSUPPORTED, "device index out of range"); }
    if (g_scratchPtr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            fail(MF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        }
        g_scratchPtr[dev] = p;
    }
    return g_scratchPtr[dev];
}

// PeakFinder on a device signal with the candidates in the scratch array:
// no device or pinned allocation; runs on the null stream.
template <typename T>
void peaksNoAlloc(const void *dx, long long n, double threshold, int minDist, int *idx, int cap, int *nPeaks)
{
    char *base = static_cast<char *>(scratchOnCurrentDevice());
    int *dCount = reinterpret_cast<int *>(base);
    int *dIdx = reinterpret_cast<int *>(base + kScratchIdxOff);
    T *dVal = reinterpret_cast<T *>(base + kScratchValOff);
    const int candCap = static_cast<int>(n / 2 + 2);
    hipStream_t stream = nullptr;
    // the counter keeps the last call's value: a fill kernel clears it
    MF_HIP(mfk::launch_fill_int(dCount, 1, 0, stream));
    MF_HIP(mfk::launch_peaks(sizeof(T) == 4 ? 0 : 1, dx, n, threshold, dCount, dIdx, dVal, candCap, stream));
    int count = 0;
    MF_HIP(hipMemcpyAsync(&count, dCount, sizeof(int), hipMemcpyDeviceToHost, stream));
    MF_HIP(hipStreamSynchronize(stream));
    if (count < 0 || count > candCap) { fail(MF_ERR_DEVICE, "peak candidate count out of range"); }
    std::vector<int> ci(count);
    std::vector<T> cv(count);
    if (count > 0) {
        MF_HIP(hipMemcpyAsync(ci.data(), dIdx, count * sizeof(int), hipMemcpyDeviceToHost, stream));
        MF_HIP(hipMemcpyAsync(cv.data(), dVal, count * sizeof(T), hipMemcpyDeviceToHost, stream));
        MF_HIP(hipStreamSynchronize(stream));
    }
    std::vector<std::pair<T, int>> cand(count);
    for (int i = 0; i < count; ++i) { cand[i] = {cv[i], ci[i]}; }
    pruneCandidates<T>(cand, minDist, idx, cap, nPeaks);
}

// Measured cost table (mf_autotune): for every FFT size and kernel variant
// the correlation launch is timed on 16 synthetic lanes x 128 blocks
// (2048 work items: the whole GPU) with 1 row, 8 rows of one template
// length and 8 rows of distinct lengths per lane; the per-block costs
// a + bShared*shared + bSolo*solo follow from the three times.
void calibrate(int precision, std::vector<double> &out)
{
    constexpr int C = 16, B = 128, R = 8;
    const size_t ts = precision == MF_PRECISION_FLOAT ? 4 : 8;
    CostTable t;
    uint64_t state = 0x9E3779B97F4A7C15ull;
    auto next = [&state]() {  // xorshift64*, deterministic
        state ^= state >> 12;
        state ^= state << 25;
        state ^= state >> 27;
        return static_cast<double>((state * 0x2545F4914F6CDD1Dull) >> 11) / 9007199254740992.0 - 0.5;
    };
    for (int lg = mfk::kMinLogM; lg <= maxLogM(precision); ++lg) {
        const int N = 2 << lg;
        const int nb = N / 4;
        const int L = blockLength(lg, nb);
        const int n = B * L + nb - 1;
        std::vector<char> sig(static_cast<size_t>(n) * ts);
        double times[3][3] = {};  // [config][variant]
        for (int cfg = 0; cfg < 3; ++cfg) {
            const int rowsPerLane = cfg == 0 ? 1 : R;
            std::unique_ptr<mf_handle> h(new mf_handle());
            h->precision = precision;
            h->mode = MF_MODE_BATCHED;
            h->tsize = ts;
            MF_HIP(hipStreamCreate(&h->stream));
            MF_HIP(hipEventCreate(&h->evA));
            MF_HIP(hipEventCreate(&h->evB));
            MF_HIP(hipEventCreate(&h->evC));
            MF_HIP(hipEventCreate(&h->evD));
            for (int c = 0; c < C; ++c) {
                for (int k = 0; k < rowsPerLane; ++k) {
                    const int Lt = cfg == 2 ? nb - 4 * k : nb;  // cfg 2: distinct lengths
                    std::vector<double> tv(Lt);
                    for (int m = 0; m < Lt; ++m) { tv[m] = next() * std::sin(M_PI * (m + 0.5) / Lt); }
                    h->tmpl.push_back(std::move(tv));
                    h->rowChIn.push_back(c);
                    h->rowGroupIn.push_back(c);
                }
            }
            h->weights.assign(h->tmpl.size(), 1.0);
            h->shifts.assign(h->tmpl.size(), 0);
            h->haveLayout = true;
            h->nChannels = C;
            h->nGroupsIn = C;
            h->chLenIn.assign(C, n);
            h->userNfft = N;
            initialize(h.get());
            for (int c = 0; c < C; ++c) {
                if (precision == MF_PRECISION_FLOAT) {
                    float *f = reinterpret_cast<float *>(sig.data());
                    for (int i = 0; i < n; ++i) { f[i] = static_cast<float>(next()); }
                } else {
                    double *d = reinterpret_cast<double *>(sig.data());
                    for (int i = 0; i < n; ++i) { d[i] = next(); }
                }
                MF_HIP(hipMemcpy(h->ch[c].buf.p, sig.data(), sig.size(), hipMemcpyHostToDevice));
                h->ch[c].dead = false;
                h->ch[c].set = true;
            }
            h->tablesDirty = true;
            for (int v = 1; v <= 2; ++v) {
                if (!variantAvailable(precision, lg, v)) { continue; }
                for (auto &la : h->launches) { la.variant = v; }
                std::vector<float> ms;
                for (int rep = 0; rep < 4; ++rep) {
                    apply(h.get());
                    MF_HIP(hipStreamSynchronize(h->stream));
                    float e = 0;
                    MF_HIP(hipEventElapsedTime(&e, h->evA, h->evB));
                    if (rep > 0) { ms.push_back(e); }
                }
                std::sort(ms.begin(), ms.end());
                times[cfg][v] = ms[ms.size() / 2];
            }
            dropDevice(h.get());
        }
        const double blocks = static_cast<double>(C) * B;
        for (int v = 1; v <= 2; ++v) {
            if (times[0][v] <= 0) { continue; }
            const double t1 = times[0][v] * 1e3 / blocks;  // us per block
            const double tS = times[1][v] * 1e3 / blocks;
            const double tD = times[2][v] * 1e3 / blocks;
            const double bSolo = std::max((tD - t1) / (R - 1), 1e-6);
            const double a = std::max(t1 - bSolo, 0.05 * t1);
            const double bShared = std::max((tS - a) / R, 1e-6);
            t.a[lg][v] = a;
            t.bShared[lg][v] = bShared;
            t.bSolo[lg][v] = bSolo;
            out.insert(out.end(), {static_cast<double>(lg), static_cast<double>(v), a, bShared, bSolo});
        }
    }
    t.measured = true;
    std::lock_guard<std::mutex> lk(costMutex());
    costTable(precision) = t;
}

}  // namespace

extern "C" {

const char *mf_last_error(void) { return g_lastError.c_str(); }

int mf_version(int *major, int *minor, int *patch)
{
    return guard([&] {
        if (major) { *major = 0; }
        if (minor) { *minor = 2; }
        if (patch) { *patch = 0; }
    });
}

int mf_create(mf_handle **h, int precision, int mode)
{
    return guard([&] {
        if (h == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "handle pointer is NULL"); }
        *h = nullptr;
        if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(precision));
        }
        if (mode != MF_MODE_SINGLE_CHANNEL && mode != MF_MODE_MULTI_CHANNEL && mode != MF_MODE_BATCHED) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown detection mode " + std::to_string(mode));
        }
        requireDevice();
        auto *p = new mf_handle();
        p->precision = precision;
        p->mode = mode;
        p->tsize = precision == MF_PRECISION_FLOAT ? 4 : 8;
        hipError_t e = hipGetDevice(&p->device);
        if (e == hipSuccess) { e = hipStreamCreate(&p->stream); }
        if (e == hipSuccess) { e = hipEventCreate(&p->evA); }
        if (e == hipSuccess) { e = hipEventCreate(&p->evB); }
        if (e == hipSuccess) { e = hipEventCreate(&p->evC); }
        if (e == hipSuccess) { e = hipEventCreate(&p->evD); }
        if (e != hipSuccess) {
            delete p;
            fail(MF_ERR_DEVICE, std::string("stream/event creation failed: ") + hipGetErrorString(e));
        }
        *h = p;
    });
}

int mf_destroy(mf_handle *h)
{
    return guard([&] {
        if (h == nullptr) { return; }
        dropDevice(h);
        delete h;
    });
}

int mf_clear(mf_handle *h)
{
    return guard([&] {
        requireHandle(h);
        dropDevice(h);
        h->tmpl.clear();
        h->shifts.clear();
        h->weights.clear();
        h->haveShifts = false;
        h->signalSize = 0;
        h->haveLayout = false;
    });
}

int mf_set_signal_size(mf_handle *h, int n)
{
    return guard([&] {
        requireHandle(h);
        if (n < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Signal size must be positive"); }
        h->signalSize = n;
    });
}

int mf_set_templates(mf_handle *h, int nTemplates, const int *lengths, const double *const *samples,
                     const int *shifts, const double *weights)
{
    return guard([&] {
        requireHandle(h);
        if (nTemplates < 1 || lengths == nullptr || samples == nullptr) {
            fail(MF_ERR_INVALID_ARGUMENT, "Waveform not yet set on template");
        }
        std::vector<std::vector<double>> t(nTemplates);
        for (int i = 0; i < nTemplates; ++i) {
            if (lengths[i] < 1 || samples[i] == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "Waveform not yet set on template"); }
            t[i].assign(samples[i], samples[i] + lengths[i]);
        }
        std::vector<double> w(nTemplates, 1.0);
        if (weights) {
            for (int i = 0; i < nTemplates; ++i) {
                if (weights[i] < 0 || weights[i] > 1) {
                    fail(MF_ERR_INVALID_ARGUMENT, "weight = " + std::to_string(weights[i]) + " must be in range [0,1]");
                }
                w[i] = weights[i];
            }
        }
        h->tmpl = std::move(t);
        h->weights = std::move(w);
        h->haveShifts = shifts != nullptr;
        h->shifts.assign(nTemplates, 0);
        if (shifts) { std::copy(shifts, shifts + nTemplates, h->shifts.begin()); }
    });
}

int mf_set_stack_absolute_values(mf_handle *h, int absval)
{
    return guard([&] {
        requireHandle(h);
        h->absval = absval ? 1 : 0;
    });
}

int mf_set_batch_layout(mf_handle *h, int nChannels, const int *channelLengths, const int *templateChannel,
                        const int *templateGroup, int nGroups)
{
    return guard([&] {
        requireHandle(h);
        if (h->mode != MF_MODE_BATCHED) { fail(MF_ERR_BAD_STATE, "batch layout needs MF_MODE_BATCHED"); }
        const int nT = static_cast<int>(h->tmpl.size());
        if (nT < 1) { fail(MF_ERR_BAD_STATE, "set the templates before the batch layout"); }
        if (nChannels < 1 || nGroups < 1 || channelLengths == nullptr || templateChannel == nullptr ||
            templateGroup == nullptr) {
            fail(MF_ERR_INVALID_ARGUMENT, "incomplete batch layout");
        }
        for (int c = 0; c < nChannels; ++c) {
            if (channelLengths[c] < 1) { fail(MF_ERR_INVALID_ARGUMENT, "channel lengths must be positive"); }
        }
        for (int r = 0; r < nT; ++r) {
            if (templateChannel[r] < 0 || templateChannel[r] >= nChannels) {
                fail(MF_ERR_INVALID_ARGUMENT, "template " + std::to_string(r) + " names channel " +
                                                  std::to_string(templateChannel[r]) + " out of range");
            }
            if (templateGroup[r] < 0 || templateGroup[r] >= nGroups) {
                fail(MF_ERR_INVALID_ARGUMENT, "template " + std::to_string(r) + " names group " +
                                                  std::to_string(templateGroup[r]) + " out of range");
            }
        }
        h->nChannels = nChannels;
        h->nGroupsIn = nGroups;
        h->chLenIn.assign(channelLengths, channelLengths + nChannels);
        h->rowChIn.assign(templateChannel, templateChannel + nT);
        h->rowGroupIn.assign(templateGroup, templateGroup + nT);
        h->haveLayout = true;
    });
}

int mf_set_kernel_variant(mf_handle *h, int variant)
{
    return guard([&] {
        requireHandle(h);
        if (variant < 0 || variant > 4) { fail(MF_ERR_INVALID_ARGUMENT, "kernel variant must be 0..4"); }
        h->userVariant = variant;
    });
}

int mf_set_fft_length(mf_handle *h, int nfft)
{
    return guard([&] {
        requireHandle(h);
        if (nfft < 0) { fail(MF_ERR_INVALID_ARGUMENT, "fftLength must be non-negative"); }
        h->userNfft = nfft;
    });
}

int mf_set_graph_launch(mf_handle *h, int on)
{
    return guard([&] {
        requireHandle(h);
        if (h->stream) { MF_HIP(hipStreamSynchronize(h->stream)); }
        h->dropGraph();
        h->useGraph = on != 0;
    });
}

int mf_initialize(mf_handle *h)
{
    return guard([&] {
        requireHandle(h);
        initialize(h);
    });
}

int mf_is_initialized(const mf_handle *h, int *yes)
{
    return guard([&] {
        requireHandle(h);
        if (yes) { *yes = h->initialized ? 1 : 0; }
    });
}

int mf_get_number_of_templates(const mf_handle *h, int *n)
{
    return guard([&] {
        requireInit(h);
        if (n) { *n = static_cast<int>(h->rows.size()); }
    });
}

int mf_get_signal_length(const mf_handle *h, int *n)
{
    return guard([&] {
        requireInit(h);
        if (n) { *n = static_cast<int>(h->ch[0].n - h->carry); }  // streaming: the chunk length
    });
}

int mf_get_filtered_length(const mf_handle *h, int row, int *n)
{
    return guard([&] {
        requireInit(h);
        checkRow(h, row);
        if (n) { *n = h->rows[row].nOut; }
    });
}

int mf_get_plan(const mf_handle *h, int row, int *nfft, int *blockLength)
{
    return guard([&] {
        requireInit(h);
        checkRow(h, row);
        const auto &ln = h->lanes[h->rows[row].lane];
        if (nfft) { *nfft = ln.td ? 0 : 2 << ln.logM; }  // 0: time-domain (no FFT)
        if (blockLength) { *blockLength = ln.L; }
    });
}

namespace {

// Streaming push: the last `carry` samples of the channel move to its
// front (through a scratch buffer: the ranges overlap when the chunk is
// shorter than the carry) and the chunk follows.  The dead-trace rule over
// the whole buffer (carry + chunk) is evaluated on the device by apply().
void pushChunk(mf_handle *h, int channel, int n, const void *x, hipMemcpyKind kind, bool wait = true)
{
    auto &c = h->ch[channel];
    const size_t ts = h->tsize;
    char *buf = static_cast<char *>(c.buf.p);
    if (h->carry > 0) {
        MF_HIP(hipMemcpyAsync(h->dCarry.p, buf + static_cast<size_t>(n) * ts, static_cast<size_t>(h->carry) * ts,
                              hipMemcpyDeviceToDevice, h->stream));
        MF_HIP(hipMemcpyAsync(buf, h->dCarry.p, static_cast<size_t>(h->carry) * ts, hipMemcpyDeviceToDevice,
                              h->stream));
    }
    MF_HIP(hipMemcpyAsync(buf + static_cast<size_t>(h->carry) * ts, x, static_cast<size_t>(n) * ts, kind, h->stream));
    c.set = true;
    // the caller may reuse its buffer once the push returns
    if (wait) { MF_HIP(hipStreamSynchronize(h->stream)); }
}

// True when kernels on `device` can read p directly: device (or managed)
// memory of that device.  Host pointers and other devices' memory go
// through the copy engine instead.
bool readableOnDevice(const void *p, int device)
{
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unknown (pageable host) pointer: clear the error it leaves
        return false;
    }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == device;
}

void checkChunkLength(const mf_handle *h, int c, int n)
{
    const long long chunk = h->ch[c].n - h->carry;
    if (n != chunk) {
        fail(MF_ERR_INVALID_ARGUMENT, "Number of samples = " + std::to_string(n) + " must equal the chunk length " +
                                          std::to_string(chunk));
    }
}

}  // namespace

int mf_push_chunks_device(mf_handle *h, int nChannels, int n, const void *const *dx)
{
    return guard([&] {
        requireInit(h);
        if (!h->streaming) { fail(MF_ERR_BAD_STATE, "mf_push_chunks_device needs streaming mode"); }
        if (nChannels != static_cast<int>(h->ch.size())) {
            fail(MF_ERR_INVALID_ARGUMENT, "a push covers all " + std::to_string(h->ch.size()) + " channels");
        }
        if (dx == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "chunk table is NULL"); }
        for (int c = 0; c < nChannels; ++c) {
            checkChunkLength(h, c, n);
            if (dx[c] == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "chunk of channel " + std::to_string(c) + " is NULL"); }
        }
        h->applied = false;
        bool direct = nChannels <= mfk::kPushMaxChannels;
        for (int c = 0; c < nChannels && direct; ++c) { direct = readableOnDevice(dx[c], h->device); }
        if (direct) {
            // two launches for every channel instead of three copies per channel
            mfk::PushTable tab{};
            for (int c = 0; c < nChannels; ++c) {
                tab.buf[c] = h->ch[c].buf.p;
                tab.src[c] = dx[c];
                h->ch[c].set = true;
            }
            const long long words = static_cast<long long>(h->tsize / 4);
            growTo(h->dPushScratch, static_cast<size_t>(nChannels) * static_cast<size_t>(h->carry) * h->tsize);
            MF_HIP(mfk::launch_push(tab, nChannels, n * words, h->carry * words, h->dPushScratch.as<unsigned>(),
                                    h->stream));
        } else {  // copies handle any pointer the runtime can reach
            for (int c = 0; c < nChannels; ++c) { pushChunk(h, c, n, dx[c], hipMemcpyDefault, false); }
        }
        MF_HIP(hipStreamSynchronize(h->stream));
    });
}

int mf_set_streaming(mf_handle *h, int on)
{
    return guard([&] {
        requireHandle(h);
        if (h->initialized) { fail(MF_ERR_BAD_STATE, "streaming must be chosen before initialize"); }
        h->streaming = on != 0;
    });
}

int mf_get_stream_state(mf_handle *h, int channel, void *out, int cap, int *len)
{
    return guard([&] {
        requireInit(h);
        if (!h->streaming) { fail(MF_ERR_BAD_STATE, "stream state needs streaming mode"); }
        checkChannel(h, channel);
        if (len) { *len = static_cast<int>(h->carry); }
        if (h->carry == 0) { return; }
        if (out == nullptr || cap < h->carry) { fail(MF_ERR_INVALID_ARGUMENT, "output holds fewer than carry samples"); }
        // the samples the next push moves to the front: the buffer's tail
        const auto &c = h->ch[channel];
        MF_HIP(hipMemcpyAsync(out, static_cast<const char *>(c.buf.p) + static_cast<size_t>(c.n - h->carry) * h->tsize,
                              static_cast<size_t>(h->carry) * h->tsize, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    });
}

int mf_set_stream_state(mf_handle *h, int channel, const void *state, int n)
{
    return guard([&] {
        requireInit(h);
        if (!h->streaming) { fail(MF_ERR_BAD_STATE, "stream state needs streaming mode"); }
        checkChannel(h, channel);
        if (n != h->carry) {
            fail(MF_ERR_INVALID_ARGUMENT, "state holds " + std::to_string(n) + " samples, the carry is " +
                                              std::to_string(h->carry));
        }
        if (n == 0) { return; }
        if (state == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "state is NULL"); }
        auto &c = h->ch[channel];
        MF_HIP(hipMemcpyAsync(static_cast<char *>(c.buf.p) + static_cast<size_t>(c.n - h->carry) * h->tsize, state,
                              static_cast<size_t>(n) * h->tsize, hipMemcpyHostToDevice, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        h->applied = false;
        // the stacks now continue another stream position: every group's
        // stream-peak finder starts over (or is restored after the carries)
        for (auto &st : h->streamPeaks) { st = mf_handle::StreamPeakState{}; }
    });
}

extern "C++" {
namespace {

// One push of the streaming PeakFinder over group g's stack.  The stream's
// lag K is the stack lag whose windows start K samples after the first
// sample of the stream's first push: push p's new stack entries (the last
// `chunk` of the stack) are K = p * chunk + (len - chunk) - carry + k, and
// lags before K = 0 (windows reaching into the carry the first push found)
// are not part of the stream.  Works on a copy of the state; the caller
// keeps it only when the peaks fit the output.
template <typename T>
void streamPeaksPush(mf_handle *h, int group, mfstream::StreamPeaks<T> &sp, long long p, bool last,
                     std::vector<long long> &out)
{
    const long long chunk = h->ch[0].n - h->carry;
    stackGroups(h, group);  // on the handle's stream; evD marks its end
    const long long len = h->stackInfo[group].len;
    const long long newStart = len - chunk;
    if (newStart < 0) { fail(MF_ERR_DEVICE, "stack shorter than the chunk"); }
    const long long base = p * chunk + newStart - h->carry;
    const long long v0 = std::min(chunk, std::max(0LL, -base));
    const long long n = chunk - v0;
    if (n > 0) {
        const char *dx = static_cast<const char *>(stackPtr(h, group)) + static_cast<size_t>(newStart + v0) * h->tsize;
        std::vector<std::pair<T, int>> cand;
        T edge[4];
        {
            std::lock_guard<std::mutex> lock(g_peakMutex);
            PeakWork &w = peakWork();
            if (w.stream == nullptr) { MF_HIP(hipStreamCreate(&w.stream)); }
            MF_HIP(hipStreamWaitEvent(w.stream, h->evD, 0));  // the scan reads the stack
            rawCandidates<T>(dx, n, static_cast<T>(sp.threshold()), cand, w.stream, edge);
        }
        sp.push(base + v0, static_cast<int>(n), std::move(cand), edge[0], edge[1], edge[2], edge[3], out);
    }
    if (last) { sp.flush(out); }
}

// One push of the streaming PeakFinder over every group at once: one stack
// launch, one candidate scan of all the new segments, one gather of their
// edge samples and one wait.  Same lag convention as streamPeaksPush.
template <typename T>
void streamPeaksPushAll(mf_handle *h, std::vector<mfstream::StreamPeaks<T>> &sps, const std::vector<long long> &p,
                        bool last, std::vector<std::vector<long long>> &out)
{
    const int nG = static_cast<int>(h->groups.size());
    const long long chunk = h->ch[0].n - h->carry;
    const int words = static_cast<int>(sizeof(T) / 4);  // 32-bit words per sample
    if (static_cast<long long>(h->ldStack) * nG * words > 0x7fffffffLL) {
        fail(MF_ERR_UNSUPPORTED, "stacks too long for the batched stream peak finder; call per group");
    }
    stackGroups(h, -1);
    std::vector<mfk::StackGroup> tab(h->hGroupTab);
    std::vector<long long> g0(nG), segN(nG);
    std::vector<int> edgeIdx(static_cast<size_t>(4) * nG * words);
    long long maxN = 0;
    for (int g = 0; g < nG; ++g) {
        const long long len = h->stackInfo[g].len;
        const long long newStart = len - chunk;
        if (newStart < 0) { fail(MF_ERR_DEVICE, "stack shorter than the chunk"); }
        const long long base = p[g] * chunk + newStart - h->carry;
        const long long v0 = std::min(chunk, std::max(0LL, -base));
        const long long n = chunk - v0;
        g0[g] = base + v0;
        segN[g] = n;
        maxN = std::max(maxN, n);
        // the segment's first sample (any sample of the group's row when it is empty)
        const long long first = static_cast<long long>(g) * h->ldStack + (n > 0 ? newStart + v0 : 0);
        tab[g].out = static_cast<char *>(tab[g].out) + static_cast<size_t>(newStart + v0) * sizeof(T);
        tab[g].len = n;
        const long long at[4] = {0, std::min(1LL, std::max(n - 1, 0LL)), std::max(n - 2, 0LL), std::max(n - 1, 0LL)};
        for (int i = 0; i < 4; ++i) {
            for (int w = 0; w < words; ++w) {
                edgeIdx[(static_cast<size_t>(g) * 4 + i) * words + w] = static_cast<int>((first + at[i]) * words + w);
            }
        }
    }
    growTo(h->dStreamTab, tab.size() * sizeof(mfk::StackGroup));
    growTo(h->dStreamIdx, edgeIdx.size() * sizeof(int));
    growTo(h->dStreamEdge, edgeIdx.size() * sizeof(int));
    MF_HIP(hipMemcpyAsync(h->dStreamTab.p, tab.data(), tab.size() * sizeof(mfk::StackGroup), hipMemcpyHostToDevice,
                          h->stream));
    MF_HIP(hipMemcpyAsync(h->dStreamIdx.p, edgeIdx.data(), edgeIdx.size() * sizeof(int), hipMemcpyHostToDevice,
                          h->stream));
    MF_HIP(mfk::launch_gather_int(h->dStacks.as<int>(), h->dStreamIdx.as<int>(), static_cast<int>(edgeIdx.size()),
                                  h->dStreamEdge.as<int>(), h->stream));
    MF_HIP(hipMemsetAsync(h->dCounts.p, 0, nG * sizeof(int), h->stream));
    const mfk::StackGroup *dTab = h->dStreamTab.as<mfk::StackGroup>();
    MF_HIP(mfk::launch_peak_groups(h->precision, dTab, nG, maxN, static_cast<double>(sps[0].threshold()), h->stream));
    MF_HIP(mfk::launch_pack_candidates(h->precision, dTab, nG, h->dPackIdx.as<int>(), h->dPackVal.p, h->stream));
    // one readback: counts, the first kPrefix candidates and the edges
    constexpr size_t kPrefix = 8192;
    const size_t pre = std::min(kPrefix, h->candCap * static_cast<size_t>(nG));
    const size_t hdrBytes = (static_cast<size_t>(nG) + pre) * sizeof(int);
    h->hPack.reserve(hdrBytes + pre * sizeof(T));
    int *hIdx = static_cast<int *>(h->hPack.p);
    T *hVal = reinterpret_cast<T *>(static_cast<char *>(h->hPack.p) + hdrBytes);
    std::vector<T> edges(static_cast<size_t>(4) * nG);
    MF_HIP(hipMemcpyAsync(hIdx, h->dPackIdx.p, hdrBytes, hipMemcpyDeviceToHost, h->stream));
    MF_HIP(hipMemcpyAsync(hVal, h->dPackVal.p, pre * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    MF_HIP(hipMemcpyAsync(edges.data(), h->dStreamEdge.p, edges.size() * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    MF_HIP(hipStreamSynchronize(h->stream));
    std::vector<size_t> first(nG + 1, 0);
    for (int g = 0; g < nG; ++g) {
        if (hIdx[g] > static_cast<int>(h->candCap)) { fail(MF_ERR_DEVICE, "peak candidate buffer overflow"); }
        first[g + 1] = first[g] + static_cast<size_t>(hIdx[g]);
    }
    const size_t total = first[nG];
    std::vector<int> ci(total);
    std::vector<T> cv(total);
    std::copy(hIdx + nG, hIdx + nG + std::min(total, pre), ci.begin());
    std::copy(hVal, hVal + std::min(total, pre), cv.begin());
    if (total > pre) {
        MF_HIP(hipMemcpyAsync(ci.data() + pre, h->dPackIdx.as<int>() + nG + pre, (total - pre) * sizeof(int),
                              hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipMemcpyAsync(cv.data() + pre, static_cast<const T *>(h->dPackVal.p) + pre, (total - pre) * sizeof(T),
                              hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    }
    out.assign(nG, {});
    for (int g = 0; g < nG; ++g) {
        if (segN[g] > 0) {
            std::vector<std::pair<T, int>> cand(first[g + 1] - first[g]);
            for (size_t i = 0; i < cand.size(); ++i) { cand[i] = {cv[first[g] + i], ci[first[g] + i]}; }
            const T *e = edges.data() + static_cast<size_t>(4) * g;
            sps[g].push(g0[g], static_cast<int>(segN[g]), std::move(cand), e[0], e[1], e[2], e[3], out[g]);
        }
        if (last) { sps[g].flush(out[g]); }
    }
}


// The checks and the state bookkeeping shared by mf_stream_peaks{,_all}:
// returns the push number of the group's stream for this apply.
long long streamPeakCheck(mf_handle *h, int group, double threshold, int minDist)
{
    auto &st = h->streamPeaks[group];
    if (st.active) {
        if (threshold != st.threshold || minDist != st.minDist) {
            fail(MF_ERR_INVALID_ARGUMENT, "threshold and minimum distance are fixed until the stream's last push");
        }
        if (h->applies != st.lastApply + 1) {
            fail(MF_ERR_BAD_STATE, "mf_stream_peaks must be called once after every apply of the stream");
        }
    }
    return h->applies - (st.active ? st.origin : h->applies);
}

void streamPeakCommit(mf_handle *h, int group, double threshold, int minDist, bool last)
{
    auto &st = h->streamPeaks[group];
    if (last) {
        st = mf_handle::StreamPeakState{};  // the next call starts a new stream
        return;
    }
    if (!st.active) { st.origin = h->applies; }
    st.active = true;
    st.lastApply = h->applies;
    st.threshold = threshold;
    st.minDist = minDist;
}

template <typename T>
std::unique_ptr<mfstream::StreamPeaks<T>> &statePtr(mf_handle::StreamPeakState &st);
template <>
std::unique_ptr<mfstream::StreamPeaks<float>> &statePtr<float>(mf_handle::StreamPeakState &st) { return st.f; }
template <>
std::unique_ptr<mfstream::StreamPeaks<double>> &statePtr<double>(mf_handle::StreamPeakState &st) { return st.d; }

template <typename T>
void streamPeaksAllT(mf_handle *h, double threshold, int minDist, bool last, long long *peaks, int cap, int *offsets)
{
    const int nG = static_cast<int>(h->groups.size());
    std::vector<long long> p(nG);
    std::vector<mfstream::StreamPeaks<T>> sps;
    sps.reserve(nG);
    for (int g = 0; g < nG; ++g) {
        p[g] = streamPeakCheck(h, g, threshold, minDist);
        auto &sp = statePtr<T>(h->streamPeaks[g]);
        sps.push_back(h->streamPeaks[g].active ? *sp
                                               : mfstream::StreamPeaks<T>(static_cast<T>(threshold), minDist));
    }
    std::vector<std::vector<long long>> out;
    streamPeaksPushAll<T>(h, sps, p, last, out);
    long long total = 0;
    for (int g = 0; g < nG; ++g) {
        offsets[g] = static_cast<int>(total);
        total += static_cast<long long>(out[g].size());
    }
    offsets[nG] = static_cast<int>(total);
    if (total > cap) {
        fail(MF_ERR_INVALID_ARGUMENT, "peaks holds " + std::to_string(cap) + " < " + std::to_string(total));
    }
    for (int g = 0; g < nG; ++g) {
        std::copy(out[g].begin(), out[g].end(), peaks + offsets[g]);
        statePtr<T>(h->streamPeaks[g]) = std::make_unique<mfstream::StreamPeaks<T>>(std::move(sps[g]));
        streamPeakCommit(h, g, threshold, minDist, last);
    }
}

template <typename T>
void streamPeaksOneT(mf_handle *h, int group, double threshold, int minDist, bool last, long long *peaks, int cap,
                     int *nPeaks)
{
    const long long p = streamPeakCheck(h, group, threshold, minDist);
    auto &sp0 = statePtr<T>(h->streamPeaks[group]);
    mfstream::StreamPeaks<T> sp = h->streamPeaks[group].active
                                      ? *sp0
                                      : mfstream::StreamPeaks<T>(static_cast<T>(threshold), minDist);
    std::vector<long long> out;
    streamPeaksPush<T>(h, group, sp, p, last, out);
    *nPeaks = static_cast<int>(out.size());
    if (static_cast<long long>(out.size()) > cap) {
        fail(MF_ERR_INVALID_ARGUMENT, "peaks holds " + std::to_string(cap) + " < " + std::to_string(out.size()));
    }
    std::copy(out.begin(), out.end(), peaks);
    sp0 = std::make_unique<mfstream::StreamPeaks<T>>(std::move(sp));
    streamPeakCommit(h, group, threshold, minDist, last);
}

}  // namespace
}  // extern "C++"

extern "C++" {
namespace {

// Checkpoint of one group's stream-peak finder: a small header (whether a
// stream is open, how many pushes it has had, the threshold and minimum
// distance it was opened with) followed by the finder's own state.
constexpr uint32_t kStreamPeakMagic = 0x4753464Du;  // "MFSG"
struct StreamPeakHeader {
    uint32_t magic, version, active, tsize;
    int64_t pushes;
    double threshold;
    int32_t minDist, pad;
};

template <typename T>
void getStreamPeakStateT(mf_handle *h, int group, std::vector<unsigned char> &out)
{
    auto &st = h->streamPeaks[group];
    StreamPeakHeader hd{kStreamPeakMagic, 1, st.active ? 1u : 0u, static_cast<uint32_t>(sizeof(T)), 0, 0.0, 0, 0};
    std::vector<unsigned char> body;
    if (st.active) {
        hd.pushes = st.lastApply - st.origin + 1;
        hd.threshold = st.threshold;
        hd.minDist = st.minDist;
        statePtr<T>(st)->save(body);
    }
    out.resize(sizeof(hd) + body.size());
    std::memcpy(out.data(), &hd, sizeof(hd));
    if (!body.empty()) { std::memcpy(out.data() + sizeof(hd), body.data(), body.size()); }
}

template <typename T>
void setStreamPeakStateT(mf_handle *h, int group, const unsigned char *p, size_t n)
{
    StreamPeakHeader hd{};
    if (n < sizeof(hd)) { fail(MF_ERR_INVALID_ARGUMENT, "stream-peak state too short"); }
    std::memcpy(&hd, p, sizeof(hd));
    if (hd.magic != kStreamPeakMagic || hd.version != 1 || hd.active > 1) {
        fail(MF_ERR_INVALID_ARGUMENT, "not a stream-peak state");
    }
    if (hd.tsize != sizeof(T)) { fail(MF_ERR_INVALID_ARGUMENT, "stream-peak state of another precision"); }
    mf_handle::StreamPeakState st{};
    if (hd.active) {
        if (hd.pushes < 1 || hd.minDist < 0 || !(hd.threshold >= 0 && hd.threshold <= 1)) {
            fail(MF_ERR_INVALID_ARGUMENT, "stream-peak state out of range");
        }
        auto sp = std::make_unique<mfstream::StreamPeaks<T>>(T(0), 0);
        if (!sp->load(p + sizeof(hd), n - sizeof(hd))) { fail(MF_ERR_INVALID_ARGUMENT, "corrupt stream-peak state"); }
        st.active = true;
        // the next call must follow the next apply, as push number `pushes`
        st.lastApply = h->applies;
        st.origin = h->applies - hd.pushes + 1;
        st.threshold = hd.threshold;
        st.minDist = hd.minDist;
        statePtr<T>(st) = std::move(sp);
    } else if (n != sizeof(hd)) {
        fail(MF_ERR_INVALID_ARGUMENT, "stream-peak state has trailing bytes");
    }
    h->streamPeaks[group] = std::move(st);
}

void checkStreamPeakGroup(mf_handle *h, int group, bool allowAll)
{
    requireInit(h);
    if (!h->streaming) { fail(MF_ERR_BAD_STATE, "stream peaks need streaming mode"); }
    const int nG = static_cast<int>(h->groups.size());
    if (group < (allowAll ? -1 : 0) || group >= nG) {
        fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " must be in [0, " + std::to_string(nG) + ")");
    }
    if (h->streamPeaks.size() != h->groups.size()) { h->streamPeaks.resize(h->groups.size()); }
}

}  // namespace
}  // extern "C++"

int mf_get_stream_peaks_state(mf_handle *h, int group, void *out, long long cap, long long *len)
{
    return guard([&] {
        checkStreamPeakGroup(h, group, false);
        std::vector<unsigned char> blob;
        if (h->precision == MF_PRECISION_FLOAT) {
            getStreamPeakStateT<float>(h, group, blob);
        } else {
            getStreamPeakStateT<double>(h, group, blob);
        }
        if (len) { *len = static_cast<long long>(blob.size()); }
        if (out == nullptr || cap < static_cast<long long>(blob.size())) {
            fail(MF_ERR_INVALID_ARGUMENT, "state needs " + std::to_string(blob.size()) + " bytes");
        }
        std::memcpy(out, blob.data(), blob.size());
    });
}

int mf_set_stream_peaks_state(mf_handle *h, int group, const void *state, long long n)
{
    return guard([&] {
        checkStreamPeakGroup(h, group, false);
        if (state == nullptr || n <= 0) { fail(MF_ERR_INVALID_ARGUMENT, "state is empty"); }
        const auto *p = static_cast<const unsigned char *>(state);
        if (h->precision == MF_PRECISION_FLOAT) {
            setStreamPeakStateT<float>(h, group, p, static_cast<size_t>(n));
        } else {
            setStreamPeakStateT<double>(h, group, p, static_cast<size_t>(n));
        }
    });
}

int mf_stream_peaks_reset(mf_handle *h, int group)
{
    return guard([&] {
        checkStreamPeakGroup(h, group, true);
        for (int g = 0; g < static_cast<int>(h->streamPeaks.size()); ++g) {
            if (group < 0 || g == group) { h->streamPeaks[g] = mf_handle::StreamPeakState{}; }
        }
    });
}

int mf_stream_peaks(mf_handle *h, int group, double threshold, int minDist, int last, long long *peaks, int cap,
                    int *nPeaks)
{
    return guard([&] {
        checkStackable(h, group);
        if (!h->streaming) { fail(MF_ERR_BAD_STATE, "mf_stream_peaks needs streaming mode"); }
        if (group < 0) { fail(MF_ERR_INVALID_ARGUMENT, "group must be non-negative"); }
        if (nPeaks == nullptr || (peaks == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "peaks is NULL"); }
        if (!(threshold >= 0 && threshold <= 1)) {
            fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
        }
        for (const auto &c : h->ch) {
            if (c.n != h->ch[0].n) { fail(MF_ERR_UNSUPPORTED, "stream peaks need one chunk length on every channel"); }
        }
        if (h->streamPeaks.size() != h->groups.size()) { h->streamPeaks.resize(h->groups.size()); }
        if (h->precision == MF_PRECISION_FLOAT) {
            streamPeaksOneT<float>(h, group, threshold, minDist, last != 0, peaks, cap, nPeaks);
        } else {
            streamPeaksOneT<double>(h, group, threshold, minDist, last != 0, peaks, cap, nPeaks);
        }
    });
}

int mf_stream_peaks_all(mf_handle *h, double threshold, int minDist, int last, long long *peaks, int cap,
                        int *offsets)
{
    return guard([&] {
        checkStackable(h, -1);
        if (!h->streaming) { fail(MF_ERR_BAD_STATE, "mf_stream_peaks_all needs streaming mode"); }
        if (offsets == nullptr || (peaks == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "peaks is NULL"); }
        if (!(threshold >= 0 && threshold <= 1)) {
            fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
        }
        for (const auto &c : h->ch) {
            if (c.n != h->ch[0].n) { fail(MF_ERR_UNSUPPORTED, "stream peaks need one chunk length on every channel"); }
        }
        if (h->streamPeaks.size() != h->groups.size()) { h->streamPeaks.resize(h->groups.size()); }
        if (h->precision == MF_PRECISION_FLOAT) {
            streamPeaksAllT<float>(h, threshold, minDist, last != 0, peaks, cap, offsets);
        } else {
            streamPeaksAllT<double>(h, threshold, minDist, last != 0, peaks, cap, offsets);
        }
    });
}

int mf_get_stream_carry(const mf_handle *h, int *carry)
{
    return guard([&] {
        requireInit(h);
        if (carry == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        *carry = static_cast<int>(h->carry);
    });
}

int mf_set_signal(mf_handle *h, int channel, int n, const void *x)
{
    return guard([&] {
        requireInit(h);
        h->applied = false;
        checkChannel(h, channel);
        if (h->streaming) {
            checkChunkLength(h, channel, n);
            if (x == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "signal is NULL"); }
            pushChunk(h, channel, n, x, hipMemcpyHostToDevice);
            return;
        }
        checkSignalLength(h, channel, n);
        if (x == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "signal is NULL"); }
        auto &c = h->ch[channel];
        const bool dead = h->precision == MF_PRECISION_FLOAT ? isDead(static_cast<const float *>(x), n)
                                                             : isDead(static_cast<const double *>(x), n);
        MF_HIP(hipMemcpyAsync(c.buf.p, x, static_cast<size_t>(n) * h->tsize, hipMemcpyHostToDevice, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        c.dead = dead;
        c.set = true;
        h->tablesDirty = true;
    });
}

int mf_set_signal_device(mf_handle *h, int channel, int n, const void *dx)
{
    return guard([&] {
        requireInit(h);
        h->applied = false;
        checkChannel(h, channel);
        if (h->streaming) {
            checkChunkLength(h, channel, n);
            if (dx == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "signal is NULL"); }
            pushChunk(h, channel, n, dx, hipMemcpyDeviceToDevice);
            return;
        }
        checkSignalLength(h, channel, n);
        if (dx == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "signal is NULL"); }
        auto &c = h->ch[channel];
        MF_HIP(hipMemcpyAsync(c.buf.p, dx, static_cast<size_t>(n) * h->tsize, hipMemcpyDeviceToDevice, h->stream));
        const int nb = static_cast<int>(std::min<long long>((n + 255) / 256, 1024));
        DevBuf pmin, pmax;
        pmin.alloc(nb * sizeof(double));
        pmax.alloc(nb * sizeof(double));
        MF_HIP(mfk::launch_minmax(h->precision, c.buf.p, n, pmin.as<double>(), pmax.as<double>(), nb, h->stream));
        std::vector<double> lo(nb), hi(nb);
        MF_HIP(hipMemcpyAsync(lo.data(), pmin.p, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipMemcpyAsync(hi.data(), pmax.p, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        const double mn = *std::min_element(lo.begin(), lo.end());
        const double mx = *std::max_element(hi.begin(), hi.end());
        c.dead = std::abs(mn) == std::abs(mx);
        c.set = true;
        h->tablesDirty = true;
    });
}

int mf_zero_signal(mf_handle *h, int channel)
{
    return guard([&] {
        requireInit(h);
        h->applied = false;
        checkChannel(h, channel);
        auto &c = h->ch[channel];
        MF_HIP(hipMemsetAsync(c.buf.p, 0, c.buf.bytes, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        c.dead = true;
        c.set = true;
        h->tablesDirty = true;
    });
}

int mf_get_signal(mf_handle *h, int channel, void *out, int cap)
{
    return guard([&] {
        requireInit(h);
        checkChannel(h, channel);
        const auto &c = h->ch[channel];
        const long long n = c.n - h->carry;  // streaming: the newest chunk
        if (out == nullptr || cap < n) { fail(MF_ERR_INVALID_ARGUMENT, "output buffer too small"); }
        MF_HIP(hipMemcpyAsync(out, static_cast<const char *>(c.buf.p) + static_cast<size_t>(h->carry) * h->tsize,
                              static_cast<size_t>(n) * h->tsize, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    });
}

int mf_apply(mf_handle *h)
{
    return guard([&] { apply(h); });
}

int mf_synchronize(mf_handle *h)
{
    return guard([&] {
        requireHandle(h);
        MF_HIP(hipStreamSynchronize(h->stream));
        if (h->applied) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, h->evA, h->evB) == hipSuccess) { h->corrMs = ms; }
        }
    });
}

int mf_have_matched_filtered_signals(const mf_handle *h, int *yes)
{
    return guard([&] {
        requireHandle(h);
        if (yes) { *yes = h->applied ? 1 : 0; }
    });
}

int mf_get_rho(mf_handle *h, int row, void *out, int cap)
{
    return guard([&] {
        requireInit(h);
        checkRow(h, row);
        if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtering not yet applied"); }
        const int nOut = h->rows[row].nOut;
        if (out == nullptr || cap < nOut) { fail(MF_ERR_INVALID_ARGUMENT, "output buffer too small"); }
        const char *src = static_cast<const char *>(h->dRho.p) + h->rows[row].rhoOff * h->tsize;
        MF_HIP(hipMemcpyAsync(out, src, static_cast<size_t>(nOut) * h->tsize, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    });
}

int mf_get_rho_device(mf_handle *h, int row, const void **dptr)
{
    return guard([&] {
        requireInit(h);
        checkRow(h, row);
        if (dptr == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "dptr is NULL"); }
        *dptr = static_cast<const char *>(h->dRho.p) + h->rows[row].rhoOff * h->tsize;
    });
}

int mf_shift_and_stack(mf_handle *h, int group, void *out, int cap, int *len)
{
    return guard([&] {
        requireInit(h);
        int n = 0;
        checkStackable(h, group);
        if (group < 0) { fail(MF_ERR_INVALID_ARGUMENT, "group must be non-negative"); }
        n = static_cast<int>(h->stackInfo[group].len);
        if (len) { *len = n; }
        if (out == nullptr || cap < n) { fail(MF_ERR_INVALID_ARGUMENT, "output buffer too small"); }
        stack(h, group, nullptr, n, &n);
        MF_HIP(hipMemcpyAsync(out, stackPtr(h, group), static_cast<size_t>(n) * h->tsize, hipMemcpyDeviceToHost,
                              h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
    });
}

int mf_shift_and_stack_device(mf_handle *h, int group, void *dout, int cap, int *len)
{
    return guard([&] {
        requireInit(h);
        if (dout == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        stack(h, group, dout, cap, len);
    });
}

int mf_detect(mf_handle *h, double threshold, int minDist, int *idx, int cap, int *offsets)
{
    return guard([&] {
        checkStackable(h, -1);
        if (threshold < 0 || threshold > 1) {
            fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
        }
        if (offsets == nullptr || (idx == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        if (h->precision == MF_PRECISION_FLOAT) {
            detectT<float>(h, threshold, minDist, idx, cap, offsets);
        } else {
            detectT<double>(h, threshold, minDist, idx, cap, offsets);
        }
    });
}

int mf_get_stack_device(mf_handle *h, int group, const void **dptr, int *len)
{
    return guard([&] {
        requireInit(h);
        if (!h->stacksReady) { fail(MF_ERR_BAD_STATE, "no stack computed yet"); }
        if (group < 0 || group >= static_cast<int>(h->groups.size())) {
            fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " out of range");
        }
        if (dptr) { *dptr = stackPtr(h, group); }
        if (len) { *len = static_cast<int>(h->stackInfo[group].len); }
    });
}

namespace {

long long reduceRows(mf_handle *h, const std::vector<int> &rows, long long len, int absval, void *dDet, int *dId,
                     long long cap);

// Detector reduce + clamp of one group's rows into dDet/dId (device),
// asynchronous on the handle's stream; returns the length.
long long reduceGroup(mf_handle *h, int group, int absval, void *dDet, int *dId, long long cap)
{
    requireInit(h);
    if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtered signals not yet computed"); }
    if (group < 0 || group >= static_cast<int>(h->groups.size())) {
        fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " out of range");
    }
    const auto &gr = h->groups[group];
    return reduceRows(h, gr.rows, gr.nOut, absval, dDet, dId, cap);
}

// The same over an explicit row list of common length len.
long long reduceRows(mf_handle *h, const std::vector<int> &rows, long long len, int absval, void *dDet, int *dId,
                     long long cap)
{
    if (cap < len) {
        fail(MF_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(cap) + " < " + std::to_string(len) + " lags");
    }
    std::vector<const void *> ptrs(rows.size());
    for (size_t k = 0; k < rows.size(); ++k) {
        ptrs[k] = static_cast<const char *>(h->dRho.p) + h->rows[rows[k]].rhoOff * h->tsize;
    }
    growTo(h->dReduceRows, ptrs.size() * sizeof(void *));
    MF_HIP(hipStreamSynchronize(h->stream));  // the table may be in use by an earlier reduce
    MF_HIP(hipMemcpy(h->dReduceRows.p, ptrs.data(), ptrs.size() * sizeof(void *), hipMemcpyHostToDevice));
    MF_HIP(mfk::launch_reduce(h->precision, static_cast<const void *const *>(h->dReduceRows.p),
                              static_cast<int>(ptrs.size()), len, absval ? 1 : 0, dDet, dId, h->stream));
    return len;
}

// group-local row index -> the handle's row (template) index
void groupRowsToTemplates(const mf_handle *h, int group, int *id, long long n)
{
    const auto &rows = h->groups[group].rows;
    for (long long k = 0; k < n; ++k) { id[k] = rows[id[k]]; }
}

}  // namespace

int mf_reduce_group(mf_handle *h, int group, int absoluteValue, void *det, int *id, int cap, int *len)
{
    return guard([&] {
        requireInit(h);
        if (group < 0 || group >= static_cast<int>(h->groups.size())) {
            fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " out of range");
        }
        const long long need = h->groups[group].nOut;
        if (len) { *len = static_cast<int>(need); }  // also when the outputs are missing (size query)
        if (det == nullptr || id == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        growTo(h->dRedDet, static_cast<size_t>(std::max(need, 1LL)) * h->tsize);
        growTo(h->dRedId, static_cast<size_t>(std::max(need, 1LL)) * sizeof(int));
        const long long n = reduceGroup(h, group, absoluteValue, h->dRedDet.p, static_cast<int *>(h->dRedId.p), cap);
        MF_HIP(hipMemcpyAsync(det, h->dRedDet.p, static_cast<size_t>(n) * h->tsize, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipMemcpyAsync(id, h->dRedId.p, static_cast<size_t>(n) * sizeof(int), hipMemcpyDeviceToHost,
                              h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        groupRowsToTemplates(h, group, id, n);
        if (len) { *len = static_cast<int>(n); }
    });
}

int mf_reduce_group_device(mf_handle *h, int group, int absoluteValue, void *dDet, int *dId, int cap, int *len)
{
    return guard([&] {
        if (dDet == nullptr || dId == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        const long long n = reduceGroup(h, group, absoluteValue, dDet, dId, cap);
        MF_HIP(hipStreamSynchronize(h->stream));
        if (len) { *len = static_cast<int>(n); }
    });
}

int mf_detect_reduced(mf_handle *h, int group, int absoluteValue, double threshold, int minDist, int *idx,
                      int *templateIds, int cap, int *nPeaks)
{
    return guard([&] {
        if (nPeaks == nullptr || ((idx == nullptr || templateIds == nullptr) && cap > 0)) {
            fail(MF_ERR_INVALID_ARGUMENT, "output is NULL");
        }
        requireInit(h);
        if (group < 0 || group >= static_cast<int>(h->groups.size())) {
            fail(MF_ERR_INVALID_ARGUMENT, "group " + std::to_string(group) + " out of range");
        }
        const long long need = h->groups[group].nOut;
        growTo(h->dRedDet, static_cast<size_t>(std::max(need, 1LL)) * h->tsize);
        growTo(h->dRedId, static_cast<size_t>(std::max(need, 1LL)) * sizeof(int));
        const long long n = reduceGroup(h, group, absoluteValue, h->dRedDet.p, static_cast<int *>(h->dRedId.p), need);
        MF_HIP(hipStreamSynchronize(h->stream));
        peaks(h->dRedDet.p, true, static_cast<int>(n), h->precision, threshold, minDist, idx, cap, nPeaks);
        const int np = *nPeaks;
        if (np > 0) {
            growTo(h->dRedPk, static_cast<size_t>(np) * sizeof(int));
            growTo(h->dRedPkId, static_cast<size_t>(np) * sizeof(int));
            MF_HIP(hipMemcpyAsync(h->dRedPk.p, idx, static_cast<size_t>(np) * sizeof(int), hipMemcpyHostToDevice,
                                  h->stream));
            MF_HIP(mfk::launch_gather_int(static_cast<const int *>(h->dRedId.p), static_cast<const int *>(h->dRedPk.p),
                                          np, static_cast<int *>(h->dRedPkId.p), h->stream));
            MF_HIP(hipMemcpyAsync(templateIds, h->dRedPkId.p, static_cast<size_t>(np) * sizeof(int),
                                  hipMemcpyDeviceToHost, h->stream));
            MF_HIP(hipStreamSynchronize(h->stream));
            groupRowsToTemplates(h, group, templateIds, np);
        }
    });
}

namespace {

// Detector::detect's SINGLE policy for template row it: the reduce kernel
// over the one row (clamp), then the standalone peak finder.
void detectOneRow(mf_handle *h, int it, int absval, double threshold, int minDist, int *idx, int cap, int *nPeaks)
{
    const long long need = h->rows[it].nOut;
    growTo(h->dRedDet, static_cast<size_t>(std::max(need, 1LL)) * h->tsize);
    growTo(h->dRedId, static_cast<size_t>(std::max(need, 1LL)) * sizeof(int));
    const long long n = reduceRows(h, std::vector<int>{it}, need, absval, h->dRedDet.p,
                                   static_cast<int *>(h->dRedId.p), need);
    MF_HIP(hipStreamSynchronize(h->stream));
    peaks(h->dRedDet.p, true, static_cast<int>(n), h->precision, threshold, minDist, idx, cap, nPeaks);
}

}  // namespace

extern "C++" {
namespace {

// Detector::detect's SINGLE policy for every template row: clamp(rho) or
// clamp(|rho|), threshold + local maximum (one launch per chunk of rows),
// the candidates packed and read back together, then the reference
// pruning per row on the host.  Each row gets room for kRowCands
// candidates (rows are usually far below); a row that overflows is redone
// alone with the full nOut/2 + 2.  Rows go in chunks of up to 256.
template <typename T>
void detectRowsT(mf_handle *h, int absval, double threshold, int minDist, int *idx, int cap, int *offsets)
{
    const int nT = static_cast<int>(h->rows.size());
    long long maxOut = 1;
    for (const auto &rp : h->rows) { maxOut = std::max<long long>(maxOut, rp.nOut); }
    constexpr size_t kRowCands = 65536;
    const size_t capR = std::min(kRowCands, static_cast<size_t>(maxOut / 2 + 2));
    const int chunk = std::min(nT, 256);
    growTo(h->dDetGroupTab, chunk * sizeof(mfk::StackGroup));
    growTo(h->dDetCounts, chunk * sizeof(int));
    growTo(h->dDetCandIdx, chunk * capR * sizeof(int));
    growTo(h->dDetCandVal, chunk * capR * sizeof(T));
    growTo(h->dDetPackIdx, (chunk + chunk * capR) * sizeof(int));
    growTo(h->dDetPackVal, chunk * capR * sizeof(T));
    constexpr size_t kPrefix = 8192;
    std::vector<mfk::StackGroup> tab(chunk);
    int off = 0;
    for (int r0 = 0; r0 < nT; r0 += chunk) {
        const int nG = std::min(chunk, nT - r0);
        long long len = 0;
        for (int g = 0; g < nG; ++g) {
            const auto &rp = h->rows[r0 + g];
            auto &t = tab[g];
            t.rows = nullptr;
            t.nRows = 0;
            t.candCap = static_cast<int>(capR);
            t.len = rp.nOut;
            t.out = static_cast<char *>(h->dRho.p) + rp.rhoOff * h->tsize;
            t.count = h->dDetCounts.as<int>() + g;
            t.candIdx = h->dDetCandIdx.as<int>() + g * capR;
            t.candVal = static_cast<char *>(h->dDetCandVal.p) + g * capR * sizeof(T);
            len = std::max<long long>(len, rp.nOut);
        }
        // tab is not touched again before this chunk's wait below
        MF_HIP(hipMemcpyAsync(h->dDetGroupTab.p, tab.data(), nG * sizeof(mfk::StackGroup), hipMemcpyHostToDevice,
                              h->stream));
        MF_HIP(hipMemsetAsync(h->dDetCounts.p, 0, nG * sizeof(int), h->stream));
        const auto *dTab = h->dDetGroupTab.as<mfk::StackGroup>();
        MF_HIP(mfk::launch_peak_rows(h->precision, dTab, nG, len, threshold, absval, h->stream));
        MF_HIP(mfk::launch_pack_candidates(h->precision, dTab, nG, h->dDetPackIdx.as<int>(), h->dDetPackVal.p,
                                           h->stream));
        const size_t pre = std::min(kPrefix, capR * static_cast<size_t>(nG));
        const size_t hdrBytes = (static_cast<size_t>(nG) + pre) * sizeof(int);
        h->hDetPack.reserve(hdrBytes + pre * sizeof(T));
        int *hIdx = static_cast<int *>(h->hDetPack.p);
        T *hVal = reinterpret_cast<T *>(static_cast<char *>(h->hDetPack.p) + hdrBytes);
        MF_HIP(hipMemcpyAsync(hIdx, h->dDetPackIdx.p, hdrBytes, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipMemcpyAsync(hVal, h->dDetPackVal.p, pre * sizeof(T), hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        std::vector<size_t> first(nG + 1, 0);
        std::vector<char> overflow(nG, 0);
        for (int g = 0; g < nG; ++g) {
            overflow[g] = hIdx[g] > static_cast<int>(capR);
            // an overflowed row packed its first capR candidates
            first[g + 1] = first[g] + std::min(static_cast<size_t>(hIdx[g]), capR);
        }
        const size_t total = first[nG];
        std::vector<int> ci(total);
        std::vector<T> cv(total);
        std::copy(hIdx + nG, hIdx + nG + std::min(total, pre), ci.begin());
        std::copy(hVal, hVal + std::min(total, pre), cv.begin());
        if (total > pre) {
            MF_HIP(hipMemcpyAsync(ci.data() + pre, h->dDetPackIdx.as<int>() + nG + pre, (total - pre) * sizeof(int),
                                  hipMemcpyDeviceToHost, h->stream));
            MF_HIP(hipMemcpyAsync(cv.data() + pre, static_cast<const T *>(h->dDetPackVal.p) + pre,
                                  (total - pre) * sizeof(T), hipMemcpyDeviceToHost, h->stream));
            MF_HIP(hipStreamSynchronize(h->stream));
        }
        for (int g = 0; g < nG; ++g) {
            offsets[r0 + g] = off;
            if (overflow[g]) {
                int np = 0;
                detectOneRow(h, r0 + g, absval, threshold, minDist, idx + off, cap - off, &np);
                off += np;
                continue;
            }
            const size_t cnt = first[g + 1] - first[g];
            std::vector<std::pair<T, int>> cand(cnt);
            for (size_t i = 0; i < cnt; ++i) { cand[i] = {cv[first[g] + i], ci[first[g] + i]}; }
            int np = 0;
            pruneCandidates<T>(cand, minDist, idx + off, cap - off, &np);
            off += np;
        }
    }
    offsets[nT] = off;
}

}  // namespace
}  // extern "C++"

int mf_detect_templates(mf_handle *h, int absoluteValue, double threshold, int minDist, int *idx, int cap,
                        int *offsets)
{
    return guard([&] {
        if (offsets == nullptr || (idx == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        requireInit(h);
        if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtered signals not yet computed"); }
        if (threshold < 0 || threshold > 1) {
            fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
        }
        if (h->precision == MF_PRECISION_FLOAT) {
            detectRowsT<float>(h, absoluteValue ? 1 : 0, threshold, minDist, idx, cap, offsets);
        } else {
            detectRowsT<double>(h, absoluteValue ? 1 : 0, threshold, minDist, idx, cap, offsets);
        }
    });
}

int mf_detect_template(mf_handle *h, int it, int absoluteValue, double threshold, int minDist, int *idx, int cap,
                       int *nPeaks)
{
    return guard([&] {
        if (nPeaks == nullptr || (idx == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        requireInit(h);
        if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtered signals not yet computed"); }
        checkRow(h, it);
        detectOneRow(h, it, absoluteValue, threshold, minDist, idx, cap, nPeaks);
    });
}

int mf_detection_products(mf_handle *h, int nPeaks, const int *peakIdx, const int *templateIds,
                          const double *samplingPeriods, int absoluteValue, int interpolation, int lanczosAlpha,
                          int lanczosPoints, double *rho, double *shift, double *alphaGR, double *alphaSR,
                          int *flags)
{
    return guard([&] {
        requireInit(h);
        if (!h->applied) { fail(MF_ERR_BAD_STATE, "Matched filtered signals not yet computed"); }
        if (nPeaks < 0) { fail(MF_ERR_INVALID_ARGUMENT, "nPeaks must be non-negative"); }
        if (nPeaks > 0 && (peakIdx == nullptr || templateIds == nullptr)) {
            fail(MF_ERR_INVALID_ARGUMENT, "peak list is NULL");
        }
        if (interpolation != mfk::kInterpLanczos && interpolation != mfk::kInterpQuadratic &&
            interpolation != mfk::kInterpNone) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown interpolation type " + std::to_string(interpolation));
        }
        // DetectionTimeInterpolationParameters' own checks
        if (interpolation == mfk::kInterpLanczos) {
            if (lanczosAlpha < 1) { fail(MF_ERR_INVALID_ARGUMENT, "alpha must be positive"); }
            if (lanczosPoints < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Number of interpolation points must be positive"); }
            if ((2LL * lanczosAlpha + 1) * (lanczosPoints - 1) > mfk::kLanczosMaxTerms) {
                fail(MF_ERR_UNSUPPORTED, "(2 alpha + 1) x (points - 1) = " +
                                             std::to_string((2LL * lanczosAlpha + 1) * (lanczosPoints - 1)) +
                                             " interpolation terms per detection exceed " +
                                             std::to_string(mfk::kLanczosMaxTerms));
            }
        }
        if (nPeaks == 0) { return; }
        const size_t jobBytes = static_cast<size_t>(nPeaks) * sizeof(mfk::ProductJob);
        const size_t outBytes = static_cast<size_t>(nPeaks) * sizeof(mfk::ProductOut);
        MF_HIP(hipStreamSynchronize(h->stream));  // the staging and job table may be in use by an earlier call
        h->hProd.reserve(std::max(jobBytes, outBytes));
        auto *jobs = static_cast<mfk::ProductJob *>(h->hProd.p);
        for (int k = 0; k < nPeaks; ++k) {
            const int it = templateIds[k];
            checkRow(h, it);
            const auto &rp = h->rows[it];
            if (peakIdx[k] < 0 || peakIdx[k] >= rp.nOut) {
                fail(MF_ERR_INVALID_ARGUMENT, "peak " + std::to_string(peakIdx[k]) + " outside [0," +
                                                  std::to_string(rp.nOut - 1) + "]");
            }
            const double dt = samplingPeriods ? samplingPeriods[it] : 1.0;
            if (!(dt > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "sampling period must be positive"); }
            auto &j = jobs[k];
            j.rho = static_cast<const char *>(h->dRho.p) + rp.rhoOff * h->tsize;
            j.sig = h->ch[rp.ch].buf.p;
            j.tmpl = h->dTmplRaw.as<double>() + rp.tmplOff;
            j.sigLen = h->ch[rp.ch].n;
            j.dt = dt;
            j.nOut = rp.nOut;
            j.peak = peakIdx[k];
            j.Lt = rp.Lt;
            j.pad = 0;
        }
        growTo(h->dProdJobs, jobBytes);
        growTo(h->dProdOut, outBytes);
        MF_HIP(hipMemcpyAsync(h->dProdJobs.p, jobs, jobBytes, hipMemcpyHostToDevice, h->stream));
        MF_HIP(mfk::launch_products(h->precision, h->dProdJobs.as<mfk::ProductJob>(), nPeaks, absoluteValue ? 1 : 0,
                                    interpolation, lanczosAlpha, lanczosPoints, h->dProdOut.as<mfk::ProductOut>(),
                                    h->stream));
        // the results reuse the staging buffer: the copy above is ordered before them on the stream
        auto *res = static_cast<mfk::ProductOut *>(h->hProd.p);
        MF_HIP(hipMemcpyAsync(res, h->dProdOut.p, outBytes, hipMemcpyDeviceToHost, h->stream));
        MF_HIP(hipStreamSynchronize(h->stream));
        for (int k = 0; k < nPeaks; ++k) {
            if (rho) { rho[k] = res[k].rho; }
            if (shift) { shift[k] = res[k].shift; }
            if (alphaGR) { alphaGR[k] = res[k].alphaGR; }
            if (alphaSR) { alphaSR[k] = res[k].alphaSR; }
            if (flags) { flags[k] = res[k].flags; }
        }
    });
}

int mf_get_device(const mf_handle *h, int *device)
{
    return guard([&] {
        requireHandle(h);
        if (device == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        *device = h->device;
    });
}

int mf_get_number_of_groups(const mf_handle *h, int *n)
{
    return guard([&] {
        requireInit(h);
        if (n) { *n = static_cast<int>(h->groups.size()); }
    });
}

int mf_template_spectrum(mf_handle *h, int row, void *outInterleaved, int cap, int *len)
{
    return guard([&] {
        requireInit(h);
        checkRow(h, row);
        const auto &rp = h->rows[row];
        const int M = 1 << rp.logM;
        const int N = 2 * M;
        if (len) { *len = M + 1; }
        if (outInterleaved == nullptr || cap < M + 1) { fail(MF_ERR_INVALID_ARGUMENT, "output buffer too small"); }
        std::vector<char> raw((M + 1) * 2 * h->tsize);
        MF_HIP(hipMemcpy(raw.data(), static_cast<const char *>(h->dSpec.p) + rp.specOff * 2 * h->tsize, raw.size(),
                         hipMemcpyDeviceToHost));
        // B[k] = exp(-2 pi i k (Lt-1)/N) * 2N * H[k]  (spectrum of the reversed template)
        for (int k = 0; k <= M; ++k) {
            double hr, hi;
            if (h->precision == MF_PRECISION_FLOAT) {
                hr = reinterpret_cast<const float *>(raw.data())[2 * k];
                hi = reinterpret_cast<const float *>(raw.data())[2 * k + 1];
            } else {
                hr = reinterpret_cast<const double *>(raw.data())[2 * k];
                hi = reinterpret_cast<const double *>(raw.data())[2 * k + 1];
            }
            const double ang = -2.0 * M_PI * static_cast<double>((static_cast<long long>(k) * (rp.Lt - 1)) % N) / N;
            const double c = std::cos(ang), s = std::sin(ang);
            const double br = 2.0 * N * (hr * c - hi * s);
            const double bi = 2.0 * N * (hr * s + hi * c);
            if (h->precision == MF_PRECISION_FLOAT) {
                static_cast<float *>(outInterleaved)[2 * k] = static_cast<float>(br);
                static_cast<float *>(outInterleaved)[2 * k + 1] = static_cast<float>(bi);
            } else {
                static_cast<double *>(outInterleaved)[2 * k] = br;
                static_cast<double *>(outInterleaved)[2 * k + 1] = bi;
            }
        }
    });
}

int mf_amplitude_factors(int precision, int Lt, const double *tmpl, int nWindows, const void *windows,
                         long long ld, double *alphaGR, double *alphaSR, int *computed)
{
    return guard([&] {
        if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(precision));
        }
        if (Lt < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Template signal length = " + std::to_string(Lt) + " must be positive"); }
        if (nWindows < 0 || ld < Lt) { fail(MF_ERR_INVALID_ARGUMENT, "nWindows must be >= 0 and ld >= Lt"); }
        if (tmpl == nullptr || (nWindows > 0 && windows == nullptr)) { fail(MF_ERR_INVALID_ARGUMENT, "input is NULL"); }
        if (nWindows == 0) { return; }
        requireDevice();
        const size_t ts = precision == MF_PRECISION_FLOAT ? 4 : 8;
        const size_t winElems = static_cast<size_t>(nWindows - 1) * static_cast<size_t>(ld) + static_cast<size_t>(Lt);
        // a process-wide workspace, grown on demand (one window per call is
        // the common case: RelativeMagnitude scores each window as it is set)
        std::lock_guard<std::mutex> lock(g_amplitudeMutex);
        AmplitudeWork &w = amplitudeWork();
        if (w.stream == nullptr) { MF_HIP(hipStreamCreate(&w.stream)); }
        growTo(w.tmpl, static_cast<size_t>(Lt) * sizeof(double));
        growTo(w.windows, winElems * ts);
        growTo(w.out, static_cast<size_t>(nWindows) * sizeof(mfk::ProductOut));
        std::vector<mfk::ProductOut> res(nWindows);
        MF_HIP(hipMemcpyAsync(w.tmpl.p, tmpl, static_cast<size_t>(Lt) * sizeof(double), hipMemcpyHostToDevice,
                              w.stream));
        MF_HIP(hipMemcpyAsync(w.windows.p, windows, winElems * ts, hipMemcpyHostToDevice, w.stream));
        MF_HIP(mfk::launch_amplitude(precision, w.tmpl.as<double>(), Lt, w.windows.p, ld, nWindows,
                                     w.out.as<mfk::ProductOut>(), w.stream));
        MF_HIP(hipMemcpyAsync(res.data(), w.out.p, res.size() * sizeof(mfk::ProductOut), hipMemcpyDeviceToHost,
                              w.stream));
        MF_HIP(hipStreamSynchronize(w.stream));
        for (int k = 0; k < nWindows; ++k) {
            if (alphaGR) { alphaGR[k] = res[k].alphaGR; }
            if (alphaSR) { alphaSR[k] = res[k].alphaSR; }
            if (computed) { computed[k] = res[k].flags & mfk::kProductAmplitude ? 1 : 0; }
        }
    });
}

int mf_peaks(const void *x, int n, int precision, double threshold, int minDist, int *idx, int cap, int *nPeaks)
{
    return guard([&] { peaks(x, false, n, precision, threshold, minDist, idx, cap, nPeaks); });
}

int mf_peaks_device(const void *dx, int n, int precision, double threshold, int minDist, int *idx, int cap,
                    int *nPeaks)
{
    return guard([&] { peaks(dx, true, n, precision, threshold, minDist, idx, cap, nPeaks); });
}

int mf_peaks_device_noalloc(const void *dx, int n, int precision, double threshold, int minDist, int *idx, int cap,
                            int *nPeaks)
{
    return guard([&] {
        if (dx == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "x is NULL"); }
        if (n < 1) { fail(MF_ERR_INVALID_ARGUMENT, "No points in x"); }
        if (nPeaks == nullptr || (idx == nullptr && cap > 0)) { fail(MF_ERR_INVALID_ARGUMENT, "peaks is NULL"); }
        if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown precision");
        }
        if (threshold < 0 || threshold > 1) {
            fail(MF_ERR_INVALID_ARGUMENT, "Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
        }
        if (n > kNoAllocMaxSamples) {
            fail(MF_ERR_INVALID_ARGUMENT, "n = " + std::to_string(n) + " exceeds the scratch limit of " +
                                              std::to_string(kNoAllocMaxSamples) +
                                              " samples (use mf_peaks_device)");
        }
        std::lock_guard<std::mutex> lock(g_noAllocMutex);
        if (precision == MF_PRECISION_FLOAT) {
            peaksNoAlloc<float>(dx, n, static_cast<float>(threshold), minDist, idx, cap, nPeaks);
        } else {
            peaksNoAlloc<double>(dx, n, threshold, minDist, idx, cap, nPeaks);
        }
    });
}

int mf_window_scale_device(const void *dx, long long n, int Lt, int scanPrecision, void *dScale, double *ms)
{
    return guard([&] {
        if (dx == nullptr || dScale == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "device pointer is NULL"); }
        if (scanPrecision != mfk::kScanF64 && scanPrecision != mfk::kScanF32c) {
            fail(MF_ERR_INVALID_ARGUMENT, "scanPrecision must be 0 (FP64) or 1 (compensated FP32)");
        }
        if (Lt < 2 || Lt > mfk::kScanMaxLt) {
            fail(MF_ERR_INVALID_ARGUMENT, "template length must be in [2, " + std::to_string(mfk::kScanMaxLt) + "]");
        }
        if (n < Lt) { fail(MF_ERR_INVALID_ARGUMENT, "signal is shorter than the template"); }
        requireDevice();
        // a blocking stream: work the caller queued on the default stream
        // (producing dx, say) completes before the scan reads it
        hipStream_t st = nullptr;
        MF_HIP(hipStreamCreate(&st));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (e0) { (void)hipEventDestroy(e0); }
            if (e1) { (void)hipEventDestroy(e1); }
            (void)hipStreamDestroy(st);
        };
        try {
            MF_HIP(hipEventCreate(&e0));
            MF_HIP(hipEventCreate(&e1));
            MF_HIP(hipEventRecord(e0, st));
            MF_HIP(mfk::launch_scan_scale(scanPrecision, static_cast<const float *>(dx), n, Lt,
                                          static_cast<float *>(dScale), st));
            MF_HIP(hipEventRecord(e1, st));
            MF_HIP(hipEventSynchronize(e1));
            float t = 0;
            MF_HIP(hipEventElapsedTime(&t, e0, e1));
            if (ms) { *ms = t; }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int mf_compute_optimal_fft_and_block_length(int nb, int nx, int *nfft, int *blockLength)
{
    return guard([&] {
        if (nfft == nullptr || blockLength == nullptr) { fail(MF_ERR_INVALID_ARGUMENT, "output is NULL"); }
        referenceWindowLength(nb, nx, nfft, blockLength);
    });
}

int mf_autotune(int precision, double *table, int cap, int *n)
{
    return guard([&] {
        if (precision != MF_PRECISION_FLOAT && precision != MF_PRECISION_DOUBLE) {
            fail(MF_ERR_INVALID_ARGUMENT, "unknown precision " + std::to_string(precision));
        }
        requireDevice();
        std::vector<double> out;
        calibrate(precision, out);
        const int rows = static_cast<int>(out.size() / 5);
        if (n) { *n = rows; }
        if (table != nullptr) {
            std::copy(out.begin(), out.begin() + std::min<size_t>(out.size(), static_cast<size_t>(std::max(cap, 0)) * 5),
                      table);
        }
    });
}

int mf_get_window_stats(mf_handle *h, long long *windows, long long *flagged, long long *full, long long *solo)
{
    return guard([&] {
        requireHandle(h);
        requireInit(h);
        unsigned long long nf[2] = {0, 0};
        MF_HIP(hipStreamSynchronize(h->stream));
        MF_HIP(hipMemcpy(nf, h->dFlagged.p, sizeof(nf), hipMemcpyDeviceToHost));
        if (windows) { *windows = h->winPerApply * h->statApplies; }
        if (flagged) { *flagged = static_cast<long long>(nf[0]); }
        if (full) { *full = static_cast<long long>(nf[1]); }
        if (solo) { *solo = h->soloPerApply * h->statApplies; }
    });
}

int mf_get_timings(mf_handle *h, double *corrMs, double *stackMs)
{
    return guard([&] {
        requireHandle(h);
        if (h->applied && hipEventQuery(h->evB) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, h->evA, h->evB) == hipSuccess) { h->corrMs = ms; }
        }
        if (corrMs) { *corrMs = h->corrMs; }
        if (stackMs) { *stackMs = h->stackMs; }
    });
}

}  // extern "C"
