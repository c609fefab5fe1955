// Time-domain correlation of short templates on MFMA (the optional
// short-template path; no FFT).
//
// For one lane (a channel and its rows, every template at most KT taps)
// the numerators form a Hankel-by-template GEMM
//     D[j][i] = sum_k t~_j[k] * x~[i + k],
// t~_j the demeaned unit-norm template j (zero past its length) and
// x~ = (x - a) / s the block's samples with the block anchor a removed and
// scaled by their RMS s.  Both operands are split into 16-bit parts:
//  - FP16 (11-bit significand): hi = round(v), lo = round(v - hi); the
//    products hi*hi, hi*lo and lo*hi;
//  - BF16 (8-bit significand): hi, mid = round(v - hi),
//    lo = round(v - hi - mid); the six products hi*hi, hi*mid, mid*hi,
//    hi*lo, lo*hi and mid*mid (the dropped ones, mid*lo, lo*mid and lo*lo,
//    are below 2^-24 of |t||x|).
// v_mfma_f32_32x32x16 accumulates them in FP32: hi*hi in its own
// accumulator, the smaller products in a second one.
// rho = s * D * scale with the window normalisation from FP64 prefix sums
// of the block.
//
// Work item: IS = 512 consecutive outputs of the lane and up to 128 rows.
// Each of the 4 waves owns 32 rows: their split taps stay in registers
// (weight-stationary, parts * KT / 4 VGPRs) while the wave walks the 16
// output tiles of 32; the Hankel operand is read from LDS, where entry p
// holds the 8 samples x~[p .. p+7] (one 16-byte read per part and k-step).
//
// Accuracy: per window |d rho| <= eps_m * sqrt(Lt S2) / den with
// eps_m = eps_split + (4 sqrt(KT/16) + 4) 2^-24 (FP32 accumulation and the
// FP32 epilogue on the FP64 window sums), S2 the
// window's sum of (x - a)^2 and den its centred norm; eps_split is the
// relative size of the dropped products and split residuals: 3 * 2^-22 for
// the FP16 pair (3 u^2, u = 2^-11) and 3 * 2^-24 for the BF16 triple
// (3 u^3, u = 2^-8).  Windows whose bound exceeds the FP32 tolerance get
// their numerator recomputed in FP64 from the anchored samples (the window
// norm from the prefix sums is accurate there); windows whose FP64
// variance lost too many digits get the exact two-pass Pearson value, as
// in the FFT path.
#include "mf_device.hpp"
#include "mf_kernels.hpp"
#include <cfloat>

using namespace mfdev;

namespace mfk {

namespace {

constexpr int NT = kThreads;
constexpr int IS = kTdOutputsPerItem;  // outputs per work item
constexpr int NTILE_I = IS / 32;       // output tiles of 32 per item

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

template <typename MT>
struct TdTraits;
template <>
struct TdTraits<_Float16> {
    using vec8 = f16x8;
    static constexpr int parts = 2;
    __device__ static f32x16 mfma(vec8 a, vec8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct TdTraits<__bf16> {
    using vec8 = bf16x8;
    static constexpr int parts = 3;
    __device__ static f32x16 mfma(vec8 a, vec8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

constexpr int kTdFixList = 64 + 16 * 64;  // list capacity of a wave's flagged windows
constexpr int kTdCoopFlags = 4;           // at most this many left at the end: the wave shares each window

// Exact FP64 value of one flagged window at item offset i, over taps
// m = m0, m0 + step, ... (one lane: 0, 1; a whole wave: lane, 64, summed
// across the wave).
//  - cancelled variance (canc): one pass over d = x - x0, x0 the window's
//    first sample (every d is exact in FP64), giving
//    rho = sum t d / sqrt(sum d^2 - (sum d)^2 / Lt); the template has zero
//    mean, so sum t d is the centred numerator, and with the shift inside
//    the window the variance keeps its digits (a constant window gives 0);
//  - noise flag: the numerator sum t (x - a) from the anchored samples with
//    the window norm from the block's FP64 prefix sums, accurate there.
// |rho| > 1.1 gives 0 (the reference's tooBig rule).
__device__ double exact_window_td(const float *raw, const double *P1, const double *P2, int i,
                                  gptr<const double> t, int Lt, double a, bool canc, int m0, int step)
{
    const double shift = canc ? static_cast<double>(raw[i]) : a;
    double s1 = 0, s2 = 0, dn = 0;
#pragma unroll 8
    for (int m = m0; m < Lt; m += step) {
        const double d = static_cast<double>(raw[i + m]) - shift;
        s1 += d;
        s2 = fma(d, d, s2);
        dn = fma(t[m], d, dn);
    }
    if (step > 1) {
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        dn = wave_sum(dn);
    }
    double rho;
    if (canc) {
        rho = dn / fmax(DBL_EPSILON, sqrt(fmax(0.0, s2 - s1 * s1 / Lt)));
    } else {
        const double S1 = P1[i + Lt] - P1[i];
        const double den2 = Lt * (P2[i + Lt] - P2[i]) - S1 * S1;
        rho = dn * sqrt(Lt / den2);
    }
    return fabs(rho) > 1.1 ? 0.0 : rho;
}

// Blocks per CU the register budget is capped for: two while the split taps
// (parts * KT / 4 VGPRs) leave room for the accumulators and the epilogue
// in 256 VGPRs, else one (the BF16 triple at KT = 256: 192 VGPRs of taps).
template <typename MT, int KT>
constexpr int td_blocks_per_cu() { return TdTraits<MT>::parts * KT / 4 <= 128 ? 2 : 1; }

template <typename MT, int KT>
__global__ __launch_bounds__(NT, (td_blocks_per_cu<MT, KT>())) void tdcorr_kernel(const TdItem *__restrict__ items,
                                                       const Seg *__restrict__ segs,
                                                       const TdRow *__restrict__ rows, FixQueue queue)
{
    using Tr = TdTraits<MT>;
    using V8 = typename Tr::vec8;
    constexpr int NP = Tr::parts;
    constexpr int NS = KT / 16;        // k-steps of 16
    constexpr int SPAN = IS + KT;      // samples one item reads
    constexpr int SPT = (SPAN + NT - 1) / NT;
    __shared__ __align__(16) MT x8[NP][SPAN][8];  // entry p of part q: x~[p .. p+7]
    __shared__ float raw[SPAN];                   // samples (FP32), for exact windows
    __shared__ double P1[SPAN + 1];               // prefix sums of d = x - a
    __shared__ double P2[SPAN + 1];               //   and of d^2
    __shared__ double red[NT / 64];
    __shared__ double wsum1[NT / 64], wsum2[NT / 64];
    // flagged windows of each wave's tiles: bit r of the low half flags the
    // lane's window r, bit r of the high half marks a cancelled variance
    __shared__ unsigned fq[NT / 64][NTILE_I][64];
    // bit p: sample p equals sample p + 1 (exactly constant windows)
    __shared__ unsigned eqw[(SPAN + 31) / 32];
    // each wave's list of flagged windows to recompute (under 64 carried
    // over plus one tile's 1024)
    __shared__ unsigned short fixList[NT / 64][kTdFixList];

    const TdItem item = items[blockIdx.x];
    const Seg sg = segs[item.seg];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar row loads and branches
    const int lane = tid & 63;
    const int half = lane >> 5;
    const int col = lane & 31;
    const long long base = item.seg0;

    if (sg.dead) {
        for (int r = 0; r < item.nRows; ++r) {
            const TdRow row = rows[item.rowBegin + r];
            const int nv = static_cast<int>(min(static_cast<long long>(IS), row.nOut - base));
            gptr<float> out = to_global(static_cast<float *>(row.out)) + base;
            for (int i = tid; i < nv; i += NT) { out[i] = 0.0f; }
        }
        return;
    }

    // 1. samples, anchor and scale
    const gptr<const float> x = to_global(static_cast<const float *>(sg.x));
    const int nSamp = static_cast<int>(min(static_cast<long long>(SPAN), sg.n - base));
    float v[SPT];
    double s = 0;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;  // contiguous run per thread (for the scan)
        const float val = (p < nSamp) ? x[base + p] : 0.0f;
        v[q] = val;
        if (p < SPAN) { raw[p] = val; }
        s += static_cast<double>(val);
    }
    const double a = block_sum<NT>(s, red) / nSamp;
    double e = 0, l1 = 0, l2 = 0;
    double d1[SPT], d2[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        const double d = (p < nSamp) ? static_cast<double>(v[q]) - a : 0.0;
        d1[q] = l1;  // exclusive local prefix
        d2[q] = l2;
        l1 += d;
        l2 += d * d;
        e += d * d;
    }
    const double energy = block_sum<NT>(e, red);
    const double sc = energy > 0 ? sqrt(energy / nSamp) : 1.0;
    const float inv = static_cast<float>(1.0 / sc);
    const float scF = static_cast<float>(sc);
    const float flK = static_cast<float>(1.0e-12 * energy);  // times Lt: the cancellation floor
    // 2. split samples into the 8-wide Hankel entries: sample p goes to
    //    entry p - k, slot k (k = 0..7), of every part.  The FP32 value
    //    xs is the sum of its parts exactly (each remainder is exact in
    //    FP32 and the parts together hold at least 24 significant bits).
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        if (p < SPAN) {
            float rest = (p < nSamp) ? static_cast<float>((static_cast<double>(v[q]) - a) * inv) : 0.0f;
#pragma unroll
            for (int pt = 0; pt < NP; ++pt) {
                const MT part = static_cast<MT>(rest);
                rest -= static_cast<float>(part);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (p - k >= 0) { x8[pt][p - k][k] = part; }
                }
            }
        }
    }
    // entries past the span read zeros in their upper slots
    for (int p = SPAN - 7 + tid; p < SPAN; p += NT) {
        for (int k = SPAN - p; k < 8; ++k) {
#pragma unroll
            for (int pt = 0; pt < NP; ++pt) { x8[pt][p][k] = static_cast<MT>(0.0f); }
        }
    }
    // 3. FP64 prefix sums P[p] = sum_{q < p} d_q (thread runs, then waves)
    double x1 = l1, x2 = l2;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y1 = __shfl_up(x1, o, 64);
        const double y2 = __shfl_up(x2, o, 64);
        if (lane >= o) {
            x1 += y1;
            x2 += y2;
        }
    }
    if (lane == 63) {
        wsum1[wave] = x1;
        wsum2[wave] = x2;
    }
    __syncthreads();
    if (tid < (SPAN + 31) / 32) {
        unsigned bits = 0;
        for (int b = 0; b < 32; ++b) {
            const int p = 32 * tid + b;
            if (p + 1 < SPAN && raw[p] == raw[p + 1]) { bits |= 1u << b; }
        }
        eqw[tid] = bits;
    }
    double w1 = 0, w2 = 0;
    for (int w = 0; w < wave; ++w) {
        w1 += wsum1[w];
        w2 += wsum2[w];
    }
    const double e1 = w1 + x1 - l1, e2 = w2 + x2 - l2;  // exclusive prefix of this thread's run
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        if (p < SPAN) {
            P1[p] = e1 + d1[q];
            P2[p] = e2 + d2[q];
        }
    }
    if (tid == NT - 1) {
        P1[SPAN] = w1 + x1;
        P2[SPAN] = w2 + x2;
    }

    // 4. this wave's 32 rows: split taps in registers (A operand: row =
    //    lane % 32, taps 16 s + 8 half + 0..7 of each part)
    const int rowBase = item.rowBegin + 32 * wave;
    const int nMine = min(32, item.nRows - 32 * wave);
    V8 ap[NP][NS];
    {
        const bool has = col < nMine;
        const MT *f = has ? static_cast<const MT *>(rows[rowBase + col].frag) : nullptr;
#pragma unroll
        for (int pt = 0; pt < NP; ++pt) {
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                if (has) {
                    const gptr<const i32x4> fh = to_global(reinterpret_cast<const i32x4 *>(f));
                    ap[pt][st] = __builtin_bit_cast(V8, fh[pt * 2 * NS + 2 * st + half]);
                } else {
                    ap[pt][st] = __builtin_bit_cast(V8, i32x4{0, 0, 0, 0});
                }
            }
        }
    }
    __syncthreads();  // Hankel entries, prefix sums and raw samples ready
    if (nMine <= 0) { return; }

    unsigned long long nFlag = 0, nCanc = 0;  // windows recomputed by this wave, in full
    unsigned tiles = 0;            // tiles with flagged windows (wave-uniform)

    // 5. output tiles: D[j][i], lane holds column i = col and rows
    //    j = 8 (r / 4) + 4 half + r % 4
    for (int tile = 0; tile < NTILE_I; ++tile) {
        const int i0 = tile * 32;
        f32x16 acc0 = {}, acc1 = {};
        const i32x4 *b0p = reinterpret_cast<const i32x4 *>(&x8[0][i0 + col + 8 * half][0]);
        const i32x4 *b1p = reinterpret_cast<const i32x4 *>(&x8[1][i0 + col + 8 * half][0]);
        const i32x4 *b2p = reinterpret_cast<const i32x4 *>(&x8[NP - 1][i0 + col + 8 * half][0]);
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const V8 b0 = __builtin_bit_cast(V8, b0p[16 * st]);  // entry + 16 st
            const V8 b1 = __builtin_bit_cast(V8, b1p[16 * st]);
            acc0 = Tr::mfma(ap[0][st], b0, acc0);
            acc1 = Tr::mfma(ap[0][st], b1, acc1);
            acc1 = Tr::mfma(ap[1][st], b0, acc1);
            if constexpr (NP == 3) {
                const V8 b2 = __builtin_bit_cast(V8, b2p[16 * st]);
                acc1 = Tr::mfma(ap[0][st], b2, acc1);
                acc1 = Tr::mfma(ap[NP - 1][st], b0, acc1);
                acc1 = Tr::mfma(ap[1][st], b1, acc1);
            }
        }
        // epilogue: normalise, flag, store; flagged windows (bit r of
        // flg, with bit r of cnc when the variance cancelled) are
        // recomputed below
        const int i = i0 + col;
        const double p1i = P1[i], p2i = P2[i];
        unsigned flg = 0, cnc = 0;
        // Branch-free over the rows: the two rows of step r (one per
        // half-wave) are wave-uniform (scalar loads); each half stores
        // through its row's range-checked buffer, so the other half's
        // lanes, flagged windows and outputs past the row's end (offset
        // out of range) are dropped by the hardware.
        // (the row loads stay inside the tile loop: hoisted out of it,
        // 32 rows' descriptors would not fit in scalar registers)
        int rowT = rowBase, nT = nMine;
        asm volatile("" : "+s"(rowT), "+s"(nT));
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int jA = 8 * (r >> 2) + (r & 3), jB = jA + 4;
            const TdRow rA = rows[rowT + min(jA, nT - 1)];
            const TdRow rB = rows[rowT + min(jB, nT - 1)];
            const int nvA = jA < nT ? static_cast<int>(max(0ll, min(static_cast<long long>(IS), rA.nOut - base))) : 0;
            const int nvB = jB < nT ? static_cast<int>(max(0ll, min(static_cast<long long>(IS), rB.nOut - base))) : 0;
            const int Lt = half ? rB.Lt : rA.Lt;
            const float kc = half ? rB.kc : rA.kc;
            const float kn = half ? rB.kn : rA.kn;
            const float sl = half ? rB.sqrtLt : rA.sqrtLt;
            const int nv = half ? nvB : nvA;
            // window sums in FP64 (differences of the block's prefix sums),
            // the rest in FP32
            const double S1 = P1[i + Lt] - p1i;
            const double S2 = P2[i + Lt] - p2i;
            const float d2 = static_cast<float>(fma(-S1, S1, Lt * S2));
            const float s2 = static_cast<float>(S2);
            const bool canc = !(d2 > fmaf(kc, s2, flK * static_cast<float>(Lt)));
            const bool flag = canc || kn * s2 > d2;
            float y = __builtin_amdgcn_rsqf(d2);
            y = y * (1.5f - 0.5f * d2 * y * y);
            const unsigned rho = __float_as_uint((acc0[r] + acc1[r]) * (scF * sl * y));
            constexpr int kDrop = 0x7ffffff0;  // beyond every row's range
            const int off = flag ? kDrop : 4 * i;
            __builtin_amdgcn_raw_buffer_store_b32(rho, buffer_rsrc(static_cast<float *>(rA.out) + base, 4 * nvA),
                                                  half ? kDrop : off, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(rho, buffer_rsrc(static_cast<float *>(rB.out) + base, 4 * nvB),
                                                  half ? off : kDrop, 0, 0);
            const bool valid = i < nv;
            flg |= (valid && flag ? 1u : 0u) << r;
            cnc |= (valid && canc ? 1u : 0u) << r;
            // two steps' loads in flight at a time: the scheduler may not
            // hoist more (their registers would not fit beside the taps)
            if ((r & 1) == 1) { __builtin_amdgcn_sched_barrier(0); }
        }
        const unsigned long long any = __ballot(flg != 0u);
        if (any != 0ull) {
            fq[wave][tile][lane] = flg | (cnc << 16);
            tiles |= 1u << tile;
        }
    }
    // 6. flagged windows, after the tile loop (the taps are dead here).
    //    Exactly constant windows (a cancelled variance over equal
    //    samples) are 0 at once; the others go to the device queue, whose
    //    windows fix_kernel recomputes after the launch.  When the queue
    //    is full they go to the wave's list, which is recomputed here one
    //    window per lane in passes of 64 (a few left at the end: the whole
    //    wave on one window at a time).  List entry: tile
    //    (4 bits), step r (4 bits), lane (6 bits), cancelled (1 bit); in
    //    tile, step, lane order.
    unsigned short *list = fixList[wave];
    int nList = 0;  // wave-uniform
    auto entryWindow = [&](unsigned e, int &wi, TdRow &row, bool &cr) {
        const int t = e >> 10, r = (e >> 6) & 15, l = e & 63;
        wi = 32 * t + (l & 31);
        row = rows[rowBase + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3)];
        cr = (e >> 14) & 1u;
    };
    auto runList = [&](bool all) {
        int done = 0;
        while (nList - done >= 64 || (all && nList - done > kTdCoopFlags)) {
            const int n = min(64, nList - done);
            if (lane < n) {
                int wi;
                TdRow row;
                bool cr;
                entryWindow(list[done + lane], wi, row, cr);
                const double rho = exact_window_td(raw, P1, P2, wi, to_global(row.tmpl), row.Lt, a, cr, 0, 1);
                to_global(static_cast<float *>(row.out))[base + wi] = static_cast<float>(rho);
            }
            done += n;
        }
        if (all) {
            for (; done < nList; ++done) {
                int wi;
                TdRow row;
                bool cr;
                entryWindow(list[done], wi, row, cr);  // uniform entry: the wave shares it
                const double rho = exact_window_td(raw, P1, P2, wi, to_global(row.tmpl), row.Lt, a, cr, lane, 64);
                if (lane == 0) { to_global(static_cast<float *>(row.out))[base + wi] = static_cast<float>(rho); }
            }
        } else if (done > 0) {
            // the remainder (under 64 entries) moves to the front: every
            // lane reads before any lane writes (one instruction each)
            const int rest = nList - done;
            const unsigned short v = lane < rest ? list[done + lane] : 0;
            if (lane < rest) { list[lane] = v; }
        }
        nList -= done;
    };
    while (tiles != 0u) {
        const int tile = __ffs(tiles) - 1;
        tiles &= tiles - 1u;
        const int i = tile * 32 + col;
        const unsigned word = fq[wave][tile][lane];
        const unsigned flg = word & 0xffffu;
        const unsigned cnc = word >> 16;
        int total = __popc(flg), full = __popc(flg & cnc);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            total += __shfl_xor(total, o, 64);
            full += __shfl_xor(full, o, 64);
        }
        nFlag += static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(total));
        nCanc += static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(full));
        // constant windows now; the others counted for the list
        unsigned keep = flg;
        for (unsigned m = flg & cnc; m != 0u; m &= m - 1u) {
            const int r = __ffs(m) - 1;
            const TdRow row = rows[rowBase + 8 * (r >> 2) + 4 * half + (r & 3)];
            if (constant_window(eqw, i, row.Lt)) {
                to_global(static_cast<float *>(row.out))[base + i] = 0.0f;
                keep &= ~(1u << r);
            }
        }
        // to the device queue when it has room for the tile's windows
        // (step by step, lanes in order)
        int n = 0;
        for (int r = 0; r < 16; ++r) { n += __popcll(__ballot((keep >> r) & 1u)); }
        if (n == 0) { continue; }
        unsigned long long q0 = 0;
        if (lane == 0) { q0 = atomicAdd(queue.count, static_cast<unsigned long long>(n)); }
        q0 = __shfl(q0, 0, 64);
        if (q0 <= queue.cap && static_cast<unsigned long long>(n) <= queue.cap - q0) {
            for (int r = 0; r < 16; ++r) {
                const bool mine = (keep >> r) & 1u;
                const unsigned long long bal = __ballot(mine);
                if (mine) {
                    const unsigned long long at = q0 + __popcll(bal & ((1ull << lane) - 1ull));
                    const int j = 8 * (r >> 2) + 4 * half + (r & 3);
                    const unsigned rowCode = static_cast<unsigned>(rowBase + j) | (((cnc >> r) & 1u) << 31);
                    gptr<unsigned long long> e = to_global(reinterpret_cast<unsigned long long *>(queue.entries));
                    e[at] = (static_cast<unsigned long long>(base + i) << 32) | rowCode;
                }
                q0 += __popcll(bal);
            }
            continue;
        }
        // queue full: the wave's own list, step by step (lanes in order):
        // the windows of a pass then share one or two template rows, whose
        // loads are broadcasts
        for (int r = 0; r < 16; ++r) {
            const bool mine = (keep >> r) & 1u;
            const unsigned long long bal = __ballot(mine);
            if (bal == 0ull) { continue; }
            if (mine) {
                const int at = nList + __popcll(bal & ((1ull << lane) - 1ull));
                list[at] = static_cast<unsigned short>((tile << 10) | (r << 6) | lane | (((cnc >> r) & 1u) << 14));
            }
            nList += __popcll(bal);
        }
        runList(false);
    }
    runList(true);
    if (sg.flagged != nullptr && nFlag != 0ull && lane == 0) {
        atomicAdd(sg.flagged, nFlag);
        if (nCanc != 0ull) { atomicAdd(sg.flagged + 1, nCanc); }
    }
}

// Split templates: frag[row] = the KT taps of each part in turn (hi, then
// lo for FP16; hi, mid, lo for BF16), zero past Lt.  Each part rounds what
// the parts before it left of the FP64 tap (the remainder is exact in FP64).
template <typename MT>
__global__ __launch_bounds__(NT) void td_prep_kernel(const TdPrepJob *__restrict__ jobs, int KT)
{
    constexpr int NP = TdTraits<MT>::parts;
    const TdPrepJob job = jobs[blockIdx.x];
    MT *f = static_cast<MT *>(job.frag);
    for (int k = threadIdx.x; k < KT; k += NT) {
        double rest = k < job.Lt ? job.tmpl[k] : 0.0;
#pragma unroll
        for (int pt = 0; pt < NP; ++pt) {
            const MT part = static_cast<MT>(static_cast<float>(rest));
            f[pt * KT + k] = part;
            rest -= static_cast<double>(static_cast<float>(part));
        }
    }
}

}  // namespace

bool td_supported(int nb) { return nb >= 1 && nb <= kTdMaxTaps; }

int td_taps(int nb) { return nb <= 64 ? 64 : (nb <= 128 ? 128 : 256); }

hipError_t launch_td_prep(int kind, const TdPrepJob *dJobs, int nJobs, int KT, hipStream_t stream)
{
    if (nJobs <= 0) { return hipSuccess; }
    if (kind == kTdFp16) {
        hipLaunchKernelGGL(td_prep_kernel<_Float16>, dim3(nJobs), dim3(NT), 0, stream, dJobs, KT);
    } else {
        hipLaunchKernelGGL(td_prep_kernel<__bf16>, dim3(nJobs), dim3(NT), 0, stream, dJobs, KT);
    }
    return hipGetLastError();
}

hipError_t launch_td(int kind, int KT, const TdItem *dItems, int nItems, const Seg *dSegs, const TdRow *dRows,
                     FixQueue queue, hipStream_t stream)
{
    if (nItems <= 0) { return hipSuccess; }
#define MF_TD_LAUNCH(MT, K) \
    hipLaunchKernelGGL((tdcorr_kernel<MT, K>), dim3(nItems), dim3(NT), 0, stream, dItems, dSegs, dRows, queue)
    if (kind == kTdFp16) {
        switch (KT) {
        case 64: MF_TD_LAUNCH(_Float16, 64); break;
        case 128: MF_TD_LAUNCH(_Float16, 128); break;
        case 256: MF_TD_LAUNCH(_Float16, 256); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (KT) {
        case 64: MF_TD_LAUNCH(__bf16, 64); break;
        case 128: MF_TD_LAUNCH(__bf16, 128); break;
        case 256: MF_TD_LAUNCH(__bf16, 256); break;
        default: return hipErrorInvalidValue;
        }
    }
#undef MF_TD_LAUNCH
    return hipGetLastError();
}

}  // namespace mfk
