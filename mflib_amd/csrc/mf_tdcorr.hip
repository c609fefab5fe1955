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
// v_mfma_f32_16x16x32 accumulates them in FP32: hi*hi in its own
// accumulator, the smaller products in a second one.
// rho = s * D * scale with the window normalisation from FP64 prefix sums
// of the block.
//
// Work item: IS = 512 consecutive outputs of the lane and up to 128 rows.
// Each of the 4 waves takes 16 rows at a time: their split taps stay in
// registers (weight-stationary, parts * KT / 8 VGPRs) while the wave walks
// the 32 output tiles of 16; the Hankel operand is read from LDS, where
// entry p holds the 8 samples x~[p .. p+7] (one 16-byte read per part and
// k-step of 32).
//
// Accuracy: per window |d rho| <= eps_m * sqrt(Lt S2) / den with
// eps_m = eps_split + (4 sqrt(KT/16) + 4) 2^-24 (FP32 accumulation over at
// most KT / 16 accumulator updates -- the kernel makes KT / 32 -- and the
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

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// The 16-bit operand type's parts and its v_mfma_f32_16x16x32 form.
template <typename MT>
struct TdTraits;
template <>
struct TdTraits<_Float16> {
    using vec8 = f16x8;
    static constexpr int parts = 2;
    __device__ static f32x4 mfma(vec8 a, vec8 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct TdTraits<__bf16> {
    using vec8 = bf16x8;
    static constexpr int parts = 3;
    __device__ static f32x4 mfma(vec8 a, vec8 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

constexpr int kTdCoopFlags = 4;  // at most this many left at the end: the wave shares each window

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

// One work item's staged samples in LDS, shared by both tile shapes.
template <typename MT, int KT>
struct TdStage {
    static constexpr int NP = TdTraits<MT>::parts;
    static constexpr int SPAN = IS + KT;  // samples one item reads
    MT x8[NP][SPAN][8];                   // entry p of part q: x~[p .. p+7]
    float raw[SPAN];                      // samples (FP32), for exact windows
    double P1[SPAN + 1];                  // prefix sums of d = x - a
    double P2[SPAN + 1];                  //   and of d^2
    double red[NT / 64];
    double wsum1[NT / 64], wsum2[NT / 64];
    unsigned eqw[(SPAN + 31) / 32];       // bit p: sample p equals sample p + 1 (constant windows)
};

// What the epilogue needs of the staged block: the anchor a, the scale s
// of x~ (FP32) and the cancellation floor 1e-12 E (per tap of Lt).
struct TdBlock {
    double a;
    float scF;
    float flK;
};

// Stages outputs [base, base + IS) of lane sg: loads the samples, takes the
// block anchor and RMS, splits x~ into the 8-wide Hankel entries of every
// part, builds the FP64 prefix sums and the equal-neighbour bits.  The
// caller synchronises the block before reading any of it.
template <typename MT, int KT>
__device__ __forceinline__ TdBlock td_stage(TdStage<MT, KT> &sm, const Seg &sg, long long base)
{
    constexpr int NP = TdTraits<MT>::parts;
    constexpr int SPAN = TdStage<MT, KT>::SPAN;
    constexpr int SPT = (SPAN + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
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
        if (p < SPAN) { sm.raw[p] = val; }
        s += static_cast<double>(val);
    }
    const double a = block_sum<NT>(s, sm.red) / nSamp;
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
    const double energy = block_sum<NT>(e, sm.red);
    const double sc = energy > 0 ? sqrt(energy / nSamp) : 1.0;
    const float inv = static_cast<float>(1.0 / sc);
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
                    if (p - k >= 0) { sm.x8[pt][p - k][k] = part; }
                }
            }
        }
    }
    // entries past the span read zeros in their upper slots
    for (int p = SPAN - 7 + tid; p < SPAN; p += NT) {
        for (int k = SPAN - p; k < 8; ++k) {
#pragma unroll
            for (int pt = 0; pt < NP; ++pt) { sm.x8[pt][p][k] = static_cast<MT>(0.0f); }
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
        sm.wsum1[wave] = x1;
        sm.wsum2[wave] = x2;
    }
    __syncthreads();
    if (tid < (SPAN + 31) / 32) {
        unsigned bits = 0;
        for (int b = 0; b < 32; ++b) {
            const int p = 32 * tid + b;
            if (p + 1 < SPAN && sm.raw[p] == sm.raw[p + 1]) { bits |= 1u << b; }
        }
        sm.eqw[tid] = bits;
    }
    double w1 = 0, w2 = 0;
    for (int w = 0; w < wave; ++w) {
        w1 += sm.wsum1[w];
        w2 += sm.wsum2[w];
    }
    const double e1 = w1 + x1 - l1, e2 = w2 + x2 - l2;  // exclusive prefix of this thread's run
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        if (p < SPAN) {
            sm.P1[p] = e1 + d1[q];
            sm.P2[p] = e2 + d2[q];
        }
    }
    if (tid == NT - 1) {
        sm.P1[SPAN] = w1 + x1;
        sm.P2[SPAN] = w2 + x2;
    }
    return TdBlock{a, static_cast<float>(sc), static_cast<float>(1.0e-12 * energy)};
}

// A dead lane's item: zeros in every row.
__device__ __forceinline__ void td_dead(const TdItem &item, const TdRow *__restrict__ rows)
{
    const long long base = item.seg0;
    for (int r = 0; r < item.nRows; ++r) {
        const TdRow row = rows[item.rowBegin + r];
        const int nv = static_cast<int>(min(static_cast<long long>(IS), row.nOut - base));
        gptr<float> out = to_global(static_cast<float *>(row.out)) + base;
        for (int i = threadIdx.x; i < nv; i += NT) { out[i] = 0.0f; }
    }
}

// Tiles of 16 rows x 16 outputs on v_mfma_f32_16x16x32: each wave holds the
// split taps of 16 rows (parts * KT / 8 VGPRs: 96 for the BF16 triple at
// KT = 256; tiles of 32 rows would need 192 and leave one wave per SIMD)
// and walks the item's 32 output tiles; the 4 waves take 64 rows at a
// time, in two row groups when the item has more.  A lane holds column
// i = lane % 16 and rows 4 (lane / 16) + r, r = 0..3: their epilogue
// constants stay in registers for the row group, and each quarter wave
// stores 16 consecutive outputs of its row.
constexpr int kTdTiles = IS / 16;        // output tiles of 16 per item
constexpr int kTdFixList = 64 + 4 * 64;  // carried over plus one tile's windows

// Blocks per CU the register budget is capped for: three for the FP16 pair
// (64 VGPRs of taps at KT = 256, 51 KB of LDS per block), two for the BF16
// triple (96 VGPRs of taps, 63 KB of LDS).
template <typename MT>
constexpr int td_blocks_per_cu() { return TdTraits<MT>::parts == 2 ? 3 : 2; }

template <typename MT, int KT>
__global__ __launch_bounds__(NT, (td_blocks_per_cu<MT>())) void tdcorr_kernel(const TdItem *__restrict__ items,
                                                       const Seg *__restrict__ segs,
                                                       const TdRow *__restrict__ rows, FixQueue queue)
{
    using Tr = TdTraits<MT>;
    using V8 = typename Tr::vec8;
    constexpr int NP = Tr::parts;
    constexpr int NS = KT / 32;  // k-steps of 32
    __shared__ __align__(16) TdStage<MT, KT> sm;
    // flagged windows of each wave's tiles: bit r flags the lane's window
    // r, bit 4 + r marks a cancelled variance
    __shared__ unsigned char fq[NT / 64][kTdTiles][64];
    // each wave's list of flagged windows to recompute when the device
    // queue is full
    __shared__ unsigned short fixList[NT / 64][kTdFixList];

    const TdItem item = items[blockIdx.x];
    const Seg sg = segs[item.seg];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int quad = lane >> 4;  // k-group of the operands, row group of the result
    // The taps of k-group quad: 8 kq + 0..7 of each k-step of 32, with kq =
    // 0, 2, 1, 3 for quad = 0..3 (A and B alike, so the products summed are
    // the same): each half wave then reads 32 consecutive Hankel entries
    // (quads 0 and 1: i0 + 0..31; 2 and 3: i0 + 8..39) instead of two
    // overlapping runs of 16 in the same LDS banks.
    const int kq = ((quad & 1) << 1) | (quad >> 1);
    const int c16 = lane & 15;   // output column
    const long long base = item.seg0;

    if (sg.dead) {
        td_dead(item, rows);
        return;
    }
    const TdBlock blk = td_stage(sm, sg, base);
    const double a = blk.a;
    const float scF = blk.scF;
    const float flK = blk.flK;
    __syncthreads();  // Hankel entries, prefix sums and raw samples ready

    unsigned long long nFlag = 0, nCanc = 0;  // windows recomputed by this wave, in full
    for (int g = 0; 64 * g < item.nRows; ++g) {
        const int rowBase = item.rowBegin + 64 * g + 16 * wave;
        const int nMine = min(16, item.nRows - 64 * g - 16 * wave);
        if (nMine <= 0) { break; }
        // split taps (A operand: row = lane % 16, taps 32 s + 8 kq + 0..7
        // of each part)
        V8 ap[NP][NS];
        {
            const bool has = c16 < nMine;
            const MT *f = has ? static_cast<const MT *>(rows[rowBase + c16].frag) : nullptr;
#pragma unroll
            for (int pt = 0; pt < NP; ++pt) {
#pragma unroll
                for (int st = 0; st < NS; ++st) {
                    if (has) {
                        const gptr<const i32x4> fh = to_global(reinterpret_cast<const i32x4 *>(f));
                        ap[pt][st] = __builtin_bit_cast(V8, fh[pt * (KT / 8) + 4 * st + kq]);
                    } else {
                        ap[pt][st] = __builtin_bit_cast(V8, i32x4{0, 0, 0, 0});
                    }
                }
            }
        }
        // the epilogue constants of the lane's 4 rows (absent rows: no
        // outputs, and a length that keeps the prefix reads in the span)
        int LtR[4], nvR[4];
        float kcR[4], knR[4], slR[4];
        gptr<float> outR[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 4 * quad + r;
            if (j < nMine) {
                const TdRow row = rows[rowBase + j];
                LtR[r] = row.Lt;
                nvR[r] = static_cast<int>(max(0ll, min(static_cast<long long>(IS), row.nOut - base)));
                kcR[r] = row.kc;
                knR[r] = row.kn;
                slR[r] = row.sqrtLt;
                outR[r] = to_global(static_cast<float *>(row.out)) + base;
            } else {
                LtR[r] = 1;
                nvR[r] = 0;
                kcR[r] = 0.0f;
                knR[r] = 0.0f;
                slR[r] = 0.0f;
                outR[r] = nullptr;
            }
        }
        unsigned tiles = 0;  // tiles with flagged windows (wave-uniform)
        for (int tile = 0; tile < kTdTiles; ++tile) {
            const int i0 = 16 * tile;
            f32x4 acc0 = {}, acc1 = {};
            // B operand: column i0 + c16, taps 32 s + 8 kq + 0..7 = entry
            // i0 + c16 + 8 kq + 32 s
            const i32x4 *b0p = reinterpret_cast<const i32x4 *>(&sm.x8[0][i0 + c16 + 8 * kq][0]);
            const i32x4 *b1p = reinterpret_cast<const i32x4 *>(&sm.x8[1][i0 + c16 + 8 * kq][0]);
            const i32x4 *b2p = reinterpret_cast<const i32x4 *>(&sm.x8[NP - 1][i0 + c16 + 8 * kq][0]);
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                const V8 b0 = __builtin_bit_cast(V8, b0p[32 * st]);
                const V8 b1 = __builtin_bit_cast(V8, b1p[32 * st]);
                acc0 = Tr::mfma(ap[0][st], b0, acc0);
                acc1 = Tr::mfma(ap[0][st], b1, acc1);
                acc1 = Tr::mfma(ap[1][st], b0, acc1);
                if constexpr (NP == 3) {
                    const V8 b2 = __builtin_bit_cast(V8, b2p[32 * st]);
                    acc1 = Tr::mfma(ap[0][st], b2, acc1);
                    acc1 = Tr::mfma(ap[NP - 1][st], b0, acc1);
                    acc1 = Tr::mfma(ap[1][st], b1, acc1);
                }
            }
            // epilogue: normalise, flag, store
            const int i = i0 + c16;
            const double p1i = sm.P1[i], p2i = sm.P2[i];
            unsigned bits = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int Lt = LtR[r];
                // window sums in FP64 (differences of the block's prefix
                // sums), the rest in FP32
                const double S1 = sm.P1[i + Lt] - p1i;
                const double S2 = sm.P2[i + Lt] - p2i;
                const float d2 = static_cast<float>(fma(-S1, S1, Lt * S2));
                const float s2 = static_cast<float>(S2);
                const bool canc = !(d2 > fmaf(kcR[r], s2, flK * static_cast<float>(Lt)));
                const bool flag = canc || knR[r] * s2 > d2;
                float y = __builtin_amdgcn_rsqf(d2);
                y = y * (1.5f - 0.5f * d2 * y * y);
                const float rho = (acc0[r] + acc1[r]) * (scF * slR[r] * y);
                const bool valid = i < nvR[r];
                if (valid && !flag) { outR[r][i] = rho; }
                bits |= (valid && flag ? 1u : 0u) << r;
                bits |= (valid && canc ? 1u : 0u) << (4 + r);
            }
            if (__ballot(bits != 0u) != 0ull) {
                fq[wave][tile][lane] = static_cast<unsigned char>(bits);
                tiles |= 1u << tile;
            }
        }
        // flagged windows of the row group, after its tile loop: exactly
        // constant windows are 0 at once, the others go to the device
        // queue (fix_kernel recomputes them after the launch), or, when it
        // is full, to the wave's list, recomputed here one window per lane
        // in passes of 64 (a few left at the end: the whole wave on one
        // window at a time).  List entry: tile (5 bits), step r (2 bits),
        // lane (6 bits), cancelled (1 bit).
        unsigned short *list = fixList[wave];
        int nList = 0;  // wave-uniform
        auto entryWindow = [&](unsigned e, int &wi, TdRow &row, bool &cr) {
            const int t = (e >> 8) & 31, r = (e >> 6) & 3, l = e & 63;
            wi = 16 * t + (l & 15);
            row = rows[rowBase + 4 * (l >> 4) + r];
            cr = (e >> 13) & 1u;
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
                    const double rho =
                        exact_window_td(sm.raw, sm.P1, sm.P2, wi, to_global(row.tmpl), row.Lt, a, cr, 0, 1);
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
                    const double rho =
                        exact_window_td(sm.raw, sm.P1, sm.P2, wi, to_global(row.tmpl), row.Lt, a, cr, lane, 64);
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
            const int i = 16 * tile + c16;
            const unsigned word = fq[wave][tile][lane];
            const unsigned flg = word & 15u;
            const unsigned cnc = word >> 4;
            int total = __popc(flg), full = __popc(flg & cnc);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                total += __shfl_xor(total, o, 64);
                full += __shfl_xor(full, o, 64);
            }
            nFlag += static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(total));
            nCanc += static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(full));
            unsigned keep = flg;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (((flg & cnc) >> r) & 1u) {
                    if (constant_window(sm.eqw, i, LtR[r])) {
                        outR[r][i] = 0.0f;
                        keep &= ~(1u << r);
                    }
                }
            }
            int n = 0;
#pragma unroll
            for (int r = 0; r < 4; ++r) { n += __popcll(__ballot((keep >> r) & 1u)); }
            if (n == 0) { continue; }
            unsigned long long q0 = 0;
            if (lane == 0) { q0 = atomicAdd(queue.count, static_cast<unsigned long long>(n)); }
            q0 = __shfl(q0, 0, 64);
            if (q0 <= queue.cap && static_cast<unsigned long long>(n) <= queue.cap - q0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const bool mine = (keep >> r) & 1u;
                    const unsigned long long bal = __ballot(mine);
                    if (mine) {
                        const unsigned long long at = q0 + __popcll(bal & ((1ull << lane) - 1ull));
                        const unsigned rowCode =
                            static_cast<unsigned>(rowBase + 4 * quad + r) | (((cnc >> r) & 1u) << 31);
                        gptr<unsigned long long> e = to_global(reinterpret_cast<unsigned long long *>(queue.entries));
                        e[at] = (static_cast<unsigned long long>(base + i) << 32) | rowCode;
                    }
                    q0 += __popcll(bal);
                }
                continue;
            }
            // The queue is full.  A reservation that straddles its end
            // leaves the slots [q0, cap) unused: they get the empty entry,
            // which fix_kernel skips (it reads min(count, cap) entries).
            if (q0 < queue.cap) {
                gptr<unsigned long long> e = to_global(reinterpret_cast<unsigned long long *>(queue.entries));
                for (unsigned long long k = q0 + lane; k < queue.cap; k += 64) { e[k] = kFixEmpty; }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bool mine = (keep >> r) & 1u;
                const unsigned long long bal = __ballot(mine);
                if (mine) {
                    const int at = nList + __popcll(bal & ((1ull << lane) - 1ull));
                    list[at] = static_cast<unsigned short>((tile << 8) | (r << 6) | lane | (((cnc >> r) & 1u) << 13));
                }
                nList += __popcll(bal);
            }
            runList(false);
        }
        runList(true);
    }
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
