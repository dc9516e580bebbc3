// engine.hip -- CDNA4 (gfx950) kernels of the spectrometer core.
//
// Hot path of TaskCoreIPP::doMaths (src/IA-32/TaskCoreIPP.cpp:394-511) as
// three fused kernels per chunk of G consecutive FFTs:
//   pass 1  : 2-bit unpack + window + N1-point column FFTs (LDS Stockham) +
//             four-step twiddle; writes Y[k1][n2] (row-contiguous).
//   pcal    : unwindowed unpack x rotator, folded into tone bins in FP64
//             (extract_PCal, :633-693); per-block partials, fixed-order reduce.
//   pass 2  : N2-point row FFTs for the row pair (k1, N1-k1), real-to-complex
//             split, |X|^2 (and X*conj(Y)) summed over the chunk in registers,
//             then one Kahan-compensated update of the resident accumulators.
// The real N-point transform is computed as an M = N/2 complex FFT of
// z[m] = x[2m] + i x[2m+1] followed by the split
//   X[k] = E[k] + W_N^k O[k],  E = (Z[k] + conj Z[M-k])/2,  O = (Z[k] - conj Z[M-k])/(2i).
// Accumulators are kept in the permuted order of pass 2 (bin k = k1 + N1 k2
// stored at k1*N2 + k2, Nyquist at M); finalize writes natural order.
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <type_traits>

namespace swspec {

// ---------------------------------------------------------------------------
// complex helpers and radix-R forward DFTs (W_R = e^{-2 pi i / R})
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R> struct Dft;

template <> struct Dft<2> {
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    }
};

template <> struct Dft<3> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float s3 = 0.866025403784438647f;
        float2 t = make_float2(v[1].x + v[2].x, v[1].y + v[2].y);
        float2 d = make_float2((v[1].x - v[2].x) * s3, (v[1].y - v[2].y) * s3);
        float2 a = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        v[0] = make_float2(v[0].x + t.x, v[0].y + t.y);
        v[1] = make_float2(a.x + d.y, a.y - d.x);
        v[2] = make_float2(a.x - d.y, a.y + d.x);
    }
};

template <> struct Dft<4> {
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 s0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        float2 d0 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        float2 s1 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        float2 d1 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
        v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
        v[1] = make_float2(d0.x + d1.y, d0.y - d1.x);   // d0 - i d1
        v[3] = make_float2(d0.x - d1.y, d0.y + d1.x);   // d0 + i d1
    }
};

template <> struct Dft<5> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = 0.309016994374947424f, c2 = -0.809016994374947424f;
        const float s1 = 0.951056516295153572f, s2 = 0.587785252292473129f;
        float2 t1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y);
        float2 t2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y);
        float2 t3 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
        float2 t4 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
        float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
        float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
        v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
        v[1] = make_float2(a1.x + b1.y, a1.y - b1.x);
        v[4] = make_float2(a1.x - b1.y, a1.y + b1.x);
        v[2] = make_float2(a2.x + b2.y, a2.y - b2.x);
        v[3] = make_float2(a2.x - b2.y, a2.y + b2.x);
    }
};

template <> struct Dft<8> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float h = 0.707106781186547524f;
        float2 a[4], b[4];
#pragma unroll
        for (int n = 0; n < 4; n++) {
            a[n] = make_float2(v[n].x + v[n + 4].x, v[n].y + v[n + 4].y);
            b[n] = make_float2(v[n].x - v[n + 4].x, v[n].y - v[n + 4].y);
        }
        // b[n] *= W_8^n
        b[1] = make_float2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
        b[2] = make_float2(b[2].y, -b[2].x);
        b[3] = make_float2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
        Dft<4>::run(a);
        Dft<4>::run(b);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[2 * k] = a[k];
            v[2 * k + 1] = b[k];
        }
    }
};

// Register radices.  R is a plain radix (2, 3, 4, 5, 8) or a composite
// R = A*B (10 = 2*5, 16 = 4*4, 20 = 4*5, 25 = 5*5, 32 = 4*8).  A stage with
// ns > 1 multiplies input r by W_D^{k r} (D = ns*R, k = q % ns).  A plain
// radix reads those R-1 twiddles from the stage table.  A composite reads A-1
// values W_D^{k B n1} and B-1 values W_D^{k n2} and applies them around its
// A-point DFTs: A+B-2 loads instead of R-1, for more multiplies (radix 16:
// 6 loads and 24 multiplies against 15 and 15).  The row pass waits on
// memory more than it computes, and R-1 loads made it 14% slower at 3.2M;
// radix_factored_tw() selects the form per radix.
__host__ __device__ constexpr int radix_a(int R)
{
    return R == 10 ? 2 : R == 16 ? 4 : R == 20 ? 4 : R == 25 ? 5 : R == 32 ? 4 : R;
}
__host__ __device__ constexpr bool radix_factored_tw(int R) { return radix_a(R) != R; }
__host__ __device__ constexpr int radix_ntw(int R)
{
    return radix_factored_tw(R) ? radix_a(R) + R / radix_a(R) - 2 : R - 1;
}

// A sub-plan of up to three stages as subplan_from builds it (the stage
// twiddle table layout included), at compile time: the plans of the common
// lengths have their own kernel instantiations in which every stride and
// bound is a constant.
__host__ __device__ constexpr SubPlan fixed_subplan(int L, int T, int padst, int pad, int r0, int r1, int r2 = 0)
{
    SubPlan sp{};
    const int rr[3] = {r0, r1, r2};
    const int n = r2 ? 3 : 2;
    int ns = 1, twl = 0;
    sp.L = L;
    sp.nst = n;
    for (int st = 0; st < n; st++) {
        sp.radix[st] = rr[st];
        sp.ns[st] = ns;
        sp.mag[st] = ns > 1 ? (unsigned)(((1ULL << 32) + (unsigned long long)ns - 1) / (unsigned long long)ns) : 0u;
        sp.twoff[st] = twl;
        twl += ns * radix_ntw(rr[st]);
        ns *= rr[st];
    }
    sp.twlen = twl;
    sp.T = T;
    sp.padst = padst;
    sp.pad = pad;
    return sp;
}
// Fixed plans.  Id 1, 3.2M points: columns 1250 = 10*25*5, 4 columns of 64
// threads; rows 1280 = 5*16*16, 96 threads per row, penultimate stage padded
// by 5.  (A 2^20 entry would not be taken: both of its passes run two
// pipelines per workgroup.)
constexpr int kFixedPlans = 1;
struct FixedPlan {
    int N1, N2, B, p1_threads;
    SubPlan cols, rows;
    int rowst;   // LDS row stride of the row pass
};
__host__ __device__ constexpr FixedPlan fixed_plan(int id)
{
    FixedPlan f{};
    if (id == 1) {
        f.N1 = 1250; f.N2 = 1280; f.B = 4; f.p1_threads = 256;
        f.cols = fixed_subplan(1250, 64, -1, 0, 10, 25, 5);
        f.rows = fixed_subplan(1280, 96, 1, 5, 5, 16, 16);
    }
    f.rowst = f.N2 + (f.rows.padst >= 0 ? f.rows.pad * f.rows.radix[f.rows.nst - 1] : 0);
    return f;
}
__host__ __device__ constexpr bool same_subplan(const SubPlan& a, const SubPlan& b)
{
    if (a.L != b.L || a.nst != b.nst || a.T != b.T || a.padst != b.padst || a.pad != b.pad) return false;
    for (int st = 0; st < a.nst; st++)
        if (a.radix[st] != b.radix[st] || a.ns[st] != b.ns[st] || a.mag[st] != b.mag[st] || a.twoff[st] != b.twoff[st])
            return false;
    return true;
}

template <bool A, class T>
__device__ __forceinline__ const T& pick_ref(const T& a, const T& b)
{
    if constexpr (A) return a;
    else return b;
}

// Roots of unity at compile time: Taylor series on the angle reduced to (-pi, pi].
constexpr double kPiD = 3.14159265358979323846264338327950288;
constexpr double cx_sin(double x)
{
    double t = x, s = x;
    for (int i = 1; i < 18; i++) { t *= -x * x / ((2.0 * i) * (2.0 * i + 1)); s += t; }
    return s;
}
constexpr double cx_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int i = 1; i < 18; i++) { t *= -x * x / ((2.0 * i - 1) * (2.0 * i)); s += t; }
    return s;
}
template <int R> struct RootTable {
    float re[R], im[R];
    constexpr RootTable() : re(), im()
    {
        for (int e = 0; e < R; e++) {
            double x = -2.0 * kPiD * e / R;       // W_R^e
            if (x <= -kPiD) x += 2.0 * kPiD;
            re[e] = (float)cx_cos(x);
            im[e] = (float)cx_sin(x);
        }
    }
};

// Twiddle (TW: a stage after the first) + butterfly; output r goes to st(r, value).
template <int R, bool COMP = (radix_a(R) != R)> struct DftStore {
    template <bool TW, class St>
    __device__ __forceinline__ static void run(float2* v, const float2* __restrict__ w, St st)
    {
        if constexpr (TW) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], w[r - 1]);
        }
        Dft<R>::run(v);
#pragma unroll
        for (int r = 0; r < R; r++) st(r, v[r]);
    }
};

// Composite R = A*B: n = B n1 + n2, k = k1 + A k2,
//   X[k1 + A k2] = sum_n2 W_B^{n2 k2} [W_D^{k n2} W_R^{n2 k1}]
//                         sum_n1 (v[B n1 + n2] W_D^{k B n1}) W_A^{n1 k1}.
// Step 1 (A-point DFTs + twiddles) runs in place; step 2 (B-point DFTs) hands
// each output straight to the store, so no second register array is needed.
// A first stage (ns = 1, k = 0) has only the internal W_R factors.  With
// R-1 table twiddles (radix_factored_tw false) the stage factors multiply the
// inputs up front, as for a plain radix.
template <int R> struct DftStore<R, true> {
    static constexpr int A = radix_a(R), B = R / A;
    static constexpr bool FT = radix_factored_tw(R);
    static constexpr RootTable<R> kRoots{};
    template <bool TW, class St>
    __device__ __forceinline__ static void run(float2* v, const float2* __restrict__ w, St st)
    {
        if constexpr (TW && !FT) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], w[r - 1]);
        }
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) {
            float2 u[A];
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) {
                u[n1] = v[B * n1 + n2];
                if (TW && FT && n1 > 0) u[n1] = cmul(u[n1], w[n1 - 1]);
            }
            Dft<A>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < A; k1++) {
                float2 x = u[k1];
                if (TW && FT && n2 > 0) x = cmul(x, w[A - 1 + n2 - 1]);
                const int e = (n2 * k1) % R;
                v[B * k1 + n2] = (e == 0) ? x : cmul(x, make_float2(kRoots.re[e], kRoots.im[e]));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) {
            float2 u[B];
#pragma unroll
            for (int n2 = 0; n2 < B; n2++) u[n2] = v[B * k1 + n2];
            Dft<B>::run(u);
#pragma unroll
            for (int k2 = 0; k2 < B; k2++) st(k1 + A * k2, u[k2]);
        }
    }
};

// ---------------------------------------------------------------------------
// Stockham stage in LDS.  Element e of the FFT lives at lds[ebase + e*es]
// (es = element stride: the B columns of a pass-1 workgroup interleaved).
// Butterfly q (q < L/R) reads q + r*L/R, writes (q - q%ns)*R + q%ns + r*ns.
// Every thread of the workgroup must call it (it contains barriers).
// Index products use 24-bit multiplies (full rate; every factor < 2^24).
// ---------------------------------------------------------------------------

// q % ns via a multiply-high with mag = ceil(2^32/ns): exact while q*ns < 2^32
// (q, ns < 16384 here); ns == 1 is handled by the caller's k = 0.
__device__ __forceinline__ int mod_magic(int q, int ns, unsigned mag)
{
    const int d = (int)__umulhi((unsigned)q, mag);
    return q - (int)__umul24((unsigned)d, (unsigned)ns);
}

// TW: apply the stage twiddles (every stage after the first; a first stage,
// ns = 1, has k = 0 and may also run twiddled with its table's ones).
// PAD (row pass): the stage's output blocks of R*ns elements are laid out
// padw elements apart beyond their length, and its inputs are read with the
// stride nb + padr (the previous stage's padded blocks, nb = its block
// length); 0 leaves the plain layout.
template <int R, int NB, bool TW, bool PAD = false>
__device__ __forceinline__ void stage_lds(float2* __restrict__ lds, int L, int ns, unsigned mag,
                                          const float2* __restrict__ tw, int ebase, int es,
                                          int q0, int T, int padw = 0, int padr = 0)
{
    const int nb = L / R;
    const int rstr = (int)__umul24((unsigned)nb, (unsigned)es) + (PAD ? padr : 0);    // uniform
    constexpr int NT = radix_ntw(R);
    // Up to 8 twiddles per thread are loaded before the barrier, so their
    // latency overlaps the LDS reads and the barrier wait (k = q % ns < ns
    // indexes the table for any q).
    constexpr bool PRE = TW && NB * NT <= 8;
    float2 v[NB][R];
    float2 wr[PRE ? NB : 1][PRE ? NT : 1];
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int q = q0 + i * T;
        if (q < nb) {
            const float2* __restrict__ src = lds + ebase + (int)__umul24((unsigned)q, (unsigned)es);
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = src[r * rstr];
        }
        if constexpr (PRE) {
            const int k = ns > 1 ? mod_magic(q, ns, mag) : 0;
            const float2* __restrict__ w = tw + (int)__umul24((unsigned)k, (unsigned)NT);
#pragma unroll
            for (int j = 0; j < NT; j++) wr[i][j] = w[j];
        }
    }
    __syncthreads();
    const int dstr = (int)__umul24((unsigned)ns, (unsigned)es);    // uniform
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int q = q0 + i * T;
        if (q < nb) {
            const int k = (TW && ns > 1) ? mod_magic(q, ns, mag) : 0;
            int o = (int)__umul24((unsigned)(q - k), (unsigned)R) + k;
            if constexpr (PAD) {
                // block q / ns of the output starts padw elements later per block
                const int d = ns > 1 ? (int)__umulhi((unsigned)q, mag) : q;
                o += (int)__umul24((unsigned)d, (unsigned)padw);
            }
            float2* __restrict__ dst = lds + ebase + (int)__umul24((unsigned)o, (unsigned)es);
            DftStore<R>::template run<TW>(v[i], PRE ? wr[PRE ? i : 0] : tw + (int)__umul24((unsigned)k, (unsigned)NT),
                                          [&](int r, float2 x) { dst[r * dstr] = x; });
        }
    }
    __syncthreads();
}

// All radices a kernel instantiation supports are compiled into one switch and
// the register allocation covers the largest, so the composite radices (CX)
// live in their own instantiations: a plan uses them only where they pay.
template <int MAXPTS, bool CX, bool TW, bool PAD = false>
__device__ __forceinline__ void stage_any(float2* lds, const SubPlan& sp, const float2* __restrict__ stw, int s,
                                          int ebase, int es, int q0, int T)
{
    const float2* tw = stw + sp.twoff[s];
    const int ns = sp.ns[s];
    const unsigned mag = sp.mag[s];
    const int padw = PAD && s == sp.padst ? sp.pad : 0;        // uniform
    const int padr = PAD && s == sp.padst + 1 ? sp.pad : 0;
    // Opaque per-stage copies: otherwise every case's per-thread index math
    // is hoisted out of the stage loop and held live across all stages.
    int q0s = q0, ebs = ebase;
    asm volatile("" : "+v"(q0s), "+v"(ebs));
    switch (sp.radix[s]) {
    case 2: stage_lds<2, MAXPTS / 2, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr); break;
    case 3: stage_lds<3, MAXPTS / 3, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr); break;
    case 4: stage_lds<4, MAXPTS / 4, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr); break;
    case 5: stage_lds<5, MAXPTS / 5, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr); break;
    case 8: stage_lds<8, MAXPTS / 8, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr); break;
    case 10:
        if constexpr (CX) stage_lds<10, MAXPTS / 10, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr);
        break;
    case 16:
        if constexpr (CX) stage_lds<16, MAXPTS / 16, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr);
        break;
    case 20:
        if constexpr (CX && MAXPTS >= 20) stage_lds<20, MAXPTS / 20, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr);
        break;
    case 25:
        if constexpr (CX && MAXPTS >= 25) stage_lds<25, MAXPTS / 25, TW, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr);
        break;
    case 32:
        // radix 32 always takes the twiddled form (its table holds ones for a
        // first stage): without the multiplies the compiler schedules its
        // 32-point butterfly into ~20 more registers, costing occupancy
        if constexpr (CX && MAXPTS >= 32) stage_lds<32, MAXPTS / 32, true, PAD>(lds, sp.L, ns, mag, tw, ebs, es, q0s, T, padw, padr);
        break;
    default: break;
    }
}

template <int MAXPTS, bool CX, bool PAD = false>
__device__ __forceinline__ void fft_lds(float2* lds, const SubPlan& sp, const float2* __restrict__ stw,
                                        int ebase, int es, int q0, int T, int nst, int s0 = 0)
{
    // stages [s0, nst); one loop (not a peeled first stage): the loop keeps
    // the compiler from interleaving the first stage with the code before it
    for (int s = s0; s < nst; s++) {
        if (s == 0) stage_any<MAXPTS, CX, false, PAD>(lds, sp, stw, s, ebase, es, q0, T);
        else stage_any<MAXPTS, CX, true, PAD>(lds, sp, stw, s, ebase, es, q0, T);
    }
}

// ---------------------------------------------------------------------------
// 2-bit decode: field s -> {+1, -1, +3.3359, -3.3359}[s]
// (src/IA-32/DataUnpackers.cpp:212, 274, 362: bit 0 = sign, bit 1 = magnitude)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float map2bit(unsigned s)
{
    // built as IEEE bits: sign from bit 0, magnitude 1.0f or 3.3359f chosen by
    // bit 1 through an integer multiply-add on the two bit patterns (exact)
    constexpr unsigned one = 0x3F800000u;
    constexpr unsigned hi = __builtin_bit_cast(unsigned, 3.3359f);
    return __uint_as_float(((s & 1u) << 31) | (__umul24(s >> 1, hi - one) + one));
}

__device__ __forceinline__ unsigned field2(const uint8_t* __restrict__ raw, const Unpack& up, long long j)
{
    if (up.kind == kUnpackSingleChannel) {
        // TwoBitSinglechannelUnpacker layout: 4 samples per byte, oldest in the MSB
        return (raw[j >> 2] >> (6 - 2 * (int)(j & 3))) & 3u;
    }
    // TwoBitUnpacker / Mk5BUnpacker: one byte group of `step` bytes per time sample
    return (raw[j * up.step + up.cbyte] >> up.shift) & 3u;
}

// VLBA level of 2-bit code l = sign bit + 2 x magnitude bit (mark5access
// initluts, mark5_format_vlba.c:90-165): {-HiMag, +1, -1, +HiMag}[l],
// HiMag = 3.3359f (mark5_stream.h:72).  Note the order differs from swspec's
// own map above.
__device__ __forceinline__ float vlba_level(unsigned l)
{
    const float mag = ((l ^ (l >> 1)) & 1u) ? 1.0f : 3.3359f;
    return (l & 1u) ? mag : -mag;
}

// Sample j of one extract_samples call (one FFT) in a one-byte-per-sample
// VLBA mode (fanout F = 2^fsh, ntrack = 8F): chunk c = j / (40128 F), payload
// segment s (0, 1, or the tail 2) of the chunk, payload byte t; frame word
// m = t / F starting at byte i0 = m F selects modulate[m] (negation) and the
// blank zone of i0 (zero outside its data range).  The word's bits per mode
// (vlba_decode_*_decimation1 of mark5_format_vlba.c: 2-bit fanout 1 :4537,
// 2 :5147, 4 :5907; 1-bit fanout 1 :1362, 2 :2021, 4 :2891), for channel c
// and sample u = t - i0 of the word:
//   2-bit F=1: byte 0, sign bit 2c, magnitude 2c+1
//   2-bit F=2: byte c/2, sign bit 4(c%2)+u, magnitude that + 2
//   2-bit F=4: byte k of the 32-bit word reordered (reorder32 :74-76),
//              k = 0, 2, 1, 3 for c = 0..3, sign bit 2u, magnitude 2u+1
//   1-bit F=1: bit c of byte 0;  F=2: byte c/4, bit 2(c%4)+u;
//   F=4: byte c%4, bit 2u + c/4;  levels {+1, -1}[bit].
// brow: the FFT's row of zones, (fsh+1) per payload segment.
__device__ __forceinline__ float vlba_sample(const uint8_t* __restrict__ raw, const Unpack& up,
                                             const int2* __restrict__ brow, unsigned j)
{
    const int fsh = up.vfsh;
    const unsigned P = (unsigned)kVlbaPayload << fsh, CH = (unsigned)kVlbaChunk << fsh;
    const unsigned c = j / CH;
    const unsigned r = j - c * CH;
    const unsigned s = r >= 2u * P ? 2u : (r >= P ? 1u : 0u);
    const unsigned t = r - s * P;
    const unsigned m = t >> fsh;
    const unsigned i0 = m << fsh;
    const unsigned u = t - i0;
    const int2 z = brow[(c * kVlbaSegPerChunk + s) * (unsigned)(fsh + 1) + (i0 >> 15)];
    if (!((int)i0 >= z.x && (int)i0 < z.y)) return 0.0f;
    const uint8_t* __restrict__ w = raw + (j - u);   // the frame word's first byte
    const unsigned ch = (unsigned)up.cbyte;
    float lv;
    if (up.vnbit == 2) {
        unsigned byte, sb, mb;
        if (fsh == 0) {
            byte = w[0];
            sb = 2u * ch;
            mb = sb + 1u;
        } else if (fsh == 1) {
            byte = w[ch >> 1];
            sb = 4u * (ch & 1u) + u;
            mb = sb + 2u;
        } else {
            const unsigned x = (unsigned)w[0] | ((unsigned)w[1] << 8) | ((unsigned)w[2] << 16) |
                               ((unsigned)w[3] << 24);
            const unsigned bits = (x & 0xAA55AA55u) | ((x & 0x55005500u) >> 7) | ((x & 0x00AA00AAu) << 7);
            const unsigned k = (ch >> 1) | ((ch & 1u) << 1);   // 0, 2, 1, 3
            byte = (bits >> (8u * k)) & 0xFFu;
            sb = 2u * u;
            mb = sb + 1u;
        }
        lv = vlba_level(((byte >> sb) & 1u) | (((byte >> mb) & 1u) << 1));
    } else {
        unsigned byte, bit;
        if (fsh == 0) {
            byte = w[0];
            bit = ch;
        } else if (fsh == 1) {
            byte = w[ch >> 2];
            bit = 2u * (ch & 3u) + u;
        } else {
            byte = w[ch & 3u];
            bit = 2u * u + (ch >> 2);
        }
        lv = ((byte >> bit) & 1u) ? -1.0f : 1.0f;
    }
    const bool flip = (up.vmod[m >> 5] >> (m & 31)) & 1u;
    return flip ? -lv : lv;
}

// Sample j of one Mark5BUnpacker::extract_samples call (one FFT) through
// mark5access's decimation-1 Mark5B decoders (mark5_format_mark5b.c:239-
// 1911): bit p = j*B + c*nbit of the FFT's bytes (LSB first), 2-bit levels
// {-HiMag, +1, -1, +HiMag}[bit p + 2 bit p+1] (initluts :59-100), 1-bit
// {+1, -1}[bit p]; zero when the sample's byte group (byte j*B/8) lies
// outside its 10000-byte payload's data zone.  Samples 0..15 of the call are
// the fixed pattern +3.3359, -3.3359, ... where the reference writes
// std::rand() signs (DataUnpackers.cpp:747-750), a stream no other program
// reproduces.
__device__ __forceinline__ float m5b_sample(const uint8_t* __restrict__ raw, const Unpack& up,
                                            const int2* __restrict__ brow, unsigned j)
{
    if (j < 16u) return (j & 1u) ? -3.3359f : 3.3359f;
    const unsigned long long p = (unsigned long long)j * (unsigned)up.vbits + (unsigned)(up.cbyte * up.vnbit);
    const unsigned i0 = (unsigned)(((unsigned long long)j * (unsigned)up.vbits) >> 3);
    const unsigned pl = i0 / (unsigned)kMark5BPayload;
    const int ip = (int)(i0 - pl * (unsigned)kMark5BPayload);
    const int2 z = brow[pl];
    if (!(ip >= z.x && ip < z.y)) return 0.0f;
    const unsigned byte = raw[p >> 3];
    const unsigned sh = (unsigned)(p & 7u);
    if (up.vnbit == 2) return vlba_level(((byte >> sh) & 1u) | (((byte >> (sh + 1u)) & 1u) << 1));
    return ((byte >> sh) & 1u) ? -1.0f : 1.0f;
}

// Either mark5access decoder.
__device__ __forceinline__ float m5a_sample(const uint8_t* __restrict__ raw, const Unpack& up,
                                            const int2* __restrict__ brow, unsigned j)
{
    return up.kind == kUnpackMark5B ? m5b_sample(raw, up, brow, j) : vlba_sample(raw, up, brow, j);
}

// Framed Mk5B: sample j of the FFT whose payload starts at byte `pay` of the
// header-stripped stream; raw is the first frame.
__device__ __forceinline__ unsigned field2_framed(const uint8_t* __restrict__ raw, const Unpack& up, long long pay,
                                                  long long j)
{
    return (raw[mk5b_frame_addr(pay + j * up.step + up.cbyte)] >> up.shift) & 3u;
}

// Any decoder, sample j of one FFT: raw is the FFT start (framed: the first
// frame, with the FFT's payload offset `pay`); brow: VLBA only.
__device__ __forceinline__ float decode_sample(const uint8_t* __restrict__ raw, const Unpack& up,
                                               const int2* __restrict__ brow, long long pay, long long j)
{
    if (is_m5a(up)) return m5a_sample(raw, up, brow, (unsigned)j);
    if (up.framed) return map2bit(field2_framed(raw, up, pay, j));
    return map2bit(field2(raw, up, j));
}

// ---------------------------------------------------------------------------
// pass 1: unpack + window + column FFTs + four-step twiddle
// ---------------------------------------------------------------------------

// Both samples of complex point n (real samples 2n, 2n+1).  MODE 0: single
// channel, 4 samples per byte (both in one byte); 1: one byte per time sample
// at an even address (one 16-bit load); 2: byte groups of `step` bytes.
template <int MODE>
__device__ __forceinline__ void fields_pair(const uint8_t* __restrict__ raw, const Unpack& up, long long n,
                                            unsigned& f0, unsigned& f1)
{
    if (MODE == 0) {
        const unsigned b = raw[n >> 1];
        const int sh = 6 - 4 * (int)(n & 1);
        f0 = (b >> sh) & 3u;
        f1 = (b >> (sh - 2)) & 3u;
    } else if (MODE == 1) {
        const unsigned w = *(const uint16_t*)(raw + 2 * n);   // time samples 2n, 2n+1
        f0 = (w >> up.shift) & 3u;
        f1 = (w >> (8 + up.shift)) & 3u;
    } else {
        f0 = (raw[2 * n * up.step + up.cbyte] >> up.shift) & 3u;
        f1 = (raw[(2 * n + 1) * up.step + up.cbyte] >> up.shift) & 3u;
    }
}

// Strided loops over [tid, total) issue their global loads in batches of U per
// thread before using any of them: the batch index is clamped into range, so
// every load is valid and a batch is one straight-line block the compiler can
// keep in flight together (a plain loop waits for each load in turn).
template <int U> __device__ __forceinline__ int batch_idx(int base, int u, int nth, int total)
{
    return min(base + u * nth, total - 1);
}

// ---------------------------------------------------------------------------
// Radix-64 first column stage on the matrix cores (split FP16, FP32 sums)
// ---------------------------------------------------------------------------

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
constexpr float kMfLo = 2048.f;   // lo parts are stored times 2^11
constexpr int kMfPairs = 8;       // (row tile, column tile) pairs per wave

// x = hi + lo / 2^11 with hi, lo in FP16 (round to nearest): the relative
// error of the pair is about 2^-22.  Returned as the bit patterns of one
// LDS element: .x = (hi re, hi im), .y = (lo re, lo im).
// lo must be taken against the hi that is stored.  Left alone, the compiler
// contracts the multiply that produced a and b into one of the two uses of
// (_Float16)a (v_fma_mixlo_f16 rounds the exact product) and not the other
// (v_cvt_pk_f16_f32 of the FP32 product); a value on an FP16 rounding tie
// then came back one FP16 ulp off.  The empty asm statements pin a, b and hi
// to single values.
__device__ __forceinline__ float2 split16(float a, float b)
{
    asm volatile("" : "+v"(a), "+v"(b));
    const h2_t h0 = {(_Float16)a, (_Float16)b};
    unsigned hb = __builtin_bit_cast(unsigned, h0);
    asm volatile("" : "+v"(hb));
    const h2_t hi = __builtin_bit_cast(h2_t, hb);
    const h2_t lo = {(_Float16)((a - (float)hi.x) * kMfLo), (_Float16)((b - (float)hi.y) * kMfLo)};
    return make_float2(__uint_as_float(hb), __builtin_bit_cast(float, lo));
}

// The stage's 64-point DFTs X[r] = sum_s x[s] W_64^{rs} of all V = (N1/64)*B
// column vectors of the workgroup form one real product
//     [X re/im interleaved] (128 x V) = F (128 x 128) [x re/im interleaved] (128 x V).
// The fill stored x split (split16); F = Fhi + Flo/2^11 comes from a host
// table in FP64.  D = Fhi xhi + (Fhi xlo + Flo xhi)/2^11: the dropped lo*lo
// term is ~2^-22 relative.  v_mfma_f32_16x16x32_f16 makes one 16 x 16 tile
// from K = 32 (16 complex inputs) per instruction.  Wave w takes row tiles
// rt = w, w + W, ... (16 rows = 8 frequencies each) and every column tile ct
// of 16 vectors; the A operands of a row tile (4 K steps x hi/lo) stay in
// registers.  Lane l covers row / column l % 16 of a tile and K slots
// 8(l/16) + t, t < 8, which both operands map to complex input
// s = 16 kc + 4(l/16) + t/2, part t % 2; the result rows of lane l are
// 4(l/16) + t, t < 4.  Vector v = q*B + b is butterfly q of column b: its
// input s sits at lp[v + s*V] (element q + s*N1/64) and output r goes to
// element 64 q + r, the first Stockham stage's order.  A last column tile
// that V does not fill computes zeros in its lanes v >= V and stores nothing
// there.  All outputs are held in registers until every wave has read its
// inputs (in-place stage).
__device__ __forceinline__ void stage_mfma64(float2* __restrict__ lp, const uint4* __restrict__ mft, int N1, int B)
{
    const int lane = (int)(threadIdx.x & 63u);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int W = (int)(blockDim.x >> 6);
    const int g = lane >> 4, j = lane & 15;
    const int V = (N1 >> 6) * B;
    const int nct = (V + 15) >> 4;
    f4_t res[kMfPairs];
    h8_t ah[4], al[4];
#pragma unroll
    for (int i = 0; i < kMfPairs; i++) {
        const int rt = wv + W * (i / nct);
        const int ct = i % nct;
        if (rt >= 8) break;
        if (ct == 0) {
#pragma unroll
            for (int kc = 0; kc < 4; kc++) {
                ah[kc] = __builtin_bit_cast(h8_t, mft[((rt * 4 + kc) * 2 + 0) * 64 + lane]);
                al[kc] = __builtin_bit_cast(h8_t, mft[((rt * 4 + kc) * 2 + 1) * 64 + lane]);
            }
        }
        const int v = ct * 16 + j;
        const bool in = v < V;
        f4_t d = {0.f, 0.f, 0.f, 0.f}, e = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            float2 x[4];
#pragma unroll
            for (int t = 0; t < 4; t++) x[t] = in ? lp[v + (16 * kc + 4 * g + t) * V] : make_float2(0.f, 0.f);
            const uint4 bh4 = make_uint4(__float_as_uint(x[0].x), __float_as_uint(x[1].x), __float_as_uint(x[2].x),
                                         __float_as_uint(x[3].x));
            const uint4 bl4 = make_uint4(__float_as_uint(x[0].y), __float_as_uint(x[1].y), __float_as_uint(x[2].y),
                                         __float_as_uint(x[3].y));
            const h8_t bh = __builtin_bit_cast(h8_t, bh4), bl = __builtin_bit_cast(h8_t, bl4);
            d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[kc], bh, d, 0, 0, 0);
            e = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[kc], bl, e, 0, 0, 0);
            e = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[kc], bh, e, 0, 0, 0);
        }
        res[i] = d + e * (1.f / kMfLo);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMfPairs; i++) {
        const int rt = wv + W * (i / nct);
        const int ct = i % nct;
        if (rt >= 8) break;
        const int v = ct * 16 + j;
        if (v >= V) continue;
        const int q = v / B, b = v - q * B;
        const int o = (q * 64 + 8 * rt + 2 * g) * B + b;
        lp[o] = make_float2(res[i].x, res[i].y);
        lp[o + B] = make_float2(res[i].z, res[i].w);
    }
    __syncthreads();
}

struct P1Args {
    int N1, N2;
    long long M;
    int B;            // columns per workgroup pipeline (any divisor of N2)
    int fpw, pth;     // FFT pipelines per workgroup, threads per pipeline (B * T1)
    int twshift;
    SubPlan sp;
    const float2* stw;
    const float2* twA;
    const float2* twB;
    const uint4* mft; // MF: DFT-64 operand table
    WindowGen win;
    int zfast;        // jobs share raw bytes: job index fastest in the workgroup order
    int rowfill;      // 1: row fill (fill_rows, k_pass1<.., RF = B>), 0: column fill
    const float4* wcol;  // window column constants (row fill)
    JobDev jobs[kMaxJobs];
};

// Workgroup id -> (x, y, z) such that consecutive x land on the same XCD.
// Workgroups are dispatched round-robin over the 8 XCDs by linear id; when the
// grid size is a multiple of 8, linear id L runs on XCD L % 8, so logical id
// (L % 8) * (total / 8) + L / 8 keeps neighbouring column blocks (which share
// cache lines of the input and of Y) in one L2.  With zfast the job index z
// varies fastest: the channels of one multi-channel source read the same raw
// bytes, and adjacent logical ids find them in the same L2 at nearly the same
// time instead of re-reading them from HBM once per channel.
__device__ __forceinline__ void xcd_remap(unsigned& x, unsigned& y, unsigned& z, bool zfast = false)
{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned total = gx * gy * gz;
    unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if ((total & 7u) == 0) L = (L & 7u) * (total >> 3) + (L >> 3);
    if (zfast) {
        z = L % gz;
        const unsigned xy = L / gz;
        x = xy % gx;
        y = xy / gx;
        return;
    }
    x = L % gx;
    const unsigned yz = L / gx;
    y = yz % gy;
    z = yz / gy;
}

// Window on the fly: real sample j = 2(N2 e + c) + r has angle
// theta = phi_e + psi_r, phi_e = 2 pi (2 N2 e) / D from a per-row table and
// psi_r = 2 pi (2c + r + delta) / D fixed per thread (its column c never
// changes: nth is a multiple of B).  Expanding the cosine sums,
//   w_r = a0 + cos(phi) P_r + sin(phi) Q_r + cos(2 phi) R_r + sin(2 phi) S_r
//   P = a1 cos psi + a2 sin psi,  Q = a2 cos psi - a1 sin psi,
//   R = a3 cos 2psi,              S = -a3 sin 2psi
// (per thread, FP64, rounded once).  phi is linear in the row, so the rows
// e0 + u*estep of a batch take (cos, sin)(phi) as the product of the table
// entries of e0 and of u*estep: one product each, nothing accumulates.
struct WinCoef {
    float P[2], Q[2], R2[2], S2[2];
    float a0;
    int mode;   // 0: no window (w = 1, no arithmetic), 1: no cos(2 theta) term, 2: all terms
};

__device__ __forceinline__ WinCoef win_coef(const WindowGen& wg, int cb)
{
    WinCoef w;
#pragma unroll
    for (int r = 0; r < 2; r++) w.P[r] = w.Q[r] = w.R2[r] = w.S2[r] = 0.f;
    if (wg.kind != 0) {
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const double x = ((double)(2 * cb + r) + wg.delta) / wg.D;
            double s, c;
            sincospi(2.0 * x, &s, &c);
            const double c2 = c * c - s * s, s2 = 2.0 * s * c;
            w.P[r] = (float)((double)wg.a1 * c + (double)wg.a2 * s);
            w.Q[r] = (float)((double)wg.a2 * c - (double)wg.a1 * s);
            w.R2[r] = (float)((double)wg.a3 * c2);
            w.S2[r] = (float)(-(double)wg.a3 * s2);
        }
    }
    w.a0 = wg.a0;
    w.mode = wg.kind == 0 ? 0 : (wg.a3 == 0.f ? 1 : 2);
    return w;
}

// Decode + window of one FFT's points into the column layout:
// lp[e*B + b] = z[N2*e + c0 + b] windowed; idx == e*B + b.  A thread keeps
// its column b = bcol (point column cb = c0 + bcol); its row e starts at et
// and advances by estep = nth / B per strided step.  Rows are clamped to
// N1 - 1 for the loads of a partial batch (values unused).  Point indices
// are < M <= 2^30 and the factors < 2^14: 24-bit products.  raw is the FFT's
// first byte (framed Mk5B: the first frame, payload offset pay); brow its
// row of blank zones (mark5access decodes).
template <bool MF>
__device__ __forceinline__ void fill_points(float2* __restrict__ lp, const uint8_t* __restrict__ raw,
                                            const Unpack& up, long long pay, const int2* __restrict__ brow,
                                            const WindowGen& wg, const WinCoef& wc, int tid, int nth, int B,
                                            int N1, int N2, int cb, int et)
{
    const int total = N1 * B;
    const int estep = nth / B;
    const bool even = ((uintptr_t)raw & 1) == 0;
    auto fill = [&](auto mode, auto wmc) __attribute__((always_inline)) {
        constexpr int MODE = decltype(mode)::value;
        constexpr int WM = decltype(wmc)::value;
        // VLBA decode holds more per sample (segment zone, modulation word), and
        // framed Mk5B 64-bit frame addresses: smaller batches keep them under
        // the register budget of the plain 2-bit modes
        constexpr int U = MODE >= 3 ? 2 : 8;
        // The thread's k-th point is on row et + k*estep.  Its window phase is
        // the table entry of row et + (k & ~7)*estep times that of row
        // (k & 7)*estep: a function of the row alone, so every decode mode and
        // batch size windows a row by the same operations.
        constexpr int G = 8;
        static_assert(G % U == 0, "batches must not straddle a phase group");
        // batch j: points k = j*U + u at base + u*nth (row e0 + u*estep); FULL:
        // all of them in range (no clamps or checks), else the thread's last,
        // partial batch
        auto batch = [&](int base, int j, auto fullc) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(fullc)::value;
            float x0[U], x1[U];
            const int k0 = j * U;
            const int e0 = et + k0 * estep;
            const float2 anc = WM ? wg.phi[et + (k0 & ~(G - 1)) * estep] : make_float2(1.f, 0.f);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = FULL ? e0 + u * estep : min(e0 + u * estep, N1 - 1);
                const unsigned n = __umul24((unsigned)N2, (unsigned)e) + (unsigned)cb;
                if constexpr (MODE == 3) {
                    x0[u] = m5a_sample(raw, up, brow, 2u * n);
                    x1[u] = m5a_sample(raw, up, brow, 2u * n + 1u);
                } else if constexpr (MODE == 4) {
                    x0[u] = map2bit(field2_framed(raw, up, pay, 2LL * n));
                    x1[u] = map2bit(field2_framed(raw, up, pay, 2LL * n + 1));
                } else {
                    unsigned f0, f1;
                    fields_pair<MODE>(raw, up, (long long)n, f0, f1);
                    x0[u] = map2bit(f0);
                    x1[u] = map2bit(f1);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int idx = base + u * nth;
                if (FULL || idx < total) {
                    float w0 = 1.f, w1 = 1.f;
                    if constexpr (WM != 0) {
                        const int g = U == G ? u : ((k0 + u) & (G - 1));   // uniform
                        float2 p = anc;
                        if (g != 0) {
                            // uniform load; rows past the table are never stored
                            const float2 r = wg.phi[min(g * estep, N1 - 1)];
                            p = cmul(anc, r);
                        }
                        w0 = wc.a0 + p.x * wc.P[0] + p.y * wc.Q[0];
                        w1 = wc.a0 + p.x * wc.P[1] + p.y * wc.Q[1];
                        if constexpr (WM == 2) {
                            const float c2 = p.x * p.x - p.y * p.y, s2 = 2.f * p.x * p.y;
                            w0 += c2 * wc.R2[0] + s2 * wc.S2[0];
                            w1 += c2 * wc.R2[1] + s2 * wc.S2[1];
                        }
                    }
                    if constexpr (MF) lp[idx] = split16(x0[u] * w0, x1[u] * w1);
                    else lp[idx] = make_float2(x0[u] * w0, x1[u] * w1);
                }
            }
        };
        const int nfull = total - (U - 1) * nth;
        int base = tid, j = 0;
        for (; base < nfull; base += U * nth, j++) batch(base, j, std::true_type());
        if (base < total) batch(base, j, std::false_type());
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    auto fill_w = [&](auto mode) __attribute__((always_inline)) {
        if (wc.mode == 0) fill(mode, I0());
        else if (wc.mode == 1) fill(mode, I1());
        else fill(mode, I2());
    };
    if (up.kind == kUnpackSingleChannel) fill_w(I0());
    else if (is_m5a(up)) fill(std::integral_constant<int, 3>(), I2());
    else if (up.framed) fill_w(std::integral_constant<int, 4>());   // same window code as mode 1
    else if (up.step == 1 && even) fill_w(I1());
    else fill_w(I2());
}

// Row fill: decode + window of one FFT's points with each thread taking whole
// rows of the workgroup's BC columns (BC = 4 or 8), for layouts with one byte
// per time sample (2-bit multi-channel, payload or framed Mk5B) and the
// single-channel 2-bit layout.  Row e's BC
// points z[N2 e + c0 + b] are 2 BC consecutive raw bytes: one 8- or 16-byte
// load (the launcher checks the alignment; framed: an aligned group never
// straddles a 10000-byte payload).  The window of sample r of column c0 + b
// in row e is
//   w = a0 + cos(phi_e) P + sin(phi_e) Q + cos(2 phi_e) R + sin(2 phi_e) S
// with (cos, sin)(phi_e) the row's table entry and (P, Q, R, S) the column's
// constants (the column fill's per-thread constants, tabulated per column on
// the host in FP64): the same for every thread of the workgroup, so they sit
// in scalar registers.  lp[e*BC + b] is written as one 32- or 64-byte row
// segment.  (A 16-entry LDS table of level pairs indexed by the two fields
// measured no faster than map2bit's integer build.)
// SC: single-channel 2-bit layout (4 samples per byte, oldest in the MSB):
// the row's 2 BC samples are BC / 2 bytes, one 16- or 32-bit load.
template <int BC, int WM, bool FR, bool SC = false>
__device__ __forceinline__ void fill_rows(float2* __restrict__ lp, const uint8_t* __restrict__ raw,
                                          const Unpack& up, long long pay,
                                          const WindowGen& wg, const float4* __restrict__ wcol, int tid, int nth,
                                          int N1, int N2, int c0)
{
    static_assert(BC == 4 || BC == 8, "row fill: 4 or 8 columns");
    static_assert(!(SC && FR), "framed input is multi-channel");
    using Raw = typename std::conditional<SC, typename std::conditional<BC == 4, uint16_t, uint32_t>::type,
                                          typename std::conditional<BC == 4, uint2, uint4>::type>::type;
    constexpr int NW = SC ? 1 : BC / 2;   // raw dwords per row segment
    float4 K[BC][2];
#pragma unroll
    for (int b = 0; b < BC; b++)
#pragma unroll
        for (int r = 0; r < 2; r++) K[b][r] = WM ? wcol[2 * (c0 + b) + r] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float a0 = wg.a0;
    const unsigned sh = (unsigned)up.shift;
    constexpr int U = 4;   // rows in flight per thread
    for (int e0 = tid; e0 < N1; e0 += U * nth) {
        Raw q[U];
        float2 z[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = min(e0 + u * nth, N1 - 1);   // clamped: values unused past N1
            const unsigned pt0 = __umul24((unsigned)N2, (unsigned)e) + (unsigned)c0;   // first point of the row segment
            const uint8_t* src;
            if constexpr (SC) src = raw + (pt0 >> 1);
            else if constexpr (FR) src = raw + mk5b_frame_addr(pay + 2LL * pt0);
            else src = raw + 2u * pt0;
            q[u] = *(const Raw*)src;
            z[u] = WM ? wg.phi[e] : make_float2(1.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + u * nth;
            if (e >= N1) break;
            float c2 = 0.f, s2 = 0.f;
            if constexpr (WM == 2) {
                c2 = z[u].x * z[u].x - z[u].y * z[u].y;
                s2 = 2.f * z[u].x * z[u].y;
            }
            float2 pt[BC];
#pragma unroll
            for (int d = 0; d < NW; d++) {
                // multi-channel: byte lanes of dword d hold the fields of samples
                // 2b, 2b+1 (b = 2d, 2d+1) of this channel; single channel: byte
                // b/2 holds points b and b+1, oldest sample in the MSB
                unsigned t;
                if constexpr (SC) t = (unsigned)q[u];
                else t = (((const unsigned*)&q[u])[d] >> sh) & 0x03030303u;
#pragma unroll
                for (int h = 0; h < (SC ? BC : 2); h++) {
                    const int b = SC ? h : 2 * d + h;
                    unsigned f0, f1;
                    if constexpr (SC) {
                        const int s0 = 8 * (b >> 1) + 6 - 4 * (b & 1);
                        f0 = (t >> s0) & 3u;
                        f1 = (t >> (s0 - 2)) & 3u;
                    } else {
                        f0 = (t >> (16 * h)) & 3u;
                        f1 = (t >> (16 * h + 8)) & 3u;
                    }
                    const float2 lv = make_float2(map2bit(f0), map2bit(f1));
                    float w0 = 1.f, w1 = 1.f;
                    if constexpr (WM != 0) {
                        w0 = a0 + z[u].x * K[b][0].x + z[u].y * K[b][0].y;
                        w1 = a0 + z[u].x * K[b][1].x + z[u].y * K[b][1].y;
                        if constexpr (WM == 2) {
                            w0 += c2 * K[b][0].z + s2 * K[b][0].w;
                            w1 += c2 * K[b][1].z + s2 * K[b][1].w;
                        }
                    }
                    pt[b] = make_float2(lv.x * w0, lv.y * w1);
                }
            }
            float4* dst = (float4*)(lp + (size_t)e * BC);
#pragma unroll
            for (int b = 0; b < BC; b += 2) dst[b / 2] = make_float4(pt[b].x, pt[b].y, pt[b + 1].x, pt[b + 1].y);
        }
    }
}

// Row fill with the first column stage fused (F1: its radix; a plain Stockham
// first stage, ns = 1).  A thread takes butterfly q of a pair of adjacent
// columns cA, cA+1 of the workgroup's BC: the F1 rows q + r*(N1/F1) of those
// two columns are F1 four-byte loads (2 points = 4 samples of one byte each;
// framed: a 4-byte group never straddles a 10000-byte payload), decoded and
// windowed as fill_rows does, and both columns' F1-point DFTs run in
// registers; output r of butterfly q is element q*F1 + r of its column,
// written for both columns as one 16-byte row piece.  This saves the fill's
// LDS write and the first stage's LDS read of every point and one barrier.
// SC: single-channel 2-bit layout, the two points one byte (oldest sample in
// the MSB).
template <int BC, int F1, int WM, bool FR, bool SC = false>
__device__ __forceinline__ void fill_stage1(float2* __restrict__ lp, const uint8_t* __restrict__ raw,
                                            const Unpack& up, long long pay, const WindowGen& wg,
                                            const float4* __restrict__ wcol, int tid, int nth, int N1, int N2,
                                            int c0)
{
    static_assert(BC == 4, "fused first stage: 4 columns");
    constexpr int NP = BC / 2;   // column pairs
    const int nb = N1 / F1;
    const unsigned sh = (unsigned)up.shift;
    const float a0 = wg.a0;
    for (int it = tid; it < nb * NP; it += nth) {
        const int q = it / NP;
        const int cA = 2 * (it - q * NP);
        float4 K[2][2];   // window constants of columns cA, cA+1, samples 0, 1
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 2; r++)
                K[b][r] = WM ? wcol[2 * (c0 + cA + b) + r] : make_float4(0.f, 0.f, 0.f, 0.f);
        unsigned w[F1];
        float2 z[F1];
#pragma unroll
        for (int r = 0; r < F1; r++) {
            const int e = q + r * nb;
            const unsigned pt = __umul24((unsigned)N2, (unsigned)e) + (unsigned)(c0 + cA);
            if constexpr (SC) w[r] = raw[pt >> 1];
            else if constexpr (FR) w[r] = *(const unsigned*)(raw + mk5b_frame_addr(pay + 2LL * pt));
            else w[r] = *(const unsigned*)(raw + 2u * pt);
            z[r] = WM ? wg.phi[e] : make_float2(1.f, 0.f);
        }
        float2 v[2][F1];
#pragma unroll
        for (int r = 0; r < F1; r++) {
            const unsigned t = SC ? w[r] : (w[r] >> sh) & 0x03030303u;
            float c2 = 0.f, s2 = 0.f;
            if constexpr (WM == 2) {
                c2 = z[r].x * z[r].x - z[r].y * z[r].y;
                s2 = 2.f * z[r].x * z[r].y;
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const unsigned f0 = SC ? (t >> (6 - 4 * b)) & 3u : (t >> (16 * b)) & 3u;
                const unsigned f1 = SC ? (t >> (4 - 4 * b)) & 3u : (t >> (16 * b + 8)) & 3u;
                float w0 = 1.f, w1 = 1.f;
                if constexpr (WM != 0) {
                    w0 = a0 + z[r].x * K[b][0].x + z[r].y * K[b][0].y;
                    w1 = a0 + z[r].x * K[b][1].x + z[r].y * K[b][1].y;
                    if constexpr (WM == 2) {
                        w0 += c2 * K[b][0].z + s2 * K[b][0].w;
                        w1 += c2 * K[b][1].z + s2 * K[b][1].w;
                    }
                }
                v[b][r] = make_float2(map2bit(f0) * w0, map2bit(f1) * w1);
            }
        }
        float2 o[2][F1];
#pragma unroll
        for (int b = 0; b < 2; b++)
            DftStore<F1>::template run<false>(v[b], nullptr, [&](int r, float2 x) { o[b][r] = x; });
        float4* dst = (float4*)(lp + (size_t)q * F1 * BC + cA);
#pragma unroll
        for (int r = 0; r < F1; r++) dst[r * (BC / 2)] = make_float4(o[0][r].x, o[0][r].y, o[1][r].x, o[1][r].y);
    }
}

// MF: the first column stage is radix 64 on the matrix cores (stage_mfma64;
// one pipeline per workgroup, whole waves)
// RF: row fill of RF = 4 or 8 columns (fill_rows; launcher: a.rowfill), 0:
// column fill.  Separate instantiations, so that each is register-allocated
// for its own fill.
// F1 (with RF): the first column stage, of radix F1, fused into the row fill
// (fill_stage1; multi-channel layouts), so the LDS stages start at stage 1.
// FX: a fixed plan (fixed_plan(FX): the 3.2M or the 2^20 column plan, one
// pipeline) with its strides and bounds as constants; 0: the plan's values.
template <int MAXPTS, bool CX, bool MF = false, int RF = 0, int F1 = 0, int FX = 0>
__global__ void __launch_bounds__(512) k_pass1(P1Args a)
{
    extern __shared__ float2 lds[];
    unsigned bx, by, bz;
    xcd_remap(bx, by, bz, a.zfast);
    const JobDev J = a.jobs[bz];
    // a workgroup runs `fpw` independent pipelines, one FFT each (short columns)
    constexpr FixedPlan kFp = fixed_plan(FX);
    const int fpw = FX ? 1 : a.fpw, pth = FX ? kFp.p1_threads : a.pth;
    if ((int)by * fpw >= J.nf) return;   // no FFT of this job here (whole workgroup)
    int pipe = 0;
    while (pipe + 1 < fpw && threadIdx.x >= (unsigned)((pipe + 1) * pth)) pipe++;
    const int f = (int)by * fpw + pipe;
    const bool active = f < J.nf;
    const int B = FX ? kFp.B : a.B, N1 = FX ? kFp.N1 : a.N1, N2 = FX ? kFp.N2 : a.N2;
    const int c0 = (int)bx * B;
    const int tid = threadIdx.x - pipe * pth, nth = pth;
    float2* __restrict__ lp = lds + pipe * (N1 * B);
    const Unpack up = J.up;
    const uint8_t* __restrict__ raw = up.framed ? J.raw : job_fft_ptr(J, active ? f : 0);
    const long long pay = J.pay0 + (long long)(active ? f : 0) * J.fft_stride;   // framed only
    const int et = tid / B;            // the thread's first row (and first butterfly)
    const int bcol = tid - et * B;     // its column, fixed
    const int cb = c0 + bcol;
    const int total = N1 * B;
    const int estep = nth / B;
    if constexpr (RF != 0 && F1 != 0) {
        if (active) {
            auto rows1 = [&](auto fr, auto sc) __attribute__((always_inline)) {
                constexpr bool FR = decltype(fr)::value;
                constexpr bool SC = decltype(sc)::value;
                const int wm = a.win.kind == 0 ? 0 : (a.win.a3 == 0.f ? 1 : 2);
                if (wm == 0) fill_stage1<RF, F1, 0, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
                else if (wm == 1) fill_stage1<RF, F1, 1, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
                else fill_stage1<RF, F1, 2, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
            };
            if (up.kind == kUnpackSingleChannel) rows1(std::false_type(), std::true_type());
            else if (up.framed) rows1(std::true_type(), std::false_type());
            else rows1(std::false_type(), std::false_type());
        }
    } else if constexpr (RF != 0) {
        if (active) {
            auto rows = [&](auto fr, auto sc) __attribute__((always_inline)) {
                constexpr bool FR = decltype(fr)::value;
                constexpr bool SC = decltype(sc)::value;
                const int wm = a.win.kind == 0 ? 0 : (a.win.a3 == 0.f ? 1 : 2);
                if (wm == 0) fill_rows<RF, 0, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
                else if (wm == 1) fill_rows<RF, 1, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
                else fill_rows<RF, 2, FR, SC>(lp, raw, up, pay, a.win, a.wcol, tid, nth, N1, N2, c0);
            };
            if (up.kind == kUnpackSingleChannel) rows(std::false_type(), std::true_type());
            else if (up.framed) rows(std::true_type(), std::false_type());
            else rows(std::false_type(), std::false_type());
        }
    } else {
        const WinCoef wc = win_coef(a.win, cb);
        // MODE 3 (VLBA): samples decoded through the FFT's row of blank zones
        const int2* __restrict__ brow = is_m5a(up) ? up.vblank + (size_t)(active ? f : 0) * up.vnseg : nullptr;
        if (active) fill_points<MF>(lp, raw, up, pay, brow, a.win, wc, tid, nth, B, N1, N2, cb, et);
    }
    __syncthreads();

    // A plain last radix (2, 3, 4, 5, 8) after at least one LDS stage runs
    // fused with the store below: its butterflies read LDS and write their
    // outputs straight to Y, saving an LDS write, a barrier and a read per
    // point.
    constexpr SubPlan kFx1 = kFp.cols;
    const SubPlan& sp = pick_ref<FX != 0>(kFx1, a.sp);
    const int rlast = sp.radix[sp.nst - 1];
    const bool fuse = sp.nst >= 2 && (rlast == 2 || rlast == 3 || rlast == 4 || rlast == 5 || rlast == 8);
    const int ebase = pipe * (N1 * B) + bcol;
    if constexpr (MF) stage_mfma64(lds, a.mft, N1, B);
    fft_lds<MAXPTS, CX>(lds, sp, a.stw, ebase, B, et, estep, fuse ? sp.nst - 1 : sp.nst, (MF || F1) ? 1 : 0);

    float2* __restrict__ Y = J.Y + (long long)(active ? f : 0) * a.M;
    const unsigned mask = (1u << a.twshift) - 1u;
    auto tw_at = [&](unsigned e) { return cmul(a.twA[e >> a.twshift], a.twB[e & mask]); };
    if (fuse) {
        if (!active) return;
        // Last stage: butterfly q < ns = N1/R (k = q) reads q + r*ns and writes
        // output row q + r*ns, times W_M^{cb (q + r ns)}.  A thread's butterflies
        // q = et + i*estep take W_M^{cb q} from the split tables at the start of
        // each chunk (<= 8 butterflies) and by the step W_M^{cb estep} within it;
        // output r steps by W_M^{cb ns}.
        auto last = [&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            constexpr int NB = MAXPTS / R;
            const int ns = N1 / R;
            const float2* __restrict__ tw = a.stw + sp.twoff[sp.nst - 1];
            const float2 s_ns = tw_at(__umul24((unsigned)cb, (unsigned)ns));       // cb*ns < M
            const float2 s_q = tw_at(__umul24((unsigned)cb, (unsigned)estep));     // cb*estep < M
            const int rstr = (int)__umul24((unsigned)ns, (unsigned)B);
            // Chunks of CH butterflies, every load of a chunk (inputs, stage
            // twiddles, four-step anchor) before its first store: Y might alias
            // the tables as far as the compiler knows, so a load placed after a
            // store is issued late and its latency is exposed.
            constexpr int CH = (16 / R) < NB ? (16 / R) : NB;   // ~50-60 VGPRs of data, <= 8
            float2 wq = make_float2(1.f, 0.f);
#pragma unroll
            for (int i0 = 0; i0 < NB; i0 += CH) {
                if (et + i0 * estep >= ns) break;
                float2 v[CH][R], w[CH][R - 1], anc = make_float2(1.f, 0.f);
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const int q = min(et + (i0 + c) * estep, ns - 1);   // clamped: values unused past ns
                    const float2* __restrict__ src = lds + ebase + (int)__umul24((unsigned)q, (unsigned)B);
#pragma unroll
                    for (int r = 0; r < R; r++) v[c][r] = src[r * rstr];
#pragma unroll
                    for (int r = 0; r < R - 1; r++) w[c][r] = tw[(int)__umul24((unsigned)q, (unsigned)(R - 1)) + r];
                    if (c == 0) anc = tw_at(__umul24((unsigned)cb, (unsigned)q));   // chunk anchor
                }
#pragma unroll
                for (int c = 0; c < CH; c++) {
                    const int i = i0 + c;
                    const int q = et + i * estep;
                    wq = (c == 0) ? anc : cmul(wq, s_q);
                    if (q >= ns) continue;
                    float2 wr = wq;
                    DftStore<R>::template run<true>(v[c], w[c], [&](int r, float2 x) {
                        if (r > 0) wr = cmul(wr, s_ns);
                        const unsigned o = __umul24((unsigned)(q + r * ns), (unsigned)N2) + (unsigned)cb;
                        Y[o] = cmul(x, wr);
                    });
                }
            }
        };
        switch (rlast) {
        case 2: last(std::integral_constant<int, 2>()); break;
        case 3: last(std::integral_constant<int, 3>()); break;
        case 4: last(std::integral_constant<int, 4>()); break;
        case 5: last(std::integral_constant<int, 5>()); break;
        default: last(std::integral_constant<int, 8>()); break;
        }
        return;
    }
    constexpr int U = 8;
    // Y[k1*N2 + n2] = lp[k1*B + bcol] * W_M^{n2 k1}, n2 = cb fixed per thread.
    // Along a thread's rows k1 = k0 + u*estep the twiddle advances by the fixed
    // factor W_M^{cb*estep}: each batch of U rows takes W_M^{cb*k0} from the
    // split tables (FP64-derived, one product) and the other U-1 by complex
    // multiplication, so the relative error stays within ~U ulp.
    const float2 tstep = tw_at(__umul24((unsigned)cb, (unsigned)estep));   // cb*estep < N2*N1 = M
    // batch at lp[base + u*nth] = rows k0 + u*estep
    auto store = [&](int base, int k0, auto fullc) __attribute__((always_inline)) {
        constexpr bool FULL = decltype(fullc)::value;   // all U rows in range
        float2 w = tw_at(__umul24((unsigned)cb, (unsigned)k0));              // n2 k1 < M
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k1 = k0 + u * estep;
            if (FULL || k1 < N1) {
                const unsigned o = __umul24((unsigned)k1, (unsigned)N2) + (unsigned)cb;
                Y[o] = cmul(lp[base + u * nth], w);
            }
            w = cmul(w, tstep);
        }
    };
    if (!active) return;
    const int nfull = total - (U - 1) * nth;
    int base = tid, k0 = et;
    for (; base < nfull; base += U * nth, k0 += U * estep) store(base, k0, std::true_type());
    if (base < total) store(base, k0, std::false_type());
}

// ---------------------------------------------------------------------------
// pass 2: row FFTs, r2c split, detection, resident accumulation
// ---------------------------------------------------------------------------

struct P2Args {
    int N1, N2;
    long long M;
    int nfft;
    int kahan;
    int S;            // slices: slice s sums FFTs [s*per, (s+1)*per) into accumulator slice s
    int fpw;          // slice pipelines per workgroup
    int pth;          // threads per pipeline (2 * NSRC * T)
    int pipe_lds;     // float2 units of LDS per pipeline
    int rowst;        // float2 units of LDS per row (N2, or more with a padded stage)
    long long accstride;   // floats between accumulator slices of a job (2*(M+1))
    long long xaccstride;  // floats between cross-power slices (4*(M+1))
    SubPlan sp;
    const float2* stw;
    const float2* rA;
    const float2* rB;
    JobDev jobs[kMaxJobs];
    XpolDev xp[kMaxJobs / 2];
};

// F1: plain first row radix run from registers (0: none; see fuse1 below)
// FX: a fixed row plan (fixed_plan(FX): 3.2M or 2^20) with its strides and
// bounds as constants, one pipeline per workgroup; 0: the plan's values.
template <int MAXPTS, int NSRC, bool CX, int F1 = 0, int FX = 0>
__global__ void __launch_bounds__(512) k_pass2(P2Args a)
{
    extern __shared__ float2 lds_all[];
    constexpr FixedPlan kFp = fixed_plan(FX);
    constexpr SubPlan kFx = kFp.rows;
    const SubPlan& SP = pick_ref<FX != 0>(kFx, a.sp);
    const int p = blockIdx.x, grp = blockIdx.y;
    const int N1 = a.N1, N2 = FX ? kFp.N2 : a.N2;
    const int RS = FX ? kFp.rowst : a.rowst;   // LDS row stride (>= N2)
    const int rowA = p;
    const int rowB = (p == 0) ? 0 : N1 - p;
    const bool self = (rowA == rowB);
    // pipeline (slice) of this thread: blockDim == fpw * pth
    const int pth = FX ? 2 * NSRC * kFx.T : a.pth;
    const int fpw = FX ? 1 : a.fpw;
    int pipe = 0;
    while (pipe + 1 < fpw && threadIdx.x >= (unsigned)((pipe + 1) * pth)) pipe++;
    const int tid = threadIdx.x - pipe * pth, nth = pth;
    const int sl_i = blockIdx.z * fpw + pipe;          // slice index
    const bool sact = sl_i < a.S;
    const int pipe_lds = FX ? 2 * NSRC * RS : a.pipe_lds;
    float2* __restrict__ lds = lds_all + pipe * pipe_lds;
    constexpr int NSL = 2 * NSRC;
    const int T = FX ? kFx.T : a.sp.T;       // pth == NSL * T
    int slot = 0;
#pragma unroll
    for (int s = 1; s < NSL; s++) slot += (tid >= s * T) ? 1 : 0;
    const int q0 = tid - slot * T;
    float2* Yp[NSRC];
    float* accp[NSRC];
#pragma unroll
    for (int s = 0; s < NSRC; s++) {
        Yp[s] = a.jobs[grp * NSRC + s].Y;
        accp[s] = a.jobs[grp * NSRC + s].acc + (sact ? sl_i : 0) * a.accstride;
    }
    // Bins in conjugate pairs: with E = (Z[k] + conj Z[M-k])/2 and
    // O = (Z[k] - conj Z[M-k])/(2i), X[k] = E + W_N^k O and X[M-k] =
    // conj(E - W_N^k O), so one pair of row values gives both bins.  Pair pi
    // of the workgroup: bin (rowA, pi) with its mirror in rowB (column
    // N2-1-pi), or within the row for a self-mirrored row; row 0's pair 0 is
    // DC with Nyquist (accumulator index M).  Pairs pi = tid + m*nth, m < MP,
    // are the same for every FFT of the slice (npair <= N2 <= MP*nth: a row FFT
    // gives a thread at most MAXPTS points; the launcher checks it).  The chunk
    // sums live in registers, so LDS holds only the rows.
    const int npair = !self ? N2 : (rowA == 0 ? N2 / 2 + 1 : (N2 + 1) / 2);
    constexpr int MP = MAXPTS / NSRC / 2;
    float acc1[NSRC][MP], acc2[NSRC][MP];
    float2 xacc1[MP], xacc2[MP];   // cross-power sums (NSRC == 2; dead otherwise)
#pragma unroll
    for (int m = 0; m < MP; m++) {
#pragma unroll
        for (int s = 0; s < NSRC; s++) acc1[s][m] = acc2[s][m] = 0.f;
        if (NSRC == 2) xacc1[m] = xacc2[m] = make_float2(0.f, 0.f);
    }
    // LDS column of the mirror of (rowA, pi) and whether the pair is one bin
    auto mirror_col = [&](int pi) { return (self && rowA == 0) ? (N2 - pi) % N2 : N2 - 1 - pi; };
    auto mirror_row = [&]() { return self ? 0 : 1; };   // row slot of the mirror (0 = rowA's slot)
    auto single = [&](int pi) {
        return self && (rowA == 0 ? (pi != 0 && 2 * pi == N2) : (2 * pi == N2 - 1));
    };

    const int nfj = a.jobs[grp * NSRC].nf;   // FFTs of this job (pair) in the chunk
    const int per = (nfj + a.S - 1) / a.S;   // FFTs per slice (uniform trip count)
    const int fb = sl_i * per;
    constexpr int U = 16 / NSL;
    // Split twiddles W_N^{rowA + N1 pi} of the thread's pairs pi = tid + m*nth,
    // the same for every FFT: two values loaded once per slice, W for m = 0
    // and the step W_N^{N1 nth} (rB holds W_N^{N1 k}); pair m takes the step
    // m times (<= 7 products, a few ulp).  The bin phase then issues no
    // global load, which would also wait for the rows being prefetched.
    // (nth >= N2 leaves a single pair per thread and the step unused.)
    const float2 wpair0 = cmul(a.rA[rowA], a.rB[min(tid, N2 - 1)]);
    const float2 wstep = a.rB[min(nth, N2 - 1)];

    auto row_src = [&](int sl, int f) {
        return Yp[sl >> 1] + (long long)f * a.M + (long long)((sl & 1) ? rowB : rowA) * N2;
    };
    // 16-point plans (rows fit one batch of U points per thread; the launcher
    // checks it) prefetch: FFT f+1's row loads are issued after FFT f's last
    // stage, into registers the FFT no longer needs, and land while FFT f's
    // bins are formed; they go to LDS after the bins.  Every path after the
    // FFT defines t, so t is never live across the FFT's registers.  Other
    // plans load each FFT's rows in batches at the top of its iteration.
    constexpr bool PREF = MAXPTS == 16;
    float2 t[NSL * U];   // prefetched row points t[sl*U + u], or first-stage inputs
    auto load_rows = [&](int f) __attribute__((always_inline)) {
#pragma unroll
        for (int sl = 0; sl < NSL; sl++) {
            const float2* __restrict__ Ys = row_src(sl, f);
#pragma unroll
            for (int u = 0; u < U; u++) t[sl * U + u] = Ys[batch_idx<U>(tid, u, nth, N2)];
        }
    };
    auto store_rows = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int sl = 0; sl < NSL; sl++)
#pragma unroll
            for (int u = 0; u < U; u++)
                if (tid + u * nth < N2) lds[sl * RS + tid + u * nth] = t[sl * U + u];
    };
    // A plain first radix F1 (16-point plans: 3.2M rows 5*16*16) runs from
    // registers: the prefetch loads each thread's first-stage butterfly inputs
    // q + r*L/F1 of its own row (q = q0 + i*T, i < 16/F1, in the 16 registers
    // of t) and the butterflies write the stage's output to LDS, saving a
    // store, a barrier and a read of every point.  A compile-time choice: a
    // kernel holding both paths is register-allocated for their union.
    constexpr bool fuse1 = PREF && F1 != 0;
    static_assert(NSL * U == 16, "first-stage registers");
    auto with_r1 = [&](auto fn) __attribute__((always_inline)) {
        if constexpr (fuse1) fn(std::integral_constant<int, F1>());
    };
    auto load_bfly = [&](int f) __attribute__((always_inline)) {
        with_r1([&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            const int nb = N2 / R;
            const float2* __restrict__ Ys = row_src(slot, f);
            // opaque: the index math of every radix case would otherwise be
            // hoisted out of the FFT loop and held live across it
            int qb = q0;
            asm volatile("" : "+v"(qb));
#pragma unroll
            for (int i = 0; i < 16 / R; i++) {
                const int q = min(qb + i * T, nb - 1);   // clamped: values unused past nb
#pragma unroll
                for (int r = 0; r < R; r++) t[i * R + r] = Ys[q + r * nb];
            }
        });
    };
    auto first_stage = [&]() __attribute__((always_inline)) {
        with_r1([&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            const int nb = N2 / R;
            float2* __restrict__ row = lds + slot * RS;
            int qb = q0;
            asm volatile("" : "+v"(qb));   // as in load_bfly
#pragma unroll
            for (int i = 0; i < 16 / R; i++) {
                const int q = qb + i * T;
                if (q < nb) {
                    float2 v[R];
#pragma unroll
                    for (int r = 0; r < R; r++) v[r] = t[i * R + r];
                    float2* __restrict__ dst = row + q * R;   // first stage: k = 0, o = q R
                    DftStore<R>::template run<false>(v, nullptr, [&](int r, float2 x) { dst[r] = x; });
                }
            }
        });
    };
    if (PREF && sact && fb < nfj) {
        if (fuse1) load_bfly(fb);
        else load_rows(fb);
    }
    for (int it = 0; it < per; it++) {
        const int f = fb + it;
        const bool act = sact && f < nfj;
        if (PREF) {
            if (act) {
                if (fuse1) first_stage();
                else store_rows();
            }
        } else if (act) {
            const float2* __restrict__ Ys[NSL];
#pragma unroll
            for (int sl = 0; sl < NSL; sl++) Ys[sl] = row_src(sl, f);
            // all rows' loads of a batch in flight together
            for (int base = tid; base < N2; base += U * nth) {
                float2 tb[NSL][U];
#pragma unroll
                for (int sl = 0; sl < NSL; sl++)
#pragma unroll
                    for (int u = 0; u < U; u++) tb[sl][u] = Ys[sl][batch_idx<U>(base, u, nth, N2)];
#pragma unroll
                for (int sl = 0; sl < NSL; sl++)
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (base + u * nth < N2) lds[sl * RS + base + u * nth] = tb[sl][u];
            }
        }
        __syncthreads();
        // padded rows (plan_geometry pads only plans that take a register first stage)
        fft_lds<MAXPTS, CX, (F1 != 0)>(lds_all, SP, a.stw, pipe * pipe_lds + slot * RS, 1, q0, T, SP.nst,
                                       fuse1 ? 1 : 0);
        if (PREF) {
            if (sact && f + 1 < nfj && it + 1 < per) {
                if (fuse1) load_bfly(f + 1);
                else load_rows(f + 1);
            } else {
#pragma unroll
                for (int u = 0; u < NSL * U; u++) t[u] = make_float2(0.f, 0.f);
            }
        }

        // each thread keeps the same bin pairs for every FFT of the slice
        float2 w = wpair0;
        asm volatile("" : "+v"(w.x), "+v"(w.y));   // not hoisted: no 8 live pair twiddles
#pragma unroll
        for (int m = 0; m < MP; m++) {
            // opaque per pair: keeps the unrolled pairs' loads and index math
            // from being hoisted and held live together
            int pi = tid + m * nth;
            asm volatile("" : "+v"(pi));
            if (m > 0) w = cmul(w, wstep);
            if (!act || pi >= npair) continue;
            const int mc = mirror_col(pi), ms = mirror_row();
            float2 X1[NSRC], X2[NSRC];
#pragma unroll
            for (int s = 0; s < NSRC; s++) {
                const float2 Z = lds[(2 * s) * RS + pi];
                const float2 Zm = lds[(2 * s + ms) * RS + mc];
                const float2 E = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
                const float2 O = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
                const float2 wO = cmul(w, O);
                X1[s] = make_float2(E.x + wO.x, E.y + wO.y);
                X2[s] = make_float2(E.x - wO.x, wO.y - E.y);   // conj(E - wO)
                acc1[s][m] += X1[s].x * X1[s].x + X1[s].y * X1[s].y;
                acc2[s][m] += X2[s].x * X2[s].x + X2[s].y * X2[s].y;
            }
            if (NSRC == 2) {
                // X0 * conj(X1) for both bins of the pair
                float2 c = xacc1[m];
                c.x += X1[0].x * X1[NSRC - 1].x + X1[0].y * X1[NSRC - 1].y;
                c.y += X1[0].y * X1[NSRC - 1].x - X1[0].x * X1[NSRC - 1].y;
                xacc1[m] = c;
                c = xacc2[m];
                c.x += X2[0].x * X2[NSRC - 1].x + X2[0].y * X2[NSRC - 1].y;
                c.y += X2[0].y * X2[NSRC - 1].x - X2[0].x * X2[NSRC - 1].y;
                xacc2[m] = c;
            }
        }
        __syncthreads();
    }

    // accumulator (sum, compensation) updates, both bins of each pair; the
    // accumulator reads of UA pairs are issued together
    float* __restrict__ xbase = NSRC == 2 ? a.xp[grp].acc + sl_i * a.xaccstride : nullptr;
    auto acc_index = [&](int pi, bool second) -> long long {
        if (!second) return (long long)rowA * N2 + pi;
        if (self && rowA == 0 && pi == 0) return a.M;               // Nyquist
        return (long long)(self ? rowA : rowB) * N2 + mirror_col(pi);
    };
    auto kahan2 = [&](float2 sc, float v) {
        if (a.kahan) {
            const float y = v - sc.y, t = sc.x + y;
            sc.y = (t - sc.x) - y;
            sc.x = t;
        } else {
            sc.x += v;
        }
        return sc;
    };
    auto kahan4 = [&](float4 x, float2 v) {
        if (a.kahan) {
            const float y0 = v.x - x.z, y1 = v.y - x.w;
            const float t0 = x.x + y0, t1 = x.y + y1;
            x.z = (t0 - x.x) - y0;
            x.w = (t1 - x.y) - y1;
            x.x = t0;
            x.y = t1;
        } else {
            x.x += v.x;
            x.y += v.y;
        }
        return x;
    };
    constexpr int UA = MP < 4 ? MP : 4;
#pragma unroll
    for (int m0 = 0; m0 < MP; m0 += UA) {
        if (!sact || tid + m0 * nth >= npair) break;
        float2 cur1[NSRC][UA], cur2[NSRC][UA];
        float4 xc1[UA], xc2[UA];
        long long o1[UA], o2[UA];
#pragma unroll
        for (int u = 0; u < UA; u++) {
            const int pi = min(tid + (m0 + u) * nth, npair - 1);
            o1[u] = acc_index(pi, false);
            o2[u] = acc_index(pi, true);
#pragma unroll
            for (int s = 0; s < NSRC; s++) {
                cur1[s][u] = *(const float2*)(accp[s] + 2 * o1[u]);
                cur2[s][u] = *(const float2*)(accp[s] + 2 * o2[u]);
            }
            if (NSRC == 2) {
                xc1[u] = *(const float4*)(xbase + 4 * o1[u]);
                xc2[u] = *(const float4*)(xbase + 4 * o2[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < UA; u++) {
            const int pi = tid + (m0 + u) * nth;
            if (pi >= npair) continue;
            const bool two = !single(pi);
#pragma unroll
            for (int s = 0; s < NSRC; s++) {
                *(float2*)(accp[s] + 2 * o1[u]) = kahan2(cur1[s][u], acc1[s][m0 + u]);
                if (two) *(float2*)(accp[s] + 2 * o2[u]) = kahan2(cur2[s][u], acc2[s][m0 + u]);
            }
            if (NSRC == 2) {
                *(float4*)(xbase + 4 * o1[u]) = kahan4(xc1[u], xacc1[m0 + u]);
                if (two) *(float4*)(xbase + 4 * o2[u]) = kahan4(xc2[u], xacc2[m0 + u]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// single pass (short FFTs): decode + window + the whole M-point FFT in LDS +
// r2c split + detection, each pipeline looping over the FFTs of one
// accumulator slice with the slice's sums in registers, so no intermediate
// leaves the workgroup.  The plan is the four-step layout with N1 = M rows
// and N2 = 1 column: the decode/window fill is pass 1's (column 0) and
// accumulator entry k is bin k (N1 = M, so the entry of bin k is k).
// ---------------------------------------------------------------------------

struct SPArgs {
    long long M;
    int T;                 // threads per FFT pipeline
    int fpw;               // pipelines per workgroup: fpw / NSRC groups of NSRC pipelines
    int S;                 // slices (workgroups) per job: slice s sums FFTs [s*per, (s+1)*per) of the chunk
    int kahan;
    long long accstride;   // floats between accumulator slices (2*(M+1))
    long long xaccstride;  // floats between cross-power slices (4*(M+1))
    SubPlan sp;
    const float2* stw;
    const float2* rA;      // W_N^k, k < M: r2c split twiddles
    WindowGen win;
    int zfast;             // jobs share raw bytes: job index fastest in the workgroup order
    JobDev jobs[kMaxJobs];
    XpolDev xp[kMaxJobs / 2];
};

__device__ __forceinline__ float2 kahan_add2(float2 sc, float v, int kahan)
{
    if (kahan) {
        const float y = v - sc.y, t = sc.x + y;
        sc.y = (t - sc.x) - y;
        sc.x = t;
    } else {
        sc.x += v;
    }
    return sc;
}

__device__ __forceinline__ float4 kahan_add4(float4 x, float2 v, int kahan)
{
    if (kahan) {
        const float y0 = v.x - x.z, y1 = v.y - x.w;
        const float t0 = x.x + y0, t1 = x.y + y1;
        x.z = (t0 - x.x) - y0;
        x.w = (t1 - x.y) - y1;
        x.x = t0;
        x.y = t1;
    } else {
        x.x += v.x;
        x.y += v.y;
    }
    return x;
}

// Grid (S, 1, job groups); block fpw * T threads.  A workgroup is one
// accumulator slice; its G = fpw / NSRC groups of NSRC pipelines take the
// slice's FFTs in turn (group g: FFTs g, g + G, ...), pipeline `pipe` holding
// source pipe % NSRC of group pipe / NSRC.  Bin pairs (k, M - k) are split
// over a group's NSRC * T threads, each thread keeping the same pairs (and
// both sources of a cross-power pair) for every FFT; at the end the groups'
// sums meet in LDS, are added in group order (a fixed order: reproducible),
// and group 0 updates the slice.
// One source at 16 points per thread is held to 128 VGPRs (4 waves per SIMD;
// a few dwords spill): 7% faster at 4096 points than 2 waves without
// spills.  At 32 points per thread the same bound spills ~250 bytes per lane
// and costs 25-35%, so those take the compiler's allocation.
template <int MAXPTS, int NSRC, bool CX>
__global__ void __launch_bounds__(512)
    __attribute__((amdgpu_waves_per_eu(NSRC == 1 && MAXPTS == 16 ? 4 : 1))) k_short(SPArgs a)
{
    extern __shared__ float2 lds[];
    const int M = (int)a.M, T = a.T;
    const int pipe = (int)threadIdx.x / T;
    const int tid = (int)threadIdx.x - pipe * T;
    const int G = a.fpw / NSRC;
    const int gi = pipe / NSRC, src = pipe - gi * NSRC;
    unsigned bx, by, bz;
    xcd_remap(bx, by, bz, a.zfast);   // grid (S, 1, job groups)
    const int sl = (int)bx;
    const int grp = (int)bz;
    JobDev J = a.jobs[grp * NSRC];
    if (NSRC == 2 && src == 1) J = a.jobs[grp * NSRC + 1];
    const int nfj = a.jobs[grp * NSRC].nf;          // FFTs of this job (pair) in the chunk
    const int per = (nfj + a.S - 1) / a.S;          // FFTs per slice
    const int fb = sl * per;
    const int trips = (per + G - 1) / G;            // uniform: barriers inside
    float2* __restrict__ lp = lds + pipe * M;
    const float2* __restrict__ lg = lds + gi * NSRC * M;   // the group's NSRC FFTs
    const int gt = src * T + tid, gth = NSRC * T;
    const WinCoef wc = win_coef(a.win, 0);
    constexpr int MP = MAXPTS / (2 * NSRC) + 1;     // pairs per thread: (M/2 + 1) / (NSRC T) <= MP
    const int npair = M / 2 + 1;
    float acc1[NSRC][MP], acc2[NSRC][MP];
    float2 xacc1[MP], xacc2[MP];
#pragma unroll
    for (int m = 0; m < MP; m++) {
#pragma unroll
        for (int s = 0; s < NSRC; s++) acc1[s][m] = acc2[s][m] = 0.f;
        xacc1[m] = xacc2[m] = make_float2(0.f, 0.f);
    }
    for (int it = 0; it < trips; it++) {
        const int k = it * G + gi;                  // FFT of the slice
        const int f = fb + k;
        const bool act = k < per && f < nfj;
        if (act) {
            const uint8_t* __restrict__ raw = J.up.framed ? J.raw : job_fft_ptr(J, f);
            const long long pay = J.pay0 + (long long)f * J.fft_stride;   // framed only
            const int2* __restrict__ brow = is_m5a(J.up) ? J.up.vblank + (size_t)f * J.up.vnseg : nullptr;
            fill_points<false>(lp, raw, J.up, pay, brow, a.win, wc, tid, T, 1, M, 1, 0, tid);
        }
        __syncthreads();
        fft_lds<MAXPTS, CX>(lds, a.sp, a.stw, pipe * M, 1, tid, T, a.sp.nst);
        if (act) {
            // X[k] = E + W_N^k O, X[M-k] = conj(E - W_N^k O) with
            // E = (Z[k] + conj Z[M-k])/2, O = (Z[k] - conj Z[M-k])/(2i);
            // pair 0 gives DC and Nyquist (accumulator entry M)
#pragma unroll
            for (int m = 0; m < MP; m++) {
                int pi = gt + m * gth;
                asm volatile("" : "+v"(pi));   // per pair: loads and index math not hoisted together
                if (pi >= npair) continue;
                const int mc = pi == 0 ? 0 : M - pi;
                const float2 w = a.rA[pi];
                float2 X1[NSRC], X2[NSRC];
#pragma unroll
                for (int s = 0; s < NSRC; s++) {
                    const float2 Z = lg[s * M + pi];
                    const float2 Zm = lg[s * M + mc];
                    const float2 E = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
                    const float2 O = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
                    const float2 wO = cmul(w, O);
                    X1[s] = make_float2(E.x + wO.x, E.y + wO.y);
                    X2[s] = make_float2(E.x - wO.x, wO.y - E.y);   // conj(E - wO)
                    acc1[s][m] += X1[s].x * X1[s].x + X1[s].y * X1[s].y;
                    acc2[s][m] += X2[s].x * X2[s].x + X2[s].y * X2[s].y;
                }
                if (NSRC == 2) {
                    // X0 * conj(X1) for both bins of the pair
                    float2 c = xacc1[m];
                    c.x += X1[0].x * X1[NSRC - 1].x + X1[0].y * X1[NSRC - 1].y;
                    c.y += X1[0].y * X1[NSRC - 1].x - X1[0].x * X1[NSRC - 1].y;
                    xacc1[m] = c;
                    c = xacc2[m];
                    c.x += X2[0].x * X2[NSRC - 1].x + X2[0].y * X2[NSRC - 1].y;
                    c.y += X2[0].y * X2[NSRC - 1].x - X2[0].x * X2[NSRC - 1].y;
                    xacc2[m] = c;
                }
            }
        }
        __syncthreads();
    }
    if (fb >= nfj) return;   // no FFT in this slice (whole workgroup): nothing to add
    // groups 1.. hand their sums to group 0 through LDS (the FFT buffers are
    // free: the loop ended with a barrier).  Group g - 1's region holds, per
    // pair, NSRC (bin k, bin M-k) sums and with cross-power the two complex
    // sums: (2 NSRC + 4 [NSRC == 2]) npair floats, which the G >= 2 groups'
    // NSRC * M * 8 bytes each of LDS hold.
    if (G > 1) {
        float* __restrict__ red = (float*)lds;
        const int gstride = (2 * NSRC + (NSRC == 2 ? 4 : 0)) * npair;
        if (gi > 0) {
            float* __restrict__ r = red + (gi - 1) * gstride;
#pragma unroll
            for (int m = 0; m < MP; m++) {
                const int pi = gt + m * gth;
                if (pi >= npair) continue;
#pragma unroll
                for (int s = 0; s < NSRC; s++) {
                    r[(2 * s) * npair + pi] = acc1[s][m];
                    r[(2 * s + 1) * npair + pi] = acc2[s][m];
                }
                if (NSRC == 2) {
                    float2* rx = (float2*)(r + 2 * NSRC * npair);
                    rx[pi] = xacc1[m];
                    rx[npair + pi] = xacc2[m];
                }
            }
        }
        __syncthreads();
        if (gi > 0) return;
        for (int g = 1; g < G; g++) {
            const float* __restrict__ r = red + (g - 1) * gstride;
#pragma unroll
            for (int m = 0; m < MP; m++) {
                const int pi = gt + m * gth;
                if (pi >= npair) continue;
#pragma unroll
                for (int s = 0; s < NSRC; s++) {
                    acc1[s][m] += r[(2 * s) * npair + pi];
                    acc2[s][m] += r[(2 * s + 1) * npair + pi];
                }
                if (NSRC == 2) {
                    const float2* rx = (const float2*)(r + 2 * NSRC * npair);
                    const float2 u = rx[pi], v = rx[npair + pi];
                    xacc1[m].x += u.x; xacc1[m].y += u.y;
                    xacc2[m].x += v.x; xacc2[m].y += v.y;
                }
            }
        }
    }
    // accumulator (sum, compensation) updates, both bins of each pair
#pragma unroll
    for (int m = 0; m < MP; m++) {
        const int pi = gt + m * gth;
        if (pi >= npair) continue;
        const long long o1 = pi, o2 = pi == 0 ? M : M - pi;
        const bool two = !(pi != 0 && 2 * pi == M);   // M/2 is its own mirror
#pragma unroll
        for (int s = 0; s < NSRC; s++) {
            float* __restrict__ accp = a.jobs[grp * NSRC + s].acc + (long long)sl * a.accstride;
            float2* e1 = (float2*)(accp + 2 * o1);
            float2* e2 = (float2*)(accp + 2 * o2);
            const float2 c1 = *e1;
            const float2 c2 = two ? *e2 : make_float2(0.f, 0.f);
            *e1 = kahan_add2(c1, acc1[s][m], a.kahan);
            if (two) *e2 = kahan_add2(c2, acc2[s][m], a.kahan);
        }
        if (NSRC == 2) {
            float* __restrict__ xb = a.xp[grp].acc + (long long)sl * a.xaccstride;
            float4* e1 = (float4*)(xb + 4 * o1);
            float4* e2 = (float4*)(xb + 4 * o2);
            const float4 c1 = *e1;
            const float4 c2 = two ? *e2 : make_float4(0.f, 0.f, 0.f, 0.f);
            *e1 = kahan_add4(c1, xacc1[m], a.kahan);
            if (two) *e2 = kahan_add4(c2, xacc2[m], a.kahan);
        }
    }
}

// ---------------------------------------------------------------------------
// finalize: permuted (sum, comp) accumulators -> natural-order float spectrum,
// scaled in FP64, accumulators cleared for the next spectrum.
// ---------------------------------------------------------------------------

// All jobs of a spectrum in one launch, through 32 x 32 tiles: a tile's
// accumulator entries (k1, k2) are read along k2 (contiguous, 256-byte row
// segments), summed over slices in slice order in FP64, scaled, and written
// along k1 to natural bin k = k2 N1 + k1 through LDS, so reads and writes are
// both coalesced; with `zero` the entries read are cleared for the next
// spectrum in the same pass.  Grid (ceil(N2/32), ceil(N1/32), jobs); the
// Nyquist entry (index M) is taken by tile (0, 0).
struct FinArgs {
    long long M;
    int N1, N2;
    double scale;
    int S;
    long long stride;   // floats between slices
    int zero;
    float* acc[kMaxJobs];
    float* out[kMaxJobs];
};

template <bool XP>
__global__ void __launch_bounds__(256) k_finalize_t(FinArgs a)
{
    constexpr int W = XP ? 4 : 2;   // floats per accumulator entry
    __shared__ float2 tile[32][33];
    float* __restrict__ acc = a.acc[blockIdx.z];
    float* __restrict__ out = a.out[blockIdx.z];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int k2b = blockIdx.x * 32, k1b = blockIdx.y * 32;
    auto entry = [&](long long idx) -> float2 {
        double re = 0.0, im = 0.0;
        for (int s = 0; s < a.S; s++) {
            float* e = acc + s * a.stride + W * idx;
            if (XP) {
                const float4 v = *(const float4*)e;
                re += (double)v.x - (double)v.z;
                im += (double)v.y - (double)v.w;
                if (a.zero) *(float4*)e = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                const float2 v = *(const float2*)e;
                re += (double)v.x - (double)v.y;
                if (a.zero) *(float2*)e = make_float2(0.f, 0.f);
            }
        }
        return make_float2((float)(re * a.scale), (float)(im * a.scale));
    };
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k1 = k1b + ty + 8 * r, k2 = k2b + tx;
        if (k1 < a.N1 && k2 < a.N2) tile[ty + 8 * r][tx] = entry((long long)k1 * a.N2 + k2);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int k2 = k2b + ty + 8 * r, k1 = k1b + tx;
        if (k1 < a.N1 && k2 < a.N2) {
            const long long k = (long long)k2 * a.N1 + k1;
            const float2 v = tile[tx][ty + 8 * r];
            if (XP) {
                out[2 * k] = v.x;
                out[2 * k + 1] = k == 0 ? 0.0f : v.y;   // DC cross product is real (TaskCoreIPP.cpp:482-485)
            } else {
                out[k] = v.x;
            }
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        const float2 v = entry(a.M);   // Nyquist
        if (XP) {
            out[2 * a.M] = v.x;
            out[2 * a.M + 1] = 0.0f;
        } else {
            out[a.M] = v.x;
        }
    }
}

// Many slices (short FFTs: up to 1024): 64 accumulator entries per block of
// 1024 threads, 16 slice groups; group h sums slices h, h+16, ... in order
// (FP64), clearing them with `zero`, and the 16 partials are added in group
// order: a fixed order, so results are reproducible.  Grid (ceil((M+1)/64),
// 1, jobs).
template <bool XP>
__global__ void __launch_bounds__(1024) k_finalize_s(FinArgs a)
{
    constexpr int W = XP ? 4 : 2;
    constexpr int NG = 16;
    __shared__ double2 part[NG][64];
    float* __restrict__ acc = a.acc[blockIdx.z];
    float* __restrict__ out = a.out[blockIdx.z];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const long long idx = (long long)blockIdx.x * 64 + lane;
    double re = 0.0, im = 0.0;
    if (idx <= a.M) {
#pragma unroll 4
        for (int s = h; s < a.S; s += NG) {
            float* e = acc + s * a.stride + W * idx;
            if (XP) {
                const float4 v = *(const float4*)e;
                re += (double)v.x - (double)v.z;
                im += (double)v.y - (double)v.w;
                if (a.zero) *(float4*)e = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
                const float2 v = *(const float2*)e;
                re += (double)v.x - (double)v.y;
                if (a.zero) *(float2*)e = make_float2(0.f, 0.f);
            }
        }
    }
    part[h][lane] = make_double2(re, im);
    __syncthreads();
    if (h == 0 && idx <= a.M) {
        for (int g = 1; g < NG; g++) { re += part[g][lane].x; im += part[g][lane].y; }
        // natural bin of accumulator entry idx = k1*N2 + k2 is k = k2*N1 + k1
        const long long k = idx == a.M ? a.M : (idx % a.N2) * a.N1 + idx / a.N2;
        if (XP) {
            out[2 * k] = (float)(re * a.scale);
            out[2 * k + 1] = (k == 0 || k == a.M) ? 0.0f : (float)(im * a.scale);
        } else {
            out[k] = (float)(re * a.scale);
        }
    }
}

// dst += src, src = 0 (combineResults on device arrays)
__global__ void k_add_clear(float* __restrict__ dst, float* __restrict__ src, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        dst[i] += src[i];
        src[i] = 0.0f;
    }
}

__global__ void k_zero(uint32_t* __restrict__ p, size_t nwords)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ---------------------------------------------------------------------------
// phase-cal: fold x[j] * rot[j mod L] into bin j mod T, FP64 accumulation
// (TaskCoreIPP.cpp:633-693 folds per FFT starting at rotator index 0)
// ---------------------------------------------------------------------------

struct PcArgs {
    PcalGeom g;
    long long N;
    JobDev jobs[kMaxJobs];
};

__global__ void __launch_bounds__(1024) k_pcal(PcArgs a)
{
    __shared__ double2 red[1024];
    const JobDev J = a.jobs[blockIdx.z];
    const int f = blockIdx.y, blk = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int T = a.g.T, L = a.g.L;
    double ar = 0.0, ai = 0.0;
    if ((J.pcal_mask >> f) & 1ull) {
        const Unpack up = J.up;
        const uint8_t* __restrict__ raw = up.framed ? J.raw : job_fft_ptr(J, f);
        const long long pay = J.pay0 + (long long)f * J.fft_stride;   // framed only
        const int2* __restrict__ brow = is_m5a(up) ? up.vblank + (size_t)f * up.vnseg : nullptr;
        const long long s0 = (long long)blk * a.g.chunk;
        const long long s1 = min(s0 + (long long)a.g.chunk, a.N);
        int ri = (int)((s0 + tid) % L);
        for (long long j = s0 + tid; j < s1; j += nth) {
            const float x = decode_sample(raw, up, brow, pay, j);
            const float2 r = a.g.d_rot[ri];
            ar += (double)x * (double)r.x;
            ai += (double)x * (double)r.y;
            ri += nth;
            if (ri >= L) ri -= L;
        }
    }
    red[tid] = make_double2(ar, ai);
    __syncthreads();
    if (tid < T) {
        double sr = 0.0, si = 0.0;
        for (int k = tid; k < nth; k += T) { sr += red[k].x; si += red[k].y; }
        double* part = J.pcal_part + 2 * (((long long)f * a.g.nblk_max + blk) * T + tid);
        part[0] = sr;
        part[1] = si;
    }
}

// Fast fold.  Thread owns groups of 4 consecutive rotator positions
// r = 4g + u (g = tid + i*nth, i < GPT); a block covers `ppb` rotator periods
// of one FFT.  Across periods the rotator value of a position is fixed, so the
// per-sample work is exact integer counting of the two 2-bit magnitudes:
//   sum_p x[pL + r] = A_r * 1 + B_r * 3.3359f   (A_r, B_r signed counts)
// and the rotator is applied once per position in FP64.  With 4*nth % T == 0
// every group of a thread folds into the same four bins, reduced in LDS in a
// fixed order.  STEP = bytes per time sample (multi-channel), 0 = 1-channel.
// FR: framed Mk5B input (its own instantiation, so the payload path's loop
// carries no frame arithmetic)
template <int STEP, int GPT, bool FR = false>
__global__ void __launch_bounds__(1024) k_pcal_fast(PcArgs a)
{
    extern __shared__ double2 red4[];   // [nth][4]
    // (job-fastest ordering as in pass 1 cut this kernel's HBM reads 4x but
    // made it slower: it is not bound by them)
    const JobDev J = a.jobs[blockIdx.z];
    const int f = blockIdx.y, blk = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int T = a.g.T, L = a.g.L;
    const int ngroups = L >> 2;
    double accr[4] = {0.0, 0.0, 0.0, 0.0}, acci[4] = {0.0, 0.0, 0.0, 0.0};
    if ((J.pcal_mask >> f) & 1ull) {
        const uint8_t* __restrict__ raw = job_fft_ptr(J, f);
        // framed Mk5B: byte P of the FFT's payload is at mk5b_frame_addr(pay + P)
        // of the first frame; a word (STEP 1) or 8-byte group (STEP 2) at an
        // aligned payload offset never straddles a frame (10000 = 8 * 1250), and
        // offsets stay below 2^32 (checked by the launcher)
        const unsigned pay = FR ? (unsigned)(J.pay0 + (long long)f * J.fft_stride) : 0u;
        auto at = [&](auto P) -> const uint8_t* {
            if constexpr (!FR) {
                return raw + P;
            } else {
                const unsigned Q = pay + (unsigned)P;   // payload index
                return J.raw + (Q + 16u * (Q / 10000u + 1u));
            }
        };
        const int nper = (int)(a.N / L);
        const int p0 = blk * a.g.ppb;
        const int p1 = min(p0 + a.g.ppb, nper);
        const int cbyte = J.up.cbyte, shift = J.up.shift;
#pragma unroll
        for (int gi = 0; gi < GPT; gi++) {
            const int g = tid + gi * nth;
            if (g >= ngroups) break;
            int A[4] = {0, 0, 0, 0}, Bc[4] = {0, 0, 0, 0};
            if constexpr (STEP == 1) {
                // One byte per time sample: the 4 positions are the 4 bytes of
                // one word, counted in parallel as byte lanes (SWAR) -- big
                // (magnitude bit), big & sign, ~big & sign -- with
                //   A = (#periods - #big) - 2 #(~big & sign),  B = #big - 2 #(big & sign)
                // flushed to integers every 255 periods (byte-lane capacity).
                constexpr unsigned M = 0x01010101u;
                for (int pc = p0; pc < p1; pc += 255) {
                    const int pe = min(pc + 255, p1);
                    unsigned cb = 0, cn1 = 0, cn0 = 0;
#pragma unroll 8
                    for (int p = pc; p < pe; p++) {
                        const unsigned w = *(const uint32_t*)at((long long)p * L + 4 * g + cbyte) >> shift;
                        const unsigned x = w >> 1;
                        cb += x & M;
                        cn1 += w & x & M;
                        cn0 += w & ~x & M;
                    }
                    const int np = pe - pc;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int nb = (int)((cb >> (8 * u)) & 0xFFu);
                        A[u] += (np - nb) - 2 * (int)((cn0 >> (8 * u)) & 0xFFu);
                        Bc[u] += nb - 2 * (int)((cn1 >> (8 * u)) & 0xFFu);
                    }
                }
            } else
            for (int p = p0; p < p1; p++) {
                const long long j0 = (long long)p * L + 4 * g;
                unsigned fld[4];
                if (STEP == 0) {
                    const unsigned b = raw[j0 >> 2];   // 4 samples, oldest in the MSB
#pragma unroll
                    for (int u = 0; u < 4; u++) fld[u] = (b >> (6 - 2 * u)) & 3u;
                } else {
                    const uint32_t* w = (const uint32_t*)at(j0 * STEP);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int bi = u * STEP + cbyte;
                        fld[u] = (w[(u * STEP) >> 2] >> (8 * (bi & 3) + shift)) & 3u;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int s = 1 - 2 * (int)(fld[u] & 1u);
                    const int big = (int)(fld[u] >> 1);
                    A[u] += s * (1 - big);
                    Bc[u] += s * big;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float2 r = a.g.d_rot[4 * g + u];
                const double v = (double)A[u] + (double)3.3359f * (double)Bc[u];
                accr[u] += v * (double)r.x;
                acci[u] += v * (double)r.y;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) red4[4 * tid + u] = make_double2(accr[u], acci[u]);
    __syncthreads();
    // bin t collects threads h with (4h mod T) == t - t%4, column t%4; the
    // thread count is a multiple of T/4 but may be below T (e.g. L = T = 256
    // runs 64 threads), so each thread writes bins t = tid, tid + nth, ...
    const int q = T >> 2;
    for (int t = tid; t < T; t += nth) {
        const int u = t & 3;
        double sr = 0.0, si = 0.0;
        for (int h = t >> 2; h < nth; h += q) { sr += red4[4 * h + u].x; si += red4[4 * h + u].y; }
        double* part = J.pcal_part + 2 * (((long long)f * a.g.nblk_max + blk) * T + t);
        part[0] = sr;
        part[1] = si;
    }
}

// Grouped fold: the NCH channels of one multi-channel source in one thread.
// Jobs z*NCH .. z*NCH+NCH-1 read the same bytes with the same layout and
// phase-cal mask and differ only in their bit shift (the launcher checks), so
// each raw word is loaded once for all of them instead of once per channel.
// One word holds the channel byte of 4 time samples (with 2 bytes per time
// sample, STEP 2, the 4 bytes are first gathered from an 8-byte group by
// v_perm_b32), and all channels of that byte are counted together in SIMD-
// within-a-register lanes.  One group of 4 rotator positions per thread, so
// the integer counts stay in registers until the FP64 rotator products at the
// end.
template <int STEP, bool FR, int NCH>
__global__ void __launch_bounds__(1024) k_pcal_group(PcArgs a)
{
    static_assert(STEP == 1 || STEP == 2, "multi-channel layouts only");
    extern __shared__ double2 red4[];   // [nth][4]
    const int z = blockIdx.z;
    const JobDev J = a.jobs[z * NCH];
    const int f = blockIdx.y, blk = blockIdx.x;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int T = a.g.T, L = a.g.L;
    const int pchunk = blk / a.g.npos_grp;
    const int g = (blk - pchunk * a.g.npos_grp) * nth + tid;
    const bool on = ((J.pcal_mask >> f) & 1ull) && g < (L >> 2);
    int A[NCH][4], Bc[NCH][4];
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int u = 0; u < 4; u++) A[c][u] = Bc[c][u] = 0;
    if (on) {
        const uint8_t* __restrict__ raw = job_fft_ptr(J, f);
        // framed Mk5B as in k_pcal_fast: an aligned word (STEP 1) or 8-byte
        // group (STEP 2) never straddles a 10000-byte payload
        const unsigned pay = FR ? (unsigned)(J.pay0 + (long long)f * J.fft_stride) : 0u;
        auto at = [&](long long P) -> const uint8_t* {
            if constexpr (!FR) {
                return raw + P;
            } else {
                const unsigned Q = pay + (unsigned)P;
                return J.raw + (Q + 16u * (Q / 10000u + 1u));
            }
        };
        const int nper = (int)(a.N / L);
        const int p0 = pchunk * a.g.ppb_grp;
        const int p1 = min(p0 + a.g.ppb_grp, nper);
        const int cbyte = J.up.cbyte;
        int sh[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) sh[c] = a.jobs[z * NCH + c].up.shift;
        // STEP 2: bytes cbyte and cbyte + 2 of each dword of the 8-byte group
        const unsigned sel = (unsigned)cbyte | ((unsigned)(cbyte + 2) << 8) | ((unsigned)(cbyte + 4) << 16) |
                             ((unsigned)(cbyte + 6) << 24);
        // The channel byte of each of the 4 time samples of a word holds up to
        // 4 two-bit fields at shifts 0, 2, 4, 6 (bit 0 sign, bit 1 magnitude).
        // Magnitude, sign and magnitude-and-sign bits of all of them are
        // counted at once: shifts 0 and 4 (E) and 2 and 6 (O) in separate
        // registers as nibble lanes (8 per register: 4 samples x 2 shifts),
        // moved into byte lanes every 15 periods (nibble capacity) and read
        // out per channel every 255 periods (byte capacity):
        //   A = (np - #big) - 2 (#neg - #big&neg),  B = #big - 2 #big&neg.
        constexpr unsigned K5 = 0x55555555u, K1 = 0x11111111u, K0F = 0x0F0F0F0Fu;
        int slot[NCH];   // byte-lane accumulator of the channel's shift: sh / 2
#pragma unroll
        for (int c = 0; c < NCH; c++) slot[c] = sh[c] >> 1;
        auto pick = [](const unsigned (&v)[4], int k) { return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3]; };
        for (int pc = p0; pc < p1; pc += 255) {
            const int pe = min(pc + 255, p1);
            unsigned bm[4] = {0u, 0u, 0u, 0u}, bs[4] = {0u, 0u, 0u, 0u}, bx[4] = {0u, 0u, 0u, 0u};
            for (int pq = pc; pq < pe; pq += 15) {
                unsigned mE = 0u, mO = 0u, sE = 0u, sO = 0u, xE = 0u, xO = 0u;
                // 15 periods, all loads issued together; periods past pe load a
                // valid word (clamped) and count it as 0 (no bits set)
#pragma unroll
                for (int k = 0; k < 15; k++) {
                    const int p = min(pq + k, pe - 1);
                    unsigned wd;
                    if constexpr (STEP == 1) {
                        wd = *(const uint32_t*)at((long long)p * L + 4 * g + cbyte);
                    } else {
                        const uint2 v = *(const uint2*)at(((long long)p * L + 4 * g) * 2);
                        wd = __builtin_amdgcn_perm(v.y, v.x, sel);
                    }
                    wd = pq + k < pe ? wd : 0u;
                    const unsigned S = wd & K5, Mg = (wd >> 1) & K5, X = S & Mg;
                    mE += Mg & K1;
                    mO += (Mg >> 2) & K1;
                    sE += S & K1;
                    sO += (S >> 2) & K1;
                    xE += X & K1;
                    xO += (X >> 2) & K1;
                }
                // byte-lane slots: 0 = shift 0 (E low nibble), 1 = shift 2 (O low),
                // 2 = shift 4 (E high), 3 = shift 6 (O high)
                bm[0] += mE & K0F; bm[2] += (mE >> 4) & K0F; bm[1] += mO & K0F; bm[3] += (mO >> 4) & K0F;
                bs[0] += sE & K0F; bs[2] += (sE >> 4) & K0F; bs[1] += sO & K0F; bs[3] += (sO >> 4) & K0F;
                bx[0] += xE & K0F; bx[2] += (xE >> 4) & K0F; bx[1] += xO & K0F; bx[3] += (xO >> 4) & K0F;
            }
            const int np = pe - pc;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const unsigned m = pick(bm, slot[c]), sg = pick(bs, slot[c]), x = pick(bx, slot[c]);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int nb = (int)((m >> (8 * u)) & 0xFFu);
                    const int nn = (int)((sg >> (8 * u)) & 0xFFu);
                    const int nbn = (int)((x >> (8 * u)) & 0xFFu);
                    A[c][u] += (np - nb) - 2 * (nn - nbn);
                    Bc[c][u] += nb - 2 * nbn;
                }
            }
        }
    }
    float2 r[4];
#pragma unroll
    for (int u = 0; u < 4; u++) r[u] = on ? a.g.d_rot[4 * g + u] : make_float2(0.f, 0.f);
    const int q = T >> 2;
    // Lanes h and h' of a wave fold into the same bins when h = h' mod q.  With
    // q dividing 64 and whole waves (the launcher's shape), a butterfly over
    // the lane bits >= log2 q sums them in registers (every lane of a class
    // ends with the same, fixed-order sum), lanes 0 .. q-1 keep one entry per
    // wave in LDS, and bin t sums the waves' entries in order.
    const bool wred = (64 % q) == 0 && (nth & 63) == 0;   // uniform
    const int lane = tid & 63, wave = tid >> 6, nw = nth >> 6;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        if (c > 0) __syncthreads();   // the previous channel's reads of red4 are done
        double vr[4], vi[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double v = (double)A[c][u] + (double)3.3359f * (double)Bc[c][u];
            vr[u] = v * (double)r[u].x;
            vi[u] = v * (double)r[u].y;
        }
        double* __restrict__ part_base = a.jobs[z * NCH + c].pcal_part;
        if (wred) {
            for (int m = q; m < 64; m <<= 1) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    vr[u] += __shfl_xor(vr[u], m);
                    vi[u] += __shfl_xor(vi[u], m);
                }
            }
            if (lane < q) {
#pragma unroll
                for (int u = 0; u < 4; u++) red4[4 * (wave * q + lane) + u] = make_double2(vr[u], vi[u]);
            }
            __syncthreads();
            for (int t = tid; t < T; t += nth) {
                const int u = t & 3, k = t >> 2;
                double sr = 0.0, si = 0.0;
                for (int w = 0; w < nw; w++) { sr += red4[4 * (w * q + k) + u].x; si += red4[4 * (w * q + k) + u].y; }
                *(double2*)(part_base + 2 * (((long long)f * a.g.nblk_max + blk) * T + t)) = make_double2(sr, si);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) red4[4 * tid + u] = make_double2(vr[u], vi[u]);
            __syncthreads();
            // bins as in k_pcal_fast (thread count a multiple of T/4)
            for (int t = tid; t < T; t += nth) {
                const int u = t & 3;
                double sr = 0.0, si = 0.0;
                for (int h = t >> 2; h < nth; h += q) { sr += red4[4 * h + u].x; si += red4[4 * h + u].y; }
                *(double2*)(part_base + 2 * (((long long)f * a.g.nblk_max + blk) * T + t)) = make_double2(sr, si);
            }
        }
    }
}

struct PcRedArgs {
    int T, nblk, stride, nfft;
    int R;      // level-1 blocks per job; their sums go to FFT 0's slots stride - R .. stride - 1
    JobDev jobs[kMaxJobs];
};

// Two-level sum of a chunk's per-block partials into the job's accumulator,
// in a fixed order (deterministic).  The (FFT, block) entries of a job,
// e = f * nblk + b, are cut into R contiguous ranges.  Level 1, block (job,
// r): lane l of bin t sums entries l, l + lanes, ... of range r (all loads of
// a lane issued together), lanes added in order, result to FFT 0's spare slot
// stride - R + r (no fold writes there: stride = nblk_max includes the R spare
// slots).  Level 2: bin t adds the R range sums in order to the accumulator.
__global__ void __launch_bounds__(1024) k_pcal_reduce1(PcRedArgs a)
{
    __shared__ double2 red[1024];
    const JobDev J = a.jobs[blockIdx.x];
    const int r = blockIdx.y;
    const int T = a.T;
    const int tid = threadIdx.x, lanes = blockDim.x / T;
    const int t = tid % T, l = tid / T;
    const int nent = a.nfft * a.nblk;
    const int per = (nent + a.R - 1) / a.R;
    const int e0 = r * per, e1 = min(e0 + per, nent);
    double sr = 0.0, si = 0.0;
    constexpr int U = 8;
    for (int e = e0 + l; e < e1; e += U * lanes) {
        double2 v[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ee = e + u * lanes;
            const int f = ee / a.nblk, b = ee - f * a.nblk;
            on[u] = ee < e1 && ((J.pcal_mask >> f) & 1ull);
            const int fe = on[u] ? f : 0, be = on[u] ? b : 0;
            v[u] = *(const double2*)(J.pcal_part + 2 * (((long long)fe * a.stride + be) * T + t));
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (on[u]) {
                sr += v[u].x;
                si += v[u].y;
            }
    }
    red[tid] = make_double2(sr, si);
    __syncthreads();
    if (l == 0) {
        for (int k = 1; k < lanes; k++) { sr += red[k * T + t].x; si += red[k * T + t].y; }
        *(double2*)(J.pcal_part + 2 * ((long long)(a.stride - a.R + r) * T + t)) = make_double2(sr, si);
    }
}

__global__ void __launch_bounds__(1024) k_pcal_reduce2(PcRedArgs a)
{
    const JobDev J = a.jobs[blockIdx.x];
    const int T = a.T;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int r0 = 0; r0 < a.R; r0 += 8) {   // 8 loads in flight, added in order
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                v[u] = *(const double2*)(J.pcal_part + 2 * ((long long)(a.stride - a.R + min(r0 + u, a.R - 1)) * T + t));
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (r0 + u < a.R) {
                    sr += v[u].x;
                    si += v[u].y;
                }
        }
        J.pcal_acc[2 * t] += sr;
        J.pcal_acc[2 * t + 1] += si;
    }
}

// Folded vector out (FP32, as the reference's pcal sink carries) and tones in
// FP64: tone n = (1/T) sum_b acc[b] e^{-2 pi i (n*step) b / T} (PCal.cpp:489-515).
__device__ __forceinline__ void pcal_emit_one(double* __restrict__ acc, float* __restrict__ out,
                                              double* __restrict__ tones, int T, int ntones, int step)
{
    __shared__ double2 s[1024];
    const int t = threadIdx.x;
    if (t < T) s[t] = make_double2(acc[2 * t], acc[2 * t + 1]);
    __syncthreads();
    if (t < T && out) {
        out[2 * t] = (float)s[t].x;
        out[2 * t + 1] = (float)s[t].y;
    }
    if (t < ntones && tones) {
        const long long bin = (long long)t * step;
        double re = 0.0, im = 0.0;
        if (bin < T) {
            for (int b = 0; b < T; b++) {
                const long long e = (bin * b) % T;
                double sn, cs;
                sincospi(-2.0 * (double)e / (double)T, &sn, &cs);
                re += s[b].x * cs - s[b].y * sn;
                im += s[b].x * sn + s[b].y * cs;
            }
        }
        tones[2 * t] = re / T;
        tones[2 * t + 1] = im / T;
    }
    __syncthreads();
    if (t < T) {
        acc[2 * t] = 0.0;
        acc[2 * t + 1] = 0.0;
    }
}

__global__ void k_pcal_emit(double* __restrict__ acc, float* __restrict__ out, double* __restrict__ tones,
                            int T, int ntones, int step)
{
    pcal_emit_one(acc, out, tones, T, ntones, step);
}

// every job of a spectrum in one launch: block j emits job j
struct EmitArgs {
    int T, ntones, step;
    double* acc[kMaxJobs];
    float* out[kMaxJobs];
    double* tones[kMaxJobs];
};

__global__ void k_pcal_emit_jobs(EmitArgs a)
{
    const int j = blockIdx.x;
    pcal_emit_one(a.acc[j], a.out[j], a.tones[j], a.T, a.ntones, a.step);
}

// Decoder only: samples [0, count) of one extract_samples call (VLBA: the
// first row of up.vblank holds its blank zones).
__global__ void k_unpack(const uint8_t* __restrict__ raw, Unpack up, long long count, float* __restrict__ out)
{
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < count;
         j += (long long)gridDim.x * blockDim.x)
        out[j] = decode_sample(raw, up, up.vblank, 0, j);
}

// ---------------------------------------------------------------------------
// VLBA blank zones (mark5access blanker_mark5.c:126-200) for every 20000-byte
// payload segment of the chunk's FFTs.  One thread per segment: the first and
// last 64-bit words decide; when exactly one is the fill word the boundary is
// found by the reference's bisection over words (findfirstinvalid :48-81,
// findfirstvalid :83-116) with its 32-bit straggler adjustment, so results
// match the reference where that bisection misses the boundary too.  Words
// not wholly inside the bytes the caller supplied count as data.
// ---------------------------------------------------------------------------

struct VbArgs {
    int nseg;
    JobDev jobs[kMaxJobs];
};

__device__ __forceinline__ bool vl_fill(const uint8_t* __restrict__ p, long long off, int bytes, long long avail)
{
    if (off < 0 || off + bytes > avail) return false;
    bool f = true;
    for (int b = 0; b < bytes; b++) {
        const unsigned want = 0x44u - 0x11u * (unsigned)(b & 3);   // 44 33 22 11: 0x11223344 little-endian
        f = f && (p[off + b] == want);
    }
    return f;
}

// One thread per (chunk, payload segment, zone) entry g of the FFT's row:
// zone z of a payload covers its 64-bit words [4096 z, min(4096 z + 4096,
// payload / 8)); bounds are in payload bytes.
__global__ void k_vlba_blank(VbArgs a)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    const JobDev& J = a.jobs[blockIdx.z];
    if (g >= a.nseg || f >= J.nf || !is_m5a(J.up)) return;
    const uint8_t* __restrict__ p = job_fft_ptr(J, f);
    const long long avail = J.raw_len - (long long)(p - J.raw);
    long long seg;
    int nword, zi;
    if (J.up.kind == kUnpackMark5B) {   // payload g of the FFT, one zone
        seg = (long long)g * kMark5BPayload;
        nword = kMark5BPayload / 8;
        zi = 0;
    } else {
        const int fsh = J.up.vfsh, nz = fsh + 1;
        const int cs = g / nz;
        zi = g - cs * nz;
        seg = (long long)(cs / kVlbaSegPerChunk) * ((long long)kVlbaChunk << fsh) +
              (long long)(cs % kVlbaSegPerChunk) * ((long long)kVlbaPayload << fsh);
        nword = (kVlbaPayload << fsh) / 8;
    }
    const int b = 4096 * zi;
    const int e = min(b + 4096, nword);
    const bool start_ok = !vl_fill(p, seg + 8LL * b, 8, avail);
    const bool end_ok = !vl_fill(p, seg + 8LL * (e - 1), 8, avail);
    int vs, ve;
    if (start_ok && end_ok) {
        vs = 0;
        ve = 1 << 30;
    } else if (!start_ok && !end_ok) {
        vs = 1 << 30;
        ve = 0;
    } else {
        const int n = e - b;
        int i = n / 2, r = n - 1;
        for (int s = n / 4; s > 0; s >>= 1) {
            if (vl_fill(p, seg + 8LL * (b + i), 8, avail) == start_ok) {
                r = i;
                i -= s;
            } else {
                i += s;
            }
        }
        int byte = 8 * (b + r);
        if (start_ok) {
            if (vl_fill(p, seg + byte - 4, 4, avail)) byte -= 4;
            vs = 0;
            ve = byte;
        } else {
            if (vl_fill(p, seg + byte - 8, 4, avail)) byte += 4;
            vs = byte;
            ve = 1 << 30;
        }
    }
    J.up.vblank[(size_t)f * a.nseg + g] = make_int2(vs, ve);
}

void vlba_modulation_bits(uint32_t* words)
{
    // mark5_format_vlba.c:167-196: register of 16 ones; each step shifts in
    // b10 ^ b12 ^ b13 ^ b15; after every 8th output one unrecorded step
    for (int w = 0; w < kVlbaModWords; w++) words[w] = 0u;
    unsigned reg = 0xFFFFu;
    for (int i = 0; i < kVlbaPayload; i++) {
        const int steps = (i % 8 == 7) ? 2 : 1;
        for (int st = 0; st < steps; st++) {
            const unsigned k = ((reg >> 10) ^ (reg >> 12) ^ (reg >> 13) ^ (reg >> 15)) & 1u;
            reg = ((reg << 1) | k) & 0xFFFFu;
            if (st == 0 && k) words[i >> 5] |= 1u << (i & 31);
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static size_t p1_lds(const Plan& p) { return (size_t)p.N1 * p.B * sizeof(float2); }
// FFT rows only (the chunk sums live in registers)
static int p2_rowst(const Plan& p)
{
    const SubPlan& sp = p.sp2;
    return p.N2 + (sp.padst >= 0 ? sp.pad * sp.radix[sp.nst - 1] : 0);
}
static size_t p2_lds(const Plan& p, int nsrc)
{
    return (size_t)2 * nsrc * p2_rowst(p) * sizeof(float2);
}

hipError_t launch_pass1(const Plan& p, const JobDev* jobs, int njobs, int nfft, hipStream_t s)
{
    if (njobs < 1 || njobs > kMaxJobs || nfft < 1) return hipErrorInvalidValue;
    if (p.single) return hipSuccess;   // single-pass plans: launch_pass2 does the whole FFT
    P1Args a;
    a.N1 = p.N1; a.N2 = p.N2; a.M = p.M; a.B = p.B; a.twshift = p.twshift;
    a.sp = p.sp1; a.stw = p.d_stw1; a.twA = p.d_twA; a.twB = p.d_twB; a.win = p.win; a.mft = p.d_mft;
    a.fpw = std::max(1, std::min(p.p1_fpw, nfft));
    if (p.mfma64) {
        // stage_mfma64 assumes one pipeline of whole waves and at most
        // kMfPairs (row tile, column tile) pairs per wave
        const int W = p.p1_threads / 64, nct = ((p.N1 / 64) * p.B + 15) / 16;
        if (!p.d_mft || p.p1_maxpts != 16 || p.p1_fpw != 1 || p.N1 % 64 || p.p1_threads % 64 || W < 1 ||
            ((8 + W - 1) / W) * nct > kMfPairs)
            return hipErrorInvalidValue;
        a.fpw = 1;
    }
    a.zfast = 0;
    for (int j = 1; j < njobs; j++)
        if (jobs[j].raw == jobs[0].raw && jobs[j].fft_stride == jobs[0].fft_stride) a.zfast = 1;
    if (const char* e = getenv("SWSPEC_P1_ZFAST")) a.zfast = atoi(e) != 0;
    a.pth = p.p1_threads;
    for (int j = 0; j < njobs; j++) a.jobs[j] = jobs[j];
    dim3 grid(p.N2 / p.B, (nfft + a.fpw - 1) / a.fpw, njobs);
    const size_t lds = p1_lds(p) * a.fpw;
    // Row fill: 4 or 8 columns, one pipeline, every job a one-byte-per-time-
    // sample 2-bit layout whose row segments are aligned (payload: raw and FFT
    // stride multiples of 16; framed: payload offsets multiples of 16, a
    // segment never straddling a 10000-byte payload).  SWSPEC_P1_ROWFILL=0
    // keeps the column fill.
    a.rowfill = 0;
    a.wcol = p.d_wcol;
    static const int rowfill_env = [] {
        const char* e = getenv("SWSPEC_P1_ROWFILL");
        return e ? atoi(e) : -1;
    }();
    if (rowfill_env != 0 && !p.mfma64 && a.fpw == 1 && (p.B == 4 || p.B == 8) && p.N2 % 8 == 0 && p.d_wcol) {
        bool ok = true;
        for (int j = 0; ok && j < njobs; j++) {
            const Unpack& u = jobs[j].up;
            if (u.kind == kUnpackSingleChannel) {   // BC / 2 bytes per row segment, aligned
                ok = !u.framed && (uintptr_t)jobs[j].raw % 16 == 0 && jobs[j].fft_stride % 16 == 0;
                continue;
            }
            ok = u.kind == kUnpackMultiChannel && u.step == 1 && u.cbyte == 0 && u.shift >= 0 && u.shift <= 6 &&
                 jobs[j].fft_stride % 16 == 0;
            if (ok && u.framed) ok = jobs[j].pay0 % 16 == 0 && (uintptr_t)jobs[j].raw % 16 == 0;
            else if (ok) ok = (uintptr_t)jobs[j].raw % 16 == 0;
        }
        a.rowfill = ok ? 1 : 0;
    }
    const dim3 block(a.fpw * a.pth);
    if (p.mfma64) {
        if (p.p1_cx) hipLaunchKernelGGL((k_pass1<16, true, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_pass1<16, false, true>), grid, block, lds, s, a);
        return hipGetLastError();
    }
    const int rf = a.rowfill ? p.B : 0;
    // First column stage fused into the row fill: a radix-10 first stage of a
    // plan with a later stage, 4 columns, jobs the row fill takes (its checks
    // give 16-byte aligned rows; fill_stage1 reads 4-byte groups of them, or
    // single bytes of the single-channel layout).  SWSPEC_P1_FUSE1=0 keeps
    // the separate first stage.
    static const bool f1_env = [] {
        const char* e = getenv("SWSPEC_P1_FUSE1");
        return !(e && atoi(e) == 0);
    }();
    // (A one-column-per-thread form for 8 columns, radix 16 or 25, measured
    // 18-21% slower at 10^6 (25*25: 128 VGPRs with spills); the 8-column
    // radix-16 plans, 2^20, run two pipelines per workgroup and so the
    // column fill.  Neither is built.)
    const int r0 = p.sp1.radix[0];
    const int f1 = (f1_env && p.sp1.nst >= 2 && p.p1_cx && rf == 4 && r0 == 10) ? 10 : 0;
    // a fixed column plan with constant strides (SWSPEC_P1_FIXED=0: the generic kernel)
    static const bool fx_env = [] {
        const char* e = getenv("SWSPEC_P1_FIXED");
        return !(e && atoi(e) == 0);
    }();
    int fx = 0;
    for (int id = 1; fx_env && f1 && !fx && id <= kFixedPlans; id++) {
        const FixedPlan fp = fixed_plan(id);
        if (p.N1 == fp.N1 && p.N2 == fp.N2 && p.B == fp.B && p.p1_threads == fp.p1_threads && a.fpw == 1 &&
            p.p1_maxpts == 32 && f1 == fp.cols.radix[0] && same_subplan(p.sp1, fp.cols))
            fx = id;
    }
#define SWSPEC_P1(MP, CXV)                                                                              \
    {                                                                                                 \
        if (CXV && MP == 32 && fx == 1)                                                               \
            hipLaunchKernelGGL((k_pass1<MP, CXV, false, 4, CXV ? 10 : 0, (CXV && MP == 32) ? 1 : 0>), grid, block, lds, s, a); \
        else if (CXV && f1 == 10) hipLaunchKernelGGL((k_pass1<MP, CXV, false, 4, CXV ? 10 : 0>), grid, block, lds, s, a); \
        else if (rf == 4) hipLaunchKernelGGL((k_pass1<MP, CXV, false, 4>), grid, block, lds, s, a);    \
        else if (rf == 8) hipLaunchKernelGGL((k_pass1<MP, CXV, false, 8>), grid, block, lds, s, a);    \
        else hipLaunchKernelGGL((k_pass1<MP, CXV, false, 0>), grid, block, lds, s, a);                 \
    }
    switch (p.p1_maxpts * 2 + (p.p1_cx ? 1 : 0)) {
    case 32: SWSPEC_P1(16, false) break;
    case 33: SWSPEC_P1(16, true) break;
    case 64: SWSPEC_P1(32, false) break;
    case 65: SWSPEC_P1(32, true) break;
    default: return hipErrorInvalidValue;
    }
#undef SWSPEC_P1
    return hipGetLastError();
}

// Slices per job for a chunk of nfft FFTs: enough pipelines to fill the
// device (~16k row-pair pipelines), at most smax (allocated accumulator slices).
static int pass2_slices(const Plan& p, int ngroups, int nfft, int smax)
{
    const long long wg = (long long)p.npairs * ngroups;
    int S = (int)std::min<long long>(std::max<long long>(1, (16384 + wg - 1) / wg), 1024);
    return std::max(1, std::min(S, std::min(smax, nfft)));
}

// Slices (workgroups) per job for the single-pass kernel: about 2048
// workgroups over the device, at least 2 FFTs per pipeline where the chunk
// allows, at most smax.
static int short_slices(int ngroups, int nfft, int groups_per_wg, int smax)
{
    const int want = std::max(1, (2048 + ngroups - 1) / ngroups);
    const int S = std::min(want, std::max(1, nfft / (2 * groups_per_wg)));
    return std::max(1, std::min(S, smax));
}

static hipError_t launch_short(const Plan& p, const JobDev* jobs, int njobs, const XpolDev* xp, int nfft, int kahan,
                               int smax, hipStream_t s)
{
    const int nsrc = xp ? 2 : 1;
    SPArgs a;
    a.M = p.M; a.T = p.sp1.T; a.kahan = kahan; a.sp = p.sp1; a.stw = p.d_stw1; a.rA = p.d_rA; a.win = p.win;
    for (int j = 0; j < njobs; j++) a.jobs[j] = jobs[j];
    if (xp) for (int g = 0; g < njobs / 2; g++) a.xp[g] = xp[g];
    a.accstride = 2 * (p.M + 1);
    a.xaccstride = 4 * (p.M + 1);
    const int gthreads = nsrc * a.T;
    const int MP = p.p1_maxpts / (2 * nsrc) + 1;   // the kernel's register pairs per thread
    if (gthreads > 512 || (p.p1_maxpts != 16 && p.p1_maxpts != 32) || p.M / 2 + 1 > (long long)MP * gthreads)
        return hipErrorInvalidValue;
    const size_t pipe_bytes = (size_t)p.M * sizeof(float2);
    // groups per workgroup: <= 512 threads, <= 64 KB of LDS (one group may
    // need more), no more than the chunk's FFTs; with G >= 2 the groups' sums
    // meet in the LDS of the FFT buffers ((2 nsrc + 4 [nsrc == 2]) (M/2 + 1)
    // floats per group after the first: within nsrc * M * 8 bytes per group)
    int gpw = std::max(1, std::min(512 / gthreads, (int)((64 * 1024) / (nsrc * pipe_bytes))));
    gpw = std::max(1, std::min(gpw, nfft));
    while (gpw > 1 && (size_t)(gpw - 1) * (2 * nsrc + (nsrc == 2 ? 4 : 0)) * (size_t)(p.M / 2 + 1) * sizeof(float) >
                          (size_t)gpw * nsrc * pipe_bytes)
        gpw--;
    a.fpw = gpw * nsrc;
    a.S = short_slices(njobs / nsrc, nfft, gpw, smax);
    const size_t lds = (size_t)a.fpw * pipe_bytes;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    using K = void (*)(SPArgs);
    // [maxpts 16|32][nsrc 1|2][composite]
    static const K table[2][2][2] = {
        {{k_short<16, 1, false>, k_short<16, 1, true>}, {k_short<16, 2, false>, k_short<16, 2, true>}},
        {{k_short<32, 1, false>, k_short<32, 1, true>}, {k_short<32, 2, false>, k_short<32, 2, true>}}};
    K k = table[p.p1_maxpts == 32][nsrc - 1][p.p1_cx ? 1 : 0];
    if (lds > 64 * 1024) {
        static std::mutex mu;
        static std::set<K> raised;
        std::lock_guard<std::mutex> l(mu);
        if (!raised.count(k)) {
            const hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     160 * 1024);
            if (e != hipSuccess) return e;
            raised.insert(k);
        }
    }
    // Slices fastest in the workgroup order, jobs slowest.  Channels of one
    // stream read the same bytes, but putting a slice's channel workgroups
    // side by side on one XCD (pass 1's order) measured 5-7% slower at
    // 4096-16384 points, 4 channels (interleaved runs on one device): the
    // bytes are few and the L2 serves them either way.  SWSPEC_SHORT_ZFAST=1
    // selects that order.
    a.zfast = 0;
    static const int zf_env = [] {
        const char* e = getenv("SWSPEC_SHORT_ZFAST");
        return e ? atoi(e) : -1;
    }();
    if (zf_env >= 0) a.zfast = zf_env != 0;
    const dim3 grid(a.S, 1, njobs / nsrc);
    hipLaunchKernelGGL(k, grid, dim3(a.fpw * a.T), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_pass2(const Plan& p, const JobDev* jobs, int njobs, const XpolDev* xp, int nfft,
                        int kahan, int smax, hipStream_t s)
{
    const int nsrc = xp ? 2 : 1;
    if (njobs < 1 || njobs > kMaxJobs || (njobs % nsrc) != 0 || nfft < 1 || smax < 1) return hipErrorInvalidValue;
    if (p.single) return launch_short(p, jobs, njobs, xp, nfft, kahan, smax, s);
    P2Args a;
    a.N1 = p.N1; a.N2 = p.N2; a.M = p.M; a.nfft = nfft; a.kahan = kahan;
    a.sp = p.sp2; a.stw = p.d_stw2; a.rA = p.d_rA; a.rB = p.d_rB;
    for (int j = 0; j < njobs; j++) a.jobs[j] = jobs[j];
    if (xp) for (int g = 0; g < njobs / 2; g++) a.xp[g] = xp[g];
    a.S = pass2_slices(p, njobs / nsrc, nfft, smax);
    a.pth = nsrc * 2 * p.sp2.T;
    // register chunk sums: each thread owns <= p2_maxpts / nsrc bins
    if ((long long)2 * p.N2 > (long long)(p.p2_maxpts / nsrc) * a.pth) return hipErrorInvalidValue;
    const size_t lds1 = p2_lds(p, nsrc);
    a.pipe_lds = (int)(lds1 / sizeof(float2));
    a.rowst = p2_rowst(p);
    // pack short rows: up to 256 threads (SWSPEC_P2_WG_THREADS) and 64 KB of LDS per workgroup
    static const int wg_threads = [] {
        const char* e = getenv("SWSPEC_P2_WG_THREADS");
        const int v = e ? atoi(e) : 256;
        return (v >= 64 && v <= 512) ? v : 256;
    }();
    a.fpw = std::max(1, std::min({wg_threads / a.pth, (int)((64 * 1024) / lds1), a.S}));
    a.accstride = 2 * (p.M + 1);
    a.xaccstride = 4 * (p.M + 1);
    const dim3 grid(p.npairs, njobs / nsrc, (a.S + a.fpw - 1) / a.fpw);
    const dim3 block(a.fpw * a.pth);
    const size_t lds = lds1 * a.fpw;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    using K = void (*)(P2Args);
    // [maxpts 16|32][nsrc 1|2][composite]
    static const K table[2][2][2] = {
        {{k_pass2<16, 1, false>, k_pass2<16, 1, true>}, {k_pass2<16, 2, false>, k_pass2<16, 2, true>}},
        {{k_pass2<32, 1, false>, k_pass2<32, 1, true>}, {k_pass2<32, 2, false>, k_pass2<32, 2, true>}}};
    // 16-point plans with a plain first radix 4, 5 or 8: [radix][nsrc][composite]
    static const K table_f1[3][2][2] = {
        {{k_pass2<16, 1, false, 4>, k_pass2<16, 1, true, 4>}, {k_pass2<16, 2, false, 4>, k_pass2<16, 2, true, 4>}},
        {{k_pass2<16, 1, false, 5>, k_pass2<16, 1, true, 5>}, {k_pass2<16, 2, false, 5>, k_pass2<16, 2, true, 5>}},
        {{k_pass2<16, 1, false, 8>, k_pass2<16, 1, true, 8>}, {k_pass2<16, 2, false, 8>, k_pass2<16, 2, true, 8>}}};
    if (p.p2_maxpts != 16 && p.p2_maxpts != 32) return hipErrorInvalidValue;
    K k = table[p.p2_maxpts == 32][nsrc - 1][p.p2_cx ? 1 : 0];
    static const bool f1_env = [] {
        const char* e = getenv("SWSPEC_P2_FUSE1");
        return !(e && atoi(e) == 0);
    }();
    const int r1 = p.sp2.radix[0];
    // a padded plan needs the register-first-stage kernel (the only one built
    // with the padded layout)
    if ((f1_env || p.sp2.padst >= 0) && p.p2_maxpts == 16 && (r1 == 4 || r1 == 5 || r1 == 8))
        k = table_f1[r1 == 4 ? 0 : (r1 == 5 ? 1 : 2)][nsrc - 1][p.p2_cx ? 1 : 0];
    // a fixed row plan with constant strides (SWSPEC_P2_FIXED=0: the generic kernel)
    static const bool fx_env = [] {
        const char* e = getenv("SWSPEC_P2_FIXED");
        return !(e && atoi(e) == 0);
    }();
    if (fx_env && a.fpw == 1 && p.p2_cx && p.p2_maxpts == 16) {
        for (int id = 1; id <= kFixedPlans; id++) {
            const FixedPlan fp = fixed_plan(id);
            const int r0 = fp.rows.radix[0];
            if (k != table_f1[r0 == 4 ? 0 : (r0 == 5 ? 1 : 2)][nsrc - 1][1]) continue;
            if (p.N2 != fp.N2 || a.rowst != fp.rowst || !same_subplan(p.sp2, fp.rows)) continue;
            k = nsrc == 1 ? k_pass2<16, 1, true, 5, 1> : k_pass2<16, 2, true, 5, 1>;
            break;
        }
    }
    hipLaunchKernelGGL(k, grid, block, lds, s, a);
    return hipGetLastError();
}

hipError_t launch_add_clear(float* dst, float* src, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_add_clear, dim3(blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s)
{
    if (bytes % 4) return hipErrorInvalidValue;
    const size_t n = bytes / 4;
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, (uint32_t*)ptr, n);
    return hipGetLastError();
}

void pcal_geometry(PcalGeom& g, long long N)
{
    const int T = g.T, L = g.L;
    // generic kernel: ~32 samples per thread per block, thread count a multiple of T
    int th = (256 / T) * T;
    if (th < T) th = T;
    while (th > L) th -= T;
    g.threads = th;
    long long chunk = (long long)th * 32;
    chunk = (chunk / T) * T;
    g.chunk = (int)chunk;
    g.nblk = (int)((N + chunk - 1) / chunk);
    g.fast = 0;
    g.nblk_grp = 0;
    // fast kernel: groups of 4 positions, 4*nth % T == 0, <= 4 groups per thread
    if ((L % 4) == 0 && (T % 4) == 0 && (N % L) == 0) {
        const int ngroups = L / 4, q = T / 4;
        int nth = std::min(ngroups, 1024);
        nth -= nth % q;
        if (nth >= q && nth >= 64) {
            const int gpt = (ngroups + nth - 1) / nth;
            if (gpt <= 4) {
                const long long nper = N / L;
                g.fast = 1;
                g.threads_fast = nth;
                g.gpt = gpt;
                const int nb = (int)std::min<long long>(nper, 50);
                g.ppb = (int)((nper + nb - 1) / nb);
                g.nblk_fast = (int)((nper + g.ppb - 1) / g.ppb);
            }
        }
        // grouped fold (a source's channels per thread): blocks of <= 256
        // threads, so that their LDS (64 B per thread) fits beside the FFT
        // passes' workgroups, a multiple of T/4 threads (every block then folds
        // into the same bins)
        int npos = (ngroups + 255) / 256;
        int thg = (ngroups + npos - 1) / npos;
        thg = ((thg + q - 1) / q) * q;
        if (64 % q == 0) {
            // whole waves (the kernel's in-register wave sums): the block size
            // of 64 .. 256 threads that pads the L/4 groups least
            int best = 0;
            for (int t : {256, 192, 128, 64}) {
                const int np = (ngroups + t - 1) / t;
                if (best == 0 || np * t < npos * best) { best = t; npos = np; }
            }
            thg = best;
            npos = (ngroups + thg - 1) / thg;
        }
        if (thg <= 1024) {
            static const int nbp_env = [] {
                const char* e = getenv("SWSPEC_PCAL_GRP_NB");
                return e ? atoi(e) : 0;
            }();
            const long long nper = N / L;
            // 5 period chunks per FFT (3.2M points, L = 3200: 210 periods per
            // thread) measured best of 2 to 40 on the 4-channel Mk5B workload
            // beside the row pass: fewer, longer waves leave the column and
            // row passes more of the device (10 chunks: 1.3% slower)
            const int nbp = (int)std::min<long long>(nper, nbp_env > 0 ? nbp_env : 5);
            g.threads_grp = thg;
            g.npos_grp = npos;
            g.ppb_grp = (int)((nper + nbp - 1) / nbp);
            g.ppb_grp = std::min<long long>(((g.ppb_grp + 14) / 15) * 15, std::max<long long>(nper, 1));   // whole load batches
            g.nblk_grp = npos * (int)((nper + g.ppb_grp - 1) / g.ppb_grp);
        }
    }
    // + the level-1 reduce's spare slots (k_pcal_reduce1)
    g.nblk_max = std::max(g.nblk, std::max(g.nblk_fast, g.nblk_grp)) + kPcalRedBlocks;
}

hipError_t launch_pcal(const PcalGeom& g, long long N, const JobDev* jobs, int njobs, int nfft, hipStream_t s)
{
    if (njobs < 1 || njobs > kMaxJobs || nfft < 1 || nfft > 64) return hipErrorInvalidValue;
    PcArgs a;
    a.g = g; a.N = N;
    bool aligned = true;
    int step = -1;
    bool framed = false;
    for (int j = 0; j < njobs; j++) {
        a.jobs[j] = jobs[j];
        aligned &= ((uintptr_t)jobs[j].raw % 4 == 0) && (jobs[j].fft_stride % 4 == 0);
        const int st = jobs[j].up.kind == kUnpackSingleChannel ? 0 : jobs[j].up.step;
        // framed Mk5B (1 or 2 bytes per time sample): payload offsets 8-byte
        // aligned and below 2^32 (32-bit frame arithmetic in the kernel)
        const bool framed_ok = jobs[j].up.framed && (st == 1 || st == 2) && jobs[j].pay0 % 8 == 0 &&
                               jobs[j].fft_stride % 8 == 0 && jobs[j].raw_len < (1LL << 32) &&
                               jobs[j].pay0 + (long long)nfft * jobs[j].fft_stride + N * st < (1LL << 32);
        framed |= jobs[j].up.framed != 0;
        if (is_m5a(jobs[j].up) || (jobs[j].up.framed && !framed_ok)) step = -2;   // generic fold
        else if (step == -1) step = st;
        else if (step != st) step = -2;
    }
    int nfr = 0;
    for (int j = 0; j < njobs; j++) nfr += jobs[j].up.framed != 0;
    if (nfr != 0 && nfr != njobs) step = -2;   // framed and payload jobs mixed: generic fold
    const bool fast = g.fast && aligned && (step == 0 || step == 1 || step == 2 || step == 4);
    // channels of one multi-channel source fold together (k_pcal_group) when
    // consecutive jobs read the same bytes and differ only in their bit shift
    int grp = 1;
    static const int group_env = [] {
        const char* e = getenv("SWSPEC_PCAL_GROUP");
        return e ? atoi(e) : -1;
    }();
    if (fast && group_env != 0 && g.nblk_grp > 0 && (step == 1 || step == 2)) {
        for (int G : {4, 2}) {
            if (njobs % G) continue;
            bool ok = true;
            for (int j = 0; ok && j < njobs; j++)   // fields within the byte (byte-lane counting)
                ok = jobs[j].up.shift >= 0 && jobs[j].up.shift <= 6 && (jobs[j].up.shift & 1) == 0;
            for (int k = 0; ok && k < njobs; k += G)
                for (int c = 1; ok && c < G; c++) {
                    const JobDev& x = jobs[k];
                    const JobDev& y = jobs[k + c];
                    ok = y.raw == x.raw && y.fft_stride == x.fft_stride && y.fft0 == x.fft0 && y.nwrap == x.nwrap &&
                         y.pcal_mask == x.pcal_mask && y.pay0 == x.pay0 && y.up.kind == x.up.kind &&
                         y.up.step == x.up.step && y.up.cbyte == x.up.cbyte && y.up.framed == x.up.framed;
                }
            if (ok) { grp = G; break; }
        }
    }
    if (fast && grp > 1) {
        const dim3 grid(g.nblk_grp, nfft, njobs / grp), block(g.threads_grp);
        const size_t lds = (size_t)g.threads_grp * 4 * sizeof(double2);
#define SWSPEC_PCAL_GRP(ST, FRM)                                                                        \
    {                                                                                                 \
        if (grp == 4) hipLaunchKernelGGL((k_pcal_group<ST, FRM, 4>), grid, block, lds, s, a);           \
        else hipLaunchKernelGGL((k_pcal_group<ST, FRM, 2>), grid, block, lds, s, a);                    \
    }
        if (step == 1) {
            if (framed) SWSPEC_PCAL_GRP(1, true)
            else SWSPEC_PCAL_GRP(1, false)
        } else {
            if (framed) SWSPEC_PCAL_GRP(2, true)
            else SWSPEC_PCAL_GRP(2, false)
        }
#undef SWSPEC_PCAL_GRP
    } else if (fast) {
        // one job (channel) per blockIdx.z
        const dim3 grid(g.nblk_fast, nfft, njobs), block(g.threads_fast);
        const size_t lds = (size_t)g.threads_fast * 4 * sizeof(double2);
#define SWSPEC_PCAL_FAST(ST, FRM)                                                                       \
    switch (g.gpt) {                                                                                  \
    case 1: hipLaunchKernelGGL((k_pcal_fast<ST, 1, FRM>), grid, block, lds, s, a); break;              \
    case 2: hipLaunchKernelGGL((k_pcal_fast<ST, 2, FRM>), grid, block, lds, s, a); break;              \
    default: hipLaunchKernelGGL((k_pcal_fast<ST, 4, FRM>), grid, block, lds, s, a); break;             \
    }
        if (framed) {   // every job framed (one core's flags), step 1 or 2
            if (step == 1) SWSPEC_PCAL_FAST(1, true)
            else SWSPEC_PCAL_FAST(2, true)
        } else {
            switch (step) {
            case 0: SWSPEC_PCAL_FAST(0, false) break;
            case 1: SWSPEC_PCAL_FAST(1, false) break;
            case 2: SWSPEC_PCAL_FAST(2, false) break;
            default: SWSPEC_PCAL_FAST(4, false) break;
            }
        }
#undef SWSPEC_PCAL_FAST
    } else {
        hipLaunchKernelGGL(k_pcal, dim3(g.nblk, nfft, njobs), dim3(g.threads), 0, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    PcRedArgs r;
    r.T = g.T; r.nblk = !fast ? g.nblk : (grp > 1 ? g.nblk_grp : g.nblk_fast); r.stride = g.nblk_max; r.nfft = nfft;
    r.R = kPcalRedBlocks;
    for (int j = 0; j < njobs; j++) r.jobs[j] = jobs[j];
    const int th = g.T <= 256 ? (256 / g.T) * g.T : g.T;   // lanes * T
    hipLaunchKernelGGL(k_pcal_reduce1, dim3(njobs, r.R), dim3(th), 0, s, r);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pcal_reduce2, dim3(njobs), dim3(std::min(g.T, 1024)), 0, s, r);
    return hipGetLastError();
}

hipError_t launch_pcal_emit(const PcalGeom& g, double* acc, float* out, double* tones, int ntones,
                            int tone_step, hipStream_t s)
{
    const int th = ((std::max(g.T, ntones) + 63) / 64) * 64;
    if (th > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pcal_emit, dim3(1), dim3(th), 0, s, acc, out, tones, g.T, ntones, tone_step);
    return hipGetLastError();
}

hipError_t launch_pcal_emit_jobs(const PcalGeom& g, double* const* acc, float* const* out, double* const* tones,
                                 int njobs, int ntones, int tone_step, hipStream_t s)
{
    const int th = ((std::max(g.T, ntones) + 63) / 64) * 64;
    if (th > 1024 || njobs < 1 || njobs > kMaxJobs) return hipErrorInvalidValue;
    EmitArgs a;
    a.T = g.T; a.ntones = ntones; a.step = tone_step;
    for (int j = 0; j < njobs; j++) { a.acc[j] = acc[j]; a.out[j] = out[j]; a.tones[j] = tones[j]; }
    hipLaunchKernelGGL(k_pcal_emit_jobs, dim3(njobs), dim3(th), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_finalize_jobs(const Plan& p, float* const* acc, float* const* out, int njobs, double scale,
                                int nslices, bool xpol, bool zero, hipStream_t s)
{
    if (njobs < 1 || njobs > kMaxJobs || nslices < 1) return hipErrorInvalidValue;
    FinArgs a;
    a.M = p.M; a.N1 = p.N1; a.N2 = p.N2; a.scale = scale; a.S = nslices;
    a.stride = (xpol ? 4 : 2) * (p.M + 1);
    a.zero = zero ? 1 : 0;
    for (int j = 0; j < njobs; j++) { a.acc[j] = acc[j]; a.out[j] = out[j]; }
    if (nslices >= 8) {   // short FFTs: parallel over slices (a tile grid would be a few blocks)
        const dim3 grid((unsigned)((p.M + 1 + 63) / 64), 1, njobs);
        if (xpol) hipLaunchKernelGGL(k_finalize_s<true>, grid, dim3(1024), 0, s, a);
        else hipLaunchKernelGGL(k_finalize_s<false>, grid, dim3(1024), 0, s, a);
        return hipGetLastError();
    }
    const dim3 grid((p.N2 + 31) / 32, (p.N1 + 31) / 32, njobs);
    if (xpol) hipLaunchKernelGGL(k_finalize_t<true>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_finalize_t<false>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vlba_blank(const JobDev* jobs, int njobs, int nfft, long long N, hipStream_t s)
{
    if (njobs < 1 || njobs > kMaxJobs || nfft < 1) return hipErrorInvalidValue;
    VbArgs a;
    a.nseg = -1;
    bool any = false;
    for (int j = 0; j < njobs; j++) {
        a.jobs[j] = jobs[j];
        if (is_m5a(jobs[j].up)) {
            const long long ns = m5a_segments(jobs[j].up, N);
            if (!jobs[j].up.vblank || jobs[j].up.vnseg != ns || (a.nseg >= 0 && a.nseg != ns))
                return hipErrorInvalidValue;
            a.nseg = (int)ns;
            any = true;
        }
    }
    if (!any) return hipSuccess;
    hipLaunchKernelGGL(k_vlba_blank, dim3((a.nseg + 63) / 64, nfft, njobs), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_unpack(const uint8_t* raw, Unpack up, long long count, float* out, hipStream_t s)
{
    const int blocks = (int)std::min<long long>((count + 255) / 256, 16384);
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3(blocks), dim3(256), 0, s, raw, up, count, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// planning (host): factorisation, thread shapes, FP64-derived tables
// ---------------------------------------------------------------------------

static bool smooth235(long long n)
{
    if (n < 1) return false;
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Threads per FFT such that each stage's butterflies per thread fit the
// register budget maxpts/R.
static int radices_threads(int L, const int* rad, int nst, int maxpts)
{
    int T = 1;
    for (int s = 0; s < nst; s++) {
        const int per = maxpts / rad[s];
        const int nb = L / rad[s];
        T = std::max(T, (nb + per - 1) / per);
    }
    return T;
}

// Radix list for an L-point sub-FFT.  Every stage is one LDS round trip, so the
// search takes the fewest stages from the register radices that fit maxpts
// points per thread (with `composite`, 10, 16, 20, 25, 32 as well as 2, 3, 4,
// 5, 8), then the fewest threads per FFT.
static void subplan_from(int L, const int* best, int bn, SubPlan& sp);

// fused_last: the column pass, whose last stage (a plain radix) runs fused
// with the Y store.  Among the shortest sequences it prefers the one with the
// largest plain radix (fewest fused rounds per thread), then the fewest
// threads, and orders composites first, smallest first, then plain radices
// smallest first, so the largest plain radix is last (3.2M columns: 10·25·5
// ran 4% faster than 25·25·2).
static bool subplan_radices(int L, SubPlan& sp, int maxpts, bool composite, bool fused_last = false)
{
    static const int kAll[] = {32, 25, 20, 16, 10, 8, 5, 4, 3, 2};
    static const int kBasic[] = {8, 5, 4, 3, 2};
    const int* set = composite ? kAll : kBasic;
    const int nset = composite ? 10 : 5;
    int best[kMaxStages], bn = 0, bT = 0, bP = 0;
    int cur[kMaxStages];
    auto plain_max = [&](int depth) {
        int m = 0;
        for (int i = 0; i < depth; i++)
            if (radix_a(cur[i]) == cur[i] && cur[i] > m) m = cur[i];
        return m;
    };
    // depth-first over non-increasing radix sequences
    auto dfs = [&](auto&& self, int n, int depth, int from) -> void {
        if (n == 1) {
            const int T = radices_threads(L, cur, depth, maxpts);
            const int P = fused_last ? plain_max(depth) : 0;
            if (bn == 0 || depth < bn || (depth == bn && (P > bP || (P == bP && T < bT)))) {
                bn = depth;
                bT = T;
                bP = P;
                for (int i = 0; i < depth; i++) best[i] = cur[i];
            }
            return;
        }
        if (depth >= kMaxStages || (bn != 0 && depth >= bn)) return;
        for (int j = from; j < nset; j++) {
            const int R = set[j];
            if (R > maxpts || n % R) continue;
            cur[depth] = R;
            self(self, n / R, depth + 1, j);
        }
    };
    dfs(dfs, L, 0, 0);
    if (bn == 0) return false;
    if (fused_last && bP > 0) {
        std::stable_sort(best, best + bn, [](int x, int y) {
            const bool cx = radix_a(x) != x, cy = radix_a(y) != y;
            if (cx != cy) return cx;
            return x < y;
        });
    } else {
        // Stage order: odd radices first (largest first), then even ones
        // smallest first.  A first stage (ns = 1) writes with a stride of R
        // elements, which for an even R maps many lanes of a b64 access onto
        // few LDS banks; an odd stride spreads them.  (Order does not change
        // the arithmetic.)
        std::stable_sort(best, best + bn, [](int x, int y) {
            const bool ox = (x & 1) != 0, oy = (y & 1) != 0;
            if (ox != oy) return ox;
            return ox ? x > y : x < y;
        });
    }
    subplan_from(L, best, bn, sp);
    return true;
}

// A sub-plan from an explicit stage sequence (tuning: SWSPEC_P2_RADICES);
// false unless the radices are supported, fit maxpts and multiply to L.
static bool subplan_forced(int L, const char* list, int maxpts, SubPlan& sp, int& cx)
{
    int rad[kMaxStages], n = 0;
    long long prod = 1;
    cx = 0;
    for (const char* q = list; *q && n < kMaxStages;) {
        const int R = atoi(q);
        const bool ok = R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 10 || R == 16 || R == 20 ||
                        R == 25 || R == 32;
        if (!ok || R > maxpts) return false;
        if (radix_a(R) != R) cx = 1;
        rad[n++] = R;
        prod *= R;
        while (*q && *q != ',') q++;
        if (*q == ',') q++;
    }
    if (n == 0 || prod != L) return false;
    subplan_from(L, rad, n, sp);
    return true;
}

static void subplan_from(int L, const int* best, int bn, SubPlan& sp)
{
    int ns = 1, twl = 0;
    sp.L = L;
    sp.nst = bn;
    sp.padst = -1;
    sp.pad = 0;
    for (int s = 0; s < bn; s++) {
        const int R = best[s];
        sp.radix[s] = R;
        sp.ns[s] = ns;
        sp.mag[s] = ns > 1 ? (unsigned)(((1ULL << 32) + (unsigned long long)ns - 1) / (unsigned long long)ns) : 0u;
        sp.twoff[s] = twl;
        twl += ns * radix_ntw(R);
        ns *= R;
    }
    sp.twlen = twl;
}

static int subplan_threads(const SubPlan& sp, int maxpts)
{
    return radices_threads(sp.L, sp.radix, sp.nst, maxpts);
}

static std::mutex g_tuned_mu;
static std::map<long long, PlanKnobs> g_tuned;

void plan_set_tuned(long long N, const PlanKnobs& k)
{
    std::lock_guard<std::mutex> l(g_tuned_mu);
    g_tuned[N] = k;
}

bool plan_get_tuned(long long N, PlanKnobs& k)
{
    std::lock_guard<std::mutex> l(g_tuned_mu);
    auto it = g_tuned.find(N);
    if (it == g_tuned.end()) return false;
    k = it->second;
    return true;
}

void plan_clear_tuned()
{
    std::lock_guard<std::mutex> l(g_tuned_mu);
    g_tuned.clear();
}

std::string plan_geometry(Plan& p, long long N, bool mfma64, const PlanKnobs* knobs)
{
    PlanKnobs tuned;
    if (!knobs && plan_get_tuned(N, tuned)) knobs = &tuned;
    if (N < 16 || (N % 2) != 0) return "FFT length must be even and >= 16";
    const long long M = N / 2;
    if (!smooth235(M)) return "FFT length must have prime factors 2, 3, 5 only";
    if (M > (1LL << 30)) return "FFT length too large";
    // N1 <= N2 closest to sqrt(M)
    long long best1 = 0;
    for (long long d = 1; d * d <= M; d++) {
        if (M % d) continue;
        if (!smooth235(d)) continue;
        best1 = d;
    }
    if (best1 < 2) return "no usable factorisation";
    // radix-64 MFMA stage: the column length must be a multiple of 64; the
    // divisor closest to sqrt(M) among those (either side of it)
    if (mfma64) {
        long long m1 = 0;
        double bestd = 0.0;
        for (long long d = 64; d <= M && d <= 8192; d += 64) {
            if (M % d || !smooth235(d) || M / d > 16384) continue;
            const double dist = std::fabs(std::log((double)d) - 0.5 * std::log((double)M));
            if (m1 == 0 || dist < bestd) { m1 = d; bestd = dist; }
        }
        if (m1 == 0) return "radix-64 MFMA stage needs a column length that is a multiple of 64";
        best1 = m1;
    }
    // tuning override: SWSPEC_N1 picks the column length when it divides M
    if (const char* e = mfma64 ? nullptr : getenv("SWSPEC_N1")) {
        const long long n1 = atoll(e);
        if (n1 >= 2 && (M % n1) == 0 && smooth235(n1) && smooth235(M / n1) && M / n1 <= 16384) best1 = n1;
    }
    p.N = N;
    p.M = M;
    p.N1 = (int)best1;
    p.N2 = (int)(M / best1);
    if (p.N2 > 16384) return "FFT length too large for two passes";
    // composite radices per pass: SWSPEC_COMPOSITE bit 0 = pass 1, bit 1 = pass 2
    int cxmask = 3;
    if (knobs && knobs->cx_mask >= 0) cxmask = knobs->cx_mask & 3;
    if (const char* e = getenv("SWSPEC_COMPOSITE")) cxmask = atoi(e) & 3;
    // sub-plan for (maxpts, composite); composite only where it saves a stage
    static const bool p1_order = [] {
        const char* e = getenv("SWSPEC_P1_ORDER");
        return !(e && atoi(e) == 0);
    }();
    auto subplan = [&](int L, int maxpts, bool cx, SubPlan& out, int& use_cx, bool fl = false) -> bool {
        SubPlan b;
        if (!subplan_radices(L, b, maxpts, false, fl)) return false;
        out = b;
        use_cx = 0;
        if (cx) {
            SubPlan c;
            if (subplan_radices(L, c, maxpts, true, fl) && c.nst < b.nst) { out = c; use_cx = 1; }
        }
        return true;
    };
    // accumulator slices: up to 1024, at most 16 MiB of slices per job (64 MiB
    // for the single pass, whose slices are its workgroups: 1024 at M = 8192)
    long long slice_mb = M <= 8192 ? 64 : 16;
    if (const char* e = getenv("SWSPEC_SLICE_MB")) slice_mb = std::max(1LL, atoll(e));
    p.smax = (int)std::max<long long>(1, std::min<long long>(1024, (slice_mb << 20) / (8 * (p.M + 1))));

    // Short FFTs (M <= 8192 complex points; SWSPEC_SHORT_MAX moves the bound,
    // 0 turns it off) run in one pass, k_short: the M-point FFT in LDS with at
    // most 256 threads per FFT, the fewest stages of 16 or 32 points per
    // thread (knobs->p1_maxpts forces one).  Its plan is the four-step layout
    // with N1 = M and N2 = 1 (the column pass's decode/window, bin k at
    // accumulator entry k).  Not with the MFMA stage.
    static const long long short_max = [] {
        const char* e = getenv("SWSPEC_SHORT_MAX");
        return e ? atoll(e) : 8192LL;
    }();
    if (!mfma64 && M <= short_max && !(knobs && knobs->single == 0)) {
        SubPlan s16, s32;
        int x16 = 0, x32 = 0;
        const bool ok16 = subplan(M, 16, cxmask & 1, s16, x16) && subplan_threads(s16, 16) <= 256;
        const bool ok32 = subplan(M, 32, cxmask & 1, s32, x32) && subplan_threads(s32, 32) <= 256;
        int mp = 0;
        if (ok16) mp = 16;
        if (ok32 && (!ok16 || s32.nst < s16.nst)) mp = 32;
        if (knobs && knobs->p1_maxpts == 16 && ok16) mp = 16;
        if (knobs && knobs->p1_maxpts == 32 && ok32) mp = 32;
        if (mp) {
            p.single = 1;
            p.N1 = (int)M;
            p.N2 = 1;
            p.B = 1;
            p.p1_maxpts = mp;
            p.sp1 = mp == 16 ? s16 : s32;
            p.p1_cx = mp == 16 ? x16 : x32;
            int T = subplan_threads(p.sp1, mp);
            // whole waves: the extra threads idle in the FFT stages but share
            // the decode/window fill and the bin pairs
            if (T > 64 && T % 64) T = std::min(256, ((T + 63) / 64) * 64);
            p.sp1.T = T;
            p.p1_threads = T;
            p.p1_fpw = 1;
            p.sp2 = SubPlan();
            p.p2_maxpts = 16;
            p.p2_cx = 0;
            p.p2_threads = 0;
            p.npairs = (int)(M / 2 + 1);
            p.twshift = 10;
            return "";
        }
    }

    SubPlan c16, c32;
    int cx16 = 0, cx32 = 0;
    if (!subplan(p.N1, 16, cxmask & 1, c16, cx16, p1_order) || !subplan(p.N1, 32, cxmask & 1, c32, cx32, p1_order))
        return "radix decomposition failed";

    // pass 1: B columns of N1 points (B a power of two dividing N2), <= 512
    // threads, LDS <= 40 KB per workgroup (80 KB when that is the only way to 4
    // columns).  Resident waves per CU are set by LDS (about 20K points per CU)
    // whatever B is, so smaller workgroups only add independent barrier domains
    // (3.2M: B = 4 beat B = 8 by 4%).  SWSPEC_P1_LDS_KB overrides the budget.
    // 32 points per thread when that saves an LDS stage without losing
    // columns, or when 16 leaves fewer than 4 columns and 32 buys more.
    static const size_t p1_lds_env = [] {
        const char* e = getenv("SWSPEC_P1_LDS_KB");
        const long v = e ? atol(e) : 0;
        return (size_t)((v >= 4 && v <= 160) ? v : 0) * 1024;
    }();
    auto pick_B = [&](int T1) {
        auto fit = [&](size_t budget) {
            for (int c : {16, 8, 4, 2, 1})
                if ((p.N2 % c) == 0 && (size_t)c * p.N1 * sizeof(float2) <= budget && c * T1 <= 512) return c;
            return 0;
        };
        if (p1_lds_env) return fit(p1_lds_env);
        const int b40 = fit(40 * 1024);
        return b40 >= 4 ? b40 : std::max(b40, fit(80 * 1024));
    };
    int T1 = subplan_threads(c16, 16);
    int B = pick_B(T1);
    p.p1_maxpts = 16;
    p.sp1 = c16;
    p.p1_cx = cx16;
    {
        const int T1b = subplan_threads(c32, 32);
        const int Bb = pick_B(T1b);
        if ((Bb > B && B < 4) || (Bb >= B && c32.nst < c16.nst)) {
            B = Bb;
            T1 = T1b;
            p.p1_maxpts = 32;
            p.sp1 = c32;
            p.p1_cx = cx32;
        }
    }
    // tuning overrides: knobs, then SWSPEC_P1_MAXPTS (16|32), SWSPEC_P1_BMAX (upper bound on B)
    if (knobs && (knobs->p1_maxpts == 16 || knobs->p1_maxpts == 32)) {
        const int mp = knobs->p1_maxpts;
        p.p1_maxpts = mp;
        p.sp1 = mp == 16 ? c16 : c32;
        p.p1_cx = mp == 16 ? cx16 : cx32;
        T1 = subplan_threads(p.sp1, mp);
        B = pick_B(T1);
    }
    if (const char* e = getenv("SWSPEC_P1_MAXPTS")) {
        const int mp = atoi(e);
        if (mp == 16 || mp == 32) {
            p.p1_maxpts = mp;
            p.sp1 = mp == 16 ? c16 : c32;
            p.p1_cx = mp == 16 ? cx16 : cx32;
            T1 = subplan_threads(p.sp1, mp);
            B = pick_B(T1);
        }
    }
    if (const char* e = getenv("SWSPEC_P1_RADICES")) {   // tuning: explicit column-stage sequence
        SubPlan f;
        int fcx = 0;
        if (subplan_forced(p.N1, e, p.p1_maxpts, f, fcx)) {
            p.sp1 = f;
            p.p1_cx = fcx;
            T1 = subplan_threads(p.sp1, p.p1_maxpts);
            B = pick_B(T1);
        }
    }
    if (knobs && knobs->p1_bmax > 0)
        while (B > 1 && B > knobs->p1_bmax) B >>= 1;
    if (const char* e = getenv("SWSPEC_P1_BMAX")) {
        const int bm = atoi(e);
        while (B > 1 && B > bm) B >>= 1;
    }
    if (B == 0) return "column FFT too large";
    // SWSPEC_P1_B forces a column count: any divisor of N2 works (the kernel
    // takes B as a stride, not a shift).  The powers of two above stay the
    // choice: B = 5 at 3.2M fills 250 of 256 lanes (B = 4: 200) and issues 20%
    // fewer VALU instructions, but its 40-byte row segments of Y cost 40% more
    // HBM write requests and its LDS stride conflicts, and it ran 19% slower.
    if (const char* e = getenv("SWSPEC_P1_B")) {
        const int c = atoi(e);
        if (c >= 1 && p.N2 % c == 0 && c * T1 <= 512 && (size_t)c * p.N1 * sizeof(float2) <= 160 * 1024) B = c;
    }
    // Whole waves: a last wave with few threads issues every instruction at
    // full cost.  Threads beyond what the FFT stages need idle there (q >=
    // L/R) but share the decode/window fill and the Y store (3.2M: 4 x 50 ->
    // 4 x 64 threads).  SWSPEC_WAVE_FILL bit 0 (pass 1), bit 1 (pass 2).
    static const int wave_fill = [] {
        const char* e = getenv("SWSPEC_WAVE_FILL");
        return e ? atoi(e) & 3 : 3;
    }();
    if ((wave_fill & 1) && B * T1 > 128 && (B * T1) % 64 != 0) {
        const int thr = ((B * T1 + 63) / 64) * 64;
        if (thr % B == 0 && thr <= 512) T1 = thr / B;
    }
    p.sp1.T = T1;
    p.B = B;
    p.p1_threads = B * T1;
    if (mfma64) {
        // stages [64, rest of N1/64]; B columns (the most that fit the LDS
        // budget), whole waves, and at most kMfPairs (row tile, column tile)
        // pairs per wave for the V = (N1/64)*B vectors in 16-column tiles
        const int L1 = p.N1 / 64;
        int list[kMaxStages], nl = 1, rcx = 0;
        list[0] = 64;
        if (L1 > 1) {
            SubPlan r;
            if (!subplan(L1, 16, cxmask & 1, r, rcx, p1_order)) return "radix decomposition failed";
            for (int s = 0; s < r.nst && nl < kMaxStages; s++) list[nl++] = r.radix[s];
        }
        subplan_from(p.N1, list, nl, p.sp1);
        const int T1m = nl > 1 ? radices_threads(p.N1, list + 1, nl - 1, 16) : 1;
        int Bm = 0, thr = 0;
        for (size_t budget : {(size_t)40 * 1024, (size_t)80 * 1024}) {
            for (int c : {16, 8, 4, 2, 1}) {
                if (p.N2 % c || (size_t)c * p.N1 * sizeof(float2) > budget) continue;
                const int nct = (L1 * c + 15) / 16;
                int t = ((c * T1m + 63) / 64) * 64;
                while (t <= 512 && ((8 + t / 64 - 1) / (t / 64)) * nct > kMfPairs) t += 64;
                if (t > 512) continue;
                Bm = c;
                thr = t;
                break;
            }
            if (Bm) break;
        }
        if (!Bm) return "radix-64 MFMA stage: no column-pass workgroup shape fits";
        // tuning overrides: SWSPEC_MF_B (columns, a divisor of N2), SWSPEC_MF_THREADS
        if (const char* e = getenv("SWSPEC_MF_B")) {
            const int c = atoi(e);
            if (c >= 1 && c <= 16 && p.N2 % c == 0 && (size_t)c * p.N1 * sizeof(float2) <= 80 * 1024) {
                const int nct = (L1 * c + 15) / 16;
                int t = ((c * T1m + 63) / 64) * 64;
                while (t % c) t += 64;
                while (t <= 512 && ((8 + t / 64 - 1) / (t / 64)) * nct > kMfPairs) t += 64;
                if (t <= 512) { Bm = c; thr = t; }
            }
        }
        if (const char* e = getenv("SWSPEC_MF_THREADS")) {
            const int t = atoi(e), nct = (L1 * Bm + 15) / 16;
            if (t >= 64 && t <= 512 && t % 64 == 0 && t % Bm == 0 && t >= Bm * T1m &&
                ((8 + t / 64 - 1) / (t / 64)) * nct <= kMfPairs)
                thr = t;
        }
        p.p1_maxpts = 16;
        p.p1_cx = rcx;
        p.B = Bm;
        p.sp1.T = thr / Bm;
        p.p1_threads = thr;
        p.mfma64 = 1;
    }

    // pass 2: 2 rows (x2 for cross-power pairs) of N2 points, <= 16 bins/thread,
    // <= 512 threads with cross-power
    SubPlan r16, r32;
    int rx16 = 0, rx32 = 0;
    if (!subplan(p.N2, 16, cxmask & 2, r16, rx16) || !subplan(p.N2, 32, cxmask & 2, r32, rx32))
        return "radix decomposition failed";
    p.p2_maxpts = 16;
    if (knobs && (knobs->p2_maxpts == 16 || knobs->p2_maxpts == 32)) p.p2_maxpts = knobs->p2_maxpts;
    if (const char* e = getenv("SWSPEC_P2_MAXPTS")) {
        const int mp = atoi(e);
        if (mp == 16 || mp == 32) p.p2_maxpts = mp;
    }
    p.sp2 = p.p2_maxpts == 16 ? r16 : r32;
    p.p2_cx = p.p2_maxpts == 16 ? rx16 : rx32;
    if (const char* e = getenv("SWSPEC_P2_RADICES")) {   // tuning: explicit row-stage sequence
        SubPlan f;
        int fcx = 0;
        if (subplan_forced(p.N2, e, p.p2_maxpts, f, fcx)) { p.sp2 = f; p.p2_cx = fcx; }
    }
    int T2 = subplan_threads(p.sp2, p.p2_maxpts);
    if (4 * T2 > 512) { p.p2_maxpts = 32; p.sp2 = r32; p.p2_cx = rx32; T2 = subplan_threads(p.sp2, 32); }
    if (4 * T2 > 512) return "row FFT too large";
    // whole waves for the two rows of a pair (the extra threads share the
    // row loads and the bin pairs), e.g. 3.2M: 2 x 86 -> 2 x 96
    if ((wave_fill & 2) && 2 * T2 > 64 && (2 * T2) % 64 != 0 && 4 * (((2 * T2 + 63) / 64) * 32) <= 512)
        T2 = ((2 * T2 + 63) / 64) * 32;
    p.sp2.T = T2;
    p.p2_threads = 2 * T2;
    // Padded penultimate row stage.  Its butterfly q = ns*j + i writes element
    // R*ns*j + i + r*ns: with ns < 16 and R*ns a multiple of 16, the 64 lanes
    // of a write land on only ns distinct bank pairs of the 32 four-byte banks
    // (3.2M, rows 5*16*16: 5 of 16, 13 lanes each).  Blocks R*ns + ns apart
    // put lane q at q mod 16, conflict-free; the last stage reads them with
    // stride R*ns + ns (it reads blocks whole, q < R*ns) and writes the plain
    // layout in place.  The padded stage is never stage 0 (a register first
    // stage writes outside the stage loop).  SWSPEC_P2_PAD=0 disables it.
    {
        static const bool pad_env = [] {
            const char* e = getenv("SWSPEC_P2_PAD");
            return !(e && atoi(e) == 0);
        }();
        SubPlan& sp = p.sp2;
        const int pen = sp.nst - 2;
        // only for the row kernels with a register first stage (16 points per
        // thread, first radix 4, 5 or 8): the others are built without it
        const bool f1k = p.p2_maxpts == 16 && (sp.radix[0] == 4 || sp.radix[0] == 5 || sp.radix[0] == 8);
        if (pad_env && f1k && pen >= 1 && sp.ns[pen] < 16 && (sp.radix[pen] * sp.ns[pen]) % 16 == 0) {
            sp.padst = pen;
            sp.pad = sp.ns[pen];
        }
    }
    p.npairs = p.N1 / 2 + 1;
    p.twshift = 10;
    // short columns: several FFTs per pass-1 workgroup (<= 256 threads, <= 64 KB LDS)
    p.p1_fpw = std::max(1, std::min(256 / p.p1_threads,
                                    (int)((64 * 1024) / ((size_t)p.N1 * p.B * sizeof(float2)))));
    if (p.mfma64) p.p1_fpw = 1;
    return "";
}

static void stage_twiddles(const SubPlan& sp, std::vector<float2>& out)
{
    out.assign(std::max(sp.twlen, 1), make_float2(1.f, 0.f));
    for (int s = 0; s < sp.nst; s++) {
        const int R = sp.radix[s], ns = sp.ns[s];
        const int A = radix_a(R), Bc = R / A, nt = radix_ntw(R);
        const long long D = (long long)ns * R;
        auto put = [&](int k, int slot, long long r) {
            const long long e = ((long long)k * r) % D;
            const double a = -2.0 * M_PI * (double)e / (double)D;
            out[sp.twoff[s] + k * nt + slot] = make_float2((float)cos(a), (float)sin(a));
        };
        for (int k = 0; k < ns; k++) {
            if (!radix_factored_tw(R)) {
                for (int r = 1; r < R; r++) put(k, r - 1, r);          // W_D^{k r}
            } else {
                for (int n1 = 1; n1 < A; n1++) put(k, n1 - 1, (long long)Bc * n1);   // W_D^{k B n1}
                for (int n2 = 1; n2 < Bc; n2++) put(k, A - 1 + n2 - 1, n2);          // W_D^{k n2}
            }
        }
    }
}

static float2 wroot(long long e, long long D)
{
    e %= D;
    const double a = -2.0 * M_PI * (double)e / (double)D;
    return make_float2((float)cos(a), (float)sin(a));
}

// The reference's windows (TaskCoreIPP.cpp:569-605) as a cosine series in
// theta = 2 pi (j + delta) / D.  Cosine: cos(pi (j - N/2 + 1/2)/N) =
// sin(pi (j + 1/2)/N); Cosine2 is its square = (1 - cos(2 pi (j + 1/2)/N))/2;
// Hamming, Hann and Blackman(alpha = 0.5) follow Intel's published symmetric
// formulas (0.54 - 0.46 cos, 0.5 - 0.5 cos, (a+1)/2 - cos/2 - (a/2) cos 2x over
// 2 pi n / (N - 1)).
static WindowGen window_gen(int type, long long n)
{
    WindowGen g;
    switch (type) {
    case 1: g.kind = 1; g.a0 = 0.f; g.a2 = 1.f; g.D = 2.0 * n; g.delta = 0.5; break;
    case 2: g.kind = 1; g.a0 = 0.5f; g.a1 = -0.5f; g.D = (double)n; g.delta = 0.5; break;
    case 3: g.kind = 1; g.a0 = 0.54f; g.a1 = -0.46f; g.D = (double)(n - 1); break;
    case 4: g.kind = 1; g.a0 = 0.5f; g.a1 = -0.5f; g.D = (double)(n - 1); break;
    case 5: g.kind = 1; g.a0 = 0.75f; g.a1 = -0.5f; g.a3 = -0.25f; g.D = (double)(n - 1); break;
    default: g.kind = 0; break;
    }
    return g;
}

template <typename T>
static hipError_t upload(T** d, const std::vector<T>& h, hipStream_t s)
{
    hipError_t e = hipMalloc((void**)d, h.size() * sizeof(T));
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// A operands of stage_mfma64: entry ((rt*4 + kc)*2 + h)*64 + lane holds the
// 8 FP16 values of lane `lane` for row tile rt, K step kc, h = 0 (hi) or 1
// (lo * 2^11).  Lane row i = lane % 16 is output row R = 16 rt + i (frequency
// r = R/2, part R % 2); slot t is input s = 16 kc + 4 (lane/16) + t/2, part
// t % 2.  With theta = 2 pi (r s mod 64)/64 the real 2 x 2 block of W_64^{rs}
// acting on (x re, x im) is [cos, sin; -sin, cos].
static std::vector<uint4> mfma64_table()
{
    std::vector<uint4> t(8 * 4 * 2 * 64);
    for (int rt = 0; rt < 8; rt++)
        for (int kc = 0; kc < 4; kc++)
            for (int lane = 0; lane < 64; lane++) {
                const int R = 16 * rt + (lane & 15), r = R >> 1, pr = R & 1;
                _Float16 hi[8], lo[8];
                for (int k = 0; k < 8; k++) {
                    const int sidx = 16 * kc + 4 * (lane >> 4) + (k >> 1), pc = k & 1;
                    const double th = 2.0 * M_PI * (double)((r * sidx) % 64) / 64.0;
                    const double c = cos(th), sn = sin(th);
                    const double val = pr == 0 ? (pc == 0 ? c : sn) : (pc == 0 ? -sn : c);
                    hi[k] = (_Float16)val;
                    lo[k] = (_Float16)((val - (double)hi[k]) * (double)kMfLo);
                }
                uint4 h, l;
                memcpy(&h, hi, 16);
                memcpy(&l, lo, 16);
                t[((rt * 4 + kc) * 2 + 0) * 64 + lane] = h;
                t[((rt * 4 + kc) * 2 + 1) * 64 + lane] = l;
            }
    return t;
}

std::string plan_create(Plan& p, long long N, int window, hipStream_t stream, bool mfma64, const PlanKnobs* knobs)
{
    std::string err = plan_geometry(p, N, mfma64, knobs);
    if (!err.empty()) return err;
    if (p.mfma64 && upload(&p.d_mft, mfma64_table(), stream) != hipSuccess)
        return "device allocation failed (MFMA stage table)";
    p.window = window;
    std::vector<float2> t;
    stage_twiddles(p.sp1, t);
    if (upload(&p.d_stw1, t, stream) != hipSuccess) return "device allocation failed (stage twiddles)";
    stage_twiddles(p.sp2, t);
    if (upload(&p.d_stw2, t, stream) != hipSuccess) return "device allocation failed (stage twiddles)";
    {
        const long long nb = 1LL << p.twshift;
        const long long na = (p.M + nb - 1) / nb + 1;
        std::vector<float2> A(na), Bt(nb);
        for (long long i = 0; i < na; i++) A[i] = wroot(i * nb, p.M);
        for (long long i = 0; i < nb; i++) Bt[i] = wroot(i, p.M);
        if (upload(&p.d_twA, A, stream) != hipSuccess || upload(&p.d_twB, Bt, stream) != hipSuccess)
            return "device allocation failed (four-step twiddles)";
    }
    {
        std::vector<float2> rA(p.N1), rB(p.N2);
        for (int k = 0; k < p.N1; k++) rA[k] = wroot(k, p.N);
        for (int k = 0; k < p.N2; k++) rB[k] = wroot((long long)p.N1 * k, p.N);
        if (upload(&p.d_rA, rA, stream) != hipSuccess || upload(&p.d_rB, rB, stream) != hipSuccess)
            return "device allocation failed (split twiddles)";
    }
    {
        // per-row window phase phi_e = 2 pi (2 N2 e mod D) / D, in FP64
        p.win = window_gen(window, p.N);
        std::vector<float2> ph(p.N1);
        for (int e = 0; e < p.N1; e++) {
            const double r = fmod(2.0 * (double)p.N2 * (double)e, p.win.D);
            const double a = 2.0 * M_PI * r / p.win.D;
            ph[e] = make_float2((float)cos(a), (float)sin(a));
        }
        if (upload(&p.d_winphi, ph, stream) != hipSuccess) return "device allocation failed (window)";
        p.win.phi = p.d_winphi;
        // per-column constants of the row fill, the column fill's win_coef in
        // FP64: sample r of column c has psi = 2 pi (2c + r + delta) / D
        if (p.win.kind != 0 && !p.single) {
            std::vector<float4> wc((size_t)2 * p.N2);
            for (int c = 0; c < p.N2; c++)
                for (int r = 0; r < 2; r++) {
                    const double x = ((double)(2 * c + r) + p.win.delta) / p.win.D;
                    const double a = 2.0 * M_PI * x;
                    const double co = cos(a), si = sin(a);
                    const double c2 = co * co - si * si, s2 = 2.0 * si * co;
                    wc[(size_t)2 * c + r] = make_float4((float)((double)p.win.a1 * co + (double)p.win.a2 * si),
                                                        (float)((double)p.win.a2 * co - (double)p.win.a1 * si),
                                                        (float)((double)p.win.a3 * c2), (float)(-(double)p.win.a3 * s2));
                }
            if (upload(&p.d_wcol, wc, stream) != hipSuccess) return "device allocation failed (window columns)";
        }
    }
    return "";
}

void plan_destroy(Plan& p)
{
    for (void* q : {(void*)p.d_stw1, (void*)p.d_stw2, (void*)p.d_twA, (void*)p.d_twB, (void*)p.d_rA,
                    (void*)p.d_rB, (void*)p.d_winphi, (void*)p.d_mft, (void*)p.d_wcol})
        if (q) (void)hipFree(q);
    p.d_stw1 = p.d_stw2 = p.d_twA = p.d_twB = p.d_rA = p.d_rB = nullptr;
    p.d_winphi = nullptr;
    p.d_mft = nullptr;
}

}  // namespace swspec
