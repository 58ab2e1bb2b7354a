// Mixed-radix (2, 3, 4, 5, 8) Stockham FFT of one row held in LDS.
//
// A length-n transform, n = prod(radix[p]), runs as `passes` autosorting
// passes.  Pass p with radix R and Ns = radix[0]*...*radix[p-1] does, for each
// butterfly j in [0, n/R), with k = j mod Ns:
//     a[r] = buf[j + r*n/R] * w^(r*k),   w = exp(DIR * 2*pi*i / (Ns*R))
//     b    = DFT_R(a)
//     buf[(j/Ns)*Ns*R + k + q*Ns] = b[q]
// which leaves the result in natural order (Stockham autosort).  Twiddles come
// from one table tw[t] = exp(-2*pi*i*t/n) (computed in FP64 on the host):
// w^(r*k) = tw[r*k*n/(Ns*R)] (conjugated for the inverse).  DIR = -1 is the
// forward transform, +1 the unnormalised inverse.
#pragma once

#include <hip/hip_runtime.h>

namespace prnu {

constexpr int kFftMaxPasses = 16;
constexpr int kFftMaxN = 8192;   // longest row held in LDS (64 KiB)
constexpr int kFftThreads = 256;

struct FftDesc {
    int n;
    int passes;
    int radix[kFftMaxPasses];
    const float2 *tw;  // n twiddles, device memory
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// i * a
__device__ __forceinline__ float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }

template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2 *tw, int idx) {
    float2 t = tw[idx];
    if (DIR > 0) t.y = -t.y;
    return t;
}

// In-register DFT of size R with exponent sign DIR.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2 *a);

template <>
__device__ __forceinline__ void dft<2, -1>(float2 *a) {
    const float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
}
template <>
__device__ __forceinline__ void dft<2, 1>(float2 *a) {
    dft<2, -1>(a);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2 *a) {
    const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    const float2 s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
    // DIR * i * (a1 - a3)
    const float2 r = DIR < 0 ? make_float2(d13.y, -d13.x) : make_float2(-d13.y, d13.x);
    a[0] = cadd(s02, s13);
    a[2] = csub(s02, s13);
    a[1] = cadd(d02, r);
    a[3] = csub(d02, r);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2 *a) { dft4<-1>(a); }
template <>
__device__ __forceinline__ void dft<4, 1>(float2 *a) { dft4<1>(a); }

template <int DIR>
__device__ __forceinline__ void dft3(float2 *a) {
    constexpr float kS = 0.866025403784438646763723170752936183f;  // sqrt(3)/2
    const float2 t1 = cadd(a[1], a[2]);
    const float2 t2 = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
    const float2 d = csub(a[1], a[2]);
    // DIR * i * sqrt(3)/2 * (a1 - a2)
    const float2 r = DIR < 0 ? make_float2(kS * d.y, -kS * d.x) : make_float2(-kS * d.y, kS * d.x);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(t2, r);
    a[2] = csub(t2, r);
}
template <>
__device__ __forceinline__ void dft<3, -1>(float2 *a) { dft3<-1>(a); }
template <>
__device__ __forceinline__ void dft<3, 1>(float2 *a) { dft3<1>(a); }

template <int DIR>
__device__ __forceinline__ void dft5(float2 *a) {
    constexpr float c1 = 0.309016994374947424102293417182819059f;   // cos(2pi/5)
    constexpr float c2 = -0.809016994374947424102293417182819059f;  // cos(4pi/5)
    constexpr float s1 = 0.951056516295153572116439333379382143f;   // sin(2pi/5)
    constexpr float s2 = 0.587785252292473129168705954639072769f;   // sin(4pi/5)
    const float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    const float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
    const float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
    const float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
    // n1 = s1*t3 + s2*t4, n2 = s2*t3 - s1*t4; multiplied by DIR * i
    const float2 n1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const float2 n2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    const float2 r1 = DIR < 0 ? make_float2(n1.y, -n1.x) : make_float2(-n1.y, n1.x);
    const float2 r2 = DIR < 0 ? make_float2(n2.y, -n2.x) : make_float2(-n2.y, n2.x);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, r1);
    a[4] = csub(m1, r1);
    a[2] = cadd(m2, r2);
    a[3] = csub(m2, r2);
}
template <>
__device__ __forceinline__ void dft<5, -1>(float2 *a) { dft5<-1>(a); }
template <>
__device__ __forceinline__ void dft<5, 1>(float2 *a) { dft5<1>(a); }

// Radix 8 as two radix-4 DFTs (even and odd inputs) joined by W8^q,
// W8 = exp(DIR * 2*pi*i / 8).
template <int DIR>
__device__ __forceinline__ void dft8(float2 *a) {
    constexpr float h = 0.707106781186547524400844362104849039f;  // sqrt(2)/2
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    dft4<DIR>(e);
    dft4<DIR>(o);
    const float2 o1 = DIR < 0 ? make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x))
                              : make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y));
    const float2 o2 = DIR < 0 ? make_float2(o[2].y, -o[2].x) : make_float2(-o[2].y, o[2].x);
    const float2 o3 = DIR < 0 ? make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y))
                              : make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
}
template <>
__device__ __forceinline__ void dft<8, -1>(float2 *a) { dft8<-1>(a); }
template <>
__device__ __forceinline__ void dft<8, 1>(float2 *a) { dft8<1>(a); }

// Radix 16 as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2): DFT-4 over n1, the
// twiddle W16^(n2 k1), DFT-4 over n2; W16 = exp(DIR * 2*pi*i / 16).
template <int DIR>
__device__ __forceinline__ void dft16(float2 *a) {
    constexpr float c1 = 0.923879532511286756128183189397f;  // cos(pi/8)
    constexpr float s1 = 0.382683432365089771728459984030f;  // sin(pi/8)
    constexpr float h = 0.707106781186547524400844362105f;   // sqrt(2)/2
    constexpr float d = DIR < 0 ? -1.0f : 1.0f;
    float2 t[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        t[n2][0] = a[n2];
        t[n2][1] = a[4 + n2];
        t[n2][2] = a[8 + n2];
        t[n2][3] = a[12 + n2];
        dft4<DIR>(t[n2]);
    }
    t[1][1] = cmul(t[1][1], make_float2(c1, d * s1));   // W16^1
    t[1][2] = cmul(t[1][2], make_float2(h, d * h));     // W16^2
    t[1][3] = cmul(t[1][3], make_float2(s1, d * c1));   // W16^3
    t[2][1] = cmul(t[2][1], make_float2(h, d * h));     // W16^2
    t[2][2] = DIR < 0 ? make_float2(t[2][2].y, -t[2][2].x) : make_float2(-t[2][2].y, t[2][2].x);  // W16^4 = DIR i
    t[2][3] = cmul(t[2][3], make_float2(-h, d * h));    // W16^6
    t[3][1] = cmul(t[3][1], make_float2(s1, d * c1));   // W16^3
    t[3][2] = cmul(t[3][2], make_float2(-h, d * h));    // W16^6
    t[3][3] = cmul(t[3][3], make_float2(-c1, -d * s1)); // W16^9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 u[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
        dft4<DIR>(u);
        a[k1] = u[0];
        a[k1 + 4] = u[1];
        a[k1 + 8] = u[2];
        a[k1 + 12] = u[3];
    }
}
template <>
__device__ __forceinline__ void dft<16, -1>(float2 *a) { dft16<-1>(a); }
template <>
__device__ __forceinline__ void dft<16, 1>(float2 *a) { dft16<1>(a); }

// j mod Ns for the butterfly's twiddle exponent; Ns is a power of two in
// every pass before the first radix-3 or radix-5 pass (fft_factor orders the
// radices 8, 4, 2, 3, 5).
__device__ __forceinline__ int fft_mod(int j, int Ns) { return (Ns & (Ns - 1)) == 0 ? (j & (Ns - 1)) : j % Ns; }

// One Stockham pass of radix R over buf[0..n) with kFftThreads threads.
template <int R, int DIR>
__device__ __forceinline__ void fft_pass(float2 *buf, int n, int Ns, const float2 *tw) {
    constexpr int kMaxB = (kFftMaxN / R + kFftThreads - 1) / kFftThreads;
    const int nb = n / R;
    const int twstep = n / (Ns * R);
    const int tid = threadIdx.x;
    float2 v[kMaxB][R];
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nb) {
            const int k = fft_mod(j, Ns);
            v[b][0] = buf[j];
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(buf[j + r * nb], twiddle<DIR>(tw, r * k * twstep));
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nb) {
            dft<R, DIR>(v[b]);
            const int k = fft_mod(j, Ns);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[base + q * Ns] = v[b][q];
        }
    }
    __syncthreads();
}

// Full in-LDS transform of buf[0..d.n).  Callers synchronise before (data in
// LDS) and may read buf after return (the last pass ends with a barrier).
template <int DIR>
__device__ inline void fft_lds(float2 *buf, const FftDesc &d) {
    int Ns = 1;
    for (int p = 0; p < d.passes; ++p) {
        const int R = d.radix[p];
        if (R == 8)
            fft_pass<8, DIR>(buf, d.n, Ns, d.tw);
        else if (R == 4)
            fft_pass<4, DIR>(buf, d.n, Ns, d.tw);
        else if (R == 2)
            fft_pass<2, DIR>(buf, d.n, Ns, d.tw);
        else if (R == 3)
            fft_pass<3, DIR>(buf, d.n, Ns, d.tw);
        else
            fft_pass<5, DIR>(buf, d.n, Ns, d.tw);
        Ns *= R;
    }
}

// One pass of fft_row: as fft_pass, but the first pass (FIRST, Ns = 1, no
// twiddles) takes its inputs from get(i) instead of LDS and the last pass
// (LAST, Ns * R = n) hands its outputs to put(i, v) instead of LDS.  Both
// touch i = j + r * (n/R) and i = j + q * Ns for consecutive j across the
// threads, so global accesses made by get/put are coalesced.
// LDS bytes fft_row needs for a row of n values.  (A pad slot per 8 values
// against bank conflicts of the strided writes measured slower: the padded
// addresses cost registers, hence occupancy.)
inline size_t fft_row_smem(int n) { return sizeof(float2) * (size_t)n; }

// Rows up to kFftSmallN (kFftMidN) values run with per-thread butterfly
// arrays sized for that length (fewer registers, more waves per SIMD).
constexpr int kFftSmallN = 2048;
constexpr int kFftMidN = 4096;

// CN > 0: the length (and CNS the pass's Ns) are compile-time constants, so
// strides, twiddle steps and the butterfly index's modulus fold into
// immediates (no integer division, LDS offsets in the instructions).
template <int R, int DIR, bool FIRST, bool LAST, int MAXN, int CN = 0, int CNS = 0, class Get, class Put>
__device__ __forceinline__ void fft_pass_io(float2 *buf, int n_rt, int Ns_rt, const float2 *tw, Get &get, Put &put) {
    constexpr int kMaxB = ((CN > 0 ? CN : MAXN) / R + kFftThreads - 1) / kFftThreads;
    const int n = CN > 0 ? CN : n_rt;
    const int Ns = CN > 0 ? CNS : Ns_rt;
    const int nb = n / R;
    const int twstep = n / (Ns * R);
    const int tid = threadIdx.x;
    float2 v[kMaxB][R];
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nb) {
            if constexpr (FIRST) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = get(j + r * nb);
            } else {
                const int k = fft_mod(j, Ns);
                v[b][0] = buf[j];
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(buf[j + r * nb], twiddle<DIR>(tw, r * k * twstep));
            }
        }
    }
    if constexpr (!FIRST) __syncthreads();  // every read of buf precedes the writes below
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int j = tid + b * kFftThreads;
        if (j < nb) {
            dft<R, DIR>(v[b]);
            const int k = FIRST ? 0 : fft_mod(j, Ns);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if constexpr (LAST)
                    put(base + q * Ns, v[b][q]);
                else
                    buf[base + q * Ns] = v[b][q];
            }
        }
    }
    __syncthreads();
}

template <int DIR, bool FIRST, bool LAST, int MAXN, class Get, class Put>
__device__ __forceinline__ void fft_pass_io_r(int R, float2 *buf, int n, int Ns, const float2 *tw, Get &get,
                                              Put &put) {
    if (R == 8)
        fft_pass_io<8, DIR, FIRST, LAST, MAXN>(buf, n, Ns, tw, get, put);
    else if (R == 4)
        fft_pass_io<4, DIR, FIRST, LAST, MAXN>(buf, n, Ns, tw, get, put);
    else if (R == 2)
        fft_pass_io<2, DIR, FIRST, LAST, MAXN>(buf, n, Ns, tw, get, put);
    else if (R == 3)
        fft_pass_io<3, DIR, FIRST, LAST, MAXN>(buf, n, Ns, tw, get, put);
    else
        fft_pass_io<5, DIR, FIRST, LAST, MAXN>(buf, n, Ns, tw, get, put);
}

// Radix plan of fft_factor (fft.hip) as a compile-time value, for the row
// lengths the chain and PCE run most.
struct RadixPlanC {
    int passes;
    int radix[kFftMaxPasses];
};
constexpr RadixPlanC radix_plan_c(int n) {
    RadixPlanC p{0, {}};
    int e2 = 0;
    while (n % 2 == 0) {
        ++e2;
        n /= 2;
    }
    int n8 = e2 / 3, n4 = 0, n2 = 0;
    if (e2 % 3 == 2) {
        n4 = 1;
    } else if (e2 % 3 == 1) {
        if (n8 > 0) {
            --n8;
            n4 = 2;
        } else {
            n2 = 1;
        }
    }
    for (int i = 0; i < n8; ++i) p.radix[p.passes++] = 8;
    for (int i = 0; i < n4; ++i) p.radix[p.passes++] = 4;
    for (int i = 0; i < n2; ++i) p.radix[p.passes++] = 2;
    while (n % 3 == 0) {
        p.radix[p.passes++] = 3;
        n /= 3;
    }
    while (n % 5 == 0) {
        p.radix[p.passes++] = 5;
        n /= 5;
    }
    return p;
}

template <int DIR, int N, int P, int NS, class Get, class Put>
__device__ __forceinline__ void fft_passes_c(float2 *buf, const float2 *tw, Get &get, Put &put) {
    constexpr RadixPlanC pl = radix_plan_c(N);
    if constexpr (P < pl.passes) {
        constexpr int R = pl.radix[P];
        fft_pass_io<R, DIR, P == 0, P == pl.passes - 1, N, N, NS>(buf, N, NS, tw, get, put);
        fft_passes_c<DIR, N, P + 1, NS * R>(buf, tw, get, put);
    }
}

// fft_row for a compile-time length N (at least two passes).
template <int DIR, int N, class Get, class Put>
__device__ __forceinline__ void fft_row_c(float2 *buf, const float2 *tw, Get &get, Put &put) {
    static_assert(radix_plan_c(N).passes >= 2, "fft_row_c: at least two passes");
    fft_passes_c<DIR, N, 0, 1>(buf, tw, get, put);
}

// ------------------------------------------------ one-wave transforms
// A transform of N values by ONE wave (a 64-thread workgroup), N / 64 values
// per lane: the same Stockham passes, with every lane busy in every pass and
// a block barrier that is a single wave's (no waiting on other waves).  LDS
// holds N values plus one pad slot per 2^SH (index p at p + p/2^SH): per 16
// for the radix-16 plans, per 8 for the radix-8 ones, so the strided writes
// of the first passes spread over all banks.
template <int N>
struct FftWavePlan;
template <>
struct FftWavePlan<1024> {  // radix 16, 16, 4
    static constexpr int kPadShift = 4;
};
template <>
struct FftWavePlan<1536> {  // radix 8, 8, 8, 3
    static constexpr int kPadShift = 3;
};
template <>
struct FftWavePlan<2048> {  // radix 16, 16, 8 on two waves
    static constexpr int kPadShift = 4;
};
template <int N>
constexpr int fft_wave_lds() { return N + (N >> FftWavePlan<N>::kPadShift); }
constexpr int fft_wave_lds(int n) { return n + n / 16; }  // the 1024-point plan's
template <int SH = 4>
__device__ __forceinline__ int fft_wave_pad(int p) { return p + (p >> SH); }

// SYNC_FIRST: a block barrier between the first pass's reads (get) and its
// writes, for callers whose get reads LDS that the waves' buffers overlay.
// Ordering of one wave's LDS accesses among its own lanes: a wave's LDS
// instructions execute in order, so the fences (which keep the compiler from
// moving memory accesses across) and a wave barrier suffice; other waves of
// the workgroup are not waited for.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// T: the threads of the transform (64: one wave; 128: two waves, whose block
// barrier then waits for the other wave).  WSYNC (T = 64 only): the
// transform's buffer is its wave's alone, so its passes order their LDS
// accesses with wave_lds_sync instead of block barriers (the caller
// synchronises the workgroup where waves share LDS).
template <int R, int DIR, bool FIRST, bool LAST, int N, int NS, bool SYNC_FIRST = false, int SH = 4, int T = 64,
          bool WSYNC = false, class Get, class Put>
__device__ __forceinline__ void fft_wave_pass(float2 *buf, const float2 *tw, Get &get, Put &put) {
    static_assert(!WSYNC || T == 64, "fft_wave_pass: wave-level ordering needs a one-wave transform");
    constexpr int nb = N / R;
    static_assert(nb % T == 0, "fft_wave_pass: N / R must be a multiple of the thread count");
    constexpr int per = nb / T;
    constexpr int twstep = N / (NS * R);
    const int lane = threadIdx.x & (T - 1);
    float2 v[per][R];
#pragma unroll
    for (int b = 0; b < per; ++b) {
        const int j = lane + T * b;
        if constexpr (FIRST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = get(j + r * nb);
        } else {
            const int k = j & (NS - 1);
            v[b][0] = buf[fft_wave_pad<SH>(j)];
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[b][r] = cmul(buf[fft_wave_pad<SH>(j + r * nb)], twiddle<DIR>(tw, r * k * twstep));
        }
    }
    if constexpr (!FIRST || SYNC_FIRST) {  // every read precedes the writes below
        if constexpr (WSYNC)
            wave_lds_sync();
        else
            __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < per; ++b) {
        const int j = lane + T * b;
        dft<R, DIR>(v[b]);
        const int k = FIRST ? 0 : (j & (NS - 1));
        const int base = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if constexpr (LAST)
                put(base + q * NS, v[b][q]);
            else
                buf[fft_wave_pad<SH>(base + q * NS)] = v[b][q];
        }
    }
    if constexpr (WSYNC)
        wave_lds_sync();
    else
        __syncthreads();
}

// 1024 points as radix 16, 16, 4 (16 values per lane); get(i) and put(i, X)
// as fft_row's.  buf: fft_wave_lds(1024) values.
template <int DIR, bool SYNC_FIRST = false, bool WSYNC = false, class Get, class Put>
__device__ __forceinline__ void fft_wave_1024(float2 *buf, const float2 *tw, Get get, Put put) {
    fft_wave_pass<16, DIR, true, false, 1024, 1, SYNC_FIRST, 4, 64, WSYNC>(buf, tw, get, put);
    fft_wave_pass<16, DIR, false, false, 1024, 16, false, 4, 64, WSYNC>(buf, tw, get, put);
    fft_wave_pass<4, DIR, false, true, 1024, 256, false, 4, 64, WSYNC>(buf, tw, get, put);
}

// 1536 points as radix 8, 8, 8, 3 (24 values per lane).  buf:
// fft_wave_lds<1536>() values.
template <int DIR, bool SYNC_FIRST = false, bool WSYNC = false, class Get, class Put>
__device__ __forceinline__ void fft_wave_1536(float2 *buf, const float2 *tw, Get get, Put put) {
    fft_wave_pass<8, DIR, true, false, 1536, 1, SYNC_FIRST, 3, 64, WSYNC>(buf, tw, get, put);
    fft_wave_pass<8, DIR, false, false, 1536, 8, false, 3, 64, WSYNC>(buf, tw, get, put);
    fft_wave_pass<8, DIR, false, false, 1536, 64, false, 3, 64, WSYNC>(buf, tw, get, put);
    fft_wave_pass<3, DIR, false, true, 1536, 512, false, 3, 64, WSYNC>(buf, tw, get, put);
}

// 2048 points as radix 16, 16, 8 on two waves (128 threads, 16 values per
// lane).  buf: fft_wave_lds<2048>() values.
template <int DIR, class Get, class Put>
__device__ __forceinline__ void fft_block_2048(float2 *buf, const float2 *tw, Get get, Put put) {
    fft_wave_pass<16, DIR, true, false, 2048, 1, false, 4, 128>(buf, tw, get, put);
    fft_wave_pass<16, DIR, false, false, 2048, 16, false, 4, 128>(buf, tw, get, put);
    fft_wave_pass<8, DIR, false, true, 2048, 256, false, 4, 128>(buf, tw, get, put);
}

// 2048 points as radix 8, 8, 8, 4 on four waves (256 threads, 8 values per
// thread), LDS padded one slot per 8: the plan of fft_row_c<2048> with a
// tenth of its bank conflicts.  buf: 2048 + 256 values.
constexpr int kFft2048PadLds = 2048 + 256;
template <int DIR, class Get, class Put>
__device__ __forceinline__ void fft_block_2048_r8(float2 *buf, const float2 *tw, Get get, Put put) {
    fft_wave_pass<8, DIR, true, false, 2048, 1, false, 3, 256>(buf, tw, get, put);
    fft_wave_pass<8, DIR, false, false, 2048, 8, false, 3, 256>(buf, tw, get, put);
    fft_wave_pass<8, DIR, false, false, 2048, 64, false, 3, 256>(buf, tw, get, put);
    fft_wave_pass<4, DIR, false, true, 2048, 512, false, 3, 256>(buf, tw, get, put);
}

// The one-wave transform of a compile-time length (1024 or 1536).
template <int DIR, int N, bool SYNC_FIRST = false, bool WSYNC = false, class Get, class Put>
__device__ __forceinline__ void fft_wave(float2 *buf, const float2 *tw, Get get, Put put) {
    if constexpr (N == 1024)
        fft_wave_1024<DIR, SYNC_FIRST, WSYNC>(buf, tw, get, put);
    else
        fft_wave_1536<DIR, SYNC_FIRST, WSYNC>(buf, tw, get, put);
}

// Transform of one row of length d.n <= MAXN from get(i) (i in [0, n), each
// read once) to put(i, X[i]) (each written once), through LDS buf
// (fft_row_smem(n) bytes) for the passes between; ends with a barrier, so
// the caller may reuse buf.
template <int DIR, int MAXN, class Get, class Put>
__device__ inline void fft_row(float2 *buf, const FftDesc &d, Get get, Put put) {
    // the hot lengths with compile-time strides
    if constexpr (MAXN >= 2048) {
        if (d.n == 2048) return fft_row_c<DIR, 2048>(buf, d.tw, get, put);
        if (d.n == 1536) return fft_row_c<DIR, 1536>(buf, d.tw, get, put);
    }
    if constexpr (MAXN >= 1024) {
        if (d.n == 1024) return fft_row_c<DIR, 1024>(buf, d.tw, get, put);
    }
    const int P = d.passes;
    if (P == 1) {
        fft_pass_io_r<DIR, true, true, MAXN>(d.radix[0], buf, d.n, 1, d.tw, get, put);
        return;
    }
    fft_pass_io_r<DIR, true, false, MAXN>(d.radix[0], buf, d.n, 1, d.tw, get, put);
    int Ns = d.radix[0];
    for (int p = 1; p < P - 1; ++p) {
        fft_pass_io_r<DIR, false, false, MAXN>(d.radix[p], buf, d.n, Ns, d.tw, get, put);
        Ns *= d.radix[p];
    }
    fft_pass_io_r<DIR, false, true, MAXN>(d.radix[P - 1], buf, d.n, Ns, d.tw, get, put);
}

// Values held in LDS by one multi-transform call (cols * n), bounded so the
// per-thread butterfly registers stay small.
constexpr int kFftMultiMax = 4096;

// One Stockham pass of radix R over `cols` independent transforms of length
// n stored back to back (buf[c*n .. c*n+n)), all butterflies of all
// transforms spread over the block: cols * n <= kFftMultiMax.
template <int R, int DIR>
__device__ __forceinline__ void fft_pass_multi(float2 *buf, int n, int cols, int Ns, const float2 *tw) {
    constexpr int kMaxB = (kFftMultiMax / R + kFftThreads - 1) / kFftThreads;
    const int nb = n / R;
    const int total = nb * cols;
    const int twstep = n / (Ns * R);
    const int tid = threadIdx.x;
    float2 v[kMaxB][R];
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int t = tid + b * kFftThreads;
        if (t < total) {
            const int c = t / nb, j = t - c * nb;
            const int k = fft_mod(j, Ns);
            const float2 *col = buf + c * n;
            v[b][0] = col[j];
#pragma unroll
            for (int r = 1; r < R; ++r) v[b][r] = cmul(col[j + r * nb], twiddle<DIR>(tw, r * k * twstep));
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxB; ++b) {
        const int t = tid + b * kFftThreads;
        if (t < total) {
            const int c = t / nb, j = t - c * nb;
            dft<R, DIR>(v[b]);
            const int k = fft_mod(j, Ns);
            const int base = c * n + (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[base + q * Ns] = v[b][q];
        }
    }
    __syncthreads();
}

// `cols` in-LDS transforms of length d.n at buf + c*d.n (cols*d.n <= kFftMultiMax).
template <int DIR>
__device__ inline void fft_lds_multi(float2 *buf, const FftDesc &d, int cols) {
    int Ns = 1;
    for (int p = 0; p < d.passes; ++p) {
        const int R = d.radix[p];
        if (R == 8)
            fft_pass_multi<8, DIR>(buf, d.n, cols, Ns, d.tw);
        else if (R == 4)
            fft_pass_multi<4, DIR>(buf, d.n, cols, Ns, d.tw);
        else if (R == 2)
            fft_pass_multi<2, DIR>(buf, d.n, cols, Ns, d.tw);
        else if (R == 3)
            fft_pass_multi<3, DIR>(buf, d.n, cols, Ns, d.tw);
        else
            fft_pass_multi<5, DIR>(buf, d.n, cols, Ns, d.tw);
        Ns *= R;
    }
}

}  // namespace prnu
