// Device helpers shared by the sweep kernels: quadruple gate kernels (the 4
// amplitudes a two-qubit matchgate couples) and wave-level FP64 reductions.
#pragma once
#include "mg_device.hpp"

namespace mg {

// ---- wave-level transposing reduction -------------------------------------------
// 32-bit exchange with a partner lane whose index differs from this lane's in bit O
// (log2 O) and possibly in lower bits, which is all that a butterfly reduction over the
// bits 5..0 needs: each step's partner mask has a new highest bit, so the masks stay
// independent and every lane still sums all 64.  O = 1, 2: DPP quad permutations
// ([1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E: lane ^ 1, lane ^ 2); O = 4: DPP row_half_mirror
// (0x141: lane ^ 7 within each 8-lane half row); O = 8: DPP row_ror:8 (0x128: lane ^ 8
// within each 16-lane row); O = 16: ds_swizzle in bit-mask mode (and 0x1F, or 0, xor 16);
// O = 32: ds_bpermute.  The DPP forms are plain vector ALU moves: no LDS-unit round trip
// and no wait on the LDS counter.  The patterns are immediates, so no per-lane address
// register stays live across the sweep.
template <int O>
__device__ inline uint32_t xlane32(uint32_t x)
{
    if constexpr (O == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);
    else if constexpr (O == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);
    else if constexpr (O == 4) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, true);
    else if constexpr (O == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true);
    else if constexpr (O < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (O << 10) | 0x1F);
    else return (uint32_t)__shfl_xor((int)x, O);
}

template <int O>
__device__ inline float xlane(float x) { return __builtin_bit_cast(float, xlane32<O>(__builtin_bit_cast(uint32_t, x))); }

template <int O>
__device__ inline double xlane(double x)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint64_t lo = xlane32<O>((uint32_t)u), hi = xlane32<O>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}

template <typename T>
struct Bits;
template <>
struct Bits<float> {
    static __device__ inline uint32_t lo(float x) { return __builtin_bit_cast(uint32_t, x); }
};

// One halving step at offset O for the value pair (a, b) = (v[k], v[k+H]):
// lanes with bit O clear return a + partner's a, lanes with it set b + partner's b.
// With the permlane swaps (offsets 32 and 16) no lane select is needed: the swap leaves
// (a_lo, b_lo) in one register and (a_hi, b_hi) in the other (halves of 32 lanes, or
// alternating rows of 16), so their sum is a + partner's a where bit O is clear and
// b + partner's b where it is set: the same operands as the select form, hence the same bits.
template <int O>
__device__ inline float halve_pair(float a, float b, int lane)
{
    if constexpr (O == 32 || O == 16) {
        const uint32_t ua = __builtin_bit_cast(uint32_t, a), ub = __builtin_bit_cast(uint32_t, b);
        uint32_t na, nb;
        if constexpr (O == 32) {
            auto r = __builtin_amdgcn_permlane32_swap(ua, ub, false, false);
            na = r[0];
            nb = r[1];
        } else {
            auto r = __builtin_amdgcn_permlane16_swap(ua, ub, false, false);
            na = r[0];
            nb = r[1];
        }
        return __builtin_bit_cast(float, na) + __builtin_bit_cast(float, nb);
    } else {
        const bool up = (lane & O) != 0;
        const float send = up ? a : b;
        const float keep = up ? b : a;
        return keep + xlane<O>(send);
    }
}

template <int O>
__device__ inline double halve_pair(double a, double b, int lane)
{
    if constexpr (O == 32 || O == 16) {
        const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
        const uint32_t al = (uint32_t)ua, ah = (uint32_t)(ua >> 32), bl = (uint32_t)ub, bh = (uint32_t)(ub >> 32);
        uint32_t nal, nbl, nah, nbh;
        if constexpr (O == 32) {
            auto r0 = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
            auto r1 = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
            nal = r0[0]; nbl = r0[1]; nah = r1[0]; nbh = r1[1];
        } else {
            auto r0 = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
            auto r1 = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
            nal = r0[0]; nbl = r0[1]; nah = r1[0]; nbh = r1[1];
        }
        const double na = __builtin_bit_cast(double, (uint64_t)nal | ((uint64_t)nah << 32));
        const double nb = __builtin_bit_cast(double, (uint64_t)nbl | ((uint64_t)nbh << 32));
        return na + nb;
    } else {
        const bool up = (lane & O) != 0;
        const double send = up ? a : b;
        const double keep = up ? b : a;
        return keep + xlane<O>(send);
    }
}

// Transposing butterfly: on return lane l holds the 64-lane sum of value
// (l >> (6 - log2 N)); N values per lane, T = float or double.
template <int N, int O, typename T>
__device__ inline void wave_halve(T* v, int lane)
{
    if constexpr (N > 1) {
        constexpr int H = N / 2;
#pragma unroll
        for (int k = 0; k < H; k++) v[k] = halve_pair<O>(v[k], v[k + H], lane);
        wave_halve<H, O / 2, T>(v, lane);
    } else if constexpr (O >= 1) {
        v[0] += xlane<O>(v[0]);
        wave_halve<1, O / 2, T>(v, lane);
    }
}

// broadcast a value that is uniform across the wave into scalar registers
__device__ inline float uniform(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ inline double uniform(double x)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
}

// n gate coefficients (wave-uniform) loaded into scalar registers
template <typename R, int N>
__device__ inline void load_coef(const cpx<R>* src, cpx<R>* dst)
{
#pragma unroll
    for (int k = 0; k < N; k++) {
        // measured on MI355X: scalar coefficients relieve the FP64 kernels' VGPR pressure but
        // slow the FP32 kernels, which have VGPRs to spare.  (Scalar loads through the constant
        // address space were slower still in the LDS-resident kernels: their waits share the
        // lgkm counter with the LDS traffic.)
        if constexpr (sizeof(R) == 8) dst[k] = {uniform(src[k].re), uniform(src[k].im)};
        else dst[k] = src[k];
    }
}

template <int N>
constexpr int log2c() { return N <= 1 ? 0 : 1 + log2c<N / 2>(); }

__device__ inline double wave_sum(double x)
{
    x += xlane<32>(x);
    x += xlane<16>(x);
    x += xlane<8>(x);
    x += xlane<4>(x);
    x += xlane<2>(x);
    x += xlane<1>(x);
    return x;
}

// ---- quadruple kernels -----------------------------------------------------------
template <typename R>
struct Quad {
    cpx<R> x, y, z, w;
};

// LDS position of amplitude a.  A wave's LDS access is served in groups of lanes
// that together cover all banks once (kLdsBankBytes per cycle): 8-byte (FP32
// complex) elements fill a group with BANK/8 lanes, 16-byte (FP64 complex) ones with
// BANK/16.  A quadruple access with a low wire bit keeps one of the group's index
// bits constant across the lanes and would collide.  XOR-ing the two index bits just
// above the group (bits B, B+1) into the low B bits with the masks below makes every
// nearest-neighbour pair (k, k+1) conflict-free: consecutive rows of the 2-column
// mask matrix are linearly independent and row B-1 has its second bit set.  Only
// the low B bits change, so the map permutes every aligned 2^(B+2)-element block.
constexpr int kLdsBankBytes = 128;
template <int B>
struct SwzMasks;
template <>
struct SwzMasks<3> { static constexpr uint32_t m0 = 5u, m1 = 6u; };
template <>
struct SwzMasks<4> { static constexpr uint32_t m0 = 5u, m1 = 10u; };
template <>
struct SwzMasks<5> { static constexpr uint32_t m0 = 21u, m1 = 26u; };

template <typename R>
__device__ inline uint32_t lds_swz(uint32_t a)
{
    constexpr int lanes = kLdsBankBytes / (int)sizeof(cpx<R>);
    constexpr int B = lanes == 8 ? 3 : (lanes == 16 ? 4 : 5);
    return a ^ (((a >> B) & 1u) * SwzMasks<B>::m0) ^ (((a >> (B + 1)) & 1u) * SwzMasks<B>::m1);
}

// SWZ: the state lives in LDS with the lds_swz layout (else plain global memory)
template <typename R, bool SWZ = false>
__device__ inline cpx<R>& sel(cpx<R>* s, uint32_t a)
{
    if constexpr (SWZ) return s[lds_swz<R>(a)];
    else return s[a];
}

template <typename R, bool SWZ = false>
__device__ inline const cpx<R>& sel(const cpx<R>* s, uint32_t a)
{
    if constexpr (SWZ) return s[lds_swz<R>(a)];
    else return s[a];
}

// element index of amplitude a: swizzled in LDS, plain in memory
template <typename R, bool SWZ>
__device__ inline uint32_t swz_idx(uint32_t a)
{
    if constexpr (SWZ) return lds_swz<R>(a);
    else return a;
}

// the quadruple whose first element sits at index sb = swz_idx(b): lds_swz is linear over GF(2)
// and b has both wire bits clear, so the other three sit at sb ^ olo, sb ^ ohi and sb ^ olo ^ ohi
// with olo, ohi the (wave-uniform) indices of the wire bits themselves
template <typename R>
__device__ inline Quad<R> load_quad_o(const cpx<R>* s, uint32_t sb, uint32_t olo, uint32_t ohi)
{
    return {s[sb], s[sb ^ olo], s[sb ^ ohi], s[sb ^ olo ^ ohi]};
}

template <typename R>
__device__ inline void store_quad_o(cpx<R>* s, uint32_t sb, uint32_t olo, uint32_t ohi, const Quad<R>& q)
{
    s[sb] = q.x;
    s[sb ^ olo] = q.y;
    s[sb ^ ohi] = q.z;
    s[sb ^ olo ^ ohi] = q.w;
}

template <typename R, bool SWZ = false>
__device__ inline Quad<R> load_quad(const cpx<R>* s, uint32_t b, int lo, int hi)
{
    return {sel<R, SWZ>(s, b), sel<R, SWZ>(s, b | (1u << lo)), sel<R, SWZ>(s, b | (1u << hi)),
            sel<R, SWZ>(s, b | (1u << lo) | (1u << hi))};
}

template <typename R, bool SWZ = false>
__device__ inline void store_quad(cpx<R>* s, uint32_t b, int lo, int hi, const Quad<R>& q)
{
    sel<R, SWZ>(s, b) = q.x;
    sel<R, SWZ>(s, b | (1u << lo)) = q.y;
    sel<R, SWZ>(s, b | (1u << hi)) = q.z;
    sel<R, SWZ>(s, b | (1u << lo) | (1u << hi)) = q.w;
}

// V q  (reference src/gate.c:641-646)
template <typename R>
__device__ inline Quad<R> gapply(const cpx<R>* g, const Quad<R>& q)
{
    return {cfma(g[4], q.x, cmul(g[5], q.w)), cfma(g[0], q.y, cmul(g[1], q.z)),
            cfma(g[2], q.y, cmul(g[3], q.z)), cfma(g[6], q.x, cmul(g[7], q.w))};
}

// V q + acc
template <typename R>
__device__ inline Quad<R> gapply_add(const cpx<R>* g, const Quad<R>& q, const Quad<R>& a)
{
    return {cfma(g[4], q.x, cfma(g[5], q.w, a.x)), cfma(g[0], q.y, cfma(g[1], q.z, a.y)),
            cfma(g[2], q.y, cfma(g[3], q.z, a.z)), cfma(g[6], q.x, cfma(g[7], q.w, a.w))};
}

// V^T q  (reference src/gate.c:760-764)
template <typename R>
__device__ inline Quad<R> gapplyT(const cpx<R>* g, const Quad<R>& q)
{
    return {cfma(g[4], q.x, cmul(g[6], q.w)), cfma(g[0], q.y, cmul(g[2], q.z)),
            cfma(g[1], q.y, cmul(g[3], q.z)), cfma(g[5], q.x, cmul(g[7], q.w))};
}

template <typename R>
__device__ inline Quad<R> gapplyT_add(const cpx<R>* g, const Quad<R>& q, const Quad<R>& a)
{
    return {cfma(g[4], q.x, cfma(g[6], q.w, a.x)), cfma(g[0], q.y, cfma(g[2], q.z, a.y)),
            cfma(g[1], q.y, cfma(g[3], q.z, a.z)), cfma(g[5], q.x, cfma(g[7], q.w, a.w))};
}

// conj(a) b + c
template <typename R>
__host__ __device__ inline cpx<R> cfma_conj(cpx<R> a, cpx<R> b, cpx<R> c)
{
    return {rfma(a.re, b.re, rfma(a.im, b.im, c.re)), rfma(a.re, b.im, rfma(-a.im, b.re, c.im))};
}

template <typename R>
__host__ __device__ inline cpx<R> cmul_conj(cpx<R> a, cpx<R> b)
{
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

// V^dagger q = conj(V)^T q: the inverse of a unitary V
template <typename R>
__device__ inline Quad<R> gapplyH(const cpx<R>* g, const Quad<R>& q)
{
    return {cfma_conj(g[4], q.x, cmul_conj(g[6], q.w)), cfma_conj(g[0], q.y, cmul_conj(g[2], q.z)),
            cfma_conj(g[1], q.y, cmul_conj(g[3], q.z)), cfma_conj(g[5], q.x, cmul_conj(g[7], q.w))};
}

// c - a b, as two fused multiply-adds per part
template <typename R>
__host__ __device__ inline cpx<R> cfms(cpx<R> a, cpx<R> b, cpx<R> c)
{
    return {rfma(-a.re, b.re, rfma(a.im, b.im, c.re)), rfma(-a.re, b.im, rfma(-a.im, b.re, c.im))};
}

// a - Z q  (the uncompute's dpsi_{g+1} - Z_g psi_g), FMAs only: no products rounded apart and
// no separate subtraction
template <typename R>
__device__ inline Quad<R> gapply_sub(const cpx<R>* g, const Quad<R>& q, const Quad<R>& a)
{
    return {cfms(g[4], q.x, cfms(g[5], q.w, a.x)), cfms(g[0], q.y, cfms(g[1], q.z, a.y)),
            cfms(g[2], q.y, cfms(g[3], q.z, a.z)), cfms(g[6], q.x, cfms(g[7], q.w, a.w))};
}

// a*b as the first term of a sum: the operations of cfma(a, b, 0) without the zero operands
template <typename R>
__device__ inline cpx<R> cfma0(cpx<R> a, cpx<R> b)
{
    return {rfma(a.re, b.re, -a.im * b.im), rfma(a.re, b.im, a.im * b.re)};
}

// acc[0..7] += d (x) p  (reference src/gate.c:753-757): center then corner entries.
// INIT: acc = d (x) p, the first term (the sums need no zeroed registers)
template <typename R, bool INIT = false>
__device__ inline void outer_acc(cpx<R>* acc, const Quad<R>& d, const Quad<R>& p)
{
    if constexpr (INIT) {
        acc[0] = cfma0(d.y, p.y);
        acc[1] = cfma0(d.y, p.z);
        acc[2] = cfma0(d.z, p.y);
        acc[3] = cfma0(d.z, p.z);
        acc[4] = cfma0(d.x, p.x);
        acc[5] = cfma0(d.x, p.w);
        acc[6] = cfma0(d.w, p.x);
        acc[7] = cfma0(d.w, p.w);
    } else {
        acc[0] = cfma(d.y, p.y, acc[0]);
        acc[1] = cfma(d.y, p.z, acc[1]);
        acc[2] = cfma(d.z, p.y, acc[2]);
        acc[3] = cfma(d.z, p.z, acc[3]);
        acc[4] = cfma(d.x, p.x, acc[4]);
        acc[5] = cfma(d.x, p.w, acc[5]);
        acc[6] = cfma(d.w, p.x, acc[6]);
        acc[7] = cfma(d.w, p.w, acc[7]);
    }
}

// One gate on N of a thread's quadruples (q0, q0 + stride, ...), which are disjoint: all loads
// first, then the arithmetic and the stores, so that the quadruples' latencies overlap.  The
// accumulators see the quadruples in increasing q, as a loop over them one at a time would.
template <typename R, int MODE, bool SW, int N>
__device__ inline void fwd_quads(uint32_t q0, uint32_t stride, int lo, int hi, uint32_t olo, uint32_t ohi,
                                 const cpx<R>* cf, const cpx<R>* pin, cpx<R>* pout, const cpx<R>* din, cpx<R>* dout)
{
    uint32_t sb[N];
    Quad<R> p[N], d[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        sb[i] = swz_idx<R, SW>(quad_base(q0 + i * stride, lo, hi));
        p[i] = load_quad_o(pin, sb[i], olo, ohi);
        if constexpr (MODE == 2) d[i] = load_quad_o(din, sb[i], olo, ohi);
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        if constexpr (MODE == 2) store_quad_o(dout, sb[i], olo, ohi, gapply_add(cf, d[i], gapply(cf + 16, p[i])));
        store_quad_o(pout, sb[i], olo, ohi, gapply(cf, p[i]));
    }
}

// the reverse step (states resident, psi recovered with V^{-1}) on N disjoint quadruples;
// INIT: the first quadruple's outer products start the sums ga, ha
template <typename R, int MODE, bool SW, int N, bool INIT = false>
__device__ inline void rev_quads(uint32_t q0, uint32_t stride, int lo, int hi, uint32_t olo, uint32_t ohi,
                                 const cpx<R>* cf, cpx<R>* lam, cpx<R>* psi, cpx<R>* dlam, cpx<R>* dpsi, cpx<R>* ga,
                                 cpx<R>* ha)
{
    uint32_t sb[N];
    Quad<R> l[N], pn[N], dn[N], dl[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        sb[i] = swz_idx<R, SW>(quad_base(q0 + i * stride, lo, hi));
        l[i] = load_quad_o(lam, sb[i], olo, ohi);
        pn[i] = load_quad_o(psi, sb[i], olo, ohi);
        if constexpr (MODE == 2) {
            dn[i] = load_quad_o(dpsi, sb[i], olo, ohi);
            dl[i] = load_quad_o(dlam, sb[i], olo, ohi);
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
        const Quad<R> p = gapply(cf + 8, pn[i]);
        store_quad_o(psi, sb[i], olo, ohi, p);
        if (INIT && i == 0) outer_acc<R, true>(ga, l[i], p);
        else outer_acc(ga, l[i], p);
        if constexpr (MODE == 2) {
            const Quad<R> dp = gapply(cf + 8, gapply_sub(cf + 16, p, dn[i]));
            store_quad_o(dpsi, sb[i], olo, ohi, dp);
            if (INIT && i == 0) outer_acc<R, true>(ha, dl[i], p);
            else outer_acc(ha, dl[i], p);
            outer_acc(ha, l[i], dp);
            store_quad_o(dlam, sb[i], olo, ohi, gapplyT_add(cf, dl[i], gapplyT(cf + 16, l[i])));
        }
        store_quad_o(lam, sb[i], olo, ohi, gapplyT(cf, l[i]));
    }
}

// One gate's reverse step over a thread's quadruples q, q + T, ... < Q, two at a time (TWO: the
// thread may have more than one), in increasing q; the first sets the sums ga, ha, which stay zero
// for a thread without quadruples
template <typename R, int MODE, bool SW, bool TWO>
__device__ inline void rev_gate(uint32_t q, uint32_t T, uint32_t Q, int lo, int hi, uint32_t olo, uint32_t ohi,
                                const cpx<R>* cf, cpx<R>* lam, cpx<R>* psi, cpx<R>* dlam, cpx<R>* dpsi, cpx<R>* ga,
                                cpx<R>* ha)
{
    if (TWO && q + T < Q) {
        rev_quads<R, MODE, SW, 2, true>(q, T, lo, hi, olo, ohi, cf, lam, psi, dlam, dpsi, ga, ha);
        for (q += 2 * T; q + T < Q; q += 2 * T)
            rev_quads<R, MODE, SW, 2>(q, T, lo, hi, olo, ohi, cf, lam, psi, dlam, dpsi, ga, ha);
        if (q < Q) rev_quads<R, MODE, SW, 1>(q, T, lo, hi, olo, ohi, cf, lam, psi, dlam, dpsi, ga, ha);
    } else if (q < Q) {
        rev_quads<R, MODE, SW, 1, true>(q, T, lo, hi, olo, ohi, cf, lam, psi, dlam, dpsi, ga, ha);
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            ga[e] = {R(0), R(0)};
            ha[e] = {R(0), R(0)};
        }
    }
}

}  // namespace mg
