// pa_rng.h -- counter-based random numbers and samplers shared by host and
// device code of the engine.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Key = the
// run's SEED.  Counter words: c0 = block of four draws, c1 = sub-stream,
// c2 = sweep, c3 = (model slot << 24) | (purpose tag << 16).  A draw is
// identified by (c1, c2, c3, slot); slot t lives in block t/4, word t%4, so
// any thread can compute any draw without a shared sequential state.  The
// reference seeds every element's GSL generator with the same SEED
// (src/PhyloAcc-ST/bpp_c.hpp:146-147); likewise the stream here depends on
// SEED only, never on an element's position in a batch.
//
// Samplers: categorical by inverse CDF over unnormalised weights; Gamma by
// Marsaglia & Tsang (ACM TOMS 26(3), 2000) with Box-Muller normals
// (boosted by U^(1/a) for shape < 1); Beta as X/(X+Y) of two Gammas.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define PA_HD __host__ __device__ inline
#else
#define PA_HD inline
#endif

namespace pa {

enum { TAG_ANC = 1, TAG_Z = 2, TAG_R2 = 3, TAG_R1 = 4, TAG_TRANS = 5, TAG_PI = 6, TAG_WL = 7, TAG_GT_BRANCH = 8,
       TAG_GT_TREE = 9, TAG_GT_HIST = 10 };

struct U4 { uint32_t x, y, z, w; };

PA_HD U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

PA_HD double u01(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

PA_HD uint32_t tagword(int model_slot, int tag) { return ((uint32_t)model_slot << 24) | ((uint32_t)tag << 16); }

PA_HD uint32_t pick(const U4 &v, int lane)
{
    return lane == 0 ? v.x : (lane == 1 ? v.y : (lane == 2 ? v.z : v.w));
}

struct Key { uint32_t k0, k1; };

PA_HD double slot_uniform(Key k, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t t)
{
    U4 v = philox(t >> 2, c1, c2, c3, k.k0, k.k1);
    return u01(pick(v, (int)(t & 3)));
}

// Sequential reader over one sub-stream (c1, c2, c3).
struct Stream {
    Key k; uint32_t c1, c2, c3, idx;
    PA_HD double next()
    {
        double u = slot_uniform(k, c1, c2, c3, idx);
        idx++;
        return u;
    }
};

#define PA_GAMMA_MAXIT 64

// capped (optional): set to 1 when the rejection loop stops at its cap
PA_HD double gamma_draw(Stream &s, double a, double b, int *capped = nullptr)
{
    double boost = 1.0;
    if (a < 1.0) {
        double u = s.next();
        boost = pow(u, 1.0 / a);
        a += 1.0;
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (int it = 0; it < PA_GAMMA_MAXIT; it++) {
        double u1 = s.next(), u2 = s.next();
        double x = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        double u = s.next();
        if (u < 1.0 - 0.0331 * (x * x) * (x * x)) return d * v * b * boost;
        if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v * b * boost;
    }
    if (capped) *capped = 1;
    return d * b * boost;
}

PA_HD double beta_draw(Key k, uint32_t cx, uint32_t cy, uint32_t c2, uint32_t c3, double a, double b)
{
    Stream sx{k, cx, c2, c3, 0}, sy{k, cy, c2, c3, 0};
    double x = gamma_draw(sx, a, 1.0), y = gamma_draw(sy, b, 1.0);
    return x / (x + y);
}

// Inverse-CDF categorical over k non-negative weights: the first i with
// u * total < p[0] + ... + p[i], or the last positive-weight index when
// rounding puts the target past the end.  Written as selects, so the lanes
// of a wave never branch apart.
template <int K>
PA_HD int categorical(const double *p, double u)
{
    double tot = 0;
    int last = 0;
#pragma unroll
    for (int i = 0; i < K; i++) {
        tot += p[i];
        last = p[i] > 0 ? i : last;
    }
    const double target = u * tot;
    double cum[K > 1 ? K - 1 : 1];
    double run = 0;
#pragma unroll
    for (int i = 0; i < K - 1; i++) {
        run += p[i];
        cum[i] = run;
    }
    int res = last;
#pragma unroll
    for (int i = K - 2; i >= 0; i--) res = target < cum[i] ? i : res;
    return res;
}

// log of the Gamma(shape a, scale b) density, closed form
PA_HD double log_gamma_pdf(double x, double a, double b)
{
    if (x < 0) return -INFINITY;
    if (x == 0) {
        if (a == 1) return -log(b);
        return a > 1 ? -INFINITY : INFINITY;
    }
    return (a - 1) * log(x) - x / b - lgamma(a) - a * log(b);
}

}  // namespace pa
