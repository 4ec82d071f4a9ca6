// Device-side building blocks shared by the slgpu kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slgpu {

// physconstmod.f:11-51
constexpr double PC_ACOEF = 7.5657e-15;
constexpr double PC_PI = 3.1415926535897932385;
constexpr double PC_PI2 = 2.0 * PC_PI;
constexpr double PC_PI4 = 4.0 * PC_PI;
constexpr double PC_C = 2.997924562e10;
constexpr double PC_KB = 1.3806505e-16;
constexpr double PC_ME = 9.1093826e-28;
constexpr double PC_H = 6.6260693e-27;
constexpr double PC_E = 4.80325e-10;
constexpr double PC_RYDBERG = 1.0973731534e5;
constexpr double PC_DEXT = 0.7104;
constexpr double PC_EV = 1.60217653e-12;
constexpr double PC_ANG = 1e-8;
constexpr double PC_MBARN = 1e-18;
constexpr double CINV = 1.0 / PC_C;
constexpr double HCK = PC_H * PC_C / PC_KB;
constexpr double FTPI4 = 15.0 / (PC_PI * PC_PI * PC_PI * PC_PI);
constexpr double FTPI4T = 15.0 / (4.0 * PC_PI * PC_PI * PC_PI * PC_PI);
constexpr int FF_NU = 81, FF_NGG = 41;

// MISC/binsrch.f90:3-77: interval index of x in arr[0..n-1], 1-based result
// (0..n with widerange, clamped to 1..n-1 without; -1 for NaN).
__device__ __forceinline__ int binsrch(double x, const double *__restrict__ arr, int n,
                                       bool widerange)
{
    if (n == 2) {
        int r;
        if (x >= arr[0] && x <= arr[1]) r = 1;
        else if (x < arr[0]) r = 0;
        else if (x > arr[1]) r = 2;
        else return -1;
        return widerange ? r : 1;
    }
    int imin = 1, imax = n, imid = (n + 1) / 2;
    while (imax - imin > 1) {
        double am = arr[imid - 1];
        if (x >= arr[imin - 1] && x < am) {
            imax = imid;
            imid = (imax + imin) / 2;
        } else if (x >= am && x <= arr[imax - 1]) {
            imin = imid;
            imid = (imax + imin) / 2;
        } else {
            if (x < arr[0]) { imin = 0; imid = 0; break; }
            else if (x > arr[n - 1]) { imin = n; imid = n; break; }
            else return -1;
        }
    }
    if (imid != imin) imid = -1;
    if (!widerange) imid = imid > n - 1 ? n - 1 : (imid < 1 ? 1 : imid);
    return imid;
}

// groupmod.f:263-270 and :395-402 integrands, x capped at 300 against overflow
__device__ __forceinline__ double fplanck(double x)
{
    if (x >= 300.0) x = 300.0;
    return x * x * x / (exp(x) - 1.0);
}

__device__ __forceinline__ double fdplanck(double x)
{
    if (x >= 300.0) x = 300.0;
    double e = exp(x);
    return (x * x * x * x * e) / ((e - 1.0) * (e - 1.0));
}

// one element of specintv mode 0 (groupmod.f:185-197) for group ig (1-based)
__device__ __forceinline__ double specintv_g(double tempinv, const double *__restrict__ wlinv,
                                             int ig)
{
    double a = HCK * tempinv;
    double x0 = a * wlinv[ig - 1], x1 = a * wlinv[ig];
    return FTPI4 * fabs(x1 - x0) * fplanck(0.5 * (x1 + x0));
}

// one element of spectintv mode 0 with the NaN guard (groupmod.f:314-327, 390-392)
__device__ __forceinline__ double spectintv_g(double tempinv, const double *__restrict__ wlinv,
                                              int ig)
{
    double a = HCK * tempinv;
    double x0 = a * wlinv[ig - 1], x1 = a * wlinv[ig];
    double v = FTPI4T * fabs(x1 - x0) * fdplanck(0.5 * (x1 + x0));
    return (v != v) ? 0.0 : v;
}

// groupmod.f:99-120
__device__ __forceinline__ double specint0(double tempinv, const double *__restrict__ wlinv,
                                           int ig)
{
    double x = HCK * tempinv;
    double dx = x * fabs(wlinv[ig] - wlinv[ig - 1]);
    x = x * 0.5 * (wlinv[ig] + wlinv[ig - 1]);
    if (x >= 300.0) x = 300.0;
    return FTPI4 * dx * (x * x * x) / (exp(x) - 1.0);
}

// groupmod.f:123-143
__device__ __forceinline__ double dopspeccalc(double tempinv, const double *__restrict__ wlinv,
                                              int ig)
{
    double x = HCK * tempinv * wlinv[ig];
    if (x >= 300.0) x = 300.0;
    return FTPI4 * (x * x * x * x) / (exp(x) - 1.0);
}

// physconstmod.f:79-93
__device__ __forceinline__ double planck(double wl, double temp)
{
    constexpr double c1 = 2.0 * PC_H * PC_C * PC_C;
    constexpr double c2 = PC_H * PC_C / PC_KB;
    double b = 1.0 / (wl * temp);
    if (c2 * b >= 300.0) b = 300.0 / c2;
    double tb = temp * b;
    return c1 * (tb * tb * tb * tb * tb) / (exp(c2 * b) - 1.0);
}

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox-4x32-10 (Salmon et al., SC'11).  One stream per particle, key
// (seed, iteration); uniforms lie strictly inside (0,1).  (The reference's rnd_r has 32
// bits of resolution, randommod.f:73-92.)
//   Philox (sources): draw k uses counter (particle, k/2, phase, 0); words (0,1) serve even
//   draws and (2,3) odd ones, 53-bit uniforms.
//   StepRng (transport): draw j of step s uses counter (particle, s, phase, j/3); a block's
//   three draws are 42-bit uniforms, word j%3 above the 10 bits of word 3 from bit
//   10(j%3).  One block covers a DDMC step's three draws, and an IMC step's two distances
//   and first event draw.  begin_step computes the step's first block at the top of the
//   step, where every lane of the pass is active.
struct Philox {
    uint32_t k0, k1, pid, phase, ndraw;
    uint32_t w2, w3;

    __device__ __forceinline__ void init(uint32_t seed, uint32_t iter, uint32_t p, uint32_t ph)
    {
        k0 = seed; k1 = iter; pid = p; phase = ph; ndraw = 0;
        w2 = w3 = 0;
    }

    __device__ __forceinline__ static void block(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1,
                                                 uint32_t out[4])
    {
        // Two rounds per loop trip: the rotation of (c0..c3) then costs no register moves,
        // and the body stays small for the instruction cache.  Each 32x32 product is formed
        // as one 64-bit product (one v_mad_u64_u32 instead of v_mul_lo_u32 + v_mul_hi_u32;
        // same bits), and each three-input XOR is one v_bitop3_b32 (truth table 0x96 =
        // a ^ b ^ c): four VALU instructions per round.
        #pragma unroll 2
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
            const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
            c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }

    // (m + 0.5) * 2^-53 for the 53-bit m = (a >> 5) * 2^26 + (b >> 6).  The fma rounds the
    // exact m + 0.5 once, as (double)m + 0.5 does (m itself is exact), without the
    // 64-bit-integer-to-double sequence.
    __device__ __forceinline__ static double u01(uint32_t a, uint32_t b)
    {
        return fma((double)(a >> 5), 67108864.0, (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
    }

    // (m + 0.5) * 2^-42 for the 42-bit m = hi * 2^10 + lo (exact in FP64)
    __device__ __forceinline__ static double u42(uint32_t hi, uint32_t lo)
    {
        return fma((double)hi, 1024.0, (double)lo + 0.5) * (1.0 / 4398046511104.0);
    }

    __device__ __forceinline__ double draw()
    {
        const uint32_t k = ndraw++;
        if ((k & 1u) == 0) {
            uint32_t o[4];
            block(pid, k >> 1, phase, 0u, k0, k1, o);
            w2 = o[2]; w3 = o[3];
            return u01(o[0], o[1]);
        }
        return u01(w2, w3);
    }
};

struct StepRng {
    static constexpr uint32_t PHASE = 1;    // the transport's counter phase (sources: 0)
    uint32_t k0, k1;        // key (seed, iteration): set once, the same in every lane
    uint32_t pid, step;
    uint32_t cnt;           // 4 * (block index) + draws taken from the current block
    uint32_t a, b, c, lo;   // a serves the next draw; lo holds its 10 low bits at the bottom

    __device__ __forceinline__ void key(uint32_t seed, uint32_t iter) { k0 = seed; k1 = iter; }

    // a new particle's stream
    __device__ __forceinline__ void init(uint32_t p)
    {
        pid = p; step = 0; cnt = 0;
        a = b = c = lo = 0;
    }

    __device__ __forceinline__ void fill(uint32_t j)
    {
        uint32_t o[4];
        Philox::block(pid, step, PHASE, j, k0, k1, o);
        a = o[0]; b = o[1]; c = o[2]; lo = o[3];
    }

    // start transport step s with its first block computed
    __device__ __forceinline__ void begin_step(uint32_t s)
    {
        step = s; cnt = 0;
        fill(0u);
    }

    __device__ __forceinline__ double draw()
    {
        if ((cnt & 3u) == 3u) { cnt += 1u; fill(cnt >> 2); }
        const double u = Philox::u42(a, lo & 0x3FFu);
        a = b; b = c; lo >>= 10; ++cnt;
        return u;
    }
};

// wave64 inclusive scan of doubles (fixed order: result is deterministic)
__device__ __forceinline__ double wave_inclusive_scan(double v)
{
    const int lane = threadIdx.x & 63;
    #pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// log2(e): the precision form evaluates exp(-chilw*hckt) as 2^(-chilw * hckt*log2(e)) with
// hckt*log2(e) rounded once to FP32 per cell (v_exp_f32 computes 2^x)
constexpr double LOG2E = 1.4426950408889634;

// Block-wide deterministic sum (fixed tree) for 256-thread blocks.
__device__ __forceinline__ double block_sum256(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// order LDS writes of a wave before its later reads of them (each wave owns its slot)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace slgpu
