// synth.hpp — seeded synthetic signals generated on the device
// (fmdbzp_synthesize; the analogue of src/generate_test_signal.jl:32-82).
//
// Every operation follows the host generator (acquisition/rng.py,
// acquisition/signals.py) in FP64 with separately rounded operations, so the
// device samples equal the host ones up to the last bit of log / sqrt / cos /
// sin, which can move a sample by one unit in the last place of the output.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct SynthEmitter {
    int64_t chips_off, sec_off, bits_off;  // offsets into the device sign table
    int32_t code_length, boc, sec_len, sec_phase, num_bits, bit_ms, bit_edge_ms, pad;
    double amplitude, code_freq, code_phase, carrier_hz, carrier_phase;
};

// Philox4x32-10 (Salmon et al., SC'11): counter (c0, c1, c2, c3), key (k0, k1).
__device__ __forceinline__ void philox4x32_10(uint32_t& x0, uint32_t& x1, uint32_t& x2, uint32_t& x3, uint32_t k0,
                                              uint32_t k1)
{
#pragma unroll
    for (int rnd = 0; rnd < 10; rnd++) {
        if (rnd) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)x0 * 0xD2511F53ull;
        const uint64_t p1 = (uint64_t)x2 * 0xCD9E8D57ull;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ x1 ^ k0, n2 = hi0 ^ x3 ^ k1;
        x0 = n0;
        x1 = lo1;
        x2 = n2;
        x3 = lo0;
    }
}

// Two standard normals from counter block `block` of stream `stream`.
__device__ __forceinline__ void philox_normal_pair(uint64_t block, uint32_t stream, uint32_t k0, uint32_t k1,
                                                   double* g0, double* g1)
{
    uint32_t x0 = (uint32_t)block, x1 = (uint32_t)(block >> 32), x2 = stream, x3 = 0;
    philox4x32_10(x0, x1, x2, x3, k0, k1);
    const double two53 = 1.0 / 9007199254740992.0;
    const double a = __dmul_rn(__dadd_rn(__dmul_rn((double)(x0 >> 5), 67108864.0), (double)(x1 >> 6)), two53);
    const double b = __dmul_rn(__dadd_rn(__dmul_rn((double)(x2 >> 5), 67108864.0), (double)(x3 >> 6)), two53);
    const double u1 = __dadd_rn(1.0, -a);
    const double r = sqrt(__dmul_rn(-2.0, log(u1)));
    const double t = __dmul_rn(6.283185307179586, b);  // 2.0 * pi
    *g0 = __dmul_rn(r, cos(t));
    *g1 = __dmul_rn(r, sin(t));
}

// Sum of the emitters' terms at sample n, added in emitter order onto (re, im).
__device__ __forceinline__ void synth_emitters(const SynthEmitter* __restrict__ em, int nem,
                                               const int8_t* __restrict__ table, int64_t n, double fs, int cplx,
                                               double* re, double* im)
{
    const double nd = (double)n;
    for (int e = 0; e < nem; e++) {
        const SynthEmitter E = em[e];
        const double pos = __dadd_rn(E.code_phase, __ddiv_rn(__dmul_rn(nd, E.code_freq), fs));
        const double fpos = floor(pos);
        const int64_t c = (int64_t)fpos;
        const int64_t L = E.code_length;
        int64_t ci = c % L;
        if (ci < 0) ci += L;
        int chip = table[E.chips_off + ci];
        if (E.boc && __dadd_rn(pos, -fpos) >= 0.5) chip = -chip;
        if (E.sec_len > 0) {
            int64_t per = c >= 0 ? c / L : -((-c + L - 1) / L);
            int64_t si = (per + E.sec_phase) % E.sec_len;
            if (si < 0) si += E.sec_len;
            chip *= table[E.sec_off + si];
        }
        if (E.num_bits > 0) {
            const int64_t ms = (int64_t)floor(__ddiv_rn(__dmul_rn(nd, 1000.0), fs));
            int64_t bi = (ms + E.bit_edge_ms) / E.bit_ms;
            if (bi >= E.num_bits) bi = E.num_bits - 1;
            chip *= table[E.bits_off + bi];
        }
        const double s = __dmul_rn(6.283185307179586, E.carrier_hz);
        const double ang = __dadd_rn(__ddiv_rn(__dmul_rn(s, nd), fs), E.carrier_phase);
        const double a = __dmul_rn(E.amplitude, (double)chip);
        *re = __dadd_rn(*re, __dmul_rn(a, cos(ang)));
        if (cplx) *im = __dadd_rn(*im, __dmul_rn(a, sin(ang)));
    }
}

__device__ __forceinline__ int16_t sat_rint16(double x)
{
    double v = rint(x);
    if (v < -32768.0) v = -32768.0;
    if (v > 32767.0) v = 32767.0;
    return (int16_t)v;
}

// Real scenes: thread g makes samples 2g and 2g + 1 (one counter block).
// Complex scenes: thread g makes sample g. Output int16 (real) or float2.
__global__ void __launch_bounds__(256)
k_synthesize(const SynthEmitter* __restrict__ em, int nem, const int8_t* __restrict__ table, int64_t n_samples,
             double fs, int cplx, int out_i16, double noise_scale, uint32_t k0, uint32_t k1, uint32_t stream,
             void* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cplx) {
        if (g >= n_samples) return;
        double g0, g1;
        philox_normal_pair((uint64_t)g, stream, k0, k1, &g0, &g1);
        double re = __dmul_rn(g0, noise_scale), im = __dmul_rn(g1, noise_scale);
        synth_emitters(em, nem, table, g, fs, 1, &re, &im);
        ((float2*)out)[g] = make_float2(__double2float_rn(re), __double2float_rn(im));
        return;
    }
    const int64_t i0 = 2 * g;
    if (i0 >= n_samples) return;
    double g0, g1;
    philox_normal_pair((uint64_t)g, stream, k0, k1, &g0, &g1);
    double v[2] = {__dmul_rn(g0, noise_scale), __dmul_rn(g1, noise_scale)};
    for (int h = 0; h < 2 && i0 + h < n_samples; h++) {
        double im = 0.0;
        synth_emitters(em, nem, table, i0 + h, fs, 0, &v[h], &im);
        if (out_i16) ((int16_t*)out)[i0 + h] = sat_rint16(v[h]);
        else ((float*)out)[i0 + h] = __double2float_rn(v[h]);
    }
}
