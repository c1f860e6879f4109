// SHA-256 as specified in FIPS 180-4: 512-bit blocks, 64 rounds, message
// padded with a 1 bit, zeros and the 64-bit big-endian bit length. The
// constants are computed from their definitions: K[t] is the first 32 bits
// of the fractional part of the cube root of the t-th prime and H0[i] those
// of the square root of the i-th prime, taken here as exact integer roots
// (floor(cbrt(p * 2^96)) and floor(sqrt(p * 2^64)), modulo 2^32).
#include "sha256.hpp"

#include <array>
#include <cstring>

namespace blitz_hip {
namespace {

using u128 = unsigned __int128;

// Largest r with r^k <= v (k = 2 or 3), by bisection.
uint64_t integer_root(u128 v, int k) {
    uint64_t lo = 0, hi = uint64_t(1) << 40;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        u128 p = mid;
        for (int i = 1; i < k; ++i) p *= mid;
        if (p <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct Constants {
    std::array<uint32_t, 64> k;
    std::array<uint32_t, 8> h;
    Constants() {
        uint32_t primes[64];
        int n = 0;
        for (uint32_t c = 2; n < 64; ++c) {
            bool prime = true;
            for (int i = 0; i < n && primes[i] * primes[i] <= c; ++i)
                if (c % primes[i] == 0) prime = false;
            if (prime) primes[n++] = c;
        }
        for (int t = 0; t < 64; ++t) k[t] = uint32_t(integer_root(u128(primes[t]) << 96, 3));
        for (int i = 0; i < 8; ++i) h[i] = uint32_t(integer_root(u128(primes[i]) << 64, 2));
    }
};

const Constants& constants() {
    static const Constants c;
    return c;
}

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void compress(std::array<uint32_t, 8>& state, const uint8_t* block, const std::array<uint32_t, 64>& k) {
    uint32_t w[64];
    for (int t = 0; t < 16; ++t)
        w[t] = uint32_t(block[4 * t]) << 24 | uint32_t(block[4 * t + 1]) << 16 | uint32_t(block[4 * t + 2]) << 8 |
               uint32_t(block[4 * t + 3]);
    for (int t = 16; t < 64; ++t) {
        const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
        const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t a = state[0], b = state[1], c = state[2], d = state[3], e = state[4], f = state[5], g = state[6],
             h = state[7];
    for (int t = 0; t < 64; ++t) {
        const uint32_t big1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        const uint32_t choose = (e & f) ^ (~e & g);
        const uint32_t t1 = h + big1 + choose + k[t] + w[t];
        const uint32_t big0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        const uint32_t majority = (a & b) ^ (a & c) ^ (b & c);
        const uint32_t t2 = big0 + majority;
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    state[0] += a;
    state[1] += b;
    state[2] += c;
    state[3] += d;
    state[4] += e;
    state[5] += f;
    state[6] += g;
    state[7] += h;
}

}  // namespace

std::string sha256_hex(const void* data, size_t n) {
    const Constants& c = constants();
    std::array<uint32_t, 8> state = c.h;
    const auto* bytes = static_cast<const uint8_t*>(data);
    size_t at = 0;
    for (; at + 64 <= n; at += 64) compress(state, bytes + at, c.k);
    // Tail: the remaining bytes, 0x80, zeros, the bit length (big-endian).
    uint8_t tail[128] = {};
    const size_t rest = n - at;
    if (rest) std::memcpy(tail, bytes + at, rest);
    tail[rest] = 0x80;
    const size_t blocks = rest + 1 + 8 <= 64 ? 1 : 2;
    const uint64_t bits = uint64_t(n) * 8;
    for (int i = 0; i < 8; ++i) tail[64 * blocks - 1 - i] = uint8_t(bits >> (8 * i));
    for (size_t b = 0; b < blocks; ++b) compress(state, tail + 64 * b, c.k);
    static const char digits[] = "0123456789abcdef";
    std::string out(64, '0');
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) out[8 * i + j] = digits[(state[i] >> (28 - 4 * j)) & 15u];
    return out;
}

}  // namespace blitz_hip
