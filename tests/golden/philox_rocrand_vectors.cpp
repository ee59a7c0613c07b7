// Writes Philox4x32-10 vectors computed by rocRAND's independent implementation
// (rocrand_device::philox4x32_10_engine, host-callable) as JSON lines
// {"ctr":[..4..],"key":[..2..],"out":[..4..]} to stdout.
// The engine's counter for (subsequence s, offset o) is (o/4 lo, o/4 hi, s lo, s hi)
// and its key is (seed lo, seed hi), so any (ctr, key) pair can be reached.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

static void emit(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    const unsigned long long seed = ((unsigned long long)k1 << 32) | k0;
    const unsigned long long subseq = ((unsigned long long)c3 << 32) | c2;
    const unsigned long long offset = (((unsigned long long)c1 << 32) | c0) * 4ull;
    rocrand_device::philox4x32_10_engine e(seed, subseq, offset);
    uint32_t o[4];
    for (int i = 0; i < 4; ++i) o[i] = e();
    std::printf("{\"ctr\":[%u,%u,%u,%u],\"key\":[%u,%u],\"out\":[%u,%u,%u,%u]}\n",
                c0, c1, c2, c3, k0, k1, o[0], o[1], o[2], o[3]);
}

int main() {
    emit(0, 0, 0, 0, 0, 0);
    // Counter words 0-1 must satisfy (c1:c0)*4 < 2^64 to be reachable through the
    // engine's 64-bit offset; the full-range KATs are checked from Random123's list.
    emit(0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    emit(0x243f6a88u, 0x05a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u);
    // Agent-stream shaped counters: (draw block, stream, step, id), key = seed.
    uint32_t x = 12345u;
    for (int i = 0; i < 64; ++i) {
        x = x * 1664525u + 1013904223u; uint32_t a = x >> 8;
        x = x * 1664525u + 1013904223u; uint32_t b = x & 0xff;
        x = x * 1664525u + 1013904223u; uint32_t c = x >> 12;
        x = x * 1664525u + 1013904223u; uint32_t d = x;
        x = x * 1664525u + 1013904223u; uint32_t k0 = x;
        emit(b & 3, a & 15, c, d, k0, (uint32_t)i);
    }
    return 0;
}
