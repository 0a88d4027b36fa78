// Writes Philox4x32-10 known-answer vectors using PyTorch's header-only
// engine (ATen/core/PhiloxRNGEngine.h) as an independent implementation.
// Build/run: see gen_golden.py (make_philox_kat).
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdint>

int main()
{
    const uint64_t seeds[] = {0ull, 1ull, 1000ull, 0xFFFFFFFFFFFFFFFFull, 0x299F31D0A4093822ull, 67280421310721ull};
    const uint64_t subs[] = {0ull, 1ull, 0x0370734413198A2Eull, 0xFFFFFFFFFFFFFFFFull, 12345ull};
    const uint64_t offs[] = {0ull, 7ull, 0x85A308D3243F6A88ull, 0xFFFFFFFFFFFFFFFFull};
    std::printf("[\n");
    bool first = true;
    for (uint64_t s : seeds)
        for (uint64_t q : subs)
            for (uint64_t o : offs) {
                at::philox_engine e(s, q, 0);
                e.set_offset(o);
                uint32_t w[4];
                for (int i = 0; i < 4; ++i) w[i] = e();
                // counter words (c0,c1,c2,c3) = (offset lo, offset hi, subsequence lo, subsequence hi), key = seed
                std::printf("%s  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}",
                            first ? "" : ",\n", (uint32_t)o, (uint32_t)(o >> 32), (uint32_t)q, (uint32_t)(q >> 32),
                            (uint32_t)s, (uint32_t)(s >> 32), w[0], w[1], w[2], w[3]);
                first = false;
            }
    std::printf("\n]\n");
    return 0;
}
