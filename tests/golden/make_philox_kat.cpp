// Known-answer vectors for Philox4x32-10 taken from rocRAND's philox4x32_10 engine
// (a library implementation, /opt/rocm/include/rocrand/rocrand_philox4x32_10.h),
// evaluated on the host.  An engine seeded with (seed, subsequence, offset = 4*q)
// holds key = (seed lo, seed hi) and counter = (q lo, q hi, subsequence lo,
// subsequence hi); its first next4() is the Philox block of that counter.
//
//   hipcc -O1 -o /tmp/make_philox_kat tests/golden/make_philox_kat.cpp
//   /tmp/make_philox_kat > tests/golden/philox_kat.txt
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

int main()
{
    const unsigned long long cases[][3] = {
        {0ull, 0ull, 0ull},
        {0xffffffffffffffffull, 0xffffffffffffffffull, 0xffffffffull},
        {0x13198a2e03707344ull, 0x243f6a8885a308d3ull, 0xa4093822ull},
        {4321ull, 20ull, 7ull},
        {31337ull, 1ull, 123456ull},
        {0xdeadbeefcafef00dull, 0x0000000100000002ull, 0x00000003ull},
    };
    std::printf("# seed ctr0 ctr1 ctr2 ctr3 -> out0 out1 out2 out3 (hex), rocRAND philox4x32_10\n");
    for (const auto& c : cases) {
        unsigned long long seed = c[0], sub = c[1], q = c[2];
        rocrand_device::philox4x32_10_engine e(seed, sub, 4ull * q);
        uint4 r = e.next4();
        std::printf("%016llx %08x %08x %08x %08x %08x %08x %08x %08x\n", seed,
                    (unsigned)q, (unsigned)(q >> 32), (unsigned)sub, (unsigned)(sub >> 32),
                    r.x, r.y, r.z, r.w);
    }
    return 0;
}
