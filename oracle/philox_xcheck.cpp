// Prints Philox4x32-10 blocks computed by rocRAND's header implementation, one
// block per input line "c0 c1 c2 c3 k0 k1" (hex), as "r0 r1 r2 r3" (hex).
// Test infrastructure: an independent implementation against which the
// oracle's Philox is checked on the CPU. Host code only.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct philox_probe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    philox_probe e;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 c; c.x = c0; c.y = c1; c.z = c2; c.w = c3;
        uint2 k; k.x = k0; k.y = k1;
        uint4 r = e.block(c, k);
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
