// Writes Philox4x32-10 output blocks computed by rocRAND's device generator
// header (rocrand_philox4x32_10.h, compiled for the host) for the counters and
// keys of philox_vectors.json.  Build and run:
//   hipcc -x hip --offload-arch=gfx950 tests/golden/make_philox_vectors.cpp -o /tmp/mkphilox
//   /tmp/mkphilox > tests/golden/philox_rocrand.txt
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    Probe p;
    unsigned long long s = 0x9E3779B97F4A7C15ull;
    for (int t = 0; t < 64; ++t) {
        // deterministic counters/keys from a splitmix64 sequence
        unsigned v[6];
        for (int j = 0; j < 6; ++j) {
            s += 0x9E3779B97F4A7C15ull;
            unsigned long long z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            v[j] = (unsigned)(z ^ (z >> 31));
        }
        uint4 r = p.block(uint4{v[0], v[1], v[2], v[3]}, uint2{v[4], v[5]});
        printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", v[0], v[1], v[2], v[3], v[4], v[5], r.x,
               r.y, r.z, r.w);
    }
    return 0;
}
