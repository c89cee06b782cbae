// Writes Philox4x32-10 vectors from rocRAND's host-compilable engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), an implementation
// independent of this repository's, to tests/golden/philox_rocrand.txt.
// Build + run: tests/golden/make_philox_rocrand.sh
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

// The raw bijection (counter, key) -> 4 words is the engine's protected
// ten_rounds(); a derived class exposes it without touching the state logic.
struct RawPhilox : rocrand_device::philox4x32_10_engine {
    uint4 raw(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

static void emit(FILE* f, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                 uint32_t k1) {
    RawPhilox e;
    uint4 r = e.raw(uint4{c0, c1, c2, c3}, uint2{k0, k1});
    fprintf(f, "%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c0, c1, c2, c3, k0, k1,
            r.x, r.y, r.z, r.w);
}

int main(int argc, char** argv) {
    FILE* f = fopen(argc > 1 ? argv[1] : "philox_rocrand.txt", "w");
    if (!f) return 1;
    fprintf(f, "# c0 c1 c2 c3 k0 k1 r0 r1 r2 r3 (hex) from rocrand philox4x32_10_engine\n");
    uint64_t x = 0x243F6A8885A308D3ull;  // xorshift64 for the counters/keys
    auto next = [&]() {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return (uint32_t)(x >> 16);
    };
    emit(f, 0, 0, 0, 0, 0, 0);
    emit(f, 1, 0, 0, 0, 1, 0);
    for (uint32_t s = 1; s <= 15; ++s) emit(f, s, 7852, 0, 0, 1, 0);  // a photon's step blocks
    for (int i = 0; i < 48; ++i) {
        uint32_t c0 = next(), c1 = next(), c2 = next(), c3 = next(), k0 = next(), k1 = next();
        emit(f, c0, c1, c2, c3, k0, k1);
    }
    fclose(f);
    return 0;
}
