// Golden-vector driver for the reference's counter-based PRNG (test infrastructure).
//
// Built only by `make ref` in this folder, against the reference's own header
// prng_engine.hpp (its class sitmo::prng_engine, prng_engine.hpp:88-380), into
// oracle/_ref/.  It prints the engine's outputs for a fixed list of seeds and
// seed/discard/draw scripts; tests/golden/make_prng_vectors.py turns the text
// into the committed fixture that pins the oracle's restatement.
//
// Each script is a sequence of steps: "d<n>" draws n outputs, "s<z>" calls
// discard(z).  Output: one line per script, "seed|script|v0 v1 ...".
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>
#include "prng_engine.hpp"

static void run(uint32_t seed, const std::vector<std::string>& steps) {
    sitmo::prng_engine e(seed);
    std::string script;
    std::printf("%u|", seed);
    for (size_t i = 0; i < steps.size(); i++) std::printf("%s%s", i ? "," : "", steps[i].c_str());
    std::printf("|");
    bool first = true;
    for (const std::string& st : steps) {
        unsigned long long n = std::strtoull(st.c_str() + 1, nullptr, 10);
        if (st[0] == 'd') {
            for (unsigned long long k = 0; k < n; k++) {
                std::printf(first ? "%u" : " %u", (unsigned) e());
                first = false;
            }
        } else {
            e.discard((uint64_t) n);
        }
    }
    std::printf("\n");
}

int main() {
    const uint64_t S = 65536;  // the IC generator's row stride (ic_basic.hpp:31-33)
    const uint32_t seeds[] = {0u, 1u, 42u, 12345u, 0xFFFFFFFFu};
    for (uint32_t s : seeds) run(s, {"d40"});
    const uint64_t skips[] = {0, 1, 5, 7, 8, 9, 15, 16, 17, 63, 64, 65, S - 2, S, S * S,
                              S * S + 3 * S + 11, ((S + 3) * S + 5) * S, ((2 * S + 7) * S + 2) * S,
                              ((3 * S + 1) * S + 1) * S, (1ull << 40) + 3, 0x7FFFFFFFFFFFFFF0ull};
    for (uint64_t z : skips) run(42u, {"s" + std::to_string(z), "d12"});
    // within-block bookkeeping: partial blocks, zero-length discards, block-crossing discards
    run(42u, {"d3", "s4", "d2", "s0", "d1", "s13", "d5", "s8", "d9", "s1", "d16"});
    run(7u, {"d1", "s6", "d1", "s1", "d1", "s7", "d2", "s65534", "d4"});
    run(2024u, {"d8", "s0", "d8", "s8", "d8", "s3", "d3"});
    return 0;
}
