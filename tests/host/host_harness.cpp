// CPU harness for the library's host-side C++ that reads untrusted input: the glTF binary reader
// (csrc/mpm_gltf.h) and the collider BVH builder (csrc/mpm_bvh.h). Built by
// tests/test_host_sanitizers.py with AddressSanitizer and UndefinedBehaviorSanitizer and run over
// the test-written GLB files, byte-level mutations of them, and random triangle lists (indices
// out of range, non-finite and degenerate coordinates). Any sanitizer report aborts the run.
//
//   host_harness FILE...                     load each GLB, then build a BVH over its arrays
//   host_harness --mutate FILE N SEED        N mutated copies of FILE through the reader
//   host_harness --bvh N SEED                N random triangle lists through the BVH builder
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "mpm_gltf.h"
#include "mpm_bvh.h"

static uint64_t rng_state = 1;
static uint32_t rnd() {   // splitmix64, high half
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((z ^ (z >> 31)) >> 32);
}

static int run_glb(const std::vector<uint8_t>& buf, bool verbose) {
    mpm_gltf::Scene sc;
    mpm_gltf::Loader ld;
    bool ok = false;
    try {
        ok = ld.load(buf.data(), buf.size(), sc);
    } catch (const std::bad_alloc&) {
        ok = false;
    }
    if (!ok) { if (verbose) std::printf("rejected: %s\n", ld.err.c_str()); return 0; }
    std::vector<mpm_bvh::BNode> nodes;
    std::vector<uint32_t> order;
    const bool built = mpm_bvh::build(sc.positions.data(), sc.positions.size() / 3, sc.indices.data(), sc.indices.size(), nodes, order);
    if (built) {
        std::vector<uint8_t> packed(nodes.size() * 32);
        mpm_bvh::pack(nodes, packed.data());
    }
    if (verbose) std::printf("loaded: %zu vertices %zu indices %zu objects, bvh %s (%zu nodes)\n", sc.positions.size() / 3,
                             sc.indices.size(), sc.objects.size(), built ? "built" : "rejected", nodes.size());
    return 0;
}

static std::vector<uint8_t> read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

int main(int argc, char** argv) {
    if (argc >= 5 && std::string(argv[1]) == "--mutate") {
        const std::vector<uint8_t> base = read_file(argv[2]);
        const long n = std::atol(argv[3]);
        rng_state = std::strtoull(argv[4], nullptr, 10);
        for (long it = 0; it < n; ++it) {
            std::vector<uint8_t> b = base;
            const int kind = rnd() % 4;
            const int edits = 1 + rnd() % 8;
            for (int e = 0; e < edits && !b.empty(); ++e) {
                const size_t at = rnd() % b.size();
                if (kind == 0) b[at] = (uint8_t)rnd();                                       // random byte
                else if (kind == 1) b[at] = (uint8_t)"0123456789-.eE[]{},:\""[rnd() % 22];  // JSON-ish byte
                else if (kind == 2) b.resize(at);                                             // truncation
                else b.insert(b.begin() + at, (uint8_t)"9e-[{\",]}"[rnd() % 9]);            // insertion
            }
            run_glb(b, false);
        }
        std::printf("mutations: %ld\n", n);
        return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "--bvh") {
        const long n = std::atol(argv[2]);
        rng_state = std::strtoull(argv[3], nullptr, 10);
        for (long it = 0; it < n; ++it) {
            const size_t nv = rnd() % 64, ni = 3 * (rnd() % 80);
            std::vector<float> pos(3 * nv + 3);
            for (float& x : pos) {
                const uint32_t r = rnd() % 100;
                x = r < 2 ? NAN : r < 4 ? INFINITY : r < 20 ? 0.5f : (float)(rnd() % 2001) / 1000.0f - 1.0f;
            }
            std::vector<uint32_t> idx(ni);
            for (uint32_t& i : idx) i = (rnd() % 50 == 0) ? (uint32_t)(nv + rnd() % 5) : (nv ? rnd() % nv : 0u);
            std::vector<mpm_bvh::BNode> nodes;
            std::vector<uint32_t> order;
            if (mpm_bvh::build(pos.data(), nv, idx.data(), idx.size(), nodes, order)) {
                std::vector<uint8_t> packed(nodes.size() * 32);
                mpm_bvh::pack(nodes, packed.data());
            }
        }
        std::printf("bvh inputs: %ld\n", n);
        return 0;
    }
    for (int i = 1; i < argc; ++i) {
        std::printf("%s: ", argv[i]);
        run_glb(read_file(argv[i]), true);
    }
    return 0;
}
