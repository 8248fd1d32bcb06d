// Mutation fuzz loop over the stream readers (test infrastructure): the
// reference's fuzz target (crates/mars-codec/fuzz/fuzz_targets/
// mars_format_read.rs:1-17) feeds arbitrary bytes to mars_format::read and,
// when that succeeds, decodes the leaves one iteration.  Here the inputs are
// mutations of seed streams (bit flips, byte overwrites, truncations,
// insertions, header-field rewrites, splices of two seeds) and every input
// also goes to the `.ifs` parser.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/fuzz/Makefile): any out-of-bounds
// access, leak or undefined operation aborts the run.  Every call must
// return 0 or a negative status.
//
// usage: fuzz_read <iterations> <seed> <stream files...>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "mars_codec.h"

namespace {

uint64_t rng_state;
uint64_t next_u64()
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return rng_state;
}
uint32_t below(uint32_t n) { return n ? (uint32_t)(next_u64() % n) : 0; }

std::vector<uint8_t> load(const char *path)
{
    std::vector<uint8_t> d;
    FILE *f = fopen(path, "rb");
    if (!f) return d;
    int c;
    while ((c = fgetc(f)) != EOF) d.push_back((uint8_t)c);
    fclose(f);
    return d;
}

std::vector<uint8_t> mutate(const std::vector<std::vector<uint8_t>> &seeds)
{
    std::vector<uint8_t> d = seeds[below((uint32_t)seeds.size())];
    const uint32_t n_ops = 1 + below(4);
    for (uint32_t k = 0; k < n_ops; k++) {
        switch (below(7)) {
        case 0: // flip a bit
            if (!d.empty()) d[below((uint32_t)d.size())] ^= (uint8_t)(1u << below(8));
            break;
        case 1: // overwrite a byte
            if (!d.empty()) d[below((uint32_t)d.size())] = (uint8_t)next_u64();
            break;
        case 2: // truncate
            d.resize(below((uint32_t)d.size() + 1));
            break;
        case 3: // insert random bytes
        {
            const uint32_t at = below((uint32_t)d.size() + 1), n = 1 + below(8);
            std::vector<uint8_t> ins(n);
            for (auto &b : ins) b = (uint8_t)next_u64();
            d.insert(d.begin() + at, ins.begin(), ins.end());
            break;
        }
        case 4: // rewrite a header field byte with an edge value
            if (d.size() >= 24) {
                static const uint8_t edge[] = {0, 1, 2, 3, 15, 16, 20, 21, 24, 31, 32, 127, 255};
                d[4 + below(20)] = edge[below(sizeof edge)];
            }
            break;
        case 5: // splice the tail of another seed
        {
            const std::vector<uint8_t> &o = seeds[below((uint32_t)seeds.size())];
            const uint32_t a = below((uint32_t)d.size() + 1), b = below((uint32_t)o.size() + 1);
            d.resize(a);
            d.insert(d.end(), o.begin() + b, o.end());
            break;
        }
        default: // rewrite the TREE length
            if (d.size() >= 29) {
                const uint32_t v = (uint32_t)next_u64() % (uint32_t)(d.size() + 8);
                for (int i = 0; i < 4; i++) d[25 + i] = (uint8_t)(v >> (8 * i));
            }
            break;
        }
    }
    return d;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: fuzz_read <iterations> <seed> <stream files...>\n");
        return 2;
    }
    const long iters = strtol(argv[1], nullptr, 10);
    rng_state = strtoull(argv[2], nullptr, 10) * 0x9E3779B97F4A7C15ull + 1;
    std::vector<std::vector<uint8_t>> seeds;
    for (int i = 3; i < argc; i++) {
        std::vector<uint8_t> d = load(argv[i]);
        if (!d.empty()) seeds.push_back(d);
    }
    if (seeds.empty()) return 2;
    std::map<int, long> read_status, ifs_status;
    long decoded = 0;
    for (long it = 0; it < iters; it++) {
        const std::vector<uint8_t> d = it < (long)seeds.size() ? seeds[it] : mutate(seeds);
        mars_header h;
        mars_leaf *leaves = nullptr;
        int32_t *res = nullptr;
        size_t n = 0, nres = 0;
        int rc = mars_read(d.data(), d.size(), &h, &leaves, &n, &res, &nres);
        read_status[rc]++;
        if (rc > 0) {
            fprintf(stderr, "mars_read returned a positive status %d\n", rc);
            return 1;
        }
        if (rc == 0 && (uint64_t)h.width * h.height <= (1u << 22)) {
            std::vector<uint8_t> img((size_t)h.width * h.height);
            const int drc = mars_decode_iterative(&h, leaves, n, res, nres, 1, img.data());
            if (drc != 0) {
                fprintf(stderr, "decode of accepted leaves failed: %d\n", drc);
                return 1;
            }
            decoded++;
        }
        mars_free(leaves);
        mars_free(res);
        rc = mars_ifs_parse(d.data(), d.size(), &h, &leaves, &n);
        ifs_status[rc]++;
        if (rc == 0 && (uint64_t)h.width * h.height <= (1u << 22)) {
            std::vector<uint8_t> img((size_t)h.width * h.height);
            if (mars_decode_iterative(&h, leaves, n, nullptr, 0, 1, img.data()) != 0) {
                fprintf(stderr, "decode of accepted .ifs leaves failed\n");
                return 1;
            }
        }
        mars_free(leaves);
    }
    printf("iterations %ld decoded %ld\nmars_read:", iters, decoded);
    for (auto &kv : read_status) printf(" %d:%ld", kv.first, kv.second);
    printf("\nifs_parse:");
    for (auto &kv : ifs_status) printf(" %d:%ld", kv.first, kv.second);
    printf("\n");
    return 0;
}
